// ffhip_core.hip — context, errors, scans and device-built meshes.
//
// cube(nx,ny,nz) and square(nx,ny) are generated on the GPU with the exact
// vertex numbering, coordinates, element vertex order, boundary elements and
// labels of FreeFEM's BuildCube (src/fflib/msh3.cpp:7879-8132) and Carre_
// (src/fflib/lgmesh.cpp:1229-1384): one thread per vertex / per hex, and a
// scan for the boundary-element offsets.
#include <hipcub/hipcub.hpp>

#include <cstdarg>
#include <map>
#include <mutex>
#include <unordered_map>
#include <cstring>
#include <vector>

#include "ffhip_internal.hpp"

namespace ffhip {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void exclusive_scan_i64(ffhip_ctx ctx, const int64_t *in, int64_t *out,
                        int64_t n) {
  size_t bytes = 0;
  FFHIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n,
                                               ctx->stream));
  void *tmp = ctx->ensure_scratch(bytes);
  FFHIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n,
                                               ctx->stream));
}

int64_t read_i64(ffhip_ctx ctx, const int64_t *p) {
  int64_t v = 0;
  FFHIP_CHECK(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
  FFHIP_CHECK(hipStreamSynchronize(ctx->stream));
  return v;
}

// ---------------------------------------------------------------- cube
// hex split selected by BuildCube for kind = 6 (hex_decoupe[70] of the table
// at msh3.cpp:7682-7742, simplexes {7,15,19,27,26,11}), hex-local numbering
// a + 2b + 4c (msh3.cpp:8039-8045)
__constant__ int c_split6[6][4] = {{4, 0, 6, 7}, {0, 4, 5, 7}, {1, 0, 5, 7},
                                   {0, 1, 3, 7}, {2, 0, 3, 7}, {0, 2, 6, 7}};
// Tet faces (Mesh3dn.cpp:72)
__constant__ int c_tetface[4][3] = {{3, 2, 1}, {0, 2, 3}, {3, 1, 0}, {0, 1, 2}};

__global__ void k_cube_vertices(int nx, int ny, int nz, double xd, double yd,
                                double zd, double *__restrict__ xyz) {
  const int64_t nv = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nv) return;
  const int i = (int)(p % (nx + 1));
  const int j = (int)((p / (nx + 1)) % (ny + 1));
  const int k = (int)(p / ((int64_t)(nx + 1) * (ny + 1)));
  xyz[3 * p + 0] = 0. + xd * i;
  xyz[3 * p + 1] = 0. + yd * j;
  xyz[3 * p + 2] = 0. + zd * k;
}

__device__ inline int cube_vlab(int64_t v, int nx, int ny, int nz) {
  const int64_t nj = nx + 1, nk = nj * (ny + 1);
  const int vi = (int)(v % nj), vj = (int)((v / nj) % (ny + 1)), vk = (int)(v / nk);
  return 1 * (vi == 0) + 2 * (vi == nx) + 4 * (vj == 0) + 8 * (vj == ny) +
         16 * (vk == 0) + 32 * (vk == nz);
}

// one thread per hex: its six tets, and its boundary triangle count
__global__ void k_cube_tets(int nx, int ny, int nz, int32_t *__restrict__ tets,
                            int64_t *__restrict__ bcount) {
  const int64_t nc = (int64_t)nx * ny * nz;
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nc) return;
  const int i = (int)(h % nx), j = (int)((h / nx) % ny), k = (int)(h / ((int64_t)nx * ny));
  const int64_t nj = nx + 1, nk = nj * (ny + 1);
  int32_t n8[8];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a)
        n8[a + 2 * b + 4 * c] = (int32_t)((i + a) + nj * (j + b) + nk * (k + c));
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    int4 t;
    t.x = n8[c_split6[d][0]]; t.y = n8[c_split6[d][1]];
    t.z = n8[c_split6[d][2]]; t.w = n8[c_split6[d][3]];
    reinterpret_cast<int4 *>(tets)[6 * h + d] = t;
  }
  // each hex face on the boundary carries two boundary triangles
  bcount[h] = 2 * ((i == 0) + (i == nx - 1) + (j == 0) + (j == ny - 1) +
                   (k == 0) + (k == nz - 1));
}

// boundary triangles in BuildCube's order: hex, tet, face (msh3.cpp:8090-8110)
__global__ void k_cube_bfaces(int nx, int ny, int nz, const int32_t *__restrict__ tets,
                              const int64_t *__restrict__ boff, const int32_t *lab6,
                              int32_t *__restrict__ btri, int32_t *__restrict__ blab) {
  const int64_t nc = (int64_t)nx * ny * nz;
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nc) return;
  const int i = (int)(h % nx), j = (int)((h / nx) % ny), k = (int)(h / ((int64_t)nx * ny));
  if (!(i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1)) return;
  const int nff[6] = {3, 1, 0, 2, 4, 5};  // msh3.cpp:8050
  int64_t o = boff[h];
  const int64_t oend = boff[h + 1];  // writes stay inside this hex's scanned range
  for (int d = 0; d < 6; ++d) {
    const int32_t *nu = tets + 4 * (6 * h + d);
    for (int f = 0; f < 4; ++f) {
      const int32_t a = nu[c_tetface[f][0]], b = nu[c_tetface[f][1]], c = nu[c_tetface[f][2]];
      const int l = cube_vlab(a, nx, ny, nz) & cube_vlab(b, nx, ny, nz) & cube_vlab(c, nx, ny, nz);
      for (int bit = 0; bit < 6; ++bit)
        if (l == (1 << bit) && o < oend) {
          btri[3 * o + 0] = a; btri[3 * o + 1] = b; btri[3 * o + 2] = c;
          blab[o] = lab6[nff[bit]];
          ++o;
        }
    }
  }
}

// ---------------------------------------------------------------- square
__global__ void k_square(int nx, int ny, double *__restrict__ xy, int32_t *__restrict__ tris) {
  const int64_t nv = (int64_t)(nx + 1) * (ny + 1), nc = (int64_t)nx * ny;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nv) {
    const int i = (int)(p % (nx + 1)), j = (int)(p / (nx + 1));
    xy[2 * p] = (double)i / nx;  // R2 P((R) i/nx,(R) j/ny)  (lgmesh.cpp:1258)
    xy[2 * p + 1] = (double)j / ny;
  }
  if (p < nc) {
    const int i = (int)(p % nx), j = (int)(p / nx);
    const int32_t i0 = i + j * (nx + 1), i1 = i0 + 1, i2 = i1 + nx + 1, i3 = i2 - 1;
    int32_t *t = tris + 6 * p;  // diag 1 (lgmesh.cpp:1300-1306)
    t[0] = i0; t[1] = i1; t[2] = i2; t[3] = i0; t[4] = i2; t[5] = i3;
  }
}

__global__ void k_square_bedges(int nx, int ny, const int32_t *lab4,
                                int32_t *__restrict__ bed, int32_t *__restrict__ blab) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbe = 2 * ((int64_t)nx + ny);
  if (b >= nbe) return;
  const int nx1 = nx + 1;
  int32_t i1, i2, l;
  if (b < nx) { i1 = (int32_t)b; i2 = i1 + 1; l = lab4[0]; }                       // bottom
  else if (b < nx + ny) { int j = (int)(b - nx); i1 = nx + j * nx1; i2 = i1 + nx1; l = lab4[1]; }  // right
  else if (b < 2 * nx + ny) { int i = (int)(b - nx - ny); i1 = i + ny * nx1; i2 = i1 + 1; l = lab4[2]; }  // up
  else { int j = (int)(b - 2 * nx - ny); i1 = j * nx1; i2 = i1 + nx1; l = lab4[3]; }  // left
  bed[2 * b] = i1; bed[2 * b + 1] = i2; blab[b] = l;
}

// plain streaming copy, 16 B per lane per step (measurement reference)
typedef int v4i __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(int64_t n16, const v4i *__restrict__ src,
                                              v4i *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = __builtin_nontemporal_load(src + i);
}

}  // namespace ffhip

using namespace ffhip;

void *ffhip_ctx_s::ensure_scratch(size_t bytes) {
  if (bytes > scratch_bytes) {
    if (scratch) (void)hipFree(scratch);
    scratch = nullptr;
    scratch_bytes = 0;
    FFHIP_CHECK(hipMalloc(&scratch, bytes));
    scratch_bytes = bytes;
  }
  return scratch;
}

void *ffhip_ctx_s::ensure_work(size_t bytes) {
  if (bytes > work_bytes) {
    if (work) (void)hipFree(work);
    work = nullptr;
    work_bytes = 0;
    FFHIP_CHECK(hipMalloc(&work, bytes));
    work_bytes = bytes;
  }
  return work;
}

void ffhip_ctx_s::tic() { FFHIP_CHECK(hipEventRecord(ev0, stream)); }
void ffhip_ctx_s::toc() {
  FFHIP_CHECK(hipEventRecord(ev1, stream));
  FFHIP_CHECK(hipEventSynchronize(ev1));
  float ms = 0.f;
  FFHIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  last_ms = ms;
}

extern "C" {

const char *ffhip_last_error(void) { return g_err; }
const char *ffhip_version(void) { return "ffhip 0.1 (gfx950)"; }

}  // extern "C"

namespace ffhip {
namespace {
struct BlockCache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void *>> free_;       // per device, by size
  std::unordered_map<void *, std::pair<int, size_t>> info;  // every block: device, size
  std::map<int, int> contexts;
  size_t cached = 0;  // bytes in the free lists
};
BlockCache &blocks() {
  static BlockCache *c = new BlockCache();  // never destroyed: used until process exit
  return *c;
}
constexpr size_t kCacheCap = size_t(32) << 30;  // cached bytes kept at most

size_t round_block(size_t b) {
  const size_t g = b >= (size_t(1) << 20) ? (size_t(2) << 20) : 256;
  return (b + g - 1) / g * g;
}

void trim_device(BlockCache &c, int dev) {  // caller holds c.mu
  auto it = c.free_.find(dev);
  if (it == c.free_.end()) return;
  for (auto &kv : it->second) {
    (void)hipFree(kv.second);
    c.info.erase(kv.second);
    c.cached -= kv.first;
  }
  it->second.clear();
}
}  // namespace

void *block_alloc(size_t bytes) {
  BlockCache &c = blocks();
  int dev = 0;
  FFHIP_CHECK(hipGetDevice(&dev));
  const size_t want = round_block(bytes);
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto &fl = c.free_[dev];
    auto it = fl.lower_bound(want);
    // a cached block at most 1/8 larger (small blocks: at most twice as large)
    if (it != fl.end() && it->first <= (want >= (size_t(1) << 20) ? want + want / 8 : 2 * want)) {
      void *p = it->second;
      c.cached -= it->first;
      fl.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {  // give the cached blocks back and try once more
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> g(c.mu);
      trim_device(c, dev);
    }
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu bytes) -> %s", want, hipGetErrorString(e));
    throw Error{FFHIP_ERR_HIP};
  }
  std::lock_guard<std::mutex> g(c.mu);
  c.info[p] = {dev, want};
  return p;
}

void block_free(void *p) {
  BlockCache &c = blocks();
  int dev = 0;
  size_t size = 0;
  bool shared = false;
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.info.find(p);
    if (it == c.info.end()) {  // not a cached block
      (void)hipFree(p);
      return;
    }
    dev = it->second.first;
    size = it->second.second;
    shared = c.contexts[dev] > 1;
    if (c.cached + size > kCacheCap) {
      c.info.erase(it);
      (void)hipFree(p);
      return;
    }
  }
  if (shared) {  // another context's stream could reuse it: let this one's work finish
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    (void)hipSetDevice(cur);
  }
  std::lock_guard<std::mutex> g(c.mu);
  c.free_[dev].emplace(size, p);
  c.cached += size;
}

void blocks_context(int device, int delta) {
  BlockCache &c = blocks();
  std::lock_guard<std::mutex> g(c.mu);
  c.contexts[device] += delta;
  if (c.contexts[device] <= 0) trim_device(c, device);
}
}  // namespace ffhip

extern "C" {

int ffhip_init(int device, ffhip_ctx *out) {
  return guard([&] {
    FFHIP_REQUIRE(out, FFHIP_ERR_ARG, "ffhip_init: null output");
    int ndev = 0;
    FFHIP_CHECK(hipGetDeviceCount(&ndev));
    FFHIP_REQUIRE(device >= 0 && device < ndev, FFHIP_ERR_ARG,
                  "ffhip_init: device %d of %d", device, ndev);
    FFHIP_CHECK(hipSetDevice(device));
    noalloc_scratch_setup(device);
    ffhip_ctx c = new ffhip_ctx_s();
    c->device = device;
    // a blocking stream: ordered with the legacy default stream, which is
    // where a host framework sharing the process (torch) enqueues its work
    FFHIP_CHECK(hipStreamCreate(&c->stream));
    FFHIP_CHECK(hipEventCreate(&c->ev0));
    FFHIP_CHECK(hipEventCreate(&c->ev1));
    hipDeviceProp_t prop;
    FFHIP_CHECK(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    blocks_context(device, +1);
    *out = c;
  });
}

int ffhip_ctx_free(ffhip_ctx c) {
  return guard([&] {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->work) (void)hipFree(c->work);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    blocks_context(c->device, -1);
    delete c;
  });
}

int ffhip_sync(ffhip_ctx c) {
  return guard([&] { FFHIP_CHECK(hipStreamSynchronize(c->stream)); });
}

double ffhip_last_ms(ffhip_ctx c) { return c ? c->last_ms : 0.; }

int ffhip_bench_copy(ffhip_ctx ctx, int64_t nbytes, int reps, double *avg_ms) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && avg_ms && reps > 0 && nbytes >= 16, FFHIP_ERR_ARG, "ffhip_bench_copy: bad arguments");
    const int64_t n16 = nbytes / 16;
    v4i *a = dalloc<v4i>(n16), *b = dalloc<v4i>(n16);
    FFHIP_CHECK(hipMemsetAsync(a, 0, 16 * n16, ctx->stream));
    // 4 blocks of 256 per CU: the fastest plain copy in a sweep of grid sizes
    // (4..32 blocks per CU) and unroll depths on MI355X
    const unsigned grid = 4 * (unsigned)ctx->num_cu;
    k_copy<<<grid, 256, 0, ctx->stream>>>(n16, a, b);
    FFHIP_CHECK(hipGetLastError());
    ctx->tic();
    for (int r = 0; r < reps; ++r) k_copy<<<grid, 256, 0, ctx->stream>>>(n16, a, b);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    *avg_ms = ctx->last_ms / reps;
    dfree(a);
    dfree(b);
  });
}

int ffhip_mesh3_cube(ffhip_ctx ctx, int nx, int ny, int nz, const int32_t labels[6],
                     ffhip_mesh *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && out, FFHIP_ERR_ARG, "ffhip_mesh3_cube: null argument");
    FFHIP_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, FFHIP_ERR_ARG,
                  "ffhip_mesh3_cube: bad sizes %d %d %d", nx, ny, nz);
    const int64_t nv = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t nc = (int64_t)nx * ny * nz, nt = 6 * nc;
    const int64_t nbe = 4 * ((int64_t)nx * ny + (int64_t)nx * nz + (int64_t)ny * nz);
    FFHIP_REQUIRE(nv < (1ll << 31) && 4 * nt < (1ll << 31), FFHIP_ERR_UNSUPPORTED,
                  "ffhip_mesh3_cube: mesh too large for int32 element ids");
    int32_t lab[6] = {1, 2, 3, 4, 5, 6};
    if (labels) memcpy(lab, labels, sizeof(lab));
    ffhip_mesh m = new ffhip_mesh_s();
    m->ctx = ctx; m->dim = 3; m->nloc = 4; m->nbloc = 3;
    m->nv = nv; m->nt = nt; m->nbe = nbe;
    m->xyz = dalloc<double>(3 * nv);
    m->elems = dalloc<int32_t>(4 * nt);
    m->belems = dalloc<int32_t>(3 * nbe);
    m->blab = dalloc<int32_t>(nbe);
    int64_t *cnt = dalloc<int64_t>(nc + 1);
    int64_t *off = dalloc<int64_t>(nc + 1);
    int32_t *dlab = dalloc<int32_t>(6);
    FFHIP_CHECK(hipMemcpyAsync(dlab, lab, sizeof(lab), hipMemcpyHostToDevice, ctx->stream));
    FFHIP_CHECK(hipMemsetAsync(cnt + nc, 0, sizeof(int64_t), ctx->stream));
    ctx->tic();
    const int B = 256;
    k_cube_vertices<<<(unsigned)((nv + B - 1) / B), B, 0, ctx->stream>>>(
        nx, ny, nz, 1. / nx, 1. / ny, 1. / nz, m->xyz);
    k_cube_tets<<<(unsigned)((nc + B - 1) / B), B, 0, ctx->stream>>>(nx, ny, nz, m->elems, cnt);
    exclusive_scan_i64(ctx, cnt, off, nc + 1);
    k_cube_bfaces<<<(unsigned)((nc + B - 1) / B), B, 0, ctx->stream>>>(
        nx, ny, nz, m->elems, off, dlab, m->belems, m->blab);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    const int64_t got = read_i64(ctx, off + nc);
    dfree(cnt); dfree(off); dfree(dlab);
    FFHIP_REQUIRE(got == nbe, FFHIP_ERR_INTERNAL, "cube: %lld boundary faces, expected %lld",
                  (long long)got, (long long)nbe);
    *out = m;
  });
}

int ffhip_mesh2_square(ffhip_ctx ctx, int nx, int ny, const int32_t labels[4],
                       ffhip_mesh *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && out && nx >= 1 && ny >= 1, FFHIP_ERR_ARG, "ffhip_mesh2_square: bad args");
    const int64_t nv = (int64_t)(nx + 1) * (ny + 1), nt = 2 * (int64_t)nx * ny;
    const int64_t nbe = 2 * ((int64_t)nx + ny);
    FFHIP_REQUIRE(nv < (1ll << 31) && 3 * nt < (1ll << 31), FFHIP_ERR_UNSUPPORTED,
                  "ffhip_mesh2_square: mesh too large");
    int32_t lab[4] = {1, 2, 3, 4};
    if (labels) memcpy(lab, labels, sizeof(lab));
    ffhip_mesh m = new ffhip_mesh_s();
    m->ctx = ctx; m->dim = 2; m->nloc = 3; m->nbloc = 2;
    m->nv = nv; m->nt = nt; m->nbe = nbe;
    m->xyz = dalloc<double>(2 * nv);
    m->elems = dalloc<int32_t>(3 * nt);
    m->belems = dalloc<int32_t>(2 * nbe);
    m->blab = dalloc<int32_t>(nbe);
    int32_t *dlab = dalloc<int32_t>(4);
    FFHIP_CHECK(hipMemcpyAsync(dlab, lab, sizeof(lab), hipMemcpyHostToDevice, ctx->stream));
    const int B = 256;
    const int64_t nn = nv > nt / 2 ? nv : nt / 2;
    ctx->tic();
    k_square<<<(unsigned)((nn + B - 1) / B), B, 0, ctx->stream>>>(nx, ny, m->xyz, m->elems);
    k_square_bedges<<<(unsigned)((nbe + B - 1) / B), B, 0, ctx->stream>>>(nx, ny, dlab, m->belems, m->blab);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    dfree(dlab);
    *out = m;
  });
}

int ffhip_mesh3_upload(ffhip_ctx ctx, int64_t nv, int64_t nt, int64_t nbe,
                       const double *xyz, const int32_t *tets, const int32_t *btri,
                       const int32_t *blab, ffhip_mesh *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && out && xyz && tets && nv > 0 && nt > 0 && nbe >= 0, FFHIP_ERR_ARG,
                  "ffhip_mesh3_upload: bad arguments");
    FFHIP_REQUIRE(nbe == 0 || (btri && blab), FFHIP_ERR_ARG, "ffhip_mesh3_upload: boundary arrays");
    FFHIP_REQUIRE(nv < (1ll << 31) && 4 * nt < (1ll << 31), FFHIP_ERR_UNSUPPORTED,
                  "ffhip_mesh3_upload: mesh too large for int32 ids");
    for (int64_t k = 0; k < 4 * nt; ++k)
      FFHIP_REQUIRE(tets[k] >= 0 && tets[k] < nv, FFHIP_ERR_ARG,
                    "ffhip_mesh3_upload: tet vertex id %d out of range", tets[k]);
    for (int64_t k = 0; k < 3 * nbe; ++k)
      FFHIP_REQUIRE(btri[k] >= 0 && btri[k] < nv, FFHIP_ERR_ARG,
                    "ffhip_mesh3_upload: boundary vertex id out of range");
    ffhip_mesh m = new ffhip_mesh_s();
    m->ctx = ctx; m->dim = 3; m->nloc = 4; m->nbloc = 3;
    m->nv = nv; m->nt = nt; m->nbe = nbe;
    m->xyz = dalloc<double>(3 * nv);
    m->elems = dalloc<int32_t>(4 * nt);
    m->belems = dalloc<int32_t>(3 * nbe);
    m->blab = dalloc<int32_t>(nbe);
    FFHIP_CHECK(hipMemcpyAsync(m->xyz, xyz, 24 * nv, hipMemcpyHostToDevice, ctx->stream));
    FFHIP_CHECK(hipMemcpyAsync(m->elems, tets, 16 * nt, hipMemcpyHostToDevice, ctx->stream));
    if (nbe) {
      FFHIP_CHECK(hipMemcpyAsync(m->belems, btri, 12 * nbe, hipMemcpyHostToDevice, ctx->stream));
      FFHIP_CHECK(hipMemcpyAsync(m->blab, blab, 4 * nbe, hipMemcpyHostToDevice, ctx->stream));
    }
    FFHIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = m;
  });
}

int ffhip_mesh2_upload(ffhip_ctx ctx, int64_t nv, int64_t nt, int64_t nbe, const double *xy,
                       const int32_t *tris, const int32_t *bedges, const int32_t *blab,
                       ffhip_mesh *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && out && xy && tris && nv > 0 && nt > 0 && nbe >= 0, FFHIP_ERR_ARG,
                  "ffhip_mesh2_upload: bad arguments");
    FFHIP_REQUIRE(nbe == 0 || (bedges && blab), FFHIP_ERR_ARG, "ffhip_mesh2_upload: boundary arrays");
    FFHIP_REQUIRE(nv < (1ll << 31) && 3 * nt < (1ll << 31), FFHIP_ERR_UNSUPPORTED,
                  "ffhip_mesh2_upload: mesh too large for int32 ids");
    for (int64_t k = 0; k < 3 * nt; ++k)
      FFHIP_REQUIRE(tris[k] >= 0 && tris[k] < nv, FFHIP_ERR_ARG,
                    "ffhip_mesh2_upload: triangle vertex id %d out of range", tris[k]);
    for (int64_t k = 0; k < 2 * nbe; ++k)
      FFHIP_REQUIRE(bedges[k] >= 0 && bedges[k] < nv, FFHIP_ERR_ARG,
                    "ffhip_mesh2_upload: boundary vertex id out of range");
    ffhip_mesh m = new ffhip_mesh_s();
    m->ctx = ctx; m->dim = 2; m->nloc = 3; m->nbloc = 2;
    m->nv = nv; m->nt = nt; m->nbe = nbe;
    m->xyz = dalloc<double>(2 * nv);
    m->elems = dalloc<int32_t>(3 * nt);
    m->belems = dalloc<int32_t>(2 * nbe);
    m->blab = dalloc<int32_t>(nbe);
    FFHIP_CHECK(hipMemcpyAsync(m->xyz, xy, 16 * nv, hipMemcpyHostToDevice, ctx->stream));
    FFHIP_CHECK(hipMemcpyAsync(m->elems, tris, 12 * nt, hipMemcpyHostToDevice, ctx->stream));
    if (nbe) {
      FFHIP_CHECK(hipMemcpyAsync(m->belems, bedges, 8 * nbe, hipMemcpyHostToDevice, ctx->stream));
      FFHIP_CHECK(hipMemcpyAsync(m->blab, blab, 4 * nbe, hipMemcpyHostToDevice, ctx->stream));
    }
    FFHIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = m;
  });
}

int ffhip_mesh_info(ffhip_mesh m, int64_t *nv, int64_t *nt, int64_t *nbe) {
  return guard([&] {
    FFHIP_REQUIRE(m, FFHIP_ERR_ARG, "ffhip_mesh_info: null mesh");
    if (nv) *nv = m->nv;
    if (nt) *nt = m->nt;
    if (nbe) *nbe = m->nbe;
  });
}

int ffhip_mesh_download(ffhip_mesh m, double *xyz, int32_t *elems, int32_t *belems,
                        int32_t *blab) {
  return guard([&] {
    FFHIP_REQUIRE(m, FFHIP_ERR_ARG, "ffhip_mesh_download: null mesh");
    hipStream_t s = m->ctx->stream;
    if (xyz) FFHIP_CHECK(hipMemcpyAsync(xyz, m->xyz, 8 * m->dim * m->nv, hipMemcpyDeviceToHost, s));
    if (elems) FFHIP_CHECK(hipMemcpyAsync(elems, m->elems, 4 * m->nloc * m->nt, hipMemcpyDeviceToHost, s));
    if (belems && m->nbe)
      FFHIP_CHECK(hipMemcpyAsync(belems, m->belems, 4 * m->nbloc * m->nbe, hipMemcpyDeviceToHost, s));
    if (blab && m->nbe) FFHIP_CHECK(hipMemcpyAsync(blab, m->blab, 4 * m->nbe, hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
  });
}

int ffhip_mesh_free(ffhip_mesh m) {
  return guard([&] {
    if (!m) return;
    (void)hipStreamSynchronize(m->ctx->stream);
    dfree(m->xyz); dfree(m->elems); dfree(m->belems); dfree(m->blab);
    dfree(m->v2e_ptr); dfree(m->v2e); dfree(m->v2e_inv);
    dfree(m->b2e_ptr); dfree(m->b2e);
    delete m;
  });
}

}  // extern "C"
