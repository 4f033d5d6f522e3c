// ffhip_internal.hpp — shared internals of libffhip (not part of the C-ABI).
//
// Device-memory objects behind the opaque handles of include/ffhip.h, error
// plumbing, and small device helpers.  Everything here targets gfx950
// (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ffhip.h"

namespace ffhip {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

struct Error {
  int code;
};

#define FFHIP_CHECK(expr)                                                      \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::ffhip::set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,         \
                         hipGetErrorString(e_));                               \
      throw ::ffhip::Error{FFHIP_ERR_HIP};                                     \
    }                                                                          \
  } while (0)

#define FFHIP_REQUIRE(cond, code, ...)                                         \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::ffhip::set_error(__VA_ARGS__);                                         \
      throw ::ffhip::Error{code};                                              \
    }                                                                          \
  } while (0)

// Runs f() and maps exceptions to C-ABI return codes.
template <class F>
int guard(F &&f) {
  try {
    f();
    return FFHIP_OK;
  } catch (const Error &e) {
    return e.code;
  } catch (...) {
    set_error("unexpected C++ exception");
    return FFHIP_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------- memory
// Device blocks: released blocks are kept per device and handed out again for
// requests of about the same size (a solve or an assembly repeated on the same
// mesh re-allocates the same sizes), so steady-state calls do not pay
// hipMalloc / hipFree (tens to hundreds of microseconds per large block).
// All work of a context is ordered on its one stream; with several contexts
// on a device, a release first waits for the device.
void *block_alloc(size_t bytes);
void block_free(void *p);
void blocks_context(int device, int delta);  // context count per device; trims at 0
template <class T>
T *dalloc(size_t n) {
  if (n == 0) n = 1;
  return (T *)block_alloc(n * sizeof(T));
}
template <class T>
void dfree(T *&p) {
  if (p) block_free(p);
  p = nullptr;
}

}  // namespace ffhip

// ---------------------------------------------------------------- handles
struct ffhip_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;  // default priority (never a priority stream)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cu = 256;
  // reusable scratch for library-internal temporaries (scans, partials)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *ensure_scratch(size_t bytes);
  // work buffer the assembly and solver-setup temporaries are carved from
  // (grows to the largest request and is kept: no allocation on repeat calls)
  void *work = nullptr;
  size_t work_bytes = 0;
  void *ensure_work(size_t bytes);
  double last_ms = 0.;  // device time of the last timed call
  void tic();           // record ev0 on the stream
  void toc();           // record ev1, synchronise, set last_ms
};

struct ffhip_mesh_s {
  ffhip_ctx ctx = nullptr;
  int dim = 3;      // 3: tetrahedra, 2: triangles
  int nloc = 4;     // vertices per element
  int nbloc = 3;    // vertices per boundary element
  int64_t nv = 0, nt = 0, nbe = 0;
  double *xyz = nullptr;    // nv * dim, vertex-major (x,y,z per vertex)
  int32_t *elems = nullptr; // nt * nloc
  int32_t *belems = nullptr;// nbe * nbloc
  int32_t *blab = nullptr;  // nbe
  // dof -> incident (element, local vertex) list, element order (built lazily)
  int64_t *v2e_ptr = nullptr;  // nv + 1
  int32_t *v2e = nullptr;      // nt * nloc entries: e * nloc + l
  int v2e_big = 0;             // some vertex has more than 32 incident elements
  int32_t *v2e_inv = nullptr;  // nt * nloc: position of item e * nloc + l in v2e
  // dof -> incident (boundary element, local vertex) list, boundary-element
  // order (built lazily by the boundary-integral calls)
  int64_t *b2e_ptr = nullptr;  // nv + 1
  int32_t *b2e = nullptr;      // nbe * nbloc entries: f * nbloc + l
};

struct ffhip_mat_s {
  ffhip_ctx ctx = nullptr;
  int64_t n = 0, nnz = 0;
  int64_t *rowptr = nullptr;   // n + 1
  int32_t *colidx = nullptr;   // nnz (sorted within each row)
  double *vals = nullptr;      // nnz
  int64_t *diag = nullptr;     // n: slot of (i,i), -1 if absent
  // element-to-row map: for incidence entry k of row i (mesh v2e order), the
  // offsets within row i of the columns of that element's vertices
  uint32_t *slotmap = nullptr; // (nt*nloc) packed 4 x uint8 (P1, nloc <= 4)
  int max_row = 0;
  ffhip_mesh mesh = nullptr;   // mesh the pattern was built from (if any)
  // P1 rows built by the lane-per-row kernels (more than 32 incident elements
  // or 64 columns); the others by the half-wave kernels
  uint8_t *big = nullptr;      // n
  int any_big = 0;
  // vector-P2 matrices: node-level rows (the scalar rows 3a..3a+2 of node a
  // hold 3 x 3 blocks of its node columns) and the byte slot map, 10 offsets
  // per node incidence
  ffhip_p2v p2v = nullptr;
  int ncomp = 0;               // P2 matrices: 3 for [P2,P2,P2], 1 for scalar P2
  int64_t *nodeptr = nullptr;  // nnodes + 1
  uint8_t *slot8 = nullptr;    // (nt * 10) * 10
  // half storage (FreeFEM's sym=1): only row >= column is stored; the full
  // view lists every full row as slots of vals (built once, by half_view)
  int half = 0;
  int64_t fnnz = 0;
  int64_t *fptr = nullptr;     // n + 1
  int32_t *fcol = nullptr;     // fnnz
  int64_t *fslot = nullptr;    // fnnz: slot in vals
};

// [P2,P2,P2] space on a tetrahedral mesh (FESpace3 with GTypeOfFESum of three
// P2): nodes numbered by first appearance over the elements, then dof
// 3 * node + component.
struct ffhip_p2v_s {
  ffhip_ctx ctx = nullptr;
  ffhip_mesh mesh = nullptr;
  int64_t nnodes = 0, ne = 0;
  int32_t *elem_nodes = nullptr;  // nt * 10: 4 vertex nodes, 6 edge nodes
  int32_t *vnode = nullptr;       // nv: node of each vertex
  uint64_t *ekeys = nullptr;      // ne: sorted edge keys (min << 32 | max)
  int32_t *enode = nullptr;       // ne: node of each edge
  int64_t *n2e_ptr = nullptr;     // nnodes + 1
  int32_t *n2e = nullptr;         // nt * 10: element * 10 + local node, element order
  int32_t *n2e_inv = nullptr;     // nt * 10: position of item element * 10 + local in n2e
};

namespace ffhip {

// ---------------------------------------------------------------- device helpers
constexpr int kWave = 64;

// wave64 sum of a double (DPP/permute based shuffles)
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block reduction of one double; result valid in every thread.
template <int BLOCK>
__device__ inline double block_sum(double v, double *lds /* BLOCK/64 */) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.;
#pragma unroll
  for (int i = 0; i < BLOCK / kWave; ++i) s += lds[i];
  return s;
}

// orders the LDS accesses of one wave (cross-lane reuse of the row staging)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware logical block id: the hardware dispatches block b to XCD b % 8;
// logical ids are renumbered so each XCD owns a contiguous block range (and
// so a contiguous row range whose x-gathers stay in that XCD's L2).
__device__ inline int xcd_block_id() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % 8) return b;
  return (b % 8) * (G / 8) + (b / 8);
}

// What the one-call P1 assembly needs for the element scales (C_e of the
// stiffness); the incidence build of a tetrahedral mesh computes them when
// given these.
struct ScaleArgs {
  const double *xyz;
  int kind;
  double c0;
  const double *kappa_e;
  double *Ce;
};

// dof -> (element, local dof) incidence lists in element order (ffhip_assembly.hip)
void build_incidence(ffhip_ctx ctx, const int32_t *elems, int64_t nt, int nloc, int64_t n,
                     int64_t **ptr_out, int32_t **list_out, int *big_out, int32_t **inv_out,
                     int32_t **bucket_out, const ScaleArgs *sc = nullptr, int *ovf_dev = nullptr);

void exclusive_scan_i64(ffhip_ctx ctx, const int64_t *in, int64_t *out,
                        int64_t n);

// 256-byte aligned temporaries carved from the context's work buffer; one
// arena at a time per context (the calls that use one are sequential)
struct Arena {
  static size_t bytes(size_t count, size_t size) { return (count * size + 255) & ~size_t(255); }
  char *base;
  size_t cap, off = 0;
  Arena(ffhip_ctx ctx, size_t total) : base((char *)ctx->ensure_work(total)), cap(total) {}
  template <class T>
  T *take(size_t count) {
    const size_t b = bytes(count ? count : 1, sizeof(T));
    FFHIP_REQUIRE(off + b <= cap, FFHIP_ERR_INTERNAL, "work arena overflow");
    T *p = (T *)(base + off);
    off += b;
    return p;
  }
};
int64_t read_i64(ffhip_ctx ctx, const int64_t *p);

// the per-device address of the no-allocation entry points' scratch array
// (ffhip_assembly.hip); set up by ffhip_init with the device current
void noalloc_scratch_setup(int device);

}  // namespace ffhip
