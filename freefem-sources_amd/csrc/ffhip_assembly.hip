// ffhip_assembly.hip — P1 sparsity pattern, atomic-free assembly, Dirichlet
// rows, right-hand side and CSR mat-vec.
//
// Design (MI355X-first, deterministic):
//   1. dof -> (element, local vertex) incidence lists, sorted by element id:
//      each block counts its items per dof in an LDS hash set and reserves
//      their places in the dofs' 32-entry buckets with one global atomic per
//      dof, a scan gives the list offsets, and every entry goes to its rank
//      among its dof's entries (k_v2e_sort_bucket, or the fused pattern pass
//      of the one-call assembly).  Lists longer than 32: an atomic fill, a
//      bitonic sort of the first 32 in a half-wave, insertion sort beyond.
//   2. pattern: one half-wave per row loads the vertices of the row's
//      incident elements (4 per lane), deduplicates them in an LDS hash set
//      and orders the unique columns by a counting sort.  One pass writes
//      each row's sorted columns to a row buffer, its length and, for every
//      incidence entry, the offsets of that element's vertices inside the row
//      (the element-to-row map); after a scan of the lengths the buffers are
//      compacted into the CSR arrays.  Rows with more than 32 incident
//      elements use a lane-per-row merge in LDS.
//   3. values: one thread per element evaluates the 14-point Tet_5 integral
//      of kappa and the P1 gradients (VALU FP64) and writes the 4 rows of its
//      symmetric 4x4 matrix where its incidence entries sit in the vertices'
//      lists; then one half-wave per row loads them contiguously (one lane
//      per element) into LDS, and the lane owning each slot adds them in
//      element order.  No atomics: every slot is summed by one lane in
//      element order, the order in which HashMatrix::operator+= accumulates
//      (src/femlib/HashMatrix.cpp:1294-1332).  2D triangles and the rows with
//      more than 32 elements use a lane-per-row walk with the row staged in
//      LDS.
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <mutex>
#include <vector>

#include "ffhip_fem.hpp"

namespace ffhip {

// T_5 (src/femlib/QuadratureFormular.cpp:172-186), filled from the host
__constant__ double c_qt5_pts[7][2];
__constant__ double c_qt5_w[7];

// ---------------------------------------------------------------- incidence
__global__ void k_v2e_count(const int32_t *__restrict__ elems, int64_t m,
                            uint32_t *__restrict__ deg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) atomicAdd(&deg[elems[i]], 1u);
}

// one pass: count and place every incidence into a 32-entry bucket of its
// dof (lists longer than 32 set *ovf and are rebuilt by count + fill)
__global__ void k_v2e_bucket(const int32_t *__restrict__ elems, int64_t m,
                             uint32_t *__restrict__ deg, int32_t *__restrict__ bucket,
                             int *__restrict__ ovf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t v = elems[i];
  const uint32_t pos = atomicAdd(&deg[v], 1u);
  if (pos < 32) bucket[32 * (int64_t)v + pos] = (int32_t)i;
  else atomicOr(ovf, 1);
}

__global__ void k_u32_to_i64(int64_t n, const uint32_t *__restrict__ a, int64_t *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_v2e_fill(const int32_t *__restrict__ elems, int64_t m,
                           const int64_t *__restrict__ ptr, int32_t *__restrict__ cur,
                           int32_t *__restrict__ v2e) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t v = elems[i];
  const int32_t pos = atomicAdd(&cur[v], 1);
  v2e[ptr[v] + pos] = (int32_t)i;
}

// ---------------------------------------------------------------- half-wave sorting networks
// A 32-lane half of a wave holds one row's list; bitonic compare-exchange
// across lanes (lane xor m, m <= 16, never leaves the half) and inside a lane.
constexpr int kHalf = 32;  // rows of up to 32 incident elements use the half-wave kernels

template <class T>
__device__ inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T>
__device__ inline T tmax(T a, T b) { return a < b ? b : a; }

// one value per lane, sequence index = lane within the half
template <class T>
__device__ inline T bitonic32(T x, int hl) {
#pragma unroll
  for (int k = 2; k <= 32; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const T p = __shfl_xor(x, j, kWave);
      const bool up = (hl & k) == 0, lower = (hl & j) == 0;
      x = (lower == up) ? tmin(x, p) : tmax(x, p);
    }
  return x;
}

// incidence lists of up to 32 entries sorted into element order, two rows
// per wave; longer lists are left to k_v2e_sort (flagged in *big)
__global__ __launch_bounds__(256) void k_v2e_sort_half(int64_t n, const int64_t *__restrict__ ptr,
                                                       int32_t *__restrict__ v2e, int *__restrict__ big) {
  const int hl = threadIdx.x % kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  int64_t k0 = 0;
  int deg = 0;
  if (v < n) { k0 = ptr[v]; deg = (int)(ptr[v + 1] - k0); }
  if (deg > kHalf) {
    if (hl == 0) atomicOr(big, 1);
    deg = 0;
  }
  int32_t x = hl < deg ? v2e[k0 + hl] : INT_MAX;
  x = bitonic32(x, hl);
  if (hl < deg) v2e[k0 + hl] = x;
}

// Incidence buckets (up to 32 entries per dof, in no particular order; the
// sort below puts them in element order) with the atomics aggregated per
// block: the block's 256 x kBucketIpt items are counted per dof in an LDS hash set (an
// item's rank among the block's items of its dof from the LDS atomicAdd),
// then each distinct dof of the block reserves its positions with one
// atomicAdd on its global count.
#ifndef FFHIP_BUCKET_IPT
#define FFHIP_BUCKET_IPT 2
#endif
// items per thread (cube(160): 1 -> 2 saves about 0.15 ms of the one-call
// assembly, 4 loses it again)
constexpr int kBucketIpt = FFHIP_BUCKET_IPT;
constexpr int kBucketLog = 9 + (kBucketIpt >= 4 ? 2 : kBucketIpt >= 2 ? 1 : 0);
constexpr int kBucketHash = 1 << kBucketLog;              // >= 2x the items of a block: probing always ends
static_assert(kBucketHash >= 2 * 256 * kBucketIpt, "hash set of twice the block's items");
__global__ __launch_bounds__(256) void k_v2e_bucket_block(const int32_t *__restrict__ elems, int64_t m,
                                                          uint32_t *__restrict__ deg,
                                                          int32_t *__restrict__ bucket,
                                                          int *__restrict__ ovf) {
  __shared__ int32_t hkey[kBucketHash];
  __shared__ uint32_t hcnt[kBucketHash];
  __shared__ uint32_t hbase[kBucketHash];
  const int t = threadIdx.x;
  for (int i = t; i < kBucketHash; i += blockDim.x) { hkey[i] = -1; hcnt[i] = 0; }
  __syncthreads();
  int32_t v[kBucketIpt];
  uint32_t h[kBucketIpt], rank[kBucketIpt];
#pragma unroll
  for (int j = 0; j < kBucketIpt; ++j) {
    const int64_t i = (int64_t)blockIdx.x * (256 * kBucketIpt) + j * 256 + t;
    v[j] = i < m ? elems[i] : -1;
  }
#pragma unroll
  for (int j = 0; j < kBucketIpt; ++j) {
    h[j] = 0;
    rank[j] = 0;
    if (v[j] >= 0) {
      h[j] = ((uint32_t)v[j] * 2654435761u) >> (32 - kBucketLog);
      for (;;) {  // ends: at most half the slots hold keys
        const int32_t old = atomicCAS(&hkey[h[j]], -1, v[j]);
        if (old == -1 || old == v[j]) break;
        h[j] = (h[j] + 1) & (kBucketHash - 1);
      }
      rank[j] = atomicAdd(&hcnt[h[j]], 1u);
    }
  }
  __syncthreads();
  for (int q = t; q < kBucketHash; q += blockDim.x) {
    const int32_t key = hkey[q];
    if (key >= 0) hbase[q] = atomicAdd(&deg[key], hcnt[q]);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kBucketIpt; ++j) {
    if (v[j] < 0) continue;
    const int64_t i = (int64_t)blockIdx.x * (256 * kBucketIpt) + j * 256 + t;
    const uint32_t pos = hbase[h[j]] + rank[j];
    if (pos < 32) bucket[32 * (int64_t)v[j] + pos] = (int32_t)i;
    else atomicOr(ovf, 1);
  }
}

// bucketed lists (up to 32 entries) put into element order and packed at
// their CSR offsets, two dofs per wave: each entry goes to its rank among
// the dof's entries (distinct items), counted over the bucket staged in LDS
__global__ __launch_bounds__(256) void k_v2e_sort_bucket(int64_t n, const int64_t *__restrict__ ptr,
                                                         const int32_t *__restrict__ bucket,
                                                         int32_t *__restrict__ v2e,
                                                         int32_t *__restrict__ inv) {
  __shared__ __align__(16) int32_t ids[256];
  const int hl = threadIdx.x % kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  int64_t k0 = 0;
  int deg = 0;
  if (v < n) { k0 = ptr[v]; deg = (int)(ptr[v + 1] - k0); }
  const int32_t x = hl < deg ? bucket[32 * v + hl] : INT_MAX;
  ids[threadIdx.x] = x;
  wave_lds_sync();
  const int4 *row = reinterpret_cast<const int4 *>(ids + (threadIdx.x & ~(kHalf - 1)));
  int rho = 0;
#pragma unroll
  for (int g = 0; g < kHalf / 4; ++g) {
    const int4 w = row[g];
    rho += (w.x < x) + (w.y < x) + (w.z < x) + (w.w < x);
  }
  if (hl < deg) {
    v2e[k0 + rho] = x;
    if (inv) inv[x] = (int32_t)(k0 + rho);
  }
}

// position of every (element, local dof) item in the incidence lists
__global__ void k_v2e_inverse(int64_t m, const int32_t *__restrict__ v2e, int32_t *__restrict__ inv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) inv[v2e[k]] = (int32_t)k;
}

// per-dof insertion sort of the lists longer than 32 entries
__global__ void k_v2e_sort(int64_t n, const int64_t *__restrict__ ptr, int32_t *__restrict__ v2e) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t a = ptr[v], b = ptr[v + 1];
  if (b - a <= kHalf) return;
  for (int64_t i = a + 1; i < b; ++i) {
    const int32_t key = v2e[i];
    int64_t j = i - 1;
    while (j >= a && v2e[j] > key) { v2e[j + 1] = v2e[j]; --j; }
    v2e[j + 1] = key;
  }
}

// ---------------------------------------------------------------- pattern (half-wave rows)
// One row per 32-lane half-wave: lane q loads incidence q's element vertices
// (4 candidates per lane), inserts them into an open-addressing set in LDS
// (atomicCAS; the lane that claims a slot appends the column), and the
// unique columns are ordered by a counting sort (rank = number of smaller
// columns).  One pass writes the row length, the sorted columns to a 64-entry
// row buffer, the relative diagonal slot and the element-to-row map (each
// candidate's offset by binary search in the sorted columns); after the scan,
// k_pattern_compact moves the buffered columns to their CSR place.  Rows with
// more than 32 incident elements or 64 columns are flagged in big[] for the
// lane-per-row merge (k_pattern_rows).
constexpr int kRowBuf = 64;

constexpr int kHash = 128;  // LDS hash slots per row (a row has at most 1 + 3 * 32 columns)

// Where the pass writes a row it builds.  Buffered: the 64-entry row
// buffers (rowbuf, rowval), the length and the diagonal offset, moved to CSR
// by k_pattern_compact after the scan of the lengths.  DIRECT: the row
// pointers are known beforehand (the vertex-link counts, k_link_len) and the
// row goes straight to colidx / vals / diag; a row whose length differs from
// its pointers sets *mismatch and is not written (the caller then takes the
// buffered pass).
struct PatternOut {
  int64_t *rowlen = nullptr;
  int32_t *rowbuf = nullptr;
  int8_t *dposr = nullptr;
  double *rowval = nullptr;
  const int64_t *rowptr = nullptr;
  int32_t *colidx = nullptr;
  int64_t *diag = nullptr;
  double *vals = nullptr;
  int *mismatch = nullptr;
};

// VALS (3D P1 stiffness): the same pass also forms the row's values from the
// element scales Ce (the rows of k_p1_values_rows).
//
// With bucket (the incidence lists unsorted, 32 entries per row, every list
// within its bucket; VALS only): each lane ranks its incidence among the
// row's and writes it to v2e at that rank, in place of k_v2e_sort_bucket.
template <int NLOC, bool VALS, bool DIRECT>
__global__ __launch_bounds__(256) void k_pattern_half(
    int64_t n, const int32_t *__restrict__ elems, const int64_t *__restrict__ v2e_ptr,
    int32_t *__restrict__ v2e, uint32_t *__restrict__ slotmap, uint8_t *__restrict__ big,
    int *__restrict__ any_big, PatternOut out, const double *__restrict__ xyz = nullptr,
    const double *__restrict__ Ce = nullptr, const int32_t *__restrict__ bucket = nullptr) {
  constexpr int NH = 256 / kHalf;
  // every LDS array below is per row (half-wave), so a wave orders its own
  // accesses (wave_lds_sync) and the waves of a block run independently
  __shared__ int32_t hk[NH][kHash];      // open-addressing set of the row's columns
  __shared__ uint8_t hslot[NH][kRowBuf];  // the set slot of each unique column
  __shared__ uint8_t hrank[NH][kHash];    // the offset in the row of the column in each set slot
  __shared__ int32_t ucol[NH][kRowBuf];  // unique columns, then sorted
  __shared__ int32_t ucnt[NH];
  __shared__ double sv[VALS ? NH : 1][VALS ? 4 * kHalf : 1];   // incidence rows (VALS)
  __shared__ __align__(16) uint32_t so[VALS ? NH : 1][VALS ? kHalf : 1];  // their slot bytes (VALS)
  __shared__ __align__(16) double dq[VALS ? NH : 1][VALS ? kHalf : 1];    // their diagonal entries (VALS)
  __shared__ uint32_t smask[VALS ? NH : 1][VALS ? kHalf : 1];  // per off-diagonal slot: its incidences (VALS)
  const int hl = threadIdx.x % kHalf, hid = threadIdx.x / kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  int64_t k0 = 0, p0 = 0, p1 = 0;
  int deg = 0;
  if (v < n) {
    k0 = v2e_ptr[v];
    deg = (int)(v2e_ptr[v + 1] - k0);
    if (DIRECT) { p0 = out.rowptr[v]; p1 = out.rowptr[v + 1]; }
  }
  bool mine = v < n && deg <= kHalf;
  if (!mine) deg = 0;
  for (int i = hl; i < kHash; i += kHalf) hk[hid][i] = INT_MAX;
  if (hl == 0) ucnt[hid] = 0;
  if constexpr (VALS) smask[hid][hl] = 0;
  int32_t c[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
  int64_t e = 0;
  int lv = 0;
  double vv[VALS ? 4 : 1][3], Cv = 0.;
  int32_t idx = INT_MAX;
  int qp = hl;  // the incidence's position in the row's list (element order)
  bool ranked = false;
  if constexpr (VALS) {
    if (bucket) {
      // unsorted bucket: each lane's rank among the row's entries (element
      // ids are distinct) places its incidence, and everything it stores
      // per incidence below, in element order
      idx = hl < deg ? bucket[32 * v + hl] : INT_MAX;
      so[hid][hl] = (uint32_t)idx;
      wave_lds_sync();
      int rho = 0;
#pragma unroll
      for (int g = 0; g < kHalf / 4; ++g) {
        const uint4 w = *reinterpret_cast<const uint4 *>(&so[hid][4 * g]);
        rho += ((int32_t)w.x < idx) + ((int32_t)w.y < idx) + ((int32_t)w.z < idx) + ((int32_t)w.w < idx);
      }
      qp = rho;
      if (hl < deg) v2e[k0 + qp] = idx;
      ranked = true;
    }
  }
  if (!ranked && hl < deg) idx = v2e[k0 + hl];
  if (hl < deg) {
    e = idx / NLOC;
    lv = idx - (int)e * NLOC;
    if (NLOC == 4) {
      const int4 t = reinterpret_cast<const int4 *>(elems)[e];
      c[0] = t.x; c[1] = t.y; c[2] = t.z; c[3] = t.w;
    } else {
#pragma unroll
      for (int m = 0; m < NLOC; ++m) c[m] = elems[e * NLOC + m];
    }
    if constexpr (VALS) {  // issued now, used after the set and sort phases (their latency hidden)
#pragma unroll
      for (int i = 0; i < (VALS ? 4 : 1); ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) vv[i][k] = xyz[3 * (int64_t)c[i] + k];
      Cv = Ce[e];
    }
  }
  wave_lds_sync();
  // insert the candidates; the lane that claims a free slot appends the column.
  // The row's own vertex is in every incident element: only lane 0 inserts it
  // (a quarter fewer atomics; c[] keeps it for the element-to-row map below)
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (c[r] == INT_MAX || (c[r] == (int32_t)v && hl != 0)) continue;
    uint32_t h = ((uint32_t)c[r] * 2654435761u) >> (32 - 7);
    for (;;) {  // ends: the set holds at most 97 columns of its 128 slots
      const int32_t old = atomicCAS(&hk[hid][h], INT_MAX, c[r]);
      if (old == INT_MAX) {
        const int idx = atomicAdd(&ucnt[hid], 1);
        if (idx < kRowBuf) { ucol[hid][idx] = c[r]; hslot[hid][idx] = (uint8_t)h; }
        break;
      }
      if (old == c[r]) break;
      h = (h + 1) & (kHash - 1);
    }
  }
  wave_lds_sync();
  const int len = ucnt[hid];
  if (len > kRowBuf) mine = false;
  const int64_t r0 = p0;
  const bool wrong = DIRECT && mine && p1 - p0 != len;  // the row's length is not its pointers' 
  // counting sort: the rank of a column is the number of smaller columns
  int32_t u[2] = {INT_MAX, INT_MAX};
  int rk[2] = {0, 0};
  if (mine) {
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int j = hl + h2 * kHalf;
      if (j < len) {
        u[h2] = ucol[hid][j];
        int r = 0;
        for (int i = 0; i < len; ++i) r += ucol[hid][i] < u[h2] ? 1 : 0;
        rk[h2] = r;
      }
    }
  }
  wave_lds_sync();
  int dpos = -1;
  if (mine) {
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2)
      if (hl + h2 * kHalf < len) {
        hrank[hid][hslot[hid][hl + h2 * kHalf]] = (uint8_t)rk[h2];
        ucol[hid][rk[h2]] = u[h2];
        if (!DIRECT) out.rowbuf[kRowBuf * v + rk[h2]] = u[h2];
        else if (!wrong) out.colidx[r0 + rk[h2]] = u[h2];
        if ((int64_t)u[h2] == v) dpos = rk[h2];
      }
  }
#pragma unroll
  for (int off = 1; off < kHalf; off <<= 1) {
    const int t = __shfl_xor(dpos, off, kWave);
    dpos = t > dpos ? t : dpos;
  }
  wave_lds_sync();
  if (v < n && hl == 0) {
    big[v] = mine ? 0 : 1;
    if (!mine) atomicOr(any_big, 1);
    if (DIRECT && wrong) atomicOr(out.mismatch, 1);
    if (mine && !DIRECT) { out.rowlen[v] = len; out.dposr[v] = (int8_t)dpos; }
    if (mine && DIRECT && !wrong) out.diag[v] = dpos < 0 ? -1 : r0 + dpos;
  }
  if (mine && hl < deg) {
    uint32_t pk = 0;
#pragma unroll
    for (int m = 0; m < NLOC; ++m) {
      // the column is in the set: its slot holds its offset in the row
      uint32_t h = ((uint32_t)c[m] * 2654435761u) >> (32 - 7);
      while (hk[hid][h] != c[m]) h = (h + 1) & (kHash - 1);
      pk |= (uint32_t)hrank[hid][h] << (8 * m);
    }
    slotmap[k0 + qp] = pk;
    if constexpr (VALS) {
      double r[4];
      p1_row_entries(vv, lv, Cv, r);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        sv[hid][4 * qp + m] = r[m];
        const uint32_t slot = (pk >> (8 * m)) & 0xffu;
        if (m != lv && slot < kHalf) atomicOr(&smask[hid][slot], 1u << qp);
      }
      dq[hid][qp] = r[lv];
      so[hid][qp] = pk;
    }
  }
  if constexpr (VALS) {  // the row's values, each slot summed by its lane in element order
    wave_lds_sync();
    const int nq = mine ? deg : 0;
    const uint32_t me0 = (uint32_t)hl * 0x01010101u, me1 = (uint32_t)(hl + kHalf) * 0x01010101u;
    double acc0 = 0., acc1 = 0.;
    // Every incidence holds the diagonal slot: its lane adds the incidences'
    // diagonal entries in order, 8 per round (reads before adds; past the
    // row's count +0.0, which leaves the sum unchanged: it starts at +0.0
    // and is never -0.0).  Another slot's lane adds the entries of the
    // incidences in its mask (bit q: incidence q holds the slot), lowest
    // first, 4 per round.  Both orders are the element order.
    if (hl == dpos) {
      for (int q0 = 0; q0 < nq; q0 += 8) {
        double a[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = dq[hid][q0 + i];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc0 += q0 + i < nq ? a[i] : 0.;
      }
    } else if (hl < len) {
      uint32_t mask = mine ? smask[hid][hl] : 0u;
      while (mask) {
        int q[4];
        bool ok[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          ok[i] = mask != 0;
          q[i] = ok[i] ? __builtin_ctz(mask) : 0;
          mask &= mask - 1;
        }
        double a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t t0 = so[hid][q[i]] ^ me0, z0 = (t0 - 0x01010101u) & ~t0 & 0x80808080u;
          a[i] = sv[hid][4 * q[i] + (z0 ? __builtin_ctz(z0) >> 3 : 0)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) acc0 += ok[i] ? a[i] : 0.;
      }
    }
    if (len > kHalf)  // rows of more than 32 columns: the second slot of each lane
      for (int q = 0; q < nq; ++q) {
        const uint32_t t1 = so[hid][q] ^ me1, z1 = (t1 - 0x01010101u) & ~t1 & 0x80808080u;
        if (z1) acc1 += sv[hid][4 * q + (__builtin_ctz(z1) >> 3)];
      }
    if (mine && !DIRECT) {
      if (hl < len) out.rowval[kRowBuf * v + hl] = acc0;
      if (hl + kHalf < len) out.rowval[kRowBuf * v + hl + kHalf] = acc1;
    }
    if (mine && DIRECT && !wrong) {
      if (hl < len) out.vals[r0 + hl] = acc0;
      if (hl + kHalf < len) out.vals[r0 + hl + kHalf] = acc1;
    }
  }
}

// The P1 row lengths a manifold simplicial mesh implies, from the counts of
// incident elements (deg) and boundary elements (b) of each vertex: the link
// of a vertex is a sphere (3D) / circle (2D) inside and a disk / path on the
// boundary, so Euler's formula gives its vertex count — 3D: 2 + deg/2 inside,
// 1 + (deg + b)/2 on the boundary; 2D: deg inside, deg + 1 on the boundary —
// and the row holds those vertices and its own.  Meshes where this does not
// hold (boundary elements inside the domain, non-manifold vertices) are found
// by the pattern pass, which compares every row with these lengths.
__global__ void k_bcount(int64_t m, const int32_t *__restrict__ belems, uint32_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) atomicAdd(&cnt[belems[i]], 1u);
}

__global__ void k_link_len(int64_t n, int dim, const int64_t *__restrict__ v2e_ptr,
                           const uint32_t *__restrict__ bcnt, int64_t *__restrict__ len) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t deg = v2e_ptr[v + 1] - v2e_ptr[v], b = bcnt[v];
  int64_t l;
  if (dim == 3) l = b == 0 ? 3 + deg / 2 : 2 + (deg + b) / 2;
  else l = b == 0 ? 1 + deg : 2 + deg;
  len[v] = deg == 0 ? 0 : l;
}

// buffered columns to their CSR place; diagonal slots.  One lane per row:
// the row's buffer is read as 16-byte groups, all issued before the stores,
// so a wave keeps 64 rows' loads in flight.
// With rowval, the row values too (2 doubles per 16-byte load).
__global__ __launch_bounds__(256) void k_pattern_compact(
    int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ rowbuf,
    const int8_t *__restrict__ dposr, const uint8_t *__restrict__ big,
    int32_t *__restrict__ colidx, int64_t *__restrict__ diag,
    const double *__restrict__ rowval = nullptr, double *__restrict__ vals = nullptr) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || big[v]) return;
  const int64_t r0 = rowptr[v];
  const int len = (int)(rowptr[v + 1] - r0);
  if (rowval) {
    const double2 *vs = reinterpret_cast<const double2 *>(rowval + kRowBuf * v);
    double *vd = vals + r0;
    for (int g0 = 0; g0 < len; g0 += 8) {  // 8 values (4 loads) per round
      double2 q[4];
#pragma unroll
      for (int g = 0; g < 4; ++g)
        if (g0 + 2 * g < len) q[g] = vs[(g0 >> 1) + g];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        if (g0 + 2 * g < len) vd[g0 + 2 * g] = q[g].x;
        if (g0 + 2 * g + 1 < len) vd[g0 + 2 * g + 1] = q[g].y;
      }
    }
  }
  const int4 *src = reinterpret_cast<const int4 *>(rowbuf + kRowBuf * v);
  int32_t *dst = colidx + r0;
  for (int g0 = 0; g0 < len; g0 += 16) {  // 16 columns (4 groups) per round
    int4 q[4];
#pragma unroll
    for (int g = 0; g < 4; ++g)
      if (g0 + 4 * g < len) q[g] = src[(g0 >> 2) + g];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int32_t e[4] = {q[g].x, q[g].y, q[g].z, q[g].w};
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (g0 + 4 * g + t < len) dst[g0 + 4 * g + t] = e[t];
    }
  }
  diag[v] = dposr[v] < 0 ? -1 : r0 + dposr[v];
}

// ---------------------------------------------------------------- element matrices
// 3D P1: Ke_ij = S grad l_i . grad l_j with S = sum_q |K| w_q kappa(x_q)
// (problem.cpp:6387-6392); the 10 unique entries per element.
__device__ inline double p1_kappa_integral(int64_t e, const int32_t id[4], const double v[4][3],
                                           double mes, int kind, double c0,
                                           const double *__restrict__ kappa_e) {
  double S = 0.;
  if (kind == FFHIP_COEF_P1) {
    const double u[4] = {kappa_e[id[0]], kappa_e[id[1]], kappa_e[id[2]], kappa_e[id[3]]};
#pragma unroll
    for (int q = 0; q < 14; ++q)
      S += (mes * qf5_weight(q)) * p1_value3(u, c_qf5_pts[q][0], c_qf5_pts[q][1], c_qf5_pts[q][2]);
  } else if (kind == FFHIP_COEF_SIN3) {
#pragma unroll 2
    for (int q = 0; q < 14; ++q) {
      const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
      const double l0 = 1. - (px + py + pz);
      double P[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) P[k] = l0 * v[0][k] + px * v[1][k] + py * v[2][k] + pz * v[3][k];
      S += (mes * qf5_weight(q)) * kappa_sin3(P[0], P[1], P[2]);
    }
  } else if (kind == FFHIP_COEF_QP) {
#pragma unroll
    for (int q = 0; q < 14; ++q) S += (mes * qf5_weight(q)) * kappa_e[14 * e + q];
  } else {
    const double kc = (kind == FFHIP_COEF_P0) ? kappa_e[e] : c0;
#pragma unroll
    for (int q = 0; q < 14; ++q) S += (mes * qf5_weight(q)) * kc;
  }
  return S;
}

__device__ inline void load_tet(const double *__restrict__ xyz, const int32_t *__restrict__ tets,
                                int64_t e, int32_t id[4], double v[4][3]) {
  const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
  id[0] = tv.x; id[1] = tv.y; id[2] = tv.z; id[3] = tv.w;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
}

// the reaction term of int3d(m*u*v): K(i,j) += sum_q |K| w_q m(x_q) l_i(q) l_j(q)
// for the 10 entries i <= j (P1 values l = barycentrics of the Tet_5 point)
__device__ inline void p1_mass_3d(int64_t e, const int32_t id[4], const double v[4][3], double mes,
                                  int mkind, double m0, const double *__restrict__ mass_e,
                                  double K[10]) {
  const double mp0 = (mkind == FFHIP_COEF_P0) ? mass_e[e] : m0;
  double mu[4] = {0., 0., 0., 0.};
  if (mkind == FFHIP_COEF_P1)
#pragma unroll
    for (int i = 0; i < 4; ++i) mu[i] = mass_e[id[i]];
#pragma unroll 2
  for (int q = 0; q < 14; ++q) {
    const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
    const double l[4] = {1. - (px + py + pz), px, py, pz};
    double mq = mp0;
    if (mkind == FFHIP_COEF_SIN3) {
      double P[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) P[k] = l[0] * v[0][k] + px * v[1][k] + py * v[2][k] + pz * v[3][k];
      mq = kappa_sin3(P[0], P[1], P[2]);
    } else if (mkind == FFHIP_COEF_QP) {
      mq = mass_e[14 * e + q];
    } else if (mkind == FFHIP_COEF_P1) {
      mq = p1_value3(mu, px, py, pz);
    }
    const double W = (mes * qf5_weight(q)) * mq;
    int o = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i; j < 4; ++j) K[o++] += (W * l[i]) * l[j];
  }
}

// entry (i, j), i <= j, of the P1 element matrix: sum_k (S G_ik) G_jk
__device__ inline double p1_entry(double S, const double G[4][3], int i, int j) {
  double a = 0.;
#pragma unroll
  for (int k = 0; k < 3; ++k) a += (S * G[i][k]) * G[j][k];
  return a;
}

// the element's 4 rows, each written where its incidence entry sits in the
// vertex's list (inv): Kinc[4k + m] = Ke(l, m) for the incidence k of (e, l),
// so the row gather reads them contiguously; with Ke != null also the 10
// unique entries per element (i <= j, row-major upper triangle) for the
// lane-per-row gather
// MASS: with the reaction term m*u*v (coefficient kind mkind, m0, mass_e)
template <bool MASS>
__global__ __launch_bounds__(256) void k_p1_element_3d(
    int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
    int kind, double c0, const double *__restrict__ kappa_e, int mkind, double m0,
    const double *__restrict__ mass_e, const int32_t *__restrict__ inv,
    double *__restrict__ Kinc, double *__restrict__ Ke) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double v[4][3], G[4][3];
  int32_t id[4];
  load_tet(xyz, tets, e, id, v);
  const double mes = tet_geom(v, G);
  const double S = p1_kappa_integral(e, id, v, mes, kind, c0, kappa_e);
  double K[10];
  int o = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) K[o++] = p1_entry(S, G, i, j);
  if (MASS) p1_mass_3d(e, id, v, mes, mkind, m0, mass_e, K);
  const int4 pos = reinterpret_cast<const int4 *>(inv)[e];
  const int32_t kk[4] = {pos.x, pos.y, pos.z, pos.w};
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    double4 row;
    row.x = K[sym4(l, 0)]; row.y = K[sym4(l, 1)]; row.z = K[sym4(l, 2)]; row.w = K[sym4(l, 3)];
    reinterpret_cast<double4 *>(Kinc)[kk[l]] = row;
  }
  if (Ke) {
    double *dst = Ke + 10 * e;
#pragma unroll
    for (int i = 0; i < 10; ++i) dst[i] = K[i];
  }
}


// 2D P1: 6 unique entries; TTriangle area and H(i) (fem.hpp:290-325)
__global__ __launch_bounds__(256) void k_p1_element_2d(
    int64_t nt, const double *__restrict__ xy, const int32_t *__restrict__ tris,
    int kind, double c0, const double *__restrict__ kappa_e, int mkind, double m0,
    const double *__restrict__ mass_e, double *__restrict__ Ke) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double v[3][2];
  int32_t id[3];
  for (int i = 0; i < 3; ++i) {
    id[i] = tris[3 * e + i];
    for (int k = 0; k < 2; ++k) v[i][k] = xy[2 * (int64_t)id[i] + k];
  }
  const double area = ((v[1][0] - v[0][0]) * (v[2][1] - v[0][1]) -
                       (v[1][1] - v[0][1]) * (v[2][0] - v[0][0])) * 0.5;
  double G[3][2];
  for (int i = 0; i < 3; ++i) {
    const int a = (i + 2) % 3, b = (i + 1) % 3;
    const double Ex = v[a][0] - v[b][0], Ey = v[a][1] - v[b][1];
    G[i][0] = -Ey / (2 * area);
    G[i][1] = Ex / (2 * area);
  }
  double S = 0.;
  for (int q = 0; q < 7; ++q) {
    double kc;
    if (kind == FFHIP_COEF_SIN3) {
      const double px = c_qt5_pts[q][0], py = c_qt5_pts[q][1];
      const double l0 = 1. - px - py;
      const double X = l0 * v[0][0] + px * v[1][0] + py * v[2][0];
      const double Y = l0 * v[0][1] + px * v[1][1] + py * v[2][1];
      const double tp = 2 * M_PI;
      kc = 1 + (0.5 * sin(tp * X)) * sin(tp * Y);
    } else if (kind == FFHIP_COEF_QP) {
      kc = kappa_e[7 * e + q];
    } else if (kind == FFHIP_COEF_P1) {
      const double u[3] = {kappa_e[id[0]], kappa_e[id[1]], kappa_e[id[2]]};
      kc = p1_value2(u, c_qt5_pts[q][0], c_qt5_pts[q][1]);
    } else {
      kc = (kind == FFHIP_COEF_P0) ? kappa_e[e] : c0;
    }
    S += (area * c_qt5_w[q]) * kc;
  }
  double K[6];
  int o = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      double a = 0.;
      for (int k = 0; k < 2; ++k) a += (S * G[i][k]) * G[j][k];
      K[o++] = a;
    }
  if (mkind >= 0) {  // the reaction term m*u*v with T_5
    for (int q = 0; q < 7; ++q) {
      const double px = c_qt5_pts[q][0], py = c_qt5_pts[q][1];
      const double l[3] = {1. - px - py, px, py};
      double mq = (mkind == FFHIP_COEF_P0) ? mass_e[e] : m0;
      if (mkind == FFHIP_COEF_SIN3) {
        const double X = l[0] * v[0][0] + px * v[1][0] + py * v[2][0];
        const double Y = l[0] * v[0][1] + px * v[1][1] + py * v[2][1];
        const double tp = 2 * M_PI;
        mq = 1 + (0.5 * sin(tp * X)) * sin(tp * Y);
      } else if (mkind == FFHIP_COEF_QP) {
        mq = mass_e[7 * e + q];
      } else if (mkind == FFHIP_COEF_P1) {
        const double u[3] = {mass_e[id[0]], mass_e[id[1]], mass_e[id[2]]};
        mq = p1_value2(u, px, py);
      }
      const double W = (area * c_qt5_w[q]) * mq;
      o = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) K[o++] += (W * l[i]) * l[j];
    }
  }
  for (int i = 0; i < 6; ++i) Ke[6 * e + i] = K[i];
}

// ---------------------------------------------------------------- row gather
// Block of BLOCK rows; the block's CSR segment is staged in LDS when it fits
// (STAGE doubles), else each lane accumulates its own row in global memory.
template <int BLOCK, int STAGE, int NLOC>
__global__ __launch_bounds__(BLOCK) void k_p1_gather(
    int64_t n, const int64_t *__restrict__ rowptr, const int64_t *__restrict__ v2e_ptr,
    const int32_t *__restrict__ v2e, const uint32_t *__restrict__ slotmap,
    const double *__restrict__ Ke, double *__restrict__ vals,
    const uint8_t *__restrict__ only = nullptr) {
  constexpr int NU = NLOC * (NLOC + 1) / 2;
  __shared__ double S[STAGE];
  const int t = threadIdx.x;
  const int64_t v0 = (int64_t)blockIdx.x * BLOCK;
  const int64_t v = v0 + t;
  const int64_t vend = (v0 + BLOCK < n) ? v0 + BLOCK : n;
  const int64_t seg0 = rowptr[v0], seg1 = rowptr[vend];
  const bool staged = (seg1 - seg0) <= STAGE;
  if (staged) {
    for (int64_t i = t; i < seg1 - seg0; i += BLOCK) S[i] = 0.;
    __syncthreads();
  }
  if (v < n && (!only || only[v])) {  // only[]: the rows this kernel owns
    const int64_t r0 = rowptr[v];
    double *acc = staged ? S + (r0 - seg0) : vals + r0;
    if (!staged)
      for (int64_t i = r0; i < rowptr[v + 1]; ++i) vals[i] = 0.;
    const int64_t k0 = v2e_ptr[v], k1 = v2e_ptr[v + 1];
    for (int64_t k = k0; k < k1; ++k) {
      const int32_t idx = v2e[k];
      const int64_t e = idx / NLOC;
      const int l = idx - (int)e * NLOC;
      const uint32_t pk = slotmap[k];
      const double *ke = Ke + NU * e;
#pragma unroll
      for (int m = 0; m < NLOC; ++m) {
        const int off = (pk >> (8 * m)) & 0xff;
        const double a = ke[NLOC == 4 ? sym4(l, m) : sym3(l, m)];
        acc[off] += a;
      }
    }
  }
  if (staged) {
    __syncthreads();
    if (!only) {
      for (int64_t i = t; i < seg1 - seg0; i += BLOCK) vals[seg0 + i] = S[i];
    } else {
      for (int64_t i = t; i < seg1 - seg0; i += BLOCK) {
        const int64_t g = seg0 + i;
        int64_t lo = v0, hi = vend - 1;  // the row of slot g
        while (lo < hi) {
          const int64_t mid = (lo + hi + 1) >> 1;
          if (rowptr[mid] <= g) lo = mid; else hi = mid - 1;
        }
        if (only[lo]) vals[g] = S[i];
      }
    }
  }
}

// 3D values, one row per 32-lane half-wave: lane q takes incidence q (element
// order), whose 4 row entries the element kernel stored at Kinc[4k..4k+3],
// and stages them and their 4 slot bytes in LDS; lane j then walks the
// incidences in order and adds the entry (if any) whose slot is j (lanes own
// slots j and j + 32) — the same sums, in the same order, as a sequential walk
// over the row's elements.  Rows flagged big[] (more than 32 elements or 64
// columns) are left to k_p1_gather.
__global__ __launch_bounds__(256) void k_p1_values_half(
    int64_t n, const int64_t *__restrict__ rowptr, const int64_t *__restrict__ v2e_ptr,
    const uint32_t *__restrict__ slotmap, const double *__restrict__ Kinc,
    const uint8_t *__restrict__ big, double *__restrict__ vals) {
  __shared__ double sv[256 / kHalf][4 * kHalf];
  __shared__ uint32_t so[256 / kHalf][kHalf];  // 4 slot bytes per incidence
  const int hl = threadIdx.x % kHalf, hid = threadIdx.x / kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  int64_t k0 = 0, r0 = 0;
  int deg = 0, len = 0;
  if (v < n && !big[v]) {
    k0 = v2e_ptr[v];
    deg = (int)(v2e_ptr[v + 1] - k0);
    r0 = rowptr[v];
    len = (int)(rowptr[v + 1] - r0);
  }
  if (hl < deg) {
    const double4 row = reinterpret_cast<const double4 *>(Kinc)[k0 + hl];
    sv[hid][4 * hl + 0] = row.x;
    sv[hid][4 * hl + 1] = row.y;
    sv[hid][4 * hl + 2] = row.z;
    sv[hid][4 * hl + 3] = row.w;
    so[hid][hl] = slotmap[k0 + hl];
  }
  __syncthreads();
  const uint32_t me0 = (uint32_t)hl * 0x01010101u, me1 = (uint32_t)(hl + kHalf) * 0x01010101u;
  double acc0 = 0., acc1 = 0.;
  for (int q = 0; q < deg; ++q) {
    // lowest byte of so[q] equal to the lane's slot (at most one is)
    const uint32_t pk = so[hid][q];
    const uint32_t t0 = pk ^ me0, z0 = (t0 - 0x01010101u) & ~t0 & 0x80808080u;
    if (z0) acc0 += sv[hid][4 * q + (__builtin_ctz(z0) >> 3)];
    if (len > kHalf) {
      const uint32_t t1 = pk ^ me1, z1 = (t1 - 0x01010101u) & ~t1 & 0x80808080u;
      if (z1) acc1 += sv[hid][4 * q + (__builtin_ctz(z1) >> 3)];
    }
  }
  if (hl < len) vals[r0 + hl] = acc0;
  if (hl + kHalf < len) vals[r0 + hl + kHalf] = acc1;
}

// ---------------------------------------------------------------- 3D stiffness in two passes
// K_e(l, m) = S_e grad l_l . grad l_m (problem.cpp:6387-6392) with
// grad l_l = N_l / (6 |K|), N_l the cross products of the edge vectors from
// vertex 0 (Tet::Gradlambda, Mesh3dn.hpp:126-137, before its division) and
// N_0 = -(N_1 + N_2 + N_3): the element pass stores the one scale
// C_e = S_e / (6 |K|)^2 per element, and the row pass forms the row's entries
// C_e N_l . N_m from the vertices — no element rows through HBM.
__global__ __launch_bounds__(256) void k_p1_scale_3d(int64_t nt, const double *__restrict__ xyz,
                                                     const int32_t *__restrict__ tets, int kind,
                                                     double c0, const double *__restrict__ kappa_e,
                                                     double *__restrict__ Ce) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double v[4][3];
  int32_t id[4];
  load_tet(xyz, tets, e, id, v);
  const double mes = tet_mesure(v);
  const double S = p1_kappa_integral(e, id, v, mes, kind, c0, kappa_e);
  const double d1 = 1. / (6. * mes);
  Ce[e] = (S * d1) * d1;
}

// The incidence buckets of k_v2e_bucket_block fused with the element scales
// of k_p1_scale_3d (tets): a block takes 256 tets (1024 items, 4 per thread,
// a 2048-slot LDS hash set); the thread's tet is loaded with its items, the
// block's global count atomics are issued before the tet's scale is
// computed (their round trip hidden by the arithmetic), and the items are
// placed after it.  The scale is the same computation as k_p1_scale_3d's.
// (5 waves per SIMD: 86 VGPRs, no spills; uncapped it takes 108 and runs 4)
constexpr int kFuseLog = 11;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5))) void k_v2e_bucket_scale(const int32_t *__restrict__ tets, int64_t nt,
                                                          ScaleArgs sc, uint32_t *__restrict__ deg,
                                                          int32_t *__restrict__ bucket,
                                                          int *__restrict__ ovf) {
  constexpr int IPT = 4, HS = 1 << kFuseLog;
  __shared__ int32_t hkey[HS];
  __shared__ uint32_t hcnt[HS];
  __shared__ uint32_t hbase[HS];
  const int t = threadIdx.x;
  const int64_t m = 4 * nt, b0 = (int64_t)blockIdx.x * (256 * IPT);
  for (int i = t; i < HS; i += blockDim.x) { hkey[i] = -1; hcnt[i] = 0; }
  int32_t v[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int64_t i = b0 + j * 256 + t;
    v[j] = i < m ? tets[i] : -1;
  }
  const int64_t e = (int64_t)blockIdx.x * 256 + t;
  double tv[4][3];
  int32_t id[4];
  if (e < nt) load_tet(sc.xyz, tets, e, id, tv);
  __syncthreads();
  uint32_t h[IPT], rank[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    h[j] = 0;
    rank[j] = 0;
    if (v[j] >= 0) {
      h[j] = ((uint32_t)v[j] * 2654435761u) >> (32 - kFuseLog);
      for (;;) {  // ends: at most half the slots hold keys
        const int32_t old = atomicCAS(&hkey[h[j]], -1, v[j]);
        if (old == -1 || old == v[j]) break;
        h[j] = (h[j] + 1) & (HS - 1);
      }
      rank[j] = atomicAdd(&hcnt[h[j]], 1u);
    }
  }
  __syncthreads();
  uint32_t hb[HS / 256];
#pragma unroll
  for (int q = 0; q < HS / 256; ++q) {
    const int32_t key = hkey[t + q * 256];
    hb[q] = key >= 0 ? atomicAdd(&deg[key], hcnt[t + q * 256]) : 0u;
  }
  if (e < nt) {
    const double mes = tet_mesure(tv);
    const double S = p1_kappa_integral(e, id, tv, mes, sc.kind, sc.c0, sc.kappa_e);
    const double d1 = 1. / (6. * mes);
    sc.Ce[e] = (S * d1) * d1;
  }
#pragma unroll
  for (int q = 0; q < HS / 256; ++q) hbase[t + q * 256] = hb[q];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (v[j] < 0) continue;
    const int64_t i = b0 + j * 256 + t;
    const uint32_t pos = hbase[h[j]] + rank[j];
    if (pos < 32) bucket[32 * (int64_t)v[j] + pos] = (int32_t)i;
    else atomicOr(ovf, 1);
  }
}

// values, one row per 32-lane half-wave (rows of up to 32 elements and 64
// columns; big[] rows are left to the lane-per-row gather): lane q takes
// incidence q (element order), forms its element's row l from the vertices
// and C_e, and stages the 4 entries with their slot bytes (the element-to-row
// map of the pattern) in LDS; lane j then adds, in element order, the entries
// whose slot is j (lanes own slots j and j + 32) — the summation order of
// HashMatrix::operator+= over the element loop.  Blocks take contiguous row
// ranges per XCD (the gathered vertices and element scales of neighbouring
// rows stay in that XCD's L2).
__global__ __launch_bounds__(256) void k_p1_values_rows(
    int64_t n, const int64_t *__restrict__ rowptr, const int64_t *__restrict__ v2e_ptr,
    const int32_t *__restrict__ v2e, const uint32_t *__restrict__ slotmap,
    const double *__restrict__ xyz, const int32_t *__restrict__ tets, const double *__restrict__ Ce,
    const uint8_t *__restrict__ big, double *__restrict__ vals) {
  __shared__ double sv[256 / kHalf][4 * kHalf];
  __shared__ uint32_t so[256 / kHalf][kHalf];
  const int hl = threadIdx.x % kHalf, hid = threadIdx.x / kHalf;
  const int64_t v = ((int64_t)xcd_block_id() * blockDim.x + threadIdx.x) / kHalf;
  int64_t k0 = 0, r0 = 0;
  int deg = 0, len = 0;
  if (v < n && !big[v]) {
    k0 = v2e_ptr[v];
    deg = (int)(v2e_ptr[v + 1] - k0);
    r0 = rowptr[v];
    len = (int)(rowptr[v + 1] - r0);
  }
  if (hl < deg) {
    const int32_t idx = v2e[k0 + hl];
    const int64_t e = idx >> 2;
    const int l = idx & 3;
    const uint32_t pk = slotmap[k0 + hl];
    const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
    const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
    const double C = Ce[e];
    double vv[4][3], r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) vv[i][k] = xyz[3 * (int64_t)id[i] + k];
    p1_row_entries(vv, l, C, r);
#pragma unroll
    for (int m = 0; m < 4; ++m) sv[hid][4 * hl + m] = r[m];
    so[hid][hl] = pk;
  }
  wave_lds_sync();  // sv / so are this row's (half-wave's) own
  const uint32_t me0 = (uint32_t)hl * 0x01010101u, me1 = (uint32_t)(hl + kHalf) * 0x01010101u;
  double acc0 = 0., acc1 = 0.;
  for (int q = 0; q < deg; ++q) {
    const uint32_t pk = so[hid][q];
    const uint32_t t0 = pk ^ me0, z0 = (t0 - 0x01010101u) & ~t0 & 0x80808080u;
    if (z0) acc0 += sv[hid][4 * q + (__builtin_ctz(z0) >> 3)];
    if (len > kHalf) {
      const uint32_t t1 = pk ^ me1, z1 = (t1 - 0x01010101u) & ~t1 & 0x80808080u;
      if (z1) acc1 += sv[hid][4 * q + (__builtin_ctz(z1) >> 3)];
    }
  }
  if (hl < len) vals[r0 + hl] = acc0;
  if (hl + kHalf < len) vals[r0 + hl + kHalf] = acc1;
}

// ---------------------------------------------------------------- Dirichlet
__global__ void k_mark_bc(int64_t nbe, int nbloc, const int32_t *__restrict__ belems,
                          const int32_t *__restrict__ blab, const int32_t *__restrict__ labs,
                          int nlab, uint8_t *__restrict__ onbc) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nbe) return;
  const int32_t lab = blab[f];
  bool on = false;
  for (int k = 0; k < nlab; ++k) on |= (lab == labs[k]);
  if (!on) return;
  for (int q = 0; q < nbloc; ++q) onbc[belems[nbloc * f + q]] = 1;  // same value from every writer
}

__global__ void k_set_diag(int64_t n, const uint8_t *__restrict__ onbc,
                           const int64_t *__restrict__ diag, double tgv,
                           double *__restrict__ vals, int *__restrict__ missing) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !onbc[i]) return;
  const int64_t s = diag[i];
  if (s < 0) { atomicMax(missing, 1); return; }
  vals[s] = tgv;
}

// ---------------------------------------------------------------- right-hand side
// b[v] = sum over the incident elements (element order) of int_K f l_v with
// the element's rule (Element_rhs, problem.cpp:7940-7978); f: CONST f0, SIN3,
// P0 fcoef[e], QP fcoef[e * nq + q], P1 fcoef[vertex] interpolated at the
// point.  Dirichlet rows: tgv * (gvals[v] or g0).
__device__ inline double rhs_coef(int kind, double f0, const double *__restrict__ fcoef, int64_t e,
                                  int q, int nq, double x, double y, double z, int dim) {
  if (kind == FFHIP_COEF_SIN3) {
    if (dim == 3) return kappa_sin3(x, y, z);
    const double tp = 2 * M_PI;
    return 1 + (0.5 * sin(tp * x)) * sin(tp * y);
  }
  if (kind == FFHIP_COEF_QP) return fcoef[e * nq + q];
  if (kind == FFHIP_COEF_P0) return fcoef[e];
  return f0;
}

// |K| of every tet, once (the right-hand side gathers it for each of the 4 incidences)
__global__ void k_tet_mes(int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
                          double *__restrict__ mes) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double v[4][3];
  const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
  mes[e] = tet_mesure(v);
}

// Sum of a row's incidence terms in element order, one row per 32-lane
// half-wave: lane q computes the term of incidence c0 + q of each chunk of 32
// into its LDS slot, and every lane then adds the 32 slots in order (the order
// of the element loop, so the sum rounds as FreeFEM's accumulation into the
// row).  Slots past the row's end hold +0.0, and adding +0.0 leaves the sum
// unchanged: a sum that starts at +0.0 is never -0.0 under round-to-nearest.
template <class F>
__device__ inline double row_sum_in_order(double *__restrict__ slot, int hl, int64_t deg, F term) {
  double s = 0.;
  for (int64_t c0 = 0; c0 < deg; c0 += kHalf) {
    slot[hl] = c0 + hl < deg ? term(c0 + hl) : 0.;
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < kHalf; ++q) s += slot[q];
    wave_lds_sync();
  }
  return s;
}

// b_v = sum over the incident tets of int_K f lambda_v (Tet_5), one row per half-wave
__global__ __launch_bounds__(256) void k_rhs_p1_3d(int64_t n, const double *__restrict__ xyz,
                            const int32_t *__restrict__ tets, const double *__restrict__ tmes,
                            const int64_t *__restrict__ v2e_ptr,
                            const int32_t *__restrict__ v2e, int kind, double f0,
                            const double *__restrict__ fcoef, const uint8_t *__restrict__ onbc,
                            double tgv, double g0, const double *__restrict__ gvals,
                            double *__restrict__ b) {
  const int hl = threadIdx.x % kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  if (v >= n) return;  // uniform in the half-wave
  if (onbc && onbc[v]) {
    if (hl == 0) b[v] = tgv * (gvals ? gvals[v] : g0);
    return;
  }
  __shared__ double slots[256];
  const int64_t k0 = v2e_ptr[v];
  const double s = row_sum_in_order(slots + (threadIdx.x & ~(kHalf - 1)), hl, v2e_ptr[v + 1] - k0,
                                    [&](int64_t k) {
    const int32_t idx = v2e[k0 + k];
    const int64_t e = idx >> 2;
    const int l = idx & 3;
    double vv[4][3], fu[4] = {0., 0., 0., 0.};
    if (kind != FFHIP_COEF_CONST) {  // the points (or the vertex values) are needed for f
      const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
      const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
      if (kind == FFHIP_COEF_P1) {
        for (int i = 0; i < 4; ++i) fu[i] = fcoef[id[i]];
      } else {
        for (int i = 0; i < 4; ++i)
          for (int kk = 0; kk < 3; ++kk) vv[i][kk] = xyz[3 * (int64_t)id[i] + kk];
      }
    }
    const double mes = tmes[e];
    double r = 0.;
    for (int q = 0; q < 14; ++q) {
      const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
      const double l0 = 1. - (px + py + pz);
      const double lam = l == 0 ? l0 : (l == 1 ? px : (l == 2 ? py : pz));
      double f = f0;
      if (kind == FFHIP_COEF_P1) {
        f = p1_value3(fu, px, py, pz);
      } else if (kind != FFHIP_COEF_CONST) {
        double P[3];
        for (int kk = 0; kk < 3; ++kk)
          P[kk] = l0 * vv[0][kk] + px * vv[1][kk] + py * vv[2][kk] + pz * vv[3][kk];
        f = rhs_coef(kind, f0, fcoef, e, q, 14, P[0], P[1], P[2], 3);
      }
      r += (mes * qf5_weight(q)) * f * lam;
    }
    return r;
  });
  if (hl == 0) b[v] = s;
}

__global__ void k_rhs_p1_2d(int64_t n, const double *__restrict__ xy,
                            const int32_t *__restrict__ tris, const int64_t *__restrict__ v2e_ptr,
                            const int32_t *__restrict__ v2e, int kind, double f0,
                            const double *__restrict__ fcoef, const uint8_t *__restrict__ onbc,
                            double tgv, double g0, const double *__restrict__ gvals,
                            double *__restrict__ b) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (onbc && onbc[v]) { b[v] = tgv * (gvals ? gvals[v] : g0); return; }
  double s = 0.;
  for (int64_t k = v2e_ptr[v]; k < v2e_ptr[v + 1]; ++k) {
    const int32_t idx = v2e[k];
    const int64_t e = idx / 3;
    const int l = idx - 3 * (int)e;
    double p[3][2], fu[3] = {0., 0., 0.};
    for (int i = 0; i < 3; ++i)
      for (int kk = 0; kk < 2; ++kk) p[i][kk] = xy[2 * (int64_t)tris[3 * e + i] + kk];
    if (kind == FFHIP_COEF_P1)
      for (int i = 0; i < 3; ++i) fu[i] = fcoef[tris[3 * e + i]];
    const double area = ((p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) -
                         (p[1][1] - p[0][1]) * (p[2][0] - p[0][0])) * 0.5;
    double r = 0.;
    for (int q = 0; q < 7; ++q) {
      const double px = c_qt5_pts[q][0], py = c_qt5_pts[q][1];
      const double l0 = 1 - px - py;
      const double lam = l == 0 ? l0 : (l == 1 ? px : py);
      double f = f0;
      if (kind == FFHIP_COEF_P1) {
        f = p1_value2(fu, px, py);
      } else if (kind != FFHIP_COEF_CONST) {
        const double X = l0 * p[0][0] + px * p[1][0] + py * p[2][0];
        const double Y = l0 * p[0][1] + px * p[1][1] + py * p[2][1];
        f = rhs_coef(kind, f0, fcoef, e, q, 7, X, Y, 0., 2);
      }
      r += (area * c_qt5_w[q]) * f * lam;
    }
    s += r;
  }
  b[v] = s;
}

// ---------------------------------------------------------------- adjoint blocks
// per element: misfit 1/2 int_K e^2 = |K|/40 (sum e_i^2 + (sum e_i)^2), e = u - ud
__global__ __launch_bounds__(256) void k_p1_misfit_elem(
    int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
    const double *__restrict__ u, const double *__restrict__ ud, double *__restrict__ part) {
  __shared__ double lds[256 / kWave];
  double acc = 0.;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nt; e += (int64_t)gridDim.x * 256) {
    const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
    const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
    double v[4][3], ev[4];
    for (int i = 0; i < 4; ++i) {
      for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
      ev[i] = u[id[i]] - ud[id[i]];
    }
    const double mes = tet_mesure(v);
    const double s1 = ev[0] + ev[1] + ev[2] + ev[3];
    const double s2 = ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2] + ev[3] * ev[3];
    acc += mes / 40. * (s2 + s1 * s1);
  }
  const double tot = block_sum<256>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// (M e)_v = sum over incident elements of |K|/20 (e_v + sum_K e_j); 0 on Dirichlet rows
__global__ __launch_bounds__(256) void k_p1_mass_apply(
    int64_t n, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
    const int64_t *__restrict__ v2e_ptr, const int32_t *__restrict__ v2e,
    const double *__restrict__ u, const double *__restrict__ ud, const uint8_t *__restrict__ onbc,
    double *__restrict__ r) {
  const int hl = threadIdx.x % kHalf;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHalf;
  if (v >= n) return;  // uniform in the half-wave
  if (onbc && onbc[v]) {
    if (hl == 0) r[v] = 0.;
    return;
  }
  const double ev = u[v] - ud[v];
  __shared__ double slots[256];
  const int64_t k0 = v2e_ptr[v];
  const double s = row_sum_in_order(slots + (threadIdx.x & ~(kHalf - 1)), hl, v2e_ptr[v + 1] - k0,
                                    [&](int64_t k) {
    const int64_t e = v2e[k0 + k] >> 2;
    const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
    const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
    double vv[4][3], se = 0.;
    for (int i = 0; i < 4; ++i) {
      for (int kk = 0; kk < 3; ++kk) vv[i][kk] = xyz[3 * (int64_t)id[i] + kk];
      se += u[id[i]] - ud[id[i]];
    }
    return tet_mesure(vv) / 20. * (ev + se);
  });
  if (hl == 0) r[v] = s;
}

// g_e = -|K| grad u_K . grad lambda_K
__global__ __launch_bounds__(256) void k_p1_grad_kappa(
    int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
    const double *__restrict__ u, const double *__restrict__ lam, double *__restrict__ g) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
  double v[4][3], G[4][3];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
  const double mes = tet_geom(v, G);
  double gu[3] = {0., 0., 0.}, gl[3] = {0., 0., 0.};
  for (int i = 0; i < 4; ++i) {
    const double ui = u[id[i]], li = lam[id[i]];
    for (int k = 0; k < 3; ++k) { gu[k] += ui * G[i][k]; gl[k] += li * G[i][k]; }
  }
  g[e] = -mes * (gu[0] * gl[0] + gu[1] * gl[1] + gu[2] * gl[2]);
}

// ---------------------------------------------------------------- CSR mat-vec
__global__ void k_spmv_csr(int64_t n, const int64_t *__restrict__ rowptr,
                           const int32_t *__restrict__ colidx, const double *__restrict__ vals,
                           const int64_t *__restrict__ tmap, const double *__restrict__ x,
                           double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    s += (tmap ? vals[tmap[k]] : vals[k]) * x[colidx[k]];
  y[i] = s;
}

// slot of the transposed entry: (i,j) at k  ->  (j,i); -1 if absent
__global__ void k_transpose_map(int64_t n, const int64_t *__restrict__ rowptr,
                                const int32_t *__restrict__ colidx, int64_t *__restrict__ tmap,
                                int *__restrict__ missing) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
    const int32_t j = colidx[k];
    int64_t lo = rowptr[j], hi = rowptr[j + 1] - 1, f = -1;
    while (lo <= hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (colidx[mid] == i) { f = mid; break; }
      if (colidx[mid] < i) lo = mid + 1; else hi = mid - 1;
    }
    tmap[k] = f;
    if (f < 0) atomicMax(missing, 1);
  }
}

static unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
int read_flag(ffhip_ctx ctx, int *d);
void half_view(ffhip_mat L);

// once per device (__constant__ memory is per device): the context's device,
// the caller's current device restored afterwards
static void upload_t5_once(int dev) {
  static bool done[64] = {};
  FFHIP_REQUIRE(dev >= 0 && dev < 64, FFHIP_ERR_UNSUPPORTED, "device id %d", dev);
  if (done[dev]) return;
  int cur = 0;
  FFHIP_CHECK(hipGetDevice(&cur));
  if (cur != dev) FFHIP_CHECK(hipSetDevice(dev));
  const double sqrt15 = 3.87298334620741688517926539978;
  const double t = 1.E0 / 3.E0, A = 0.225E0;
  const double r = (6 - sqrt15) / 21, s = (9 + 2 * sqrt15) / 21, B = (155 - sqrt15) / 1200;
  const double u = (6 + sqrt15) / 21, v = (9 - 2 * sqrt15) / 21, C = (155 + sqrt15) / 1200;
  const double P[7][2] = {{t, t}, {r, r}, {r, s}, {s, r}, {u, u}, {u, v}, {v, u}};
  const double W[7] = {A, B, B, B, C, C, C};
  FFHIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_qt5_pts), P, sizeof(P)));
  FFHIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_qt5_w), W, sizeof(W)));
  if (cur != dev) FFHIP_CHECK(hipSetDevice(cur));
  done[dev] = true;
}

// dof -> (element * nloc + local) incidence lists in element order, for any
// element-to-dof table (P1 vertices, P2 nodes)
// The lists when some exceeds its 32-entry bucket: an atomic fill at the
// list offsets, then the lists of up to 32 entries sorted in a half-wave and
// the longer ones per dof (temporaries from the block cache: callable while
// an arena is held).
void v2e_fill_sort(ffhip_ctx ctx, const int32_t *elems, int64_t M, int64_t n, const int64_t *ptr,
                   int32_t *list, int32_t *inv) {
  hipStream_t s = ctx->stream;
  int32_t *cur = dalloc<int32_t>(n);
  int *flag = dalloc<int>(1);
  FFHIP_CHECK(hipMemsetAsync(cur, 0, 4 * n, s));
  FFHIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  k_v2e_fill<<<nblocks(M, 256), 256, 0, s>>>(elems, M, ptr, cur, list);
  k_v2e_sort_half<<<nblocks(n * kHalf, 256), 256, 0, s>>>(n, ptr, list, flag);
  k_v2e_sort<<<nblocks(n, 256), 256, 0, s>>>(n, ptr, list);
  if (inv) k_v2e_inverse<<<nblocks(M, 256), 256, 0, s>>>(M, list, inv);
  FFHIP_CHECK(hipGetLastError());
  dfree(cur);
  dfree(flag);
}

void build_incidence(ffhip_ctx ctx, const int32_t *elems, int64_t nt, int nloc, int64_t n,
                     int64_t **ptr_out, int32_t **list_out, int *big_out, int32_t **inv_out,
                     int32_t **bucket_out, const ScaleArgs *sc, int *ovf_dev) {
  hipStream_t s = ctx->stream;
  const int64_t M = nt * nloc;
  Arena ar(ctx, Arena::bytes(n, 4) + Arena::bytes(n + 1, 8) + (bucket_out ? 0 : Arena::bytes(32 * n, 4)) +
                    Arena::bytes(1, 4) + Arena::bytes(n, 4));
  uint32_t *deg = ar.take<uint32_t>(n);
  int64_t *deg64 = ar.take<int64_t>(n + 1);
  int32_t *bucket = bucket_out ? dalloc<int32_t>(32 * n) : ar.take<int32_t>(32 * n);
  if (bucket_out) *bucket_out = nullptr;
  // with ovf_dev (and bucket_out) the overflow flag is the caller's, read by
  // the caller later: nothing here waits for the device
  int *flag = ovf_dev && bucket_out ? ovf_dev : ar.take<int>(1);
  int64_t *ptr = dalloc<int64_t>(n + 1);
  int32_t *list = dalloc<int32_t>(M);
  int32_t *inv = inv_out ? dalloc<int32_t>(M) : nullptr;
  FFHIP_CHECK(hipMemsetAsync(deg, 0, 4 * n, s));
  FFHIP_CHECK(hipMemsetAsync(deg64 + n, 0, 8, s));
  FFHIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  if (sc && nloc == 4)  // the element scales computed in the same pass
    k_v2e_bucket_scale<<<nblocks(nt, 256), 256, 0, s>>>(elems, nt, *sc, deg, bucket, flag);
  else
    k_v2e_bucket_block<<<nblocks(M, 256 * kBucketIpt), 256, 0, s>>>(elems, M, deg, bucket, flag);
  k_u32_to_i64<<<nblocks(n, 256), 256, 0, s>>>(n, deg, deg64);
  exclusive_scan_i64(ctx, deg64, ptr, n + 1);
  FFHIP_CHECK(hipGetLastError());
  if (ovf_dev && bucket_out) {  // the caller checks the flag and sorts the buckets
    *bucket_out = bucket;
    if (big_out) *big_out = 0;
    if (inv_out) *inv_out = inv;
    *ptr_out = ptr;
    *list_out = list;
    return;
  }
  const int has_big = read_flag(ctx, flag);
  if (!has_big && bucket_out) {
    *bucket_out = bucket;  // the caller sorts the buckets into the lists
  } else if (!has_big) {
    k_v2e_sort_bucket<<<nblocks(n * kHalf, 256), 256, 0, s>>>(n, ptr, bucket, list, inv);
  } else {  // some list exceeds its bucket: fill by atomics, then sort
    if (bucket_out) dfree(bucket);
    v2e_fill_sort(ctx, elems, M, n, ptr, list, inv);
  }
  FFHIP_CHECK(hipGetLastError());
  FFHIP_CHECK(hipStreamSynchronize(s));
  if (big_out) *big_out = has_big;
  if (inv_out) *inv_out = inv;
  *ptr_out = ptr;
  *list_out = list;
}

// vertex -> incidence lists of a mesh (built once per mesh), and the position
// of every (element, local vertex) item in them
void build_v2e(ffhip_mesh m) {
  if (m->v2e_ptr) return;
  build_incidence(m->ctx, m->elems, m->nt, m->nloc, m->nv, &m->v2e_ptr, &m->v2e, &m->v2e_big,
                  nullptr, nullptr, nullptr);
}

// As build_v2e, but where every list fits its 32-entry bucket the buckets are
// returned unsorted and m->v2e is left unfilled: the caller must sort them
// into m->v2e (k_pattern_half with the buckets does, k_v2e_sort_bucket
// otherwise) and free the buckets before anything else reads m->v2e.
// With sc (a tetrahedral mesh), the incidence build also computes the
// element scales (*scaled = true); when the lists exist already it does not.
// ovf (device int): raised when some list exceeds its bucket; the caller
// reads it and then builds the lists with v2e_fill_sort instead.
int32_t *build_v2e_unsorted(ffhip_mesh m, const ScaleArgs *sc, bool *scaled, int *ovf) {
  *scaled = false;
  if (m->v2e_ptr) return nullptr;
  int32_t *bucket = nullptr;
  build_incidence(m->ctx, m->elems, m->nt, m->nloc, m->nv, &m->v2e_ptr, &m->v2e, &m->v2e_big,
                  nullptr, &bucket, sc, ovf);
  *scaled = sc && m->nloc == 4;
  return bucket;
}

// the position of every (element, local vertex) item in the incidence lists
// (built on first use: only the element-row assembly path scatters by it)
void build_v2e_inv(ffhip_mesh m) {
  build_v2e(m);
  if (m->v2e_inv) return;
  const int64_t M = m->nt * m->nloc;
  m->v2e_inv = dalloc<int32_t>(M);
  k_v2e_inverse<<<nblocks(M, 256), 256, 0, m->ctx->stream>>>(M, m->v2e, m->v2e_inv);
  FFHIP_CHECK(hipGetLastError());
}

// a_ii = tgv on the marked rows (HashMatrix::SetBC); *missing set when a
// marked row has no diagonal slot
void set_diag_rows(ffhip_ctx ctx, int64_t n, const uint8_t *onbc, const int64_t *diag, double tgv,
                   double *vals, int *missing) {
  k_set_diag<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, onbc, diag, tgv, vals, missing);
  FFHIP_CHECK(hipGetLastError());
}

// tmap[k] = slot of the transposed entry; false if the pattern is not symmetric
bool transpose_map(ffhip_ctx ctx, ffhip_mat A, int64_t *tmap) {
  int *missing = dalloc<int>(1);
  FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), ctx->stream));
  k_transpose_map<<<nblocks(A->n, 256), 256, 0, ctx->stream>>>(A->n, A->rowptr, A->colidx, tmap,
                                                                missing);
  FFHIP_CHECK(hipGetLastError());
  const int miss = read_flag(ctx, missing);
  dfree(missing);
  return miss == 0;
}

int read_flag(ffhip_ctx ctx, int *d) {
  int h = 0;
  FFHIP_CHECK(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FFHIP_CHECK(hipStreamSynchronize(ctx->stream));
  return h;
}

// ---------------------------------------------------------------- misfit
// J = 1/2 int (u - ud)^2 (partials of kMisfitBlocks blocks, summed on the host
// in block order) and r = M (u - ud) with the Dirichlet rows zeroed, on the
// given buffers: part (kMisfitBlocks doubles, when J), onbc (nv bytes) and
// labs (nlab int32, when r and nlab > 0)
constexpr int kMisfitBlocks = 1024;
struct MisfitBuffers {
  double *part = nullptr;
  uint8_t *onbc = nullptr;
  int32_t *labs = nullptr;
};

__global__ void k_fill_u8(int64_t n, uint8_t *__restrict__ a, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

void misfit_run(ffhip_ctx ctx, ffhip_mesh m, const double *u, const double *ud,
                const int32_t *labels, int nlab, double *J, double *r, const MisfitBuffers &b) {
  hipStream_t s = ctx->stream;
  ctx->tic();
  if (J) {
    k_p1_misfit_elem<<<kMisfitBlocks, 256, 0, s>>>(m->nt, m->xyz, m->elems, u, ud, b.part);
    FFHIP_CHECK(hipGetLastError());
    double h[kMisfitBlocks];
    FFHIP_CHECK(hipMemcpyAsync(h, b.part, sizeof(h), hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    double tot = 0.;
    for (int i = 0; i < kMisfitBlocks; ++i) tot += h[i];
    *J = tot;
  }
  if (r) {
    const uint8_t *onbc = nullptr;
    if (nlab > 0) {
      FFHIP_CHECK(hipMemcpyAsync(b.labs, labels, 4 * nlab, hipMemcpyHostToDevice, s));
      k_fill_u8<<<nblocks(m->nv, 256), 256, 0, s>>>(m->nv, b.onbc, 0);
      if (m->nbe)
        k_mark_bc<<<nblocks(m->nbe, 256), 256, 0, s>>>(m->nbe, m->nbloc, m->belems, m->blab, b.labs,
                                                       nlab, b.onbc);
      onbc = b.onbc;
    }
    k_p1_mass_apply<<<nblocks(m->nv * kHalf, 256), 256, 0, s>>>(m->nv, m->xyz, m->elems, m->v2e_ptr, m->v2e,
                                                        u, ud, onbc, r);
    FFHIP_CHECK(hipGetLastError());
  }
  ctx->toc();
}

// ---------------------------------------------------------------- no-allocation scratch
// The scratch of ffhip_p1_misfit_noalloc: a file-scope device array, one copy
// per device, reached through the address hipGetSymbolAddress gives on that
// device (the array's name is a host-side stand-in in host code).  Layout:
// misfit partials | labels | Dirichlet marks.  Every call on a device shares
// it: overlapping calls on one device race.
constexpr int64_t kNoallocMaxVertices = (int64_t)1 << 24;
constexpr int kNoallocMaxLabels = 64;
constexpr size_t kScratchPartOff = 0;
constexpr size_t kScratchLabelsOff = 8 * kMisfitBlocks;
constexpr size_t kScratchOnbcOff = kScratchLabelsOff + 4 * kNoallocMaxLabels;
constexpr size_t kScratchBytes = kScratchOnbcOff + (size_t)kNoallocMaxVertices;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex &scratch_mutex() {
  static std::mutex mu;
  return mu;
}
unsigned char *&scratch_ptr(int device) {
  static unsigned char *ptr[kMaxDevices] = {};
  return ptr[device];
}

// resolves the device address of the scratch array on `device`, which must be
// the current device (called by ffhip_init, once per device)
void noalloc_scratch_setup(int device) {
  FFHIP_REQUIRE(device >= 0 && device < kMaxDevices, FFHIP_ERR_UNSUPPORTED, "device %d", device);
  std::lock_guard<std::mutex> g(scratch_mutex());
  if (scratch_ptr(device)) return;
  void *p = nullptr;
  FFHIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  scratch_ptr(device) = static_cast<unsigned char *>(p);
}

unsigned char *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> g(scratch_mutex());
  unsigned char *p = (device >= 0 && device < kMaxDevices) ? scratch_ptr(device) : nullptr;
  FFHIP_REQUIRE(p, FFHIP_ERR_INTERNAL, "no-allocation scratch not set up on device %d", device);
  return p;
}

// The P1 pattern (ffhip_pattern_p1).  With Ce (3D: the element scales,
// computed from sc by the incidence build, or by k_p1_scale_3d when the
// mesh's lists exist) the half-wave pattern pass also forms the values of
// the rows it builds; *values_done tells whether that covered every row (no
// lane-per-row rows).
//
// First try: the row pointers from the vertex-link counts (k_link_len), the
// pattern pass writing straight to CSR.  If some row is not what they
// predict, or needs the lane-per-row kernels, the buffered passes run.
ffhip_mat pattern_p1(ffhip_ctx ctx, ffhip_mesh m, const double *Ce, bool *values_done,
                     const ScaleArgs *sc) {
  hipStream_t s = ctx->stream;
  const bool vals_pass = Ce && m->nloc == 4;
  // the fused pass sorts the incidence buckets into the lists itself; with
  // sc the incidence build computes the element scales Ce as well (or, on a
  // mesh whose lists exist, k_p1_scale_3d does)
  int32_t *bucket = nullptr;
  int *ovf = nullptr;  // bucket overflow, read back with the row pointers
  if (vals_pass) {
    bool scaled = false;
    ovf = dalloc<int>(1);
    FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
    bucket = build_v2e_unsorted(m, sc, &scaled, ovf);
    if (sc && !scaled) {
      k_p1_scale_3d<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, sc->xyz, m->elems, sc->kind, sc->c0,
                                                        sc->kappa_e, sc->Ce);
      FFHIP_CHECK(hipGetLastError());
    }
  } else {
    build_v2e(m);
  }
  const int64_t n = m->nv;
  ffhip_mat A = new ffhip_mat_s();
  A->ctx = ctx; A->n = n; A->mesh = m;
  // the lists sorted from the buckets by k_v2e_sort_bucket (when the fused
  // pass did not run, or not for every row)
  auto sort_buckets = [&] {
    if (!bucket) return;
    k_v2e_sort_bucket<<<nblocks(n * kHalf, 256), 256, 0, s>>>(n, m->v2e_ptr, bucket, m->v2e, nullptr);
    FFHIP_CHECK(hipGetLastError());
    dfree(bucket);
  };
  try {
    A->rowptr = dalloc<int64_t>(n + 1);
    A->diag = dalloc<int64_t>(n);
    A->slotmap = dalloc<uint32_t>(m->nt * m->nloc);
    A->big = dalloc<uint8_t>(n);
    bool direct_ok = false;
    {  // direct pass
      Arena ar(ctx, Arena::bytes(n, 4) + Arena::bytes(n + 1, 8) + Arena::bytes(2, 4));
      uint32_t *bcnt = ar.take<uint32_t>(n);
      int64_t *len = ar.take<int64_t>(n + 1);
      int *flags = ar.take<int>(2);  // any big row, any row unlike its link count
      FFHIP_CHECK(hipMemsetAsync(bcnt, 0, 4 * n, s));
      FFHIP_CHECK(hipMemsetAsync(len + n, 0, 8, s));
      FFHIP_CHECK(hipMemsetAsync(flags, 0, 2 * sizeof(int), s));
      if (m->nbe)
        k_bcount<<<nblocks(m->nbe * m->nbloc, 256), 256, 0, s>>>(m->nbe * m->nbloc, m->belems, bcnt);
      k_link_len<<<nblocks(n, 256), 256, 0, s>>>(n, m->dim, m->v2e_ptr, bcnt, len);
      FFHIP_CHECK(hipGetLastError());
      exclusive_scan_i64(ctx, len, A->rowptr, n + 1);
      if (bucket) {  // the entry count and the bucket overflow flag in one wait
        int64_t nnz = 0;
        int of = 0;
        FFHIP_CHECK(hipMemcpyAsync(&nnz, A->rowptr + n, sizeof(nnz), hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipMemcpyAsync(&of, ovf, sizeof(of), hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipStreamSynchronize(s));
        A->nnz = nnz;
        if (of) {  // some list exceeds its bucket: the lists by the atomic fill
          dfree(bucket);
          v2e_fill_sort(ctx, m->elems, m->nt * m->nloc, n, m->v2e_ptr, m->v2e, nullptr);
          m->v2e_big = 1;
        }
      } else {
        A->nnz = read_i64(ctx, A->rowptr + n);
      }
      A->colidx = dalloc<int32_t>(A->nnz);
      A->vals = dalloc<double>(A->nnz);
      PatternOut out;
      out.rowptr = A->rowptr; out.colidx = A->colidx; out.diag = A->diag; out.vals = A->vals;
      out.mismatch = flags + 1;
      if (vals_pass)
        k_pattern_half<4, true, true><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
            n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, flags, out, m->xyz, Ce, bucket);
      else if (m->nloc == 4)
        k_pattern_half<4, false, true><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
            n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, flags, out);
      else
        k_pattern_half<3, false, true><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
            n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, flags, out);
      FFHIP_CHECK(hipGetLastError());
      int h[2] = {0, 0};
      FFHIP_CHECK(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, s));
      FFHIP_CHECK(hipStreamSynchronize(s));
      direct_ok = !h[0] && !h[1];
      if (!direct_ok) { dfree(A->colidx); dfree(A->vals); }
      if (direct_ok) dfree(bucket);  // every list sorted by the pass
      else sort_buckets();
      dfree(ovf);
    }
    if (direct_ok) {
      if (!vals_pass) FFHIP_CHECK(hipMemsetAsync(A->vals, 0, 8 * A->nnz, s));
      A->any_big = 0;
      A->max_row = kRowBuf;
      if (values_done) *values_done = vals_pass;
      return A;
    }
    // buffered passes
    Arena ar(ctx, Arena::bytes(n + 1, 8) + Arena::bytes(1, 4) + Arena::bytes(kRowBuf * n, 4) +
                      Arena::bytes(n, 1) + Arena::bytes(1, 4) +
                      (vals_pass ? Arena::bytes(kRowBuf * n, 8) : 0));
    int64_t *len = ar.take<int64_t>(n + 1);
    int *ovf = ar.take<int>(1);
    FFHIP_CHECK(hipMemsetAsync(len, 0, 8 * (n + 1), s));
    FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
    // rows with up to 32 incident elements and 64 columns: one half-wave
    // pass into row buffers (big = 0); the rest: the lane-per-row merge with
    // 64 LDS columns per lane (big = 1), then with 256 (big = 2: the slot
    // offsets of the element-to-row map are bytes)
    constexpr int B = 128, CAP = 64, BW = 32, CAPW = 256;
    PatternOut out;
    out.rowlen = len;
    out.rowbuf = ar.take<int32_t>(kRowBuf * n);
    out.dposr = ar.take<int8_t>(n);
    int *anyb = ar.take<int>(1);
    out.rowval = vals_pass ? ar.take<double>(kRowBuf * n) : nullptr;
    FFHIP_CHECK(hipMemsetAsync(anyb, 0, sizeof(int), s));
    if (vals_pass)
      k_pattern_half<4, true, false><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
          n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, anyb, out, m->xyz, Ce);
    else if (m->nloc == 4)
      k_pattern_half<4, false, false><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
          n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, anyb, out);
    else
      k_pattern_half<3, false, false><<<nblocks(n * kHalf, 256), 256, 0, s>>>(
          n, m->elems, m->v2e_ptr, m->v2e, A->slotmap, A->big, anyb, out);
    FFHIP_CHECK(hipGetLastError());
    A->any_big = read_flag(ctx, anyb);
    bool wide = false;
    if (A->any_big) {
      k_pattern_rows<B, CAP, false><<<nblocks(n, B), B, 0, s>>>(
          n, m->nloc, m->elems, m->v2e_ptr, m->v2e, len, nullptr, nullptr, nullptr, nullptr,
          nullptr, ovf, A->big, 1);
      FFHIP_CHECK(hipGetLastError());
      wide = read_flag(ctx, ovf) != 0;
    }
    if (wide) {
      FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
      k_pattern_rows<BW, CAPW, false><<<nblocks(n, BW), BW, 0, s>>>(
          n, m->nloc, m->elems, m->v2e_ptr, m->v2e, len, nullptr, nullptr, nullptr, nullptr,
          nullptr, ovf, A->big, 2);
      FFHIP_CHECK(hipGetLastError());
      FFHIP_REQUIRE(read_flag(ctx, ovf) == 0, FFHIP_ERR_UNSUPPORTED,
                    "ffhip_pattern_p1: a row has more than %d columns", CAPW);
    }
    exclusive_scan_i64(ctx, len, A->rowptr, n + 1);
    A->nnz = read_i64(ctx, A->rowptr + n);
    A->colidx = dalloc<int32_t>(A->nnz);
    A->vals = dalloc<double>(A->nnz);
    k_pattern_compact<<<nblocks(n, 256), 256, 0, s>>>(n, A->rowptr, out.rowbuf, out.dposr, A->big,
                                                      A->colidx, A->diag, out.rowval,
                                                      vals_pass ? A->vals : nullptr);
    if (A->any_big)
      k_pattern_rows<B, CAP, true><<<nblocks(n, B), B, 0, s>>>(
          n, m->nloc, m->elems, m->v2e_ptr, m->v2e, nullptr, A->rowptr, A->colidx, A->diag,
          A->slotmap, nullptr, ovf, A->big, 1);
    if (wide)
      k_pattern_rows<BW, CAPW, true><<<nblocks(n, BW), BW, 0, s>>>(
          n, m->nloc, m->elems, m->v2e_ptr, m->v2e, nullptr, A->rowptr, A->colidx, A->diag,
          A->slotmap, nullptr, ovf, A->big, 2);
    FFHIP_CHECK(hipGetLastError());
    if (!vals_pass) FFHIP_CHECK(hipMemsetAsync(A->vals, 0, 8 * A->nnz, s));
    if (values_done) *values_done = vals_pass && !A->any_big;
    A->max_row = wide ? CAPW : CAP;
  } catch (...) {
    (void)hipStreamSynchronize(s);
    dfree(ovf);
    if (bucket) {  // the lists were never sorted: the mesh builds them again when next needed
      dfree(bucket);
      dfree(m->v2e_ptr);
      dfree(m->v2e);
    }
    dfree(A->rowptr); dfree(A->colidx); dfree(A->vals); dfree(A->diag);
    dfree(A->slotmap); dfree(A->big);
    delete A;
    throw;
  }
  return A;
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_pattern_p1(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && out, FFHIP_ERR_ARG, "ffhip_pattern_p1: null argument");
    FFHIP_REQUIRE(m->nloc <= 4, FFHIP_ERR_UNSUPPORTED, "ffhip_pattern_p1: P1 on simplices only");
    ctx->tic();
    ffhip_mat A = pattern_p1(ctx, m, nullptr, nullptr, nullptr);
    ctx->toc();
    *out = A;
  });
}

int ffhip_pattern_assemble_p1(ffhip_ctx ctx, ffhip_mesh m, int kind, double c0, const double *kappa_e,
                              int kappa_on_device, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && out, FFHIP_ERR_ARG, "ffhip_pattern_assemble_p1: null argument");
    FFHIP_REQUIRE(m->dim == 3 && m->nloc == 4, FFHIP_ERR_UNSUPPORTED,
                  "ffhip_pattern_assemble_p1: 3D P1 meshes (2D: ffhip_pattern_p1 + ffhip_assemble_p1)");
    FFHIP_REQUIRE(kind == FFHIP_COEF_CONST || kind == FFHIP_COEF_SIN3 || kind == FFHIP_COEF_P0 ||
                      kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1,
                  FFHIP_ERR_ARG, "ffhip_pattern_assemble_p1: coefficient kind %d", kind);
    const bool arr = kind == FFHIP_COEF_P0 || kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1;
    FFHIP_REQUIRE(!arr || kappa_e, FFHIP_ERR_ARG, "ffhip_pattern_assemble_p1: kappa_e");
    hipStream_t s = ctx->stream;
    const double *kd = kappa_e;
    double *kbuf = nullptr, *Ce = nullptr;
    ffhip_mat A = nullptr;
    try {
      if (arr && !kappa_on_device) {
        const int64_t cnt = kind == FFHIP_COEF_P1 ? m->nv : kind == FFHIP_COEF_P0 ? m->nt : 14 * m->nt;
        kbuf = dalloc<double>(cnt);
        FFHIP_CHECK(hipMemcpyAsync(kbuf, kappa_e, 8 * cnt, hipMemcpyHostToDevice, s));
        kd = kbuf;
      }
      Ce = dalloc<double>(m->nt);
      ctx->tic();
      // the element scales are computed by the incidence build (one pass:
      // the coefficient at 14 points per tet hides the atomics' round trips),
      // then the pattern pass forms the row values from them
      const ScaleArgs sc{m->xyz, kind, c0, kd, Ce};
      bool done = false;
      A = pattern_p1(ctx, m, Ce, &done, &sc);
      ctx->toc();
      if (!done) {  // rows for the lane-per-row kernels: all values by the element-row path
        const int rc = ffhip_assemble_p1(ctx, m, A, kind, c0, kd, 1);
        FFHIP_REQUIRE(rc == FFHIP_OK, rc, "ffhip_pattern_assemble_p1: %s", ffhip_last_error());
      }
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dfree(Ce); dfree(kbuf);
      if (A) ffhip_mat_free(A);
      throw;
    }
    dfree(Ce); dfree(kbuf);
    *out = A;
  });
}

int ffhip_mat_info(ffhip_mat A, int64_t *n, int64_t *nnz) {
  return guard([&] {
    FFHIP_REQUIRE(A, FFHIP_ERR_ARG, "ffhip_mat_info: null matrix");
    if (n) *n = A->n;
    if (nnz) *nnz = A->nnz;
  });
}

int ffhip_assemble_p1(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat A, int kind, double c0,
                      const double *kappa_e, int kappa_on_device) {
  return ffhip_assemble_p1_form(ctx, m, A, kind, c0, kappa_e, FFHIP_COEF_NONE, 0., nullptr,
                                kappa_on_device);
}

int ffhip_assemble_p1_form(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat A, int kind, double c0,
                           const double *kappa_e, int mkind, double m0, const double *mass_e,
                           int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && A && A->mesh == m && A->slotmap, FFHIP_ERR_ARG,
                  "ffhip_assemble_p1: matrix pattern was not built from this mesh");
    auto valid = [](int k) {
      return k == FFHIP_COEF_CONST || k == FFHIP_COEF_SIN3 || k == FFHIP_COEF_P0 ||
             k == FFHIP_COEF_QP || k == FFHIP_COEF_P1;
    };
    FFHIP_REQUIRE(valid(kind), FFHIP_ERR_ARG, "ffhip_assemble_p1: coefficient kind %d", kind);
    FFHIP_REQUIRE(mkind == FFHIP_COEF_NONE || valid(mkind), FFHIP_ERR_ARG,
                  "ffhip_assemble_p1_form: mass coefficient kind %d", mkind);
    auto is_array = [](int k) {
      return k == FFHIP_COEF_P0 || k == FFHIP_COEF_QP || k == FFHIP_COEF_P1;
    };
    const bool arr = is_array(kind), marr = is_array(mkind);
    FFHIP_REQUIRE(!arr || kappa_e, FFHIP_ERR_ARG, "ffhip_assemble_p1: kappa_e");
    FFHIP_REQUIRE(!marr || mass_e, FFHIP_ERR_ARG, "ffhip_assemble_p1_form: mass_e");
    const int kappa_on_device = on_device;
    hipStream_t s = ctx->stream;
    const int nu = m->nloc == 4 ? 10 : 6;
    // 3D stiffness without a reaction term and with every row in the
    // half-wave kernels: element scales, then rows formed from the vertices.
    // Otherwise the element rows in incidence order (3D), and Ke for the
    // lane-per-row gather of long rows and 2D.
    const bool rows_path = m->dim == 3 && mkind == FFHIP_COEF_NONE && !A->any_big;
    const bool needKe = m->nloc != 4 || A->any_big;
    Arena ar(ctx, (rows_path ? Arena::bytes(m->nt, 8) : 0) +
                      (!rows_path && m->dim == 3 ? Arena::bytes(16 * m->nt, 8) : 0) +
                      (needKe ? Arena::bytes(nu * m->nt, 8) : 0));
    double *Ke = needKe ? ar.take<double>(nu * m->nt) : nullptr;
    const double *kd = kappa_e, *md = mass_e;
    double *kbuf = nullptr, *mbuf = nullptr;
    auto count = [&](int k) {
      return k == FFHIP_COEF_P1 ? m->nv : k == FFHIP_COEF_P0 ? m->nt : m->nt * (m->dim == 3 ? 14 : 7);
    };
    if (arr && !kappa_on_device) {
      const int64_t cnt = count(kind);
      kbuf = dalloc<double>(cnt);
      FFHIP_CHECK(hipMemcpyAsync(kbuf, kappa_e, 8 * cnt, hipMemcpyHostToDevice, s));
      kd = kbuf;
    }
    if (marr && !on_device) {
      const int64_t cnt = count(mkind);
      mbuf = dalloc<double>(cnt);
      FFHIP_CHECK(hipMemcpyAsync(mbuf, mass_e, 8 * cnt, hipMemcpyHostToDevice, s));
      md = mbuf;
    }
    if (m->dim == 2) upload_t5_once(ctx->device);
    if (!rows_path && m->dim == 3) build_v2e_inv(m);
    ctx->tic();
    double *Kinc = nullptr;
    if (rows_path) {
      double *Ce = ar.take<double>(m->nt);
      k_p1_scale_3d<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, m->xyz, m->elems, kind, c0, kd, Ce);
      const int64_t g = (nblocks(A->n * kHalf, 256) + 7) / 8 * 8;  // a multiple of the 8 XCDs
      k_p1_values_rows<<<(unsigned)g, 256, 0, s>>>(A->n, A->rowptr, m->v2e_ptr, m->v2e, A->slotmap,
                                                   m->xyz, m->elems, Ce, A->big, A->vals);
    } else if (m->dim == 3) {
      Kinc = ar.take<double>(16 * m->nt);
      if (mkind == FFHIP_COEF_NONE)
        k_p1_element_3d<false><<<nblocks(m->nt, 256), 256, 0, s>>>(
            m->nt, m->xyz, m->elems, kind, c0, kd, mkind, m0, md, m->v2e_inv, Kinc,
            A->any_big ? Ke : nullptr);
      else
        k_p1_element_3d<true><<<nblocks(m->nt, 256), 256, 0, s>>>(
            m->nt, m->xyz, m->elems, kind, c0, kd, mkind, m0, md, m->v2e_inv, Kinc,
            A->any_big ? Ke : nullptr);
    } else
      k_p1_element_2d<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, m->xyz, m->elems, kind, c0, kd,
                                                           mkind, m0, md, Ke);
    constexpr int B = 256, STAGE = 6144;
    if (!rows_path && m->nloc == 4) {
      k_p1_values_half<<<nblocks(A->n * kHalf, 256), 256, 0, s>>>(
          A->n, A->rowptr, m->v2e_ptr, A->slotmap, Kinc, A->big, A->vals);
      if (A->any_big)
        k_p1_gather<B, STAGE, 4><<<nblocks(A->n, B), B, 0, s>>>(
            A->n, A->rowptr, m->v2e_ptr, m->v2e, A->slotmap, Ke, A->vals, A->big);
    } else if (m->nloc == 3) {
      k_p1_gather<B, STAGE, 3><<<nblocks(A->n, B), B, 0, s>>>(A->n, A->rowptr, m->v2e_ptr, m->v2e,
                                                               A->slotmap, Ke, A->vals);
    }
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    dfree(kbuf);
    dfree(mbuf);
  });
}

int ffhip_set_bc(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat A, const int32_t *labels, int nlab,
                 double tgv, uint8_t *onbc_out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && A && A->n == m->nv, FFHIP_ERR_ARG, "ffhip_set_bc: bad arguments");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_set_bc: labels");
    FFHIP_REQUIRE(tgv > 0, FFHIP_ERR_UNSUPPORTED, "ffhip_set_bc: only tgv > 0 is supported");
    hipStream_t s = ctx->stream;
    uint8_t *onbc = dalloc<uint8_t>(A->n);
    int32_t *labs = dalloc<int32_t>(nlab > 0 ? nlab : 1);
    int *missing = dalloc<int>(1);
    if (nlab) FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * nlab, hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemsetAsync(onbc, 0, A->n, s));
    FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), s));
    ctx->tic();
    if (m->nbe)
      k_mark_bc<<<nblocks(m->nbe, 256), 256, 0, s>>>(m->nbe, m->nbloc, m->belems, m->blab, labs,
                                                     nlab, onbc);
    k_set_diag<<<nblocks(A->n, 256), 256, 0, s>>>(A->n, onbc, A->diag, tgv, A->vals, missing);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    if (onbc_out) FFHIP_CHECK(hipMemcpy(onbc_out, onbc, A->n, hipMemcpyDeviceToHost));
    const int miss = read_flag(ctx, missing);
    dfree(onbc); dfree(labs); dfree(missing);
    FFHIP_REQUIRE(miss == 0, FFHIP_ERR_INTERNAL, "ffhip_set_bc: a Dirichlet row has no diagonal");
  });
}

int ffhip_rhs_p1_form(ffhip_ctx ctx, ffhip_mesh m, int kind, double f0, const double *fcoef,
                      const int32_t *labels, int nlab, double tgv, double g0, const double *gvals,
                      double *b, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && b, FFHIP_ERR_ARG, "ffhip_rhs_p1: bad arguments");
    FFHIP_REQUIRE(kind == FFHIP_COEF_CONST || kind == FFHIP_COEF_SIN3 || kind == FFHIP_COEF_P0 ||
                      kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1,
                  FFHIP_ERR_ARG, "ffhip_rhs_p1_form: coefficient kind");
    const bool arr = kind == FFHIP_COEF_P0 || kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1;
    FFHIP_REQUIRE(!arr || fcoef, FFHIP_ERR_ARG, "ffhip_rhs_p1_form: coefficient array missing");
    hipStream_t s = ctx->stream;
    build_v2e(m);
    const int64_t n = m->nv;
    uint8_t *onbc = nullptr;
    int32_t *labs = nullptr;
    if (nlab > 0) {
      onbc = dalloc<uint8_t>(n);
      labs = dalloc<int32_t>(nlab);
      FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * nlab, hipMemcpyHostToDevice, s));
      FFHIP_CHECK(hipMemsetAsync(onbc, 0, n, s));
      if (m->nbe)
        k_mark_bc<<<nblocks(m->nbe, 256), 256, 0, s>>>(m->nbe, m->nbloc, m->belems, m->blab, labs,
                                                       nlab, onbc);
    }
    double *bd = on_device ? b : dalloc<double>(n);
    const double *fd = fcoef, *gd = gvals;
    double *fown = nullptr, *gown = nullptr;
    if (arr && !on_device) {
      const int64_t cnt = kind == FFHIP_COEF_P1   ? m->nv
                          : kind == FFHIP_COEF_P0 ? m->nt
                                                  : m->nt * (m->dim == 3 ? 14 : 7);
      fown = dalloc<double>(cnt);
      FFHIP_CHECK(hipMemcpyAsync(fown, fcoef, 8 * cnt, hipMemcpyHostToDevice, s));
      fd = fown;
    }
    if (gvals && !on_device) {
      gown = dalloc<double>(n);
      FFHIP_CHECK(hipMemcpyAsync(gown, gvals, 8 * n, hipMemcpyHostToDevice, s));
      gd = gown;
    }
    if (m->dim == 2) upload_t5_once(ctx->device);
    double *tmes = m->dim == 3 ? dalloc<double>(m->nt) : nullptr;
    ctx->tic();
    if (m->dim == 3) {
      k_tet_mes<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, m->xyz, m->elems, tmes);
      k_rhs_p1_3d<<<nblocks(n * kHalf, 256), 256, 0, s>>>(n, m->xyz, m->elems, tmes, m->v2e_ptr, m->v2e, kind,
                                                   f0, fd, onbc, tgv, g0, gd, bd);
    }
    else
      k_rhs_p1_2d<<<nblocks(n, 256), 256, 0, s>>>(n, m->xyz, m->elems, m->v2e_ptr, m->v2e, kind, f0,
                                                   fd, onbc, tgv, g0, gd, bd);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    if (!on_device) {
      FFHIP_CHECK(hipMemcpy(b, bd, 8 * n, hipMemcpyDeviceToHost));
      dfree(bd);
    }
    FFHIP_CHECK(hipStreamSynchronize(s));
    dfree(onbc);
    dfree(labs);
    dfree(fown);
    dfree(gown);
    dfree(tmes);
  });
}

int ffhip_rhs_p1(ffhip_ctx ctx, ffhip_mesh m, double f, const int32_t *labels, int nlab,
                 double tgv, double g, double *b, int on_device) {
  return ffhip_rhs_p1_form(ctx, m, FFHIP_COEF_CONST, f, nullptr, labels, nlab, tgv, g, nullptr, b,
                           on_device);
}

int ffhip_p1_misfit(ffhip_ctx ctx, ffhip_mesh m, const double *u, const double *ud,
                    const int32_t *labels, int nlab, double *J, double *r) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && u && ud && m->dim == 3, FFHIP_ERR_ARG, "ffhip_p1_misfit: bad arguments");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_p1_misfit: labels");
    build_v2e(m);
    MisfitBuffers b;
    b.part = J ? dalloc<double>(kMisfitBlocks) : nullptr;
    if (r && nlab > 0) {
      b.onbc = dalloc<uint8_t>(m->nv);
      b.labs = dalloc<int32_t>(nlab);
    }
    try {
      misfit_run(ctx, m, u, ud, labels, nlab, J, r, b);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      dfree(b.part); dfree(b.onbc); dfree(b.labs);
      throw;
    }
    dfree(b.part); dfree(b.onbc); dfree(b.labs);
  });
}

int ffhip_p1_misfit_noalloc(ffhip_ctx ctx, ffhip_mesh m, const double *u, const double *ud,
                            const int32_t *labels, int nlab, double *J, double *r) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && u && ud && m->dim == 3, FFHIP_ERR_ARG,
                  "ffhip_p1_misfit_noalloc: bad arguments");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_p1_misfit_noalloc: labels");
    FFHIP_REQUIRE(m->nv <= kNoallocMaxVertices, FFHIP_ERR_ARG,
                  "ffhip_p1_misfit_noalloc: %lld vertices, more than its maximum %lld",
                  (long long)m->nv, (long long)kNoallocMaxVertices);
    FFHIP_REQUIRE(nlab <= kNoallocMaxLabels, FFHIP_ERR_ARG,
                  "ffhip_p1_misfit_noalloc: %d labels, more than its maximum %d", nlab, kNoallocMaxLabels);
    FFHIP_REQUIRE(!r || m->v2e_ptr, FFHIP_ERR_ARG,
                  "ffhip_p1_misfit_noalloc: the mesh's vertex incidence is not built yet (an assembly "
                  "or ffhip_p1_misfit on this mesh builds it; this entry point allocates nothing)");
    unsigned char *base = noalloc_scratch(ctx->device);
    MisfitBuffers b;
    b.part = reinterpret_cast<double *>(base + kScratchPartOff);
    b.labs = reinterpret_cast<int32_t *>(base + kScratchLabelsOff);
    b.onbc = base + kScratchOnbcOff;
    misfit_run(ctx, m, u, ud, labels, nlab, J, r, b);
  });
}

int ffhip_p1_grad_kappa(ffhip_ctx ctx, ffhip_mesh m, const double *u, const double *lambda,
                        double *g) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && u && lambda && g && m->dim == 3, FFHIP_ERR_ARG,
                  "ffhip_p1_grad_kappa: bad arguments");
    ctx->tic();
    k_p1_grad_kappa<<<nblocks(m->nt, 256), 256, 0, ctx->stream>>>(m->nt, m->xyz, m->elems, u, lambda, g);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
  });
}

int ffhip_mat_download(ffhip_mat A, int64_t *rowptr, int32_t *colidx, double *vals) {
  return guard([&] {
    FFHIP_REQUIRE(A, FFHIP_ERR_ARG, "ffhip_mat_download: null matrix");
    hipStream_t s = A->ctx->stream;
    if (rowptr) FFHIP_CHECK(hipMemcpyAsync(rowptr, A->rowptr, 8 * (A->n + 1), hipMemcpyDeviceToHost, s));
    if (colidx) FFHIP_CHECK(hipMemcpyAsync(colidx, A->colidx, 4 * A->nnz, hipMemcpyDeviceToHost, s));
    if (vals) FFHIP_CHECK(hipMemcpyAsync(vals, A->vals, 8 * A->nnz, hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
  });
}

int ffhip_mat_upload(ffhip_ctx ctx, int64_t n, int64_t nnz, const int64_t *rowptr,
                     const int32_t *colidx, const double *vals, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && out && rowptr && colidx && vals && n > 0 && nnz > 0, FFHIP_ERR_ARG,
                  "ffhip_mat_upload: bad arguments");
    FFHIP_REQUIRE(rowptr[0] == 0 && rowptr[n] == nnz, FFHIP_ERR_ARG, "ffhip_mat_upload: rowptr");
    FFHIP_REQUIRE(n < (1ll << 31), FFHIP_ERR_UNSUPPORTED, "ffhip_mat_upload: n too large");
    for (int64_t i = 0; i < n; ++i) {
      FFHIP_REQUIRE(rowptr[i + 1] >= rowptr[i], FFHIP_ERR_ARG, "ffhip_mat_upload: rowptr not monotone");
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        FFHIP_REQUIRE(colidx[k] >= 0 && colidx[k] < n, FFHIP_ERR_ARG, "ffhip_mat_upload: column out of range");
        FFHIP_REQUIRE(k == rowptr[i] || colidx[k] > colidx[k - 1], FFHIP_ERR_ARG,
                      "ffhip_mat_upload: columns must be sorted and unique (HashMatrix::CSR order)");
      }
    }
    hipStream_t s = ctx->stream;
    ffhip_mat A = new ffhip_mat_s();
    A->ctx = ctx; A->n = n; A->nnz = nnz;
    A->rowptr = dalloc<int64_t>(n + 1);
    A->colidx = dalloc<int32_t>(nnz);
    A->vals = dalloc<double>(nnz);
    A->diag = dalloc<int64_t>(n);
    std::vector<int64_t> d(n, -1);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
        if (colidx[k] == i) d[i] = k;
    FFHIP_CHECK(hipMemcpyAsync(A->rowptr, rowptr, 8 * (n + 1), hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(A->colidx, colidx, 4 * nnz, hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(A->vals, vals, 8 * nnz, hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(A->diag, d.data(), 8 * n, hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    *out = A;
  });
}

int ffhip_mat_set_values(ffhip_mat A, const double *vals) {
  return guard([&] {
    FFHIP_REQUIRE(A && vals, FFHIP_ERR_ARG, "ffhip_mat_set_values: null argument");
    FFHIP_CHECK(hipMemcpyAsync(A->vals, vals, 8 * A->nnz, hipMemcpyHostToDevice, A->ctx->stream));
    FFHIP_CHECK(hipStreamSynchronize(A->ctx->stream));
  });
}

int ffhip_mat_device_ptrs(ffhip_mat A, int64_t **rowptr, int32_t **colidx, double **vals) {
  return guard([&] {
    FFHIP_REQUIRE(A, FFHIP_ERR_ARG, "ffhip_mat_device_ptrs: null matrix");
    if (rowptr) *rowptr = A->rowptr;
    if (colidx) *colidx = A->colidx;
    if (vals) *vals = A->vals;
  });
}

int ffhip_mat_free(ffhip_mat A) {
  return guard([&] {
    if (!A) return;
    (void)hipStreamSynchronize(A->ctx->stream);
    dfree(A->rowptr); dfree(A->colidx); dfree(A->vals); dfree(A->diag); dfree(A->slotmap);
    dfree(A->nodeptr); dfree(A->slot8); dfree(A->big);
    dfree(A->fptr); dfree(A->fcol); dfree(A->fslot);
    delete A;
  });
}

int ffhip_spmv(ffhip_ctx ctx, ffhip_mat A, const double *x, double *y, int trans) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && A && x && y, FFHIP_ERR_ARG, "ffhip_spmv: null argument");
    hipStream_t s = ctx->stream;
    if (A->half) {  // y = (L + L^T - diag L) x over the full view; symmetric: trans alike
      half_view(A);
      ctx->tic();
      k_spmv_csr<<<nblocks(A->n, 256), 256, 0, s>>>(A->n, A->fptr, A->fcol, A->vals, A->fslot, x, y);
      FFHIP_CHECK(hipGetLastError());
      ctx->toc();
      return;
    }
    int64_t *tmap = nullptr;
    if (trans) {
      tmap = dalloc<int64_t>(A->nnz);
      const bool ok = transpose_map(ctx, A, tmap);
      if (!ok) dfree(tmap);
      FFHIP_REQUIRE(ok, FFHIP_ERR_UNSUPPORTED, "ffhip_spmv: transpose needs a symmetric pattern");
    }
    ctx->tic();
    k_spmv_csr<<<nblocks(A->n, 256), 256, 0, s>>>(A->n, A->rowptr, A->colidx, A->vals, tmap, x, y);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    dfree(tmap);
  });
}

}  // extern "C"
