// ffhip_cg.hip — Jacobi-preconditioned conjugate gradient on one MI355X.
//
// Replaces SolverCG / HMatVirtPrecon / ConjugueGradient
// (src/femlib/VirtualSolverCG.hpp:13-192, src/femlib/CG.cpp:195-265).
//
// Layout: the solver streams the matrix in sliced ELL with 64-row slices
// (one wavefront per slice, one lane per row, column-major inside a slice),
// so every matrix load of a wave is one contiguous 256 B (int32) / 512 B
// (FP64) segment and no cross-lane reduction is needed per row.
//
// Iteration (same recurrences as ConjugueGradient):
//   K1: H = gamma*H_prev - C g  (formed on the fly for every gathered column)
//       AH = A H, partial sums of (H,AH) and (g,H)
//   K2: rho = -(g,H)/(H,AH); x += rho H; g += rho AH; Cg = D^-1 g; (g,Cg)
// Each kernel starts by re-reducing the previous kernel's per-block partial
// sums in a fixed order, so all blocks derive identical scalars with no
// cross-block atomics or fences, and the result is bitwise reproducible.
// Convergence is decided on the device (sticky done flag); the host only
// polls every few iterations.
//
// MIXED mode: the FP32 copy of D^-1/2 A D^-1/2 is streamed by an FP32 CG
// whose iterate is folded into the FP64 solution and whose residual is
// replaced by the FP64 residual b - A x at reliable-update points
// (iterative refinement that keeps the Krylov direction); dot products are
// accumulated in FP64 and convergence is only accepted on the FP64 residual.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ffhip_internal.hpp"

namespace ffhip {
void build_v2e(ffhip_mesh m);
int read_flag(ffhip_ctx ctx, int *d);
bool transpose_map(ffhip_ctx ctx, ffhip_mat A, int64_t *tmap);

constexpr int kSlice = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kSlice;

// Tuning of the streaming mat-vec kernels: unroll depth of a row (sliced ELL;
// 3x3-block rows with FP64 and with FP32 values), the waves per SIMD the
// register allocation targets, and resident blocks per CU.
#ifndef FFHIP_SELL_U
#define FFHIP_SELL_U 16
#endif
#ifndef FFHIP_WPE
#define FFHIP_WPE 3
#endif
#ifndef FFHIP_BSR_U
#define FFHIP_BSR_U 2
#endif
#ifndef FFHIP_BSR_U32
#define FFHIP_BSR_U32 8
#endif
#ifndef FFHIP_GRID_PER_CU
#define FFHIP_GRID_PER_CU (FFHIP_WPE)
#endif
#define FFHIP_MATVEC_ATTR __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(FFHIP_WPE, FFHIP_WPE)))
constexpr int kU = FFHIP_SELL_U;

// matrix layouts of the mat-vec kernels (described at MatFmt below)
constexpr int kSell = 0;
constexpr int kCsrV = 1;
constexpr int kSell16 = 2;  // sliced ELL with 16-bit column deltas (col - row)
constexpr int kBsr3 = 3;    // 3x3 blocks of the [P2,P2,P2] node rows
// MIXED: the FP32 residual is replaced by the FP64 one once (r32, r32) fell by
// this factor since the last replacement (every two decades of the residual
// norm: FP32 recurrences of ill-conditioned systems, e.g. vector P2
// elasticity, drift if left longer; P1 Poisson is insensitive to it)
constexpr double kMixUpdate = 1e-2;
constexpr int kLanesPerRow = 16;

struct CgState {
  double gcg_prev[2];  // (g, Cg) at the start of the iteration, by parity
  double gcg0, eps2;
  double sigma;        // MIXED: scale of the FP32 residual
  double rho32;        // MIXED: (r32, r32) at the start of the last executed step
  int done;            // 0 run, 1 converged, 2 gcg0 < 1e-30, 3 itmax, 4 NaN
  int iters;
  int need_update;     // MIXED: reliable update requested
  int refinements;
  // iteration of the first launch of the current chunk (the kernels add their
  // position in the chunk; k_cg_advance moves it on), so one captured chunk
  // serves every chunk of a solve
  int it0;
  int itmax;
  double eps;
  // FP64 PCG: the stop decision of the direction kernel (published as done by
  // the next mat-vec launch, so every block of the deciding launch still runs)
  // and the step length of the last update, whose x += rho H the next
  // direction kernel applies
  int stop;
  double rho;
};

// ---------------------------------------------------------------- reductions
// Sum of P partials (stride `stride`, component `c`), identical in every block.
__device__ inline double reduce_partials(const double *__restrict__ p, int P, int stride, int c,
                                         double *lds) {
  double s = 0.;
  for (int i = threadIdx.x; i < P; i += kBlock) s += p[(int64_t)i * stride + c];
  return block_sum<kBlock>(s, lds);
}

template <class T>
__device__ inline T ntload(const T *p) { return __builtin_nontemporal_load(p); }

// One row of a slice: sum_k a_k * X(j_k) over the L (wave-uniform) entries of
// this lane's row.  Unrolled by U: the U column and value loads of a round
// are all issued before any gather, so every wave keeps 2U streaming loads
// and U (or 2U) gathers in flight instead of one dependent chain per entry.
// IT = int32_t: absolute columns; IT = int16_t: deltas to the row (row0 + d).
template <int U, class V, class ACC, class IT, class F>
__device__ inline ACC sell_row(const IT *__restrict__ colp, const V *__restrict__ valp, int L,
                               int32_t row0, F xval) {
  ACC acc = 0;
  for (int k = 0; k < L; k += U) {
    int32_t j[U];
    V a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      j[u] = row0;  // past the row end: a valid column with a zero value
      a[u] = 0;
      if (k + u < L) {
        const IT c = ntload(colp + (int64_t)(k + u) * kSlice);
        j[u] = sizeof(IT) == 2 ? row0 + (int32_t)c : (int32_t)c;
        a[u] = ntload(valp + (int64_t)(k + u) * kSlice);
      }
    }
    ACC xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = xval(j[u]);  // all gathers issued together
#pragma unroll
    for (int u = 0; u < U; ++u) acc += (ACC)a[u] * xv[u];
  }
  return acc;
}

// 16-bit column deltas packed four per lane: the deltas of entries 4g..4g+3 of
// a lane's row are one 8-byte load (layout [g][lane][4] inside a slice), so a
// round of U entries issues U/4 index loads ahead of the U value loads.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP64 values of a kSell16 slice of width L are stored in entry pairs: pair m
// (entries 2m, 2m+1) of lane l at 2 (64 m + l), so one 16-byte load per lane
// reads two entries of its row; for odd L the last entry follows the L/2 pair
// planes as one plane of single values.  Position of entry k of lane l,
// relative to the slice start:
__host__ __device__ inline int64_t pair_pos(int L, int k, int lane) {
  const int h = L >> 1;
  return (k >> 1) < h ? ((int64_t)(k >> 1) * kSlice + lane) * 2 + (k & 1) : (int64_t)h * 2 * kSlice + lane;
}

template <int U, class V, class ACC, class F>
__device__ inline ACC sell_row16(const int16_t *__restrict__ colp, const V *__restrict__ valp,
                                 int L, int32_t row0, F xval) {
  static_assert(U % 4 == 0, "unroll by whole index groups");
  ACC acc = 0;
  for (int k = 0; k < L; k += U) {
    int32_t j[U];
    V a[U];
#pragma unroll
    for (int g = 0; g < U / 4; ++g) {
      u32x2 pk = {0u, 0u};  // past the row end: deltas 0 (the row itself)
      if (k + 4 * g < L)
        pk = ntload(reinterpret_cast<const u32x2 *>(colp + (int64_t)((k >> 2) + g) * (4 * kSlice)));
      j[4 * g + 0] = row0 + (int32_t)(int16_t)(pk.x & 0xffffu);
      j[4 * g + 1] = row0 + (int32_t)(int16_t)(pk.x >> 16);
      j[4 * g + 2] = row0 + (int32_t)(int16_t)(pk.y & 0xffffu);
      j[4 * g + 3] = row0 + (int32_t)(int16_t)(pk.y >> 16);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = 0;
      if (k + u < L) a[u] = ntload(valp + (int64_t)(k + u) * kSlice);
    }
    ACC xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = xval(j[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += (ACC)a[u] * xv[u];
  }
  return acc;
}

// wave-uniform slice index (kept in scalar registers)
__device__ inline int64_t uniform(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return ((int64_t)hi << 32) | (uint32_t)lo;
}
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kSlice); }

// Slices of this wave in the streaming mat-vec kernels.  CYCLIC: each XCD's
// contiguous slice range is walked by all of its waves together, wave q taking
// slices q, q + Q, q + 2Q, ... (Q waves on the XCD), so the XCD streams one
// contiguous window of Q slices at a time and a column's x value, needed by
// rows up to one window-width apart (the plane neighbours of a 3D mesh), is
// re-read while still in that XCD's L2 (cube(160) P1: 8 % less HBM fetch, the
// FP32 mat-vec 6 % faster).  Otherwise each block walks its own contiguous
// chunk of slices — kept for the [P2,P2,P2] block layout, whose node rows are
// sorted by length in windows and ran 1.8x slower with the cyclic walk.
constexpr bool kCyclicSell = true, kCyclicBsr = false;
template <bool CYCLIC>
__device__ inline void wave_slices(int64_t S, int64_t &first, int64_t &end, int64_t &stride) {
  if (CYCLIC) {
    const int G = gridDim.x, b = blockIdx.x;
    const int groups = (G % 8) ? 1 : 8;
    const int grp = b % groups, loc = b / groups, nloc = G / groups;
    first = (grp * S) / groups + (int64_t)loc * kWavesPerBlock + wave_id();
    end = ((grp + 1) * S) / groups;
    stride = (int64_t)nloc * kWavesPerBlock;
  } else {
    const int64_t G = gridDim.x, lb = xcd_block_id();
    first = (lb * S) / G + wave_id();
    end = ((lb + 1) * S) / G;
    stride = kWavesPerBlock;
  }
}

// ---------------------------------------------------------------- SELL build
__global__ void k_sell_len(int64_t n, int64_t S, const int64_t *__restrict__ rowptr,
                           int32_t *__restrict__ len, int64_t *__restrict__ cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  int64_t m = 0;
  for (int l = 0; l < kSlice; ++l) {
    const int64_t r = s * kSlice + l;
    if (r < n) {
      const int64_t c = rowptr[r + 1] - rowptr[r];
      m = c > m ? c : m;
    }
  }
  len[s] = (int32_t)m;
  cnt[s] = m * kSlice;
}

// vals_src may be permuted through tmap (transposed values)
// One block per 64-row slice: the slice's CSR segment is read once, coalesced,
// into LDS (when it fits in kFillStage entries; otherwise the rows are read
// from memory where they lie), then the block writes the slice layout plane
// by plane, consecutive threads on consecutive rows, so every store is
// coalesced (FP64 entry pairs as one 16-byte store).  Entries past a row's
// end are zeros at the row's own column.
// 1280 entries (a P1 slice of 64 rows of 15 holds 960): 15 KB of LDS, 10
// blocks per CU (2048 entries, 6 blocks per CU: cube(160) 401 -> 330 us)
constexpr int kFillStage = 1280;
__global__ __launch_bounds__(256) void k_sell_fill(
    int64_t n, int64_t S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const double *__restrict__ vals, const int64_t *__restrict__ tmap,
    const int64_t *__restrict__ sptr, const int32_t *__restrict__ len,
    const double *__restrict__ dsc, int32_t *__restrict__ scol,
    const int64_t *__restrict__ sptr16, int16_t *__restrict__ scol16, double *__restrict__ sval,
    float *__restrict__ sval32, bool pairs, const int64_t *__restrict__ diag,
    double *__restrict__ dout, uint8_t *__restrict__ hout, int *__restrict__ far) {
  __shared__ int32_t lc[kFillStage];
  __shared__ double lv[kFillStage];
  __shared__ int64_t rs[kSlice];  // row start (in the stage, or in the CSR)
  __shared__ int32_t rl[kSlice];  // row length
  const int t = threadIdx.x;
  const int64_t sl = blockIdx.x;
  const int64_t r0 = sl * kSlice;
  const int64_t rend = r0 + kSlice < n ? r0 + kSlice : n;
  const int64_t seg0 = r0 < n ? rowptr[r0] : 0, seg1 = r0 < n ? rowptr[rend] : 0;
  const bool staged = seg1 - seg0 <= kFillStage;
  if (t < kSlice) {
    const int64_t r = r0 + t;
    const int64_t p = r < n ? rowptr[r] : 0;
    rs[t] = staged ? p - seg0 : p;
    rl[t] = r < n ? (int32_t)(rowptr[r + 1] - p) : 0;
  }
  if (staged)
    for (int64_t i = t; i < seg1 - seg0; i += blockDim.x) {
      lc[i] = colidx[seg0 + i];
      lv[i] = tmap ? vals[tmap[seg0 + i]] : vals[seg0 + i];
    }
  __syncthreads();
  if (dout && t < kSlice && r0 + t < n) {  // the diagonal of the Jacobi set-up (as k_diag_vals)
    const int64_t r = r0 + t, q = diag[r];
    dout[r] = q < 0 ? 0. : (staged && !tmap && q >= seg0 && q < seg1 ? lv[q - seg0] : vals[q]);
    hout[r] = q >= 0;
  }
  const int L = len[sl];
  const int64_t base = sptr[sl];
  // entry k of lane li's row; past its end a zero at the row's own column
  auto col = [&](int li, int k) -> int32_t {
    if (k < rl[li]) return staged ? lc[rs[li] + k] : colidx[rs[li] + k];
    const int64_t r = r0 + li;
    return r < n ? (int32_t)r : 0;
  };
  auto val = [&](int li, int k) -> double {
    if (k >= rl[li]) return 0.;
    const int64_t q = rs[li] + k;
    return staged ? lv[q] : (tmap ? vals[tmap[q]] : vals[q]);
  };
  if (sval && pairs) {
    const int Lh = L >> 1;
    for (int it = t; it < (Lh + (L & 1)) * kSlice; it += blockDim.x) {
      const int li = it % kSlice, kp = it / kSlice;
      if (kp < Lh)
        *reinterpret_cast<double2 *>(sval + base + ((int64_t)kp * kSlice + li) * 2) =
            double2{val(li, 2 * kp), val(li, 2 * kp + 1)};
      else  // the odd last entry: its own plane after the pairs
        sval[base + pair_pos(L, L - 1, li)] = val(li, L - 1);
    }
  }
  for (int it = t; it < L * kSlice; it += blockDim.x) {
    const int li = it % kSlice, k = it / kSlice;
    const int64_t idx = base + it;  // entry k of lane li: k kSlice + li
    if (sval && !pairs) sval[idx] = val(li, k);
    if (scol || sval32) {
      const int32_t c = col(li, k);
      if (scol) scol[idx] = c;
      if (sval32)
        sval32[idx] = k < rl[li] ? (float)(val(li, k) * dsc[r0 + li] * dsc[c]) : 0.f;
    }
  }
  if (scol16) {  // [g][lane][4] packed deltas, the group tail padded with 0: one 8-byte store per group
    for (int it = t; it < ((L + 3) >> 2) * kSlice; it += blockDim.x) {
      const int li = it % kSlice, g = it / kSlice;
      const int64_t r = r0 + li;
      uint32_t h[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * g + q;
        const int64_t dl = k < rl[li] ? (int64_t)col(li, k) - r : 0;
        if (far && (dl < -32768 || dl > 32767)) atomicOr(far, 1);  // then refilled with int32 columns
        h[q] = (uint32_t)(uint16_t)(int16_t)dl;
      }
      const u32x2 pk = {h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
      *reinterpret_cast<u32x2 *>(scol16 + sptr16[sl] + (int64_t)it * 4) = pk;
    }
  }
}

// slice sizes of the packed 16-bit index array (widths rounded up to 4)
__global__ void k_sell_len4(int64_t S, const int32_t *__restrict__ len, int64_t *__restrict__ cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) cnt[s] = (int64_t)((len[s] + 3) & ~3) * kSlice;
}

// node-row sort keys: window of 4096 rows in the high bits, longer rows first
__global__ void k_bsr_keys(int64_t nb, const int64_t *__restrict__ bptr, uint32_t *__restrict__ key,
                           int32_t *__restrict__ val) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nb) return;
  const int64_t L = bptr[a + 1] - bptr[a];
  key[a] = ((uint32_t)(a >> 12) << 8) | (uint32_t)(255 - (L < 255 ? L : 255));
  val[a] = (int32_t)a;
}

// slice widths of the sorted node rows
__global__ void k_bsr_len(int64_t nb, int64_t S3, const int64_t *__restrict__ bptr,
                          const int32_t *__restrict__ perm, int32_t *__restrict__ len,
                          int64_t *__restrict__ cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S3) return;
  int64_t m = 0;
  for (int l = 0; l < kSlice; ++l) {
    const int64_t i = s * kSlice + l;
    if (i < nb) {
      const int32_t a = perm[i];
      const int64_t c = bptr[a + 1] - bptr[a];
      m = c > m ? c : m;
    }
  }
  len[s] = (int32_t)m;
  cnt[s] = m * kSlice;
}

// 3x3 blocks of a vector-P2 matrix from its scalar CSR: entry k of sorted node
// row i (node a = perm[i]) holds rows 3a..3a+2 at offsets 3k..3k+2 of those
// rows (values optionally through tmap; FP32 copy of D^-1/2 A D^-1/2).
// One block per 64-row slice: the slice's rows are read KC blocks at a time,
// each of its 3 x 64 scalar rows as one contiguous run of 3 KC values, into
// LDS, and written out entry-major (the 9 x 64 values of block entry k of the
// slice at 9 ptr[s] + 9 64 k, placed by bsr_pos); padding entries get value 0
// and the row's own block column.
constexpr int kBsrKC = 8;
// Position of value q (3c + c') of lane li inside one slice entry of the
// 3x3-block values (9 x 64 values per entry): values 0..7 in vectors of W = 16
// bytes / sizeof(value) per lane (FP64 pairs, FP32 quads), [vector][lane][W],
// then value 8 of the 64 lanes.
template <class V>
__host__ __device__ constexpr int bsr_pos(int q, int li) {
  constexpr int W = 16 / (int)sizeof(V);
  return q < 8 ? ((q / W) * kSlice + li) * W + q % W : 8 * kSlice + li;
}
__global__ __launch_bounds__(256) void k_bsr_fill(
    int64_t nb, int64_t ents, const int64_t *__restrict__ bptr, const int32_t *__restrict__ perm,
    const int64_t *__restrict__ sptr, const int32_t *__restrict__ slen,
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const double *__restrict__ vals, const int64_t *__restrict__ tmap,
    const double *__restrict__ dsc, int32_t *__restrict__ bcol, double *__restrict__ bval,
    float *__restrict__ bval32) {
  constexpr int KC = kBsrKC, P = kSlice + 1;  // P: padded lane stride (LDS banks)
  __shared__ double T[9 * KC * P];
  __shared__ int32_t Cb[KC * kSlice];
  __shared__ int32_t An[kSlice], Ln[kSlice];
  __shared__ int64_t R0[3 * kSlice];
  const int64_t s = blockIdx.x;
  const int t = threadIdx.x;
  if (t < kSlice) {
    const int64_t i = s * kSlice + t;
    const int32_t a = i < nb ? perm[i] : -1;
    An[t] = a;
    Ln[t] = a >= 0 ? (int32_t)(bptr[a + 1] - bptr[a]) : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) R0[3 * t + c] = a >= 0 ? rowptr[3 * (int64_t)a + c] : 0;
  }
  __syncthreads();
  const int W = slen[s];
  const int64_t base = sptr[s];
  for (int k0 = 0; k0 < W; k0 += KC) {
    const int kc = W - k0 < KC ? W - k0 : KC;
    // read: item = ((li * 3 + c) * 3 KC + x), x = 3 kk + cp
    for (int it = t; it < kSlice * 9 * KC; it += 256) {
      const int x = it % (3 * KC), rc = it / (3 * KC), c = rc % 3, li = rc / 3;
      const int kk = x / 3, cp = x % 3, k = k0 + kk;
      double v = 0.;
      if (kk < kc && k < Ln[li]) {
        const int64_t src = R0[3 * li + c] + 3 * (int64_t)k + cp;
        v = tmap ? vals[tmap[src]] : vals[src];
      }
      T[((3 * c + cp) * KC + kk) * P + li] = v;
    }
    for (int it = t; it < KC * kSlice; it += 256) {
      const int li = it % kSlice, kk = it / kSlice, k = k0 + kk;
      Cb[it] = kk < kc && k < Ln[li] ? colidx[R0[3 * li] + 3 * (int64_t)k] / 3 : -1;
    }
    __syncthreads();
    // write: entry kk, value q, lane li (one contiguous run of 9 x 64 kc values)
    for (int it = t; it < 9 * kc * kSlice; it += 256) {
      const int li = it % kSlice, q = (it / kSlice) % 9, kk = it / (9 * kSlice);
      const int64_t eb = 9 * (base + (int64_t)(k0 + kk) * kSlice);
      const double v = T[(q * KC + kk) * P + li];
      if (bval) bval[eb + bsr_pos<double>(q, li)] = v;
      if (bval32) {
        const int cb = Cb[kk * kSlice + li];
        bval32[eb + bsr_pos<float>(q, li)] =
            cb < 0 ? 0.f : (float)(v * dsc[3 * (int64_t)An[li] + q / 3] * dsc[3 * (int64_t)cb + q % 3]);
      }
    }
    if (bcol)
      for (int it = t; it < kc * kSlice; it += 256) {
        const int li = it % kSlice, kk = it / kSlice;
        const int cb = Cb[kk * kSlice + li];
        bcol[base + (int64_t)(k0 + kk) * kSlice + li] = cb >= 0 ? cb : (An[li] >= 0 ? An[li] : 0);
      }
    __syncthreads();
  }
}

// First slice of each of nw waves when the S3 slices are split over them by
// entries: wave w takes the slices that start in [w E / nw, (w+1) E / nw),
// E = sptr[S3].
__global__ void k_bsr_waves(int64_t S3, const int64_t *__restrict__ sptr, int nw,
                            int32_t *__restrict__ wstart) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > nw) return;
  const int64_t e = w == nw ? sptr[S3] : (int64_t)w * sptr[S3] / nw;
  int64_t lo = 0, hi = S3;  // first slice s with sptr[s] >= e
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (sptr[mid] < e) lo = mid + 1;
    else hi = mid;
  }
  wstart[w] = (int32_t)lo;
}

// flags a row with a column farther than 16 bits from the row index
// the columns of a row are sorted: its first and last bound the deltas
// CSR-layout values: A (optionally transposed through tmap) and/or the FP32
// copy of D^-1/2 A D^-1/2, one 16-lane group per row
__global__ void k_csr_fill(int64_t n, const int64_t *__restrict__ rowptr,
                           const int32_t *__restrict__ colidx, const double *__restrict__ vals,
                           const int64_t *__restrict__ tmap, const double *__restrict__ dsc,
                           double *__restrict__ dv, float *__restrict__ fv) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  if (r >= n) return;
  const int64_t k1 = rowptr[r + 1];
  for (int64_t k = rowptr[r] + lane; k < k1; k += kLanesPerRow) {
    const double a = tmap ? vals[tmap[k]] : vals[k];
    if (dv) dv[k] = a;
    if (fv) fv[k] = (float)(a * dsc[r] * dsc[colidx[k]]);
  }
}

// ---------------------------------------------------------------- Jacobi / tgv
__global__ void k_diag_vals(int64_t n, const int64_t *__restrict__ diag,
                            const double *__restrict__ vals, double *__restrict__ d,
                            uint8_t *__restrict__ has) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = diag[i];
  d[i] = s >= 0 ? vals[s] : 0.;
  has[i] = s >= 0;
}

// What the solver set-up reads back to the host, in one copy: the tgv
// figures (k_tgv_reduce) and the layout sizes (k_setup_sizes).
struct SetupRb {
  double ttgv, ntgv;      // the tgv value (0: none) and the count of rows holding it
  int64_t ents, ents16;   // entries of the sliced layout, of its packed 16-bit deltas
  int far;                // some column is more than 2^15 from its row
};

// pass 1: max diagonal; pass 2: count of the max (*top), max of the rest
// (HashMatrix::gettgv, HashMatrix.cpp:1340-1371)
__global__ void k_tgv_pass(int64_t n, const double *__restrict__ d, const uint8_t *__restrict__ has,
                           const double *__restrict__ topp, int pass, double *__restrict__ out) {
  __shared__ double lds[kBlock / kWave];
  const double top = pass == 2 ? *topp : 0.;
  double m = 0., c = 0.;  // the reference starts from ttgv = max1 = 0
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    if (!has[i]) continue;
    const double a = d[i];
    if (pass == 1) m = fmax(m, a);
    else if (a == top) c += 1.;
    else m = fmax(m, a);
  }
  // block max via shuffles
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) lds[w] = m;
  __syncthreads();
  double bm = 0.;
  for (int i = 0; i < kBlock / kWave; ++i) bm = fmax(bm, lds[i]);
  const double bc = block_sum<kBlock>(c, lds);
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = bm; out[2 * blockIdx.x + 1] = bc; }
}

// the partials of a k_tgv_pass, one block: pass 1 puts the max in rb->ttgv;
// pass 2 applies gettgv's rule (a max not 1e6 x above the next distinct
// value is no tgv) to that max, the count and the max of the rest.  Maxima
// and integral counts are exact in any order.
__global__ __launch_bounds__(kBlock) void k_tgv_reduce(int GB, const double *__restrict__ part,
                                                       int pass, SetupRb *__restrict__ rb) {
  __shared__ double lds[kBlock / kWave];
  double m = 0., c = 0.;
  for (int i = threadIdx.x; i < GB; i += kBlock) {
    m = fmax(m, part[2 * i]);
    c += part[2 * i + 1];
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) lds[w] = m;
  __syncthreads();
  double bm = 0.;
  for (int i = 0; i < kBlock / kWave; ++i) bm = fmax(bm, lds[i]);
  const double count = block_sum<kBlock>(c, lds);
  if (threadIdx.x != 0) return;
  if (pass == 1) {
    rb->ttgv = bm;
    return;
  }
  double ttgv = rb->ttgv, ntgv = count;
  if (bm * 1e6 > ttgv) { ttgv = 0.; ntgv = 0.; }
  rb->ttgv = ttgv;
  rb->ntgv = ntgv;
}

// the layout sizes for the host (entries of the sliced layout and of the
// packed deltas: the last slice offsets) and the far-column flag
__global__ void k_setup_sizes(SetupRb *__restrict__ rb, const int64_t *__restrict__ ents,
                              const int64_t *__restrict__ ents16, const int *__restrict__ far) {
  rb->ents = ents ? *ents : 0;
  rb->ents16 = ents16 ? *ents16 : 0;
  rb->far = far ? *far : 1;
}

__global__ void k_jacobi(int64_t n, const double *__restrict__ d, const uint8_t *__restrict__ has,
                         const SetupRb *__restrict__ rb, double *__restrict__ dinv,
                         double *__restrict__ dsc, uint8_t *__restrict__ wcl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tgve = rb->ttgv <= 0 ? 1e200 : rb->ttgv;
  const bool with_wcl = rb->ntgv > 0;
  const double a = has[i] ? d[i] : 0.;  // a missing diagonal reads as 0 (then 1)
  dinv[i] = (a * a < 1e-60) ? 1. : 1. / a;
  dsc[i] = a > 0. ? 1. / sqrt(a) : 1.;
  wcl[i] = with_wcl ? (has[i] && d[i] == tgve) : 0;
}

// SetInitWithBC (VirtualSolverCG.hpp:93-99)
__global__ void k_init_bc(int64_t n, const uint8_t *__restrict__ wcl, const double *__restrict__ b,
                          double tgv, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && wcl[i]) x[i] = b[i] / tgv;
}

// ---------------------------------------------------------------- matrix formats
// Every mat-vec kernel streams the matrix through one of these layouts:
//   kSell:   sliced ELL, 64-row slices (one lane per row, no reduction) — for
//            matrices whose rows have nearly equal length (P1);
//   kSell16: the same with 16-bit column deltas packed four per lane;
//   kCsrV:   CSR with 16 lanes per row and a shuffle reduction — for rows of
//            very different lengths, where sliced ELL would pad;
//   kBsr3:   3x3 blocks of a vector-P2 matrix, sliced ELL over node rows:
//            the node rows are sorted by length inside windows of 4096 rows
//            (so a 64-row slice pads little), one lane per node row (3 scalar
//            rows), one int32 block column and 9 values per slice entry, the
//            values of a slice entry-major (9 x 64 per entry) — 76 B per 9
//            FP64 entries instead of 108, every load coalesced.
template <class V>
struct MatFmt {
  int64_t n, S;
  const int64_t *ptr;  // kSell*: slice offsets (S+1); kCsrV: row pointers (n+1)
  const int32_t *len;  // kSell*: slice widths
  const int32_t *col;  // kSell, kCsrV
  const int16_t *col16;  // kSell16: col - row, packed [group of 4][lane][4] per slice
  const int64_t *ptr16;  // kSell16: slice offsets of col16
  const V *val;          // kBsr3: 9 ents values, entry k of slice s at 9 ptr[s] + 9 64 k (bsr_pos)
  int64_t nb, ents;      // kBsr3: node rows, slice entries (ptr/len/col: slices, perm: rows)
  const int32_t *perm;   // kBsr3: node row of each sorted position
  const int32_t *wstart; // kBsr3: first slice of each of the nw waves (nw + 1), split by entries
  int nw;
};

// Buffer resource over [p, p + bytes): a buffer load at a byte offset past
// `bytes` returns 0 and touches no memory (raw buffer, stride 0).
__device__ inline __amdgpu_buffer_rsrc_t slice_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
constexpr int kBufNT = 2;  // cache policy of the streamed matrix loads: non-temporal

// Entries k >= U of a kSell16 FP64 row (rows longer than U), summed into acc
// in entry order, four per round (one packed delta group).
template <class ACC, class F>
__device__ inline ACC sell16_tail(const int16_t *__restrict__ col16, const double *__restrict__ val,
                                  int L, int lane, int32_t row0, F xval, ACC acc) {
  for (int k = kU; k < L; k += 4) {
    const u32x2 pk = ntload(reinterpret_cast<const u32x2 *>(col16 + (int64_t)(k >> 2) * (4 * kSlice) + 4 * lane));
    const int32_t j[4] = {row0 + (int32_t)(int16_t)(pk.x & 0xffffu), row0 + (int32_t)(int16_t)(pk.x >> 16),
                          row0 + (int32_t)(int16_t)(pk.y & 0xffffu), row0 + (int32_t)(int16_t)(pk.y >> 16)};
    double a[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) a[t] = k + t < L ? ntload(val + pair_pos(L, k + t, lane)) : 0.;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc += (ACC)a[t] * xval(j[t]);
  }
  return acc;
}

// The FP64 kSell16 walk over this wave's slices.  A slice's values (entry
// pairs, 16-byte loads) and packed deltas are read through buffer resources
// that end at the slice's last entry, so the loads of a round are issued
// unconditionally: entries past the slice width read as value 0, delta 0
// (the row itself).  pre(r) (the row's own epilogue operands) is issued with
// the gathers.  epi(r, acc, pre(r), r < n) is called for every lane (rows
// past n only in the last slice, whose pre was given a clamped row).  The
// products are summed in entry order.
template <class ACC, class F, class P, class E>
__device__ inline void sell16_walk(const MatFmt<double> &M, F xval, P pre, E epi) {
  constexpr int U = kU, G4 = kU / 4, NP = kU / 2;
  static_assert(kU % 4 == 0, "unroll by whole index groups");
  int64_t s, s1, ds;
  wave_slices<kCyclicSell>(M.S, s, s1, ds);
  const int lane = threadIdx.x % kSlice;
  for (; s < s1; s += ds) {
    const int64_t vb = uniform(M.ptr[s]), cb = uniform(M.ptr16[s]);
    const int L = __builtin_amdgcn_readfirstlane(M.len[s]);
    const int Lh = L >> 1;
    const int64_t r = s * kSlice + lane;
    // rows past n (last slice) have only padding: their column is clamped
    const int32_t r32 = (int32_t)(r < M.n ? r : M.n - 1);
    const auto rc = slice_rsrc(M.col16 + cb, (uint32_t)((L + 3) >> 2) * (4 * kSlice) * 2);
    const auto rp = slice_rsrc(M.val + vb, (uint32_t)Lh * (2 * kSlice) * 8);
    const auto r1 = slice_rsrc(M.val + vb + (int64_t)Lh * (2 * kSlice), (uint32_t)(L & 1) * kSlice * 8);
    u32x2 pk[G4];
#pragma unroll
    for (int g = 0; g < G4; ++g)
      pk[g] = __builtin_amdgcn_raw_buffer_load_b64(rc, (g * kSlice + lane) * 8, 0, kBufNT);
    u32x4 q[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m)
      q[m] = __builtin_amdgcn_raw_buffer_load_b128(rp, (m * kSlice + lane) * 16, 0, kBufNT);
    const u32x2 odd = __builtin_amdgcn_raw_buffer_load_b64(r1, lane * 8, 0, kBufNT);
    int32_t j[U];
#pragma unroll
    for (int g = 0; g < G4; ++g) {
      j[4 * g + 0] = r32 + (int32_t)(int16_t)(pk[g].x & 0xffffu);
      j[4 * g + 1] = r32 + (int32_t)(int16_t)(pk[g].x >> 16);
      j[4 * g + 2] = r32 + (int32_t)(int16_t)(pk[g].y & 0xffffu);
      j[4 * g + 3] = r32 + (int32_t)(int16_t)(pk[g].y >> 16);
    }
    double a[U];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      const u32x2 lo = m == Lh ? odd : u32x2{q[m].x, q[m].y};
      a[2 * m] = __builtin_bit_cast(double, lo);
      a[2 * m + 1] = __builtin_bit_cast(double, u32x2{q[m].z, q[m].w});
    }
    ACC xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = xval(j[u]);
    const auto pv = pre(r32);
    ACC acc = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) acc += (ACC)a[u] * xv[u];
    if (L > U) acc = sell16_tail(M.col16 + cb, M.val + vb, L, lane, r32, xval, acc);
    epi(r, acc, pv, r < M.n);
  }
}

// Calls epi(r, sum_k a_rk X(j_k)) once for every row r of this block's range;
// for kCsrV only the first lane of each 16-lane group calls it.
template <int KIND, class V, class ACC, class F, class E>
__device__ inline void matvec_rows(const MatFmt<V> &M, F xval, E epi) {
  if constexpr (KIND == kSell16 && sizeof(V) == 8) {
    sell16_walk<ACC>(M, xval, [](int32_t) { return 0; },
                     [&](int64_t r, ACC acc, int, bool row) { if (row) epi(r, acc); });
  } else if constexpr (KIND == kSell || KIND == kSell16) {
    int64_t s0, s1, ds;
    wave_slices<kCyclicSell>(M.S, s0, s1, ds);
    const int lane = threadIdx.x % kSlice;
    for (int64_t s = s0; s < s1; s += ds) {
      const int64_t base = uniform(M.ptr[s]) + lane;
      const int L = __builtin_amdgcn_readfirstlane(M.len[s]);
      const int64_t r = s * kSlice + lane;
      // rows past n (last slice) have only padding: their column is clamped
      const int32_t r32 = (int32_t)(r < M.n ? r : M.n - 1);
      ACC acc;
      if constexpr (KIND == kSell16)
        acc = sell_row16<kU, V, ACC>(M.col16 + uniform(M.ptr16[s]) + 4 * lane, M.val + base, L, r32,
                                     xval);
      else
        acc = sell_row<kU, V, ACC, int32_t>(M.col + base, M.val + base, L, r32, xval);
      if (r < M.n) epi(r, acc);
    }
  } else if constexpr (KIND == kBsr3) {
    // Groups of U block entries; the next group's block columns are loaded
    // with the current group's values, so the gathers of x need no second
    // round trip.  The slice's columns and values are read through buffer
    // resources bounded to the slice: past its width they read 0 without a
    // memory access, and those blocks' sums are dropped by a uniform select.
    constexpr int U = sizeof(V) == 8 ? FFHIP_BSR_U : FFHIP_BSR_U32;
    const int64_t S3 = (M.nb + kSlice - 1) / kSlice;
    int64_t s0, s1, ds;
    if (M.wstart && (int)gridDim.x * kWavesPerBlock == M.nw) {
      // slices split over the waves by entries (rows of vertex and of edge
      // nodes differ in length by about 2.5x)
      const int wg = xcd_block_id() * kWavesPerBlock + wave_id();
      s0 = __builtin_amdgcn_readfirstlane(M.wstart[wg]);
      s1 = __builtin_amdgcn_readfirstlane(M.wstart[wg + 1]);
      ds = 1;
    } else {
      wave_slices<kCyclicBsr>(S3, s0, s1, ds);
    }
    const int lane = threadIdx.x % kSlice;
    for (int64_t s = s0; s < s1; s += ds) {
      const int64_t i = s * kSlice + lane;
      const int32_t a = i < M.nb ? M.perm[i] : -1;
      const int64_t sb = uniform(M.ptr[s]);
      const int L = __builtin_amdgcn_readfirstlane(M.len[s]);
      const auto rc = slice_rsrc(M.col + sb, (uint32_t)L * kSlice * 4);
      const auto rv = slice_rsrc(M.val + 9 * sb, (uint32_t)L * (9 * kSlice) * (uint32_t)sizeof(V));
      ACC s0v = 0, s1v = 0, s2v = 0;
      int32_t bn[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        bn[u] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rc, (u * kSlice + lane) * 4, 0, kBufNT);
      for (int k = 0; k < L; k += U) {
        int32_t b[U];
        ACC xv[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          b[u] = bn[u];
#pragma unroll
          for (int c = 0; c < 3; ++c) xv[u][c] = xval(3 * b[u] + c);
        }
        V wv[U][9];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          constexpr int W = 16 / (int)sizeof(V), Z = (int)sizeof(V);
          const int eb = 9 * kSlice * (k + u);
#pragma unroll
          for (int g = 0; g < 8 / W; ++g) {
            const u32x4 r = __builtin_amdgcn_raw_buffer_load_b128(rv, (eb + bsr_pos<V>(g * W, lane)) * Z, 0, kBufNT);
            if constexpr (sizeof(V) == 8) {
              wv[u][2 * g] = __builtin_bit_cast(V, u32x2{r.x, r.y});
              wv[u][2 * g + 1] = __builtin_bit_cast(V, u32x2{r.z, r.w});
            } else {
              // the whole vector is reinterpreted (a bit cast of one component
              // lvalue, r.y, reads the vector's first component)
              const f32x4 f = __builtin_bit_cast(f32x4, r);
              wv[u][4 * g] = f.x;
              wv[u][4 * g + 1] = f.y;
              wv[u][4 * g + 2] = f.z;
              wv[u][4 * g + 3] = f.w;
            }
          }
          const int o8 = (eb + bsr_pos<V>(8, lane)) * Z;
          if constexpr (sizeof(V) == 8)
            wv[u][8] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rv, o8, 0, kBufNT));
          else
            wv[u][8] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(rv, o8, 0, kBufNT));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          bn[u] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rc, ((k + U + u) * kSlice + lane) * 4, 0,
                                                                kBufNT);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const ACC t0 = (ACC)wv[u][0] * xv[u][0] + (ACC)wv[u][1] * xv[u][1] + (ACC)wv[u][2] * xv[u][2];
          const ACC t1 = (ACC)wv[u][3] * xv[u][0] + (ACC)wv[u][4] * xv[u][1] + (ACC)wv[u][5] * xv[u][2];
          const ACC t2 = (ACC)wv[u][6] * xv[u][0] + (ACC)wv[u][7] * xv[u][1] + (ACC)wv[u][8] * xv[u][2];
          const bool in = k + u < L;
          s0v += in ? t0 : (ACC)0;
          s1v += in ? t1 : (ACC)0;
          s2v += in ? t2 : (ACC)0;
        }
        // keeps the next group's column loads in this iteration (the compiler
        // would otherwise load them again at the top of the next one and wait
        // for them there before the gathers)
#pragma unroll
        for (int u = 0; u < U; ++u) asm volatile("" : "+v"(bn[u]));
      }
      if (a >= 0) {
        epi(3 * (int64_t)a, s0v);
        epi(3 * (int64_t)a + 1, s1v);
        epi(3 * (int64_t)a + 2, s2v);
      }
    }
  } else {
    const int64_t G = gridDim.x, lb = xcd_block_id();
    const int64_t r0 = (lb * M.n) / G, r1 = ((lb + 1) * M.n) / G;
    const int g = threadIdx.x / kLanesPerRow, lane = threadIdx.x % kLanesPerRow;
    constexpr int NG = kBlock / kLanesPerRow;
    for (int64_t rb = r0; rb < r1; rb += NG) {
      const int64_t r = rb + g;
      ACC acc = 0;
      if (r < r1) {
        const int64_t k0 = M.ptr[r], k1 = M.ptr[r + 1];
        for (int64_t k = k0 + lane; k < k1; k += 4 * kLanesPerRow) {
          int32_t j[4];
          V a[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int64_t kk = k + u * kLanesPerRow;
            j[u] = 0;
            a[u] = 0;
            if (kk < k1) { j[u] = ntload(M.col + kk); a[u] = ntload(M.val + kk); }
          }
          ACC xv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) xv[u] = xval(j[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u) acc += (ACC)a[u] * xv[u];
        }
      }
#pragma unroll
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
      if (r < r1 && lane == 0) epi(r, acc);
    }
  }
}

// Calls epi(r, sum_k a_rk X(j_k), pre(r), row) for the rows r of this block's
// range, row = false for the padding lanes of the last kSell16 slice (their
// r is past n: epi stores nothing there, and pre was given a clamped row);
// pre(r) (operands of the epilogue) is loaded ahead where the layout allows.
// For kCsrV only the first lane of each 16-lane group calls epi.
template <int KIND, class V, class ACC, class F, class P, class E>
__device__ inline void matvec_rows_pre(const MatFmt<V> &M, F xval, P pre, E epi) {
  if constexpr (KIND == kSell16 && sizeof(V) == 8) {
    sell16_walk<ACC>(M, xval, pre, epi);
  } else {
    matvec_rows<KIND, V, ACC>(M, xval, [&](int64_t r, ACC acc) { epi(r, acc, pre(r), true); });
  }
}

// ---------------------------------------------------------------- FP64 PCG
// G = A x - b, CG = D^-1 G, H0 = 0, partial (G, CG)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_pcg_init(MatFmt<double> M, const double *__restrict__ x,
                                             const double *__restrict__ b,
                                             const double *__restrict__ dinv, double *__restrict__ G,
                                             double *__restrict__ CG, double *__restrict__ H0,
                                             double *__restrict__ pout, CgState *st) {
  __shared__ double lds[kBlock / kWave];
  double part = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return x[j]; },
      [&](int64_t r, double acc) {
        const double g = acc + -1. * b[r];
        const double c = dinv[r] * g;
        G[r] = g; CG[r] = c; H0[r] = 0.;
        part += g * c;
      });
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) { st->done = 0; st->stop = 0; st->iters = 0; st->rho = 0.; }
}

// Start of FP64 PCG iteration it = it0 + c: (g, Cg) from the last update's
// partials, the stop tests of ConjugueGradient (CG.cpp:226-252) and gamma.
// Returns false when the solve stops (the lead thread records why in stop;
// the next launch turns it into done).
__device__ inline bool pcg_begin(int c, const double *__restrict__ pin, CgState *st, double *lds,
                                 double &gamma) {
  const int it = st->it0 + c, itmax = st->itmax;
  const double eps = st->eps;
  const double gcg = reduce_partials(pin, gridDim.x, 1, 0, lds);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  gamma = 0.;
  int done = 0;
  if (it == 1) {
    const double eps2 = eps > 0 ? eps * eps * gcg : eps * eps;
    if (gcg != gcg) done = 4;
    else if (gcg < 1e-30) done = 2;
    if (lead) {
      st->gcg0 = gcg; st->eps2 = eps2; st->gcg_prev[1] = gcg;
      if (done) { st->stop = done; st->iters = 0; }
    }
  } else {
    const double gcgp = st->gcg_prev[(it - 1) & 1];
    gamma = gcg / gcgp;
    if (gcg != gcg) done = 4;
    else if (gcg < st->eps2) done = 1;
    else if (it - 1 >= itmax) done = 3;
    if (lead) {
      st->gcg_prev[it & 1] = gcg;
      if (done) { st->stop = done; st->iters = it - 1; }
    }
  }
  return done == 0;
}

// this block's row range for the row-wise vector kernels
__device__ inline void block_rows(int64_t n, int64_t &r0, int64_t &r1) {
  const int64_t G = gridDim.x, lb = xcd_block_id();
  r0 = (lb * n) / G;
  r1 = ((lb + 1) * n) / G;
}

// the 4-aligned middle [a4, b4) of [r0, r1), for 16-byte FP32 accesses
// (vectors start 256-byte aligned); head [r0, a4) and tail [b4, r1) go scalar
__device__ inline void block_rows4(int64_t r0, int64_t r1, int64_t &a4, int64_t &b4) {
  a4 = (r0 + 3) & ~int64_t(3);
  if (a4 > r1) a4 = r1;
  b4 = r1 & ~int64_t(3);
  if (b4 < a4) b4 = a4;
}

// K1: the direction H = gamma H_prev - Cg, written by an elementwise pass so
// that the mat-vec gathers one vector (forming H on the fly from two gathered
// vectors per column made the cube(160) mat-vec 17 % slower and the
// [P2,P2,P2] block mat-vec 50 % slower).
// It also applies the last update's step x += rho H_prev (ConjugueGradient's
// x += ro*h, CG.cpp:241), moved here from the update kernel: this pass reads
// H_prev anyway, so the step costs 16 instead of 24 bytes per row.  x enters
// nothing else of the iteration; the stopping launch still applies it.
__global__ __launch_bounds__(kBlock) void k_pcg_dirvec(int c, int64_t n,
                                                       const double *__restrict__ Hold,
                                                       const double *__restrict__ dinv,
                                                       const double *__restrict__ G,
                                                       double *__restrict__ Hnew,
                                                       double *__restrict__ x,
                                                       const double *__restrict__ pin,
                                                       double *__restrict__ pout, CgState *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done) return;  // set by an earlier launch only: every block of this one agrees
  double gamma;
  const bool go = pcg_begin(c, pin, st, lds, gamma);
  const bool step = st->it0 + c > 1;  // an update ran before this direction
  const double rho = st->rho;
  int64_t r0, r1;
  block_rows(n, r0, r1);
  if (go) {
    // the new direction and its partial (g, h), the numerator of rho; C g is
    // formed here from D^-1 and g (the update keeps no copy of it), the
    // product rounded on its own as the stored copy was
    double part = 0.;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
      const double h = Hold[r];
      if (step) x[r] += rho * h;
      const double g = G[r];
      const double hn = gamma * h - __dmul_rn(dinv[r], g);
      Hnew[r] = hn;
      part += g * hn;
    }
    const double gh = block_sum<kBlock>(part, lds);
    if (threadIdx.x == 0) pout[2 * xcd_block_id() + 1] = gh;
  } else if (step) {
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) x[r] += rho * Hold[r];
  }
}

// K2: AH = A H, partials (H, AH), (g, H)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_pcg_matvec(MatFmt<double> M, const double *__restrict__ H,
                                               double *__restrict__ AH, double *__restrict__ pout,
                                               CgState *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->stop) {  // the direction kernel stopped the solve: publish it
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = st->stop;
    return;
  }
  if (st->done) return;
  double p_hah = 0.;
  auto epi = [&](int64_t r, double acc, double h, bool row) {
    if (row) AH[r] = acc;
    p_hah += (row ? h : 0.) * acc;
  };
  if constexpr (KIND == kSell16) {
    // gathers through a buffer resource over H (n < 2^28 in this layout)
    const auto hr = slice_rsrc(H, (uint32_t)(M.n * 8));
    matvec_rows_pre<KIND, double, double>(
        M, [&](int32_t j) { return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(hr, j * 8, 0, 0)); },
        [&](int64_t r) { return H[r]; }, epi);
  } else {
    matvec_rows_pre<KIND, double, double>(M, [&](int32_t j) { return H[j]; }, [&](int64_t r) { return H[r]; },
                                          epi);
  }
  const double a = block_sum<kBlock>(p_hah, lds);
  if (threadIdx.x == 0) pout[2 * xcd_block_id()] = a;  // (g, h) comes from the direction kernel
}

__global__ __launch_bounds__(kBlock) void k_pcg_update(
    int64_t n, const double *__restrict__ AH, const double *__restrict__ dinv,
    double *__restrict__ G, const double *__restrict__ pin, double *__restrict__ pout, CgState *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done) return;
  const int P = gridDim.x;
  const double hah = reduce_partials(pin, P, 2, 0, lds);
  const double gh = reduce_partials(pin, P, 2, 1, lds);
  const double rho = -gh / hah;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rho = rho;  // x += rho H: next direction kernel
  int64_t r0, r1;
  block_rows(n, r0, r1);
  double part = 0.;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const double g = G[r] + rho * AH[r];
    G[r] = g;
    const double c = __dmul_rn(dinv[r], g);  // C g, formed again by the next direction kernel
    part += g * c;
  }
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// ---------------------------------------------------------------- residual check
// r = b - A x (FP64); partials (r,r), (b,b), (r, D^-1 r)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_residual(MatFmt<double> M, const double *__restrict__ x,
                                             const double *__restrict__ b,
                                             const double *__restrict__ dinv,
                                             double *__restrict__ rout, double *__restrict__ pout) {
  __shared__ double lds[kBlock / kWave];
  double prr = 0., pbb = 0., prc = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return x[j]; },
      [&](int64_t r, double acc) {
        const double rr = b[r] - acc;
        if (rout) rout[r] = rr;
        prr += rr * rr;
        pbb += b[r] * b[r];
        prc += rr * (dinv[r] * rr);
      });
  const double a = block_sum<kBlock>(prr, lds);
  const double c = block_sum<kBlock>(pbb, lds);
  const double e = block_sum<kBlock>(prc, lds);
  if (threadIdx.x == 0) {
    const int lb = xcd_block_id();
    pout[3 * lb] = a; pout[3 * lb + 1] = c; pout[3 * lb + 2] = e;
  }
}

// ---------------------------------------------------------------- affine CG
// LinearCG / AffineCG with the matrix product as the operator:
// ConjuguedGradient2 (src/femlib/MatriceCreuse.hpp:180-240).  Iteration it:
//   k_cg2_step:   (Cg, g) of the last update, stop test, gamma = g2/g2p;
//                 h = gamma h - Cg (it > 0); x += rop h
//   k_cg2_dir:    Ah = A x - b - g; partials (h, Ah), (g, h)
//   k_cg2_update: ro = -(g, h) rop / (h, Ah); x += (ro - rop) h;
//                 g += (ro / rop) Ah; Cg = C g; partial (Cg, g)
// C is the identity (dinv null) or the Jacobi diagonal inverse.  Scalars
// written by a kernel's lead thread are kept by parity of the iteration, so
// no block reads a slot another block of the same kernel writes.
struct Cg2State {
  double g2[2];    // (Cg, g) reduced by the step kernel of iteration it, slot it & 1
  double ro[2];    // ro of iteration it, slot it & 1
  double g20, reps2, g2last;
  int done;        // 0 run, 1 converged, 2 g2_0 < 1e-30, 3 itmax, 5 (h,Ah)^2 < 1e-60, 6 ro NaN
  int iters;
};

// g = A x - b, Cg = C g, h = -Cg; partial (Cg, g)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_cg2_init(MatFmt<double> M, const double *__restrict__ x,
                                             const double *__restrict__ b,
                                             const double *__restrict__ dinv, double *__restrict__ g,
                                             double *__restrict__ Cg, double *__restrict__ h,
                                             double *__restrict__ pout, Cg2State *st) {
  __shared__ double lds[kBlock / kWave];
  double part = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return x[j]; },
      [&](int64_t r, double acc) {
        const double gv = b ? acc - b[r] : acc;
        const double c = dinv ? dinv[r] * gv : gv;
        g[r] = gv; Cg[r] = c; h[r] = -c;
        part += c * gv;
      });
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) { st->done = 0; st->iters = 0; }
}

__global__ __launch_bounds__(kBlock) void k_cg2_step(int it, int itmax, double eps, int64_t n,
                                                     double *__restrict__ h,
                                                     const double *__restrict__ Cg,
                                                     double *__restrict__ x,
                                                     const double *__restrict__ pin, Cg2State *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done) return;
  const double g2 = reduce_partials(pin, gridDim.x, 1, 0, lds);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  double gamma = 0., rop = 1.;
  if (it == 0) {
    int done = 0;
    if (g2 != g2) done = 6;  // the reference's first ro would be NaN
    else if (g2 < 1e-30) done = 2;
    if (lead) {
      st->g20 = g2; st->g2last = g2; st->g2[0] = g2;
      st->reps2 = eps > 0 ? eps * eps * g2 : -eps;
      if (done) { st->done = done; st->iters = 0; }
    }
    if (done) return;
  } else {
    int done = 0;
    if (g2 < st->reps2) done = 1;
    else if (it - 1 >= itmax) done = 3;
    if (lead) {
      st->g2[it & 1] = g2; st->g2last = g2;
      if (done) { st->done = done; st->iters = it; }
    }
    if (done) return;
    gamma = g2 / st->g2[(it - 1) & 1];
    rop = st->ro[(it - 1) & 1];
  }
  int64_t r0, r1;
  block_rows(n, r0, r1);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double hv = h[r];
    if (it > 0) { hv = hv * gamma - Cg[r]; h[r] = hv; }
    x[r] += rop * hv;
  }
}

template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_cg2_dir(MatFmt<double> M, const double *__restrict__ x,
                                            const double *__restrict__ b,
                                            const double *__restrict__ g,
                                            const double *__restrict__ h, double *__restrict__ Ah,
                                            double *__restrict__ pout, const Cg2State *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done) return;
  double p_hah = 0., p_gh = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return x[j]; },
      [&](int64_t r, double acc) {
        const double a = (b ? acc - b[r] : acc) - g[r];
        Ah[r] = a;
        p_hah += h[r] * a;
        p_gh += g[r] * h[r];
      });
  const double a = block_sum<kBlock>(p_hah, lds);
  const double c = block_sum<kBlock>(p_gh, lds);
  if (threadIdx.x == 0) {
    const int lb = xcd_block_id();
    pout[2 * lb] = a;
    pout[2 * lb + 1] = c;
  }
}

__global__ __launch_bounds__(kBlock) void k_cg2_update(
    int it, int64_t n, const double *__restrict__ h, const double *__restrict__ Ah,
    const double *__restrict__ dinv, double *__restrict__ x, double *__restrict__ g,
    double *__restrict__ Cg, const double *__restrict__ pin, double *__restrict__ pout,
    Cg2State *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done) return;
  const int P = gridDim.x;
  const double hah = reduce_partials(pin, P, 2, 0, lds);
  const double gh = reduce_partials(pin, P, 2, 1, lds);
  const double rop = it == 0 ? 1. : st->ro[(it - 1) & 1];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (hah * hah < 1e-60) {
    if (lead) { st->done = 5; st->iters = it; }
    return;
  }
  const double ro = -gh * rop / hah;
  if (ro != ro) {
    if (lead) { st->done = 6; st->iters = it; }
    return;
  }
  if (lead) st->ro[it & 1] = ro;
  const double dx = ro - rop, dg = ro / rop;
  int64_t r0, r1;
  block_rows(n, r0, r1);
  double part = 0.;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    x[r] += dx * h[r];
    const double gv = g[r] + dg * Ah[r];
    g[r] = gv;
    const double c = dinv ? dinv[r] * gv : gv;
    Cg[r] = c;
    part += c * gv;
  }
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// ---------------------------------------------------------------- GMRES
// Restarted right-preconditioned GMRES with modified Gram-Schmidt
// (SolverGMRES::dosolver + fgmres, VirtualSolverCG.hpp:236-255, CG.cpp:347-517).
// Per inner iteration it: k_gm_prec (V_it = w / ||w|| and Vp_it = C V_it),
// k_gm_matvec (w = A Vp_it, partial (w, V_0)), then one k_gm_mgs per basis
// vector i <= it (h_i from the previous partials, w -= h_i V_i, partial
// (w, V_i+1) or (w, w)); the Hessenberg column, its Givens rotations and the
// convergence test are the reference's scalar code on the host.

// partial (r, r) of the right-hand side off the tgv rows: ||b|| of fgmres
__global__ __launch_bounds__(kBlock) void k_gm_normb(int64_t n, const double *__restrict__ b,
                                                     const uint8_t *__restrict__ wcl,
                                                     double *__restrict__ pout) {
  __shared__ double lds[kBlock / kWave];
  int64_t r0, r1;
  block_rows(n, r0, r1);
  double part = 0.;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const double v = (wcl && wcl[r]) ? 0. : b[r];
    part += v * v;
  }
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// z = -(A u - b) (fgmres' ri, pCl the identity); partial (z, z)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_gm_residual(MatFmt<double> M, const double *__restrict__ u,
                                                const double *__restrict__ b,
                                                double *__restrict__ z,
                                                double *__restrict__ pout) {
  __shared__ double lds[kBlock / kWave];
  double part = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return u[j]; },
      [&](int64_t r, double acc) {
        const double v = (acc + -1. * b[r]) * -1.;
        z[r] = v;
        part += v * v;
      });
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// V = inv * w; Vp = C V (dinv null: the identity)
__global__ __launch_bounds__(kBlock) void k_gm_prec(int64_t n, double inv,
                                                    const double *__restrict__ w,
                                                    const double *__restrict__ dinv,
                                                    double *__restrict__ V, double *__restrict__ Vp) {
  int64_t r0, r1;
  block_rows(n, r0, r1);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const double v = inv * w[r];
    V[r] = v;
    Vp[r] = dinv ? dinv[r] * v : v;
  }
}

// w = A Vp; partial (w, V_0)
template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_gm_matvec(MatFmt<double> M, const double *__restrict__ Vp,
                                              const double *__restrict__ V0,
                                              double *__restrict__ w, double *__restrict__ pout) {
  __shared__ double lds[kBlock / kWave];
  double part = 0.;
  matvec_rows<KIND, double, double>(
      M, [&](int32_t j) { return Vp[j]; },
      [&](int64_t r, double acc) {
        w[r] = acc;
        part += acc * V0[r];
      });
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// h_i = (w, V_i) from the partials; w -= h_i V_i; partial (w, Vnext), or
// (w, w) when Vnext is null (the last basis vector)
__global__ __launch_bounds__(kBlock) void k_gm_mgs(int i, int64_t n, const double *__restrict__ Vi,
                                                   const double *__restrict__ Vnext,
                                                   double *__restrict__ w,
                                                   const double *__restrict__ pin,
                                                   double *__restrict__ pout,
                                                   double *__restrict__ hcol) {
  __shared__ double lds[kBlock / kWave];
  const double h = reduce_partials(pin, gridDim.x, 1, 0, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) hcol[i] = h;
  const double a = -h;
  int64_t r0, r1;
  block_rows(n, r0, r1);
  double part = 0.;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    const double v = w[r] + a * Vi[r];
    w[r] = v;
    part += v * (Vnext ? Vnext[r] : v);
  }
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// x = (sum_i y_i Vp_i) + x0, the sum taken in order i = 0..k-1
__global__ __launch_bounds__(kBlock) void k_gm_update(int64_t n, int k, const double *__restrict__ y,
                                                      const double *const *__restrict__ Vp,
                                                      const double *__restrict__ x0,
                                                      double *__restrict__ x) {
  int64_t r0, r1;
  block_rows(n, r0, r1);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
    double acc = 0.;
    for (int i = 0; i < k; ++i) acc += y[i] * Vp[i][r];
    x[r] = acc + x0[r];
  }
}

// plain mat-vec for measurement: y = A x
template <int KIND, class V, class X>
__global__ FFHIP_MATVEC_ATTR void k_matvec(MatFmt<V> M, const X *__restrict__ x, X *__restrict__ y) {
  if constexpr (KIND == kSell16 && sizeof(X) == 8) {
    const auto xr = slice_rsrc(x, (uint32_t)(M.n * 8));  // n < 2^28 in this layout
    matvec_rows<KIND, V, X>(
        M, [&](int32_t j) { return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b64(xr, j * 8, 0, 0)); },
        [&](int64_t r, X acc) { y[r] = acc; });
  } else {
    matvec_rows<KIND, V, X>(M, [&](int32_t j) { return x[j]; }, [&](int64_t r, X acc) { y[r] = acc; });
  }
}

// ---------------------------------------------------------------- MIXED (FP32 inner)
// Mixed-precision state lives in CgState: sigma (scale of r32), rho32 (last
// (r32,r32)), and the per-parity (r32,r32) history below.
struct MixState {
  double rho_prev[2];   // (r32, r32) at the start of inner iteration, by parity
  double cbeta;         // coefficient of p_old in the first iteration after an update
  int first_it;         // inner iteration that uses cbeta
  int since_update;     // host-side bookkeeping copy
};

// K_m1: p = r32 + cb*p_old (on the fly for every gathered column); q = Ahat p;
// partial (p, q).  Requests a reliable update when (r32, r32) fell by
// kMixUpdate
// since the last update, or the recurrence says the FP64 criterion is met.
// Start of inner iteration it = it0 + c: rho = (r32, r32) from the last
// update's partials, the requests (reliable update, itmax, NaN) and the
// coefficient of p_old.  Returns false when the chunk stops.
__device__ inline bool mix_begin(int c, const double *__restrict__ pin, CgState *st, MixState *ms,
                                 double *lds, float &cbf) {
  const int it = st->it0 + c, itmax = st->itmax;
  const double rho = reduce_partials(pin, gridDim.x, 1, 0, lds);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double sig = st->sigma;
  int req = 0;
  if (rho != rho) req = 4;
  else if (rho * sig * sig < st->eps2 || rho < kMixUpdate) req = 1;  // check on the FP64 residual
  else if (it - 1 >= itmax) req = 3;
  if (req) {
    if (lead) {
      if (req == 4) st->done = 4;
      else if (req == 3) st->done = 3;
      else st->need_update = 1;
      // (r32, r32) of the last executed iteration: the denominator of beta once
      // the newest recurrence residual is replaced by the FP64 one
      st->rho32 = ms->rho_prev[(it - 1) & 1];
      st->iters = it - 1;
    }
    return false;
  }
  const double cb = (it == ms->first_it) ? ms->cbeta : rho / ms->rho_prev[(it - 1) & 1];
  if (lead) { ms->rho_prev[it & 1] = rho; st->iters = it - 1; }
  cbf = (float)cb;
  return true;
}

// K_m1: p = r32 + cb p_old (elementwise); K_m2: q = Ahat p, partial (p, q).
// Requests a reliable update when (r32, r32) fell by kMixUpdate since the last
// update, or the recurrence says the FP64 criterion is met.
__global__ __launch_bounds__(kBlock) void k_mix_dirvec(int c, int64_t n,
                                                       const float *__restrict__ pold,
                                                       const float *__restrict__ r32,
                                                       float *__restrict__ pnew,
                                                       const double *__restrict__ pin, CgState *st,
                                                       MixState *ms) {
  __shared__ double lds[kBlock / kWave];
  if (st->done || st->need_update) return;
  float cbf;
  if (!mix_begin(c, pin, st, ms, lds, cbf)) return;
  int64_t r0, r1, a4, b4;
  block_rows(n, r0, r1);
  block_rows4(r0, r1, a4, b4);
  for (int64_t r = r0 + threadIdx.x; r < a4; r += kBlock) pnew[r] = r32[r] + cbf * pold[r];
  for (int64_t i = a4 / 4 + threadIdx.x; i < b4 / 4; i += kBlock) {
    const float4 rv = reinterpret_cast<const float4 *>(r32)[i];
    const float4 po = reinterpret_cast<const float4 *>(pold)[i];
    float4 pn;
    pn.x = rv.x + cbf * po.x; pn.y = rv.y + cbf * po.y;
    pn.z = rv.z + cbf * po.z; pn.w = rv.w + cbf * po.w;
    reinterpret_cast<float4 *>(pnew)[i] = pn;
  }
  for (int64_t r = b4 + threadIdx.x; r < r1; r += kBlock) pnew[r] = r32[r] + cbf * pold[r];
}

template <int KIND>
__global__ FFHIP_MATVEC_ATTR void k_mix_matvec(MatFmt<float> M, const float *__restrict__ p,
                                               float *__restrict__ q, double *__restrict__ pout,
                                               const CgState *st) {
  __shared__ double lds[kBlock / kWave];
  if (st->done || st->need_update) return;
  double part = 0.;
  matvec_rows<KIND, float, float>(
      M, [&](int32_t j) { return p[j]; },
      [&](int64_t r, float acc) {
        q[r] = acc;
        part += (double)p[r] * (double)acc;
      });
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// K_m3: alpha = rho/(p,q); y32 += alpha p; r32 -= alpha q; partial (r32, r32)
__global__ __launch_bounds__(kBlock) void k_mix_update(
    int c, int64_t n, const float *__restrict__ p, const float *__restrict__ q,
    float *__restrict__ y32, float *__restrict__ r32, const double *__restrict__ pin,
    double *__restrict__ pout, CgState *st, MixState *ms) {
  __shared__ double lds[kBlock / kWave];
  if (st->done || st->need_update) return;
  const int it = st->it0 + c;
  const int P = gridDim.x;
  const double pq = reduce_partials(pin, P, 1, 0, lds);
  const double alpha = ms->rho_prev[it & 1] / pq;
  const float af = (float)alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->iters = it;
  int64_t r0, r1, a4, b4;
  block_rows(n, r0, r1);
  block_rows4(r0, r1, a4, b4);
  double part = 0.;
  auto one = [&](int64_t r) {
    y32[r] += af * p[r];
    const float rr = r32[r] - af * q[r];
    r32[r] = rr;
    part += (double)rr * (double)rr;
  };
  for (int64_t r = r0 + threadIdx.x; r < a4; r += kBlock) one(r);
  for (int64_t i = a4 / 4 + threadIdx.x; i < b4 / 4; i += kBlock) {
    const float4 pv = reinterpret_cast<const float4 *>(p)[i];
    const float4 qv = reinterpret_cast<const float4 *>(q)[i];
    float4 yv = reinterpret_cast<const float4 *>(y32)[i];
    float4 rv = reinterpret_cast<const float4 *>(r32)[i];
    yv.x += af * pv.x; yv.y += af * pv.y; yv.z += af * pv.z; yv.w += af * pv.w;
    rv.x -= af * qv.x; rv.y -= af * qv.y; rv.z -= af * qv.z; rv.w -= af * qv.w;
    reinterpret_cast<float4 *>(y32)[i] = yv;
    reinterpret_cast<float4 *>(r32)[i] = rv;
    part += (double)rv.x * (double)rv.x;
    part += (double)rv.y * (double)rv.y;
    part += (double)rv.z * (double)rv.z;
    part += (double)rv.w * (double)rv.w;
  }
  for (int64_t r = b4 + threadIdx.x; r < r1; r += kBlock) one(r);
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

// end of a chunk of CG launches: the next chunk starts `chunk` iterations on
__global__ void k_cg_advance(CgState *st, int chunk) {
  if (blockIdx.x == 0 && threadIdx.x == 0) st->it0 += chunk;
}

// fold the FP32 correction into the FP64 solution: x += D^-1/2 (sigma y32); y32 = 0
__global__ void k_mix_fold(int64_t n, double sigma, const double *__restrict__ dsc,
                           float *__restrict__ y32, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += dsc[i] * (sigma * (double)y32[i]);
  y32[i] = 0.f;
}

// new scaled residual r32 = D^-1/2 r / sigma (r = b - A x in FP64); partial (r32,r32)
__global__ __launch_bounds__(kBlock) void k_mix_rescale(
    int64_t n, double inv_sigma, const double *__restrict__ dsc,
    const double *__restrict__ r, float *__restrict__ r32, double *__restrict__ pout) {
  __shared__ double lds[kBlock / kWave];
  int64_t r0, r1;
  block_rows(n, r0, r1);
  double part = 0.;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    const float v = (float)(dsc[i] * r[i] * inv_sigma);
    r32[i] = v;
    part += (double)v * (double)v;
  }
  const double tot = block_sum<kBlock>(part, lds);
  if (threadIdx.x == 0) pout[xcd_block_id()] = tot;
}

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// CG iterations launched per host poll of the device state; one chunk is
// captured once into a hipGraph and replayed (even, so the two-buffer
// parities of a chunk repeat)
constexpr int kChunk = 16;

// stream-captures the launches of f() into an executable graph
template <class F>
hipGraphExec_t capture_graph(hipStream_t s, F f) {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  FFHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  try {
    f();
  } catch (...) {  // the stream must leave capture mode whatever happened
    if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    throw;
  }
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(s, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    FFHIP_CHECK(le);
    FFHIP_CHECK(ce);
  }
  const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  FFHIP_CHECK(ie);
  return e;
}

}  // namespace ffhip

using namespace ffhip;

struct ffhip_cg_s {
  ffhip_ctx ctx = nullptr;
  ffhip_mat A = nullptr;
  int mode = 0;
  int fmt = kSell;
  int far_pending = 0;  // kSell16 chosen before the fill has checked every column delta
  int64_t n = 0, S = 0, ents = 0;
  int grid = 0;
  // kSell layout (kCsrV streams A's own rowptr/colidx)
  int64_t *sptr = nullptr;
  int32_t *slen = nullptr, *scol = nullptr;
  int16_t *scol16 = nullptr;  // kSell16
  int64_t *sptr16 = nullptr;  // kSell16 slice offsets of scol16
  int32_t *bcol = nullptr;    // kBsr3 block columns (slices over sorted node rows)
  int32_t *bperm = nullptr;   // kBsr3 node row of each sorted position
  int32_t *bwave = nullptr;   // kBsr3 first slice of each mat-vec wave (grid x 4 + 1)
  // values in the solver's layout; for kCsrV, sval is A's value array (borrowed)
  double *sval = nullptr;    // A
  double *svalT = nullptr;   // A^T (built on first transposed solve)
  float *sval32 = nullptr;   // D^-1/2 A D^-1/2 (MIXED)
  float *sval32T = nullptr;
  // Jacobi / tgv
  double *dinv = nullptr, *dsc = nullptr;
  uint8_t *wcl = nullptr;
  double tgvm = 0.;
  int ntgv = 0;
  SetupRb *rb = nullptr;  // set-up figures for the host (one copy per set-up)
  // work
  double *xw = nullptr, *bw = nullptr, *G = nullptr, *CG = nullptr, *H[2] = {nullptr, nullptr},
         *AH = nullptr;
  float *r32 = nullptr, *p32[2] = {nullptr, nullptr}, *q32 = nullptr, *y32 = nullptr;
  double *pA = nullptr, *pB = nullptr, *pC = nullptr;
  CgState *st = nullptr;
  MixState *ms = nullptr;
  Cg2State *st2 = nullptr;  // affine CG (LinearCG)
  // GMRES: Krylov basis V_i and C V_i, allocated vector by vector as a solve
  // first reaches them (as the reference's KN<KN<K>> are) and kept; the
  // device table of the C V_i pointers; w, x0, the Hessenberg column, y
  std::vector<double *> gmV, gmVp;
  double **gmVpDev = nullptr;
  int gmCap = 0;  // entries of gmVpDev, gmH and gmY
  double *gmW = nullptr, *gmX0 = nullptr, *gmH = nullptr, *gmY = nullptr;
  void gm_basis(int k) {
    if (k > gmCap) {
      const int cap = std::max(k, 2 * gmCap);
      double **p = dalloc<double *>(cap);
      double *h = dalloc<double>(cap + 1), *y = dalloc<double>(cap + 1);
      dfree(gmVpDev); dfree(gmH); dfree(gmY);
      gmVpDev = p; gmH = h; gmY = y; gmCap = cap;
      FFHIP_CHECK(hipMemcpy(gmVpDev, gmVp.data(), sizeof(double *) * gmVp.size(),
                            hipMemcpyHostToDevice));
    }
    while ((int)gmV.size() < k) {
      gmV.push_back(dalloc<double>(n));
      gmVp.push_back(dalloc<double>(n));
      FFHIP_CHECK(hipMemcpy(gmVpDev + gmVp.size() - 1, &gmVp.back(), sizeof(double *),
                            hipMemcpyHostToDevice));
    }
    if (!gmW) { gmW = dalloc<double>(n); gmX0 = dalloc<double>(n); }
  }
  // captured chunks of CG launches, keyed by (mode, value array, parity of the
  // chunk's first iteration); kernel by kernel when eager or profiling (event
  // records captured in a graph carry no timestamps hipEventElapsedTime reads)
  struct Chunk {
    int kind;
    const void *val;
    int par;
    hipGraphExec_t exec;
  };
  std::vector<Chunk> graphs;
  int use_graph = 1;
  template <class F>
  hipGraphExec_t graph(int kind, const void *val, int par, F launch) {
    if (!use_graph || profile) return nullptr;
    for (const Chunk &c : graphs)
      if (c.kind == kind && c.val == val && c.par == par) return c.exec;
    graphs.push_back({kind, val, par, capture_graph(ctx->stream, launch)});
    return graphs.back().exec;
  }
  // per-kernel timing (measurement only): 4 events per iteration of a chunk,
  // around the direction, mat-vec and update kernels
  int profile = 0;
  std::vector<hipEvent_t> ev;
  double t_dv = 0., t_mv = 0., t_up = 0.;
  int n_timed = 0;
  void ensure_events(int k) {
    while ((int)ev.size() < k) {
      hipEvent_t e;
      FFHIP_CHECK(hipEventCreate(&e));
      ev.push_back(e);
    }
  }
  void mark(int slot) {
    if (profile) FFHIP_CHECK(hipEventRecord(ev[slot], ctx->stream));
  }
  // after a synchronised chunk: accumulate the iterations [it0, it1] that ran
  void collect(int chunk_first_it, int nits, int last_exec_it) {
    if (!profile) return;
    for (int c = 0; c < nits; ++c) {
      const int it = chunk_first_it + c;
      if (it > last_exec_it) break;
      float d = 0.f, a = 0.f, b = 0.f;
      FFHIP_CHECK(hipEventElapsedTime(&d, ev[4 * c], ev[4 * c + 1]));
      FFHIP_CHECK(hipEventElapsedTime(&a, ev[4 * c + 1], ev[4 * c + 2]));
      FFHIP_CHECK(hipEventElapsedTime(&b, ev[4 * c + 2], ev[4 * c + 3]));
      t_dv += d; t_mv += a; t_up += b; ++n_timed;
    }
  }
  template <class V>
  MatFmt<V> fmt_of(const V *val) const {
    MatFmt<V> M;
    M.n = n;
    M.S = S;
    M.col16 = scol16;
    M.ptr16 = sptr16;
    M.nb = 0;
    M.ents = ents;
    M.perm = bperm;
    if (fmt == kSell || fmt == kSell16) { M.ptr = sptr; M.len = slen; M.col = scol; }
    else if (fmt == kBsr3) { M.ptr = sptr; M.len = slen; M.col = bcol; M.nb = n / 3; }
    else { M.ptr = A->rowptr; M.len = nullptr; M.col = A->colidx; }
    M.wstart = fmt == kBsr3 ? bwave : nullptr;
    M.nw = fmt == kBsr3 && bwave ? grid * kWavesPerBlock : 0;
    M.val = val;
    return M;
  }
};

// launch a layout-templated mat-vec kernel with the solver's layout
#define FFHIP_LAUNCH_FMT(cg, kern, grid, stream, ...)                        \
  do {                                                                       \
    if ((cg)->fmt == kSell) kern<kSell><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); \
    else if ((cg)->fmt == kSell16) kern<kSell16><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); \
    else if ((cg)->fmt == kBsr3) kern<kBsr3><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); \
    else kern<kCsrV><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__);           \
  } while (0)

#define FFHIP_LAUNCH_MATVEC(cg, T, M, x, y, stream)                                          \
  do {                                                                                      \
    if ((cg)->fmt == kSell)                                                                 \
      k_matvec<kSell, T, T><<<(cg)->grid, kBlock, 0, (stream)>>>(M, (const T *)(x), (T *)(y)); \
    else if ((cg)->fmt == kSell16)                                                          \
      k_matvec<kSell16, T, T><<<(cg)->grid, kBlock, 0, (stream)>>>(M, (const T *)(x), (T *)(y)); \
    else if ((cg)->fmt == kBsr3)                                                            \
      k_matvec<kBsr3, T, T><<<(cg)->grid, kBlock, 0, (stream)>>>(M, (const T *)(x), (T *)(y)); \
    else                                                                                    \
      k_matvec<kCsrV, T, T><<<(cg)->grid, kBlock, 0, (stream)>>>(M, (const T *)(x), (T *)(y)); \
  } while (0)

// the set-up figures to the host (one copy, after the stream's work so far);
// the tgv figures into the solver
static SetupRb setup_readback(ffhip_cg cg) {
  SetupRb h;
  FFHIP_CHECK(hipMemcpyAsync(&h, cg->rb, sizeof(h), hipMemcpyDeviceToHost, cg->ctx->stream));
  FFHIP_CHECK(hipStreamSynchronize(cg->ctx->stream));
  cg->tgvm = h.ttgv;
  cg->ntgv = (int)h.ntgv;
  return h;
}

// Layout choice: sliced ELL unless its slices pad the matrix by more than 25 %
// (rows of very different lengths), then CSR with 16 lanes per row.
static void cg_plan_layout(ffhip_cg cg, int layout, bool index32) {
  ffhip_ctx ctx = cg->ctx;
  ffhip_mat A = cg->A;
  hipStream_t s = ctx->stream;
  cg->S = (A->n + kSlice - 1) / kSlice;
  if (A->nodeptr && A->ncomp == 3 &&
      (layout == FFHIP_CG_LAYOUT_AUTO || layout == FFHIP_CG_LAYOUT_BSR3)) {
    // vector P2: the node rows are 3x3 blocks; slices over node rows sorted by
    // length inside windows of 4096 rows
    cg->fmt = kBsr3;
    const int64_t nb = A->n / 3, S3 = (nb + kSlice - 1) / kSlice;
    uint32_t *keys = dalloc<uint32_t>(nb), *skeys = dalloc<uint32_t>(nb);
    int32_t *vals = dalloc<int32_t>(nb);
    cg->bperm = dalloc<int32_t>(nb);
    k_bsr_keys<<<nblk(nb, 256), 256, 0, s>>>(nb, A->nodeptr, keys, vals);
    size_t tb = 0;
    FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, skeys, vals, cg->bperm, (int)nb,
                                                   0, 32, s));
    char *tmp = dalloc<char>(tb);
    FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, skeys, vals, cg->bperm, (int)nb, 0,
                                                   32, s));
    cg->sptr = dalloc<int64_t>(S3 + 1);
    cg->slen = dalloc<int32_t>(S3);
    int64_t *cnt = dalloc<int64_t>(S3 + 1);
    FFHIP_CHECK(hipMemsetAsync(cnt + S3, 0, 8, s));
    k_bsr_len<<<nblk(S3, 256), 256, 0, s>>>(nb, S3, A->nodeptr, cg->bperm, cg->slen, cnt);
    exclusive_scan_i64(ctx, cnt, cg->sptr, S3 + 1);
    k_setup_sizes<<<1, 1, 0, s>>>(cg->rb, cg->sptr + S3, nullptr, nullptr);
    cg->ents = setup_readback(cg).ents;
    dfree(cnt); dfree(tmp); dfree(keys); dfree(skeys); dfree(vals);
    cg->bcol = dalloc<int32_t>(cg->ents);
  } else if (layout == FFHIP_CG_LAYOUT_BSR3) {
    FFHIP_REQUIRE(false, FFHIP_ERR_UNSUPPORTED, "ffhip_cg_create: 3x3 blocks need a [P2,P2,P2] matrix");
  } else if (layout != FFHIP_CG_LAYOUT_CSR) {
    // slice widths and offsets and the offsets of the packed 16-bit deltas,
    // then one read-back decides among sliced ELL with 16-bit deltas (the fill
    // then checks them), with int32 columns, and CSR
    Arena ar(ctx, 3 * Arena::bytes(cg->S + 1, 8) + Arena::bytes(1, 4));
    cg->sptr = dalloc<int64_t>(cg->S + 1);
    cg->slen = dalloc<int32_t>(cg->S);
    cg->sptr16 = dalloc<int64_t>(cg->S + 1);
    int64_t *cnt = ar.take<int64_t>(cg->S + 1);
    int64_t *cnt4 = ar.take<int64_t>(cg->S + 1);
    int *fard = ar.take<int>(1);
    FFHIP_CHECK(hipMemsetAsync(cnt + cg->S, 0, 8, s));
    FFHIP_CHECK(hipMemsetAsync(cnt4 + cg->S, 0, 8, s));
    FFHIP_CHECK(hipMemsetAsync(fard, 0, sizeof(int), s));
    k_sell_len<<<nblk(cg->S, 256), 256, 0, s>>>(A->n, cg->S, A->rowptr, cg->slen, cnt);
    exclusive_scan_i64(ctx, cnt, cg->sptr, cg->S + 1);
    // 16-bit column deltas unless index32: the fill checks every delta
    // against 2^15 and falls back to int32 columns (far_pending)
    k_sell_len4<<<nblk(cg->S, 256), 256, 0, s>>>(cg->S, cg->slen, cnt4);
    exclusive_scan_i64(ctx, cnt4, cg->sptr16, cg->S + 1);
    k_setup_sizes<<<1, 1, 0, s>>>(cg->rb, cg->sptr + cg->S, cg->sptr16 + cg->S, index32 ? nullptr : fard);
    FFHIP_CHECK(hipGetLastError());
    const SetupRb h = setup_readback(cg);
    cg->ents = h.ents;
    if (layout == FFHIP_CG_LAYOUT_SELL || 4 * cg->ents <= 5 * A->nnz) {
      // 16-bit deltas, and gathers through buffer resources with 32-bit byte offsets
      if (h.far || A->n >= ((int64_t)1 << 28)) {
        cg->fmt = kSell;
        cg->scol = dalloc<int32_t>(cg->ents);
        dfree(cg->sptr16);
      } else {
        cg->fmt = kSell16;
        cg->scol16 = dalloc<int16_t>(h.ents16);
        cg->far_pending = 1;
      }
    } else {
      dfree(cg->sptr); dfree(cg->slen); dfree(cg->sptr16);
      cg->fmt = kCsrV;
    }
  } else {
    cg->fmt = kCsrV;
    setup_readback(cg);  // the tgv figures
  }
  int64_t g;
  if (cg->fmt == kBsr3) {
    g = ((A->n / 3 + kSlice - 1) / kSlice + kWavesPerBlock - 1) / kWavesPerBlock;
  } else if (cg->fmt != kCsrV) {
    g = (cg->S + kWavesPerBlock - 1) / kWavesPerBlock;
  } else {
    cg->ents = A->nnz;
    g = (A->n + kBlock / kLanesPerRow - 1) / (kBlock / kLanesPerRow);
  }
  const int64_t gmax = FFHIP_GRID_PER_CU * (int64_t)ctx->num_cu;
  if (g > gmax) g = gmax;
  if (g >= 64) g -= g % 8;  // whole XCD groups
  if (g < 1) g = 1;
  cg->grid = (int)g;
  if (cg->fmt == kBsr3) {
    const int nw = cg->grid * kWavesPerBlock;
    cg->bwave = dalloc<int32_t>(nw + 1);
    k_bsr_waves<<<nblk(nw + 1, 256), 256, 0, s>>>((A->n / 3 + kSlice - 1) / kSlice, cg->sptr, nw,
                                                  cg->bwave);
    FFHIP_CHECK(hipGetLastError());
  }
}

// The diagonal values and presence flags the Jacobi set-up starts from, and
// the partials of its tgv passes.
struct DiagBufs {
  static constexpr int GB = 512;  // blocks of a tgv pass
  double *d;
  uint8_t *has;
  double *part;
  static size_t bytes(int64_t n) { return Arena::bytes(n, 8) + Arena::bytes(n, 1) + Arena::bytes(2 * GB, 8); }
  DiagBufs(Arena &ar, int64_t n) : d(ar.take<double>(n)), has(ar.take<uint8_t>(n)), part(ar.take<double>(2 * GB)) {}
};
static void cg_jacobi_tail(ffhip_cg cg, const DiagBufs &b);

// values in the solver's layout (A or A^T; FP32 scaled copy for MIXED).  With
// db the sliced-ELL fill also writes the diagonal and the Jacobi set-up runs
// behind it; with read_rb the tgv figures come back with the fill's wait.
static void cg_fill_values(ffhip_cg cg, bool transposed, const DiagBufs *db = nullptr,
                           bool read_rb = false) {
  ffhip_ctx ctx = cg->ctx;
  ffhip_mat A = cg->A;
  hipStream_t s = ctx->stream;
  int64_t *tmap = nullptr;
  if (transposed) {
    tmap = dalloc<int64_t>(A->nnz);
    // transposed values through the symmetric-pattern slot map
    const bool ok = transpose_map(ctx, A, tmap);
    if (!ok) dfree(tmap);
    FFHIP_REQUIRE(ok, FFHIP_ERR_UNSUPPORTED, "transposed solve needs a symmetric pattern");
  }
  const bool mixed = cg->mode == FFHIP_CG_MIXED;
  const bool check_far = !transposed && cg->far_pending && cg->fmt == kSell16;
  double *&dv = transposed ? cg->svalT : cg->sval;
  float *&fv = transposed ? cg->sval32T : cg->sval32;
  const int64_t nval = cg->fmt == kBsr3 ? 9 * cg->ents : cg->ents;
  if (mixed && !fv) fv = dalloc<float>(nval);
  if (cg->fmt == kBsr3) {
    if (!dv) dv = dalloc<double>(nval);
    const int64_t nb = A->n / 3, S3 = (nb + kSlice - 1) / kSlice;
    k_bsr_fill<<<S3, 256, 0, s>>>(
        nb, cg->ents, A->nodeptr, cg->bperm, cg->sptr, cg->slen, A->rowptr, A->colidx, A->vals, tmap,
        cg->dsc, transposed ? nullptr : cg->bcol, dv, mixed ? fv : nullptr);
  } else if (cg->fmt != kCsrV) {
    if (!dv) dv = dalloc<double>(cg->ents);
    k_sell_fill<<<cg->S, 256, 0, s>>>(
        A->n, cg->S, A->rowptr, A->colidx, A->vals, tmap, cg->sptr, cg->slen, cg->dsc,
        transposed ? nullptr : cg->scol, cg->sptr16, transposed ? nullptr : cg->scol16, dv,
        mixed ? fv : nullptr, cg->fmt == kSell16, A->diag, db ? db->d : nullptr,
        db ? db->has : nullptr, check_far ? &cg->rb->far : nullptr);
    if (db) {
      FFHIP_CHECK(hipGetLastError());
      cg_jacobi_tail(cg, *db);
    }
  } else {
    double *dst = nullptr;
    if (transposed) {
      if (!dv) dv = dalloc<double>(cg->ents);
      dst = dv;
    } else {
      dv = A->vals;  // borrowed: A's CSR order is the layout
    }
    if (dst || mixed)
      k_csr_fill<<<nblk(A->n * kLanesPerRow, 256), 256, 0, s>>>(
          A->n, A->rowptr, A->colidx, A->vals, tmap, cg->dsc, dst, mixed ? fv : nullptr);
  }
  FFHIP_CHECK(hipGetLastError());
  SetupRb h{};
  if (read_rb || check_far) FFHIP_CHECK(hipMemcpyAsync(&h, cg->rb, sizeof(h), hipMemcpyDeviceToHost, s));
  FFHIP_CHECK(hipStreamSynchronize(s));
  if (read_rb) {
    cg->tgvm = h.ttgv;
    cg->ntgv = (int)h.ntgv;
  }
  dfree(tmap);
  if (check_far) {
    cg->far_pending = 0;
    if (h.far) {  // some column more than 2^15 from its row: int32 columns, filled again
      dfree(cg->scol16);
      dfree(cg->sptr16);
      cg->fmt = kSell;
      cg->scol = dalloc<int32_t>(cg->ents);
      cg_fill_values(cg, transposed, db, read_rb);
    }
  }
}

// Jacobi scaling and the tgv rows, on the device (the tgv figures reach the
// host with the next setup_readback)
static void cg_setup_jacobi(ffhip_cg cg) {
  const int64_t n = cg->A->n;
  Arena ar(cg->ctx, DiagBufs::bytes(n));
  const DiagBufs b(ar, n);
  k_diag_vals<<<nblk(n, 256), 256, 0, cg->ctx->stream>>>(n, cg->A->diag, cg->A->vals, b.d, b.has);
  cg_jacobi_tail(cg, b);
}

// the tgv passes and the Jacobi scaling from the diagonal in b
static void cg_jacobi_tail(ffhip_cg cg, const DiagBufs &b) {
  hipStream_t s = cg->ctx->stream;
  const int64_t n = cg->A->n;
  const int GB = DiagBufs::GB;
  const double *d = b.d;
  const uint8_t *has = b.has;
  double *part = b.part;
  k_tgv_pass<<<GB, kBlock, 0, s>>>(n, d, has, nullptr, 1, part);
  k_tgv_reduce<<<1, kBlock, 0, s>>>(GB, part, 1, cg->rb);
  k_tgv_pass<<<GB, kBlock, 0, s>>>(n, d, has, &cg->rb->ttgv, 2, part);
  k_tgv_reduce<<<1, kBlock, 0, s>>>(GB, part, 2, cg->rb);
  k_jacobi<<<nblk(n, 256), 256, 0, s>>>(n, d, has, cg->rb, cg->dinv, cg->dsc, cg->wcl);
  FFHIP_CHECK(hipGetLastError());
}

// r = b - A x (rout may be null); the three partial sums of k_residual added on
// the host in block order
static void residual_sums(ffhip_cg cg, const double *val, const double *x, const double *b,
                          double *rout, double out[3]) {
  hipStream_t s = cg->ctx->stream;
  FFHIP_LAUNCH_FMT(cg, k_residual, cg->grid, s, cg->fmt_of(val), x, b, cg->dinv, rout, cg->pC);
  FFHIP_CHECK(hipGetLastError());
  std::vector<double> h(3 * (size_t)cg->grid);
  FFHIP_CHECK(hipMemcpyAsync(h.data(), cg->pC, 24 * cg->grid, hipMemcpyDeviceToHost, s));
  FFHIP_CHECK(hipStreamSynchronize(s));
  out[0] = out[1] = out[2] = 0.;
  for (int i = 0; i < cg->grid; ++i)
    for (int c = 0; c < 3; ++c) out[c] += h[3 * i + c];
}

// residual r = b - A x into cg->G; returns (r, D^-1 r)
static double mix_residual(ffhip_cg cg, const double *sval) {
  double r[3];
  residual_sums(cg, sval, cg->xw, cg->bw, cg->G, r);
  return r[2];
}

// One chunk of FP64 PCG iterations it0 + c, c < kChunk, it0 = 1 (mod kChunk):
// the H buffer parity follows c.
static void launch_pcg_chunk(ffhip_cg cg, const double *sval, bool marks) {
  hipStream_t s = cg->ctx->stream;
  const int G = cg->grid;
  const MatFmt<double> M = cg->fmt_of(sval);
  for (int c = 0; c < kChunk; ++c) {
    if (marks) cg->mark(4 * c);
    k_pcg_dirvec<<<G, kBlock, 0, s>>>(c, cg->n, cg->H[c & 1], cg->dinv, cg->G, cg->H[(c + 1) & 1],
                                      cg->xw, cg->pA, cg->pB, cg->st);
    if (marks) cg->mark(4 * c + 1);
    FFHIP_LAUNCH_FMT(cg, k_pcg_matvec, G, s, M, cg->H[(c + 1) & 1], cg->AH, cg->pB, cg->st);
    if (marks) cg->mark(4 * c + 2);
    k_pcg_update<<<G, kBlock, 0, s>>>(cg->n, cg->AH, cg->dinv, cg->G, cg->pB, cg->pA, cg->st);
    if (marks) cg->mark(4 * c + 3);
  }
  k_cg_advance<<<1, 64, 0, s>>>(cg->st, kChunk);
}

// One chunk of MIXED inner iterations it0 + c, par = it0 & 1.
static void launch_mix_chunk(ffhip_cg cg, const float *sval32, int par, bool marks) {
  hipStream_t s = cg->ctx->stream;
  const int G = cg->grid;
  const MatFmt<float> M32 = cg->fmt_of(sval32);
  for (int c = 0; c < kChunk; ++c) {
    const float *pold = cg->p32[(par + c + 1) & 1];
    float *pnew = cg->p32[(par + c) & 1];
    if (marks) cg->mark(4 * c);
    k_mix_dirvec<<<G, kBlock, 0, s>>>(c, cg->n, pold, cg->r32, pnew, cg->pA, cg->st, cg->ms);
    if (marks) cg->mark(4 * c + 1);
    FFHIP_LAUNCH_FMT(cg, k_mix_matvec, G, s, M32, pnew, cg->q32, cg->pB, cg->st);
    if (marks) cg->mark(4 * c + 2);
    k_mix_update<<<G, kBlock, 0, s>>>(c, cg->n, pnew, cg->q32, cg->y32, cg->r32, cg->pB, cg->pA,
                                      cg->st, cg->ms);
    if (marks) cg->mark(4 * c + 3);
  }
  k_cg_advance<<<1, 64, 0, s>>>(cg->st, kChunk);
}

static hipGraphExec_t pcg_chunk_graph(ffhip_cg cg, const double *sval) {
  return cg->graph(0, sval, 1, [&] { launch_pcg_chunk(cg, sval, false); });
}
static hipGraphExec_t mix_chunk_graph(ffhip_cg cg, const float *sval32, int par) {
  return cg->graph(1, sval32, par, [&] { launch_mix_chunk(cg, sval32, par, false); });
}

// FP32 Jacobi-scaled CG with FP64 reliable updates (see the file comment).
int ffhip_cg_solve_mixed(ffhip_cg cg, const double *sval, const float *sval32, double eps,
                         int itmax, CgState &hs) {
  ffhip_ctx ctx = cg->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = cg->n;
  const int G = cg->grid;
  const int chunk = kChunk;
  hipGraphExec_t gx[2] = {mix_chunk_graph(cg, sval32, 0), mix_chunk_graph(cg, sval32, 1)};
  FFHIP_CHECK(hipMemsetAsync(cg->p32[0], 0, 4 * n, s));
  FFHIP_CHECK(hipMemsetAsync(cg->p32[1], 0, 4 * n, s));
  FFHIP_CHECK(hipMemsetAsync(cg->y32, 0, 4 * n, s));
  memset(&hs, 0, sizeof(hs));
  MixState hm;
  memset(&hm, 0, sizeof(hm));
  double gcg = mix_residual(cg, sval);
  hs.refinements = 1;
  hs.gcg0 = gcg;
  hs.eps2 = eps > 0 ? eps * eps * gcg : eps * eps;
  if (gcg != gcg) { hs.done = 4; return 0; }
  if (gcg < 1e-30) { hs.done = 2; return 0; }
  double sigma = std::sqrt(gcg);
  double rho_last = 0.;
  int it = 1;
  hm.cbeta = 0.;
  hm.first_it = 1;
  for (;;) {
    // (re)start the FP32 recurrence from the FP64 residual
    k_mix_rescale<<<G, kBlock, 0, s>>>(n, 1. / sigma, cg->dsc, cg->G, cg->r32, cg->pA);
    hs.sigma = sigma;
    hs.need_update = 0;
    hs.iters = it - 1;
    hs.it0 = it;
    hs.itmax = itmax;
    hs.eps = eps;
    FFHIP_CHECK(hipMemcpyAsync(cg->st, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(cg->ms, &hm, sizeof(hm), hipMemcpyHostToDevice, s));
    cg->ensure_events(4 * chunk);
    const int par = it & 1;
    for (;;) {
      const int it0 = it;
      if (gx[par]) FFHIP_CHECK(hipGraphLaunch(gx[par], s));
      else launch_mix_chunk(cg, sval32, par, cg->profile != 0);
      it += chunk;
      FFHIP_CHECK(hipGetLastError());
      FFHIP_CHECK(hipMemcpyAsync(&hs, cg->st, sizeof(hs), hipMemcpyDeviceToHost, s));
      FFHIP_CHECK(hipStreamSynchronize(s));
      cg->collect(it0, chunk, hs.iters);
      if (hs.done || hs.need_update) break;
    }
    if (hs.done) {  // NaN or itmax: fold what we have
      k_mix_fold<<<nblk(n, 256), 256, 0, s>>>(n, sigma, cg->dsc, cg->y32, cg->xw);
      return 0;
    }
    // reliable update: fold, recompute the FP64 residual, accept or rescale
    if (hs.iters + 1 == hm.first_it && hs.refinements > 1) {  // no inner progress since the last update
      hs.done = 3;
      k_mix_fold<<<nblk(n, 256), 256, 0, s>>>(n, sigma, cg->dsc, cg->y32, cg->xw);
      return 0;
    }
    it = hs.iters + 1;
    rho_last = hs.rho32;
    k_mix_fold<<<nblk(n, 256), 256, 0, s>>>(n, sigma, cg->dsc, cg->y32, cg->xw);
    gcg = mix_residual(cg, sval);
    hs.refinements += 1;
    if (gcg != gcg) { hs.done = 4; return 0; }
    if (gcg < hs.eps2) { hs.done = 1; return 0; }
    if (it - 1 >= itmax) { hs.done = 3; return 0; }
    const double sig_new = std::sqrt(gcg);
    // p_true = r_true + beta p_true_old, beta = gcg / (rho_last sigma^2), in units of sig_new
    hm.cbeta = (gcg / (rho_last * sigma * sigma)) * (sigma / sig_new);
    hm.first_it = it;
    sigma = sig_new;
  }
}

static void cg_release(ffhip_cg cg) {
  for (hipEvent_t e : cg->ev) (void)hipEventDestroy(e);
  for (auto &g : cg->graphs) (void)hipGraphExecDestroy(g.exec);
  if (cg->fmt == kCsrV) cg->sval = nullptr;  // A's array
  dfree(cg->sptr); dfree(cg->slen); dfree(cg->scol); dfree(cg->scol16); dfree(cg->sptr16); dfree(cg->bcol); dfree(cg->bperm); dfree(cg->bwave); dfree(cg->sval); dfree(cg->rb);
  dfree(cg->svalT);
  dfree(cg->sval32); dfree(cg->sval32T); dfree(cg->dinv); dfree(cg->dsc); dfree(cg->wcl);
  dfree(cg->xw); dfree(cg->bw); dfree(cg->G); dfree(cg->CG); dfree(cg->H[0]); dfree(cg->H[1]);
  dfree(cg->AH); dfree(cg->r32); dfree(cg->q32); dfree(cg->y32); dfree(cg->p32[0]);
  dfree(cg->p32[1]); dfree(cg->pA); dfree(cg->pB); dfree(cg->pC); dfree(cg->st); dfree(cg->ms);
  dfree(cg->st2);
  for (double *&p : cg->gmV) dfree(p);
  for (double *&p : cg->gmVp) dfree(p);
  dfree(cg->gmVpDev); dfree(cg->gmW); dfree(cg->gmX0); dfree(cg->gmH); dfree(cg->gmY);
  delete cg;
}

extern "C" {

int ffhip_cg_create(ffhip_ctx ctx, ffhip_mat A, int mode, ffhip_cg *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && A && out, FFHIP_ERR_ARG, "ffhip_cg_create: null argument");
    FFHIP_REQUIRE(!A->half, FFHIP_ERR_UNSUPPORTED,
                  "ffhip_cg_create: a half (sym=1) matrix; expand it with ffhip_mat_full");
    const int prec = mode & 0xf, layout = mode & 0x130;
    const bool index32 = (mode & FFHIP_CG_INDEX32) != 0;
    FFHIP_REQUIRE((prec == FFHIP_CG_FP64 || prec == FFHIP_CG_MIXED) &&
                      (mode & ~(0x13f | FFHIP_CG_INDEX32 | FFHIP_CG_EAGER)) == 0 &&
                      (layout == 0 || layout == FFHIP_CG_LAYOUT_SELL || layout == FFHIP_CG_LAYOUT_CSR ||
                       layout == FFHIP_CG_LAYOUT_BSR3),
                  FFHIP_ERR_ARG, "ffhip_cg_create: mode %d", mode);
    ffhip_cg cg = new ffhip_cg_s();
    cg->ctx = ctx; cg->A = A; cg->mode = prec; cg->n = A->n;
    cg->use_graph = (mode & FFHIP_CG_EAGER) ? 0 : 1;
    try {
      const int64_t n = A->n;
      cg->dinv = dalloc<double>(n);
      cg->dsc = dalloc<double>(n);
      cg->wcl = dalloc<uint8_t>(n);
      cg->rb = dalloc<SetupRb>(1);
      if (prec == FFHIP_CG_MIXED) {  // the FP32 fill scales by the Jacobi diagonal
        cg_setup_jacobi(cg);
        cg_plan_layout(cg, layout, index32);  // reads the tgv figures back with its sizes
        cg_fill_values(cg, false);
      } else {  // FP64: the Jacobi set-up after the layout, from the fill's diagonal on sliced ELL
        FFHIP_CHECK(hipMemsetAsync(cg->rb, 0, sizeof(SetupRb), ctx->stream));
        cg_plan_layout(cg, layout, index32);
        Arena ar(ctx, DiagBufs::bytes(n));
        const DiagBufs b(ar, n);
        const bool in_fill = cg->fmt == kSell || cg->fmt == kSell16;
        if (!in_fill) {
          k_diag_vals<<<nblk(n, 256), 256, 0, ctx->stream>>>(n, A->diag, A->vals, b.d, b.has);
          cg_jacobi_tail(cg, b);
        }
        cg_fill_values(cg, false, in_fill ? &b : nullptr, true);
      }
      cg->xw = dalloc<double>(n); cg->bw = dalloc<double>(n);
      cg->G = dalloc<double>(n); cg->CG = dalloc<double>(n);
      cg->H[0] = dalloc<double>(n); cg->H[1] = dalloc<double>(n);
      cg->AH = dalloc<double>(n);
      if (prec == FFHIP_CG_MIXED) {
        cg->r32 = dalloc<float>(n); cg->q32 = dalloc<float>(n); cg->y32 = dalloc<float>(n);
        cg->p32[0] = dalloc<float>(n); cg->p32[1] = dalloc<float>(n);
      }
      cg->pA = dalloc<double>(3 * (int64_t)cg->grid);
      cg->pB = dalloc<double>(3 * (int64_t)cg->grid);
      cg->pC = dalloc<double>(3 * (int64_t)cg->grid);
      cg->st = dalloc<CgState>(1);
      cg->ms = dalloc<MixState>(1);
      cg->st2 = dalloc<Cg2State>(1);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      cg_release(cg);
      throw;
    }
    *out = cg;
  });
}

int ffhip_cg_layout(ffhip_cg cg, int *layout, int64_t *entries) {
  return guard([&] {
    FFHIP_REQUIRE(cg, FFHIP_ERR_ARG, "ffhip_cg_layout: null solver");
    if (layout)
      *layout = cg->fmt == kCsrV    ? FFHIP_CG_LAYOUT_CSR
                : cg->fmt == kBsr3 ? FFHIP_CG_LAYOUT_BSR3
                                   : (FFHIP_CG_LAYOUT_SELL | (cg->fmt == kSell16 ? FFHIP_CG_INDEX16 : 0));
    if (entries) *entries = cg->fmt == kBsr3 ? 9 * cg->ents : cg->ents;
  });
}

int ffhip_cg_update_values(ffhip_cg cg) {
  return guard([&] {
    FFHIP_REQUIRE(cg, FFHIP_ERR_ARG, "ffhip_cg_update_values: null solver");
    cg_setup_jacobi(cg);
    setup_readback(cg);
    cg_fill_values(cg, false);
    if (cg->svalT) cg_fill_values(cg, true);
  });
}

int ffhip_cg_solve(ffhip_cg cg, const double *b, double *x, int on_device, double eps, int itmax,
                   double tgv, int trans, ffhip_cg_stats *stats) {
  return guard([&] {
    FFHIP_REQUIRE(cg && b && x, FFHIP_ERR_ARG, "ffhip_cg_solve: null argument");
    ffhip_ctx ctx = cg->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = cg->n;
    if (trans && !cg->svalT) cg_fill_values(cg, true);
    const double *sval = trans ? cg->svalT : cg->sval;
    const float *sval32 = trans ? cg->sval32T : cg->sval32;
    if (itmax <= 0) itmax = (int)n;  // SolverCG(ds): itmax > 0 ? itmax : n
    // captured before this solve's first launch (kept for later solves)
    if (cg->mode == FFHIP_CG_FP64) (void)pcg_chunk_graph(cg, sval);
    else { (void)mix_chunk_graph(cg, sval32, 0); (void)mix_chunk_graph(cg, sval32, 1); }
    FFHIP_CHECK(hipMemcpyAsync(cg->bw, b, 8 * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(cg->xw, x, 8 * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    ctx->tic();
    if (cg->ntgv > 0) k_init_bc<<<nblk(n, 256), 256, 0, s>>>(n, cg->wcl, cg->bw, tgv, cg->xw);
    CgState hs;
    memset(&hs, 0, sizeof(hs));
    cg->t_dv = cg->t_mv = cg->t_up = 0.;
    cg->n_timed = 0;
    if (cg->mode == FFHIP_CG_FP64) {
      const int G = cg->grid;
      const MatFmt<double> M = cg->fmt_of(sval);
      const int chunk = kChunk;
      hipGraphExec_t gx = pcg_chunk_graph(cg, sval);
      hs.it0 = 1;
      hs.itmax = itmax;
      hs.eps = eps;
      FFHIP_CHECK(hipMemcpyAsync(cg->st, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
      FFHIP_LAUNCH_FMT(cg, k_pcg_init, G, s, M, cg->xw, cg->bw, cg->dinv, cg->G, cg->CG, cg->H[0],
                       cg->pA, cg->st);
      FFHIP_CHECK(hipGetLastError());
      cg->ensure_events(4 * chunk);
      for (int it = 1;; it += chunk) {
        if (gx) FFHIP_CHECK(hipGraphLaunch(gx, s));
        else launch_pcg_chunk(cg, sval, cg->profile != 0);
        FFHIP_CHECK(hipGetLastError());
        FFHIP_CHECK(hipMemcpyAsync(&hs, cg->st, sizeof(hs), hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipStreamSynchronize(s));
        cg->collect(it, chunk, hs.done ? hs.iters : it + chunk - 1);
        if (hs.done) break;
      }
    } else {
      ffhip_cg_solve_mixed(cg, sval, sval32, eps, itmax, hs);
    }
    ctx->toc();
    const double ms = ctx->last_ms;
    double res[3];
    residual_sums(cg, sval, cg->xw, cg->bw, nullptr, res);
    FFHIP_CHECK(hipMemcpyAsync(x, cg->xw, 8 * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    if (stats) {
      stats->iters = hs.iters;
      stats->converged = hs.done == 1 ? 1 : (hs.done == 2 ? 2 : 0);
      stats->refinements = hs.refinements;
      stats->gcg0 = hs.gcg0;
      stats->gcg = cg->mode == FFHIP_CG_FP64 ? hs.gcg_prev[(hs.iters + 1) & 1] : res[2];
      stats->true_gcg = res[2];
      stats->true_relres = res[1] > 0 ? std::sqrt(res[0] / res[1]) : std::sqrt(res[0]);
      stats->ms_solve = ms;
      stats->timed = cg->n_timed;
      stats->ms_matvec_avg = cg->n_timed ? cg->t_mv / cg->n_timed : 0.;
      stats->ms_update_avg = cg->n_timed ? cg->t_up / cg->n_timed : 0.;
      stats->ms_direction_avg = cg->n_timed ? cg->t_dv / cg->n_timed : 0.;
    }
    FFHIP_REQUIRE(hs.done != 4, FFHIP_ERR_NAN, "ffhip_cg_solve: NaN in CG (bad matrix)");
    FFHIP_REQUIRE(hs.done != 3, FFHIP_ERR_NOCONV,
                  "ffhip_cg_solve: ConjugueGradient did not converge in %d iterations", itmax);
  });
}

int ffhip_cg_solve_affine(ffhip_cg cg, const double *b, double *x, int on_device, double eps,
                          int itmax, int precon, double *eps_out, ffhip_cg_stats *stats) {
  return guard([&] {
    FFHIP_REQUIRE(cg && x, FFHIP_ERR_ARG, "ffhip_cg_solve_affine: null argument");
    FFHIP_REQUIRE(precon == FFHIP_CG2_IDENTITY || precon == FFHIP_CG2_JACOBI, FFHIP_ERR_ARG,
                  "ffhip_cg_solve_affine: precon %d", precon);
    FFHIP_REQUIRE(itmax >= 0, FFHIP_ERR_ARG, "ffhip_cg_solve_affine: nbiter %d", itmax);
    ffhip_ctx ctx = cg->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = cg->n;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (b) FFHIP_CHECK(hipMemcpyAsync(cg->bw, b, 8 * n, kin, s));
    FFHIP_CHECK(hipMemcpyAsync(cg->xw, x, 8 * n, kin, s));
    const double *bd = b ? cg->bw : nullptr;
    const double *dinv = precon == FFHIP_CG2_JACOBI ? cg->dinv : nullptr;
    double *h = cg->H[0];
    const int G = cg->grid;
    const MatFmt<double> M = cg->fmt_of(cg->sval);
    ctx->tic();
    FFHIP_LAUNCH_FMT(cg, k_cg2_init, G, s, M, cg->xw, bd, dinv, cg->G, cg->CG, h, cg->pA, cg->st2);
    FFHIP_CHECK(hipGetLastError());
    Cg2State hs;
    const int chunk = 16;
    for (int it = 0;; ++it) {
      k_cg2_step<<<G, kBlock, 0, s>>>(it, itmax, eps, n, h, cg->CG, cg->xw, cg->pA, cg->st2);
      FFHIP_LAUNCH_FMT(cg, k_cg2_dir, G, s, M, cg->xw, bd, cg->G, h, cg->AH, cg->pB, cg->st2);
      k_cg2_update<<<G, kBlock, 0, s>>>(it, n, h, cg->AH, dinv, cg->xw, cg->G, cg->CG, cg->pB,
                                        cg->pA, cg->st2);
      if ((it + 1) % chunk == 0) {
        FFHIP_CHECK(hipGetLastError());
        FFHIP_CHECK(hipMemcpyAsync(&hs, cg->st2, sizeof(hs), hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipStreamSynchronize(s));
        if (hs.done) break;
      }
    }
    ctx->toc();
    const double ms = ctx->last_ms;
    double res[3] = {0., 0., 0.};
    if (!b) FFHIP_CHECK(hipMemsetAsync(cg->bw, 0, 8 * n, s));
    residual_sums(cg, cg->sval, cg->xw, cg->bw, nullptr, res);
    FFHIP_CHECK(hipMemcpyAsync(x, cg->xw, 8 * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    if (eps_out) *eps_out = hs.done == 2 ? eps : hs.reps2;
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->iters = hs.iters;
      stats->converged = hs.done == 1 ? 1 : (hs.done == 2 ? 2 : 0);
      stats->refinements = 0;
      stats->gcg0 = hs.g20;
      stats->gcg = hs.g2last;
      stats->true_gcg = dinv ? res[2] : res[0];
      stats->true_relres = res[1] > 0 ? std::sqrt(res[0] / res[1]) : std::sqrt(res[0]);
      stats->ms_solve = ms;
    }
    FFHIP_REQUIRE(hs.done != 5, FFHIP_ERR_NAN, "CG2: Matrix is not defined (/0), sorry");
    FFHIP_REQUIRE(hs.done != 6, FFHIP_ERR_NAN, "CG2: Bug : ro is NaN");
  });
}

int ffhip_gmres_solve(ffhip_cg cg, const double *b, double *x, int on_device, double eps,
                      int itmax, int restart, double tgv, int trans, ffhip_cg_stats *stats) {
  return guard([&] {
    FFHIP_REQUIRE(cg && b && x, FFHIP_ERR_ARG, "ffhip_gmres_solve: null argument");
    FFHIP_REQUIRE(restart >= 1, FFHIP_ERR_ARG, "ffhip_gmres_solve: restart %d", restart);
    ffhip_ctx ctx = cg->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = cg->n;
    const int G = cg->grid, m = restart;
    if (trans && !cg->svalT) cg_fill_values(cg, true);
    const double *sval = trans ? cg->svalT : cg->sval;
    const MatFmt<double> M = cg->fmt_of(sval);
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    cg->gm_basis(1);
    FFHIP_CHECK(hipMemcpyAsync(cg->bw, b, 8 * n, kin, s));
    FFHIP_CHECK(hipMemcpyAsync(cg->xw, x, 8 * n, kin, s));
    std::vector<double> part(G), hcol(m + 1);
    // block partials written by the last kernel, added on the host in block order
    auto host_sum = [&](const double *p) {
      FFHIP_CHECK(hipMemcpyAsync(part.data(), p, 8 * G, hipMemcpyDeviceToHost, s));
      FFHIP_CHECK(hipStreamSynchronize(s));
      double t = 0.;
      for (int i = 0; i < G; ++i) t += part[i];
      return t;
    };
    ctx->tic();
    if (cg->ntgv > 0) k_init_bc<<<nblk(n, 256), 256, 0, s>>>(n, cg->wcl, cg->bw, tgv, cg->xw);
    const int nbitermx = itmax > 0 ? itmax : (int)n;  // SolverGMRES(ds)
    // fgmres (CG.cpp:347-517) with leftC = 0: the scalar part as the reference's
    std::vector<double> rot0(m + 2), rot1(m + 2), g(m + 1), g1(m + 1), y(m + 1),
        Hn((size_t)(m + 2) * (m + 1), 0.);
    auto HN = [&](int i, int j) -> double & { return Hn[(size_t)i * (m + 1) + j]; };
    k_gm_normb<<<G, kBlock, 0, s>>>(n, cg->bw, cg->ntgv > 0 ? cg->wcl : nullptr, cg->pA);
    double normb = std::sqrt(host_sum(cg->pA));
    double relres = 1e100;
    bool noconv = true;
    int iter = 0, nrestart = 0;
    while (noconv) {
      FFHIP_CHECK(hipMemcpyAsync(cg->gmX0, cg->xw, 8 * n, hipMemcpyDeviceToDevice, s));
      FFHIP_LAUNCH_FMT(cg, k_gm_residual, G, s, M, cg->xw, cg->bw, cg->gmW, cg->pA);
      FFHIP_CHECK(hipGetLastError());
      g[0] = std::sqrt(host_sum(cg->pA));
      if (normb < 1.e-20 || eps < 0) normb = 1.;
      double inv = 1. / g[0];
      int it;
      for (it = 0; it < m; it++, iter++) {
        cg->gm_basis(it + 1);
        k_gm_prec<<<G, kBlock, 0, s>>>(n, inv, cg->gmW, cg->dinv, cg->gmV[it], cg->gmVp[it]);
        FFHIP_LAUNCH_FMT(cg, k_gm_matvec, G, s, M, cg->gmVp[it], cg->gmV[0], cg->gmW, cg->pB);
        double *pin = cg->pB, *pout = cg->pA;
        for (int i = 0; i <= it; ++i) {  // modified Gram-Schmidt
          k_gm_mgs<<<G, kBlock, 0, s>>>(i, n, cg->gmV[i], i < it ? cg->gmV[i + 1] : nullptr,
                                        cg->gmW, pin, pout, cg->gmH);
          std::swap(pin, pout);
        }
        FFHIP_CHECK(hipGetLastError());
        FFHIP_CHECK(hipMemcpyAsync(hcol.data(), cg->gmH, 8 * (it + 1), hipMemcpyDeviceToHost, s));
        const double ww = host_sum(pin);
        for (int i = 0; i <= it; ++i) HN(i, it) = hcol[i];
        const double aux = HN(it + 1, it) = std::sqrt(ww);
        inv = 1. / aux;
        for (int i = 0; i < it; i++) {
          const double aa = rot0[i] * HN(i, it) + rot1[i] * HN(i + 1, it);
          const double bb = -rot1[i] * HN(i, it) + rot0[i] * HN(i + 1, it);
          HN(i, it) = aa;
          HN(i + 1, it) = bb;
        }
        const double sq = std::sqrt(HN(it, it) * HN(it, it) + HN(it + 1, it) * HN(it + 1, it));
        rot0[it] = HN(it, it) / sq;
        rot1[it] = HN(it + 1, it) / sq;
        HN(it, it) = rot0[it] * HN(it, it) + rot1[it] * HN(it + 1, it);
        HN(it + 1, it) = 0.;
        g[it + 1] = -rot1[it] * g[it];
        g[it] = rot0[it] * g[it];
        relres = std::fabs(g[it + 1]);
        if (relres / normb < std::fabs(eps)) { noconv = false; break; }
        if (it > nbitermx) break;
      }
      it = std::min(it, m - 1);
      for (int i = it; i >= 0; i--) {  // reconstruction
        g1[i] = g[i];
        for (int j = i + 1; j < it + 1; j++) g1[i] = g1[i] - HN(i, j) * y[j];
        y[i] = g1[i] / HN(i, i);
      }
      FFHIP_CHECK(hipMemcpyAsync(cg->gmY, y.data(), 8 * (it + 1), hipMemcpyHostToDevice, s));
      k_gm_update<<<G, kBlock, 0, s>>>(n, it + 1, cg->gmY, cg->gmVpDev, cg->gmX0, cg->xw);
      FFHIP_CHECK(hipGetLastError());
      if (!noconv) break;
      if (iter > nbitermx) break;
      nrestart++;
    }
    ctx->toc();
    const double ms = ctx->last_ms;
    double res[3];
    residual_sums(cg, sval, cg->xw, cg->bw, nullptr, res);
    FFHIP_CHECK(hipMemcpyAsync(x, cg->xw, 8 * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->iters = iter;
      stats->converged = noconv ? 0 : 1;
      stats->refinements = nrestart;
      stats->gcg0 = normb;
      stats->gcg = relres / normb;
      stats->true_gcg = res[2];
      stats->true_relres = res[1] > 0 ? std::sqrt(res[0] / res[1]) : std::sqrt(res[0]);
      stats->ms_solve = ms;
    }
    FFHIP_REQUIRE(!noconv, FFHIP_ERR_NOCONV, "ffhip_gmres_solve: fgmres did not converge in %d iterations",
                  iter);
  });
}

int ffhip_cg_set_profiling(ffhip_cg cg, int on) {
  return guard([&] {
    FFHIP_REQUIRE(cg, FFHIP_ERR_ARG, "ffhip_cg_set_profiling: null solver");
    cg->profile = on ? 1 : 0;
  });
}

int ffhip_cg_free(ffhip_cg cg) {
  return guard([&] {
    if (!cg) return;
    (void)hipStreamSynchronize(cg->ctx->stream);
    cg_release(cg);
  });
}

int ffhip_bench_spmv(ffhip_cg cg, int mode, int reps, const void *x, void *y, double *avg_ms) {
  return guard([&] {
    FFHIP_REQUIRE(cg && x && y && avg_ms && reps > 0, FFHIP_ERR_ARG, "ffhip_bench_spmv: bad arguments");
    FFHIP_REQUIRE(mode == FFHIP_CG_FP64 || cg->sval32, FFHIP_ERR_ARG,
                  "ffhip_bench_spmv: FP32 matrix only exists for a MIXED solver");
    ffhip_ctx ctx = cg->ctx;
    hipStream_t s = ctx->stream;
    auto launch = [&] {
      if (mode == FFHIP_CG_FP64) {
        const MatFmt<double> M = cg->fmt_of<double>(cg->sval);
        FFHIP_LAUNCH_MATVEC(cg, double, M, x, y, s);
      } else {
        const MatFmt<float> M = cg->fmt_of<float>(cg->sval32);
        FFHIP_LAUNCH_MATVEC(cg, float, M, x, y, s);
      }
    };
    launch();  // warm
    FFHIP_CHECK(hipGetLastError());
    ctx->tic();
    for (int i = 0; i < reps; ++i) launch();
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    *avg_ms = ctx->last_ms / reps;
  });
}

}  // extern "C"
