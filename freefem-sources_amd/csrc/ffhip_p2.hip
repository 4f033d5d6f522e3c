// ffhip_p2.hip — vector P2 ([P2,P2,P2]) linear elasticity on tetrahedra:
// node numbering, matrix pattern, Lame element matrices on the matrix cores,
// atomic-free row assembly, clamped boundary rows and body-force right-hand
// side.
//
// Replaces, for `varf va([u1,u2,u3],[v1,v2,v3]) = int3d(Th)(lambda*div(u)*div(v)
// + 2*mu*(epsilon(u)'*epsilon(v))) + int3d(Th)(f'*v) + on(4, u1=0,u2=0,u3=0)`
// on Vh(Th,[P2,P2,P2]):
//   node numbering   GenericMesh::BuildDFNumbering (GenericMesh.hpp:1880-1920)
//                    and FESpacen.cpp:403-428 (dof = 3*node + component)
//   element matrix   Element_Op (problem.cpp:6337-6437) with the P2 basis of
//                    TypeOfFE_P2Lagrange3d::FB (P012_3d.cpp:199-270), Tet_5
//   accumulation     HashMatrix::operator+= (HashMatrix.cpp:1294-1332)
//   Dirichlet        AssembleBC (problem.cpp:10120-10164), SetBC tgv
//   right-hand side  Element_rhs (problem.cpp:7940-7978)
//
// Design:
//   numbering  every (element, local node) item carries a key (vertex or
//              edge); a node's number is the rank of its first item, found
//              with atomicMin over vertices, a radix sort of the edge keys and
//              one scan over the items — equal to FreeFEM's sequential
//              first-appearance numbering.
//   pattern    node rows (sorted unique node columns, LDS merge per lane, as
//              P1), expanded to scalar rows of 3 x 3 blocks.
//   element    one wave per element: M = D W D^T (30 x 30, D = P2 gradients
//              at the 14 points, W = |K| w_q) as three 16x16 tiles of
//              v_mfma_f64_16x16x4f64 over k = 16 (14 points, zero padded);
//              K[(c,i),(c',j)] = lambda M[ic][jc'] + mu M[ic'][jc]
//              + mu delta_cc' sum_a M[ia][ja] is formed from M in LDS and
//              written node-block major (90 doubles per local node).
//   gather     one wave per node: the incident elements' 90-value blocks are
//              added in element order into the node's 3 rows staged in LDS,
//              then written back contiguously.  No atomics, deterministic.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

#include "ffhip_fem.hpp"

namespace ffhip {
void set_diag_rows(ffhip_ctx ctx, int64_t n, const uint8_t *onbc, const int64_t *diag, double tgv,
                   double *vals, int *missing);
int read_flag(ffhip_ctx ctx, int *d);

// Tet::nvedge (local vertex pairs of the 6 edges)
static __constant__ int c_edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ inline uint64_t edge_key(int32_t a, int32_t b) {
  if (a > b) { const int32_t t = a; a = b; b = t; }
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// ---------------------------------------------------------------- numbering
__global__ void k_fill_i32(int64_t n, int32_t v, int32_t *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// first item (10 k + i) of every vertex
__global__ void k_p2_vfirst(int64_t nt, const int32_t *__restrict__ tets, int32_t *__restrict__ firstv) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= 4 * nt) return;
  const int64_t k = m / 4;
  const int i = (int)(m % 4);
  atomicMin(&firstv[tets[m]], (int32_t)(10 * k + i));
}

// edge keys and their items, in item order (the radix sort is stable)
// sort keys of the edge items: min << vb | max (vb the bits of a vertex id,
// so the radix sort runs over 2 vb bits, not 64)
__global__ void k_p2_edge_keys(int64_t nt, const int32_t *__restrict__ tets, int vb,
                               uint64_t *__restrict__ keys, int32_t *__restrict__ items) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= 6 * nt) return;
  const int64_t k = m / 6;
  const int j = (int)(m % 6);
  const int32_t *t = tets + 4 * k;
  const int32_t a = t[c_edge[j][0]], b = t[c_edge[j][1]];
  keys[m] = ((uint64_t)(uint32_t)(a < b ? a : b) << vb) | (uint32_t)(a < b ? b : a);
  items[m] = (int32_t)(10 * k + 4 + j);
}

// head flags of the sorted edge keys; first-appearance marks of edge items
__global__ void k_p2_edge_heads(int64_t E, const uint64_t *__restrict__ skeys,
                                const int32_t *__restrict__ sitems, int64_t *__restrict__ head,
                                int64_t *__restrict__ mark) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const int64_t h = (p == 0 || skeys[p] != skeys[p - 1]) ? 1 : 0;
  head[p] = h;
  mark[sitems[p]] = h;
}

// edge segment of every edge item; unique keys and their first items
__global__ void k_p2_edge_seg(int64_t E, const uint64_t *__restrict__ skeys, int vb,
                              const int32_t *__restrict__ sitems, const int64_t *__restrict__ hscan,
                              const int64_t *__restrict__ head, int32_t *__restrict__ item_seg,
                              uint64_t *__restrict__ ukeys, int32_t *__restrict__ segfirst) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const int64_t seg = hscan[p] + head[p] - 1;  // inclusive scan - 1
  const int32_t item = sitems[p];
  item_seg[(item / 10) * 6 + (item % 10 - 4)] = (int32_t)seg;
  if (head[p]) {  // the unique keys in edge_key form (min << 32 | max)
    const uint64_t k = skeys[p];
    ukeys[seg] = ((k >> vb) << 32) | (k & ((uint64_t(1) << vb) - 1));
    segfirst[seg] = item;
  }
}

__global__ void k_p2_vmark(int64_t nt, const int32_t *__restrict__ tets,
                           const int32_t *__restrict__ firstv, int64_t *__restrict__ mark) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= 4 * nt) return;
  const int64_t k = m / 4;
  const int i = (int)(m % 4);
  const int32_t item = (int32_t)(10 * k + i);
  mark[10 * k + i] = firstv[tets[m]] == item ? 1 : 0;
}

__global__ void k_p2_assign(int64_t nt, const int32_t *__restrict__ tets,
                            const int32_t *__restrict__ firstv, const int32_t *__restrict__ item_seg,
                            const int32_t *__restrict__ segfirst, const int64_t *__restrict__ rank,
                            int32_t *__restrict__ elem_nodes) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= 10 * nt) return;
  const int64_t k = m / 10;
  const int i = (int)(m % 10);
  const int32_t first = i < 4 ? firstv[tets[4 * k + i]] : segfirst[item_seg[6 * k + i - 4]];
  elem_nodes[m] = (int32_t)rank[first];
}

__global__ void k_p2_vnode(int64_t nv, const int32_t *__restrict__ firstv,
                           const int64_t *__restrict__ rank, int32_t *__restrict__ vnode) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int32_t f = firstv[v];
  vnode[v] = f == INT_MAX ? -1 : (int32_t)rank[f];
}

__global__ void k_p2_enode(int64_t ne, const int32_t *__restrict__ segfirst,
                           const int64_t *__restrict__ rank, int32_t *__restrict__ enode) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ne) enode[s] = (int32_t)rank[segfirst[s]];
}

// ---------------------------------------------------------------- pattern
constexpr int kNodeCap = 96;   // node columns per node row
constexpr int kNodeHash = 256;  // LDS hash slots per node row

// One node row per wave: lane q takes incidence q's 10 element nodes and
// inserts them into an open-addressing set in LDS (atomicCAS, bounded
// probing); the unique node columns are ordered by a counting sort and
// written to the row's buffer (kNodeCap entries), with the row length, the
// relative diagonal position and the 10 slot bytes of every incidence
// (from a rank table indexed by the column's hash slot).  A row with more than 64 incident
// elements, more than kNodeCap columns or a full set raises *ovf (the
// lane-per-row merge then builds the whole pattern).
__global__ __launch_bounds__(256) void k_p2_pattern_wave(
    int64_t nn, const int32_t *__restrict__ elem_nodes, const int64_t *__restrict__ n2e_ptr,
    const int32_t *__restrict__ n2e, int64_t *__restrict__ nlen, int32_t *__restrict__ rowbuf,
    int8_t *__restrict__ dposr, uint8_t *__restrict__ slot8, int *__restrict__ ovf) {
  constexpr int NW = 256 / kWave;
  __shared__ int32_t hk[NW][kNodeHash];
  __shared__ int32_t ucol[NW][kNodeCap];
  __shared__ uint8_t hslot[NW][kNodeCap];   // the set slot of each unique column
  __shared__ uint8_t hrank[NW][kNodeHash];  // the offset in the row of the column in each set slot
  __shared__ int32_t ucnt[NW];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int64_t a = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  int64_t k0 = 0;
  int deg = 0;
  if (a < nn) { k0 = n2e_ptr[a]; deg = (int)(n2e_ptr[a + 1] - k0); }
  bool ok = deg <= kWave;
  if (!ok) deg = 0;
  for (int i = lane; i < kNodeHash; i += kWave) hk[w][i] = INT_MAX;
  if (lane == 0) ucnt[w] = 0;
  int32_t c[10];
  if (lane < deg) {
    const int64_t e = n2e[k0 + lane] / 10;
#pragma unroll
    for (int i = 0; i < 10; ++i) c[i] = elem_nodes[10 * e + i];
  }
  __syncthreads();
  bool full = false;
  if (lane < deg) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      uint32_t h = ((uint32_t)c[i] * 2654435761u) >> (32 - 8);
      bool placed = false;
      for (int probe = 0; probe < kNodeHash && !placed; ++probe) {
        const int32_t old = atomicCAS(&hk[w][h], INT_MAX, c[i]);
        if (old == INT_MAX) {
          const int idx = atomicAdd(&ucnt[w], 1);
          if (idx < kNodeCap) { ucol[w][idx] = c[i]; hslot[w][idx] = (uint8_t)h; }
          placed = true;
        } else if (old == c[i]) {
          placed = true;
        }
        h = (h + 1) & (kNodeHash - 1);
      }
      full |= !placed;
    }
  }
  __syncthreads();
  const int len = ucnt[w];
  if (__any(full) || len > kNodeCap) ok = false;
  if (a < nn && !ok && lane == 0) atomicOr(ovf, 1);
  int32_t u[2] = {INT_MAX, INT_MAX};
  int rk[2] = {0, 0};
  if (ok) {
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int j = lane + h2 * kWave;
      if (j < len) {
        u[h2] = ucol[w][j];
        int r = 0;
        for (int i = 0; i < len; ++i) r += ucol[w][i] < u[h2] ? 1 : 0;
        rk[h2] = r;
      }
    }
  }
  __syncthreads();
  int dpos = -1;
  if (ok) {
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2)
      if (lane + h2 * kWave < len) {
        hrank[w][hslot[w][lane + h2 * kWave]] = (uint8_t)rk[h2];
        ucol[w][rk[h2]] = u[h2];
        rowbuf[kNodeCap * a + rk[h2]] = u[h2];
        if ((int64_t)u[h2] == a) dpos = rk[h2];
      }
  }
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_xor(dpos, off, kWave);
    dpos = t > dpos ? t : dpos;
  }
  __syncthreads();
  if (ok && a < nn && lane == 0) { nlen[a] = len; dposr[a] = (int8_t)dpos; }
  if (ok && lane < deg) {
    uint8_t *dst = slot8 + (k0 + lane) * 10;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      // the column is in the set: its slot holds its offset in the row
      uint32_t h = ((uint32_t)c[i] * 2654435761u) >> (32 - 8);
      while (hk[w][h] != c[i]) h = (h + 1) & (kNodeHash - 1);
      dst[i] = hrank[w][h];
    }
  }
}

// scalar rows 3a + c of node a: for every node column b, columns 3b, 3b+1, 3b+2
// (node columns from ncol at nodeptr, or from the row buffers when rowbuf != null).
// One wave per node: the node's 3 rows are one run of 9 L columns at 9 p0,
// written with consecutive lanes on consecutive entries.
__global__ __launch_bounds__(256) void k_p2v_expand(int64_t nn, const int64_t *__restrict__ nodeptr,
                                                    const int32_t *__restrict__ ncol,
                                                    const int64_t *__restrict__ ndiag,
                                                    const int32_t *__restrict__ rowbuf,
                                                    const int8_t *__restrict__ dposr,
                                                    int64_t *__restrict__ rowptr,
                                                    int32_t *__restrict__ colidx,
                                                    int64_t *__restrict__ diag) {
  const int lane = threadIdx.x % kWave;
  const int64_t a = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (a >= nn) return;  // uniform in the wave
  const int64_t p0 = nodeptr[a], L = nodeptr[a + 1] - p0;
  const int32_t *cols = rowbuf ? rowbuf + kNodeCap * a : ncol + p0;
  const int64_t base0 = 9 * p0;
  for (int c = 0; c < 3; ++c)
    for (int64_t j = lane; j < 3 * L; j += kWave) {
      const int64_t q = j / 3;
      colidx[base0 + c * 3 * L + j] = 3 * cols[q] + (int32_t)(j - 3 * q);
    }
  if (lane < 3) {
    const int c = lane;
    const int64_t r = 3 * a + c, base = base0 + c * 3 * L;
    rowptr[r] = base;
    if (r == 3 * nn - 1) rowptr[3 * nn] = 9 * nodeptr[nn];
    const int64_t d = rowbuf ? (dposr[a] < 0 ? -1 : dposr[a]) : (ndiag[a] < 0 ? -1 : ndiag[a] - p0);
    diag[r] = d < 0 ? -1 : base + 3 * d + c;
  }
}

// scalar P2: the node rows are the matrix rows (columns from ncol at nodeptr,
// or from the row buffers when rowbuf != null)
__global__ void k_p2_compact(int64_t nn, const int64_t *__restrict__ nodeptr,
                             const int32_t *__restrict__ ncol, const int64_t *__restrict__ ndiag,
                             const int32_t *__restrict__ rowbuf, const int8_t *__restrict__ dposr,
                             int64_t *__restrict__ rowptr, int32_t *__restrict__ colidx,
                             int64_t *__restrict__ diag) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nn) return;
  const int64_t p0 = nodeptr[a], L = nodeptr[a + 1] - p0;
  rowptr[a] = p0;
  if (a == nn - 1) rowptr[nn] = nodeptr[nn];
  if (rowbuf)
    for (int64_t q = 0; q < L; ++q) colidx[p0 + q] = rowbuf[kNodeCap * a + q];
  const int64_t d = rowbuf ? (dposr[a] < 0 ? -1 : dposr[a]) : (ndiag[a] < 0 ? -1 : ndiag[a] - p0);
  diag[a] = d < 0 ? -1 : p0 + d;
}

// ---------------------------------------------------------------- element matrices
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline double pick3(const double v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }
__device__ inline double pick4(const double v[4], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}


// element rows a node-row gather keeps in flight
constexpr int kGatherD = 4;

constexpr int kElemBlock = 256;
constexpr int kElemWaves = kElemBlock / kWave;
constexpr int kMs = 33;  // LDS row stride of M (doubles)

// Row r = 3 i + a of D (r < 30): gradient component a of P2 basis i,
//   vertex i:        G_i[a] (4 l_i - 1)
//   edge (i0, i1):   4 (G_i1[a] l_i0 + G_i0[a] l_i1)        (P012_3d.cpp:236-262)
struct P2Row {
  int vert;      // 1: vertex basis
  int i0, i1;    // vertex index (i0) or edge end points
  double g0, g1; // G_i0[a], G_i1[a] of the current element
};

__device__ inline double p2_grad(const P2Row &R, const double l[4]) {
  if (R.vert) return R.g0 * (4 * pick4(l, R.i0) - 1);
  return 4 * (R.g1 * pick4(l, R.i0) + R.g0 * pick4(l, R.i1));
}

// LAP = false: K[(c,i),(c',j)] (row = test function) of the Lame form as 10
// node blocks of 90 doubles [c][j][c'], block i stored at Ke[90 k] for the
// incidence k of (e, i).  LAP = true: the scalar P2 stiffness of
// int3d(kappa grad u . grad v), K(i,j) = sum_a M[(i,a),(j,a)] with kappa(x_q)
// in the weights, as 10 rows of 10 doubles at Ke[10 k].
template <bool LAP>
__global__ __launch_bounds__(kElemBlock) void k_p2_element(
    int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets, double lambda,
    double mu, int kind, double c0, const double *__restrict__ kappa_e,
    const int32_t *__restrict__ inv, double *__restrict__ Ke) {
  __shared__ double sM[kElemWaves][32 * kMs];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int col = lane % 16, kq = lane / 16;
  // the two rows of D this lane feeds to the matrix cores (r = col, 16 + col)
  int rr[2] = {col, 16 + col};
  P2Row R[2];
  int rvalid[2];
#pragma unroll
  for (int I = 0; I < 2; ++I) {
    const int r = rr[I];
    rvalid[I] = r < 30;
    const int i = r < 30 ? r / 3 : 0;
    R[I].vert = i < 4;
    const int ed = i < 4 ? 0 : i - 4;
    R[I].i0 = i < 4 ? i : c_edge[ed][0];
    R[I].i1 = i < 4 ? i : c_edge[ed][1];
  }
  double* M = sM[w];
  for (int64_t base = (int64_t)blockIdx.x * kElemWaves; base < nt; base += (int64_t)gridDim.x * kElemWaves) {
    const int64_t e = base + w;  // wave-uniform
    if (e < nt) {
      const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
      const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
      double v[4][3];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
      double G[4][3];
      const double mes = tet_geom(v, G);
#pragma unroll
      for (int I = 0; I < 2; ++I) {
        const int a = rr[I] < 30 ? rr[I] % 3 : 0;
        double ga[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) ga[t] = pick3(G[t], a);
        R[I].g0 = pick4(ga, R[I].i0);
        R[I].g1 = pick4(ga, R[I].i1);
      }
      d4 acc00 = {0., 0., 0., 0.}, acc01 = {0., 0., 0., 0.}, acc11 = {0., 0., 0., 0.};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int q = 4 * ks + kq;  // quadrature point of this lane's k index
        double W = 0., l[4] = {0., 0., 0., 0.};
        if (q < 14) {
          const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
          l[0] = 1. - (px + py + pz); l[1] = px; l[2] = py; l[3] = pz;
          W = mes * qf5_weight(q);
          if (LAP) {
            double kq = c0;
            if (kind == FFHIP_COEF_P0) kq = kappa_e[e];
            else if (kind == FFHIP_COEF_QP) kq = kappa_e[14 * e + q];
            else if (kind == FFHIP_COEF_P1) {
              const double u[4] = {kappa_e[id[0]], kappa_e[id[1]], kappa_e[id[2]], kappa_e[id[3]]};
              kq = p1_value3(u, px, py, pz);
            }
            else if (kind == FFHIP_COEF_SIN3) {
              double P[3];
#pragma unroll
              for (int k = 0; k < 3; ++k)
                P[k] = l[0] * v[0][k] + px * v[1][k] + py * v[2][k] + pz * v[3][k];
              kq = kappa_sin3(P[0], P[1], P[2]);
            }
            W *= kq;
          }
        }
        const double g0 = rvalid[0] ? p2_grad(R[0], l) : 0.;
        const double g1 = rvalid[1] ? p2_grad(R[1], l) : 0.;
        acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(W * g0, g0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(W * g0, g1, acc01, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(W * g1, g1, acc11, 0, 0, 0);
      }
      // D[i][j] of a 16x16 tile in lane l, register v: j = l % 16, i = l / 16 + 4 v
      // (A[i][k] and B[k][j] are taken from lane i + 16 k and j + 16 k)
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        const int i = kq + 4 * vv, j = col;
        M[i * kMs + j] = acc00[vv];
        M[i * kMs + 16 + j] = acc01[vv];
        M[(16 + j) * kMs + i] = acc01[vv];
        M[(16 + i) * kMs + 16 + j] = acc11[vv];
      }
    }
    wave_lds_sync();  // M is this wave's own: the waves of the block run independently
    if (LAP && e < nt) {
      for (int o = lane; o < 100; o += kWave) {
        const int i = o / 10, j = o % 10;
        const double k = M[(3 * i) * kMs + 3 * j] + M[(3 * i + 1) * kMs + 3 * j + 1] +
                         M[(3 * i + 2) * kMs + 3 * j + 2];
        Ke[(int64_t)inv[10 * e + i] * 10 + j] = k;
      }
    } else if (e < nt) {
      // entry pairs (rem, rem + 1): a pair stays in its 90-value node-block
      // row, and rows start at multiples of 720 bytes, so one 16-byte store
      auto entry = [&](int i, int rem) {
        const int c = rem / 30, r2 = rem % 30, j = r2 / 3, cp = r2 % 3;
        double k = lambda * M[(3 * i + c) * kMs + 3 * j + cp] + mu * M[(3 * i + cp) * kMs + 3 * j + c];
        if (c == cp)
          k += mu * (M[(3 * i) * kMs + 3 * j] + M[(3 * i + 1) * kMs + 3 * j + 1] +
                     M[(3 * i + 2) * kMs + 3 * j + 2]);
        return k;
      };
      for (int o = lane; o < 450; o += kWave) {
        const int i = o / 45, rem = 2 * (o % 45);
        typedef double f64x2 __attribute__((ext_vector_type(2)));
        const f64x2 kk = {entry(i, rem), entry(i, rem + 1)};
        // node block i goes where its incidence entry sits in node i's list
        // (streamed: non-temporal, the gather reads it once)
        __builtin_nontemporal_store(kk, reinterpret_cast<f64x2 *>(&Ke[(int64_t)inv[10 * e + i] * 90 + rem]));
      }
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------- row gather

__global__ __launch_bounds__(256) void k_p2v_gather(
    int64_t nn, const int64_t *__restrict__ nodeptr, const int64_t *__restrict__ n2e_ptr,
    const uint8_t *__restrict__ slot8, const double *__restrict__ Ke, double *__restrict__ vals) {
  __shared__ double S[256 / kWave][9 * kNodeCap];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t a = (int64_t)blockIdx.x * (256 / kWave) + w;
  if (a >= nn) return;  // wave-uniform; the kernel has no block barrier
  double *acc = S[w];
  const int64_t p0 = nodeptr[a];
  const int L = (int)(nodeptr[a + 1] - p0), Wd = 9 * L;
  for (int t = lane; t < Wd; t += kWave) acc[t] = 0.;
  // this lane's two entries of a 90-value block: (c, j, c') of t = lane, lane + 64
  int tj[2], tdst[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = lane + h * kWave;
    const int c = t / 30, r2 = t % 30;
    tj[h] = r2 / 3;
    tdst[h] = c * 3 * L + r2 % 3;  // + 3 * offset of node column j
  }
  const bool two = lane < 90 - kWave;
  wave_lds_sync();
  const int64_t k0 = n2e_ptr[a], k1 = n2e_ptr[a + 1];
  // the node's blocks are contiguous (incidence order = element order) and
  // are added one after the other; they are loaded kGatherD at a time, the
  // next group's loads issued before the current group is added
  constexpr int D = kGatherD;
  double va[D], vb[D], na[D], nb[D];
  int oa[D], ob[D], ma[D], mb[D];
  auto load = [&](int64_t kg, double (&A)[D], double (&B)[D], int (&OA)[D], int (&OB)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t k = kg + d;
      A[d] = 0.; B[d] = 0.; OA[d] = 0; OB[d] = 0;
      if (k < k1) {
        A[d] = __builtin_nontemporal_load(&Ke[k * 90 + lane]);
        OA[d] = slot8[k * 10 + tj[0]];
        if (two) { B[d] = __builtin_nontemporal_load(&Ke[k * 90 + lane + kWave]); OB[d] = slot8[k * 10 + tj[1]]; }
      }
    }
  };
  load(k0, va, vb, oa, ob);
  for (int64_t kg = k0; kg < k1; kg += D) {
    load(kg + D, na, nb, ma, mb);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (kg + d < k1) {  // wave-uniform
        acc[tdst[0] + 3 * oa[d]] += va[d];
        if (two) acc[tdst[1] + 3 * ob[d]] += vb[d];
        wave_lds_sync();
      }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) { va[d] = na[d]; vb[d] = nb[d]; oa[d] = ma[d]; ob[d] = mb[d]; }
  }
  double *out = vals + 9 * p0;
  for (int t = lane; t < Wd; t += kWave) out[t] = acc[t];
}

// scalar P2: one wave per node row; the row's incident element rows (10
// values each, element order) added slot by slot in LDS
__global__ __launch_bounds__(256) void k_p2_gather(int64_t nn, const int64_t *__restrict__ nodeptr,
                                                   const int64_t *__restrict__ n2e_ptr,
                                                   const uint8_t *__restrict__ slot8,
                                                   const double *__restrict__ Ke,
                                                   double *__restrict__ vals) {
  __shared__ double S[256 / kWave][kNodeCap];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t a = (int64_t)blockIdx.x * (256 / kWave) + w;
  if (a >= nn) return;  // wave-uniform; the kernel has no block barrier
  double *acc = S[w];
  const int64_t p0 = nodeptr[a];
  const int L = (int)(nodeptr[a + 1] - p0);
  for (int t = lane; t < L; t += kWave) acc[t] = 0.;
  wave_lds_sync();
  const int64_t k0 = n2e_ptr[a], k1 = n2e_ptr[a + 1];
  // six incidences per step: lane = 10 u + j takes value j of incidence k + u;
  // the steps keep element order, and within a step the incidences are added
  // one after the other
  const int u = lane / 10, j = lane % 10;
  for (int64_t k = k0; k < k1; k += 6) {
    const bool mine = u < 6 && k + u < k1;
    double v = 0.;
    int o = 0;
    if (mine) { v = Ke[(k + u) * 10 + j]; o = slot8[(k + u) * 10 + j]; }
    for (int uu = 0; uu < 6; ++uu) {
      if (mine && u == uu) acc[o] += v;
      wave_lds_sync();
    }
  }
  for (int t = lane; t < L; t += kWave) vals[p0 + t] = acc[t];
}

// ---------------------------------------------------------------- Dirichlet
__device__ inline int32_t find_edge_node(const uint64_t *__restrict__ ekeys,
                                         const int32_t *__restrict__ enode, int64_t ne, uint64_t key) {
  int64_t lo = 0, hi = ne - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint64_t k = ekeys[mid];
    if (k == key) return enode[mid];
    if (k < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// dofs of the labelled boundary triangles: 3 vertex nodes and 3 edge nodes,
// all components
__global__ void k_p2v_mark_bc(int64_t nbe, const int32_t *__restrict__ btri,
                              const int32_t *__restrict__ blab, const int32_t *__restrict__ labs,
                              int nlab, const int32_t *__restrict__ vnode,
                              const uint64_t *__restrict__ ekeys, const int32_t *__restrict__ enode,
                              int64_t ne, int ncomp, uint8_t *__restrict__ onbc,
                              int *__restrict__ missing) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nbe) return;
  int on = 0;
  for (int q = 0; q < nlab; ++q) on |= blab[f] == labs[q];
  if (!on) return;
  const int32_t *t = btri + 3 * f;
  int32_t nodes[6] = {vnode[t[0]], vnode[t[1]], vnode[t[2]],
                      find_edge_node(ekeys, enode, ne, edge_key(t[0], t[1])),
                      find_edge_node(ekeys, enode, ne, edge_key(t[1], t[2])),
                      find_edge_node(ekeys, enode, ne, edge_key(t[0], t[2]))};
  for (int i = 0; i < 6; ++i) {
    if (nodes[i] < 0) { atomicMax(missing, 1); continue; }
    for (int c = 0; c < ncomp; ++c) onbc[ncomp * (int64_t)nodes[i] + c] = 1;
  }
}

// ---------------------------------------------------------------- right-hand side
// b_(3a+c) = sum over incident elements (element order) and points q of
// (|K| w_q f_c) phi_i(x_q), as Element_rhs accumulates it; tgv * g_c on
// Dirichlet rows
__global__ void k_p2v_rhs(int64_t nn, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
                          const int64_t *__restrict__ n2e_ptr, const int32_t *__restrict__ n2e,
                          int ncomp, double f0, double f1, double f2,
                          const uint8_t *__restrict__ onbc, double tgv, double g0, double g1,
                          double g2, double *__restrict__ b) {
#pragma clang fp contract(off)
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nn) return;
  const double fv[3] = {f0, f1, f2};
  double acc[3] = {0., 0., 0.};
  for (int64_t k = n2e_ptr[a]; k < n2e_ptr[a + 1]; ++k) {
    const int32_t item = n2e[k];
    const int64_t e = item / 10;
    const int i = item - (int)e * 10;
    double v[4][3];
    for (int t = 0; t < 4; ++t)
      for (int d = 0; d < 3; ++d) v[t][d] = xyz[3 * (int64_t)tets[4 * e + t] + d];
    const double mes = tet_mesure(v);
    const int ed = i < 4 ? 0 : i - 4;
    const int i0 = i < 4 ? i : c_edge[ed][0], i1 = i < 4 ? i : c_edge[ed][1];
    for (int q = 0; q < 14; ++q) {
      const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
      const double l[4] = {1. - (px + py + pz), px, py, pz};
      const double coef = mes * qf5_weight(q);
      const double phi = i < 4 ? l[i] * (2 * l[i] - 1.) : 4. * l[i0] * l[i1];
      for (int c = 0; c < ncomp; ++c) acc[c] += coef * fv[c] * phi;
    }
  }
  const double g[3] = {g0, g1, g2};
  for (int c = 0; c < ncomp; ++c)
    b[ncomp * a + c] = onbc[ncomp * a + c] ? tgv * g[c] : acc[c];
}

// onbc (3 nnodes bytes) of the labelled boundary; false if a node was missing
static bool p2v_mark(ffhip_ctx ctx, ffhip_p2v sp, const int32_t *labels, int nlab, int ncomp,
                     uint8_t *onbc) {
  hipStream_t s = ctx->stream;
  ffhip_mesh m = sp->mesh;
  int32_t *labs = dalloc<int32_t>(nlab > 0 ? nlab : 1);
  int *missing = dalloc<int>(1);
  if (nlab) FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * nlab, hipMemcpyHostToDevice, s));
  FFHIP_CHECK(hipMemsetAsync(onbc, 0, ncomp * sp->nnodes, s));
  FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), s));
  if (m->nbe)
    k_p2v_mark_bc<<<nblk(m->nbe, 256), 256, 0, s>>>(m->nbe, m->belems, m->blab, labs, nlab,
                                                    sp->vnode, sp->ekeys, sp->enode, sp->ne, ncomp,
                                                    onbc, missing);
  FFHIP_CHECK(hipGetLastError());
  const int miss = read_flag(ctx, missing);
  dfree(labs);
  dfree(missing);
  return miss == 0;
}

static void p2v_release(ffhip_p2v sp) {
  dfree(sp->elem_nodes); dfree(sp->vnode); dfree(sp->ekeys); dfree(sp->enode);
  dfree(sp->n2e_ptr); dfree(sp->n2e); dfree(sp->n2e_inv);
  delete sp;
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_p2v_create(ffhip_ctx ctx, ffhip_mesh m, ffhip_p2v *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && out, FFHIP_ERR_ARG, "ffhip_p2v_create: null argument");
    FFHIP_REQUIRE(m->dim == 3 && m->nloc == 4, FFHIP_ERR_UNSUPPORTED,
                  "ffhip_p2v_create: needs a tetrahedral mesh");
    FFHIP_REQUIRE(m->nt > 0 && 10 * m->nt < INT_MAX, FFHIP_ERR_UNSUPPORTED,
                  "ffhip_p2v_create: %lld elements (need 10 nt < 2^31)", (long long)m->nt);
    hipStream_t s = ctx->stream;
    const int64_t nt = m->nt, nv = m->nv, E = 6 * nt, NI = 10 * nt;
    int vb = 1;  // bits of a vertex id
    while (vb < 31 && (int64_t(1) << vb) < nv) ++vb;
    ffhip_p2v sp = new ffhip_p2v_s();
    sp->ctx = ctx; sp->mesh = m;
    try {
      size_t tb = 0;
      FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                                     (int32_t *)nullptr, (int32_t *)nullptr, (int)E, 0,
                                                     2 * vb, s));
      Arena ar(ctx, Arena::bytes(nv, 4) + 2 * Arena::bytes(E, 8) + 4 * Arena::bytes(E, 4) +
                        Arena::bytes(tb, 1) + 2 * Arena::bytes(E + 1, 8) + 2 * Arena::bytes(NI + 1, 8));
      ctx->tic();
      int32_t *firstv = ar.take<int32_t>(nv);
      k_fill_i32<<<nblk(nv, 256), 256, 0, s>>>(nv, INT_MAX, firstv);
      k_p2_vfirst<<<nblk(4 * nt, 256), 256, 0, s>>>(nt, m->elems, firstv);
      uint64_t *keys = ar.take<uint64_t>(E), *skeys = ar.take<uint64_t>(E);
      int32_t *items = ar.take<int32_t>(E), *sitems = ar.take<int32_t>(E);
      k_p2_edge_keys<<<nblk(E, 256), 256, 0, s>>>(nt, m->elems, vb, keys, items);
      void *tmp = ar.take<char>(tb);
      FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, skeys, items, sitems, (int)E, 0,
                                                     2 * vb, s));
      int64_t *head = ar.take<int64_t>(E + 1), *hscan = ar.take<int64_t>(E + 1);
      int64_t *mark = ar.take<int64_t>(NI + 1), *rank = ar.take<int64_t>(NI + 1);
      FFHIP_CHECK(hipMemsetAsync(head + E, 0, 8, s));
      FFHIP_CHECK(hipMemsetAsync(mark + NI, 0, 8, s));
      k_p2_edge_heads<<<nblk(E, 256), 256, 0, s>>>(E, skeys, sitems, head, mark);
      k_p2_vmark<<<nblk(4 * nt, 256), 256, 0, s>>>(nt, m->elems, firstv, mark);
      exclusive_scan_i64(ctx, head, hscan, E + 1);
      sp->ne = read_i64(ctx, hscan + E);
      int32_t *item_seg = ar.take<int32_t>(E);
      int32_t *segfirst = ar.take<int32_t>(E);  // ne <= E
      sp->ekeys = dalloc<uint64_t>(sp->ne);
      k_p2_edge_seg<<<nblk(E, 256), 256, 0, s>>>(E, skeys, vb, sitems, hscan, head, item_seg, sp->ekeys,
                                                 segfirst);
      exclusive_scan_i64(ctx, mark, rank, NI + 1);
      sp->nnodes = read_i64(ctx, rank + NI);
      sp->elem_nodes = dalloc<int32_t>(NI);
      k_p2_assign<<<nblk(NI, 256), 256, 0, s>>>(nt, m->elems, firstv, item_seg, segfirst, rank,
                                                sp->elem_nodes);
      sp->vnode = dalloc<int32_t>(nv);
      k_p2_vnode<<<nblk(nv, 256), 256, 0, s>>>(nv, firstv, rank, sp->vnode);
      sp->enode = dalloc<int32_t>(sp->ne);
      k_p2_enode<<<nblk(sp->ne, 256), 256, 0, s>>>(sp->ne, segfirst, rank, sp->enode);
      FFHIP_CHECK(hipGetLastError());
      FFHIP_CHECK(hipStreamSynchronize(s));
      // the arena's temporaries are dead from here on (build_incidence reuses the work buffer)
      build_incidence(ctx, sp->elem_nodes, nt, 10, sp->nnodes, &sp->n2e_ptr, &sp->n2e, nullptr,
                      &sp->n2e_inv, nullptr);
      ctx->toc();
    } catch (...) {
      (void)hipStreamSynchronize(s);
      p2v_release(sp);
      throw;
    }
    *out = sp;
  });
}

int ffhip_p2v_info(ffhip_p2v sp, int64_t *nnodes, int64_t *ndof) {
  return guard([&] {
    FFHIP_REQUIRE(sp, FFHIP_ERR_ARG, "ffhip_p2v_info: null space");
    if (nnodes) *nnodes = sp->nnodes;
    if (ndof) *ndof = 3 * sp->nnodes;
  });
}

int ffhip_p2v_download(ffhip_p2v sp, int32_t *elem_nodes) {
  return guard([&] {
    FFHIP_REQUIRE(sp && elem_nodes, FFHIP_ERR_ARG, "ffhip_p2v_download: null argument");
    hipStream_t s = sp->ctx->stream;
    FFHIP_CHECK(hipMemcpyAsync(elem_nodes, sp->elem_nodes, 4 * 10 * sp->mesh->nt,
                               hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
  });
}

int ffhip_p2v_free(ffhip_p2v sp) {
  return guard([&] {
    if (!sp) return;
    (void)hipStreamSynchronize(sp->ctx->stream);
    p2v_release(sp);
  });
}

}  // extern "C"

namespace ffhip {
// the node pattern of the P2 space (one wave per node row, the lane-per-row
// merge if a row does not fit), then the matrix rows: 3 x 3 blocks of the node
// columns (ncomp = 3) or the node rows themselves (ncomp = 1)
static void p2_pattern(ffhip_ctx ctx, ffhip_p2v sp, int ncomp, ffhip_mat *out) {
    hipStream_t s = ctx->stream;
    ffhip_mesh m = sp->mesh;
    const int64_t nn = sp->nnodes, n = ncomp * nn;
    FFHIP_REQUIRE(n < INT_MAX, FFHIP_ERR_UNSUPPORTED, "ffhip_pattern_p2: too many dofs");
    ffhip_mat A = new ffhip_mat_s();
    A->ctx = ctx; A->n = n; A->p2v = sp; A->ncomp = ncomp;
    try {
      Arena ar(ctx, Arena::bytes(nn + 1, 8) + Arena::bytes(1, 4) + Arena::bytes(nn, 8) +
                        Arena::bytes(kNodeCap * nn, 4) + Arena::bytes(nn, 1));
      ctx->tic();
      int64_t *len = ar.take<int64_t>(nn + 1);
      int *ovf = ar.take<int>(1);
      A->nodeptr = dalloc<int64_t>(nn + 1);
      int64_t *ndiag = ar.take<int64_t>(nn);
      int32_t *rowbuf = ar.take<int32_t>(kNodeCap * nn);
      int8_t *dposr = ar.take<int8_t>(nn);
      A->slot8 = dalloc<uint8_t>(m->nt * 100);
      FFHIP_CHECK(hipMemsetAsync(len, 0, 8 * (nn + 1), s));
      FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
      // one wave per node row (hash set in LDS); the lane-per-row merge if a
      // row does not fit
      k_p2_pattern_wave<<<nblk(nn * kWave, 256), 256, 0, s>>>(nn, sp->elem_nodes, sp->n2e_ptr, sp->n2e,
                                                              len, rowbuf, dposr, A->slot8, ovf);
      FFHIP_CHECK(hipGetLastError());
      const bool wave_ok = read_flag(ctx, ovf) == 0;
      constexpr int B = 128, CAP = kNodeCap;
      if (!wave_ok) {
        FFHIP_CHECK(hipMemsetAsync(len, 0, 8 * (nn + 1), s));
        FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
        k_pattern_rows<B, CAP, false><<<nblk(nn, B), B, 0, s>>>(
            nn, 10, sp->elem_nodes, sp->n2e_ptr, sp->n2e, len, nullptr, nullptr, nullptr, nullptr,
            nullptr, ovf);
        FFHIP_CHECK(hipGetLastError());
        FFHIP_REQUIRE(read_flag(ctx, ovf) == 0, FFHIP_ERR_UNSUPPORTED,
                      "ffhip_pattern_p2v: a node couples to more than %d nodes", CAP);
      }
      exclusive_scan_i64(ctx, len, A->nodeptr, nn + 1);
      const int64_t nnz_nodes = read_i64(ctx, A->nodeptr + nn);
      int32_t *ncol = nullptr;
      if (!wave_ok) {
        ncol = rowbuf;  // the row buffers are free: nnz_nodes <= kNodeCap * nn
        k_pattern_rows<B, CAP, true><<<nblk(nn, B), B, 0, s>>>(
            nn, 10, sp->elem_nodes, sp->n2e_ptr, sp->n2e, nullptr, A->nodeptr, ncol, ndiag, nullptr,
            A->slot8, ovf);
      }
      A->nnz = (int64_t)ncomp * ncomp * nnz_nodes;
      A->rowptr = dalloc<int64_t>(n + 1);
      A->colidx = dalloc<int32_t>(A->nnz);
      A->diag = dalloc<int64_t>(n);
      A->vals = dalloc<double>(A->nnz);
      if (ncomp == 3)
        k_p2v_expand<<<nblk(nn * kWave, 256), 256, 0, s>>>(nn, A->nodeptr, ncol, ndiag,
                                                  wave_ok ? rowbuf : nullptr, dposr, A->rowptr,
                                                  A->colidx, A->diag);
      else {
        if (!wave_ok)  // the merge wrote the columns at their CSR positions
          FFHIP_CHECK(hipMemcpyAsync(A->colidx, ncol, 4 * nnz_nodes, hipMemcpyDeviceToDevice, s));
        k_p2_compact<<<nblk(nn, 256), 256, 0, s>>>(nn, A->nodeptr, ncol, ndiag,
                                                   wave_ok ? rowbuf : nullptr, dposr, A->rowptr,
                                                   A->colidx, A->diag);
      }
      FFHIP_CHECK(hipMemsetAsync(A->vals, 0, 8 * A->nnz, s));
      FFHIP_CHECK(hipGetLastError());
      ctx->toc();
      A->max_row = ncomp * CAP;
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dfree(A->rowptr); dfree(A->colidx); dfree(A->vals); dfree(A->diag);
      dfree(A->nodeptr); dfree(A->slot8);
      delete A;
      throw;
    }
    *out = A;
}
}  // namespace ffhip

extern "C" {

int ffhip_pattern_p2v(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && out, FFHIP_ERR_ARG, "ffhip_pattern_p2v: null argument");
    p2_pattern(ctx, sp, 3, out);
  });
}

int ffhip_pattern_p2(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && out, FFHIP_ERR_ARG, "ffhip_pattern_p2: null argument");
    p2_pattern(ctx, sp, 1, out);
  });
}

int ffhip_assemble_p2v_elasticity(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat A, double lambda,
                                  double mu) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && A && A->p2v == sp && A->slot8 && A->ncomp == 3, FFHIP_ERR_ARG,
                  "ffhip_assemble_p2v_elasticity: matrix pattern was not built from this space");
    hipStream_t s = ctx->stream;
    ffhip_mesh m = sp->mesh;
    Arena ar(ctx, Arena::bytes(900 * m->nt, 8));
    double *Ke = ar.take<double>(900 * m->nt);
    int64_t g = (m->nt + kElemWaves - 1) / kElemWaves;
    const int64_t gmax = 8 * (int64_t)ctx->num_cu;
    if (g > gmax) g = gmax;
    ctx->tic();
    k_p2_element<false><<<(unsigned)g, kElemBlock, 0, s>>>(m->nt, m->xyz, m->elems, lambda, mu, 0,
                                                          0., nullptr, sp->n2e_inv, Ke);
    k_p2v_gather<<<nblk(sp->nnodes, 256 / kWave), 256, 0, s>>>(sp->nnodes, A->nodeptr, sp->n2e_ptr,
                                                              A->slot8, Ke, A->vals);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
  });
}

int ffhip_set_bc_p2v(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat A, const int32_t *labels, int nlab,
                     double tgv, uint8_t *onbc_out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && A && A->n == 3 * sp->nnodes, FFHIP_ERR_ARG,
                  "ffhip_set_bc_p2v: bad arguments");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_set_bc_p2v: labels");
    FFHIP_REQUIRE(tgv > 0, FFHIP_ERR_UNSUPPORTED, "ffhip_set_bc_p2v: only tgv > 0 is supported");
    uint8_t *onbc = dalloc<uint8_t>(A->n);
    int *missing = dalloc<int>(1);
    bool ok = false;
    int miss = 0;
    try {
      ctx->tic();
      ok = p2v_mark(ctx, sp, labels, nlab, 3, onbc);
      FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), ctx->stream));
      set_diag_rows(ctx, A->n, onbc, A->diag, tgv, A->vals, missing);
      ctx->toc();
      if (onbc_out) {
        FFHIP_CHECK(hipMemcpyAsync(onbc_out, onbc, A->n, hipMemcpyDeviceToHost, ctx->stream));
      }
      miss = read_flag(ctx, missing);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      dfree(onbc); dfree(missing);
      throw;
    }
    dfree(onbc); dfree(missing);
    FFHIP_REQUIRE(ok, FFHIP_ERR_INTERNAL, "ffhip_set_bc_p2v: a boundary edge is no element edge");
    FFHIP_REQUIRE(miss == 0, FFHIP_ERR_INTERNAL, "ffhip_set_bc_p2v: a Dirichlet row has no diagonal");
  });
}

int ffhip_rhs_p2v(ffhip_ctx ctx, ffhip_p2v sp, const double f[3], const int32_t *labels, int nlab,
                  double tgv, const double g[3], double *b, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && f && b, FFHIP_ERR_ARG, "ffhip_rhs_p2v: null argument");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_rhs_p2v: labels");
    hipStream_t s = ctx->stream;
    ffhip_mesh m = sp->mesh;
    const int64_t n = 3 * sp->nnodes;
    const double gz[3] = {0., 0., 0.};
    const double *gg = g ? g : gz;
    uint8_t *onbc = dalloc<uint8_t>(n);
    double *bd = on_device ? b : nullptr;
    bool ok = false;
    try {
      ok = p2v_mark(ctx, sp, labels, nlab, 3, onbc);
      if (!on_device) bd = dalloc<double>(n);
      ctx->tic();
      k_p2v_rhs<<<nblk(sp->nnodes, 256), 256, 0, s>>>(sp->nnodes, m->xyz, m->elems, sp->n2e_ptr,
                                                      sp->n2e, 3, f[0], f[1], f[2], onbc, tgv, gg[0],
                                                      gg[1], gg[2], bd);
      FFHIP_CHECK(hipGetLastError());
      ctx->toc();
      if (!on_device) {
        FFHIP_CHECK(hipMemcpyAsync(b, bd, 8 * n, hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipStreamSynchronize(s));
      }
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dfree(onbc);
      if (!on_device) dfree(bd);
      throw;
    }
    dfree(onbc);
    if (!on_device) dfree(bd);
    FFHIP_REQUIRE(ok, FFHIP_ERR_INTERNAL, "ffhip_rhs_p2v: a boundary edge is no element edge");
  });
}

// ---- scalar P2 (Vh(Th,P2), dof = node)
int ffhip_assemble_p2(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat A, int kind, double c0,
                      const double *kappa_e, int kappa_on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && A && A->p2v == sp && A->slot8 && A->ncomp == 1, FFHIP_ERR_ARG,
                  "ffhip_assemble_p2: matrix pattern was not built by ffhip_pattern_p2 from this space");
    FFHIP_REQUIRE(kind == FFHIP_COEF_CONST || kind == FFHIP_COEF_SIN3 || kind == FFHIP_COEF_P0 ||
                      kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1,
                  FFHIP_ERR_ARG, "ffhip_assemble_p2: coefficient kind %d", kind);
    const bool arr = kind == FFHIP_COEF_P0 || kind == FFHIP_COEF_QP || kind == FFHIP_COEF_P1;
    FFHIP_REQUIRE(!arr || kappa_e, FFHIP_ERR_ARG, "ffhip_assemble_p2: kappa_e");
    hipStream_t s = ctx->stream;
    ffhip_mesh m = sp->mesh;
    const int64_t cnt = kind == FFHIP_COEF_P1 ? m->nv : kind == FFHIP_COEF_P0 ? m->nt : 14 * m->nt;
    Arena ar(ctx, Arena::bytes(100 * m->nt, 8) + (arr && !kappa_on_device ? Arena::bytes(cnt, 8) : 0));
    double *Ke = ar.take<double>(100 * m->nt);
    const double *kd = kappa_e;
    if (arr && !kappa_on_device) {
      double *kb = ar.take<double>(cnt);
      FFHIP_CHECK(hipMemcpyAsync(kb, kappa_e, 8 * cnt, hipMemcpyHostToDevice, s));
      kd = kb;
    }
    int64_t g = (m->nt + kElemWaves - 1) / kElemWaves;
    const int64_t gmax = 8 * (int64_t)ctx->num_cu;
    if (g > gmax) g = gmax;
    ctx->tic();
    k_p2_element<true><<<(unsigned)g, kElemBlock, 0, s>>>(m->nt, m->xyz, m->elems, 0., 0., kind, c0,
                                                          kd, sp->n2e_inv, Ke);
    k_p2_gather<<<nblk(sp->nnodes, 256 / kWave), 256, 0, s>>>(sp->nnodes, A->nodeptr, sp->n2e_ptr,
                                                             A->slot8, Ke, A->vals);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
  });
}

int ffhip_set_bc_p2(ffhip_ctx ctx, ffhip_p2v sp, ffhip_mat A, const int32_t *labels, int nlab,
                    double tgv, uint8_t *onbc_out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && A && A->ncomp == 1 && A->n == sp->nnodes, FFHIP_ERR_ARG,
                  "ffhip_set_bc_p2: bad arguments");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_set_bc_p2: labels");
    FFHIP_REQUIRE(tgv > 0, FFHIP_ERR_UNSUPPORTED, "ffhip_set_bc_p2: only tgv > 0 is supported");
    uint8_t *onbc = dalloc<uint8_t>(A->n);
    int *missing = dalloc<int>(1);
    bool ok = false;
    int miss = 0;
    try {
      ctx->tic();
      ok = p2v_mark(ctx, sp, labels, nlab, 1, onbc);
      FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), ctx->stream));
      set_diag_rows(ctx, A->n, onbc, A->diag, tgv, A->vals, missing);
      ctx->toc();
      if (onbc_out)
        FFHIP_CHECK(hipMemcpyAsync(onbc_out, onbc, A->n, hipMemcpyDeviceToHost, ctx->stream));
      miss = read_flag(ctx, missing);
    } catch (...) {
      (void)hipStreamSynchronize(ctx->stream);
      dfree(onbc); dfree(missing);
      throw;
    }
    dfree(onbc); dfree(missing);
    FFHIP_REQUIRE(ok, FFHIP_ERR_INTERNAL, "ffhip_set_bc_p2: a boundary edge is no element edge");
    FFHIP_REQUIRE(miss == 0, FFHIP_ERR_INTERNAL, "ffhip_set_bc_p2: a Dirichlet row has no diagonal");
  });
}

int ffhip_rhs_p2(ffhip_ctx ctx, ffhip_p2v sp, double f, const int32_t *labels, int nlab, double tgv,
                 double g, double *b, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && sp && b, FFHIP_ERR_ARG, "ffhip_rhs_p2: null argument");
    FFHIP_REQUIRE(nlab >= 0 && (nlab == 0 || labels), FFHIP_ERR_ARG, "ffhip_rhs_p2: labels");
    hipStream_t s = ctx->stream;
    ffhip_mesh m = sp->mesh;
    const int64_t n = sp->nnodes;
    uint8_t *onbc = dalloc<uint8_t>(n);
    double *bd = on_device ? b : nullptr;
    bool ok = false;
    try {
      ok = p2v_mark(ctx, sp, labels, nlab, 1, onbc);
      if (!on_device) bd = dalloc<double>(n);
      ctx->tic();
      k_p2v_rhs<<<nblk(n, 256), 256, 0, s>>>(n, m->xyz, m->elems, sp->n2e_ptr, sp->n2e, 1, f, 0., 0.,
                                             onbc, tgv, g, 0., 0., bd);
      FFHIP_CHECK(hipGetLastError());
      ctx->toc();
      if (!on_device) {
        FFHIP_CHECK(hipMemcpyAsync(b, bd, 8 * n, hipMemcpyDeviceToHost, s));
        FFHIP_CHECK(hipStreamSynchronize(s));
      }
    } catch (...) {
      (void)hipStreamSynchronize(s);
      dfree(onbc);
      if (!on_device) dfree(bd);
      throw;
    }
    dfree(onbc);
    if (!on_device) dfree(bd);
    FFHIP_REQUIRE(ok, FFHIP_ERR_INTERNAL, "ffhip_rhs_p2: a boundary edge is no element edge");
  });
}

}  // extern "C"
