// ffhip_assembly.hip — P1 sparsity pattern, atomic-free assembly, Dirichlet
// rows, right-hand side and CSR mat-vec.
//
// Design (MI355X-first, deterministic):
//   1. dof -> (element, local vertex) incidence lists, sorted by element id
//      (counts by integer atomics, a scan, and a per-dof sort: deterministic).
//   2. pattern: one lane per row merges the vertex sets of its incident
//      elements into a sorted unique list held in LDS; a count pass, a scan,
//      a fill pass.  The fill pass also records, for every incidence entry,
//      the offsets of that element's vertices inside the row — the
//      element-to-row map.
//   3. values: one thread per element evaluates the 14-point Tet_5 integral
//      of kappa and the P1 gradients (VALU FP64) and writes the 10 unique
//      entries of its symmetric 4x4 matrix; then one lane per row walks its
//      incidence list in element order and accumulates into the row's slots
//      staged in LDS, which the block writes back coalesced.  No atomics:
//      every slot is owned by one lane and summed in element order, the
//      order in which HashMatrix::operator+= accumulates
//      (src/femlib/HashMatrix.cpp:1294-1332).
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <vector>

#include "ffhip_fem.hpp"

namespace ffhip {

// T_5 (src/femlib/QuadratureFormular.cpp:172-186), filled from the host
__constant__ double c_qt5_pts[7][2];
__constant__ double c_qt5_w[7];

// ---------------------------------------------------------------- incidence
__global__ void k_v2e_count(const int32_t *__restrict__ elems, int64_t m,
                            unsigned long long *__restrict__ deg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) atomicAdd(&deg[elems[i]], 1ull);
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

// per-dof insertion sort: incidence entries in element order
__global__ void k_v2e_sort(int64_t n, const int64_t *__restrict__ ptr, int32_t *__restrict__ v2e) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t a = ptr[v], b = ptr[v + 1];
  for (int64_t i = a + 1; i < b; ++i) {
    const int32_t key = v2e[i];
    int64_t j = i - 1;
    while (j >= a && v2e[j] > key) { v2e[j + 1] = v2e[j]; --j; }
    v2e[j + 1] = key;
  }
}

// ---------------------------------------------------------------- element matrices
// 3D P1: 10 unique entries of Ke = sum_q |K| w_q kappa(x_q) grad l_i . grad l_j
__global__ __launch_bounds__(256) void k_p1_element_3d(
    int64_t nt, const double *__restrict__ xyz, const int32_t *__restrict__ tets,
    int kind, double c0, const double *__restrict__ kappa_e, double *__restrict__ Ke) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
  double v[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[i][k] = xyz[3 * (int64_t)id[i] + k];
  double G[4][3];
  const double mes = tet_geom(v, G);
  // S = sum_q coef_q * kappa(x_q), coef_q = |K| w_q  (problem.cpp:6387-6392)
  double S;
  if (kind == FFHIP_COEF_SIN3) {
    S = 0.;
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
    S = 0.;
#pragma unroll
    for (int q = 0; q < 14; ++q) S += (mes * qf5_weight(q)) * kappa_e[14 * e + q];
  } else {
    const double kc = (kind == FFHIP_COEF_P0) ? kappa_e[e] : c0;
    S = 0.;
#pragma unroll
    for (int q = 0; q < 14; ++q) S += (mes * qf5_weight(q)) * kc;
  }
  double out[10];
  int o = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) {
      double a = 0.;
#pragma unroll
      for (int k = 0; k < 3; ++k) a += (S * G[i][k]) * G[j][k];
      out[o++] = a;
    }
  double *dst = Ke + 10 * e;
#pragma unroll
  for (int i = 0; i < 10; ++i) dst[i] = out[i];
}

// 2D P1: 6 unique entries; TTriangle area and H(i) (fem.hpp:290-325)
__global__ __launch_bounds__(256) void k_p1_element_2d(
    int64_t nt, const double *__restrict__ xy, const int32_t *__restrict__ tris,
    int kind, double c0, const double *__restrict__ kappa_e, double *__restrict__ Ke) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  double v[3][2];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 2; ++k) v[i][k] = xy[2 * (int64_t)tris[3 * e + i] + k];
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
    } else {
      kc = (kind == FFHIP_COEF_P0) ? kappa_e[e] : c0;
    }
    S += (area * c_qt5_w[q]) * kc;
  }
  int o = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      double a = 0.;
      for (int k = 0; k < 2; ++k) a += (S * G[i][k]) * G[j][k];
      Ke[6 * e + o++] = a;
    }
}

// ---------------------------------------------------------------- row gather
// Block of BLOCK rows; the block's CSR segment is staged in LDS when it fits
// (STAGE doubles), else each lane accumulates its own row in global memory.
template <int BLOCK, int STAGE, int NLOC>
__global__ __launch_bounds__(BLOCK) void k_p1_gather(
    int64_t n, const int64_t *__restrict__ rowptr, const int64_t *__restrict__ v2e_ptr,
    const int32_t *__restrict__ v2e, const uint32_t *__restrict__ slotmap,
    const double *__restrict__ Ke, double *__restrict__ vals) {
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
  if (v < n) {
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
    for (int64_t i = t; i < seg1 - seg0; i += BLOCK) vals[seg0 + i] = S[i];
  }
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
// b_v = sum over incident elements (element order) of
//       sum_q (|K| w_q f) lambda_l(x_q)   (Element_rhs, problem.cpp:7940-7978)
__global__ void k_rhs_p1_3d(int64_t n, const double *__restrict__ xyz,
                            const int32_t *__restrict__ tets, const int64_t *__restrict__ v2e_ptr,
                            const int32_t *__restrict__ v2e, double f,
                            const uint8_t *__restrict__ onbc, double tgvg, double *__restrict__ b) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (onbc && onbc[v]) { b[v] = tgvg; return; }
  double s = 0.;
  for (int64_t k = v2e_ptr[v]; k < v2e_ptr[v + 1]; ++k) {
    const int32_t idx = v2e[k];
    const int64_t e = idx >> 2;
    const int l = idx & 3;
    const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
    const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
    double vv[4][3];
    for (int i = 0; i < 4; ++i)
      for (int kk = 0; kk < 3; ++kk) vv[i][kk] = xyz[3 * (int64_t)id[i] + kk];
    const double mes = tet_mesure(vv);
    double r = 0.;
    for (int q = 0; q < 14; ++q) {
      const double px = c_qf5_pts[q][0], py = c_qf5_pts[q][1], pz = c_qf5_pts[q][2];
      const double lam = l == 0 ? 1. - (px + py + pz) : (l == 1 ? px : (l == 2 ? py : pz));
      r += (mes * qf5_weight(q)) * f * lam;
    }
    s += r;
  }
  b[v] = s;
}

__global__ void k_rhs_p1_2d(int64_t n, const double *__restrict__ xy,
                            const int32_t *__restrict__ tris, const int64_t *__restrict__ v2e_ptr,
                            const int32_t *__restrict__ v2e, double f,
                            const uint8_t *__restrict__ onbc, double tgvg, double *__restrict__ b) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (onbc && onbc[v]) { b[v] = tgvg; return; }
  double s = 0.;
  for (int64_t k = v2e_ptr[v]; k < v2e_ptr[v + 1]; ++k) {
    const int32_t idx = v2e[k];
    const int64_t e = idx / 3;
    const int l = idx - 3 * (int)e;
    double p[3][2];
    for (int i = 0; i < 3; ++i)
      for (int kk = 0; kk < 2; ++kk) p[i][kk] = xy[2 * (int64_t)tris[3 * e + i] + kk];
    const double area = ((p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) -
                         (p[1][1] - p[0][1]) * (p[2][0] - p[0][0])) * 0.5;
    double r = 0.;
    for (int q = 0; q < 7; ++q) {
      const double px = c_qt5_pts[q][0], py = c_qt5_pts[q][1];
      const double lam = l == 0 ? 1 - px - py : (l == 1 ? px : py);
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
__global__ void k_p1_mass_apply(int64_t n, const double *__restrict__ xyz,
                                const int32_t *__restrict__ tets, const int64_t *__restrict__ v2e_ptr,
                                const int32_t *__restrict__ v2e, const double *__restrict__ u,
                                const double *__restrict__ ud, const uint8_t *__restrict__ onbc,
                                double *__restrict__ r) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (onbc && onbc[v]) { r[v] = 0.; return; }
  const double ev = u[v] - ud[v];
  double s = 0.;
  for (int64_t k = v2e_ptr[v]; k < v2e_ptr[v + 1]; ++k) {
    const int64_t e = v2e[k] >> 2;
    const int4 tv = reinterpret_cast<const int4 *>(tets)[e];
    const int32_t id[4] = {tv.x, tv.y, tv.z, tv.w};
    double vv[4][3], se = 0.;
    for (int i = 0; i < 4; ++i) {
      for (int kk = 0; kk < 3; ++kk) vv[i][kk] = xyz[3 * (int64_t)id[i] + kk];
      se += u[id[i]] - ud[id[i]];
    }
    s += tet_mesure(vv) / 20. * (ev + se);
  }
  r[v] = s;
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

static void upload_t5_once() {
  static bool done = false;
  if (done) return;
  const double sqrt15 = 3.87298334620741688517926539978;
  const double t = 1.E0 / 3.E0, A = 0.225E0;
  const double r = (6 - sqrt15) / 21, s = (9 + 2 * sqrt15) / 21, B = (155 - sqrt15) / 1200;
  const double u = (6 + sqrt15) / 21, v = (9 - 2 * sqrt15) / 21, C = (155 + sqrt15) / 1200;
  const double P[7][2] = {{t, t}, {r, r}, {r, s}, {s, r}, {u, u}, {u, v}, {v, u}};
  const double W[7] = {A, B, B, B, C, C, C};
  FFHIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_qt5_pts), P, sizeof(P)));
  FFHIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_qt5_w), W, sizeof(W)));
  done = true;
}

// dof -> (element * nloc + local) incidence lists in element order, for any
// element-to-dof table (P1 vertices, P2 nodes)
void build_incidence(ffhip_ctx ctx, const int32_t *elems, int64_t nt, int nloc, int64_t n,
                     int64_t **ptr_out, int32_t **list_out) {
  hipStream_t s = ctx->stream;
  const int64_t M = nt * nloc;
  unsigned long long *deg = dalloc<unsigned long long>(n + 1);
  int32_t *cur = dalloc<int32_t>(n);
  int64_t *ptr = dalloc<int64_t>(n + 1);
  int32_t *list = dalloc<int32_t>(M);
  FFHIP_CHECK(hipMemsetAsync(deg, 0, 8 * (n + 1), s));
  FFHIP_CHECK(hipMemsetAsync(cur, 0, 4 * n, s));
  k_v2e_count<<<nblocks(M, 256), 256, 0, s>>>(elems, M, deg);
  exclusive_scan_i64(ctx, (const int64_t *)deg, ptr, n + 1);
  k_v2e_fill<<<nblocks(M, 256), 256, 0, s>>>(elems, M, ptr, cur, list);
  k_v2e_sort<<<nblocks(n, 256), 256, 0, s>>>(n, ptr, list);
  FFHIP_CHECK(hipGetLastError());
  FFHIP_CHECK(hipStreamSynchronize(s));
  dfree(deg);
  dfree(cur);
  *ptr_out = ptr;
  *list_out = list;
}

// vertex -> incidence lists of a mesh (built once per mesh)
void build_v2e(ffhip_mesh m) {
  if (m->v2e_ptr) return;
  build_incidence(m->ctx, m->elems, m->nt, m->nloc, m->nv, &m->v2e_ptr, &m->v2e);
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

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_pattern_p1(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && out, FFHIP_ERR_ARG, "ffhip_pattern_p1: null argument");
    FFHIP_REQUIRE(m->nloc <= 4, FFHIP_ERR_UNSUPPORTED, "ffhip_pattern_p1: P1 on simplices only");
    hipStream_t s = ctx->stream;
    ctx->tic();
    build_v2e(m);
    const int64_t n = m->nv;
    ffhip_mat A = new ffhip_mat_s();
    A->ctx = ctx; A->n = n; A->mesh = m;
    int64_t *len = dalloc<int64_t>(n + 1);
    int *ovf = dalloc<int>(1);
    A->rowptr = dalloc<int64_t>(n + 1);
    A->diag = dalloc<int64_t>(n);
    A->slotmap = dalloc<uint32_t>(m->nt * m->nloc);
    FFHIP_CHECK(hipMemsetAsync(len, 0, 8 * (n + 1), s));
    FFHIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), s));
    constexpr int B = 128, CAP = 64;
    k_pattern_rows<B, CAP, false><<<nblocks(n, B), B, 0, s>>>(
        n, m->nloc, m->elems, m->v2e_ptr, m->v2e, len, nullptr, nullptr, nullptr, nullptr, nullptr,
        ovf);
    FFHIP_CHECK(hipGetLastError());
    FFHIP_REQUIRE(read_flag(ctx, ovf) == 0, FFHIP_ERR_UNSUPPORTED,
                  "ffhip_pattern_p1: a row has more than %d columns", CAP);
    exclusive_scan_i64(ctx, len, A->rowptr, n + 1);
    A->nnz = read_i64(ctx, A->rowptr + n);
    A->colidx = dalloc<int32_t>(A->nnz);
    A->vals = dalloc<double>(A->nnz);
    k_pattern_rows<B, CAP, true><<<nblocks(n, B), B, 0, s>>>(
        n, m->nloc, m->elems, m->v2e_ptr, m->v2e, nullptr, A->rowptr, A->colidx, A->diag,
        A->slotmap, nullptr, ovf);
    FFHIP_CHECK(hipGetLastError());
    FFHIP_CHECK(hipMemsetAsync(A->vals, 0, 8 * A->nnz, s));
    ctx->toc();
    A->max_row = CAP;
    dfree(len);
    dfree(ovf);
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
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && A && A->mesh == m && A->slotmap, FFHIP_ERR_ARG,
                  "ffhip_assemble_p1: matrix pattern was not built from this mesh");
    FFHIP_REQUIRE(kind == FFHIP_COEF_CONST || kind == FFHIP_COEF_SIN3 || kind == FFHIP_COEF_P0 ||
                      kind == FFHIP_COEF_QP,
                  FFHIP_ERR_ARG, "ffhip_assemble_p1: coefficient kind %d", kind);
    const bool arr = kind == FFHIP_COEF_P0 || kind == FFHIP_COEF_QP;
    FFHIP_REQUIRE(!arr || kappa_e, FFHIP_ERR_ARG, "ffhip_assemble_p1: kappa_e");
    hipStream_t s = ctx->stream;
    const int nu = m->nloc == 4 ? 10 : 6;
    double *Ke = dalloc<double>(nu * m->nt);
    const double *kd = kappa_e;
    double *kbuf = nullptr;
    if (arr && !kappa_on_device) {
      const int64_t cnt = kind == FFHIP_COEF_P0 ? m->nt : m->nt * (m->dim == 3 ? 14 : 7);
      kbuf = dalloc<double>(cnt);
      FFHIP_CHECK(hipMemcpyAsync(kbuf, kappa_e, 8 * cnt, hipMemcpyHostToDevice, s));
      kd = kbuf;
    }
    if (m->dim == 2) upload_t5_once();
    ctx->tic();
    if (m->dim == 3)
      k_p1_element_3d<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, m->xyz, m->elems, kind, c0, kd, Ke);
    else
      k_p1_element_2d<<<nblocks(m->nt, 256), 256, 0, s>>>(m->nt, m->xyz, m->elems, kind, c0, kd, Ke);
    constexpr int B = 256, STAGE = 6144;
    if (m->nloc == 4)
      k_p1_gather<B, STAGE, 4><<<nblocks(A->n, B), B, 0, s>>>(A->n, A->rowptr, m->v2e_ptr, m->v2e,
                                                               A->slotmap, Ke, A->vals);
    else
      k_p1_gather<B, STAGE, 3><<<nblocks(A->n, B), B, 0, s>>>(A->n, A->rowptr, m->v2e_ptr, m->v2e,
                                                               A->slotmap, Ke, A->vals);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    dfree(Ke);
    dfree(kbuf);
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

int ffhip_rhs_p1(ffhip_ctx ctx, ffhip_mesh m, double f, const int32_t *labels, int nlab,
                 double tgv, double g, double *b, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && b, FFHIP_ERR_ARG, "ffhip_rhs_p1: bad arguments");
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
    if (m->dim == 2) upload_t5_once();
    ctx->tic();
    if (m->dim == 3)
      k_rhs_p1_3d<<<nblocks(n, 256), 256, 0, s>>>(n, m->xyz, m->elems, m->v2e_ptr, m->v2e, f, onbc,
                                                   tgv * g, bd);
    else
      k_rhs_p1_2d<<<nblocks(n, 256), 256, 0, s>>>(n, m->xyz, m->elems, m->v2e_ptr, m->v2e, f, onbc,
                                                   tgv * g, bd);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    if (!on_device) {
      FFHIP_CHECK(hipMemcpy(b, bd, 8 * n, hipMemcpyDeviceToHost));
      dfree(bd);
    }
    dfree(onbc);
    dfree(labs);
  });
}

int ffhip_p1_misfit(ffhip_ctx ctx, ffhip_mesh m, const double *u, const double *ud,
                    const int32_t *labels, int nlab, double *J, double *r) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && u && ud && m->dim == 3, FFHIP_ERR_ARG, "ffhip_p1_misfit: bad arguments");
    hipStream_t s = ctx->stream;
    build_v2e(m);
    ctx->tic();
    if (J) {
      const int GB = 1024;
      double *part = dalloc<double>(GB);
      k_p1_misfit_elem<<<GB, 256, 0, s>>>(m->nt, m->xyz, m->elems, u, ud, part);
      FFHIP_CHECK(hipGetLastError());
      std::vector<double> h(GB);
      FFHIP_CHECK(hipMemcpyAsync(h.data(), part, 8 * GB, hipMemcpyDeviceToHost, s));
      FFHIP_CHECK(hipStreamSynchronize(s));
      double tot = 0.;
      for (int i = 0; i < GB; ++i) tot += h[i];
      *J = tot;
      dfree(part);
    }
    if (r) {
      uint8_t *onbc = nullptr;
      int32_t *labs = nullptr;
      if (nlab > 0) {
        onbc = dalloc<uint8_t>(m->nv);
        labs = dalloc<int32_t>(nlab);
        FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * nlab, hipMemcpyHostToDevice, s));
        FFHIP_CHECK(hipMemsetAsync(onbc, 0, m->nv, s));
        if (m->nbe)
          k_mark_bc<<<nblocks(m->nbe, 256), 256, 0, s>>>(m->nbe, m->nbloc, m->belems, m->blab, labs,
                                                         nlab, onbc);
      }
      k_p1_mass_apply<<<nblocks(m->nv, 256), 256, 0, s>>>(m->nv, m->xyz, m->elems, m->v2e_ptr, m->v2e,
                                                          u, ud, onbc, r);
      FFHIP_CHECK(hipGetLastError());
      FFHIP_CHECK(hipStreamSynchronize(s));
      dfree(onbc);
      dfree(labs);
    }
    ctx->toc();
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
    dfree(A->nodeptr); dfree(A->slot8);
    delete A;
  });
}

int ffhip_spmv(ffhip_ctx ctx, ffhip_mat A, const double *x, double *y, int trans) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && A && x && y, FFHIP_ERR_ARG, "ffhip_spmv: null argument");
    hipStream_t s = ctx->stream;
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
