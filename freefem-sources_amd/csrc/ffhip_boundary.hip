// ffhip_boundary.hip — boundary integrals of P1 varfs and linear forms:
//   int2d(Th3, labels)(alpha*u*v)  /  int1d(Th2, labels)(alpha*u*v)   (matrix)
//   int2d(Th3, labels)(g*v)        /  int1d(Th2, labels)(g*v)         (right-hand side)
// i.e. the Robin / Neumann terms of a FreeFEM problem.
//
// FreeFEM visits the boundary elements whose label is listed, in boundary-
// element order, and adds the face (edge) element matrix of the adjacent
// element (AssembleBilinearForm 3D, src/fflib/problem.cpp:1273-1327; the face
// branch of Element_Op, problem.cpp:6515-6561: coef = |N(ie)|/2 * w_q at the
// T_5 points of the face; 2D edges: problem.cpp:6223-6240, coef = |E| w_q at
// the QF_GaussLegendre3 points) and of the linear form (Element_rhs,
// problem.cpp:8517-8590).  On a face only the basis functions of its own
// vertices are non-zero, so the contribution is the boundary element's own
// P1 mass matrix weighted by the coefficient.
//
// GPU mapping: one lane per vertex row walks the vertex's boundary-element
// incidences in boundary-element order and adds that element's row into the
// CSR slots of the row (found by binary search in the sorted columns): every
// slot is owned by one lane and receives its contributions in FreeFEM's order,
// after the volume terms already in the matrix — deterministic, no atomics.
// The boundary is small next to the volume (a surface of the mesh), so the
// element rows are recomputed per incidence rather than stored.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ffhip_fem.hpp"

namespace ffhip {

int read_flag(ffhip_ctx ctx, int *d);

namespace {

unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Quadrature of a boundary element: T_5 (QuadratureFormular.cpp:172-186) on
// triangles, QF_GaussLegendre3 (QuadratureFormular.cpp:320-343) on edges.
// Point q of element e is x_q = (1 - s_q - t_q) B0 + s_q B1 + t_q B2
// (edges: (1 - s_q) B0 + s_q B1) for its vertices B in mesh order.
struct BndRule {
  int nq;
  double s[7], t[7], w[7];
};

BndRule rule(int dim) {
  BndRule R{};
  if (dim == 3) {
    const double sqrt15 = 3.87298334620741688517926539978;
    const double a = 1.E0 / 3.E0, A = 0.225E0;
    const double r = (6 - sqrt15) / 21, s = (9 + 2 * sqrt15) / 21, B = (155 - sqrt15) / 1200;
    const double u = (6 + sqrt15) / 21, v = (9 - 2 * sqrt15) / 21, C = (155 + sqrt15) / 1200;
    const double P[7][2] = {{a, a}, {r, r}, {r, s}, {s, r}, {u, u}, {u, v}, {v, u}};
    const double W[7] = {A, B, B, B, C, C, C};
    R.nq = 7;
    for (int q = 0; q < 7; ++q) { R.s[q] = P[q][0]; R.t[q] = P[q][1]; R.w[q] = W[q]; }
  } else {
    const double g1 = (1 - sqrt(3. / 5.)) / 2;
    R.nq = 3;
    R.s[0] = 0.5; R.s[1] = g1; R.s[2] = 1 - g1;
    R.w[0] = 8. / 18.; R.w[1] = 5. / 18.; R.w[2] = 5. / 18.;
  }
  return R;
}

__device__ inline double coef_at(int kind, double c0, const double *__restrict__ coef, int64_t e,
                                 int q, int nq, const double *P, int dim) {
  if (kind == FFHIP_COEF_SIN3) {
    double s = 0.5 * sin2pi(P[0]);
    s = s * sin2pi(P[1]);
    if (dim == 3) s = s * sin2pi(P[2]);
    return 1 + s;
  }
  if (kind == FFHIP_COEF_QP) return coef[e * nq + q];
  if (kind == FFHIP_COEF_P0) return coef[e];
  return c0;
}

// One lane per vertex v.  MAT: add the rows of alpha*u*v into vals (slots of
// row v); otherwise add g*v into b[v] (rows with skip[v] set are left alone:
// they are the Dirichlet rows that on(...) overwrites afterwards).
template <int NB, bool MAT>
__global__ __launch_bounds__(256) void k_p1_boundary(
    int64_t n, const double *__restrict__ xyz, const int32_t *__restrict__ belems,
    const int32_t *__restrict__ blab, const int32_t *__restrict__ labs, int nlab,
    const int64_t *__restrict__ bptr, const int32_t *__restrict__ blist, BndRule R, int kind,
    double c0, const double *__restrict__ coef, const int64_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, double *__restrict__ vals, const uint8_t *__restrict__ skip,
    double *__restrict__ b, int *__restrict__ missing) {
  constexpr int D = NB;  // coordinates per vertex: 3 for faces of a 3D mesh, 2 for edges
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t k0 = bptr[v], k1 = bptr[v + 1];
  if (k0 == k1) return;
  if (!MAT && skip && skip[v]) return;
  double bv = MAT ? 0. : b[v];
  for (int64_t k = k0; k < k1; ++k) {
    const int32_t item = blist[k];
    const int64_t e = item / NB;
    const int l = item - (int)e * NB;
    bool on = false;
    const int32_t lab = blab[e];
    for (int i = 0; i < nlab; ++i) on |= lab == labs[i];
    if (!on) continue;
    double B[NB][D];
    int32_t vid[NB];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      vid[m] = belems[NB * e + m];
#pragma unroll
      for (int c = 0; c < D; ++c) B[m][c] = xyz[(int64_t)D * vid[m] + c];
    }
    double meas;
    if constexpr (NB == 3) {  // |N| / 2 with N the cross product of two edges
      const double a0 = B[1][0] - B[0][0], a1 = B[1][1] - B[0][1], a2 = B[1][2] - B[0][2];
      const double c0_ = B[2][0] - B[0][0], c1 = B[2][1] - B[0][1], c2 = B[2][2] - B[0][2];
      const double N0 = a1 * c2 - a2 * c1, N1 = a2 * c0_ - a0 * c2, N2 = a0 * c1 - a1 * c0_;
      meas = 0.5 * sqrt(N0 * N0 + N1 * N1 + N2 * N2);
    } else {
      const double E0 = B[1][0] - B[0][0], E1 = B[1][1] - B[0][1];
      meas = sqrt(E0 * E0 + E1 * E1);
    }
    double row[NB];
#pragma unroll
    for (int m = 0; m < NB; ++m) row[m] = 0.;
    double r = 0.;
    for (int q = 0; q < R.nq; ++q) {
      double lam[NB];
      if constexpr (NB == 3) {
        lam[0] = 1. - R.s[q] - R.t[q]; lam[1] = R.s[q]; lam[2] = R.t[q];
      } else {
        lam[0] = 1. - R.s[q]; lam[1] = R.s[q];
      }
      double ll = lam[0];  // lam[l] without a dynamically indexed register array
#pragma unroll
      for (int m = 1; m < NB; ++m) ll = l == m ? lam[m] : ll;
      double P[3] = {0., 0., 0.};
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double x = 0.;
#pragma unroll
        for (int m = 0; m < NB; ++m) x += lam[m] * B[m][c];
        P[c] = x;
      }
      const double W = meas * R.w[q] * coef_at(kind, c0, coef, e, q, R.nq, P, D);
      if (MAT) {
#pragma unroll
        for (int m = 0; m < NB; ++m) row[m] += W * ll * lam[m];
      } else {
        r += W * ll;
      }
    }
    if (MAT) {
      const int64_t r0 = rowptr[v], r1 = rowptr[v + 1];
#pragma unroll
      for (int m = 0; m < NB; ++m) {
        int64_t lo = r0, hi = r1;  // first column >= vid[m]
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (colidx[mid] < vid[m]) lo = mid + 1; else hi = mid;
        }
        if (lo < r1 && colidx[lo] == vid[m]) vals[lo] += row[m];
        else atomicOr(missing, 1);
      }
    } else {
      bv += r;
    }
  }
  if (!MAT) b[v] = bv;
}

__global__ void k_mark_labels(int64_t nbe, int nbloc, const int32_t *__restrict__ belems,
                              const int32_t *__restrict__ blab, const int32_t *__restrict__ labs,
                              int nlab, uint8_t *__restrict__ onbc) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nbe) return;
  bool on = false;
  for (int k = 0; k < nlab; ++k) on |= blab[f] == labs[k];
  if (!on) return;
  for (int q = 0; q < nbloc; ++q) onbc[belems[nbloc * f + q]] = 1;  // same value from every writer
}

// vertex -> (boundary element * nbloc + local) lists, boundary-element order
void build_b2v(ffhip_mesh m) {
  if (m->b2e_ptr) return;
  build_incidence(m->ctx, m->belems, m->nbe, m->nbloc, m->nv, &m->b2e_ptr, &m->b2e, nullptr,
                  nullptr, nullptr);
}

// device copy of a coefficient array (nullptr when none is needed)
const double *coef_device(ffhip_mesh m, int kind, const double *coef, int on_device,
                          double **owned) {
  *owned = nullptr;
  if (kind != FFHIP_COEF_P0 && kind != FFHIP_COEF_QP) return nullptr;
  FFHIP_REQUIRE(coef, FFHIP_ERR_ARG, "boundary term: coefficient array missing");
  if (on_device) return coef;
  const int64_t cnt = kind == FFHIP_COEF_P0 ? m->nbe : m->nbe * (m->dim == 3 ? 7 : 3);
  *owned = dalloc<double>(cnt);
  FFHIP_CHECK(hipMemcpyAsync(*owned, coef, 8 * cnt, hipMemcpyHostToDevice, m->ctx->stream));
  return *owned;
}

bool valid_kind(int k) {
  return k == FFHIP_COEF_CONST || k == FFHIP_COEF_SIN3 || k == FFHIP_COEF_P0 || k == FFHIP_COEF_QP;
}

}  // namespace
}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_assemble_p1_boundary(ffhip_ctx ctx, ffhip_mesh m, ffhip_mat A, const int32_t *labels,
                               int nlab, int kind, double c0, const double *coef, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && A && A->n == m->nv && A->ncomp == 0 && A->vals, FFHIP_ERR_ARG,
                  "ffhip_assemble_p1_boundary: needs the P1 matrix of this mesh");
    FFHIP_REQUIRE(nlab > 0 && labels, FFHIP_ERR_ARG, "ffhip_assemble_p1_boundary: labels");
    FFHIP_REQUIRE(valid_kind(kind), FFHIP_ERR_ARG, "ffhip_assemble_p1_boundary: coefficient kind");
    if (m->nbe == 0) return;
    hipStream_t s = ctx->stream;
    build_b2v(m);
    double *owned;
    const double *cd = coef_device(m, kind, coef, on_device, &owned);
    int32_t *labs = dalloc<int32_t>(nlab);
    int *missing = dalloc<int>(1);
    FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * (size_t)nlab, hipMemcpyHostToDevice, s));
    FFHIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), s));
    const BndRule R = rule(m->dim);
    ctx->tic();
    if (m->dim == 3)
      k_p1_boundary<3, true><<<nblocks(m->nv, 256), 256, 0, s>>>(
          m->nv, m->xyz, m->belems, m->blab, labs, nlab, m->b2e_ptr, m->b2e, R, kind, c0, cd,
          A->rowptr, A->colidx, A->vals, nullptr, nullptr, missing);
    else
      k_p1_boundary<2, true><<<nblocks(m->nv, 256), 256, 0, s>>>(
          m->nv, m->xyz, m->belems, m->blab, labs, nlab, m->b2e_ptr, m->b2e, R, kind, c0, cd,
          A->rowptr, A->colidx, A->vals, nullptr, nullptr, missing);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    const int miss = read_flag(ctx, missing);
    dfree(labs); dfree(missing); dfree(owned);
    FFHIP_REQUIRE(miss == 0, FFHIP_ERR_ARG,
                  "ffhip_assemble_p1_boundary: a boundary element is not a face of the mesh");
  });
}

int ffhip_rhs_p1_boundary(ffhip_ctx ctx, ffhip_mesh m, const int32_t *labels, int nlab, int kind,
                          double g0, const double *coef, const int32_t *dlabels, int ndlab,
                          double *b, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && m && b, FFHIP_ERR_ARG, "ffhip_rhs_p1_boundary: bad arguments");
    FFHIP_REQUIRE(nlab > 0 && labels, FFHIP_ERR_ARG, "ffhip_rhs_p1_boundary: labels");
    FFHIP_REQUIRE(ndlab >= 0 && (ndlab == 0 || dlabels), FFHIP_ERR_ARG,
                  "ffhip_rhs_p1_boundary: Dirichlet labels");
    FFHIP_REQUIRE(valid_kind(kind), FFHIP_ERR_ARG, "ffhip_rhs_p1_boundary: coefficient kind");
    if (m->nbe == 0) return;
    hipStream_t s = ctx->stream;
    const int64_t n = m->nv;
    build_b2v(m);
    double *owned;
    const double *cd = coef_device(m, kind, coef, on_device, &owned);
    int32_t *labs = dalloc<int32_t>(nlab + ndlab);
    FFHIP_CHECK(hipMemcpyAsync(labs, labels, 4 * (size_t)nlab, hipMemcpyHostToDevice, s));
    if (ndlab)
      FFHIP_CHECK(hipMemcpyAsync(labs + nlab, dlabels, 4 * (size_t)ndlab, hipMemcpyHostToDevice, s));
    uint8_t *skip = nullptr;
    if (ndlab) {
      skip = dalloc<uint8_t>(n);
      FFHIP_CHECK(hipMemsetAsync(skip, 0, n, s));
      k_mark_labels<<<nblocks(m->nbe, 256), 256, 0, s>>>(m->nbe, m->nbloc, m->belems, m->blab,
                                                         labs + nlab, ndlab, skip);
    }
    double *bd = b;
    if (!on_device) {
      bd = dalloc<double>(n);
      FFHIP_CHECK(hipMemcpyAsync(bd, b, 8 * n, hipMemcpyHostToDevice, s));
    }
    const BndRule R = rule(m->dim);
    ctx->tic();
    if (m->dim == 3)
      k_p1_boundary<3, false><<<nblocks(n, 256), 256, 0, s>>>(
          n, m->xyz, m->belems, m->blab, labs, nlab, m->b2e_ptr, m->b2e, R, kind, g0, cd, nullptr,
          nullptr, nullptr, skip, bd, nullptr);
    else
      k_p1_boundary<2, false><<<nblocks(n, 256), 256, 0, s>>>(
          n, m->xyz, m->belems, m->blab, labs, nlab, m->b2e_ptr, m->b2e, R, kind, g0, cd, nullptr,
          nullptr, nullptr, skip, bd, nullptr);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    if (!on_device) {
      FFHIP_CHECK(hipMemcpy(b, bd, 8 * n, hipMemcpyDeviceToHost));
      dfree(bd);
    }
    dfree(labs); dfree(skip); dfree(owned);
  });
}

}  // extern "C"
