// ffhip_fem.hpp — finite-element device pieces shared by the P1 and P2 paths:
// the Tet_5 quadrature table, tetrahedron geometry as FreeFEM computes it, and
// the row-wise sparsity-pattern kernel.
#pragma once

#include "ffhip_internal.hpp"

namespace ffhip {

// Tet_5 (src/femlib/QuadratureFormular.cpp:725-743)
static __constant__ double c_qf5_pts[14][3] = {
    {0.7217942490673263207930282587889082, 0.0927352503108912264023239137370306, 0.0927352503108912264023239137370306},
    {0.0927352503108912264023239137370306, 0.7217942490673263207930282587889082, 0.0927352503108912264023239137370306},
    {0.0927352503108912264023239137370306, 0.0927352503108912264023239137370306, 0.7217942490673263207930282587889082},
    {0.0927352503108912264023239137370306, 0.0927352503108912264023239137370306, 0.0927352503108912264023239137370306},
    {0.067342242210098170607962798709629, 0.310885919263300609797345733763457, 0.310885919263300609797345733763457},
    {0.310885919263300609797345733763457, 0.067342242210098170607962798709629, 0.310885919263300609797345733763457},
    {0.310885919263300609797345733763457, 0.310885919263300609797345733763457, 0.067342242210098170607962798709629},
    {0.310885919263300609797345733763457, 0.310885919263300609797345733763457, 0.310885919263300609797345733763457},
    {0.454496295874350350508119473720660, 0.454496295874350350508119473720660, 0.045503704125649649491880526279339},
    {0.454496295874350350508119473720660, 0.045503704125649649491880526279339, 0.454496295874350350508119473720660},
    {0.045503704125649649491880526279339, 0.454496295874350350508119473720660, 0.454496295874350350508119473720660},
    {0.045503704125649649491880526279339, 0.045503704125649649491880526279339, 0.454496295874350350508119473720660},
    {0.045503704125649649491880526279339, 0.454496295874350350508119473720660, 0.045503704125649649491880526279339},
    {0.454496295874350350508119473720660, 0.045503704125649649491880526279339, 0.045503704125649649491880526279339}};
static __constant__ double c_qf5_w[3] = {0.0122488405193936582572850342477212 * 6.,
                                  0.0187813209530026417998642753888810 * 6.,
                                  7.09100346284691107301157135337624E-3 * 6.};

__device__ inline double qf5_weight(int q) { return c_qf5_w[q < 4 ? 0 : (q < 8 ? 1 : 2)]; }

// det with x-column pivoting, as R3.hpp:90-104
__device__ inline double det3_pivot(double A[3], double B[3], double C[3]) {
  double s = 1.;
  if (fabs(A[0]) < fabs(B[0])) {
    for (int k = 0; k < 3; ++k) { double t = A[k]; A[k] = B[k]; B[k] = t; }
    s = -s;
  }
  if (fabs(A[0]) < fabs(C[0])) {
    for (int k = 0; k < 3; ++k) { double t = A[k]; A[k] = C[k]; C[k] = t; }
    s = -s;
  }
  if (fabs(A[0]) > 1e-50) {
    s *= A[0];
    A[1] /= A[0]; A[2] /= A[0];
    B[1] -= A[1] * B[0]; B[2] -= A[2] * B[0];
    C[1] -= A[1] * C[0]; C[2] -= A[2] * C[0];
    return s * (B[1] * C[2] - B[2] * C[1]);
  }
  return 0.;
}

// |K| (Mesh3dn.hpp:65-71) and grad lambda_i (Mesh3dn.hpp:126-137)
__device__ inline double tet_geom(const double v[4][3], double G[4][3]) {
  double AB[3], AC[3], AD[3];
  for (int k = 0; k < 3; ++k) {
    AB[k] = v[1][k] - v[0][k]; AC[k] = v[2][k] - v[0][k]; AD[k] = v[3][k] - v[0][k];
  }
  double V1[3] = {AB[0], AB[1], AB[2]}, V2[3] = {AC[0], AC[1], AC[2]}, V3[3] = {AD[0], AD[1], AD[2]};
  const double mes = det3_pivot(AB, AC, AD) / 6.;
  const double det1 = 1. / (6. * mes);
  G[1][0] = (V2[1] * V3[2] - V2[2] * V3[1]) * det1;
  G[1][1] = (V3[0] * V2[2] - V2[0] * V3[2]) * det1;
  G[1][2] = (V2[0] * V3[1] - V2[1] * V3[0]) * det1;
  G[2][0] = (V3[1] * V1[2] - V3[2] * V1[1]) * det1;
  G[2][1] = (V1[0] * V3[2] - V3[0] * V1[2]) * det1;
  G[2][2] = (V3[0] * V1[1] - V3[1] * V1[0]) * det1;
  G[3][0] = (V1[1] * V2[2] - V1[2] * V2[1]) * det1;
  G[3][1] = (V2[0] * V1[2] - V1[0] * V2[2]) * det1;
  G[3][2] = (V1[0] * V2[1] - V1[1] * V2[0]) * det1;
  for (int k = 0; k < 3; ++k) G[0][k] = -G[1][k] - G[2][k] - G[3][k];
  return mes;
}

// the 4 cross products N_l of a tet's edge vectors from vertex 0, grad l_l
// times 6 |K| (Tet::Gradlambda, Mesh3dn.hpp:126-137, before its division),
// with N_0 = -(N_1 + N_2 + N_3)
__device__ inline void tet_normals(const double v[4][3], double N[4][3]) {
  double V1[3], V2[3], V3[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    V1[k] = v[1][k] - v[0][k]; V2[k] = v[2][k] - v[0][k]; V3[k] = v[3][k] - v[0][k];
  }
  N[1][0] = V2[1] * V3[2] - V2[2] * V3[1];
  N[1][1] = V3[0] * V2[2] - V2[0] * V3[2];
  N[1][2] = V2[0] * V3[1] - V2[1] * V3[0];
  N[2][0] = V3[1] * V1[2] - V3[2] * V1[1];
  N[2][1] = V1[0] * V3[2] - V3[0] * V1[2];
  N[2][2] = V3[0] * V1[1] - V3[1] * V1[0];
  N[3][0] = V1[1] * V2[2] - V1[2] * V2[1];
  N[3][1] = V2[0] * V1[2] - V1[0] * V2[2];
  N[3][2] = V1[0] * V2[1] - V1[1] * V2[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) N[0][k] = -N[1][k] - N[2][k] - N[3][k];
}

// the 4 entries K_e(l, m) = C N_l . N_m of row l of a P1 stiffness element
// matrix whose scale is C = S_e / (6 |K|)^2 (K_e = S_e grad l . grad l)
__device__ inline void p1_row_entries(const double v[4][3], int l, double C, double r[4]) {
  double N[4][3];
  tet_normals(v, N);
  double Nl[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    Nl[k] = l == 0 ? N[0][k] : (l == 1 ? N[1][k] : (l == 2 ? N[2][k] : N[3][k]));
#pragma unroll
  for (int m = 0; m < 4; ++m) r[m] = C * (Nl[0] * N[m][0] + Nl[1] * N[m][1] + Nl[2] * N[m][2]);
}

__device__ inline double tet_mesure(const double v[4][3]) {
  double AB[3], AC[3], AD[3];
  for (int k = 0; k < 3; ++k) {
    AB[k] = v[1][k] - v[0][k]; AC[k] = v[2][k] - v[0][k]; AD[k] = v[3][k] - v[0][k];
  }
  return det3_pivot(AB, AC, AD) / 6.;
}

// sin(2 pi x): t = 2x = k + r with k = rint(t) and |r| <= 1/2 (exact), then
// sin(pi r) from its Taylor series truncated after r^21 (coefficients
// (-1)^j pi^(2j+1) / (2j+1)!, Horner in r^2; truncation below 3e-16 on
// |r| <= 1/2) and the sign (-1)^k.  About 16 FP64 operations, a third of a
// general sinpi with its quadrant branches: the 42 evaluations per tet are
// most of the element kernel's work.
__device__ inline double sin2pi(double x) {
  const double t = 2. * x;
  const double k = rint(t);
  const double r = t - k;
  const double s = r * r;
  double p = 5.39266466260812848935e-10;
  p = fma(p, s, -2.29484289972698731102e-8);
  p = fma(p, s, 7.95205400147551278478e-7);
  p = fma(p, s, -2.19153534478302158274e-5);
  p = fma(p, s, 4.66302805767612564421e-4);
  p = fma(p, s, -7.37043094571435077726e-3);
  p = fma(p, s, 8.21458866111282287988e-2);
  p = fma(p, s, -5.99264529320792076888e-1);
  p = fma(p, s, 2.55016403987734544386);
  p = fma(p, s, -5.16771278004997002925);
  p = fma(p, s, 3.14159265358979323846);
  const double y = r * p;
  return ((int64_t)k & 1) ? -y : y;
}

// kappa = 1+0.5*sin(2*pi*x)*sin(2*pi*y)*sin(2*pi*z), the script's product
// order (no range reduction of a rounded 2 pi x: within about an ulp of the
// script's sin(2*pi*x))
__device__ inline double kappa_sin3(double x, double y, double z) {
  double s = 0.5 * sin2pi(x);
  s = s * sin2pi(y);
  s = s * sin2pi(z);
  return 1 + s;
}

// unique index of (l,m), l <= m, in a packed symmetric 4x4 / 3x3
__device__ inline int sym4(int l, int m) {
  if (l > m) { int t = l; l = m; m = t; }
  return l * 4 - (l * (l - 1)) / 2 + (m - l);
}
__device__ inline int sym3(int l, int m) {
  if (l > m) { int t = l; l = m; m = t; }
  return l * 3 - (l * (l - 1)) / 2 + (m - l);
}

// A P1 finite element function at a reference point of an element, from its
// vertex values, in FreeFEM's order of operations: 3D
// TypeOfFE_P1Lagrange3d::operator() (P012_3d.cpp:98-106), 2D
// TypeOfFE_P1Lagrange::operator() (FESpace.cpp:1078-1085).
__device__ inline double p1_value3(const double u[4], double x, double y, double z) {
  const double l0 = 1 - x - y - z;
  return u[0] * l0 + u[1] * x + y * u[2] + z * u[3];
}
__device__ inline double p1_value2(const double u[3], double x, double y) {
  const double l0 = 1 - x - y;
  return u[0] * l0 + u[1] * x + y * u[2];
}

// ---------------------------------------------------------------- pattern
// One lane per row; the row's sorted unique column list lives in LDS, column
// c of lane t at lds[c * BLOCK + t] (consecutive lanes in consecutive banks).
// With only != null the kernel takes the rows whose only[] byte equals want;
// a row of more than CAP columns sets *overflow and, with only != null, its
// byte to want + 1 (the next, wider pass takes it).
template <int BLOCK, int CAP, bool FILL>
__global__ __launch_bounds__(BLOCK) void k_pattern_rows(
    int64_t n, int nloc, const int32_t *__restrict__ elems,
    const int64_t *__restrict__ v2e_ptr, const int32_t *__restrict__ v2e,
    int64_t *__restrict__ rowlen, const int64_t *__restrict__ rowptr,
    int32_t *__restrict__ colidx, int64_t *__restrict__ diag,
    uint32_t *__restrict__ slotmap, uint8_t *__restrict__ slot8, int *__restrict__ overflow,
    uint8_t *__restrict__ only = nullptr, uint8_t want = 1) {
  __shared__ int32_t L[CAP * BLOCK];
  const int t = threadIdx.x;
  const int64_t v = (int64_t)blockIdx.x * BLOCK + t;
  if (v >= n) return;
  int u = 0;
  bool ovf = false;
  const int64_t k0 = v2e_ptr[v], k1 = v2e_ptr[v + 1];
  if (only && only[v] != want) return;  // rows of another pass
  for (int64_t k = k0; k < k1 && !ovf; ++k) {
    const int64_t e = v2e[k] / nloc;
    for (int mloc = 0; mloc < nloc; ++mloc) {
      const int32_t c = elems[e * nloc + mloc];
      int pos = 0;
      while (pos < u && L[pos * BLOCK + t] < c) ++pos;
      if (pos < u && L[pos * BLOCK + t] == c) continue;
      if (u == CAP) { ovf = true; break; }
      for (int q = u; q > pos; --q) L[q * BLOCK + t] = L[(q - 1) * BLOCK + t];
      L[pos * BLOCK + t] = c;
      ++u;
    }
  }
  if (ovf) {
    if (only) only[v] = (uint8_t)(want + 1);
    atomicMax(overflow, 1);
    return;
  }
  if (!FILL) { rowlen[v] = u; return; }
  const int64_t r0 = rowptr[v];
  int dpos = -1;
  for (int q = 0; q < u; ++q) {
    const int32_t c = L[q * BLOCK + t];
    colidx[r0 + q] = c;
    if (c == v) dpos = q;
  }
  diag[v] = dpos < 0 ? -1 : r0 + dpos;
  // element-to-row map: offsets of the element's dofs within this row, packed
  // 4 x uint8 per incidence (nloc <= 4) or nloc bytes per incidence (slot8)
  for (int64_t k = k0; k < k1; ++k) {
    const int64_t e = v2e[k] / nloc;
    uint32_t packed = 0;
    for (int mloc = 0; mloc < nloc; ++mloc) {
      const int32_t c = elems[e * nloc + mloc];
      int lo = 0, hi = u - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L[mid * BLOCK + t] < c) lo = mid + 1; else hi = mid;
      }
      if (slot8) slot8[k * nloc + mloc] = (uint8_t)lo;
      else packed |= (uint32_t)lo << (8 * mloc);
    }
    if (!slot8) slotmap[k] = packed;
  }
}

}  // namespace ffhip
