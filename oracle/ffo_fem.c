/*
 * ffo_fem.c — oracle restatement of FreeFEM's quadrature, P1 element matrices,
 * HashMatrix pattern/accumulation, tgv Dirichlet and right-hand side.
 * TEST INFRASTRUCTURE ONLY (see ffo.h): what the tests compare against, never the product.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ffo.h"

/* ---- quadrature --------------------------------------------------------- */
/* QF_TET_5 (QuadratureFormular.cpp:725-743, Grundmann-Moeller 14-point rule,
 * weights given per unit-volume reference tet, i.e. 6 x the raw weight). */
int ffo_qf_tet5(double *pts, double *w) {
  const double a = 0.7217942490673263207930282587889082,
               b = 0.0927352503108912264023239137370306,
               c = 0.067342242210098170607962798709629,
               d = 0.310885919263300609797345733763457,
               e = 0.454496295874350350508119473720660,
               f = 0.045503704125649649491880526279339;
  const double wa = 0.0122488405193936582572850342477212 * 6.,
               wc = 0.0187813209530026417998642753888810 * 6.,
               we = 7.09100346284691107301157135337624E-3 * 6.;
  const double P[14][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b},
                           {c, d, d}, {d, c, d}, {d, d, c}, {d, d, d},
                           {e, e, f}, {e, f, e}, {f, e, e}, {f, f, e},
                           {f, e, f}, {e, f, f}};
  for (int q = 0; q < 14; ++q) {
    pts[3 * q] = P[q][0]; pts[3 * q + 1] = P[q][1]; pts[3 * q + 2] = P[q][2];
    w[q] = q < 4 ? wa : (q < 8 ? wc : we);
  }
  return 14;
}

/* QuadratureFormular_T_5 (QuadratureFormular.cpp:172-186, Stroud p.314). */
int ffo_qf_tri5(double *pts, double *w) {
  const double sqrt15 = 3.87298334620741688517926539978;
  const double t = 1.E0 / 3.E0, A = 0.225E0;
  const double r = (6 - sqrt15) / 21, s = (9 + 2 * sqrt15) / 21,
               B = (155 - sqrt15) / 1200;
  const double u = (6 + sqrt15) / 21, v = (9 - 2 * sqrt15) / 21,
               C = (155 + sqrt15) / 1200;
  const double P[7][2] = {{t, t}, {r, r}, {r, s}, {s, r}, {u, u}, {u, v}, {v, u}};
  const double W[7] = {A, B, B, B, C, C, C};
  for (int q = 0; q < 7; ++q) {
    pts[2 * q] = P[q][0]; pts[2 * q + 1] = P[q][1]; w[q] = W[q];
  }
  return 7;
}

/* ---- geometry ----------------------------------------------------------- */
/* det(A,B,C) with partial pivoting on the x column (R3.hpp:90-104). */
static double det3(double A[3], double B[3], double C[3]) {
  double s = 1., T[3];
  if (fabs(A[0]) < fabs(B[0])) { memcpy(T, A, 24); memcpy(A, B, 24); memcpy(B, T, 24); s = -s; }
  if (fabs(A[0]) < fabs(C[0])) { memcpy(T, A, 24); memcpy(A, C, 24); memcpy(C, T, 24); s = -s; }
  if (fabs(A[0]) > 1e-50) {
    s *= A[0];
    A[1] /= A[0]; A[2] /= A[0];
    B[1] -= A[1] * B[0]; B[2] -= A[2] * B[0];
    C[1] -= A[1] * C[0]; C[2] -= A[2] * C[0];
    return s * (B[1] * C[2] - B[2] * C[1]);
  }
  return 0.;
}

double ffo_tet_mesure(const double *v0, const double *v1, const double *v2,
                      const double *v3) {
  double AB[3], AC[3], AD[3];
  for (int c = 0; c < 3; ++c) {
    AB[c] = v1[c] - v0[c]; AC[c] = v2[c] - v0[c]; AD[c] = v3[c] - v0[c];
  }
  return det3(AB, AC, AD) / 6.;
}

static void cross(const double *u, const double *P, double *r) {
  /* R3::operator^ (R3.hpp:68) */
  r[0] = u[1] * P[2] - u[2] * P[1];
  r[1] = P[0] * u[2] - u[0] * P[2];
  r[2] = u[0] * P[1] - u[1] * P[0];
}

void ffo_tet_gradlambda(const double *v, double mes, double *G) {
  double V1[3], V2[3], V3[3], c[3];
  for (int k = 0; k < 3; ++k) {
    V1[k] = v[3 + k] - v[k]; V2[k] = v[6 + k] - v[k]; V3[k] = v[9 + k] - v[k];
  }
  const double det1 = 1. / (6. * mes);
  cross(V2, V3, c); for (int k = 0; k < 3; ++k) G[3 + k] = c[k] * det1;
  cross(V3, V1, c); for (int k = 0; k < 3; ++k) G[6 + k] = c[k] * det1;
  cross(V1, V2, c); for (int k = 0; k < 3; ++k) G[9 + k] = c[k] * det1;
  for (int k = 0; k < 3; ++k) G[k] = -G[3 + k] - G[6 + k] - G[9 + k];
}

/* kappa at a physical point, evaluated as the script expression
 * 1+0.5*sin(2*pi*x)*sin(2*pi*y)*sin(2*pi*z) (left to right). */
double ffo_coef_eval(int kind, double c0, double ke, const double *P, int dim) {
  if (kind == FFO_COEF_CONST) return c0;
  if (kind == FFO_COEF_P0) return ke;
  const double tp = 2 * M_PI;
  double s = 0.5 * sin(tp * P[0]);
  s = s * sin(tp * P[1]);
  if (dim == 3) s = s * sin(tp * P[2]);
  return 1 + s;
}

/* Element_Op 3D (problem.cpp:6384-6437) for three terms dx*dx, dy*dy, dz*dz
 * with coefficient kappa; mapped point T(Pt) per GenericElement::operator()
 * (GenericMesh.hpp:372-377). */
int ffo_p1_element_3d(const double *xyz, const int32_t *tet, int kind,
                      double c0, double kappa_e, double *Ke) {
  return ffo_p1_form_element_3d(xyz, tet, kind, c0, kappa_e, -1, 0., 0., Ke);
}

/* The same Element_Op loop with a fourth term m*u*v (op_id x op_id, P1
 * values l_i = barycentrics of the point, P012_3d.cpp:123-175) after the
 * three derivative terms at each quadrature point; mkind < 0: no such term. */
int ffo_p1_form_element_3d(const double *xyz, const int32_t *tet, int kind, double c0,
                           double kappa_e, int mkind, double m0, double mass_e, double *Ke) {
  double v[12], qp[42], qw[14], G[12];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = xyz[3 * (int64_t)tet[i] + k];
  const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
  ffo_tet_gradlambda(v, mes, G);
  ffo_qf_tet5(qp, qw);
  for (int i = 0; i < 16; ++i) Ke[i] = 0.;
  for (int q = 0; q < 14; ++q) {
    const double coef = mes * qw[q];
    const double *Ph = qp + 3 * q;
    double P[3];
    const double l0 = 1. - (Ph[0] + Ph[1] + Ph[2]);
    for (int k = 0; k < 3; ++k) P[k] = l0 * v[k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[0] * v[3 + k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[1] * v[6 + k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[2] * v[9 + k];
    for (int t = 0; t < 3; ++t) {
      double ccc = ffo_coef_eval(kind, c0, kappa_e, P, 3);
      ccc *= coef;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          Ke[4 * i + j] += ccc * G[3 * i + t] * G[3 * j + t];
    }
    if (mkind >= 0) {
      const double l[4] = {l0, Ph[0], Ph[1], Ph[2]};
      double ccc = ffo_coef_eval(mkind, m0, mass_e, P, 3);
      ccc *= coef;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Ke[4 * i + j] += ccc * l[i] * l[j];
    }
  }
  return 0;
}

/* 2D TTriangle area/H (fem.hpp:290-325), P1 FB (FESpace.cpp:1100-1135),
 * Element_Op 2D with terms dx*dx, dy*dy and QF T_5. */
int ffo_p1_element_2d(const double *xy, const int32_t *tri, int kind,
                      double c0, double kappa_e, double *Ke) {
  return ffo_p1_form_element_2d(xy, tri, kind, c0, kappa_e, -1, 0., 0., Ke);
}

int ffo_p1_form_element_2d(const double *xy, const int32_t *tri, int kind, double c0,
                           double kappa_e, int mkind, double m0, double mass_e, double *Ke) {
  double v[6], qp[14], qw[7], G[6];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 2; ++k) v[2 * i + k] = xy[2 * (int64_t)tri[i] + k];
  const double BAx = v[2] - v[0], BAy = v[3] - v[1];
  const double CAx = v[4] - v[0], CAy = v[5] - v[1];
  const double area = (BAx * CAy - BAy * CAx) * 0.5;
  for (int i = 0; i < 3; ++i) {
    const int a = (i + 2) % 3, b = (i + 1) % 3;
    const double Ex = v[2 * a] - v[2 * b], Ey = v[2 * a + 1] - v[2 * b + 1];
    G[2 * i] = -Ey / (2 * area);
    G[2 * i + 1] = Ex / (2 * area);
  }
  ffo_qf_tri5(qp, qw);
  for (int i = 0; i < 9; ++i) Ke[i] = 0.;
  for (int q = 0; q < 7; ++q) {
    const double coef = area * qw[q];
    const double *Ph = qp + 2 * q;
    double P[2];
    const double l0 = 1. - Ph[0] - Ph[1];
    for (int k = 0; k < 2; ++k)
      P[k] = l0 * v[k] + Ph[0] * v[2 + k] + Ph[1] * v[4 + k];
    for (int t = 0; t < 2; ++t) {
      double ccc = ffo_coef_eval(kind, c0, kappa_e, P, 2);
      ccc *= coef;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          Ke[3 * i + j] += ccc * G[2 * i + t] * G[2 * j + t];
    }
    if (mkind >= 0) {
      const double l[3] = {l0, Ph[0], Ph[1]};
      double ccc = ffo_coef_eval(mkind, m0, mass_e, P, 2);
      ccc *= coef;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ke[3 * i + j] += ccc * l[i] * l[j];
    }
  }
  return 0;
}

/* ---- pattern ------------------------------------------------------------ */
typedef struct { int64_t *ptr; int64_t *list; } adj_t;

/* dof -> (element, local index) adjacency, elements in increasing order */
static int build_adj(int n, int64_t nel, int nloc, const int32_t *elems, adj_t *a) {
  a->ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  a->list = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nel * nloc + 1));
  if (!a->ptr || !a->list) return -1;
  for (int64_t e = 0; e < nel * nloc; ++e) a->ptr[elems[e] + 1]++;
  for (int i = 0; i < n; ++i) a->ptr[i + 1] += a->ptr[i];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  memcpy(fill, a->ptr, sizeof(int64_t) * (size_t)n);
  for (int64_t e = 0; e < nel; ++e)
    for (int l = 0; l < nloc; ++l) a->list[fill[elems[e * nloc + l]]++] = e * nloc + l;
  free(fill);
  return 0;
}

static int cmp_i32(const void *x, const void *y) {
  int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
  return (a > b) - (a < b);
}

/* sorted unique column set of row i into buf; returns count */
static int64_t row_cols(const adj_t *a, int i, int nloc, const int32_t *elems,
                        int32_t **buf, int64_t *cap) {
  int64_t m = 0;
  for (int64_t k = a->ptr[i]; k < a->ptr[i + 1]; ++k) {
    const int64_t e = a->list[k] / nloc;
    if (m + nloc > *cap) {
      *cap = 2 * (*cap) + nloc;
      *buf = (int32_t *)realloc(*buf, sizeof(int32_t) * (size_t)(*cap));
    }
    for (int l = 0; l < nloc; ++l) (*buf)[m++] = elems[e * nloc + l];
  }
  qsort(*buf, (size_t)m, sizeof(int32_t), cmp_i32);
  int64_t u = 0;
  for (int64_t k = 0; k < m; ++k)
    if (u == 0 || (*buf)[k] != (*buf)[u - 1]) (*buf)[u++] = (*buf)[k];
  return u;
}

int64_t ffo_pattern_rowptr(int n, int64_t nel, int nloc, const int32_t *elems,
                           int64_t *rowptr) {
  adj_t a;
  if (build_adj(n, nel, nloc, elems, &a)) return -1;
  int64_t cap = 256;
  int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * cap);
  rowptr[0] = 0;
  for (int i = 0; i < n; ++i) rowptr[i + 1] = rowptr[i] + row_cols(&a, i, nloc, elems, &buf, &cap);
  free(buf); free(a.ptr); free(a.list);
  return rowptr[n];
}

int ffo_pattern_colidx(int n, int64_t nel, int nloc, const int32_t *elems,
                       const int64_t *rowptr, int32_t *colidx) {
  adj_t a;
  if (build_adj(n, nel, nloc, elems, &a)) return -1;
  int64_t cap = 256;
  int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * cap);
  int err = 0;
  for (int i = 0; i < n; ++i) {
    int64_t m = row_cols(&a, i, nloc, elems, &buf, &cap);
    if (m != rowptr[i + 1] - rowptr[i]) { err = -2; break; }
    memcpy(colidx + rowptr[i], buf, sizeof(int32_t) * (size_t)m);
  }
  free(buf); free(a.ptr); free(a.list);
  return err;
}

static int cmp_i64(const void *x, const void *y) {
  int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return (a > b) - (a < b);
}

int64_t ffo_pattern_bruteforce(int n, int64_t nel, int nloc,
                               const int32_t *elems, int64_t *rowptr,
                               int32_t *colidx) {
  const int64_t m = nel * nloc * nloc;
  int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
  int64_t k = 0;
  for (int64_t e = 0; e < nel; ++e)
    for (int i = 0; i < nloc; ++i)
      for (int j = 0; j < nloc; ++j)
        keys[k++] = (int64_t)elems[e * nloc + i] * n + elems[e * nloc + j];
  qsort(keys, (size_t)m, sizeof(int64_t), cmp_i64);
  int64_t u = 0;
  for (int i = 0; i <= n; ++i) rowptr[i] = 0;
  for (k = 0; k < m; ++k)
    if (k == 0 || keys[k] != keys[k - 1]) {
      colidx[u++] = (int32_t)(keys[k] % n);
      rowptr[keys[k] / n + 1]++;
    }
  for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  free(keys);
  return u;
}

/* slot of column c in row r (binary search over the sorted row) */
static int64_t find_slot(const int64_t *rowptr, const int32_t *colidx, int r, int32_t c) {
  int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    if (colidx[mid] == c) return mid;
    if (colidx[mid] < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* ---- assembly ----------------------------------------------------------- */
static int assemble(int nv, int64_t nel, int nloc, int dim, const double *X,
                    const int32_t *el, int kind, double c0, const double *kappa_e,
                    int mkind, double m0, const double *mass_e,
                    const int64_t *rowptr, const int32_t *colidx, double *vals) {
  (void)nv;
  memset(vals, 0, sizeof(double) * (size_t)rowptr[nv]);
  double Ke[16];
  for (int64_t e = 0; e < nel; ++e) {
    const int32_t *ni = el + e * nloc;
    const double ke = kappa_e ? kappa_e[e] : 0.;
    const double me = mass_e ? mass_e[e] : 0.;
    if (dim == 3) ffo_p1_form_element_3d(X, ni, kind, c0, ke, mkind, m0, me, Ke);
    else ffo_p1_form_element_2d(X, ni, kind, c0, ke, mkind, m0, me, Ke);
    /* HashMatrix::operator+= full case: il outer, jl inner (HashMatrix.cpp:1311-1317) */
    for (int il = 0; il < nloc; ++il)
      for (int jl = 0; jl < nloc; ++jl) {
        int64_t s = find_slot(rowptr, colidx, ni[il], ni[jl]);
        if (s < 0) return -3;
        vals[s] += Ke[nloc * il + jl];
      }
  }
  return 0;
}

int ffo_assemble_p1_3d(int nv, int64_t nt, const double *xyz,
                       const int32_t *tets, int kind, double c0,
                       const double *kappa_e, const int64_t *rowptr,
                       const int32_t *colidx, double *vals) {
  return assemble(nv, nt, 4, 3, xyz, tets, kind, c0, kappa_e, -1, 0., NULL, rowptr, colidx, vals);
}

int ffo_assemble_p1_form_3d(int nv, int64_t nt, const double *xyz, const int32_t *tets, int kind,
                            double c0, const double *kappa_e, int mkind, double m0,
                            const double *mass_e, const int64_t *rowptr, const int32_t *colidx,
                            double *vals) {
  return assemble(nv, nt, 4, 3, xyz, tets, kind, c0, kappa_e, mkind, m0, mass_e, rowptr, colidx,
                  vals);
}

int ffo_assemble_p1_2d(int nv, int64_t nt, const double *xy,
                       const int32_t *tris, int kind, double c0,
                       const double *kappa_e, const int64_t *rowptr,
                       const int32_t *colidx, double *vals) {
  return assemble(nv, nt, 3, 2, xy, tris, kind, c0, kappa_e, -1, 0., NULL, rowptr, colidx, vals);
}

int ffo_assemble_p1_form_2d(int nv, int64_t nt, const double *xy, const int32_t *tris, int kind,
                            double c0, const double *kappa_e, int mkind, double m0,
                            const double *mass_e, const int64_t *rowptr, const int32_t *colidx,
                            double *vals) {
  return assemble(nv, nt, 3, 2, xy, tris, kind, c0, kappa_e, mkind, m0, mass_e, rowptr, colidx,
                  vals);
}

/* Element_rhs 3D (problem.cpp:7940-7978): B[Kv(i)] += coef*c*w_i per point. */
int ffo_rhs_p1_3d(int nv, int64_t nt, const double *xyz, const int32_t *tets,
                  double f, double *b) {
  double qp[42], qw[14];
  ffo_qf_tet5(qp, qw);
  memset(b, 0, sizeof(double) * (size_t)nv);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = tets + 4 * e;
    const double *v0 = xyz + 3 * (int64_t)t[0], *v1 = xyz + 3 * (int64_t)t[1],
                 *v2 = xyz + 3 * (int64_t)t[2], *v3 = xyz + 3 * (int64_t)t[3];
    const double mes = ffo_tet_mesure(v0, v1, v2, v3);
    for (int q = 0; q < 14; ++q) {
      const double coef = mes * qw[q];
      const double *Ph = qp + 3 * q;
      const double l[4] = {1. - (Ph[0] + Ph[1] + Ph[2]), Ph[0], Ph[1], Ph[2]};
      for (int i = 0; i < 4; ++i) b[t[i]] += coef * f * l[i];
    }
  }
  return 0;
}

int ffo_rhs_p1_2d(int nv, int64_t nt, const double *xy, const int32_t *tris,
                  double f, double *b) {
  double qp[14], qw[7];
  ffo_qf_tri5(qp, qw);
  memset(b, 0, sizeof(double) * (size_t)nv);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = tris + 3 * e;
    const double *A = xy + 2 * (int64_t)t[0], *B = xy + 2 * (int64_t)t[1],
                 *C = xy + 2 * (int64_t)t[2];
    const double area = ((B[0] - A[0]) * (C[1] - A[1]) - (B[1] - A[1]) * (C[0] - A[0])) * 0.5;
    for (int q = 0; q < 7; ++q) {
      const double coef = area * qw[q];
      const double *Ph = qp + 2 * q;
      const double l[3] = {1 - Ph[0] - Ph[1], Ph[0], Ph[1]};
      for (int i = 0; i < 3; ++i) b[t[i]] += coef * f * l[i];
    }
  }
  return 0;
}

/* Element_rhs with a coefficient (problem.cpp:7940-7978): per element and
 * point, B[K(i)] += coef * f(T(Pt)) * w_i, coef = |K| w_q, T(Pt) as
 * GenericElement::operator() (GenericMesh.hpp:372-377); f: CONST f0, SIN3,
 * P0 fcoef[e], QP fcoef[e * nq + q].  dim 3: Tet_5, dim 2: T_5.  b zeroed first. */
int ffo_rhs_p1_coef(int dim, int nv, int64_t nt, const double *X, const int32_t *elems, int kind,
                    double f0, const double *fcoef, double *b) {
  double qp[42], qw[14];
  const int nq = dim == 3 ? ffo_qf_tet5(qp, qw) : ffo_qf_tri5(qp, qw);
  const int nloc = dim + 1;
  memset(b, 0, sizeof(double) * (size_t)nv);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = elems + nloc * e;
    const double *v[4];
    for (int i = 0; i < nloc; ++i) v[i] = X + dim * (int64_t)t[i];
    double mes;
    if (dim == 3) {
      mes = ffo_tet_mesure(v[0], v[1], v[2], v[3]);
    } else {
      mes = ((v[1][0] - v[0][0]) * (v[2][1] - v[0][1]) - (v[1][1] - v[0][1]) * (v[2][0] - v[0][0])) * 0.5;
    }
    for (int q = 0; q < nq; ++q) {
      const double *Ph = qp + dim * q;
      /* l0: 1 - PHat.sum() in 3D (P012_3d.cpp), 1 - x - y in 2D (P012_2d.cpp) */
      const double l0 = dim == 3 ? 1. - (Ph[0] + Ph[1] + Ph[2]) : 1 - Ph[0] - Ph[1];
      double P[3] = {0, 0, 0}, l[4];
      for (int k = 0; k < dim; ++k) {
        double r = l0 * v[0][k];
        for (int i = 1; i < nloc; ++i) r += Ph[i - 1] * v[i][k];
        P[k] = r;
      }
      l[0] = l0;
      for (int i = 1; i < nloc; ++i) l[i] = Ph[i - 1];
      double f;
      if (kind == FFO_COEF_QP) f = fcoef[e * nq + q];
      else f = ffo_coef_eval(kind, f0, kind == FFO_COEF_P0 ? fcoef[e] : 0., P, dim);
      const double coef = mes * qw[q];
      for (int i = 0; i < nloc; ++i) b[t[i]] += coef * f * l[i];
    }
  }
  return 0;
}

/* A P1 finite element function (vertex values u) at the quadrature points
 * of every element, as FreeFEM evaluates one inside Element_Op / Element_rhs
 * (the MeshPoint is on the element itself, pf3r2R lgmesh3.cpp:2169-2256):
 * 3D TypeOfFE_P1Lagrange3d::operator() (P012_3d.cpp:98-106)
 *   l0 = 1 - x - y - z;  r = u0*l0 + u1*l1 + l2*u2 + l3*u3,
 * 2D TypeOfFE_P1Lagrange::operator() (FESpace.cpp:1078-1085)
 *   l0 = 1 - x - y;  r = u0*l0 + u1*l1 + l2*u2,
 * at the Tet_5 (14) / T_5 (7) points; out[e * nq + q]. */
int ffo_p1_at_qp(int dim, int64_t nt, const int32_t *elems, const double *u, double *out) {
  double qp[42], qw[14];
  const int nq = dim == 3 ? ffo_qf_tet5(qp, qw) : ffo_qf_tri5(qp, qw);
  const int nloc = dim + 1;
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = elems + nloc * e;
    for (int q = 0; q < nq; ++q) {
      const double *Ph = qp + dim * q;
      double r;
      if (dim == 3) {
        const double l0 = 1 - Ph[0] - Ph[1] - Ph[2], l1 = Ph[0], l2 = Ph[1], l3 = Ph[2];
        r = u[t[0]] * l0 + u[t[1]] * l1 + l2 * u[t[2]] + l3 * u[t[3]];
      } else {
        const double l0 = 1 - Ph[0] - Ph[1], l1 = Ph[0], l2 = Ph[1];
        r = u[t[0]] * l0 + u[t[1]] * l1 + l2 * u[t[2]];
      }
      out[e * nq + q] = r;
    }
  }
  return nq;
}

/* ---- Dirichlet ---------------------------------------------------------- */
int ffo_mark_bc(int n, int64_t nbf, int nvf, const int32_t *bfaces,
                const int32_t *blab, int nlab, const int32_t *labs,
                uint8_t *onbc) {
  memset(onbc, 0, (size_t)n);
  for (int64_t f = 0; f < nbf; ++f) {
    int on = 0;
    for (int k = 0; k < nlab; ++k) on |= (blab[f] == labs[k]);
    if (!on) continue;
    for (int v = 0; v < nvf; ++v) onbc[bfaces[f * nvf + v]] = 1;
  }
  return 0;
}

int ffo_set_bc(int n, const int64_t *rowptr, const int32_t *colidx,
               double *vals, const uint8_t *onbc, double tgv) {
  for (int i = 0; i < n; ++i)
    if (onbc[i]) {
      int64_t s = find_slot(rowptr, colidx, i, i);
      if (s < 0) return -3;
      vals[s] = tgv;
    }
  return 0;
}

int ffo_rhs_bc(int n, const uint8_t *onbc, double tgv, double g, double *b) {
  for (int i = 0; i < n; ++i)
    if (onbc[i]) b[i] = tgv * g;
  return 0;
}

/* ---- adjoint building blocks ------------------------------------------- */
int ffo_assemble_p1_mass_3d(int nv, int64_t nt, const double *xyz,
                            const int32_t *tets, const int64_t *rowptr,
                            const int32_t *colidx, double *vals) {
  double qp[42], qw[14];
  ffo_qf_tet5(qp, qw);
  memset(vals, 0, sizeof(double) * (size_t)rowptr[nv]);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = tets + 4 * e;
    double v[12], Ke[16];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] = xyz[3 * (int64_t)t[i] + k];
    const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
    for (int i = 0; i < 16; ++i) Ke[i] = 0.;
    for (int q = 0; q < 14; ++q) {
      const double coef = mes * qw[q];
      const double *Ph = qp + 3 * q;
      const double l[4] = {1. - (Ph[0] + Ph[1] + Ph[2]), Ph[0], Ph[1], Ph[2]};
      double ccc = 1.;
      ccc *= coef;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Ke[4 * i + j] += ccc * l[i] * l[j];
    }
    for (int il = 0; il < 4; ++il)
      for (int jl = 0; jl < 4; ++jl) {
        int64_t s = find_slot(rowptr, colidx, t[il], t[jl]);
        if (s < 0) return -3;
        vals[s] += Ke[4 * il + jl];
      }
  }
  return 0;
}

double ffo_p1_misfit_3d(int64_t nt, const double *xyz, const int32_t *tets,
                        const double *u, const double *ud) {
  double qp[42], qw[14], J = 0.;
  ffo_qf_tet5(qp, qw);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = tets + 4 * e;
    const double mes = ffo_tet_mesure(xyz + 3 * (int64_t)t[0], xyz + 3 * (int64_t)t[1],
                                      xyz + 3 * (int64_t)t[2], xyz + 3 * (int64_t)t[3]);
    for (int q = 0; q < 14; ++q) {
      const double *Ph = qp + 3 * q;
      const double l[4] = {1. - (Ph[0] + Ph[1] + Ph[2]), Ph[0], Ph[1], Ph[2]};
      double eq = 0.;
      for (int i = 0; i < 4; ++i) eq += (u[t[i]] - ud[t[i]]) * l[i];
      J += mes * qw[q] * (eq * eq);
    }
  }
  return 0.5 * J;
}

int ffo_p0_grad_3d(int64_t nt, const double *xyz, const int32_t *tets,
                   const double *u, const double *lam, double *g) {
  double qp[42], qw[14];
  ffo_qf_tet5(qp, qw);
  for (int64_t e = 0; e < nt; ++e) {
    const int32_t *t = tets + 4 * e;
    double v[12], G[12];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] = xyz[3 * (int64_t)t[i] + k];
    const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
    ffo_tet_gradlambda(v, mes, G);
    double gu[3] = {0, 0, 0}, gl[3] = {0, 0, 0};
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k) { gu[k] += u[t[i]] * G[3 * i + k]; gl[k] += lam[t[i]] * G[3 * i + k]; }
    const double c = -(gu[0] * gl[0] + gu[1] * gl[1] + gu[2] * gl[2]);
    double b = 0.;
    for (int q = 0; q < 14; ++q) b += mes * qw[q] * c * 1.;
    g[e] = b;
  }
  return 0;
}
