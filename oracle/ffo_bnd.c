/*
 * ffo_bnd.c — oracle restatement of FreeFEM's boundary integrals on P1:
 * int2d(Th3, labels)(alpha*u*v) in a varf and int2d(Th3, labels)(g*v) in a
 * linear form, and their int1d forms on 2D meshes (Robin / Neumann terms).
 * TEST INFRASTRUCTURE ONLY (see ffo.h): what the tests compare against, never the product.
 *
 * Follows the reference literally: each listed boundary element is mapped to
 * its element and local face (Mesh::BoundaryElement), the element's four
 * (three) basis functions are evaluated at the face points T.PBord(ie, pi)
 * and the whole element matrix is added (problem.cpp:1273-1327 with
 * Element_Op's face branch problem.cpp:6515-6561; 2D edges problem.cpp:
 * 6223-6240; Element_rhs on a face problem.cpp:8517-8590).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ffo.h"

/* QF_GaussLegendre3 (QuadratureFormular.cpp:320-343), in its point order. */
int ffo_qf_gl3(double *x, double *w) {
  const double g1 = (1 - sqrt(3. / 5.)) / 2;
  x[0] = 0.5; x[1] = g1; x[2] = 1 - g1;
  w[0] = 8. / 18.; w[1] = 5. / 18.; w[2] = 5. / 18.;
  return 3;
}

static const int nvfaceTet[4][3] = {{3, 2, 1}, {0, 2, 3}, {3, 1, 0}, {0, 1, 2}}; /* Mesh3dn.cpp:72 */
static const int VerticesOfTriangularEdge[3][2] = {{1, 2}, {2, 0}, {0, 1}};        /* fem.hpp:95 */
static const double KHat3[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const double KHat2[3][2] = {{0, 0}, {1, 0}, {0, 1}};

static int64_t find_slot(const int64_t *rowptr, const int32_t *colidx, int r, int32_t c) {
  for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
    if (colidx[k] == c) return k;
  return -1;
}

/* the element holding boundary element f and the local number ie of the
 * vertex not on it (the face / edge opposite vertex ie) */
static int64_t owner(int nloc, int nb, const int32_t *elems, const int64_t *vptr,
                     const int64_t *vlist, const int32_t *bv, int *ie) {
  for (int64_t k = vptr[bv[0]]; k < vptr[bv[0] + 1]; ++k) {
    const int64_t t = vlist[k];
    const int32_t *el = elems + nloc * t;
    int found = 0, mask = 0;
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nloc; ++i)
        if (el[i] == bv[j]) { found++; mask |= 1 << i; }
    if (found == nb) {
      for (int i = 0; i < nloc; ++i)
        if (!(mask & (1 << i))) *ie = i;
      return t;
    }
  }
  return -1;
}

/* index of the point of the boundary element's own rule (vertices in mesh
 * order) at barycentric weights lb[] of its vertices */
static int bq_index(int dim, const double *qs, const double *qt, int nq, const double *lb) {
  for (int q = 0; q < nq; ++q) {
    if (dim == 3 && fabs(qs[q] - lb[1]) < 1e-12 && fabs(qt[q] - lb[2]) < 1e-12) return q;
    if (dim == 2 && fabs(qs[q] - lb[1]) < 1e-12) return q;
  }
  return -1;
}

int ffo_p1_boundary(int dim, int nv, int64_t nel, const double *X, const int32_t *elems,
                    int64_t nbe, const int32_t *belems, const int32_t *blab, int nlab,
                    const int32_t *labs, int kind, double c0, const double *coef,
                    const int64_t *rowptr, const int32_t *colidx, double *vals, double *b) {
  const int nloc = dim + 1, nb = dim;
  /* vertex -> elements */
  int64_t *vptr = calloc((size_t)nv + 1, sizeof(int64_t));
  int64_t *vlist = malloc(sizeof(int64_t) * (size_t)(nel * nloc + 1));
  int64_t *cur = calloc((size_t)nv, sizeof(int64_t));
  for (int64_t i = 0; i < nel * nloc; ++i) vptr[elems[i] + 1]++;
  for (int i = 0; i < nv; ++i) vptr[i + 1] += vptr[i];
  for (int64_t i = 0; i < nel * nloc; ++i) vlist[vptr[elems[i]] + cur[elems[i]]++] = i / nloc;
  free(cur);
  /* face rule (T_5 on triangles, Gauss-Legendre 3 on edges) */
  double qp[14], qw[7], qs[7], qt[7];
  int nq;
  if (dim == 3) {
    nq = ffo_qf_tri5(qp, qw);
    for (int q = 0; q < nq; ++q) { qs[q] = qp[2 * q]; qt[q] = qp[2 * q + 1]; }
  } else {
    nq = ffo_qf_gl3(qs, qw);
    for (int q = 0; q < nq; ++q) qt[q] = 0.;
  }
  int rc = 0;
  for (int64_t f = 0; f < nbe && rc == 0; ++f) {
    int on = 0;
    for (int k = 0; k < nlab; ++k) on |= blab[f] == labs[k];
    if (!on) continue;
    const int32_t *bv = belems + nb * f;
    int ie = -1;
    const int64_t t = owner(nloc, nb, elems, vptr, vlist, bv, &ie);
    if (t < 0) { rc = -4; break; }
    const int32_t *el = elems + nloc * t;
    const double *v[4];
    for (int i = 0; i < nloc; ++i) v[i] = X + dim * (int64_t)el[i];
    double M[4][4];
    memset(M, 0, sizeof M);
    for (int q = 0; q < nq; ++q) {
      double Pt[3], P[3] = {0, 0, 0}, phi[4], lb[3], coefq;
      if (dim == 3) {
        const int *fv = nvfaceTet[ie];
        /* N(ie) = (v[f0] -> v[f2]) ^ (v[f0] -> v[f1])  (GenericMesh.hpp:272, 369) */
        double a[3], c[3], N[3];
        for (int k = 0; k < 3; ++k) { a[k] = v[fv[2]][k] - v[fv[0]][k]; c[k] = v[fv[1]][k] - v[fv[0]][k]; }
        N[0] = a[1] * c[2] - a[2] * c[1];
        N[1] = a[2] * c[0] - a[0] * c[2];
        N[2] = a[0] * c[1] - a[1] * c[0];
        const double mes = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
        coefq = 0.5 * mes * qw[q];
        const double s = qs[q], tt = qt[q];
        /* PBord (Mesh3dn.hpp:76-79) */
        for (int k = 0; k < 3; ++k)
          Pt[k] = KHat3[fv[0]][k] * (1 - s - tt) + KHat3[fv[1]][k] * s + KHat3[fv[2]][k] * tt;
        const double sum = Pt[0] + Pt[1] + Pt[2];
        for (int k = 0; k < 3; ++k) {
          double r = (1 - sum) * v[0][k];
          r += Pt[0] * v[1][k];
          r += Pt[1] * v[2][k];
          r += Pt[2] * v[3][k];
          P[k] = r;
        }
        phi[0] = 1. - sum; phi[1] = Pt[0]; phi[2] = Pt[1]; phi[3] = Pt[2];
        const double lf[3] = {1 - s - tt, s, tt};
        for (int j = 0; j < 3; ++j)
          for (int m = 0; m < 3; ++m)
            if (el[fv[m]] == bv[j]) lb[j] = lf[m];
      } else {
        const int a0 = VerticesOfTriangularEdge[ie][0], a1 = VerticesOfTriangularEdge[ie][1];
        const double E0 = v[a1][0] - v[a0][0], E1 = v[a1][1] - v[a0][1];
        const double le = sqrt(E0 * E0 + E1 * E1);
        coefq = le * qw[q];
        const double sa = qs[q], sb = 1 - sa;
        for (int k = 0; k < 2; ++k) Pt[k] = KHat2[a0][k] * sa + KHat2[a1][k] * sb;
        for (int k = 0; k < 2; ++k) {
          double r = (1 - (Pt[0] + Pt[1])) * v[0][k];
          r += Pt[0] * v[1][k];
          r += Pt[1] * v[2][k];
          P[k] = r;
        }
        phi[0] = 1 - Pt[0] - Pt[1]; phi[1] = Pt[0]; phi[2] = Pt[1];
        for (int j = 0; j < 2; ++j) lb[j] = el[a0] == bv[j] ? sa : sb;
      }
      double cq;
      if (kind == FFO_COEF_QP) {
        const int q2 = bq_index(dim, qs, qt, nq, lb);
        if (q2 < 0) { rc = -5; break; }
        cq = coef[f * nq + q2];
      } else {
        cq = ffo_coef_eval(kind, c0, kind == FFO_COEF_P0 ? coef[f] : 0., P, dim);
      }
      for (int i = 0; i < nloc; ++i) {
        if (b) b[el[i]] += coefq * cq * phi[i];
        for (int j = 0; j < nloc; ++j) M[i][j] += coefq * cq * phi[i] * phi[j];
      }
    }
    if (vals && rc == 0)
      for (int i = 0; i < nloc; ++i)
        for (int j = 0; j < nloc; ++j) {
          const int64_t sl = find_slot(rowptr, colidx, el[i], el[j]);
          if (sl < 0) { rc = -3; break; }
          vals[sl] += M[i][j];
        }
  }
  free(vptr);
  free(vlist);
  return rc;
}
