/*
 * ffo_cg.c — oracle restatement of FreeFEM's sparse mat-vec and Jacobi-
 * preconditioned conjugate gradient (SolverCG + ConjugueGradient).
 * TEST INFRASTRUCTURE ONLY (see ffo.h): what the tests compare against, never the product.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ffo.h"

/* y = A x in CSR row order (the HashMatrix non-half product
 * HashMatrix.cpp:1138-1139, summed per row in column order). */
void ffo_spmv(int n, const int64_t *rowptr, const int32_t *colidx,
              const double *vals, const double *x, double *y) {
  for (int i = 0; i < n; ++i) {
    double s = 0.;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += vals[k] * x[colidx[k]];
    y[i] = s;
  }
}

/* y = A x for a half matrix (lower triangle stored, HashMatrix::half): the
 * half branch of addMatMul (HashMatrix.cpp:1128-1136) over the entries in
 * row-major order: y_i += a x_j, and y_j += a x_i off the diagonal. */
void ffo_spmv_half(int n, const int64_t *rowptr, const int32_t *colidx,
                   const double *vals, const double *x, double *y) {
  memset(y, 0, sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      y[i] += vals[k] * x[j];
      if (i != j) y[j] += vals[k] * x[i];
    }
}

/* y = A^T x (addMatMul with Transpose: i/j swapped, HashMatrix.cpp:1094). */
void ffo_spmv_t(int n, const int64_t *rowptr, const int32_t *colidx,
                const double *vals, const double *x, double *y) {
  memset(y, 0, sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) y[colidx[k]] += vals[k] * x[i];
}

static double diag_of(const int64_t *rowptr, const int32_t *colidx,
                      const double *vals, int i, int *found) {
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (colidx[k] == i) { *found = 1; return vals[k]; }
  *found = 0;
  return 0.;
}

double ffo_gettgv(int n, const int64_t *rowptr, const int32_t *colidx,
                  const double *vals, int *pntgv) {
  double ttgv = 0, max1 = 0;
  int ntgv = 0;
  for (int ii = 0; ii < n; ++ii) {
    int f;
    double a = diag_of(rowptr, colidx, vals, ii, &f);
    if (!f) continue;
    if (a > ttgv) { max1 = ttgv; ttgv = a; ntgv = 1; }
    else if (a == ttgv) ++ntgv;
    else if (a > max1) max1 = a;
  }
  if (max1 * 1e6 > ttgv) { ttgv = 0; ntgv = 0; }
  if (pntgv) *pntgv = ntgv;
  return ttgv;
}

static double dot(int n, const double *x, const double *y) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

int ffo_pcg(int n, const int64_t *rowptr, const int32_t *colidx,
            const double *vals, const double *b, double *x, double eps,
            int itmax, double tgv, int trans, double *stats) {
  /* HMatVirtPrecon (VirtualSolverCG.hpp:27-69): tgv rows, Jacobi diag1 */
  int ntgv1 = 0;
  const double tgvm = ffo_gettgv(n, rowptr, colidx, vals, &ntgv1);
  uint8_t *wcl = NULL;
  double *d1 = (double *)malloc(sizeof(double) * (size_t)n);
  if (ntgv1) {
    double tgve = tgvm;
    if (tgve <= 0) tgve = 1e200;
    wcl = (uint8_t *)malloc((size_t)n);
    for (int i = 0; i < n; ++i) {
      int f;
      wcl[i] = diag_of(rowptr, colidx, vals, i, &f) == tgve;
    }
  }
  for (int i = 0; i < n; ++i) {
    int f;
    d1[i] = diag_of(rowptr, colidx, vals, i, &f);
  }
  for (int i = 0; i < n; ++i) d1[i] = (d1[i] * d1[i] < 1e-60) ? 1 : 1. / d1[i];
  /* SetInitWithBC (VirtualSolverCG.hpp:93-99) */
  if (wcl)
    for (int i = 0; i < n; ++i)
      if (wcl[i]) x[i] = b[i] / tgv;
  if (itmax <= 0) itmax = n; /* SolverCG(ds): itmax>0 ? itmax : n */

  /* ConjugueGradient (CG.cpp:195-265) */
  double *G = (double *)malloc(sizeof(double) * (size_t)n);
  double *CG = (double *)malloc(sizeof(double) * (size_t)n);
  double *H = (double *)malloc(sizeof(double) * (size_t)n);
  double *AH = CG; /* same array as CG in the reference */
  void (*mv)(int, const int64_t *, const int32_t *, const double *, const double *, double *) =
      trans ? ffo_spmv_t : ffo_spmv;
  mv(n, rowptr, colidx, vals, x, G);
  for (int i = 0; i < n; ++i) G[i] += -1. * b[i];          /* G = Ax - b */
  for (int i = 0; i < n; ++i) CG[i] = d1[i] * G[i];         /* C.matmul */
  double gCg = dot(n, G, CG), gCgp;
  const double gCg0 = gCg;
  for (int i = 0; i < n; ++i) H[i] = CG[i];
  for (int i = 0; i < n; ++i) H[i] *= -1.;                  /* H = -CG */
  double eps2 = eps * eps;
  if (eps > 0) { eps2 *= gCg; }
  int nret = 0, iters = itmax;
  if (gCg != gCg) nret = -1;
  else if (gCg < 1e-30) { nret = 2; iters = 0; }
  else
    for (int iter = 1; iter <= itmax; ++iter) {
      gCgp = gCg;
      const double gh = dot(n, G, H);
      mv(n, rowptr, colidx, vals, H, AH);
      const double rho = -gh / dot(n, H, AH);
      for (int i = 0; i < n; ++i) x[i] += rho * H[i];
      for (int i = 0; i < n; ++i) G[i] += rho * AH[i];
      for (int i = 0; i < n; ++i) CG[i] = d1[i] * G[i];
      gCg = dot(n, G, CG);
      const double gamma = gCg / gCgp;
      for (int i = 0; i < n; ++i) H[i] *= gamma;
      for (int i = 0; i < n; ++i) H[i] += -1. * CG[i];
      if (gCg < eps2) { nret = 1; iters = iter; break; }
    }
  if (stats) {
    stats[0] = iters; stats[1] = gCg0; stats[2] = gCg; stats[3] = nret;
  }
  free(G); free(CG); free(H); free(d1); free(wcl);
  return nret;
}

/* LinearCG / AffineCG with a matrix operator: ConjuguedGradient2
 * (MatriceCreuse.hpp:180-240).  The operator A(x) is the matrix product, so
 * Ah = A x - b - g of the reference is rop * A h up to rounding; the
 * restatement keeps the reference's affine form (recomputed from x). */
int ffo_cg2(int n, const int64_t *rowptr, const int32_t *colidx,
            const double *vals, const double *b, double *x, double *eps,
            int itmax, int precon, double *stats) {
  double *d1 = NULL;
  if (precon) { /* the Jacobi func of a precon= argument */
    d1 = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
      int f;
      const double a = diag_of(rowptr, colidx, vals, i, &f);
      d1[i] = (a * a < 1e-60) ? 1. : 1. / a;
    }
  }
  double *g = (double *)malloc(sizeof(double) * (size_t)n);
  double *h = (double *)malloc(sizeof(double) * (size_t)n);
  double *Ah = (double *)malloc(sizeof(double) * (size_t)n);
  double *Cg = Ah; /* "Rn ... Ah(n), & Cg(Ah)": the same array */
  double ro = 1.;
  ffo_spmv(n, rowptr, colidx, vals, x, g); /* g = A*x; g -= b */
  if (b) for (int i = 0; i < n; ++i) g[i] -= b[i];
  for (int i = 0; i < n; ++i) Cg[i] = d1 ? 0. + d1[i] * g[i] : 0. + g[i]; /* Cg = C*g */
  for (int i = 0; i < n; ++i) h[i] = -Cg[i];
  double g2 = dot(n, Cg, g);
  const double g20 = g2;
  int nret = 0, iters = 0;
  if (g2 < 1e-30) {
    nret = 2;
  } else {
    const double reps2 = *eps > 0 ? *eps * *eps * g2 : -*eps;
    *eps = reps2;
    iters = itmax + 1;
    for (int iter = 0; iter <= itmax; iter++) {
      const double rop = ro;
      for (int i = 0; i < n; ++i) x[i] += rop * h[i];
      ffo_spmv(n, rowptr, colidx, vals, x, Ah); /* Ah = A*x; Ah -= b; Ah -= g */
      if (b) for (int i = 0; i < n; ++i) Ah[i] -= b[i];
      for (int i = 0; i < n; ++i) Ah[i] -= g[i];
      const double hAh = dot(n, h, Ah);
      if (hAh * hAh < 1e-60) { nret = -1; iters = iter; break; }
      ro = -dot(n, g, h) * rop / hAh;
      if (ro != ro) { nret = -2; iters = iter; break; }
      for (int i = 0; i < n; ++i) x[i] += (ro - rop) * h[i];
      for (int i = 0; i < n; ++i) g[i] += (ro / rop) * Ah[i];
      for (int i = 0; i < n; ++i) Cg[i] = d1 ? 0. + d1[i] * g[i] : 0. + g[i];
      const double g2p = g2;
      g2 = dot(n, Cg, g);
      if (g2 < reps2) { nret = 1; iters = iter + 1; break; }
      const double gamma = g2 / g2p;
      for (int i = 0; i < n; ++i) h[i] *= gamma;
      for (int i = 0; i < n; ++i) h[i] -= Cg[i];
    }
  }
  if (stats) { stats[0] = iters; stats[1] = g20; stats[2] = g2; stats[3] = nret; }
  free(g); free(h); free(Ah); free(d1);
  return nret;
}

/* ---- restarted flexible GMRES: SolverGMRES::dosolver (VirtualSolverCG.hpp:
 * 236-255) with HMatVirtPrecon (13-111: tgv rows, Jacobi diag1, SetInitWithBC)
 * over fgmres (CG.cpp:347-517, right preconditioning: leftC is forced to 0). */
static double nrm2(int n, const double *x) {
  double s = 0.;
  for (int i = 0; i < n; ++i) s += x[i] * x[i];
  return sqrt(s);
}

int ffo_fgmres(int n, const int64_t *rowptr, const int32_t *colidx, const double *vals,
               const double *b, double *x, double eps, int itmax, int restart, double tgv,
               double *stats) {
  const int m = restart;
  int ntgv1 = 0;
  const double tgvm = ffo_gettgv(n, rowptr, colidx, vals, &ntgv1);
  int *wbc = NULL;
  if (ntgv1) {
    const double tgve = tgvm <= 0 ? 1e200 : tgvm;
    wbc = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) {
      int f;
      wbc[i] = diag_of(rowptr, colidx, vals, i, &f) == tgve;
    }
  }
  double *d1 = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    int f;
    const double a = diag_of(rowptr, colidx, vals, i, &f);
    d1[i] = (a * a < 1e-60) ? 1. : 1. / a;
  }
  if (wbc) /* SetInitWithBC */
    for (int i = 0; i < n; ++i)
      if (wbc[i]) x[i] = b[i] / tgv;
  int nbitermx = itmax > 0 ? itmax : n;
  /* fgmres */
  const size_t N = (size_t)n;
  double *rot0 = (double *)malloc(sizeof(double) * (m + 2)), *rot1 = (double *)malloc(sizeof(double) * (m + 2));
  double *g = (double *)malloc(sizeof(double) * (m + 1)), *g1 = (double *)malloc(sizeof(double) * (m + 1));
  double *Hn = (double *)calloc((size_t)(m + 2) * (m + 1), sizeof(double));
#define HN(i, j) Hn[(size_t)(i) * (m + 1) + (j)]
  double *uni = (double *)malloc(sizeof(double) * N), *x0 = (double *)malloc(sizeof(double) * N);
  double *ri = (double *)malloc(sizeof(double) * N), *zi = (double *)malloc(sizeof(double) * N);
  double *vi = (double *)malloc(sizeof(double) * N), *wi = (double *)malloc(sizeof(double) * N);
  double *V = (double *)malloc(sizeof(double) * N * (m + 1));
  double *Vp = (double *)malloc(sizeof(double) * N * (m + 1));
  double *y = (double *)malloc(sizeof(double) * (m + 1));
  double relres = 1e100, normb;
  for (int i = 0; i < n; ++i) wi[i] = b[i];
  if (wbc)
    for (int i = 0; i < n; ++i)
      if (wbc[i]) wi[i] = 0.;
  for (int i = 0; i < n; ++i) vi[i] = 0. + wi[i]; /* pCl = identity */
  normb = nrm2(n, vi);
  for (int i = 0; i < n; ++i) uni[i] = x[i];
  int noconv = 1, iter = 0, nrestart = 0;
  while (noconv) {
    for (int i = 0; i < n; ++i) x0[i] = uni[i];
    ffo_spmv(n, rowptr, colidx, vals, uni, ri); /* ri = -(A uni - rhs) */
    for (int i = 0; i < n; ++i) ri[i] += -1. * b[i];
    for (int i = 0; i < n; ++i) ri[i] *= -1.;
    for (int i = 0; i < n; ++i) zi[i] = 0. + ri[i];
    g[0] = nrm2(n, zi);
    if (normb < 1.e-20 || eps < 0) normb = 1.;
    {
      const double inv = 1. / g[0];
      for (int i = 0; i < n; ++i) V[i] = inv * zi[i];
    }
    int it;
    for (it = 0; it < m; it++, iter++) {
      double *Vit = V + N * it;
      for (int i = 0; i < n; ++i) vi[i] = 0. + d1[i] * Vit[i]; /* C.matmul */
      for (int i = 0; i < n; ++i) Vp[N * it + i] = vi[i];
      ffo_spmv(n, rowptr, colidx, vals, vi, wi);
      for (int i = 0; i < it + 1; i++) { /* modified Gram-Schmidt */
        const double *Vi = V + N * i;
        HN(i, it) = dot(n, wi, Vi);
        const double a = -HN(i, it);
        for (int k = 0; k < n; ++k) wi[k] += a * Vi[k];
      }
      const double aux = HN(it + 1, it) = nrm2(n, wi);
      {
        const double inv = 1. / aux;
        for (int i = 0; i < n; ++i) V[N * (it + 1) + i] = inv * wi[i];
      }
      for (int i = 0; i < it; i++) {
        const double aa = rot0[i] * HN(i, it) + rot1[i] * HN(i + 1, it);
        const double bb = -rot1[i] * HN(i, it) + rot0[i] * HN(i + 1, it);
        HN(i, it) = aa;
        HN(i + 1, it) = bb;
      }
      const double sq = sqrt(HN(it, it) * HN(it, it) + HN(it + 1, it) * HN(it + 1, it));
      rot0[it] = HN(it, it) / sq;
      rot1[it] = HN(it + 1, it) / sq;
      HN(it, it) = rot0[it] * HN(it, it) + rot1[it] * HN(it + 1, it);
      HN(it + 1, it) = 0.;
      g[it + 1] = -rot1[it] * g[it];
      g[it] = rot0[it] * g[it];
      relres = fabs(g[it + 1]);
      if (relres / normb < fabs(eps)) { noconv = 0; break; }
      if (it > nbitermx) break;
    }
    if (it > m - 1) it = m - 1;
    for (int i = it; i >= 0; i--) { /* reconstruction */
      g1[i] = g[i];
      for (int j = i + 1; j < it + 1; j++) g1[i] = g1[i] - HN(i, j) * y[j];
      y[i] = g1[i] / HN(i, i);
    }
    for (int k = 0; k < n; ++k) wi[k] = 0.;
    for (int i = 0; i < it + 1; i++)
      for (int k = 0; k < n; ++k) wi[k] += y[i] * Vp[N * i + k];
    for (int k = 0; k < n; ++k) uni[k] = wi[k];
    for (int k = 0; k < n; ++k) uni[k] += x0[k];
    for (int k = 0; k < n; ++k) x[k] = uni[k];
    if (!noconv) break;
    if (iter > nbitermx) break;
    nrestart++;
  }
#undef HN
  if (stats) { stats[0] = iter; stats[1] = relres / normb; stats[2] = normb; stats[3] = nrestart; }
  free(rot0); free(rot1); free(g); free(g1); free(Hn); free(uni); free(x0); free(ri); free(zi);
  free(vi); free(wi); free(V); free(Vp); free(y); free(d1); free(wbc);
  return !noconv;
}
