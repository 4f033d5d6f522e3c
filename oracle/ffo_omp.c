/*
 * ffo_omp.c — the oracle's 3D P1 path (pattern, element-order values,
 * Dirichlet rows, Jacobi-CG) with its loops split over all host cores with
 * OpenMP: the all-core CPU baseline of the benchmark.
 *
 * TEST INFRASTRUCTURE ONLY (see ffo.h): loaded by tests/ and bench.py's
 * cpu_baseline leg, never by the product.
 *
 * Same results as the serial restatement: the pattern is the sorted
 * HashMatrix pattern (HashMatrix.cpp:1294-1332 inserts, Sortij/Buildp
 * HashMatrix.cpp:670-682, 992-1030); each matrix entry receives its element
 * contributions in element order (HashMatrix::operator+=, the order of
 * ffo_assemble_p1_3d), here by rows: the element matrices
 * (ffo_p1_form_element_3d, problem.cpp:6384-6437) are formed in parallel,
 * then every row adds its incident elements' rows in element order, so the
 * values are bit-identical to ffo_assemble_p1_3d.  The CG is ConjugueGradient
 * (CG.cpp:195-265) with the Jacobi preconditioner of SolverCG
 * (VirtualSolverCG.hpp:27-99), its dot products summed per thread and then
 * over threads, so its iterates differ from ffo_pcg's in the last bits.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "ffo.h"

struct ffo_omp_sys {
  int n;
  int64_t nnz;
  int64_t *rowptr;
  int32_t *colidx;
  double *vals;
};

int ffo_omp_max_threads(void) { return omp_get_max_threads(); }

/* vertex -> incidences e*4+l, in element order */
static int build_v2e(int n, int64_t nt, const int32_t *tets, int64_t **pptr, int64_t **plist) {
  int64_t *ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (size_t)(4 * nt + 1));
  int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!ptr || !list || !cur) { free(ptr); free(list); free(cur); return -1; }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < 4 * nt; ++k) {
#pragma omp atomic
    ptr[tets[k] + 1]++;
  }
  for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  memcpy(cur, ptr, sizeof(int64_t) * (size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < 4 * nt; ++k) {
    int64_t at;
#pragma omp atomic capture
    at = cur[tets[k]]++;
    list[at] = k;
  }
  /* the parallel fill leaves each list unordered: restore element order */
#pragma omp parallel for schedule(dynamic, 4096)
  for (int i = 0; i < n; ++i) {
    int64_t *a = list + ptr[i];
    const int64_t m = ptr[i + 1] - ptr[i];
    for (int64_t p = 1; p < m; ++p) {
      const int64_t v = a[p];
      int64_t q = p - 1;
      while (q >= 0 && a[q] > v) { a[q + 1] = a[q]; --q; }
      a[q + 1] = v;
    }
  }
  free(cur);
  *pptr = ptr;
  *plist = list;
  return 0;
}

/* sorted unique columns of row i (at most 4 m entries for m incidences) */
static int row_cols(const int64_t *ptr, const int64_t *list, const int32_t *tets, int i,
                    int32_t *buf) {
  int m = 0;
  for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int64_t e = list[k] / 4;
    for (int l = 0; l < 4; ++l) buf[m++] = tets[4 * e + l];
  }
  for (int p = 1; p < m; ++p) {
    const int32_t v = buf[p];
    int q = p - 1;
    while (q >= 0 && buf[q] > v) { buf[q + 1] = buf[q]; --q; }
    buf[q + 1] = v;
  }
  int u = 0;
  for (int k = 0; k < m; ++k)
    if (u == 0 || buf[k] != buf[u - 1]) buf[u++] = buf[k];
  return u;
}

struct ffo_omp_sys *ffo_omp_assemble_p1_3d(int nv, int64_t nt, const double *xyz,
                                           const int32_t *tets, int kind, double c0,
                                           const double *kappa_e) {
  int64_t *ptr = NULL, *list = NULL;
  if (build_v2e(nv, nt, tets, &ptr, &list)) return NULL;
  int maxinc = 0;
  for (int i = 0; i < nv; ++i)
    if (ptr[i + 1] - ptr[i] > maxinc) maxinc = (int)(ptr[i + 1] - ptr[i]);
  struct ffo_omp_sys *S = (struct ffo_omp_sys *)calloc(1, sizeof(*S));
  if (S) S->rowptr = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nv + 1));
  if (!S || !S->rowptr) { free(S); free(ptr); free(list); return NULL; }
  S->n = nv;
  int err = 0;
  /* pattern: row lengths, prefix sum, columns */
#pragma omp parallel
  {
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (4 * (size_t)maxinc + 4));
#pragma omp for schedule(dynamic, 4096)
    for (int i = 0; i < nv; ++i) S->rowptr[i + 1] = row_cols(ptr, list, tets, i, buf);
    free(buf);
  }
  S->rowptr[0] = 0;
  for (int i = 0; i < nv; ++i) S->rowptr[i + 1] += S->rowptr[i];
  S->nnz = S->rowptr[nv];
  S->colidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)S->nnz + 4);
  S->vals = (double *)malloc(sizeof(double) * (size_t)S->nnz + 8);
  double *Ke = (double *)malloc(sizeof(double) * 16 * (size_t)nt + 8);
  if (!S->colidx || !S->vals || !Ke) err = -1;
  if (!err) {
#pragma omp parallel
    {
      int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (4 * (size_t)maxinc + 4));
#pragma omp for schedule(dynamic, 4096)
      for (int i = 0; i < nv; ++i) {
        const int m = row_cols(ptr, list, tets, i, buf);
        memcpy(S->colidx + S->rowptr[i], buf, sizeof(int32_t) * (size_t)m);
      }
      free(buf);
    }
    /* element matrices, then rows in element order */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nt; ++e)
      ffo_p1_form_element_3d(xyz, tets + 4 * e, kind, c0, kappa_e ? kappa_e[e] : 0., -1, 0., 0.,
                             Ke + 16 * e);
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : err)
    for (int i = 0; i < nv; ++i) {
      const int64_t r0 = S->rowptr[i], r1 = S->rowptr[i + 1];
      double *v = S->vals + r0;
      const int32_t *c = S->colidx + r0;
      for (int64_t s = 0; s < r1 - r0; ++s) v[s] = 0.;
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int64_t e = list[k] / 4;
        const int l = (int)(list[k] % 4);
        for (int jl = 0; jl < 4; ++jl) {
          const int32_t col = tets[4 * e + jl];
          int64_t lo = 0, hi = r1 - r0 - 1, at = -1;
          while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (c[mid] == col) { at = mid; break; }
            if (c[mid] < col) lo = mid + 1; else hi = mid - 1;
          }
          if (at < 0) { err |= 1; continue; }
          v[at] += Ke[16 * e + 4 * l + jl];
        }
      }
    }
  }
  free(Ke); free(ptr); free(list);
  if (err) { ffo_omp_free(S); return NULL; }
  return S;
}

int64_t ffo_omp_nnz(const struct ffo_omp_sys *S) { return S->nnz; }

void ffo_omp_arrays(const struct ffo_omp_sys *S, int64_t *rowptr, int32_t *colidx, double *vals) {
  memcpy(rowptr, S->rowptr, sizeof(int64_t) * ((size_t)S->n + 1));
  memcpy(colidx, S->colidx, sizeof(int32_t) * (size_t)S->nnz);
  memcpy(vals, S->vals, sizeof(double) * (size_t)S->nnz);
}

/* SetBC (HashMatrix.cpp:1194-1238): the diagonal of each marked row = tgv */
void ffo_omp_setbc(struct ffo_omp_sys *S, const uint8_t *onbc, double tgv) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S->n; ++i) {
    if (!onbc[i]) continue;
    for (int64_t k = S->rowptr[i]; k < S->rowptr[i + 1]; ++k)
      if (S->colidx[k] == i) S->vals[k] = tgv;
  }
}

static void spmv(const struct ffo_omp_sys *S, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S->n; ++i) {
    double s = 0.;
    for (int64_t k = S->rowptr[i]; k < S->rowptr[i + 1]; ++k) s += S->vals[k] * x[S->colidx[k]];
    y[i] = s;
  }
}

static double dot(int n, const double *x, const double *y) {
  double s = 0.;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

int ffo_omp_pcg(const struct ffo_omp_sys *S, const double *b, double *x, double eps, int itmax,
                double tgv, double *stats) {
  const int n = S->n;
  double *d1 = (double *)malloc(sizeof(double) * (size_t)n);
  double ttgv = 0., max1 = 0.;
  /* gettgv (VirtualSolverCG.hpp:27-69) over the diagonal */
  for (int i = 0; i < n; ++i) {
    double a = 0.;
    for (int64_t k = S->rowptr[i]; k < S->rowptr[i + 1]; ++k)
      if (S->colidx[k] == i) a = S->vals[k];
    d1[i] = a;
    if (a > ttgv) { max1 = ttgv; ttgv = a; }
    else if (a < ttgv && a > max1) max1 = a;
  }
  if (max1 * 1e6 > ttgv) ttgv = 0.;
  const double tgve = ttgv <= 0 ? 1e200 : ttgv;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (ttgv > 0 && d1[i] == tgve) x[i] = b[i] / tgv; /* SetInitWithBC */
    d1[i] = (d1[i] * d1[i] < 1e-60) ? 1 : 1. / d1[i];
  }
  if (itmax <= 0) itmax = n;
  double *G = (double *)malloc(sizeof(double) * (size_t)n);
  double *CG = (double *)malloc(sizeof(double) * (size_t)n);
  double *H = (double *)malloc(sizeof(double) * (size_t)n);
  double *AH = CG;
  spmv(S, x, G);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    G[i] += -1. * b[i];
    CG[i] = d1[i] * G[i];
    H[i] = -CG[i];
  }
  double gCg = dot(n, G, CG), gCgp;
  const double gCg0 = gCg;
  double eps2 = eps * eps;
  if (eps > 0) eps2 *= gCg;
  int nret = 0, iters = itmax;
  if (gCg != gCg) nret = -1;
  else if (gCg < 1e-30) { nret = 2; iters = 0; }
  else
    for (int iter = 1; iter <= itmax; ++iter) {
      gCgp = gCg;
      const double gh = dot(n, G, H);
      spmv(S, H, AH);
      const double rho = -gh / dot(n, H, AH);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        x[i] += rho * H[i];
        G[i] += rho * AH[i];
        CG[i] = d1[i] * G[i];
      }
      gCg = dot(n, G, CG);
      const double gamma = gCg / gCgp;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) H[i] = H[i] * gamma + -1. * CG[i];
      if (gCg < eps2) { nret = 1; iters = iter; break; }
    }
  if (stats) { stats[0] = iters; stats[1] = gCg0; stats[2] = gCg; stats[3] = nret; }
  free(G); free(CG); free(H); free(d1);
  return nret;
}

void ffo_omp_free(struct ffo_omp_sys *S) {
  if (!S) return;
  free(S->rowptr); free(S->colidx); free(S->vals); free(S);
}
