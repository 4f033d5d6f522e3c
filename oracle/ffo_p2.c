/*
 * ffo_p2.c — oracle restatement of FreeFEM's P2 node numbering, P2 basis and
 * the vector-P2 Lame (linear elasticity) element matrix, right-hand side and
 * clamped boundary.
 * TEST INFRASTRUCTURE ONLY (see ffo.h): what the tests compare against, never the product.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ffo.h"

static const int kEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

/* open-addressing hash of 64-bit keys -> node id */
typedef struct { uint64_t *k; int32_t *v; uint64_t cap; } hmap_t;

static uint64_t mix(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}

static int32_t *hfind(hmap_t *h, uint64_t key, int *fresh) {
  uint64_t i = mix(key) & (h->cap - 1);
  for (;;) {
    if (h->k[i] == UINT64_MAX) { h->k[i] = key; *fresh = 1; return &h->v[i]; }
    if (h->k[i] == key) { *fresh = 0; return &h->v[i]; }
    i = (i + 1) & (h->cap - 1);
  }
}

static uint64_t vkey(int32_t v) { return ((uint64_t)(uint32_t)v << 32) | 0xffffffffULL; }
static uint64_t ekey(int32_t a, int32_t b) {
  if (a > b) { int32_t t = a; a = b; b = t; }
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

int64_t ffo_p2_nodes(int nv, int64_t nt, const int32_t *tets, int32_t *elem_nodes) {
  (void)nv;
  hmap_t h;
  h.cap = 1;
  while (h.cap < (uint64_t)(nt * 10 * 2 + 16)) h.cap <<= 1;
  h.k = (uint64_t *)malloc(8 * h.cap);
  h.v = (int32_t *)malloc(4 * h.cap);
  memset(h.k, 0xff, 8 * h.cap);
  int32_t nn = 0;
  for (int64_t k = 0; k < nt; ++k) {
    const int32_t *t = tets + 4 * k;
    uint64_t keys[10];
    for (int i = 0; i < 4; ++i) keys[i] = vkey(t[i]);
    for (int e = 0; e < 6; ++e) keys[4 + e] = ekey(t[kEdge[e][0]], t[kEdge[e][1]]);
    for (int i = 0; i < 10; ++i) {
      int fresh;
      int32_t *p = hfind(&h, keys[i], &fresh);
      if (fresh) *p = nn++;
      elem_nodes[10 * k + i] = *p;
    }
  }
  free(h.k);
  free(h.v);
  return nn;
}

void ffo_p2_basis(const double *phat, const double *G, double *f, double *df) {
  const double l[4] = {1. - (phat[0] + phat[1] + phat[2]), phat[0], phat[1], phat[2]};
  const double l4[4] = {(4 * l[0] - 1), (4 * l[1] - 1), (4 * l[2] - 1), (4 * l[3] - 1)};
  for (int i = 0; i < 4; ++i) {
    f[i] = l[i] * (2 * l[i] - 1.);
    for (int k = 0; k < 3; ++k) df[3 * i + k] = G[3 * i + k] * l4[i];
  }
  for (int e = 0; e < 6; ++e) {
    const int i0 = kEdge[e][0], i1 = kEdge[e][1];
    f[4 + e] = 4. * l[i0] * l[i1];
    for (int k = 0; k < 3; ++k)
      df[3 * (4 + e) + k] = 4 * (G[3 * i1 + k] * l[i0] + G[3 * i0 + k] * l[i1]);
  }
}

/* K[(c,i),(c',j)] += coef (lambda d_c phi_i d_c' phi_j
 *                          + mu (d_c' phi_i d_c phi_j + delta_cc' grad phi_i . grad phi_j))
 * per Tet_5 point: lambda div u div v + 2 mu eps(u):eps(v) for u = phi_j e_c',
 * v = phi_i e_c. */
int ffo_p2_elast_element(const double *xyz, const int32_t *tet, double lambda,
                         double mu, double *Ke) {
  double v[12], G[12], qp[42], qw[14];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = xyz[3 * (int64_t)tet[i] + k];
  const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
  ffo_tet_gradlambda(v, mes, G);
  ffo_qf_tet5(qp, qw);
  memset(Ke, 0, sizeof(double) * 900);
  for (int q = 0; q < 14; ++q) {
    double f[10], d[30];
    ffo_p2_basis(qp + 3 * q, G, f, d);
    const double coef = mes * qw[q];
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < 10; ++i)
        for (int cp = 0; cp < 3; ++cp)
          for (int j = 0; j < 10; ++j) {
            double a = lambda * d[3 * i + c] * d[3 * j + cp] + mu * d[3 * i + cp] * d[3 * j + c];
            if (c == cp)
              a += mu * (d[3 * i] * d[3 * j] + d[3 * i + 1] * d[3 * j + 1] + d[3 * i + 2] * d[3 * j + 2]);
            Ke[30 * (10 * c + i) + 10 * cp + j] += coef * a;
          }
  }
  return 0;
}

int ffo_p2v_elem_dofs(int64_t nt, const int32_t *elem_nodes, int32_t *elem_dofs) {
  for (int64_t k = 0; k < nt; ++k)
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < 10; ++i) elem_dofs[30 * k + 10 * c + i] = 3 * elem_nodes[10 * k + i] + c;
  return 0;
}

static int64_t slot_of(const int64_t *rowptr, const int32_t *colidx, int32_t r, int32_t c) {
  int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (colidx[mid] == c) return mid;
    if (colidx[mid] < c) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

int ffo_assemble_p2_elast(int64_t nt, const double *xyz, const int32_t *tets,
                          const int32_t *elem_dofs, double lambda, double mu,
                          int64_t ndof, const int64_t *rowptr,
                          const int32_t *colidx, double *vals) {
  memset(vals, 0, sizeof(double) * (size_t)rowptr[ndof]);
  double *Ke = (double *)malloc(sizeof(double) * 900);
  for (int64_t k = 0; k < nt; ++k) {
    ffo_p2_elast_element(xyz, tets + 4 * k, lambda, mu, Ke);
    const int32_t *d = elem_dofs + 30 * k;
    for (int il = 0; il < 30; ++il)
      for (int jl = 0; jl < 30; ++jl) {
        const int64_t s = slot_of(rowptr, colidx, d[il], d[jl]);
        if (s < 0) { free(Ke); return -3; }
        vals[s] += Ke[30 * il + jl];
      }
  }
  free(Ke);
  return 0;
}

int ffo_rhs_p2_elast(int64_t nt, const double *xyz, const int32_t *tets,
                     const int32_t *elem_dofs, const double *fv, int64_t ndof,
                     double *b) {
  double qp[42], qw[14];
  ffo_qf_tet5(qp, qw);
  memset(b, 0, sizeof(double) * (size_t)ndof);
  for (int64_t k = 0; k < nt; ++k) {
    const int32_t *t = tets + 4 * k;
    double v[12], G[12];
    for (int i = 0; i < 4; ++i)
      for (int kk = 0; kk < 3; ++kk) v[3 * i + kk] = xyz[3 * (int64_t)t[i] + kk];
    const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
    ffo_tet_gradlambda(v, mes, G);
    for (int q = 0; q < 14; ++q) {
      double f[10], d[30];
      ffo_p2_basis(qp + 3 * q, G, f, d);
      const double coef = mes * qw[q];
      for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 10; ++i) b[elem_dofs[30 * k + 10 * c + i]] += coef * fv[c] * f[i];
    }
  }
  return 0;
}

int ffo_p2v_mark_bc(int nv, int64_t nt, const int32_t *tets,
                    const int32_t *elem_nodes, int64_t nbe, const int32_t *btri,
                    const int32_t *blab, int nlab, const int32_t *labs,
                    int64_t ndof, uint8_t *onbc) {
  return ffo_p2_mark_bc(nv, nt, tets, elem_nodes, nbe, btri, blab, nlab, labs, 3, ndof, onbc);
}

/* the nodes of the labelled boundary triangles (3 vertices, 3 edge midpoints),
 * every one of the ncomp components (dof ncomp * node + c) */
int ffo_p2_mark_bc(int nv, int64_t nt, const int32_t *tets, const int32_t *elem_nodes,
                   int64_t nbe, const int32_t *btri, const int32_t *blab, int nlab,
                   const int32_t *labs, int ncomp, int64_t ndof, uint8_t *onbc) {
  (void)nv;
  /* edge key -> node, vertex -> node from the element tables */
  hmap_t h;
  h.cap = 1;
  while (h.cap < (uint64_t)(nt * 10 * 2 + 16)) h.cap <<= 1;
  h.k = (uint64_t *)malloc(8 * h.cap);
  h.v = (int32_t *)malloc(4 * h.cap);
  memset(h.k, 0xff, 8 * h.cap);
  for (int64_t k = 0; k < nt; ++k) {
    const int32_t *t = tets + 4 * k;
    for (int i = 0; i < 10; ++i) {
      const uint64_t key = i < 4 ? vkey(t[i]) : ekey(t[kEdge[i - 4][0]], t[kEdge[i - 4][1]]);
      int fresh;
      *hfind(&h, key, &fresh) = elem_nodes[10 * k + i];
    }
  }
  memset(onbc, 0, (size_t)ndof);
  for (int64_t f = 0; f < nbe; ++f) {
    int on = 0;
    for (int q = 0; q < nlab; ++q) on |= blab[f] == labs[q];
    if (!on) continue;
    const int32_t *tr = btri + 3 * f;
    uint64_t keys[6] = {vkey(tr[0]), vkey(tr[1]), vkey(tr[2]),
                        ekey(tr[0], tr[1]), ekey(tr[1], tr[2]), ekey(tr[0], tr[2])};
    for (int i = 0; i < 6; ++i) {
      int fresh;
      const int32_t node = *hfind(&h, keys[i], &fresh);
      if (fresh) { free(h.k); free(h.v); return -3; }
      for (int c = 0; c < ncomp; ++c) onbc[ncomp * (int64_t)node + c] = 1;
    }
  }
  free(h.k);
  free(h.v);
  return 0;
}

/* ---- scalar P2: int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v))) on
 * Vh(Th,P2) (dof = node), Element_Op (problem.cpp:6384-6437) with
 * TypeOfFE_P2Lagrange3d::FB (P012_3d.cpp:199-270) and Tet_5. */
int ffo_p2_lap_element(const double *xyz, const int32_t *tet, int kind, double c0,
                       double kappa_e, double *Ke) {
  double v[12], G[12], qp[42], qw[14];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = xyz[3 * (int64_t)tet[i] + k];
  const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
  ffo_tet_gradlambda(v, mes, G);
  ffo_qf_tet5(qp, qw);
  memset(Ke, 0, sizeof(double) * 100);
  for (int q = 0; q < 14; ++q) {
    double f[10], d[30];
    const double *Ph = qp + 3 * q;
    ffo_p2_basis(Ph, G, f, d);
    const double coef = mes * qw[q];
    const double l0 = 1. - (Ph[0] + Ph[1] + Ph[2]);
    double P[3];
    for (int k = 0; k < 3; ++k) P[k] = l0 * v[k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[0] * v[3 + k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[1] * v[6 + k];
    for (int k = 0; k < 3; ++k) P[k] += Ph[2] * v[9 + k];
    for (int t = 0; t < 3; ++t) {
      double ccc = ffo_coef_eval(kind, c0, kappa_e, P, 3);
      ccc *= coef;
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) Ke[10 * i + j] += ccc * d[3 * i + t] * d[3 * j + t];
    }
  }
  return 0;
}

int ffo_assemble_p2_lap(int64_t nt, const double *xyz, const int32_t *tets,
                        const int32_t *elem_nodes, int kind, double c0, const double *kappa_e,
                        int64_t nn, const int64_t *rowptr, const int32_t *colidx, double *vals) {
  memset(vals, 0, sizeof(double) * (size_t)rowptr[nn]);
  double Ke[100];
  for (int64_t k = 0; k < nt; ++k) {
    ffo_p2_lap_element(xyz, tets + 4 * k, kind, c0, kappa_e ? kappa_e[k] : 0., Ke);
    const int32_t *d = elem_nodes + 10 * k;
    for (int il = 0; il < 10; ++il)
      for (int jl = 0; jl < 10; ++jl) {
        const int64_t s = slot_of(rowptr, colidx, d[il], d[jl]);
        if (s < 0) return -3;
        vals[s] += Ke[10 * il + jl];
      }
  }
  return 0;
}

/* Element_rhs for int3d(Th)(f*v) on Vh(Th,P2), f constant */
int ffo_rhs_p2(int64_t nt, const double *xyz, const int32_t *tets, const int32_t *elem_nodes,
               double fv, int64_t nn, double *b) {
  double qp[42], qw[14];
  ffo_qf_tet5(qp, qw);
  memset(b, 0, sizeof(double) * (size_t)nn);
  for (int64_t k = 0; k < nt; ++k) {
    const int32_t *t = tets + 4 * k;
    double v[12], G[12];
    for (int i = 0; i < 4; ++i)
      for (int kk = 0; kk < 3; ++kk) v[3 * i + kk] = xyz[3 * (int64_t)t[i] + kk];
    const double mes = ffo_tet_mesure(v, v + 3, v + 6, v + 9);
    ffo_tet_gradlambda(v, mes, G);
    for (int q = 0; q < 14; ++q) {
      double f[10], d[30];
      ffo_p2_basis(qp + 3 * q, G, f, d);
      const double coef = mes * qw[q];
      for (int i = 0; i < 10; ++i) b[elem_nodes[10 * k + i]] += coef * fv * f[i];
    }
  }
  return 0;
}
