// ref_driver.cpp — host driver around FreeFEM's own femlib sources, compiled by
// oracle/build_ref.sh into oracle/_ref/libffref_ref.so.
//
// TEST INFRASTRUCTURE ONLY: it is used by tests/ to pin the oracle against the
// reference's compiled code (HashMatrix accumulation / CSR sort / SetBC /
// gettgv / COO addMatMul, ConjugueGradient, fgmres, the Tet_5 and T_5
// tables, Tet geometry).  It is never part of the product and never travels as a
// dependency of libffhip.
//
// The reference pieces linked in: src/femlib/HashMatrix.cpp, CG.cpp,
// QuadratureFormular.cpp, and the mesh / finite-element sources that build a
// Mesh3 and its FESpace3 (Mesh*dn.cpp, FESpacen.cpp, P012_*.cpp, ...; the
// list is in oracle/build_ref.sh), compiled from /root/reference with
// -DCMAKE, the reference's own switch that skips the autoconf config.h.  What the driver
// restates itself (the interpreter-bound parts that cannot be compiled alone):
// the Element_Op quadrature loop (problem.cpp:6384-6437) and the two tiny
// CGMatVirt adaptors of VirtualSolverCG.hpp (HMatVirt 148-153, the diag1 branch
// of HMatVirtPrecon 55-76), whose header pulls the interpreter in.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "CG.hpp"
#include "HashMatrix.hpp"
#include "Mesh3dn.hpp"
#include "QuadratureFormular.hpp"
#include "FESpacen.hpp"

// The host-program globals (verbosity, mpirank, ShowDebugStack) come from the
// reference's src/Graphics/gggg.cpp, linked in.

using namespace Fem2D;

namespace {

typedef HashMatrix<int, double> HMat;

// A full element matrix carrier for HashMatrix::operator+=(MatriceElementaire&)
// (HashMatrix.cpp:1294-1332 only reads n, m, ni, nj, a and mtype).  The base
// destructor owns ni (delete[]), as MatriceElementairePleine's does.
struct ElemFull : public MatriceElementaire<double> {
  // (datasize, llga, nni, lk = 0, Full): ni == nj, no face data
  ElemFull(int nloc)
      : MatriceElementaire<double>(0, nloc * nloc, new int[nloc], 0, Full) {
    n = m = nloc;
  }
  double &operator()(int i, int j) { return a[i * m + j]; }
  void call(int, int, int, void *, void *) {}
};

struct HMatVirt : public CGMatVirt<int, double> {
  HMat *A;
  int t;
  HMatVirt(HMat *AA, int trans) : CGMatVirt<int, double>(AA->n), A(AA), t(trans) {}
  double *addmatmul(double *x, double *Ax) const { return A->addMatMul(x, Ax, t); }
};

struct Jacobi : public CGMatVirt<int, double> {
  std::vector<double> d;
  Jacobi(HMat *A) : CGMatVirt<int, double>(A->n), d(A->n) {
    for (int i = 0; i < A->n; ++i) d[i] = (*A)(i, i);
    for (int i = 0; i < A->n; ++i)
      d[i] = (std::norm(d[i]) < 1e-60) ? 1 : 1. / d[i];
  }
  double *addmatmul(double *x, double *Ax) const {
    for (int i = 0; i < n; ++i) Ax[i] += d[i] * x[i];
    return Ax;
  }
};

double coef_eval(int kind, double c0, double ke, const R3 &P) {
  if (kind == 0) return c0;
  if (kind == 2) return ke;
  return 1 + 0.5 * sin(2 * M_PI * P.x) * sin(2 * M_PI * P.y) * sin(2 * M_PI * P.z);
}

}  // namespace

extern "C" {

int ref_qf_tet5(double *pts, double *w) {
  const GQuadratureFormular<R3> &Q = QuadratureFormular_Tet_5;
  for (int q = 0; q < Q.n; ++q) {
    pts[3 * q] = Q[q].x; pts[3 * q + 1] = Q[q].y; pts[3 * q + 2] = Q[q].z;
    w[q] = Q[q].a;
  }
  return Q.n;
}

int ref_qf_tri5(double *pts, double *w) {
  const GQuadratureFormular<R2> &Q = QuadratureFormular_T_5;
  for (int q = 0; q < Q.n; ++q) {
    pts[2 * q] = Q[q].x; pts[2 * q + 1] = Q[q].y; w[q] = Q[q].a;
  }
  return Q.n;
}

// QF_GaussLegendre3 (QuadratureFormular.cpp:320-343): points on [0,1], weights
int ref_qf_gl3(double *x, double *w) {
  const QuadratureFormular1d &Q = QF_GaussLegendre3;
  for (int q = 0; q < Q.n; ++q) { x[q] = Q[q].x; w[q] = Q[q].a; }
  return Q.n;
}

int ref_tet_geom(const double *v12, double *mes, double *G12) {
  Vertex3 v[4];
  for (int i = 0; i < 4; ++i) {
    v[i].x = v12[3 * i]; v[i].y = v12[3 * i + 1]; v[i].z = v12[3 * i + 2]; v[i].lab = 0;
  }
  int iv[4] = {0, 1, 2, 3};
  Tet T;
  T.set(v, iv, 0);
  *mes = T.mesure();
  R3 G[4];
  T.Gradlambda(G);
  for (int i = 0; i < 4; ++i) { G12[3 * i] = G[i].x; G12[3 * i + 1] = G[i].y; G12[3 * i + 2] = G[i].z; }
  return 0;
}

void *ref_hm_new(int n) {
  verbosity = 0;
  // as creationBlockOfMatrixToBilinearForm does (problem.hpp:1688)
  return new HMat(n, n, 2 * n, 0);
}
void ref_hm_free(void *h) { delete (HMat *)h; }

// Element loop (problem.cpp:1398-1405) with the Element_Op 3D quadrature loop
// (problem.cpp:6384-6437) for kappa*(dx dx + dy dy + dz dz), accumulated by the
// reference HashMatrix::operator+=.
int ref_hm_add_p1_3d(void *h, int nv, int64_t nt, const double *xyz,
                     const int32_t *tets, int kind, double c0, const double *kappa_e) {
  HMat &A = *(HMat *)h;
  std::vector<Vertex3> V(nv);
  for (int i = 0; i < nv; ++i) {
    V[i].x = xyz[3 * i]; V[i].y = xyz[3 * i + 1]; V[i].z = xyz[3 * i + 2]; V[i].lab = 0;
  }
  const GQuadratureFormular<R3> &FI = QuadratureFormular_Tet_5;
  ElemFull me(4);
  int *ni = me.ni;
  for (int64_t e = 0; e < nt; ++e) {
    int iv[4];
    for (int k = 0; k < 4; ++k) ni[k] = iv[k] = tets[4 * e + k];
    Tet T;
    T.set(V.data(), iv, 0);
    R3 Dl[4];
    T.Gradlambda(Dl);
    double *a = me.a;
    for (int i = 0; i < 16; ++i) a[i] = 0.;
    for (int npi = 0; npi < FI.n; npi++) {
      GQuadraturePoint<R3> pi(FI[npi]);
      double coef = T.mesure() * pi.a;
      R3 Pt(pi);
      R3 P = T(Pt);
      for (int t = 0; t < 3; ++t) {
        double ccc = coef_eval(kind, c0, kappa_e ? kappa_e[e] : 0., P);
        ccc *= coef;
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            double w_i = Dl[i][t], w_j = Dl[j][t];
            me(i, j) += ccc * w_i * w_j;
          }
      }
    }
    A += me;
  }
  return 0;
}

int ref_hm_setbc(void *h, const uint8_t *onbc, double tgv) {
  HMat &A = *(HMat *)h;
  std::vector<char> w(A.n);
  for (int i = 0; i < A.n; ++i) w[i] = onbc[i] ? '1' : '\0';
  A.SetBC(w.data(), tgv);
  return 0;
}

int64_t ref_hm_nnz(void *h) { return (int64_t)((HMat *)h)->nnz; }

double ref_hm_gettgv(void *h, int *ntgv) { return ((HMat *)h)->gettgv(ntgv); }

// y = A x with the COO product in current storage order (HashMatrix.cpp:1086-1154)
int ref_hm_matmul(void *h, const double *x, double *y, int trans) {
  HMat &A = *(HMat *)h;
  std::fill(y, y + A.n, 0.);
  A.addMatMul(const_cast<double *>(x), y, trans);
  return 0;
}

// HashMatrix::CSR (Sortij + Buildp, HashMatrix.cpp:858-867); copies out.
int ref_hm_csr(void *h, int64_t *rowptr, int32_t *colidx, double *vals) {
  HMat &A = *(HMat *)h;
  int *ia, *ja;
  double *aa;
  A.CSR(ia, ja, aa);
  for (int i = 0; i <= A.n; ++i) rowptr[i] = ia[i];
  for (size_t k = 0; k < A.nnz; ++k) { colidx[k] = ja[k]; vals[k] = aa[k]; }
  return 0;
}

// SolverCG::dosolver (VirtualSolverCG.hpp:169-190): SetInitWithBC, then the
// reference ConjugueGradient with HMatVirt + Jacobi.
int ref_hm_cg(void *h, const double *b, double *x, double eps, int itmax,
              double tgv, int trans, double *stats) {
  HMat &A = *(HMat *)h;
  int ntgv1 = 0;
  double tgvm = A.gettgv(&ntgv1);
  if (ntgv1) {
    double tgve = tgvm <= 0 ? 1e200 : tgvm;
    for (int i = 0; i < A.n; ++i)
      if (A(i, i) == tgve) x[i] = b[i] / tgv;
  }
  HMatVirt AA(&A, trans);
  Jacobi C(&A);
  int itermx = itmax > 0 ? itmax : A.n;
  double epss = eps;
  std::vector<double> bb(b, b + A.n);
  int res = ConjugueGradient<int, double>(AA, C, bb.data(), x, itermx, epss, 0);
  if (stats) { stats[0] = itermx; stats[1] = epss; stats[3] = res; }
  return res;
}

// a HashMatrix holding a given CSR matrix (HashMatrix::set with CSR input,
// HashMatrix.cpp:697-728, as a plugin hands a device matrix back)
void *ref_hm_from_csr(int n, const int64_t *rowptr, const int32_t *colidx, const double *vals) {
  verbosity = 0;
  const size_t nnz = (size_t)rowptr[n];
  std::vector<int> ia(n + 1), ja(colidx, colidx + nnz);
  for (int i = 0; i <= n; ++i) ia[i] = (int)rowptr[i];
  std::vector<double> aa(vals, vals + nnz);
  HMat *A = new HMat(n, n, (int)nnz, 0);
  A->set(n, n, 0, nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  return A;
}

// the same for a half (lower-triangle, sym=1) matrix: HashMatrix::half = 1
void *ref_hm_from_csr_half(int n, const int64_t *rowptr, const int32_t *colidx, const double *vals) {
  verbosity = 0;
  const size_t nnz = (size_t)rowptr[n];
  std::vector<int> ia(n + 1), ja(colidx, colidx + nnz);
  for (int i = 0; i <= n; ++i) ia[i] = (int)rowptr[i];
  std::vector<double> aa(vals, vals + nnz);
  HMat *A = new HMat(n, n, (int)nnz, 1);
  A->set(n, n, 1, nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  return A;
}

// SolverGMRES::dosolver (VirtualSolverCG.hpp:236-255): SetInitWithBC, then the
// reference fgmres with HMatVirt + the Jacobi diag1 of HMatVirtPrecon and its
// tgv-row mask.
int ref_hm_gmres(void *h, const double *b, double *x, double eps, int itmax, int restart,
                 double tgv, double *stats) {
  HMat &A = *(HMat *)h;
  int ntgv1 = 0;
  double tgvm = A.gettgv(&ntgv1);
  std::vector<int> wcl;
  if (ntgv1) {
    double tgve = tgvm <= 0 ? 1e200 : tgvm;
    wcl.resize(A.n);
    for (int i = 0; i < A.n; ++i) wcl[i] = A(i, i) == tgve;
    for (int i = 0; i < A.n; ++i)
      if (wcl[i]) x[i] = b[i] / tgv;
  }
  HMatVirt AA(&A, 0);
  Jacobi C(&A);
  int itermx = itmax > 0 ? itmax : A.n;
  double veps = eps;
  std::vector<double> bb(b, b + A.n);
  bool res = fgmres<double, int>(AA, C, 1, bb.data(), x, veps, itermx, restart, 0,
                                 wcl.empty() ? nullptr : wcl.data());
  if (stats) { stats[0] = itermx; stats[3] = res; }
  return res ? 1 : 0;
}

// Mesh3 (Mesh3dn.cpp:1416) from arrays, owned by the returned object
static Mesh3 *make_mesh3(int nv, const double *xyz, int64_t nt, const int32_t *tets) {
  verbosity = 0;
  Vertex3 *v = new Vertex3[nv];
  for (int i = 0; i < nv; ++i) {
    v[i].x = xyz[3 * i]; v[i].y = xyz[3 * i + 1]; v[i].z = xyz[3 * i + 2]; v[i].lab = 0;
  }
  Tet *t = new Tet[nt];
  for (int64_t k = 0; k < nt; ++k) {
    int iv[4] = {tets[4 * k], tets[4 * k + 1], tets[4 * k + 2], tets[4 * k + 3]};
    t[k].set(v, iv, 0);
  }
  return new Mesh3(nv, (int)nt, 0, v, t, new Triangle3[0]);
}

// Global dofs of every element of an FESpace3 built by FreeFEM's own
// BuildDFNumbering (GenericMesh.hpp:1711-1954): kind 1 = P1, 2 = P2,
// 3 = [P2,P2,P2] (GTypeOfFESum, FESpacen.cpp:403-428).  Returns NbOfDF.
int64_t ref_fespace_dofs(int nv, const double *xyz, int64_t nt, const int32_t *tets, int kind,
                         int32_t *elem_dofs, int *ndofloc) {
  Mesh3 *Th = make_mesh3(nv, xyz, nt, tets);
  FESpace3 *Vh = nullptr;
  GTypeOfFESum<Mesh3> *sum = nullptr;
  if (kind == 1) Vh = new FESpace3(*Th, DataFE<Mesh3>::P1);
  else if (kind == 2) Vh = new FESpace3(*Th, DataFE<Mesh3>::P2);
  else {
    KN<GTypeOfFE<Mesh3> const *> t(3);
    t[0] = t[1] = t[2] = &DataFE<Mesh3>::P2;
    sum = new GTypeOfFESum<Mesh3>(t);
    Vh = new FESpace3(*Th, *sum);
  }
  const int nd = (*Vh)[0].NbDoF();
  *ndofloc = nd;
  if (elem_dofs)
    for (int64_t k = 0; k < nt; ++k) {
      FElement3 K((*Vh)[(int)k]);
      for (int i = 0; i < nd; ++i) elem_dofs[k * nd + i] = K(i);
    }
  const int64_t ndf = Vh->NbOfDF;
  delete Vh;  // releases its counted reference to the mesh
  delete sum;
  Th->destroy();  // RefCounter (RefCounter.hpp:55-66)
  return ndf;
}

// `ofstream f(path); f << A;` (HashMatrix.hpp:485-521): the Morse text after
// HashMatrix::CSR when csr != 0, else the hash-state COO dump
int ref_hm_write(void *h, const char *path, int csr) {
  HMat &A = *(HMat *)h;
  if (csr) A.CSR();
  std::ofstream f(path);
  f << A;
  return f.good() ? 0 : -1;
}

// HashMatrix(istream, -1) (HashMatrix.cpp:7-30, 136-214) then CSR: nmh =
// {n, m, half}; fills rowptr/colidx/vals when nnz <= cap; returns nnz
int64_t ref_hm_read(const char *path, int64_t cap, int64_t *nmh, int64_t *rowptr, int32_t *colidx,
                    double *vals) {
  std::ifstream f(path);
  HMat A(f, -1);
  int *ia, *ja;
  double *aa;
  A.CSR(ia, ja, aa);
  nmh[0] = A.n; nmh[1] = A.m; nmh[2] = A.half;
  if ((int64_t)A.nnz <= cap) {
    for (int i = 0; i <= A.n; ++i) rowptr[i] = ia[i];
    for (size_t k = 0; k < A.nnz; ++k) { colidx[k] = ja[k]; vals[k] = aa[k]; }
  }
  return (int64_t)A.nnz;
}

// the HashMatrix::COO() arrays (HashMatrix.cpp:837-843) that [I, J, C] = A
// copies (lgmat.cpp:2690-2700)
int64_t ref_hm_coo(void *h, int64_t *I, int64_t *J, double *C) {
  HMat &A = *(HMat *)h;
  A.COO();
  for (size_t k = 0; k < A.nnz; ++k) { I[k] = A.i[k]; J[k] = A.j[k]; C[k] = A.aij[k]; }
  return (int64_t)A.nnz;
}

// P2 basis values and gradients on one tet at a reference point
// (TypeOfFE_P2Lagrange3d::FB, P012_3d.cpp:199-270): f[10], df[10*3]
int ref_p2_fb(const double *v12, const double *phat, double *f, double *df) {
  const int32_t tet[4] = {0, 1, 2, 3};
  Mesh3 *Th = make_mesh3(4, v12, 1, tet);
  {
    FESpace3 Vh(*Th, DataFE<Mesh3>::P2);
    FElement3 K(Vh[0]);
    KNMK<double> val(10, 1, last_operatortype);  // FB zeroes every operator slot
    K.BF(Fop_D0 | Fop_D1, R3(phat[0], phat[1], phat[2]), val);
    for (int i = 0; i < 10; ++i) {
      f[i] = val(i, 0, op_id);
      df[3 * i] = val(i, 0, op_dx); df[3 * i + 1] = val(i, 0, op_dy); df[3 * i + 2] = val(i, 0, op_dz);
    }
  }
  Th->destroy();
  return 0;
}

}  // extern "C"
