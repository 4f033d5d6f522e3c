// ffhip.cpp — FreeFEM plugin: `load "ffhip"`.
//
// Drop-in MI355X path for
//   matrix A = va(Vh, Vh, solver=HIPCG);   u[] = A^-1 * b;
// on 3D P1 Poisson-type forms, through libffhip's C-ABI (include/ffhip.h):
//
//  * solvers "HIPCG" (FP64 Jacobi-PCG) and "HIPCGMIXED" (FP32 matrix and
//    inner CG, FP64 residual refinement), registered with addsolver
//    (src/femlib/SparseLinearSolver.hpp:96-102) with the orTypeSol of SolverCG
//    (src/femlib/VirtualSolverCG.hpp:116) so the matrix gets the same full
//    (non-half) storage as with solver=CG.  `ffhipreplacecg(1)` re-points
//    solver=CG itself (TheFFSolver::ChangeSolver, SparseLinearSolver.hpp:59-71).
//  * solver "HIPGMRES": restarted GMRES with the Jacobi diagonal
//    (SolverGMRES, VirtualSolverCG.hpp:196-255) for non-symmetric matrices.
//  * a higher-preference "=" / "<-" operator for
//    (Matrice_Creuse<double>*, Call_FormBilinear<v_fes3,v_fes3>*)
//    (the overload of OpMatrixtoBilinearForm, src/fflib/problem.hpp:897-927,
//    registered at src/fflib/lgfem.cpp:6669-6683; the highest pref wins,
//    src/fflib/AFunction2.cpp:87-107).  It recognises
//        int3d(Th)(c*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v))) [+ on(labels,u=g)]
//    [+ int3d(Th)(m*u*v)] (or the reaction term alone)
//    [+ int2d(Th, labels)(alpha*u*v) ...] (Robin terms, P1 only, default T_5)
//    on a scalar P1 space (the Laplacian also on a scalar P2 space, whose
//    dof numbering is checked against the device's first-appearance node
//    numbering) with the default Tet_5 rule; the coefficient c is
//    evaluated by FreeFEM at the 14 quadrature points of every element (as
//    Element_Op does, problem.cpp:6392-6395) unless it is a constant, and the
//    pattern, values and tgv rows are built on the GPU.  It also recognises
//    linear elasticity on Vh(Th,[P2,P2,P2]):
//        int3d(Th)(lambda*div(u)*div(v) + 2*mu*(epsilon(u)'*epsilon(v)))
//        [+ on(labels, u1=.., u2=.., u3=..)]
//    in any spelling whose first-derivative terms have mesh-independent
//    coefficients that add up to the Lame table
//        C[c][a][c'][b] = lambda d_ca d_c'b + mu (d_cb d_ac' + d_cc' d_ab)
//    (test component c, derivative a; trial component c', derivative b).
//    With sym=1 the P1 forms keep the lower triangle (HashMatrix::half), as
//    FreeFEM's symmetric element matrices do.
//    Any other form falls back to FreeFEM's own
//    creationBlockOfMatrixToBilinearForm (problem.hpp:1663-1714), as the
//    PETSc plugin does for its Mat type.
//  * the same for 2D P1 forms (Vh on a mesh):
//        int2d(Th)(c*(dx(u)*dx(v)+dy(u)*dy(v))) [+ int2d(Th)(m*u*v)]
//        [+ int1d(Th, labels)(alpha*u*v)] [+ on(labels, u=g)]
//  * a higher-preference "=" / "<-" for (real[int], varf(0, Vh)) on 3D P1:
//        int3d(Th)(f*v) [+ int2d(Th, labels)(g*v)] [+ on(labels, u=g)]
//    with f, g evaluated by FreeFEM at the quadrature points (the on() value at
//    the boundary vertices) and the vector built on the GPU; on P2 and
//    [P2,P2,P2] spaces with mesh-independent f and on() values; anything else
//    goes to FreeFEM's OpArraytoLinearForm code path.
//  * LinearCG / AffineCG overloads whose operator is a matrix (the affine
//    ConjuguedGradient2 of LinearCG on the GPU, see HipLinearCG).
//
// Build (FreeFEM's plugin script):  ff-c++ -auto ffhip.cpp -I<repo>/include
//   -L<repo>/freefem-sources_amd -lffhip
#include <ff++.hpp>

#include <cstring>
#include <set>
#include <vector>

#include "AFunction_ext.hpp"
#include "problem.hpp"

#include "ffhip.h"

using namespace std;
using namespace Fem2D;

// defined in src/fflib/lgfem.cpp:7792 (namespace Fem2D), exported by the
// FreeFEM executable; declared where it is used, as src/fflib/problem.cpp:109 does
namespace Fem2D {
void Expandsetoflab(Stack stack, const BC_set &bc, set<long> &setoflab);
void Expandsetoflab(Stack stack, const CDomainOfIntegration &di, set<int> &setoflab, bool &all);
}

namespace {

ffhip_ctx hip_ctx() {
  static ffhip_ctx ctx = nullptr;
  if (!ctx) {
    if (ffhip_init(0, &ctx) != FFHIP_OK) {
      cerr << "ffhip: " << ffhip_last_error() << endl;
      ExecError("ffhip: no usable MI355X device");
    }
  }
  return ctx;
}

void hip_check(int rc, const char *what) {
  if (rc == FFHIP_OK) return;
  cerr << "ffhip: " << what << ": " << ffhip_last_error() << endl;
  ExecError(what);
}

// ---------------------------------------------------------------- solver
template <int MODE>
class SolverHIPCG : public VirtualSolver<int, double> {
 public:
  // 1 unsym, 2 herm, 4 sym, 8 pos, 32 seq: as SolverCG
  static const int orTypeSol = 1 | 2 | 4 | 8 | 32;
  typedef HashMatrix<int, double> HMat;
  HMat *ptA;
  ffhip_mat M;
  ffhip_cg cg;
  int verb, itermax;
  double eps, tgv;
  double *veps;
  long *getnbiter;
  long cn, cs;

  SolverHIPCG(HMat &A, const Data_Sparse_Solver &ds, Stack)
      : ptA(&A), M(nullptr), cg(nullptr), verb(ds.verb), itermax(ds.itmax > 0 ? ds.itmax : A.n),
        eps(ds.epsilon), tgv(ds.tgv), veps(ds.veps), getnbiter(ds.getnbiter), cn(0), cs(0) {
    ffassert(A.n == A.m);
    if (verb > 4)
      cout << " ## SolverHIPCG" << (MODE ? "MIXED " : " ") << A.n << " eps " << eps << " itmax "
           << itermax << endl;
  }

  void UpdateState() {
    if (ptA->GetReDoNumerics()) cn++;
    if (ptA->GetReDoSymbolic()) cs++;
    this->ChangeCodeState(ptA->n, cs, cn);
  }

  // the pattern: HashMatrix::CSR (sorted, HashMatrix.cpp:858-867) to the GPU
  void fac_symbolic() {
    int *ia = 0, *ja = 0;
    double *a = 0;
    ptA->CSR(ia, ja, a);
    const int64_t n = ptA->n, nnz = ia[n];
    vector<int64_t> rp(ia, ia + n + 1);
    if (cg) { ffhip_cg_free(cg); cg = nullptr; }
    if (M) { ffhip_mat_free(M); M = nullptr; }
    hip_check(ffhip_mat_upload(hip_ctx(), n, nnz, rp.data(), ja, a, &M), "HIPCG: matrix upload");
    hip_check(ffhip_cg_create(hip_ctx(), M, MODE ? FFHIP_CG_MIXED : FFHIP_CG_FP64, &cg),
              "HIPCG: solver setup");
  }

  void fac_numeric() {
    int *ia = 0, *ja = 0;
    double *a = 0;
    ptA->CSR(ia, ja, a);
    hip_check(ffhip_mat_set_values(M, a), "HIPCG: matrix values");
    hip_check(ffhip_cg_update_values(cg), "HIPCG: solver values");
  }

  // SolverCG::dosolver semantics (VirtualSolverCG.hpp:169-190)
  void dosolver(double *x, double *b, int N, int trans) {
    int err = 0;
    if (getnbiter) *getnbiter = 0;
    for (int k = 0; k < N; ++k) {
      ffhip_cg_stats st;
      const int rc = ffhip_cg_solve(cg, b + (size_t)k * ptA->n, x + (size_t)k * ptA->n, 0, eps,
                                    itermax, tgv, trans, &st);
      if (rc == FFHIP_ERR_NOCONV) err++;
      else hip_check(rc, "HIPCG: solve");
      if (rc == FFHIP_OK && getnbiter) *getnbiter += st.iters;
      if (veps) *veps = st.gcg;
      if (verb > 1)
        cout << " HIPCG: " << (st.converged ? "converge after " : "no convergence after ") << st.iters
             << " g=" << st.gcg << " (" << st.ms_solve << " ms on the GPU)" << endl;
    }
    if (err) {
      cerr << "Error: ConjugueGradient do not converge nb end =" << err << endl;
      ffassert(0);
    }
  }

  ~SolverHIPCG() {
    // device buffers only: the HashMatrix may already be deleted
    // (type_isdeleted, HashMatrix.hpp:62) and is not touched here
    if (cg) ffhip_cg_free(cg);
    if (M) ffhip_mat_free(M);
  }
};

// solver=HIPGMRES: SolverGMRES (VirtualSolverCG.hpp:196-255) — restarted
// right-preconditioned GMRES (fgmres, CG.cpp:347-517) with the Jacobi
// diagonal, for non-symmetric matrices; restart = dimKrylov (NbSpace)
class SolverHIPGMRES : public VirtualSolver<int, double> {
 public:
  // 1 unsym, 2 herm, 4 sym, 8 pos, 16 nopos, 32 seq: as SolverGMRES
  static const int orTypeSol = 1 | 2 | 4 | 8 | 16 | 32;
  typedef HashMatrix<int, double> HMat;
  HMat *ptA;
  ffhip_mat M;
  ffhip_cg cg;
  int verb, itermax, restart;
  double eps, tgv;
  double *veps;
  long *getnbiter;
  long cn, cs;

  SolverHIPGMRES(HMat &A, const Data_Sparse_Solver &ds, Stack)
      : ptA(&A), M(nullptr), cg(nullptr), verb(ds.verb), itermax(ds.itmax > 0 ? ds.itmax : A.n),
        restart(ds.NbSpace), eps(ds.epsilon), tgv(ds.tgv), veps(ds.veps),
        getnbiter(ds.getnbiter), cn(0), cs(0) {
    ffassert(A.n == A.m);
  }

  void UpdateState() {
    if (ptA->GetReDoNumerics()) cn++;
    if (ptA->GetReDoSymbolic()) cs++;
    this->ChangeCodeState(ptA->n, cs, cn);
  }

  void fac_symbolic() {
    int *ia = 0, *ja = 0;
    double *a = 0;
    ptA->CSR(ia, ja, a);
    const int64_t n = ptA->n, nnz = ia[n];
    vector<int64_t> rp(ia, ia + n + 1);
    if (cg) { ffhip_cg_free(cg); cg = nullptr; }
    if (M) { ffhip_mat_free(M); M = nullptr; }
    hip_check(ffhip_mat_upload(hip_ctx(), n, nnz, rp.data(), ja, a, &M), "HIPGMRES: matrix upload");
    hip_check(ffhip_cg_create(hip_ctx(), M, FFHIP_CG_FP64, &cg), "HIPGMRES: solver setup");
  }

  void fac_numeric() {
    int *ia = 0, *ja = 0;
    double *a = 0;
    ptA->CSR(ia, ja, a);
    hip_check(ffhip_mat_set_values(M, a), "HIPGMRES: matrix values");
    hip_check(ffhip_cg_update_values(cg), "HIPGMRES: solver values");
  }

  // SolverGMRES::dosolver semantics (VirtualSolverCG.hpp:236-255)
  void dosolver(double *x, double *b, int N, int trans) {
    int err = 0;
    for (int k = 0; k < N; ++k) {
      ffhip_cg_stats st;
      const int rc = ffhip_gmres_solve(cg, b + (size_t)k * ptA->n, x + (size_t)k * ptA->n, 0, eps,
                                       itermax, restart, tgv, trans, &st);
      if (rc == FFHIP_ERR_NOCONV) err++;
      else hip_check(rc, "HIPGMRES: solve");
      if (veps) *veps = eps;
      if (getnbiter) *getnbiter = st.iters;
      if (verb > 1)
        cout << " HIPGMRES: " << (st.converged ? "converged in " : "no convergence after ") << st.iters
             << " iterations, relative residual " << st.gcg << " (" << st.ms_solve << " ms on the GPU)"
             << endl;
    }
    if (err) {
      cerr << "Error: fgmres do not converge nb end =" << err << endl;
      ffassert(0);
    }
  }

  ~SolverHIPGMRES() {
    if (cg) ffhip_cg_free(cg);
    if (M) ffhip_mat_free(M);
  }
};

long ffhip_replace_cg(long on) {
  if (on) TheFFSolver<int, double>::ChangeSolver("CG", "HIPCG");
  return on;
}

// ---------------------------------------------------------------- assembly
typedef Call_FormBilinear<v_fes3, v_fes3> CFB3;

struct Recognised {
  const FormBilinear *bf = nullptr;
  int order = 1;              // P1 or P2 (scalar)
  bool has_grad = false;      // the three derivative terms
  C_F0 coef;                  // their common coefficient
  bool coef_const = false;
  double c0 = 0.;
  bool has_mass = false;      // a reaction term m*u*v
  C_F0 mcoef;
  bool mcoef_const = false;
  double m0 = 0.;
  set<long> labels;           // Dirichlet labels (all BC_set of the form)
  struct Boundary {           // int2d(Th, labels)(alpha*u*v)
    C_F0 coef;
    bool is_const = false;
    double c0 = 0.;
    set<int> labels;
    bool all = true;          // no label list: every boundary element
  };
  vector<Boundary> bnd;
};

// int2d(Th, labels)(alpha*u*v) on P1 with the default face rule: appended to r.bnd
bool recognise_boundary(Stack stack, const FESpace3 &Uh, const FormBilinear *bf, Recognised &r) {
  if (bf->di->islevelset()) return false;
  if (GetAny<const Mesh3 *>((*bf->di->Th)(stack)) != &Uh.Th) return false;
  if (bf->di->FIT(stack).n != QuadratureFormular_T_5.n) return false;  // default T_5 only
  const BilinearOperator &Op = *bf->b;
  if (Op.v.size() != 1) return false;
  const BilinearOperator::K &ll = *Op.v.begin();
  if (ll.first.first.first || ll.first.second.first || ll.first.first.second != op_id ||
      ll.first.second.second != op_id)
    return false;
  Recognised::Boundary t;
  t.coef = ll.second;
  if (t.coef.EvaluableWithOutStack()) {
    t.is_const = true;
    t.c0 = GetAny<double>(t.coef.eval(stack));
  }
  Fem2D::Expandsetoflab(stack, *bf->di, t.labels, t.all);
  r.bnd.push_back(t);
  return true;
}

// int3d(Th)(c*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v)) + m*u*v) + on(...) on P1
// (either part may be absent), or false
bool recognise(Stack stack, const FESpace3 &Uh, const FESpace3 &Vh, const list<C_F0> &largs,
               Recognised &r) {
  if (&Uh != &Vh || Uh.N != 1 || Uh.TFE.N() != 1) return false;
  if (Uh.TFE[0] == &DataFE<Mesh3>::P1) {
    if (Uh.NbOfDF != Uh.Th.nv) return false;
    r.order = 1;
  } else if (Uh.TFE[0] == &DataFE<Mesh3>::P2) {
    r.order = 2;
  } else {
    return false;
  }
  FreeFempp::TypeVarForm<double> *tvf = FreeFempp::TypeVarForm<double>::Global;
  for (list<C_F0>::const_iterator ii = largs.begin(); ii != largs.end(); ++ii) {
    const aType t = ii->left();
    const Expression e = ii->LeftValue();
    if (t == tvf->tFL) continue;  // linear terms do not enter the matrix
    if (t == tvf->tBC) {
      const BC_set *bc = dynamic_cast<const BC_set *>(e);
      if (!bc) return false;
      for (size_t k = 0; k < bc->bc.size(); ++k)
        if (bc->bc[k].first != 0) return false;
      set<long> labs;
      Fem2D::Expandsetoflab(stack, *bc, labs);
      r.labels.insert(labs.begin(), labs.end());
      continue;
    }
    if (t != tvf->tFB) return false;
    const FormBilinear *bf = dynamic_cast<const FormBilinear *>(e);
    if (!bf) return false;
    if (bf->di->kind == CDomainOfIntegration::int2d && r.order == 1) {
      if (!recognise_boundary(stack, Uh, bf, r)) return false;
      continue;
    }
    if (r.bf) return false;  // exactly one volume term
    if (bf->di->kind != CDomainOfIntegration::int3d || !bf->di->what.empty()) return false;
    if (GetAny<const Mesh3 *>((*bf->di->Th)(stack)) != &Uh.Th) return false;
    if (bf->di->FIV(stack).n != QuadratureFormular_Tet_5.n) return false;  // default Tet_5 only
    const BilinearOperator &Op = *bf->b;
    bool seen[4] = {false, false, false, false};
    int ngrad = 0;
    for (BilinearOperator::const_iterator l = Op.v.begin(); l != Op.v.end(); ++l) {
      const BilinearOperator::K &ll = *l;
      const int jcomp = ll.first.first.first, jop = ll.first.first.second;
      const int icomp = ll.first.second.first, iop = ll.first.second.second;
      if (jcomp || icomp || jop != iop || jop < op_id || jop > op_dz || seen[jop]) return false;
      seen[jop] = true;
      if (jop == op_id) {
        r.has_mass = true;
        r.mcoef = ll.second;
      } else if (ngrad++ == 0) {
        r.coef = ll.second;
      } else if (ll.second.LeftValue() != r.coef.LeftValue()) {
        return false;
      }
    }
    if (ngrad != 0 && ngrad != 3) return false;
    r.has_grad = ngrad == 3;
    if (!r.has_grad && !r.has_mass) return false;
    r.bf = bf;
  }
  if (!r.bf) return false;
  if (r.order == 2 && (r.has_mass || !r.has_grad)) return false;  // P2: the Laplacian only
  if (r.has_grad && r.coef.EvaluableWithOutStack()) {
    r.coef_const = true;
    r.c0 = GetAny<double>(r.coef.eval(stack));
  }
  if (!r.has_grad) {  // no derivative terms: kappa = 0
    r.coef_const = true;
    r.c0 = 0.;
  }
  if (r.has_mass && r.mcoef.EvaluableWithOutStack()) {
    r.mcoef_const = true;
    r.m0 = GetAny<double>(r.mcoef.eval(stack));
  }
  return true;
}

// ---------------------------------------------------------------- [P2,P2,P2]
struct RecognisedP2 {
  const FormBilinear *bf = nullptr;
  double lambda = 0., mu = 0.;
  set<long> labels;
};

// Vh(Th,[P2,P2,P2]): a GTypeOfFESum of three P2 (FESpacen.cpp:403-428, 531-575)
bool is_p2v(const FESpace3 &Uh) {
  if (Uh.N != 3 || Uh.TFE.N() != 1) return false;
  const GTypeOfFESum<Mesh3> *sum = dynamic_cast<const GTypeOfFESum<Mesh3> *>(Uh.TFE[0]);
  if (!sum || sum->k != 3) return false;
  for (int c = 0; c < 3; ++c)
    if (sum->teb[c] != &DataFE<Mesh3>::P2) return false;
  return true;
}

bool recognise_p2v(Stack stack, const FESpace3 &Uh, const FESpace3 &Vh, const list<C_F0> &largs,
                   RecognisedP2 &r) {
  if (&Uh != &Vh || !is_p2v(Uh)) return false;
  FreeFempp::TypeVarForm<double> *tvf = FreeFempp::TypeVarForm<double>::Global;
  double C[3][3][3][3];
  memset(C, 0, sizeof(C));
  for (list<C_F0>::const_iterator ii = largs.begin(); ii != largs.end(); ++ii) {
    const aType t = ii->left();
    const Expression e = ii->LeftValue();
    if (t == tvf->tFL) continue;
    if (t == tvf->tBC) {  // every component of the listed labels
      const BC_set *bc = dynamic_cast<const BC_set *>(e);
      if (!bc) return false;
      set<int> comps;
      for (size_t k = 0; k < bc->bc.size(); ++k) comps.insert(bc->bc[k].first);
      if (comps.size() != 3 || *comps.begin() != 0 || *comps.rbegin() != 2) return false;
      set<long> labs;
      Fem2D::Expandsetoflab(stack, *bc, labs);
      r.labels.insert(labs.begin(), labs.end());
      continue;
    }
    if (t != tvf->tFB || r.bf) return false;
    const FormBilinear *bf = dynamic_cast<const FormBilinear *>(e);
    if (!bf || bf->di->kind != CDomainOfIntegration::int3d || !bf->di->what.empty()) return false;
    if (GetAny<const Mesh3 *>((*bf->di->Th)(stack)) != &Uh.Th) return false;
    if (bf->di->FIV(stack).n != QuadratureFormular_Tet_5.n) return false;
    const BilinearOperator &Op = *bf->b;
    for (BilinearOperator::const_iterator l = Op.v.begin(); l != Op.v.end(); ++l) {
      const BilinearOperator::K &ll = *l;
      // trial (u) component/derivative first, test (v) second (problem.cpp:6404-6405)
      const int jcomp = ll.first.first.first, jop = ll.first.first.second;
      const int icomp = ll.first.second.first, iop = ll.first.second.second;
      if (jop < op_dx || jop > op_dz || iop < op_dx || iop > op_dz) return false;
      if (jcomp < 0 || jcomp > 2 || icomp < 0 || icomp > 2) return false;
      if (!ll.second.LeftValue()->MeshIndependent()) return false;
      C[icomp][iop - op_dx][jcomp][jop - op_dx] += GetAny<double>(ll.second.eval(stack));
    }
    r.bf = bf;
  }
  if (!r.bf) return false;
  const double lam = C[0][0][1][1], mu = C[0][1][1][0];
  const double scale = max(fabs(lam), fabs(mu));
  if (!(scale > 0)) return false;
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a)
      for (int cp = 0; cp < 3; ++cp)
        for (int b = 0; b < 3; ++b) {
          const double want = lam * (c == a) * (cp == b) + mu * ((c == b) * (a == cp) + (c == cp) * (a == b));
          if (fabs(C[c][a][cp][b] - want) > 1e-12 * scale) return false;
        }
  r.lambda = lam;
  r.mu = mu;
  return true;
}

// true when built on the GPU; false (nothing changed) when FreeFEM's dof
// numbering of Uh is not the first-appearance numbering the device builds
bool assemble_p2v_on_gpu(const FESpace3 &Uh, const RecognisedP2 &r, double tgv,
                         Matrice_Creuse<double> &A) {
  const Mesh3 &Th = Uh.Th;
  vector<double> xyz(3 * (size_t)Th.nv);
  vector<int32_t> tets(4 * (size_t)Th.nt), btri(3 * (size_t)Th.nbe), blab(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    xyz[3 * i] = Th(i).x; xyz[3 * i + 1] = Th(i).y; xyz[3 * i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) tets[4 * k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) btri[3 * k + j] = Th(Th.be(k)[j]);
    blab[k] = Th.be(k).lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  ffhip_p2v sp = nullptr;
  hip_check(ffhip_mesh3_upload(ctx, Th.nv, Th.nt, Th.nbe, xyz.data(), tets.data(), btri.data(),
                               blab.data(), &m), "ffhip: mesh upload");
  hip_check(ffhip_p2v_create(ctx, m, &sp), "ffhip: P2 numbering");
  vector<int32_t> en(10 * (size_t)Th.nt);
  hip_check(ffhip_p2v_download(sp, en.data()), "ffhip: P2 numbering");
  bool same = true;
  for (int k = 0; k < Th.nt && same; ++k) {
    const FElement3 K(Uh[k]);
    for (int df = 0; df < 30 && same; ++df)  // local dofs comp-blocked (c*10 + i)
      same = K(df) == 3 * en[10 * (size_t)k + df % 10] + df / 10;
  }
  if (!same) {
    ffhip_p2v_free(sp);
    ffhip_mesh_free(m);
    return false;
  }
  ffhip_mat M = nullptr;
  hip_check(ffhip_pattern_p2v(ctx, sp, &M), "ffhip: P2 pattern");
  hip_check(ffhip_assemble_p2v_elasticity(ctx, sp, M, r.lambda, r.mu), "ffhip: P2 assembly");
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc_p2v(ctx, sp, M, labs.data(), (int)labs.size(), tgv, nullptr), "ffhip: on()");
  }
  int64_t n = 0, nnz = 0;
  ffhip_mat_info(M, &n, &nnz);
  vector<int64_t> rp(n + 1);
  vector<int> ia(n + 1), ja(nnz);
  vector<double> aa(nnz);
  hip_check(ffhip_mat_download(M, rp.data(), ja.data(), aa.data()), "ffhip: download");
  for (int64_t i = 0; i <= n; ++i) ia[i] = (int)rp[i];
  ffhip_mat_free(M);
  ffhip_p2v_free(sp);
  ffhip_mesh_free(m);
  MatriceMorse<double> *pH = new MatriceMorse<double>((int)n, (int)n, (int)nnz, 0);
  pH->set((int)n, (int)n, 0, (size_t)nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  A.A.master(pH);
  return true;
}

// the coefficient at the 14 Tet_5 points of every element, evaluated by the
// interpreter as Element_Op does (problem.cpp:6385-6395)
void eval_coef_qp(Stack stack, const FESpace3 &Uh, const C_F0 &coef, vector<double> &kq) {
  const Mesh3 &Th = Uh.Th;
  const GQuadratureFormular<R3> &FI = QuadratureFormular_Tet_5;
  kq.resize((size_t)Th.nt * FI.n);
  MeshPoint mp = *MeshPointStack(stack);
  for (int it = 0; it < Th.nt; ++it) {
    const FElement3 K(Uh[it]);
    for (int q = 0; q < FI.n; ++q) {
      const R3 Pt(FI[q]);
      MeshPointStack(stack)->set(K.T(Pt), Pt, K);
      kq[(size_t)it * FI.n + q] = GetAny<double>(coef.eval(stack));
    }
  }
  *MeshPointStack(stack) = mp;
}

// a boundary coefficient at the 7 T_5 points of every boundary element, in the
// order of the device rule: point q of boundary element k is
// (1 - s_q - t_q) B0 + s_q B1 + t_q B2 for its vertices B in Th.be(k) order,
// located in the adjacent tetrahedron (Mesh3::BoundaryElement) and evaluated
// with the label and the outward normal set as Element_Op does on a face
// (problem.cpp:6515-6528); zero for boundary elements whose label is not listed
void eval_coef_bqp(Stack stack, const FESpace3 &Uh, const Recognised::Boundary &t,
                   vector<double> &gq) {
  const Mesh3 &Th = Uh.Th;
  const GQuadratureFormular<R2> &FI = QuadratureFormular_T_5;
  gq.assign((size_t)Th.nbe * FI.n, 0.);
  MeshPoint mp = *MeshPointStack(stack);
  for (int k = 0; k < Th.nbe; ++k) {
    const int lab = Th.be(k).lab;
    if (!t.all && t.labels.find(lab) == t.labels.end()) continue;
    int ie = 0;
    const int it = Th.BoundaryElement(k, ie);
    const FElement3 K(Uh[it]);
    const Tet &T = Th[it];
    R3 NN = T.N(ie);
    NN /= NN.norme();
    int loc[3] = {0, 0, 0};
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 4; ++i)
        if (Th(T[i]) == Th(Th.be(k)[j])) loc[j] = i;
    for (int q = 0; q < FI.n; ++q) {
      const double s = FI[q].x, tt = FI[q].y;
      const R3 Pt = R3::KHat[loc[0]] * (1 - s - tt) + R3::KHat[loc[1]] * s + R3::KHat[loc[2]] * tt;
      MeshPointStack(stack)->set(T(Pt), Pt, K, lab, NN, ie);
      gq[(size_t)k * FI.n + q] = GetAny<double>(t.coef.eval(stack));
    }
  }
  *MeshPointStack(stack) = mp;
}

void assemble_on_gpu(Stack stack, const FESpace3 &Uh, const Recognised &r, double tgv, int sym,
                     Matrice_Creuse<double> &A) {
  const Mesh3 &Th = Uh.Th;
  vector<double> xyz(3 * (size_t)Th.nv);
  vector<int32_t> tets(4 * (size_t)Th.nt), btri(3 * (size_t)Th.nbe), blab(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    xyz[3 * i] = Th(i).x; xyz[3 * i + 1] = Th(i).y; xyz[3 * i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) tets[4 * k + j] = Th(Th[k][j]);  // P1 dof = vertex id
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) btri[3 * k + j] = Th(Th.be(k)[j]);
    blab[k] = Th.be(k).lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  ffhip_mat M = nullptr;
  hip_check(ffhip_mesh3_upload(ctx, Th.nv, Th.nt, Th.nbe, xyz.data(), tets.data(), btri.data(),
                               blab.data(), &m), "ffhip: mesh upload");
  hip_check(ffhip_pattern_p1(ctx, m, &M), "ffhip: pattern");
  {
    vector<double> kq, mq;
    if (!r.coef_const) eval_coef_qp(stack, Uh, r.coef, kq);
    if (r.has_mass && !r.mcoef_const) eval_coef_qp(stack, Uh, r.mcoef, mq);
    const int kind = r.coef_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP;
    const int mkind = !r.has_mass ? FFHIP_COEF_NONE : (r.mcoef_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP);
    hip_check(ffhip_assemble_p1_form(ctx, m, M, kind, r.c0, kq.empty() ? nullptr : kq.data(), mkind,
                                     r.m0, mq.empty() ? nullptr : mq.data(), 0),
              "ffhip: assembly");
  }
  for (const Recognised::Boundary &t : r.bnd) {  // in the order of the varf's terms
    vector<int32_t> labs;
    if (t.all) {
      set<int> present;
      for (int k = 0; k < Th.nbe; ++k) present.insert(Th.be(k).lab);
      labs.assign(present.begin(), present.end());
    } else {
      labs.assign(t.labels.begin(), t.labels.end());
    }
    if (labs.empty()) continue;
    vector<double> gq;
    if (!t.is_const) eval_coef_bqp(stack, Uh, t, gq);
    hip_check(ffhip_assemble_p1_boundary(ctx, m, M, labs.data(), (int)labs.size(),
                                         t.is_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP, t.c0,
                                         gq.empty() ? nullptr : gq.data(), 0),
              "ffhip: boundary term");
  }
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc(ctx, m, M, labs.data(), (int)labs.size(), tgv, nullptr), "ffhip: on()");
  }
  if (sym) {  // sym=1: the lower triangle, HashMatrix::half (HashMatrix.cpp:1312-1318)
    ffhip_mat L = nullptr;
    hip_check(ffhip_mat_lower(M, &L), "ffhip: sym=1 lower triangle");
    ffhip_mat_free(M);
    M = L;
  }
  int64_t n = 0, nnz = 0;
  ffhip_mat_info(M, &n, &nnz);
  vector<int64_t> rp(n + 1);
  vector<int> ia(n + 1), ja(nnz);
  vector<double> aa(nnz);
  hip_check(ffhip_mat_download(M, rp.data(), ja.data(), aa.data()), "ffhip: download");
  for (int64_t i = 0; i <= n; ++i) ia[i] = (int)rp[i];
  ffhip_mat_free(M);
  ffhip_mesh_free(m);
  // the result is an ordinary HashMatrix (HashMatrix::set with CSR input,
  // HashMatrix.cpp:697-728) so every later FreeFEM operation works on it
  const int half = sym ? 1 : 0;
  MatriceMorse<double> *pH = new MatriceMorse<double>((int)n, (int)n, (int)nnz, half);
  pH->set((int)n, (int)n, half, (size_t)nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  A.A.master(pH);
}

// Vh(Th,P2): true when built on the GPU; false (nothing changed) when FreeFEM's
// dof numbering of Uh is not the first-appearance node numbering of the device
bool assemble_p2_on_gpu(Stack stack, const FESpace3 &Uh, const Recognised &r, double tgv,
                        Matrice_Creuse<double> &A) {
  const Mesh3 &Th = Uh.Th;
  vector<double> xyz(3 * (size_t)Th.nv);
  vector<int32_t> tets(4 * (size_t)Th.nt), btri(3 * (size_t)Th.nbe), blab(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    xyz[3 * i] = Th(i).x; xyz[3 * i + 1] = Th(i).y; xyz[3 * i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) tets[4 * k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) btri[3 * k + j] = Th(Th.be(k)[j]);
    blab[k] = Th.be(k).lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  ffhip_p2v sp = nullptr;
  hip_check(ffhip_mesh3_upload(ctx, Th.nv, Th.nt, Th.nbe, xyz.data(), tets.data(), btri.data(),
                               blab.data(), &m), "ffhip: mesh upload");
  hip_check(ffhip_p2v_create(ctx, m, &sp), "ffhip: P2 numbering");
  vector<int32_t> en(10 * (size_t)Th.nt);
  hip_check(ffhip_p2v_download(sp, en.data()), "ffhip: P2 numbering");
  bool same = true;
  for (int k = 0; k < Th.nt && same; ++k) {
    const FElement3 K(Uh[k]);
    for (int df = 0; df < 10 && same; ++df) same = K(df) == en[10 * (size_t)k + df];
  }
  if (!same) {
    ffhip_p2v_free(sp);
    ffhip_mesh_free(m);
    return false;
  }
  ffhip_mat M = nullptr;
  hip_check(ffhip_pattern_p2(ctx, sp, &M), "ffhip: P2 pattern");
  if (r.coef_const) {
    hip_check(ffhip_assemble_p2(ctx, sp, M, FFHIP_COEF_CONST, r.c0, nullptr, 0), "ffhip: P2 assembly");
  } else {
    vector<double> kq;
    eval_coef_qp(stack, Uh, r.coef, kq);
    hip_check(ffhip_assemble_p2(ctx, sp, M, FFHIP_COEF_QP, 0., kq.data(), 0), "ffhip: P2 assembly");
  }
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc_p2(ctx, sp, M, labs.data(), (int)labs.size(), tgv, nullptr), "ffhip: on()");
  }
  int64_t n = 0, nnz = 0;
  ffhip_mat_info(M, &n, &nnz);
  vector<int64_t> rp(n + 1);
  vector<int> ia(n + 1), ja(nnz);
  vector<double> aa(nnz);
  hip_check(ffhip_mat_download(M, rp.data(), ja.data(), aa.data()), "ffhip: download");
  for (int64_t i = 0; i <= n; ++i) ia[i] = (int)rp[i];
  ffhip_mat_free(M);
  ffhip_p2v_free(sp);
  ffhip_mesh_free(m);
  MatriceMorse<double> *pH = new MatriceMorse<double>((int)n, (int)n, (int)nnz, 0);
  pH->set((int)n, (int)n, 0, (size_t)nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  A.A.master(pH);
  return true;
}

class OpHipMatrixP1 : public OneOperator {
 public:
  int init;
  class Op : public E_F0mps {
   public:
    CFB3 *b;
    Expression a;
    int init;
    Op(Expression aa, Expression bb, int ii)
        : b(new CFB3(*dynamic_cast<const CFB3 *>(bb))), a(aa), init(ii) {
      ffassert(b && b->nargs);
    }
    operator aType() const { return atype<Matrice_Creuse<double> *>(); }
    AnyType operator()(Stack stack) const {
      pfes3 *pUh = GetAny<pfes3 *>((*b->euh)(stack));
      pfes3 *pVh = GetAny<pfes3 *>((*b->evh)(stack));
      const FESpace3 *PUh = (FESpace3 *)**pUh;
      const FESpace3 *PVh = (FESpace3 *)**pVh;
      const bool square = (void *)PUh == (void *)PVh || (PUh && PVh && PUh->NbOfDF == PVh->NbOfDF);
      const bool VF = isVF(b->largs);
      Data_Sparse_Solver ds;
      ds.factorize = 0;
      ds.initmat = true;
      const int np = OpCall_FormBilinear_np::n_name_param - NB_NAME_PARM_HMAT;
      SetEnd_Data_Sparse_Solver<double>(stack, ds, b->nargs, np);
      WhereStackOfPtr2Free(stack) = new StackOfPtr2Free(stack);
      Matrice_Creuse<double> &A(*GetAny<Matrice_Creuse<double> *>((*a)(stack)));
      if (init) A.init();
      if (!PUh || !PVh) return SetAny<Matrice_Creuse<double> *>(&A);
      Recognised r;
      RecognisedP2 r2;
      bool done = false;
      if (ds.tgv > 0 && recognise(stack, *PUh, *PVh, b->largs, r) && (ds.sym == 0 || r.order == 1)) {
        A.A = 0;
        A.Uh = *PUh;
        A.Vh = *PVh;
        if (r.order == 1) {  // the recognised P1 forms are symmetric: sym=1 keeps the lower half
          assemble_on_gpu(stack, *PUh, r, ds.tgv, ds.sym, A);
          done = true;
        } else {
          done = assemble_p2_on_gpu(stack, *PUh, r, ds.tgv, A);
        }
      } else if (ds.sym == 0 && ds.tgv > 0 && recognise_p2v(stack, *PUh, *PVh, b->largs, r2)) {
        A.A = 0;
        A.Uh = *PUh;
        A.Vh = *PVh;
        done = assemble_p2v_on_gpu(*PUh, r2, ds.tgv, A);
      }
      if (!done) {
        creationBlockOfMatrixToBilinearForm<double, Mesh3, FESpace3, FESpace3>(PUh, PVh, ds.sym, ds.tgv,
                                                                               b->largs, stack, A);
      }
      A.pHM()->half = ds.sym;
      if (square) SetSolver(stack, VF, *A.A, ds);  // as OpMatrixtoBilinearForm (problem.hpp:1654-1656)
      return SetAny<Matrice_Creuse<double> *>(&A);
    }
  };
  E_F0 *code(const basicAC_F0 &args) const {
    return new Op(to<Matrice_Creuse<double> *>(args[0]), args[1], init);
  }
  OpHipMatrixP1(int ii)
      : OneOperator(atype<Matrice_Creuse<double> *>(), atype<Matrice_Creuse<double> *>(),
                    atype<const CFB3 *>()),
        init(ii) {
    pref = 10;  // preferred over OpMatrixtoBilinearForm (pref 0)
  }
};

// ---------------------------------------------------------------- 2D P1
typedef Call_FormBilinear<v_fes, v_fes> CFB2;

// int2d(Th)(c*(dx(u)*dx(v)+dy(u)*dy(v)) + m*u*v) [+ int1d(Th, labels)(alpha*u*v)]
// [+ on(labels, u=g)] on a 2D P1 space, the default T_5 / Gauss-Legendre 3 rules
bool recognise2(Stack stack, const FESpace &Uh, const FESpace &Vh, const list<C_F0> &largs,
                Recognised &r) {
  if (&Uh != &Vh || Uh.N != 1 || Uh.TFE.N() != 1 || Uh.TFE[0] != &P1Lagrange) return false;
  if (Uh.NbOfDF != Uh.Th.nv) return false;
  r.order = 1;
  FreeFempp::TypeVarForm<double> *tvf = FreeFempp::TypeVarForm<double>::Global;
  for (list<C_F0>::const_iterator ii = largs.begin(); ii != largs.end(); ++ii) {
    const aType t = ii->left();
    const Expression e = ii->LeftValue();
    if (t == tvf->tFL) continue;
    if (t == tvf->tBC) {
      const BC_set *bc = dynamic_cast<const BC_set *>(e);
      if (!bc) return false;
      for (size_t k = 0; k < bc->bc.size(); ++k)
        if (bc->bc[k].first != 0) return false;
      set<long> labs;
      Fem2D::Expandsetoflab(stack, *bc, labs);
      r.labels.insert(labs.begin(), labs.end());
      continue;
    }
    if (t != tvf->tFB) return false;
    const FormBilinear *bf = dynamic_cast<const FormBilinear *>(e);
    if (!bf || bf->di->islevelset()) return false;
    if (GetAny<pmesh>((*bf->di->Th)(stack)) != &Uh.Th) return false;
    const BilinearOperator &Op = *bf->b;
    if (bf->di->kind == CDomainOfIntegration::int1d) {  // Robin term
      if (bf->di->FIE(stack).n != QF_GaussLegendre3.n || Op.v.size() != 1) return false;
      const BilinearOperator::K &ll = *Op.v.begin();
      if (ll.first.first.first || ll.first.second.first || ll.first.first.second != op_id ||
          ll.first.second.second != op_id)
        return false;
      Recognised::Boundary b;
      b.coef = ll.second;
      if (b.coef.EvaluableWithOutStack()) {
        b.is_const = true;
        b.c0 = GetAny<double>(b.coef.eval(stack));
      }
      Fem2D::Expandsetoflab(stack, *bf->di, b.labels, b.all);
      r.bnd.push_back(b);
      continue;
    }
    if (r.bf || bf->di->kind != CDomainOfIntegration::int2d || !bf->di->what.empty()) return false;
    if (bf->di->FIT(stack).n != QuadratureFormular_T_5.n) return false;
    bool seen[3] = {false, false, false};
    int ngrad = 0;
    for (BilinearOperator::const_iterator l = Op.v.begin(); l != Op.v.end(); ++l) {
      const BilinearOperator::K &ll = *l;
      const int jcomp = ll.first.first.first, jop = ll.first.first.second;
      const int icomp = ll.first.second.first, iop = ll.first.second.second;
      if (jcomp || icomp || jop != iop || jop < op_id || jop > op_dy || seen[jop]) return false;
      seen[jop] = true;
      if (jop == op_id) {
        r.has_mass = true;
        r.mcoef = ll.second;
      } else if (ngrad++ == 0) {
        r.coef = ll.second;
      } else if (ll.second.LeftValue() != r.coef.LeftValue()) {
        return false;
      }
    }
    if (ngrad != 0 && ngrad != 2) return false;
    r.has_grad = ngrad == 2;
    if (!r.has_grad && !r.has_mass) return false;
    r.bf = bf;
  }
  if (!r.bf) return false;
  if (r.has_grad && r.coef.EvaluableWithOutStack()) {
    r.coef_const = true;
    r.c0 = GetAny<double>(r.coef.eval(stack));
  }
  if (!r.has_grad) {
    r.coef_const = true;
    r.c0 = 0.;
  }
  if (r.has_mass && r.mcoef.EvaluableWithOutStack()) {
    r.mcoef_const = true;
    r.m0 = GetAny<double>(r.mcoef.eval(stack));
  }
  return true;
}

// a coefficient at the 7 T_5 points of every triangle (Element_Op 2D)
void eval_coef_qp2(Stack stack, const FESpace &Uh, const C_F0 &coef, vector<double> &kq) {
  const Mesh &Th = Uh.Th;
  const QuadratureFormular &FI = QuadratureFormular_T_5;
  kq.resize((size_t)Th.nt * FI.n);
  MeshPoint mp = *MeshPointStack(stack);
  for (int it = 0; it < Th.nt; ++it) {
    const FElement K(Uh[it]);
    for (int q = 0; q < FI.n; ++q) {
      const R2 Pt(FI[q]);
      MeshPointStack(stack)->set(K.T(Pt), Pt, K);
      kq[(size_t)it * FI.n + q] = GetAny<double>(coef.eval(stack));
    }
  }
  *MeshPointStack(stack) = mp;
}

// a boundary coefficient at the 3 Gauss-Legendre points of every boundary
// edge in the device rule's order, x_q = (1 - s_q) B0 + s_q B1 for its
// vertices in Th.bedges order, located in the adjacent triangle with the
// label and the outward normal set as Element_Op does on an edge
// (problem.cpp:6226-6237); zero for edges whose label is not listed
void eval_coef_bqp2(Stack stack, const FESpace &Uh, const Recognised::Boundary &t,
                    vector<double> &gq) {
  const Mesh &Th = Uh.Th;
  const QuadratureFormular1d &FI = QF_GaussLegendre3;
  gq.assign((size_t)Th.neb * FI.n, 0.);
  // the device rule's points, in its order (s = 1/2, (1 - sqrt(3/5))/2, (1 + sqrt(3/5))/2)
  const double g1 = (1 - sqrt(3. / 5.)) / 2;
  const double sq[3] = {0.5, g1, 1 - g1};
  MeshPoint mp = *MeshPointStack(stack);
  for (int k = 0; k < Th.neb; ++k) {
    const int lab = Th.bedges[k].lab;
    if (!t.all && t.labels.find(lab) == t.labels.end()) continue;
    int ie = 0;
    const int it = Th.BoundaryElement(k, ie);
    const FElement K(Uh[it]);
    const Triangle &T = Th[it];
    const R2 E = T.Edge(ie);
    const double le = sqrt((E, E));
    const R2 NN(E.y / le, -E.x / le);
    int loc[2] = {0, 0};
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 3; ++i)
        if (&T[i] == &Th.bedges[k][j]) loc[j] = i;
    for (int q = 0; q < FI.n; ++q) {
      const R2 Pt = TriangleHat[loc[0]] * (1 - sq[q]) + TriangleHat[loc[1]] * sq[q];
      MeshPointStack(stack)->set(T(Pt), Pt, K, lab, NN, ie);
      gq[(size_t)k * FI.n + q] = GetAny<double>(t.coef.eval(stack));
    }
  }
  *MeshPointStack(stack) = mp;
}

void assemble2_on_gpu(Stack stack, const FESpace &Uh, const Recognised &r, double tgv, int sym,
                      Matrice_Creuse<double> &A) {
  const Mesh &Th = Uh.Th;
  vector<double> xy(2 * (size_t)Th.nv);
  vector<int32_t> tris(3 * (size_t)Th.nt), bed(2 * (size_t)Th.neb), blab(Th.neb);
  for (int i = 0; i < Th.nv; ++i) {
    xy[2 * i] = Th(i).x; xy[2 * i + 1] = Th(i).y;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 3; ++j) tris[3 * k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.neb; ++k) {
    for (int j = 0; j < 2; ++j) bed[2 * k + j] = Th(Th.bedges[k][j]);
    blab[k] = Th.bedges[k].lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  ffhip_mat M = nullptr;
  hip_check(ffhip_mesh2_upload(ctx, Th.nv, Th.nt, Th.neb, xy.data(), tris.data(), bed.data(),
                               blab.data(), &m), "ffhip: 2D mesh upload");
  hip_check(ffhip_pattern_p1(ctx, m, &M), "ffhip: pattern");
  {
    vector<double> kq, mq;
    if (!r.coef_const) eval_coef_qp2(stack, Uh, r.coef, kq);
    if (r.has_mass && !r.mcoef_const) eval_coef_qp2(stack, Uh, r.mcoef, mq);
    const int kind = r.coef_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP;
    const int mkind = !r.has_mass ? FFHIP_COEF_NONE : (r.mcoef_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP);
    hip_check(ffhip_assemble_p1_form(ctx, m, M, kind, r.c0, kq.empty() ? nullptr : kq.data(), mkind,
                                     r.m0, mq.empty() ? nullptr : mq.data(), 0),
              "ffhip: assembly");
  }
  for (const Recognised::Boundary &t : r.bnd) {
    vector<int32_t> labs;
    if (t.all) {
      set<int> present;
      for (int k = 0; k < Th.neb; ++k) present.insert(Th.bedges[k].lab);
      labs.assign(present.begin(), present.end());
    } else {
      labs.assign(t.labels.begin(), t.labels.end());
    }
    if (labs.empty()) continue;
    vector<double> gq;
    if (!t.is_const) eval_coef_bqp2(stack, Uh, t, gq);
    hip_check(ffhip_assemble_p1_boundary(ctx, m, M, labs.data(), (int)labs.size(),
                                         t.is_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP, t.c0,
                                         gq.empty() ? nullptr : gq.data(), 0),
              "ffhip: boundary term");
  }
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc(ctx, m, M, labs.data(), (int)labs.size(), tgv, nullptr), "ffhip: on()");
  }
  if (sym) {
    ffhip_mat L = nullptr;
    hip_check(ffhip_mat_lower(M, &L), "ffhip: sym=1 lower triangle");
    ffhip_mat_free(M);
    M = L;
  }
  int64_t n = 0, nnz = 0;
  ffhip_mat_info(M, &n, &nnz);
  vector<int64_t> rp(n + 1);
  vector<int> ia(n + 1), ja(nnz);
  vector<double> aa(nnz);
  hip_check(ffhip_mat_download(M, rp.data(), ja.data(), aa.data()), "ffhip: download");
  for (int64_t i = 0; i <= n; ++i) ia[i] = (int)rp[i];
  ffhip_mat_free(M);
  ffhip_mesh_free(m);
  const int half = sym ? 1 : 0;
  MatriceMorse<double> *pH = new MatriceMorse<double>((int)n, (int)n, (int)nnz, half);
  pH->set((int)n, (int)n, half, (size_t)nnz, ia.data(), ja.data(), aa.data(), 0, 1);
  A.A.master(pH);
}

// "=" / "<-" for (matrix, varf(Vh, Vh)) on 2D P1 — preferred over
// OpMatrixtoBilinearForm<double, Mesh, v_fes, v_fes> (problem.hpp:1611-1660);
// other forms take its code path unchanged
class OpHipMatrix2D : public OneOperator {
 public:
  int init;
  class Op : public E_F0mps {
   public:
    CFB2 *b;
    Expression a;
    int init;
    Op(Expression aa, Expression bb, int ii)
        : b(new CFB2(*dynamic_cast<const CFB2 *>(bb))), a(aa), init(ii) {
      ffassert(b && b->nargs);
    }
    operator aType() const { return atype<Matrice_Creuse<double> *>(); }
    AnyType operator()(Stack stack) const {
      pfes *pUh = GetAny<pfes *>((*b->euh)(stack));
      pfes *pVh = GetAny<pfes *>((*b->evh)(stack));
      const FESpace *PUh = (FESpace *)**pUh;
      const FESpace *PVh = (FESpace *)**pVh;
      const bool square = (void *)PUh == (void *)PVh || (PUh && PVh && PUh->NbOfDF == PVh->NbOfDF);
      const bool VF = isVF(b->largs);
      Data_Sparse_Solver ds;
      ds.factorize = 0;
      ds.initmat = true;
      const int np = OpCall_FormBilinear_np::n_name_param - NB_NAME_PARM_HMAT;
      SetEnd_Data_Sparse_Solver<double>(stack, ds, b->nargs, np);
      WhereStackOfPtr2Free(stack) = new StackOfPtr2Free(stack);
      Matrice_Creuse<double> &A(*GetAny<Matrice_Creuse<double> *>((*a)(stack)));
      if (init) A.init();
      if (!PUh || !PVh) return SetAny<Matrice_Creuse<double> *>(&A);
      Recognised r;
      if (ds.tgv > 0 && recognise2(stack, *PUh, *PVh, b->largs, r)) {
        A.A = 0;
        A.Uh = *PUh;
        A.Vh = *PVh;
        assemble2_on_gpu(stack, *PUh, r, ds.tgv, ds.sym, A);
      } else {
        creationBlockOfMatrixToBilinearForm<double, Mesh, FESpace, FESpace>(PUh, PVh, ds.sym, ds.tgv,
                                                                            b->largs, stack, A);
      }
      A.pHM()->half = ds.sym;
      if (square) SetSolver(stack, VF, *A.A, ds);
      return SetAny<Matrice_Creuse<double> *>(&A);
    }
  };
  E_F0 *code(const basicAC_F0 &args) const {
    return new Op(to<Matrice_Creuse<double> *>(args[0]), args[1], init);
  }
  OpHipMatrix2D(int ii)
      : OneOperator(atype<Matrice_Creuse<double> *>(), atype<Matrice_Creuse<double> *>(),
                    atype<const CFB2 *>()),
        init(ii) {
    pref = 10;  // preferred over OpMatrixtoBilinearForm (pref 0)
  }
};

// ---------------------------------------------------------------- linear forms
typedef Call_FormLinear<v_fes3> CFL3;

// b = va(0, Vh) on P1: int3d(Th)(f*v) [+ int2d(Th, labels)(g*v) ...]
// [+ on(labels, u=g) ...]; bilinear terms do not enter the vector
// (AssembleVarForm with no matrix, problem.hpp:1385-1386)
struct RecognisedRhs {
  bool has_f = false, f_const = false;
  C_F0 f;
  double f0 = 0.;
  vector<Recognised::Boundary> bnd;
  vector<const BC_set *> bcs;  // in the order of the form
  set<long> dlabels;
};

bool recognise_rhs(Stack stack, const FESpace3 &Vh, const list<C_F0> &largs, RecognisedRhs &r) {
  if (Vh.N != 1 || Vh.TFE.N() != 1 || Vh.TFE[0] != &DataFE<Mesh3>::P1 || Vh.NbOfDF != Vh.Th.nv)
    return false;
  FreeFempp::TypeVarForm<double> *tvf = FreeFempp::TypeVarForm<double>::Global;
  for (list<C_F0>::const_iterator ii = largs.begin(); ii != largs.end(); ++ii) {
    const aType t = ii->left();
    const Expression e = ii->LeftValue();
    if (t == tvf->tFB) continue;
    if (t == tvf->tBC) {
      const BC_set *bc = dynamic_cast<const BC_set *>(e);
      if (!bc || bc->complextype || bc->bc.size() != 1 || bc->bc[0].first != 0) return false;
      set<long> labs;
      Fem2D::Expandsetoflab(stack, *bc, labs);
      r.dlabels.insert(labs.begin(), labs.end());
      r.bcs.push_back(bc);
      continue;
    }
    if (t != tvf->tFL) return false;
    const FormLinear *fl = dynamic_cast<const FormLinear *>(e);
    if (!fl || fl->di->islevelset()) return false;
    if (GetAny<const Mesh3 *>((*fl->di->Th)(stack)) != &Vh.Th) return false;
    const Ftest &L = *fl->l;
    if (L.v.size() != 1) return false;
    const MDroit &op = L.v.begin()->first;
    if (op.first != 0 || op.second != op_id) return false;
    const C_F0 &coef = L.v.begin()->second;
    if (fl->di->kind == CDomainOfIntegration::int3d) {
      if (r.has_f || !fl->di->what.empty()) return false;
      if (fl->di->FIV(stack).n != QuadratureFormular_Tet_5.n) return false;
      r.has_f = true;
      r.f = coef;
      if (coef.EvaluableWithOutStack()) {
        r.f_const = true;
        r.f0 = GetAny<double>(coef.eval(stack));
      }
    } else if (fl->di->kind == CDomainOfIntegration::int2d) {
      if (fl->di->FIT(stack).n != QuadratureFormular_T_5.n) return false;
      Recognised::Boundary b;
      b.coef = coef;
      if (coef.EvaluableWithOutStack()) {
        b.is_const = true;
        b.c0 = GetAny<double>(coef.eval(stack));
      }
      Fem2D::Expandsetoflab(stack, *fl->di, b.labels, b.all);
      r.bnd.push_back(b);
    } else {
      return false;
    }
  }
  return true;
}

// on(labels, u=g): g at the vertices of the labelled boundary triangles,
// evaluated as AssembleBC does for P1 (problem.cpp:10143-10176: the vertex,
// the label and the outward normal of the face set), later on() terms
// overwriting earlier ones
void eval_dirichlet(Stack stack, const FESpace3 &Vh, const RecognisedRhs &r, vector<double> &gv) {
  const Mesh3 &Th = Vh.Th;
  gv.assign(Th.nv, 0.);
  MeshPoint mp = *MeshPointStack(stack);
  for (const BC_set *bc : r.bcs) {
    set<long> labs;
    Fem2D::Expandsetoflab(stack, *bc, labs);
    for (int k = 0; k < Th.nbe; ++k) {
      const int lab = Th.be(k).lab;
      if (labs.find(lab) == labs.end()) continue;
      int ie = 0;
      const int it = Th.BoundaryElement(k, ie);
      const FElement3 K(Vh[it]);
      const Tet &T = Th[it];
      R3 NN = T.N(ie);
      NN /= NN.norme();
      for (int i = 0; i < 4; ++i) {
        if (i == ie) continue;  // the vertex opposite the face
        const R3 PtHat = R3::KHat[i];
        MeshPointStack(stack)->set(T(PtHat), PtHat, K, lab, NN, ie);
        gv[Th(T[i])] = GetAny<double>((*bc->bc[0].second)(stack));
      }
    }
  }
  *MeshPointStack(stack) = mp;
}

void rhs_on_gpu(Stack stack, const FESpace3 &Vh, const RecognisedRhs &r, double tgv, KN_<double> &b) {
  const Mesh3 &Th = Vh.Th;
  vector<double> xyz(3 * (size_t)Th.nv);
  vector<int32_t> tets(4 * (size_t)Th.nt), btri(3 * (size_t)Th.nbe), blab(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    xyz[3 * i] = Th(i).x; xyz[3 * i + 1] = Th(i).y; xyz[3 * i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) tets[4 * k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) btri[3 * k + j] = Th(Th.be(k)[j]);
    blab[k] = Th.be(k).lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  hip_check(ffhip_mesh3_upload(ctx, Th.nv, Th.nt, Th.nbe, xyz.data(), tets.data(), btri.data(),
                               blab.data(), &m), "ffhip: mesh upload");
  vector<double> fq, gv, out(Th.nv);
  if (r.has_f && !r.f_const) eval_coef_qp(stack, Vh, r.f, fq);
  if (!r.bcs.empty()) eval_dirichlet(stack, Vh, r, gv);
  vector<int32_t> dl(r.dlabels.begin(), r.dlabels.end());
  const int kind = !r.has_f ? FFHIP_COEF_CONST : (r.f_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP);
  const double f0 = r.has_f ? r.f0 : 0.;
  hip_check(ffhip_rhs_p1_form(ctx, m, kind, f0, fq.empty() ? nullptr : fq.data(), dl.data(),
                              (int)dl.size(), tgv, 0., gv.empty() ? nullptr : gv.data(), out.data(), 0),
            "ffhip: linear form");
  for (const Recognised::Boundary &t : r.bnd) {
    vector<int32_t> labs;
    if (t.all) {
      set<int> present;
      for (int k = 0; k < Th.nbe; ++k) present.insert(Th.be(k).lab);
      labs.assign(present.begin(), present.end());
    } else {
      labs.assign(t.labels.begin(), t.labels.end());
    }
    if (labs.empty()) continue;
    vector<double> gq;
    if (!t.is_const) eval_coef_bqp(stack, Vh, t, gq);
    hip_check(ffhip_rhs_p1_boundary(ctx, m, labs.data(), (int)labs.size(),
                                    t.is_const ? FFHIP_COEF_CONST : FFHIP_COEF_QP, t.c0,
                                    gq.empty() ? nullptr : gq.data(), dl.data(), (int)dl.size(),
                                    out.data(), 0),
              "ffhip: boundary linear form");
  }
  ffhip_mesh_free(m);
  for (int i = 0; i < Th.nv; ++i) b[i] = out[i];
}

// b = va(0, Vh) on Vh(Th,[P2,P2,P2]) or Vh(Th,P2): int3d(Th)(f_c*v_c) terms with
// mesh-independent coefficients and on(labels, u_c=g_c) with mesh-independent
// values, the same g for every label (ffhip_rhs_p2v / ffhip_rhs_p2)
struct RecognisedRhsP2 {
  int ncomp = 0;
  double f[3] = {0., 0., 0.}, g[3] = {0., 0., 0.};
  bool gset = false;
  set<long> labels;
};

bool recognise_rhs_p2(Stack stack, const FESpace3 &Vh, const list<C_F0> &largs, RecognisedRhsP2 &r) {
  if (is_p2v(Vh)) r.ncomp = 3;
  else if (Vh.N == 1 && Vh.TFE.N() == 1 && Vh.TFE[0] == &DataFE<Mesh3>::P2) r.ncomp = 1;
  else return false;
  FreeFempp::TypeVarForm<double> *tvf = FreeFempp::TypeVarForm<double>::Global;
  for (list<C_F0>::const_iterator ii = largs.begin(); ii != largs.end(); ++ii) {
    const aType t = ii->left();
    const Expression e = ii->LeftValue();
    if (t == tvf->tFB) continue;
    if (t == tvf->tBC) {
      const BC_set *bc = dynamic_cast<const BC_set *>(e);
      if (!bc || bc->complextype || (int)bc->bc.size() != r.ncomp) return false;
      double g[3] = {0., 0., 0.};
      for (int k = 0; k < r.ncomp; ++k) {
        if (bc->bc[k].first != k || !bc->bc[k].second->MeshIndependent()) return false;
        g[k] = GetAny<double>((*bc->bc[k].second)(stack));
      }
      if (r.gset && (g[0] != r.g[0] || g[1] != r.g[1] || g[2] != r.g[2])) return false;
      for (int k = 0; k < 3; ++k) r.g[k] = g[k];
      r.gset = true;
      set<long> labs;
      Fem2D::Expandsetoflab(stack, *bc, labs);
      r.labels.insert(labs.begin(), labs.end());
      continue;
    }
    if (t != tvf->tFL) return false;
    const FormLinear *fl = dynamic_cast<const FormLinear *>(e);
    if (!fl || fl->di->islevelset() || fl->di->kind != CDomainOfIntegration::int3d ||
        !fl->di->what.empty())
      return false;
    if (GetAny<const Mesh3 *>((*fl->di->Th)(stack)) != &Vh.Th) return false;
    if (fl->di->FIV(stack).n != QuadratureFormular_Tet_5.n) return false;
    for (Ftest::const_iterator l = fl->l->v.begin(); l != fl->l->v.end(); ++l) {
      const MDroit &op = l->first;
      if (op.first < 0 || op.first >= r.ncomp || op.second != op_id) return false;
      if (!l->second.EvaluableWithOutStack()) return false;
      r.f[op.first] += GetAny<double>(l->second.eval(stack));
    }
  }
  return true;
}

// true when built on the GPU; false (b untouched) when FreeFEM's dof numbering
// is not the device's first-appearance node numbering
bool rhs_p2_on_gpu(const FESpace3 &Vh, const RecognisedRhsP2 &r, double tgv, KN_<double> &b) {
  const Mesh3 &Th = Vh.Th;
  vector<double> xyz(3 * (size_t)Th.nv);
  vector<int32_t> tets(4 * (size_t)Th.nt), btri(3 * (size_t)Th.nbe), blab(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    xyz[3 * i] = Th(i).x; xyz[3 * i + 1] = Th(i).y; xyz[3 * i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) tets[4 * k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) btri[3 * k + j] = Th(Th.be(k)[j]);
    blab[k] = Th.be(k).lab;
  }
  ffhip_ctx ctx = hip_ctx();
  ffhip_mesh m = nullptr;
  ffhip_p2v sp = nullptr;
  hip_check(ffhip_mesh3_upload(ctx, Th.nv, Th.nt, Th.nbe, xyz.data(), tets.data(), btri.data(),
                               blab.data(), &m), "ffhip: mesh upload");
  hip_check(ffhip_p2v_create(ctx, m, &sp), "ffhip: P2 numbering");
  vector<int32_t> en(10 * (size_t)Th.nt);
  hip_check(ffhip_p2v_download(sp, en.data()), "ffhip: P2 numbering");
  bool same = true;
  const int nd = 10 * r.ncomp;
  for (int k = 0; k < Th.nt && same; ++k) {
    const FElement3 K(Vh[k]);
    for (int df = 0; df < nd && same; ++df)  // local dofs comp-blocked (c*10 + i)
      same = K(df) == r.ncomp * en[10 * (size_t)k + df % 10] + df / 10;
  }
  if (same) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    vector<double> out(b.N());
    if (r.ncomp == 3)
      hip_check(ffhip_rhs_p2v(ctx, sp, r.f, labs.data(), (int)labs.size(), tgv, r.g, out.data(), 0),
                "ffhip: P2 linear form");
    else
      hip_check(ffhip_rhs_p2(ctx, sp, r.f[0], labs.data(), (int)labs.size(), tgv, r.g[0], out.data(), 0),
                "ffhip: P2 linear form");
    for (long i = 0; i < b.N(); ++i) b[i] = out[i];
  }
  ffhip_p2v_free(sp);
  ffhip_mesh_free(m);
  return same;
}

// "=" / "<-" for (real[int], varf(0, Vh)) on 3D P1 — preferred over
// OpArraytoLinearForm<double, Mesh3, v_fes3> (problem.hpp:853-927, registered
// at lgfem.cpp:6672 and 6688); other forms take FreeFEM's own path, with the
// same array semantics (problem.hpp:1351-1388)
class OpHipRhsP1 : public OneOperator {
 public:
  const bool isptr, init;
  class Op : public E_F0mps {
   public:
    CFL3 *l;
    Expression x;
    const bool isptr, init;
    Op(Expression xx, Expression ll, bool p, bool i)
        : l(new CFL3(*dynamic_cast<const CFL3 *>(ll))), x(xx), isptr(p), init(i) {
      ffassert(l);
    }
    operator aType() const { return atype<KN<double> *>(); }
    AnyType operator()(Stack stack) const {
      pfes3 &pp = *GetAny<pfes3 *>((*l->ppfes)(stack));
      FESpace3 *pVh = *pp;
      double tgv = ff_tgv;
      if (l->nargs[0]) tgv = GetAny<double>((*l->nargs[0])(stack));
      const long NbOfDF = pVh ? pVh->NbOfDF : 0;
      KN<double> *px = 0;
      if (isptr) {
        px = GetAny<KN<double> *>((*x)(stack));
        if (init) px->init(NbOfDF);
        if (px->N() != NbOfDF) px->resize(NbOfDF);
      }
      KN_<double> xx(px ? *(KN_<double> *)px : GetAny<KN_<double> >((*x)(stack)));
      if (NbOfDF) xx = 0.;
      RecognisedRhs r;
      RecognisedRhsP2 r2;
      bool done = false;
      if (pVh && tgv > 0 && xx.N() == NbOfDF && recognise_rhs(stack, *pVh, l->largs, r)) {
        rhs_on_gpu(stack, *pVh, r, tgv, xx);
        done = true;
      } else if (pVh && tgv > 0 && xx.N() == NbOfDF && recognise_rhs_p2(stack, *pVh, l->largs, r2)) {
        done = rhs_p2_on_gpu(*pVh, r2, tgv, xx);
      }
      if (!done && pVh &&
          AssembleVarForm<double, MatriceCreuse<double>, Mesh3, FESpace3, FESpace3>(
              stack, pVh->Th, *pVh, *pVh, false, 0, &xx, l->largs)) {
        AssembleBC<double, Mesh3, FESpace3, FESpace3>(stack, pVh->Th, *pVh, *pVh, false, 0, &xx, 0,
                                                      l->largs, tgv);
      }
      return SetAny<KN_<double> >(xx);
    }
  };
  E_F0 *code(const basicAC_F0 &args) const {
    if (isptr) return new Op(to<KN<double> *>(args[0]), args[1], isptr, init);
    return new Op(to<KN_<double> >(args[0]), args[1], isptr, init);
  }
  OpHipRhsP1(const basicForEachType *tt, bool p, bool i)
      : OneOperator(atype<KN_<double> >(), tt, atype<const CFL3 *>()), isptr(p), init(i) {
    pref = 10;  // preferred over OpArraytoLinearForm (pref 0)
  }
};

// ---------------------------------------------------------------- LinearCG
// LinearCG(A, x, b, ...) and AffineCG(A, x, ...) / LinearCG(A, x, ...) with a
// matrix as the operator: ConjuguedGradient2 (MatriceCreuse.hpp:180-240) on
// the GPU in place of LinearCG<R>::E_LCG (lgfem.cpp:705-758) calling back
//   func real[int] op(real[int] &u) { real[int] v = A*u; return v; }
// for every product.  Named parameters as LinearCG's (eps, nbiter, veps,
// verbosity); the preconditioner is named by jacobi=true (C g = D^-1 g, the
// diagonal of A) instead of a precon= function.  Returns LinearCG's value:
// 1 converged, 2 nothing to do, 0 not converged.
class HipLinearCG : public OneOperator {
 public:
  const int cas;  // 2: with a right-hand side, 1: without (AffineCG)
  class Op : public E_F0mps {
   public:
    static const int n_name_param = 5;
    static basicAC_F0::name_and_type name_param[];
    Expression nargs[n_name_param];
    Expression A, X, B;
    Op(const basicAC_F0 &args, int cas) : B(0) {
      args.SetNameParam(n_name_param, name_param, nargs);
      A = to<Matrice_Creuse<double> *>(args[0]);
      X = to<KN<double> *>(args[1]);
      if (cas == 2) B = to<KN<double> *>(args[2]);
    }
    operator aType() const { return atype<long>(); }
    AnyType operator()(Stack stack) const {
      Matrice_Creuse<double> &MA = *GetAny<Matrice_Creuse<double> *>((*A)(stack));
      KN<double> &x = *GetAny<KN<double> *>((*X)(stack));
      KN<double> *b = B ? GetAny<KN<double> *>((*B)(stack)) : 0;
      double eps = 1.0e-6, *veps = 0;
      long nbitermax = 100, verb = verbosity;
      bool jacobi = false;
      if (nargs[0]) eps = GetAny<double>((*nargs[0])(stack));
      if (nargs[1]) nbitermax = GetAny<long>((*nargs[1])(stack));
      if (nargs[2]) veps = GetAny<double *>((*nargs[2])(stack));
      if (nargs[3]) verb = Abs(GetAny<long>((*nargs[3])(stack)));
      if (nargs[4]) jacobi = GetAny<bool>((*nargs[4])(stack));
      if (veps) eps = *veps;
      HashMatrix<int, double> *H = MA.pHM();
      if (!H) ExecError("LinearCG: empty matrix");
      ffassert(H->n == H->m && H->n == x.N() && (!b || b->N() == x.N()));
      int *ia = 0, *ja = 0;
      double *a = 0;
      H->CSR(ia, ja, a);
      const int64_t n = H->n, nnz = ia[n];
      vector<int64_t> rp(ia, ia + n + 1);
      ffhip_mat M = nullptr;
      ffhip_cg cg = nullptr;
      hip_check(ffhip_mat_upload(hip_ctx(), n, nnz, rp.data(), ja, a, &M), "LinearCG: matrix upload");
      int rc = ffhip_cg_create(hip_ctx(), M, FFHIP_CG_FP64, &cg);
      if (rc != FFHIP_OK) { ffhip_mat_free(M); hip_check(rc, "LinearCG: solver setup"); }
      double eps_out = eps;
      ffhip_cg_stats st;
      rc = ffhip_cg_solve_affine(cg, b ? (double *)*b : nullptr, (double *)x, 0, eps, (int)nbitermax,
                                 jacobi ? FFHIP_CG2_JACOBI : FFHIP_CG2_IDENTITY, &eps_out, &st);
      ffhip_cg_free(cg);
      ffhip_mat_free(M);
      hip_check(rc, "CG2: Matrix is not defined (/0) or ro is NaN");
      if (veps) *veps = -eps_out;
      if (verb > 1 || (verb && st.converged != 1))
        cout << (st.converged ? "CG converges " : "CG doesn't converge: ") << st.iters
             << " ||g||^2 = " << st.gcg << " (" << st.ms_solve << " ms on the GPU)" << endl;
      return SetAny<long>(st.converged);
    }
  };
  E_F0 *code(const basicAC_F0 &args) const { return new Op(args, cas); }
  HipLinearCG()
      : OneOperator(atype<long>(), atype<Matrice_Creuse<double> *>(), atype<KN<double> *>(),
                    atype<KN<double> *>()),
        cas(2) {}
  explicit HipLinearCG(int)
      : OneOperator(atype<long>(), atype<Matrice_Creuse<double> *>(), atype<KN<double> *>()),
        cas(1) {}
};

basicAC_F0::name_and_type HipLinearCG::Op::name_param[] = {
    {"eps", &typeid(double)},    {"nbiter", &typeid(long)},     {"veps", &typeid(double *)},
    {"verbosity", &typeid(long)}, {"jacobi", &typeid(bool)}};

}  // namespace

static void Load_Init() {
  addsolver<SolverHIPCG<0> >("HIPCG", 50);
  addsolver<SolverHIPCG<1> >("HIPCGMIXED", 50);
  addsolver<SolverHIPGMRES>("HIPGMRES", 50);
  Global.Add("ffhipreplacecg", "(", new OneOperator1<long, long>(ffhip_replace_cg));
  TheOperators->Add("=", new OpHipMatrixP1(0));
  TheOperators->Add("<-", new OpHipMatrixP1(1));
  TheOperators->Add("=", new OpHipMatrix2D(0));
  TheOperators->Add("<-", new OpHipMatrix2D(1));
  TheOperators->Add("=", new OpHipRhsP1(atype<KN_<double> >(), false, false));
  TheOperators->Add("<-", new OpHipRhsP1(atype<KN<double> *>(), true, true));
  Global.Add("LinearCG", "(", new HipLinearCG());
  Global.Add("LinearCG", "(", new HipLinearCG(1));
  Global.Add("AffineCG", "(", new HipLinearCG(1));
  if (verbosity > 1) cout << " load: ffhip (" << ffhip_version() << ")" << endl;
}

LOADFUNC(Load_Init)
