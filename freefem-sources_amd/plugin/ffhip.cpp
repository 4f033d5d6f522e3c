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

#include <dlfcn.h>

#include <cstring>
#include <memory>
#include <set>
#include <vector>

#include "AFunction_ext.hpp"
#include "problem.hpp"

#include "ffhip.h"
#include "ffhip_glue.hpp"

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
typedef HashMatrix<int, double> HMat;
typedef ffhip_glue::SolverCore<HMat> Core;

// The device state an assembly operator hands to the solver SetSolver creates
// next for the same matrix (set just before SetSolver, claimed by the
// solver's constructor, freed by the operator if no ffhip solver took it).
struct Handover {
  const HMat *H = nullptr;
  ffhip_glue::Resident *r = nullptr;
};
Handover &handover() {
  static Handover h;
  return h;
}
ffhip_glue::Resident *claim_handover(const HMat &A) {
  Handover &h = handover();
  if (h.H != &A) return nullptr;
  ffhip_glue::Resident *r = h.r;
  h = Handover();
  return r;
}

// common part of the ffhip solvers: the device side (SolverCore) and the
// redo bookkeeping of UpdateState
class HipSolverBase : public VirtualSolver<int, double> {
 public:
  HMat *ptA;
  Core core;
  long cn, cs;
  HipSolverBase(HMat &A, int mode) : ptA(&A), core(&A, hip_ctx(), mode), cn(0), cs(0) {
    core.adopt(claim_handover(A));
  }
  // the matrix's redo flags (HashMatrix.hpp:278-279) as the reference's
  // sparse solvers read them; the device copy is refreshed before the next
  // solve when they are set (a script that changed A after the solver was made)
  void UpdateState() {
    const bool num = ptA->GetReDoNumerics(), sym = ptA->GetReDoSymbolic();
    core.note_changes(num, sym);
    if (num) cn++;
    if (sym) cs++;
    this->ChangeCodeState(ptA->n, cs, cn);
  }
  void fac_symbolic() { hip_check(core.symbolic(), "ffhip solver: matrix to the GPU"); }
  void fac_numeric() { hip_check(core.numeric(), "ffhip solver: matrix values"); }
  // the assembly's device state, for a re-assembly that keeps the pattern
  // (only when this solver is the matrix's sole owner)
  ffhip_glue::Resident *release_resident() { return this->count == 0 ? core.take_resident() : nullptr; }
};

// solver=HIPCG / HIPCGMIXED: SolverCG (VirtualSolverCG.hpp:112-192) on the GPU
template <int MODE>
class SolverHIPCG : public HipSolverBase {
 public:
  // 1 unsym, 2 herm, 4 sym, 8 pos, 32 seq: as SolverCG
  static const int orTypeSol = 1 | 2 | 4 | 8 | 32;
  int verb, itermax;
  double eps, tgv;
  double *veps;
  long *getnbiter;

  SolverHIPCG(HMat &A, const Data_Sparse_Solver &ds, Stack)
      : HipSolverBase(A, MODE ? FFHIP_CG_MIXED : FFHIP_CG_FP64), verb(ds.verb),
        itermax(ds.itmax > 0 ? ds.itmax : A.n), eps(ds.epsilon), tgv(ds.tgv), veps(ds.veps),
        getnbiter(ds.getnbiter) {
    ffassert(A.n == A.m);
    if (ds.precon) ExecError("HIPCG: precon= is not supported (the GPU solver applies Jacobi)");
    if (verb > 4)
      cout << " ## SolverHIPCG" << (MODE ? "MIXED " : " ") << A.n << " eps " << eps << " itmax "
           << itermax << (core.res ? " (device copy from the assembly)" : "") << endl;
  }

  // SolverCG::dosolver semantics (VirtualSolverCG.hpp:169-190): x holds the
  // initial guess, tgv rows start at b_i / tgv, veps receives the absolute
  // threshold ConjugueGradient leaves in eps (CG.cpp:226)
  void dosolver(double *x, double *b, int N, int trans) {
    hip_check(core.refresh(), "HIPCG: matrix to the GPU");
    int err = 0;
    if (getnbiter) *getnbiter = 0;
    for (int k = 0; k < N; ++k) {
      ffhip_cg_stats st;
      const int rc = ffhip_cg_solve(core.cg, b + (size_t)k * ptA->n, x + (size_t)k * ptA->n, 0, eps,
                                    itermax, tgv, trans, &st);
      if (rc == FFHIP_ERR_NOCONV) err++;
      else hip_check(rc, "HIPCG: solve");
      if (rc == FFHIP_OK && getnbiter) *getnbiter += st.iters;
      if (veps) *veps = ffhip_glue::cg_written_back_eps(eps, st.gcg0);
      if (verb > 1)
        cout << " HIPCG: " << (st.converged ? "converge after " : "no convergence after ") << st.iters
             << " g=" << st.gcg << " (" << st.ms_solve << " ms on the GPU)" << endl;
    }
    if (err) {
      cerr << "Error: ConjugueGradient do not converge nb end =" << err << endl;
      ffassert(0);
    }
  }
};

// solver=HIPGMRES: SolverGMRES (VirtualSolverCG.hpp:196-255) — restarted
// right-preconditioned GMRES (fgmres, CG.cpp:347-517) with the Jacobi
// diagonal, for non-symmetric matrices; restart = dimKrylov (NbSpace)
class SolverHIPGMRES : public HipSolverBase {
 public:
  // 1 unsym, 2 herm, 4 sym, 8 pos, 16 nopos, 32 seq: as SolverGMRES
  static const int orTypeSol = 1 | 2 | 4 | 8 | 16 | 32;
  int verb, itermax, restart;
  double eps, tgv;
  double *veps;
  long *getnbiter;

  SolverHIPGMRES(HMat &A, const Data_Sparse_Solver &ds, Stack)
      : HipSolverBase(A, FFHIP_CG_FP64), verb(ds.verb), itermax(ds.itmax > 0 ? ds.itmax : A.n),
        restart(ds.NbSpace), eps(ds.epsilon), tgv(ds.tgv), veps(ds.veps), getnbiter(ds.getnbiter) {
    ffassert(A.n == A.m);
    if (ds.precon) ExecError("HIPGMRES: precon= is not supported (the GPU solver applies Jacobi)");
  }

  // SolverGMRES::dosolver semantics (VirtualSolverCG.hpp:236-255): veps
  // receives eps unchanged, the iteration count is the last right-hand side's
  void dosolver(double *x, double *b, int N, int trans) {
    hip_check(core.refresh(), "HIPGMRES: matrix to the GPU");
    int err = 0;
    for (int k = 0; k < N; ++k) {
      ffhip_cg_stats st;
      const int rc = ffhip_gmres_solve(core.cg, b + (size_t)k * ptA->n, x + (size_t)k * ptA->n, 0,
                                       eps, itermax, restart, tgv, trans, &st);
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

// kinds of forms a Resident was built for (reuse needs the same kind)
enum { FORM_P1_3D = 1, FORM_P1_2D = 2, FORM_P2_3D = 3, FORM_P2V_3D = 4 };

typedef std::unique_ptr<ffhip_glue::Resident> ResidentPtr;

// the device mesh of Th (3D) into r
void upload_mesh3(const Mesh3 &Th, ffhip_glue::Resident &r) {
  ffhip_glue::MeshArrays a;
  ffhip_glue::mesh3_arrays(Th, a);
  hip_check(ffhip_glue::upload_mesh(hip_ctx(), 3, a, &r.mesh), "ffhip: mesh upload");
  r.Th = &Th;
}

// the device P2 numbering of r's mesh; false when FreeFEM's dof numbering of
// Uh is not the device's first-appearance numbering (ncomp = 1: Vh(Th,P2),
// 3: [P2,P2,P2], local dofs comp-blocked c*10 + i)
bool p2_numbering(const FESpace3 &Uh, int ncomp, ffhip_glue::Resident &r) {
  hip_check(ffhip_p2v_create(hip_ctx(), r.mesh, &r.sp), "ffhip: P2 numbering");
  const Mesh3 &Th = Uh.Th;
  vector<int32_t> en(10 * (size_t)Th.nt);
  hip_check(ffhip_p2v_download(r.sp, en.data()), "ffhip: P2 numbering");
  for (int k = 0; k < Th.nt; ++k) {
    const FElement3 K(Uh[k]);
    for (int df = 0; df < 10 * ncomp; ++df)
      if (K(df) != ncomp * en[10 * (size_t)k + df % 10] + df / 10) return false;
  }
  return true;
}

// the device matrix of res into A: a new HashMatrix (res->nnz_host set), or,
// when A's HashMatrix already holds this pattern (reuse), its values only
void device_into_matrix(ffhip_mat M, int half, bool reuse, ffhip_glue::Resident &res,
                        Matrice_Creuse<double> &A) {
  if (reuse) {
    hip_check(ffhip_glue::values_into_hashmatrix(M, *A.pHM()), "ffhip: values to the matrix");
    return;
  }
  int64_t n = 0, nnz = 0;
  ffhip_mat_info(M, &n, &nnz);
  // an ordinary HashMatrix (HashMatrix::set with CSR input, HashMatrix.cpp:697-728),
  // so every later FreeFEM operation works on it
  MatriceMorse<double> *pH = new MatriceMorse<double>((int)n, (int)n, (int)nnz, half);
  const int rc = ffhip_glue::csr_into_hashmatrix(M, *pH, half);
  if (rc != FFHIP_OK) {
    delete pH;
    hip_check(rc, "ffhip: matrix to FreeFEM");
  }
  A.A.master(pH);
  res.nnz_host = (int64_t)pH->nnz;
}

// [P2,P2,P2] Lame form; true when built on the GPU, false (A untouched) when
// FreeFEM's dof numbering is not the device's.  res: in, the previous
// assembly's device state to reuse (pattern kept), or null; out, this one's.
bool assemble_p2v_on_gpu(const FESpace3 &Uh, const RecognisedP2 &r, double tgv,
                         Matrice_Creuse<double> &A, ResidentPtr &res) {
  ffhip_ctx ctx = hip_ctx();
  const bool reuse = (bool)res;
  if (!reuse) {
    res.reset(new ffhip_glue::Resident);
    res->kind = FORM_P2V_3D;
    res->Vh = &Uh;
    upload_mesh3(Uh.Th, *res);
    if (!p2_numbering(Uh, 3, *res)) {
      res.reset();
      return false;
    }
    hip_check(ffhip_pattern_p2v(ctx, res->sp, &res->M), "ffhip: P2 pattern");
  }
  hip_check(ffhip_assemble_p2v_elasticity(ctx, res->sp, res->M, r.lambda, r.mu), "ffhip: P2 assembly");
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc_p2v(ctx, res->sp, res->M, labs.data(), (int)labs.size(), tgv, nullptr),
              "ffhip: on()");
  }
  device_into_matrix(res->M, 0, reuse, *res, A);
  return true;
}

// ---------------------------------------------------------------- coefficients
// The functions FreeFEM compiles a real finite element function into when a
// script uses it as a value: the casts pf3r -> double (3D, lgmesh3.cpp:
// 3093-3095, over pf3r2R<double,0,v_fes3>, lgmesh3.cpp:2169-2256) and
// pfer -> double (2D, lgfem.cpp:6993, over pfer2R<double,0>, lgfem.cpp:2054).
// They are looked up by name in the running executable (FreeFEM links with
// -rdynamic); where one is not found, coefficients go through the
// interpreter at the quadrature points as before.
Function1 fe_value_function(int dim) {
  static Function1 f3 = nullptr, f2 = nullptr;
  static bool looked = false;
  if (!looked) {
    looked = true;
    const char *n3[2] = {"_Z6pf3r2RIdLi0E6v_fes3E19AnyTypeWithOutCheckPvRKS1_",
                         "_Z6pf3r2RIdLi0E6v_fes3E16AnyTypeWithCheckPvRKS1_"};
    const char *n2[2] = {"_Z6pfer2RIdLi0EE19AnyTypeWithOutCheckPvRKS0_",
                         "_Z6pfer2RIdLi0EE16AnyTypeWithCheckPvRKS0_"};
    for (int k = 0; k < 2 && !f3; ++k) f3 = (Function1)dlsym(RTLD_DEFAULT, n3[k]);
    for (int k = 0; k < 2 && !f2; ++k) f2 = (Function1)dlsym(RTLD_DEFAULT, n2[k]);
  }
  return dim == 3 ? f3 : f2;
}

// A volume coefficient as the device takes it: CONST (c0) when it does not
// depend on the point, P0 / P1 (the array of the function's values) when it
// is a P0 or P1 finite element function of the form's own mesh with the
// standard numbering, else QP with the interpreter's values at the points.
struct CoefArg {
  int kind = FFHIP_COEF_CONST;
  double c0 = 0.;
  const double *ptr = nullptr;
  vector<double> own;
};

// the spaces whose numbering was checked (dof = element for P0, dof = vertex
// number for P1), so the O(nt) check runs once per space
bool numbering_checked(const UniqueffId &id, bool set) {
  static vector<UniqueffId> ok;
  for (const UniqueffId &u : ok)
    if (u == id) return true;
  if (set) {
    if (ok.size() > 16) ok.erase(ok.begin());
    ok.push_back(id);
  }
  return false;
}

// c is a P0 / P1 function of Th (3D): its kind and values (true), else false
bool fe_coef3(Stack stack, const C_F0 &c, const Mesh3 &Th, CoefArg &out) {
  const Function1 f = fe_value_function(3);
  const E_F0_Func1 *ef = dynamic_cast<const E_F0_Func1 *>(c.LeftValue());
  if (!f || !ef || ef->f != f) return false;
  pf3r p = GetAny<pf3r>((*ef->a)(stack));
  FEbase<double, v_fes3> *fe = p.first;
  if (!fe || p.second != 0) return false;
  if (!fe->x()) {  // an unset function evaluates to 0 (pf3r2R)
    out.kind = FFHIP_COEF_CONST;
    out.c0 = 0.;
    return true;
  }
  const FESpace3 *Vh = (FESpace3 *)fe->Vh;
  if (!Vh || &Vh->Th != &Th || Vh->N != 1 || Vh->TFE.N() != 1) return false;
  const bool p1 = Vh->TFE[0] == &DataFE<Mesh3>::P1, p0 = Vh->TFE[0] == &DataFE<Mesh3>::P0;
  if (!p1 && !p0) return false;
  if (Vh->NbOfDF != (p1 ? Th.nv : Th.nt)) return false;
  if (!numbering_checked(*Vh, false)) {
    for (int k = 0; k < Th.nt; ++k) {
      const FElement3 K((*Vh)[k]);
      if (p0 && K(0) != k) return false;
      if (p1)
        for (int i = 0; i < 4; ++i)
          if (K(i) != Th(Th[k][i])) return false;
    }
    numbering_checked(*Vh, true);
  }
  out.kind = p1 ? FFHIP_COEF_P1 : FFHIP_COEF_P0;
  out.ptr = (const double *)*fe->x();
  return true;
}

// the same for a 2D mesh
bool fe_coef2(Stack stack, const C_F0 &c, const Mesh &Th, CoefArg &out) {
  typedef pair<FEbase<double, v_fes> *, int> pfe2;
  const Function1 f = fe_value_function(2);
  const E_F0_Func1 *ef = dynamic_cast<const E_F0_Func1 *>(c.LeftValue());
  if (!f || !ef || ef->f != f) return false;
  pfe2 p = GetAny<pfe2>((*ef->a)(stack));
  FEbase<double, v_fes> *fe = p.first;
  if (!fe || p.second != 0) return false;
  if (!fe->x()) {
    out.kind = FFHIP_COEF_CONST;
    out.c0 = 0.;
    return true;
  }
  const FESpace *Vh = (FESpace *)fe->Vh;
  if (!Vh || &Vh->Th != &Th || Vh->N != 1 || Vh->TFE.N() != 1) return false;
  const bool p1 = Vh->TFE[0] == &P1Lagrange, p0 = Vh->TFE[0] == &P0Lagrange;
  if (!p1 && !p0) return false;
  if (Vh->NbOfDF != (p1 ? Th.nv : Th.nt)) return false;
  if (!numbering_checked(*Vh, false)) {
    for (int k = 0; k < Th.nt; ++k) {
      const FElement K((*Vh)[k]);
      if (p0 && K(0) != k) return false;
      if (p1)
        for (int i = 0; i < 3; ++i)
          if (K(i) != Th(Th[k][i])) return false;
    }
    numbering_checked(*Vh, true);
  }
  out.kind = p1 ? FFHIP_COEF_P1 : FFHIP_COEF_P0;
  out.ptr = (const double *)*fe->x();
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

// a volume coefficient of a 3D form as the device takes it (CoefArg)
void coef_arg3(Stack stack, const FESpace3 &Uh, bool is_const, double c0, const C_F0 &c,
               CoefArg &out) {
  if (is_const) {
    out.kind = FFHIP_COEF_CONST;
    out.c0 = c0;
  } else if (!fe_coef3(stack, c, Uh.Th, out)) {
    eval_coef_qp(stack, Uh, c, out.own);
    out.kind = FFHIP_COEF_QP;
    out.ptr = out.own.data();
  }
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

// the P1 form of r on the GPU into A; false (A untouched) when the device
// pattern cannot hold the mesh (a vertex row of more than 256 columns): the
// caller then assembles with FreeFEM.  res: in, the previous assembly's
// device state to reuse (pattern kept), or null; out, this one's.
bool assemble_on_gpu(Stack stack, const FESpace3 &Uh, const Recognised &r, double tgv, int sym,
                     Matrice_Creuse<double> &A, ResidentPtr &res) {
  const Mesh3 &Th = Uh.Th;
  ffhip_ctx ctx = hip_ctx();
  const bool reuse = (bool)res;
  if (!reuse) {
    res.reset(new ffhip_glue::Resident);
    res->kind = FORM_P1_3D;
    res->Vh = &Uh;
    res->half = sym ? 1 : 0;
    upload_mesh3(Th, *res);
    const int rc = ffhip_pattern_p1(ctx, res->mesh, &res->M);
    if (rc == FFHIP_ERR_UNSUPPORTED) {
      if (verbosity > 1) cout << " ffhip: " << ffhip_last_error() << ": FreeFEM assembles" << endl;
      res.reset();
      return false;
    }
    hip_check(rc, "ffhip: pattern");
  }
  ffhip_mesh m = res->mesh;
  ffhip_mat M = res->M;
  {
    CoefArg k, mc;
    coef_arg3(stack, Uh, r.coef_const, r.c0, r.coef, k);
    if (r.has_mass) coef_arg3(stack, Uh, r.mcoef_const, r.m0, r.mcoef, mc);
    else mc.kind = FFHIP_COEF_NONE;
    hip_check(ffhip_assemble_p1_form(ctx, m, M, k.kind, k.c0, k.ptr, mc.kind, mc.c0, mc.ptr, 0),
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
    struct FreeL { ffhip_mat L; ~FreeL() { ffhip_mat_free(L); } } guard{L};
    device_into_matrix(L, 1, reuse, *res, A);
  } else {
    device_into_matrix(M, 0, reuse, *res, A);
  }
  return true;
}

// Vh(Th,P2): true when built on the GPU; false (A untouched) when FreeFEM's
// dof numbering of Uh is not the first-appearance node numbering of the device
bool assemble_p2_on_gpu(Stack stack, const FESpace3 &Uh, const Recognised &r, double tgv,
                        Matrice_Creuse<double> &A, ResidentPtr &res) {
  ffhip_ctx ctx = hip_ctx();
  const bool reuse = (bool)res;
  if (!reuse) {
    res.reset(new ffhip_glue::Resident);
    res->kind = FORM_P2_3D;
    res->Vh = &Uh;
    upload_mesh3(Uh.Th, *res);
    if (!p2_numbering(Uh, 1, *res)) {
      res.reset();
      return false;
    }
    hip_check(ffhip_pattern_p2(ctx, res->sp, &res->M), "ffhip: P2 pattern");
  }
  {
    CoefArg k;
    coef_arg3(stack, Uh, r.coef_const, r.c0, r.coef, k);
    hip_check(ffhip_assemble_p2(ctx, res->sp, res->M, k.kind, k.c0, k.ptr, 0), "ffhip: P2 assembly");
  }
  if (!r.labels.empty()) {
    vector<int32_t> labs(r.labels.begin(), r.labels.end());
    hip_check(ffhip_set_bc_p2(ctx, res->sp, res->M, labs.data(), (int)labs.size(), tgv, nullptr),
              "ffhip: on()");
  }
  device_into_matrix(res->M, 0, reuse, *res, A);
  return true;
}

// A's previous GPU assembly, when it can serve again for this one: the same
// spaces (A.Uh, A.Vh: FreeFEM then keeps the HashMatrix and resets its values,
// problem.hpp:1679-1691), the same kind of form and storage, A's solver one of
// ffhip's and not shared, and the HashMatrix's pattern untouched since (no
// symbolic redo pending, still in CSR order, the same entry count)
template <class FES>
ResidentPtr reusable(Matrice_Creuse<double> &A, const FES &Uh, const FES &Vh, int kind, int half) {
  if (!A.A || A.Uh != Uh || A.Vh != Vh) return ResidentPtr();
  HMat *H = A.pHM();
  if (!H || !H->vsolver || H->re_do_symbolic || H->state != HMat::sorted_ij || H->half != half)
    return ResidentPtr();
  HipSolverBase *sb = dynamic_cast<HipSolverBase *>(H->vsolver);
  if (!sb || !sb->core.res) return ResidentPtr();
  const ffhip_glue::Resident &c = *sb->core.res;
  if (c.kind != kind || c.half != half || c.Vh != (const void *)&Uh || (int64_t)H->nnz != c.nnz_host)
    return ResidentPtr();
  return ResidentPtr(sb->release_resident());
}

// SetSolver with the device state handed to the solver it creates; what no
// ffhip solver claimed is freed
void set_solver_with(Stack stack, bool VF, Matrice_Creuse<double> &A, const Data_Sparse_Solver &ds,
                     ResidentPtr &res) {
  struct Clear {
    ~Clear() {
      delete handover().r;
      handover() = Handover();
    }
  } clear;
  handover().H = A.pHM();
  handover().r = res.release();
  SetSolver(stack, VF, *A.A, ds);  // as OpMatrixtoBilinearForm (problem.hpp:1654-1656)
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
      ResidentPtr res;
      if (ds.tgv > 0 && recognise(stack, *PUh, *PVh, b->largs, r) && (ds.sym == 0 || r.order == 1)) {
        const int kind = r.order == 1 ? FORM_P1_3D : FORM_P2_3D;
        res = reusable(A, *PUh, *PVh, kind, r.order == 1 ? ds.sym : 0);
        if (!res) {
          A.A = 0;
          A.Uh = *PUh;
          A.Vh = *PVh;
        }
        if (r.order == 1)  // the recognised P1 forms are symmetric: sym=1 keeps the lower half
          done = assemble_on_gpu(stack, *PUh, r, ds.tgv, ds.sym, A, res);
        else
          done = assemble_p2_on_gpu(stack, *PUh, r, ds.tgv, A, res);
      } else if (ds.sym == 0 && ds.tgv > 0 && recognise_p2v(stack, *PUh, *PVh, b->largs, r2)) {
        res = reusable(A, *PUh, *PVh, FORM_P2V_3D, 0);
        if (!res) {
          A.A = 0;
          A.Uh = *PUh;
          A.Vh = *PVh;
        }
        done = assemble_p2v_on_gpu(*PUh, r2, ds.tgv, A, res);
      }
      if (!done) {
        creationBlockOfMatrixToBilinearForm<double, Mesh3, FESpace3, FESpace3>(PUh, PVh, ds.sym, ds.tgv,
                                                                               b->largs, stack, A);
      }
      A.pHM()->half = ds.sym;
      if (square) set_solver_with(stack, VF, A, ds, res);
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

// a volume coefficient of a 2D form as the device takes it (CoefArg)
void coef_arg2(Stack stack, const FESpace &Uh, bool is_const, double c0, const C_F0 &c,
               CoefArg &out) {
  if (is_const) {
    out.kind = FFHIP_COEF_CONST;
    out.c0 = c0;
  } else if (!fe_coef2(stack, c, Uh.Th, out)) {
    eval_coef_qp2(stack, Uh, c, out.own);
    out.kind = FFHIP_COEF_QP;
    out.ptr = out.own.data();
  }
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

// the 2D P1 form of r on the GPU into A; false (A untouched) when the device
// pattern cannot hold the mesh.  res as for assemble_on_gpu.
bool assemble2_on_gpu(Stack stack, const FESpace &Uh, const Recognised &r, double tgv, int sym,
                      Matrice_Creuse<double> &A, ResidentPtr &res) {
  const Mesh &Th = Uh.Th;
  ffhip_ctx ctx = hip_ctx();
  const bool reuse = (bool)res;
  if (!reuse) {
    res.reset(new ffhip_glue::Resident);
    res->kind = FORM_P1_2D;
    res->Vh = &Uh;
    res->Th = &Th;
    res->half = sym ? 1 : 0;
    ffhip_glue::MeshArrays a;
    ffhip_glue::mesh2_arrays(Th, a);
    hip_check(ffhip_glue::upload_mesh(ctx, 2, a, &res->mesh), "ffhip: 2D mesh upload");
    const int rc = ffhip_pattern_p1(ctx, res->mesh, &res->M);
    if (rc == FFHIP_ERR_UNSUPPORTED) {
      if (verbosity > 1) cout << " ffhip: " << ffhip_last_error() << ": FreeFEM assembles" << endl;
      res.reset();
      return false;
    }
    hip_check(rc, "ffhip: pattern");
  }
  ffhip_mesh m = res->mesh;
  ffhip_mat M = res->M;
  {
    CoefArg k, mc;
    coef_arg2(stack, Uh, r.coef_const, r.c0, r.coef, k);
    if (r.has_mass) coef_arg2(stack, Uh, r.mcoef_const, r.m0, r.mcoef, mc);
    else mc.kind = FFHIP_COEF_NONE;
    hip_check(ffhip_assemble_p1_form(ctx, m, M, k.kind, k.c0, k.ptr, mc.kind, mc.c0, mc.ptr, 0),
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
    struct FreeL { ffhip_mat L; ~FreeL() { ffhip_mat_free(L); } } guard{L};
    device_into_matrix(L, 1, reuse, *res, A);
  } else {
    device_into_matrix(M, 0, reuse, *res, A);
  }
  return true;
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
      ResidentPtr res;
      bool done = false;
      if (ds.tgv > 0 && recognise2(stack, *PUh, *PVh, b->largs, r)) {
        res = reusable(A, *PUh, *PVh, FORM_P1_2D, ds.sym);
        if (!res) {
          A.A = 0;
          A.Uh = *PUh;
          A.Vh = *PVh;
        }
        done = assemble2_on_gpu(stack, *PUh, r, ds.tgv, ds.sym, A, res);
      }
      if (!done) {
        creationBlockOfMatrixToBilinearForm<double, Mesh, FESpace, FESpace>(PUh, PVh, ds.sym, ds.tgv,
                                                                            b->largs, stack, A);
      }
      A.pHM()->half = ds.sym;
      if (square) set_solver_with(stack, VF, A, ds, res);
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
  ffhip_ctx ctx = hip_ctx();
  ffhip_glue::Resident dev;  // the device mesh, freed on return
  upload_mesh3(Th, dev);
  ffhip_mesh m = dev.mesh;
  vector<double> gv, out(Th.nv);
  CoefArg f;
  if (r.has_f) coef_arg3(stack, Vh, r.f_const, r.f0, r.f, f);
  if (!r.bcs.empty()) eval_dirichlet(stack, Vh, r, gv);
  vector<int32_t> dl(r.dlabels.begin(), r.dlabels.end());
  hip_check(ffhip_rhs_p1_form(ctx, m, f.kind, f.c0, f.ptr, dl.data(), (int)dl.size(), tgv, 0.,
                              gv.empty() ? nullptr : gv.data(), out.data(), 0),
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
  ffhip_ctx ctx = hip_ctx();
  ffhip_glue::Resident dev;  // the device mesh and P2 numbering, freed on return
  upload_mesh3(Vh.Th, dev);
  if (!p2_numbering(Vh, r.ncomp, dev)) return false;
  vector<int32_t> labs(r.labels.begin(), r.labels.end());
  vector<double> out(b.N());
  if (r.ncomp == 3)
    hip_check(ffhip_rhs_p2v(ctx, dev.sp, r.f, labs.data(), (int)labs.size(), tgv, r.g, out.data(), 0),
              "ffhip: P2 linear form");
  else
    hip_check(ffhip_rhs_p2(ctx, dev.sp, r.f[0], labs.data(), (int)labs.size(), tgv, r.g[0], out.data(), 0),
              "ffhip: P2 linear form");
  for (long i = 0; i < b.N(); ++i) b[i] = out[i];
  return true;
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
      // the device copy of A's ffhip solver when it has one (refreshed if the
      // script changed A since), else a temporary upload (half matrices
      // expanded to the full symmetric operator)
      HipSolverBase *sb = H->vsolver ? dynamic_cast<HipSolverBase *>(H->vsolver) : nullptr;
      std::unique_ptr<Core> tmp;
      Core *core = nullptr;
      if (sb) {
        core = &sb->core;
      } else {
        tmp.reset(new Core(H, hip_ctx(), FFHIP_CG_FP64));
        core = tmp.get();
      }
      core->note_changes(H->GetReDoNumerics(), H->GetReDoSymbolic());
      hip_check(core->refresh(), "LinearCG: matrix to the GPU");
      double eps_out = eps;
      ffhip_cg_stats st;
      const int rc = ffhip_cg_solve_affine(core->cg, b ? (double *)*b : nullptr, (double *)x, 0, eps,
                                           (int)nbitermax, jacobi ? FFHIP_CG2_JACOBI : FFHIP_CG2_IDENTITY,
                                           &eps_out, &st);
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
