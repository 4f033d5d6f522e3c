// ffhip_glue.hpp — the marshalling between FreeFEM's objects and libffhip's
// C-ABI, shared by the plugin (ffhip.cpp) and the test driver that runs it
// against the compiled reference objects (oracle/seam_driver.cpp).
//
// It depends on FreeFEM's femlib types only through templates (a Mesh3 /
// Mesh, a HashMatrix<int,double>), never on the interpreter, so it compiles
// with the femlib headers alone.  The includer includes the FreeFEM headers
// (Mesh3dn.hpp, HashMatrix.hpp) and ffhip.h first.
//
//  * mesh3_arrays / mesh2_arrays: a FreeFEM mesh as the flat arrays
//    ffhip_mesh3_upload / ffhip_mesh2_upload take (P1 dof = vertex number,
//    GenericMesh.hpp:1762-1772);
//  * hashmatrix_from_csr / csr_into_hashmatrix: a (device) CSR into a HashMatrix through
//    HashMatrix::set(n, n, half, nnz, ia, ja, a, 0, tcsr = 1)
//    (HashMatrix.cpp:697-728); the entries arrive sorted by (i, j), so the
//    matrix is marked sorted (state = sorted_ij) and HashMatrix::CSR() skips
//    its heap sort (HashMatrix.cpp:670-682, 858-867);
//  * values_into_hashmatrix: new values on an unchanged pattern, written over
//    aij in place (the pattern-keeping re-assembly of
//    creationBlockOfMatrixToBilinearForm, problem.hpp:1679-1691);
//  * upload_for_solver: a HashMatrix as the full device matrix the solvers
//    take — HashMatrix::CSR() uploaded, and for a half matrix (sym = 1, the
//    lower triangle; HashMatrix::half) the full symmetric matrix
//    L + L^T - diag L built on the device (the operator of the half branch of
//    HashMatrix::addMatMul, HashMatrix.cpp:1128-1136);
//  * cg_written_back_eps: the eps SolverCG::dosolver writes to veps: the
//    absolute threshold ConjugueGradient leaves in its eps argument,
//    sqrt(eps^2 (g0, C g0)) for eps > 0, eps itself otherwise (CG.cpp:226,
//    VirtualSolverCG.hpp:180-186);
//  * Resident: the device state of a form assembled on the GPU (mesh,
//    pattern, values), handed from the assembly operator to the solver that
//    SetSolver creates for the same matrix, so the solver starts from the
//    device copy instead of uploading the HashMatrix again.
#ifndef FFHIP_GLUE_HPP
#define FFHIP_GLUE_HPP

#include <cmath>
#include <cstdint>
#include <vector>

namespace ffhip_glue {

struct MeshArrays {
  std::vector<double> xyz;      // dim per vertex
  std::vector<int32_t> elems;   // dim + 1 per element
  std::vector<int32_t> belems;  // dim per boundary element
  std::vector<int32_t> blab;    // label per boundary element
  long nv = 0, nt = 0, nbe = 0;
};

// Mesh3: vertices, tets, boundary triangles and their labels
template <class Mesh3>
void mesh3_arrays(const Mesh3 &Th, MeshArrays &a) {
  a.nv = Th.nv; a.nt = Th.nt; a.nbe = Th.nbe;
  a.xyz.resize(3 * (size_t)Th.nv);
  a.elems.resize(4 * (size_t)Th.nt);
  a.belems.resize(3 * (size_t)Th.nbe);
  a.blab.resize(Th.nbe);
  for (int i = 0; i < Th.nv; ++i) {
    a.xyz[3 * (size_t)i] = Th(i).x;
    a.xyz[3 * (size_t)i + 1] = Th(i).y;
    a.xyz[3 * (size_t)i + 2] = Th(i).z;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 4; ++j) a.elems[4 * (size_t)k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.nbe; ++k) {
    for (int j = 0; j < 3; ++j) a.belems[3 * (size_t)k + j] = Th(Th.be(k)[j]);
    a.blab[k] = Th.be(k).lab;
  }
}

// the 2D Mesh: vertices, triangles, boundary edges and their labels
template <class Mesh2>
void mesh2_arrays(const Mesh2 &Th, MeshArrays &a) {
  a.nv = Th.nv; a.nt = Th.nt; a.nbe = Th.neb;
  a.xyz.resize(2 * (size_t)Th.nv);
  a.elems.resize(3 * (size_t)Th.nt);
  a.belems.resize(2 * (size_t)Th.neb);
  a.blab.resize(Th.neb);
  for (int i = 0; i < Th.nv; ++i) {
    a.xyz[2 * (size_t)i] = Th(i).x;
    a.xyz[2 * (size_t)i + 1] = Th(i).y;
  }
  for (int k = 0; k < Th.nt; ++k)
    for (int j = 0; j < 3; ++j) a.elems[3 * (size_t)k + j] = Th(Th[k][j]);
  for (int k = 0; k < Th.neb; ++k) {
    for (int j = 0; j < 2; ++j) a.belems[2 * (size_t)k + j] = Th(Th.bedges[k][j]);
    a.blab[k] = Th.bedges[k].lab;
  }
}

inline int upload_mesh(ffhip_ctx ctx, int dim, const MeshArrays &a, ffhip_mesh *m) {
  if (dim == 3)
    return ffhip_mesh3_upload(ctx, a.nv, a.nt, a.nbe, a.xyz.data(), a.elems.data(),
                              a.belems.data(), a.blab.data(), m);
  return ffhip_mesh2_upload(ctx, a.nv, a.nt, a.nbe, a.xyz.data(), a.elems.data(), a.belems.data(),
                            a.blab.data(), m);
}

// a CSR with sorted rows (int64 row pointers, as the device keeps them) into
// H; half = 1 for a lower triangle
template <class HMat>
int hashmatrix_from_csr(HMat &H, int64_t n, int half, const int64_t *rp, const int *ci,
                        const double *va) {
  const int64_t nnz = rp[n];
  if (n >= (int64_t)1 << 31 || nnz >= (int64_t)1 << 31) return FFHIP_ERR_UNSUPPORTED;
  std::vector<int> ia(n + 1);
  for (int64_t i = 0; i <= n; ++i) ia[i] = (int)rp[i];
  H.set((int)n, (int)n, half, (size_t)nnz, ia.data(), const_cast<int *>(ci), const_cast<double *>(va),
        0, 1);
  H.state = HMat::sorted_ij;  // set() kept the (i, j) order of the CSR input
  H.re_do_numerics = 0;       // the values are the device's: nothing to redo
  H.re_do_symbolic = 0;
  return FFHIP_OK;
}

// the device matrix M (CSR, sorted rows) into H; half = 1 for a lower triangle
template <class HMat>
int csr_into_hashmatrix(ffhip_mat M, HMat &H, int half) {
  int64_t n = 0, nnz = 0;
  int rc = ffhip_mat_info(M, &n, &nnz);
  if (rc != FFHIP_OK) return rc;
  if (n >= (int64_t)1 << 31 || nnz >= (int64_t)1 << 31) return FFHIP_ERR_UNSUPPORTED;
  std::vector<int64_t> rp(n + 1);
  std::vector<int> ja(nnz);
  std::vector<double> aa(nnz);
  rc = ffhip_mat_download(M, rp.data(), ja.data(), aa.data());
  if (rc != FFHIP_OK) return rc;
  return hashmatrix_from_csr(H, n, half, rp.data(), ja.data(), aa.data());
}

// M's values over H's, H holding M's pattern in CSR order (checked: sizes and
// sort state; the caller knows the pattern came from M)
template <class HMat>
int values_into_hashmatrix(ffhip_mat M, HMat &H) {
  int64_t n = 0, nnz = 0;
  int rc = ffhip_mat_info(M, &n, &nnz);
  if (rc != FFHIP_OK) return rc;
  if (H.n != n || (int64_t)H.nnz != nnz || H.state != HMat::sorted_ij) return FFHIP_ERR_ARG;
  rc = ffhip_mat_download(M, nullptr, nullptr, H.aij);
  if (rc != FFHIP_OK) return rc;
  H.re_do_numerics = 0;
  return FFHIP_OK;
}

// H as the full device matrix of the solvers (a new device matrix in *out)
template <class HMat>
int upload_for_solver(ffhip_ctx ctx, HMat &H, ffhip_mat *out) {
  *out = nullptr;
  int *ia = 0, *ja = 0;
  double *a = 0;
  H.CSR(ia, ja, a);
  const int64_t n = H.n, nnz = ia[n];
  std::vector<int64_t> rp(ia, ia + n + 1);
  ffhip_mat M = nullptr;
  int rc = ffhip_mat_upload(ctx, n, nnz, rp.data(), ja, a, &M);
  if (rc != FFHIP_OK) return rc;
  if (H.half > 0) {  // the stored lower triangle: expand on the device
    ffhip_mat F = nullptr;
    rc = ffhip_mat_set_half(M, 1);
    if (rc == FFHIP_OK) rc = ffhip_mat_full(M, &F);
    ffhip_mat_free(M);
    if (rc != FFHIP_OK) return rc;
    M = F;
  }
  *out = M;
  return FFHIP_OK;
}

// new values of a full H (same pattern) into the device matrix M made by
// upload_for_solver; FFHIP_ERR_UNSUPPORTED for a half H, whose device matrix
// is the expanded one: the caller uploads it again (upload_for_solver)
template <class HMat>
int values_for_solver(HMat &H, ffhip_mat M) {
  if (H.half > 0) return FFHIP_ERR_UNSUPPORTED;
  int *ia = 0, *ja = 0;
  double *a = 0;
  H.CSR(ia, ja, a);
  return ffhip_mat_set_values(M, a);
}

inline double cg_written_back_eps(double eps, double gcg0) {
  return eps > 0 ? std::sqrt(eps * eps * gcg0) : eps;
}

// device state of a form assembled on the GPU: the mesh, and the full matrix
// (pattern + values) whose entries the HashMatrix holds (its lower triangle
// when half)
struct Resident {
  ffhip_mesh mesh = nullptr;
  ffhip_p2v sp = nullptr;    // P2 node numbering (P2 forms)
  ffhip_mat M = nullptr;
  const void *Th = nullptr;  // the FreeFEM mesh it was built from
  const void *Vh = nullptr;  // the FreeFEM space
  int kind = 0;              // which assembly built it (the plugin's form kind)
  int half = 0;
  int64_t nnz_host = 0;      // entries of the HashMatrix (lower triangle when half)
  Resident() = default;
  Resident(const Resident &) = delete;
  Resident &operator=(const Resident &) = delete;
  ~Resident() {
    if (M) ffhip_mat_free(M);
    if (sp) ffhip_p2v_free(sp);
    if (mesh) ffhip_mesh_free(mesh);
  }
};

// The device side of SolverHIPCG / SolverHIPGMRES for one HashMatrix: the
// device matrix (adopted from the assembly, or uploaded from the HashMatrix)
// and the solver built on it.  The solver class calls symbolic() / numeric()
// from fac_symbolic / fac_numeric, note_changes() from UpdateState with the
// matrix's redo flags (HashMatrix::GetReDoNumerics / GetReDoSymbolic,
// HashMatrix.hpp:278-279), and refresh() before each solve, so a HashMatrix
// changed by the script after the device copy was made is read again.
template <class HMat>
struct SolverCore {
  HMat *A;
  ffhip_ctx ctx;
  int mode;                 // ffhip_cg_create mode
  Resident *res = nullptr;  // adopted device state (owned)
  ffhip_mat M = nullptr;    // the solver's matrix: res->M or own
  bool own_M = false;
  ffhip_cg cg = nullptr;
  bool ready = false, need_sym = false, need_num = false;
  int64_t uploads = 0;      // host-to-device matrix uploads (0 when the assembly's copy served)

  SolverCore(HMat *a, ffhip_ctx c, int m) : A(a), ctx(c), mode(m) {}
  SolverCore(const SolverCore &) = delete;
  SolverCore &operator=(const SolverCore &) = delete;
  ~SolverCore() { release(); delete res; }

  void release() {
    if (cg) ffhip_cg_free(cg);
    cg = nullptr;
    if (own_M && M) ffhip_mat_free(M);
    M = nullptr;
    own_M = false;
    ready = false;
  }
  // the adopted device copy matches A while A has not changed since it was
  // written (its redo flags were cleared then) and A is the full / half
  // matrix the copy was made for
  void adopt(Resident *r) {
    delete res;
    res = r;
  }
  // the adopted device state, for a re-assembly that keeps the pattern; null
  // when the HashMatrix's pattern may have changed since (a noted symbolic
  // redo).  The solver built on it is released.
  Resident *take_resident() {
    Resident *r = res;
    if (r && M == r->M) release();
    res = nullptr;
    if (r && need_sym) {
      delete r;
      r = nullptr;
    }
    return r;
  }
  void note_changes(bool numerics, bool symbolic) {
    if (symbolic) need_sym = true;
    else if (numerics) need_num = true;
  }
  int symbolic() {
    release();
    int rc;
    if (res && !need_sym && !need_num && (int64_t)A->nnz == res->nnz_host &&
        (A->half > 0) == (res->half > 0)) {
      M = res->M;
    } else {
      delete res;
      res = nullptr;
      rc = upload_for_solver(ctx, *A, &M);
      if (rc != FFHIP_OK) return rc;
      own_M = true;
      ++uploads;
    }
    rc = ffhip_cg_create(ctx, M, mode, &cg);
    if (rc != FFHIP_OK) { release(); return rc; }
    ready = true;
    need_sym = need_num = false;
    return FFHIP_OK;
  }
  int numeric() {
    if (!cg) return symbolic();
    if (res && M == res->M && !need_num) return FFHIP_OK;  // the device values are current
    if (res && M == res->M) return symbolic();  // the assembly's copy is stale: upload A
    int rc = values_for_solver(*A, M);
    if (rc == FFHIP_ERR_UNSUPPORTED) {
      need_sym = true;
      return symbolic();
    }
    if (rc == FFHIP_OK) {
      ++uploads;
      rc = ffhip_cg_update_values(cg);
    }
    need_num = false;
    return rc;
  }
  int refresh() {
    if (!ready || need_sym) return symbolic();
    if (need_num) return numeric();
    return FFHIP_OK;
  }
};

}  // namespace ffhip_glue

#endif
