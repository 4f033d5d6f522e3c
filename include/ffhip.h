/*
 * ffhip.h — C-ABI of libffhip: MI355X-native replacement for FreeFEM's
 * bilinear-form assembly (AssembleBilinearForm / Element_Op -> HashMatrix),
 * the CSR build (HashMatrix::CSR) and the Jacobi-preconditioned CG solve
 * (SolverCG / ConjugueGradient).
 *
 * Plain C types only (no torch, no C++ classes).  Every entry point returns 0
 * on success or a negative FFHIP_ERR_* code; ffhip_last_error() gives the
 * message.  All calls are synchronous from the caller's point of view (the
 * library's stream is synchronised before returning), matching the
 * single-threaded, synchronous solver contract of FreeFEM's plugins
 * (src/femlib/VirtualSolver.hpp:190-243).
 *
 * Citations are to the FreeFEM source tree (src/...).
 */
#ifndef FFHIP_H
#define FFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FFHIP_OK = 0,
  FFHIP_ERR_ARG = -1,        /* bad argument / size mismatch (ExecError)   */
  FFHIP_ERR_HIP = -2,        /* HIP runtime failure                        */
  FFHIP_ERR_NOCONV = -3,     /* CG did not converge (VirtualSolverCG.hpp:188) */
  FFHIP_ERR_NAN = -4,        /* NaN in CG (CG.cpp:228 assert)             */
  FFHIP_ERR_UNSUPPORTED = -5,/* mesh / form outside the supported cases    */
  FFHIP_ERR_INTERNAL = -6
};

typedef struct ffhip_ctx_s *ffhip_ctx;   /* device, stream, scratch        */
typedef struct ffhip_mesh_s *ffhip_mesh; /* device-resident mesh            */
typedef struct ffhip_mat_s *ffhip_mat;   /* device-resident CSR matrix      */
typedef struct ffhip_cg_s *ffhip_cg;     /* device-resident CG solver       */
typedef struct ffhip_p2v_s *ffhip_p2v;   /* [P2,P2,P2] space on a 3D mesh   */

const char *ffhip_last_error(void);
const char *ffhip_version(void);

/* Library init on one GPU (plugin Load_Init, load.cpp:93-105 / InitFunct.hpp:76-98).
 * All work of a context is ordered on its one stream.  Device blocks that
 * handles release are kept per device and reused by later allocations of
 * about the same size; they go back to the device when the last context of
 * the device is freed.  Free every handle of a context before the context.
 * ffhip_init makes `device` the calling thread's current device; the calls on
 * a context allocate on the current device, so a program that drives several
 * devices selects the context's device (hipSetDevice) before calling on it. */
int ffhip_init(int device, ffhip_ctx *ctx);
int ffhip_ctx_free(ffhip_ctx ctx);
int ffhip_sync(ffhip_ctx ctx);

/* ---- meshes ------------------------------------------------------------ */
/* cube(nx,ny,nz) built on the device, identical to BuildCube kind=6
 * (src/fflib/msh3.cpp:7879-8132; labels as Cube_Op 8134-8164: labels[6] in
 * the order of the script's label=[...] argument, default {1,...,6}). */
int ffhip_mesh3_cube(ffhip_ctx ctx, int nx, int ny, int nz,
                     const int32_t labels[6], ffhip_mesh *mesh);
/* square(nx,ny) built on the device, identical to Carre_ flags=0
 * (src/fflib/lgmesh.cpp:1229-1384; labels bottom,right,top,left). */
int ffhip_mesh2_square(ffhip_ctx ctx, int nx, int ny, const int32_t labels[4],
                       ffhip_mesh *mesh);
/* Upload a Mesh3 (src/femlib/Mesh3dn.hpp:143): vertices xyz[3*nv], tets[4*nt]
 * (0-based vertex ids, Tet vertex order), boundary triangles btri[3*nbe] with
 * labels blab[nbe].  Host pointers. */
int ffhip_mesh3_upload(ffhip_ctx ctx, int64_t nv, int64_t nt, int64_t nbe,
                       const double *xyz, const int32_t *tets,
                       const int32_t *btri, const int32_t *blab,
                       ffhip_mesh *mesh);
/* Upload a 2D Mesh (src/femlib/fem.hpp, class Mesh): vertices xy[2*nv],
 * triangles tris[3*nt] (0-based vertex ids, Triangle vertex order), boundary
 * edges bedges[2*nbe] with labels blab[nbe].  Host pointers. */
int ffhip_mesh2_upload(ffhip_ctx ctx, int64_t nv, int64_t nt, int64_t nbe,
                       const double *xy, const int32_t *tris,
                       const int32_t *bedges, const int32_t *blab,
                       ffhip_mesh *mesh);
int ffhip_mesh_info(ffhip_mesh mesh, int64_t *nv, int64_t *nt, int64_t *nbe);
/* Copy the mesh back to host arrays (any pointer may be NULL). */
int ffhip_mesh_download(ffhip_mesh mesh, double *xyz, int32_t *elems,
                        int32_t *belems, int32_t *blab);
int ffhip_mesh_free(ffhip_mesh mesh);

/* ---- P1 matrix: pattern, values, Dirichlet ----------------------------- */
/* Integer CSR pattern of the P1 stiffness matrix, bit-exact with the
 * HashMatrix that `A += mate` builds over the element loop
 * (src/femlib/HashMatrix.cpp:1294-1332) once sorted by HashMatrix::CSR
 * (HashMatrix.cpp:858-867, Sortij 670-682, Buildp 992-1030).  Also builds the
 * element-to-row map used by the atomic-free assembly. */
int ffhip_pattern_p1(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_mat *mat);
/* `matrix A = va(Vh,Vh)` of int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+
 * dz(u)*dz(v))) on a 3D P1 space in one call: the pattern of
 * ffhip_pattern_p1 and the values of ffhip_assemble_p1 (kappa described as
 * there), the same matrix; what creationBlockOfMatrixToBilinearForm does the
 * first time a matrix is assembled on a space (problem.hpp:1663-1714).  The
 * element pass runs beside the incidence and pattern passes, which also form
 * the row values. */
int ffhip_pattern_assemble_p1(ffhip_ctx ctx, ffhip_mesh mesh, int kind, double c0,
                              const double *kappa_e, int kappa_on_device, ffhip_mat *mat);
int ffhip_mat_info(ffhip_mat mat, int64_t *n, int64_t *nnz);

enum { FFHIP_COEF_NONE = -1, FFHIP_COEF_CONST = 0, FFHIP_COEF_SIN3 = 1, FFHIP_COEF_P0 = 2,
       FFHIP_COEF_QP = 3, FFHIP_COEF_P1 = 4 };
/* Values of int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v))) with the
 * default Tet_5 rule (src/fflib/problem.cpp:14102-14124), replacing
 * AssembleBilinearForm 3D (problem.cpp:1116-1417) + Element_Op
 * (problem.cpp:6337-6437).  kappa: CONST -> c0; SIN3 ->
 * 1+0.5*sin(2 pi x)sin(2 pi y)sin(2 pi z); P0 -> kappa_e[element] (nt values);
 * QP -> kappa_e[element * nq + q], the coefficient already evaluated at the
 * nq quadrature points of each element (nq = 14 for Tet_5, 7 for T_5) — the
 * values FreeFEM's interpreter produces for any scalar coefficient expression
 * (problem.cpp:6392-6395); P1 -> kappa_e[vertex] (nv values), a P1 finite
 * element function interpolated at each quadrature point as FreeFEM evaluates
 * one there (TypeOfFE_P1Lagrange3d::operator(), P012_3d.cpp:98-106:
 * u0 l0 + u1 l1 + u2 l2 + u3 l3 with l0 = 1 - x - y - z; 2D:
 * TypeOfFE_P1Lagrange::operator(), FESpace.cpp:1078-1085).  kappa_e is a host
 * pointer, or a device pointer when kappa_on_device != 0.  2D meshes: int2d
 * with T_5, terms dx,dy. */
int ffhip_assemble_p1(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_mat mat, int kind,
                      double c0, const double *kappa_e, int kappa_on_device);
/* Values of int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v)) + m*u*v)
 * (int2d with dx, dy in 2D): the same Element_Op loop with a reaction term
 * (op_id x op_id, problem.cpp:6384-6437; P1 values P012_3d.cpp:123-175), e.g.
 * the u*v/dt of an implicit heat step.  m is described like kappa: mkind
 * FFHIP_COEF_NONE (no reaction term: ffhip_assemble_p1), CONST (m0), SIN3,
 * P0 (mass_e[element]), QP (mass_e at the quadrature points) or P1
 * (mass_e[vertex], interpolated at the points).  kappa_e and
 * mass_e are device pointers when on_device != 0, else host pointers. */
int ffhip_assemble_p1_form(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_mat mat, int kind, double c0,
                           const double *kappa_e, int mkind, double m0, const double *mass_e,
                           int on_device);
/* on(labels..., u=g) with tgv > 0: AssembleBC (problem.cpp:10039-10194) then
 * HashMatrix::SetBC (HashMatrix.cpp:1194-1238): a_ii = tgv on the dofs of
 * boundary elements whose label is listed.  onbc (optional, host, n bytes)
 * receives the marked dofs. */
int ffhip_set_bc(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_mat mat,
                 const int32_t *labels, int nlab, double tgv, uint8_t *onbc);
/* Right-hand side of int3d(Th)(f*v) (+ on(...)=g rows set to tgv*g):
 * AssembleLinearForm/Element_rhs 3D (problem.cpp:10878, 7917-7985) and the B
 * part of AssembleBC (problem.cpp:10176).  b is a device pointer (n doubles)
 * when on_device != 0, else a host pointer. */
int ffhip_rhs_p1(ffhip_ctx ctx, ffhip_mesh mesh, double f,
                 const int32_t *labels, int nlab, double tgv, double g,
                 double *b, int on_device);
/* The same linear form with a coefficient: int3d(Th)(f*v) (int2d in 2D) with
 * f described like kappa in ffhip_assemble_p1 (kind CONST f0, SIN3, P0
 * fcoef[element], QP fcoef[element * nq + q] at the Tet_5 / T_5 points — the
 * values FreeFEM's interpreter gives any expression there, Element_rhs
 * problem.cpp:7940-7978 — or P1 fcoef[vertex] interpolated at the points),
 * and on(labels, u=g) rows set to tgv * g with g per
 * dof (gvals, n values: g at the dof's vertex, as AssembleBC evaluates it,
 * problem.cpp:10120-10176) or the constant g0 when gvals is NULL.  fcoef,
 * gvals and b are device pointers when on_device != 0, else host pointers. */
int ffhip_rhs_p1_form(ffhip_ctx ctx, ffhip_mesh mesh, int kind, double f0,
                      const double *fcoef, const int32_t *labels, int nlab,
                      double tgv, double g0, const double *gvals, double *b,
                      int on_device);
/* ---- boundary integrals (P1, Robin / Neumann terms) -------------------- */
/* int2d(Th3, labels)(alpha*u*v) (2D meshes: int1d(Th2, labels)(alpha*u*v))
 * added to the values of mat, the P1 matrix of this mesh (ffhip_pattern_p1):
 * AssembleBilinearForm's boundary branch (problem.cpp:1273-1327) over the
 * boundary elements whose label is listed, with the face rule of Element_Op
 * (problem.cpp:6515-6561: T_5 on a face, coef = |N|/2 w_q; edges
 * problem.cpp:6223-6240: QF_GaussLegendre3, coef = |E| w_q).  Call it after
 * the volume terms and before ffhip_set_bc, the order FreeFEM adds them in.
 * alpha: kind CONST (c0), SIN3, P0 (coef[boundary element], nbe values) or
 * QP (coef[f * nq + q], nq = 7 on faces, 3 on edges, at the points
 * x_q = (1 - s_q - t_q) B0 + s_q B1 + t_q B2 of boundary element f with
 * vertices B0, B1, B2 in mesh order and (s_q, t_q) the T_5 points; edges:
 * x_q = (1 - s_q) B0 + s_q B1 with s_q = 1/2, (1 - sqrt(3/5))/2,
 * (1 + sqrt(3/5))/2).  coef is a device pointer when on_device != 0. */
int ffhip_assemble_p1_boundary(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_mat mat,
                               const int32_t *labels, int nlab, int kind,
                               double c0, const double *coef, int on_device);
/* b += int2d(Th3, labels)(g*v) (int1d in 2D): the boundary terms of
 * AssembleLinearForm (Element_rhs on a face, problem.cpp:8517-8590), g
 * described like alpha above (g0 for CONST).  The rows of the vertices of
 * boundary elements labelled dlabels (the Dirichlet labels of the problem,
 * whose rows on(...) sets to tgv*g afterwards) are left unchanged, so the
 * call can follow ffhip_rhs_p1.  b: n doubles, device pointer when
 * on_device != 0. */
int ffhip_rhs_p1_boundary(ffhip_ctx ctx, ffhip_mesh mesh, const int32_t *labels,
                          int nlab, int kind, double g0, const double *coef,
                          const int32_t *dlabels, int ndlab, double *b,
                          int on_device);
/* ---- symmetric half storage (sym=1) ------------------------------------- */
/* The lower triangle (row >= column) of a symmetric device matrix, as a new
 * half matrix: what `matrix A = va(Vh, Vh, sym=1)` holds (HashMatrix::
 * operator+= with MatriceElementaire::Symmetric stores (max(i,j), min(i,j)),
 * HashMatrix.cpp:1312-1318; HashMatrix::half = 1).  ffhip_spmv on a half
 * matrix computes y = (L + L^T - diag L) x, the half branch of
 * HashMatrix::addMatMul (HashMatrix.cpp:1128-1136), trans alike;
 * ffhip_mat_download / ffhip_mat_coo / ffhip_mat_write_morse give the stored
 * triangle (the text with half = 1). */
int ffhip_mat_lower(ffhip_mat mat, ffhip_mat *half_mat);
/* Declare an uploaded matrix (ffhip_mat_upload) half: its entries must all
 * have column <= row (checked; FFHIP_ERR_ARG otherwise).  half = 0 is
 * accepted for a full matrix only. */
int ffhip_mat_set_half(ffhip_mat mat, int half);
/* The full symmetric matrix of a half one (new device matrix; the solvers
 * take full matrices). */
int ffhip_mat_full(ffhip_mat half_mat, ffhip_mat *mat);
/* *half = 1 for a half matrix, 0 otherwise. */
int ffhip_mat_half(ffhip_mat mat, int *half);
/* ---- matrix exchange: triplets and FreeFEM's text files ---------------- */
/* [I, J, C] = A (lgmat.cpp:2670-2710): the entries in HashMatrix::COO()
 * order (rows, then columns, ascending; HashMatrix.cpp:837-843), 0-based, and
 * — as FreeFEM does when A holds no (n-1, n-1) entry — one more entry
 * (n, n, 0).  *count receives the number of entries; with I, J, C all NULL
 * only the count is returned.  I, J: int64 (FreeFEM's KN<long>), C: double;
 * device pointers when on_device != 0, else host pointers. */
int ffhip_mat_coo(ffhip_mat mat, int64_t *count, int64_t *I, int64_t *J,
                  double *C, int on_device);
/* `ofstream f(path); f << A;` for a CSR-state HashMatrix (HashMatrix.hpp:
 * 485-508): three '#' lines, "n m half  nnz", then one line per entry,
 * setw(9) i+1, setw(9) j+1, the value with setprecision(20) (|a| < 1e-305
 * as 0, RNM_tpl.hpp:60).  Host arrays (a sorted CSR); the first comment
 * line carries an address, as FreeFEM's does. */
int ffhip_csr_write_morse(int64_t n, int64_t m, int half, const int64_t *rowptr,
                          const int32_t *colidx, const double *vals,
                          const char *path);
/* The same text for a device matrix (downloaded, then written). */
int ffhip_mat_write_morse(ffhip_mat mat, const char *path);
/* `ifstream f(path); f >> A;`: HashMatrix(istream) (HashMatrix.cpp:7-30,
 * 136-214) reading the Morse text or the COO dump ("(Morse)" / "(COO)" in
 * the first comment line; other '#' lines skipped; a repeated (i,j) keeps
 * its last value), as the sorted CSR HashMatrix::CSR gives.  n, m, half and
 * nnz are always returned; rowptr (n+1), colidx, vals (host) are filled when
 * given and nnz <= cap (else FFHIP_ERR_ARG). */
int ffhip_csr_read_text(const char *path, int64_t cap, int64_t *n, int64_t *m,
                        int *half, int64_t *nnz, int64_t *rowptr,
                        int32_t *colidx, double *vals);
/* Read a square matrix file into a device matrix; a half (lower-triangle)
 * file gives the full symmetric matrix it stands for. */
int ffhip_mat_read_text(ffhip_ctx ctx, const char *path, ffhip_mat *mat);
/* ---- adjoint building blocks (3D P1) ----------------------------------- */
/* J = 1/2 int_Th (u - ud)^2 and the adjoint right-hand side
 * r = M (u - ud) with the rows of the Dirichlet dofs (boundary labels listed)
 * set to 0 — the mass varf int3d(Th)(u*v) of FreeFEM (Element_Op with op_id
 * terms, problem.cpp:6384-6437; Tet_5 is exact for it) applied to u - ud, and
 * the zeroed Dirichlet rows of the adjoint problem.  u, ud, r: device
 * pointers (n doubles); J: host pointer. FreeFEM has no adjoint routine; this
 * is the continuous adjoint of its composed varfs. */
int ffhip_p1_misfit(ffhip_ctx ctx, ffhip_mesh mesh, const double *u,
                    const double *ud, const int32_t *labels, int nlab,
                    double *J, double *r);
/* ffhip_p1_misfit without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * temporaries (the J partials, the label list, the Dirichlet marks) live in a
 * fixed device array of the library, one per device, of 16.8 MB: meshes of
 * at most 2^24 vertices, at most 64 labels (FFHIP_ERR_ARG beyond).  The
 * mesh's vertex incidence must already exist when r is wanted (any assembly
 * or ffhip_p1_misfit call on the mesh builds it).  Calls that overlap on one
 * device share that array and race; the results equal ffhip_p1_misfit's. */
int ffhip_p1_misfit_noalloc(ffhip_ctx ctx, ffhip_mesh mesh, const double *u,
                            const double *ud, const int32_t *labels, int nlab,
                            double *J, double *r);
/* g[e] = dJ/dkappa_e = - int_K grad u . grad lambda  (= -|K| grad u_K . grad
 * lambda_K for P1, exact), the P0 linear form int3d(Th)(-(grad u.grad l)*w)
 * of the reference's composed varfs.  Gradients are recomputed from the
 * coordinates, not stored.  u, lambda, g: device pointers. */
int ffhip_p1_grad_kappa(ffhip_ctx ctx, ffhip_mesh mesh, const double *u,
                        const double *lambda, double *g);
/* CSR out (host): rowptr[n+1] (int64), colidx[nnz], vals[nnz]; any NULL skipped. */
int ffhip_mat_download(ffhip_mat mat, int64_t *rowptr, int32_t *colidx,
                       double *vals);
/* CSR in (host) — the solver path of a FreeFEM HashMatrix: A->CSR(ia,ja,a)
 * in SolverHIPCG::fac_symbolic/fac_numeric (HashMatrix.cpp:858-867). */
int ffhip_mat_upload(ffhip_ctx ctx, int64_t n, int64_t nnz,
                     const int64_t *rowptr, const int32_t *colidx,
                     const double *vals, ffhip_mat *mat);
/* New values for an existing pattern (host pointer, nnz doubles, CSR order):
 * the numeric phase of a matrix whose HashMatrix values changed
 * (HashMatrix::GetReDoNumerics, HashMatrix.hpp:278). */
int ffhip_mat_set_values(ffhip_mat mat, const double *vals);
/* Device pointers of the CSR arrays (owned by mat). */
int ffhip_mat_device_ptrs(ffhip_mat mat, int64_t **rowptr, int32_t **colidx,
                          double **vals);
int ffhip_mat_free(ffhip_mat mat);

/* y = A x (trans = 0) or A^T x (trans = 1) on device vectors:
 * HashMatrix::addMatMul (HashMatrix.cpp:1086-1154) without the accumulate. */
int ffhip_spmv(ffhip_ctx ctx, ffhip_mat mat, const double *x, double *y,
               int trans);

/* ---- vector P2 linear elasticity (3D) ------------------------------------ */
/* Vh(Th,[P2,P2,P2]): P2 nodes numbered by first appearance over the elements
 * (4 vertices, then the 6 edges of Tet::nvedge per element), as
 * GenericMesh::BuildDFNumbering (GenericMesh.hpp:1880-1920) does; dof of
 * component c at node a = 3 a + c (FESpacen.cpp:403-428). */
int ffhip_p2v_create(ffhip_ctx ctx, ffhip_mesh mesh, ffhip_p2v *space);
int ffhip_p2v_info(ffhip_p2v space, int64_t *nnodes, int64_t *ndof);
/* element nodes (nt x 10, host): FESpace::operator()(k, i) of the P2 space */
int ffhip_p2v_download(ffhip_p2v space, int32_t *elem_nodes);
int ffhip_p2v_free(ffhip_p2v space);
/* Pattern of `matrix A = va(Vh,Vh)`: every (dof, dof) pair of each element's
 * 30 dofs (HashMatrix::operator+=, HashMatrix.cpp:1294-1332), CSR-sorted. */
int ffhip_pattern_p2v(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat *mat);
/* Values of int3d(Th)(lambda*div(u)*div(v) + 2*mu*(epsilon(u)'*epsilon(v)))
 * with Tet_5: Element_Op (problem.cpp:6337-6437) with the P2 basis of
 * TypeOfFE_P2Lagrange3d (P012_3d.cpp:199-270), accumulated in element order.
 * Local matrices on the FP64 matrix cores. */
int ffhip_assemble_p2v_elasticity(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat mat,
                                  double lambda, double mu);
/* on(labels, u1=.., u2=.., u3=..): all components of the vertex and edge nodes
 * of the labelled boundary triangles (AssembleBC, problem.cpp:10120-10164);
 * a_ii = tgv (HashMatrix::SetBC).  onbc_out (host, ndof bytes) optional. */
int ffhip_set_bc_p2v(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat mat,
                     const int32_t *labels, int nlab, double tgv,
                     uint8_t *onbc_out);
/* b = int3d(Th)(f . v) for a constant body force f[3] (Element_rhs,
 * problem.cpp:7940-7978), with b_i = tgv * g[c] on Dirichlet dofs (g may be
 * null for 0).  b: device pointer when on_device != 0, else host. */
int ffhip_rhs_p2v(ffhip_ctx ctx, ffhip_p2v space, const double f[3],
                  const int32_t *labels, int nlab, double tgv, const double g[3],
                  double *b, int on_device);

/* ---- scalar P2 on a tetrahedral mesh (Vh(Th,P2), dof = node) ----------- */
/* The same node numbering as ffhip_p2v_create (FreeFEM's BuildDFNumbering
 * for P2, GenericMesh.hpp:1880-1920).  Pattern: the node rows, bit-exact
 * with the HashMatrix of the element loop (HashMatrix.cpp:1294-1332) sorted
 * by CSR(). */
int ffhip_pattern_p2(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat *mat);
/* int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v))) with
 * TypeOfFE_P2Lagrange3d (P012_3d.cpp:199-270) and Tet_5 — Element_Op
 * (problem.cpp:6384-6437); kappa as in ffhip_assemble_p1 (QP: 14 values per
 * element). */
int ffhip_assemble_p2(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat mat, int kind, double c0,
                      const double *kappa_e, int kappa_on_device);
/* on(labels, u=0): the nodes of the labelled boundary triangles (3 vertices
 * and 3 edges) get a_ii = tgv (AssembleBC, problem.cpp:10120-10164). */
int ffhip_set_bc_p2(ffhip_ctx ctx, ffhip_p2v space, ffhip_mat mat, const int32_t *labels,
                    int nlab, double tgv, uint8_t *onbc);
/* int3d(Th)(f*v) (Element_rhs, problem.cpp:7940-7978), f constant; tgv*g on
 * the Dirichlet rows. */
int ffhip_rhs_p2(ffhip_ctx ctx, ffhip_p2v space, double f, const int32_t *labels, int nlab,
                 double tgv, double g, double *b, int on_device);

/* ---- Jacobi-preconditioned CG ------------------------------------------ */
enum { FFHIP_CG_FP64 = 0, FFHIP_CG_MIXED = 1 };
/* Layout of the streamed matrix, OR'ed into mode: sliced ELL with 64-row
 * slices (one lane per row), CSR with 16 lanes per row, or (matrices of
 * ffhip_pattern_p2v only) 3x3 blocks of the node rows.  AUTO takes the 3x3
 * blocks for a [P2,P2,P2] matrix, else sliced ELL unless its slices would pad
 * the matrix by more than 25 %, then CSR. */
enum { FFHIP_CG_LAYOUT_AUTO = 0, FFHIP_CG_LAYOUT_SELL = 0x10, FFHIP_CG_LAYOUT_CSR = 0x20,
       FFHIP_CG_LAYOUT_BSR3 = 0x100 };
/* Sliced ELL streams 16-bit column deltas (column - row) when every column is
 * within 2^15 of its row, else 32-bit columns; FFHIP_CG_INDEX32 in mode forces
 * 32-bit columns.  ffhip_cg_layout reports FFHIP_CG_INDEX16 with the layout
 * when the deltas are used. */
enum { FFHIP_CG_INDEX16 = 0x40, FFHIP_CG_INDEX32 = 0x80 };
/* The solve loops replay one captured chunk of 16 iterations (a hipGraph per
 * solver, value array and parity, captured on first use); FFHIP_CG_EAGER in
 * mode launches the kernels one by one instead (as does per-kernel
 * profiling). */
enum { FFHIP_CG_EAGER = 0x200 };
typedef struct {
  int iters;          /* CG iterations (inner FP32 iterations for MIXED)   */
  int converged;      /* ConjugueGradient's nret: 1 converged, 2 gCg0<1e-30, 0 no */
  int refinements;    /* MIXED: FP64 residual recomputations              */
  int reserved;
  double gcg0;        /* (g0, D^-1 g0), g0 = A x0 - b                       */
  double gcg;         /* final (g, D^-1 g) (recurrence for FP64, true for MIXED) */
  double true_gcg;    /* final (g, D^-1 g) recomputed in FP64 from x         */
  double true_relres; /* ||b - A x||_2 / ||b||_2 in FP64                     */
  double ms_solve;    /* device time of the solve (HIP events)              */
  /* with ffhip_cg_set_profiling(cg, 1): average device time of one launch of
   * the iteration's mat-vec kernel and of its vector-update kernel, over the
   * `timed` iterations of this solve (HIP events on the library stream) */
  double ms_matvec_avg;
  double ms_update_avg;
  int timed;
  int reserved2;
  double ms_direction_avg; /* the same for the direction kernel */
} ffhip_cg_stats;
/* Solver setup for a matrix (SolverCG ctor + HMatVirtPrecon,
 * src/femlib/VirtualSolverCG.hpp:27-69, 157-167): tgv detection
 * (HashMatrix::gettgv HashMatrix.cpp:1340-1371), Jacobi diagonal, and the
 * device layouts the kernels stream (FP64 values in the chosen layout; for
 * MIXED also the FP32 Jacobi-scaled copy).  mode = FFHIP_CG_FP64 or
 * FFHIP_CG_MIXED, optionally OR'ed with a FFHIP_CG_LAYOUT_* value. */
int ffhip_cg_create(ffhip_ctx ctx, ffhip_mat mat, int mode, ffhip_cg *cg);
/* Layout the solver chose (FFHIP_CG_LAYOUT_SELL, possibly | FFHIP_CG_INDEX16,
 * FFHIP_CG_LAYOUT_CSR or FFHIP_CG_LAYOUT_BSR3) and the number of
 * matrix entries one mat-vec streams (padding included).  No reference
 * counterpart (a property of this solver's device layout). */
int ffhip_cg_layout(ffhip_cg cg, int *layout, int64_t *entries);
/* Re-read the values of mat (same pattern) — SolverHIPCG::fac_numeric. */
int ffhip_cg_update_values(ffhip_cg cg);
/* Solve A x = b (SolverCG::dosolver, VirtualSolverCG.hpp:169-190, one rhs):
 * x holds the initial guess on entry (warm start, RNM.hpp:568-569); tgv rows
 * are initialised to b_i/tgv (SetInitWithBC); stop when
 * (g, D^-1 g) < eps^2 (g0, D^-1 g0) for eps > 0, or < eps^2 for eps < 0
 * (CG.cpp:226,247).  itmax <= 0 means n.  b, x: device pointers when
 * on_device != 0, else host.  trans != 0 solves A^T x = b (SolveT). */
int ffhip_cg_solve(ffhip_cg cg, const double *b, double *x, int on_device,
                   double eps, int itmax, double tgv, int trans,
                   ffhip_cg_stats *stats);
/* LinearCG / AffineCG whose operator is the matrix product A*x
 * (LinearCG<R>::E_LCG::operator(), src/fflib/lgfem.cpp:705-758, over
 * ConjuguedGradient2, src/femlib/MatriceCreuse.hpp:180-240), in FP64 with the
 * solver's layout whatever its mode: affine CG on g = A x - b (b NULL: the
 * zero right-hand side of AffineCG / LinearCG without b).  precon:
 * FFHIP_CG2_IDENTITY (no precon=, MatriceIdentite) or FFHIP_CG2_JACOBI
 * (C g = D^-1 g with the Jacobi diagonal of ffhip_cg_solve).  Stop when
 * (Cg, g) < eps^2 (Cg0, g0) for eps > 0, < -eps for eps < 0; the loop runs
 * iter = 0..itmax (nbiter).  *eps_out (may be NULL) receives the threshold
 * the reference writes back into eps (LinearCG stores its negation in
 * veps).  stats->converged is ConjuguedGradient2's return value, 1
 * converged, 2 (Cg0, g0) < 1e-30, 0 not converged: not converging is no
 * error (LinearCG returns 0 to the script).  FFHIP_ERR_NAN for the
 * reference's two ExecErrors ((h, Ah)^2 < 1e-60; ro NaN).  stats->iters
 * counts executed iterations (iter + 1 at convergence). */
enum { FFHIP_CG2_IDENTITY = 0, FFHIP_CG2_JACOBI = 1 };
int ffhip_cg_solve_affine(ffhip_cg cg, const double *b, double *x, int on_device,
                          double eps, int itmax, int precon, double *eps_out,
                          ffhip_cg_stats *stats);
/* solver=GMRES: SolverGMRES::dosolver (VirtualSolverCG.hpp:236-255) over
 * fgmres (CG.cpp:347-517) — restarted GMRES, right-preconditioned by the
 * solver's Jacobi diagonal (HMatVirtPrecon), modified Gram-Schmidt, in FP64
 * with the solver's layout whatever its mode.  tgv rows start at b_i/tgv
 * (SetInitWithBC); stop when |g_{k+1}| / ||b|| < |eps|, ||b|| taken with the
 * tgv rows zeroed (and 1 for eps < 0 or ||b|| < 1e-20); restart = the Krylov
 * dimension (NbSpace, dimKrylov=, 1000 by default in FreeFEM); itmax <= 0
 * means n.  stats: iters = fgmres' iteration count, refinements = restarts,
 * gcg0 = ||b||, gcg = the final |g_{k+1}| / ||b||.  FFHIP_ERR_NOCONV when it
 * does not converge (SolverGMRES, erronerr = 1).  The basis vectors are
 * allocated as a solve first reaches them and kept by the solver. */
int ffhip_gmres_solve(ffhip_cg cg, const double *b, double *x, int on_device,
                      double eps, int itmax, int restart, double tgv, int trans,
                      ffhip_cg_stats *stats);
int ffhip_cg_free(ffhip_cg cg);
/* Per-kernel HIP-event timing of the solve loop (measurement only). */
int ffhip_cg_set_profiling(ffhip_cg cg, int on);

/* ---- measurement helpers ---------------------------------------------- */
/* Average device time (ms) of `reps` back-to-back launches of the solver's
 * SpMV kernel (mode FP64 or MIXED's FP32), timed with HIP events on the
 * library stream; x, y device vectors of n doubles (FP64) / floats (FP32). */
int ffhip_bench_spmv(ffhip_cg cg, int mode, int reps, const void *x, void *y,
                     double *avg_ms);
/* Average device time (ms) of `reps` launches of a plain streaming copy of
 * nbytes (library-owned buffers): the achievable-bandwidth reference and the
 * scale for hardware byte counters. */
int ffhip_bench_copy(ffhip_ctx ctx, int64_t nbytes, int reps, double *avg_ms);
/* Device time (ms) of the last pattern / assembly / bc / rhs call. */
double ffhip_last_ms(ffhip_ctx ctx);

#ifdef __cplusplus
}
#endif
#endif
