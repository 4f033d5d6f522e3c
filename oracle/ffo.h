/*
 * ffo.h — CPU restatement ("oracle") of FreeFEM's assembly + Jacobi-CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the tests compare against: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * to compare against or to time as the CPU baseline.  The product (libffhip)
 * never links, loads or calls it.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the FreeFEM source tree).  Arithmetic order follows the reference wherever the
 * reference fixes it (element loop order, quadrature loop, term loop, HashMatrix
 * accumulation order), so values agree with the compiled reference subset in
 * oracle/_ref bit for bit where both are evaluated the same way.
 */
#ifndef FFO_H
#define FFO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- meshes ------------------------------------------------------------ */
/* cube(nx,ny,nz): msh3.cpp:7879-8132 (BuildCube, kind=6), labels as Cube_Op
 * 8134-8164.  out = {nv, nt, nbe}. */
void ffo_cube_counts(int nx, int ny, int nz, int64_t out[3]);
int ffo_cube(int nx, int ny, int nz, const int32_t labels[6], double *xyz,
             int32_t *tets, int32_t *btri, int32_t *blab);
/* square(nx,ny): lgmesh.cpp:1229-1384 (Carre_, flags=0).  out = {nv, nt, nbe}. */
void ffo_square_counts(int nx, int ny, int64_t out[3]);
int ffo_square(int nx, int ny, const int32_t labels[4], double *xy,
               int32_t *tris, int32_t *bedges, int32_t *blab);

/* ---- quadrature --------------------------------------------------------- */
/* QuadratureFormular_Tet_5 (QuadratureFormular.cpp:725-743): 14 points,
 * barycentric (x,y,z) and weight (sum 1). */
int ffo_qf_tet5(double *pts, double *w);
/* QuadratureFormular_T_5 (QuadratureFormular.cpp:172-186): 7 points. */
int ffo_qf_tri5(double *pts, double *w);

/* ---- geometry ----------------------------------------------------------- */
/* Tet mesure: Mesh3dn.hpp:65-71 with R3.hpp:90-106 det(). */
double ffo_tet_mesure(const double *v0, const double *v1, const double *v2,
                      const double *v3);
/* Tet::Gradlambda: Mesh3dn.hpp:126-137.  v = 4 vertices xyz, G = 4 gradients */
void ffo_tet_gradlambda(const double *v, double mes, double *G);

/* ---- sparsity pattern (HashMatrix CSR view) ----------------------------- */
/* The pattern of HashMatrix after `A += mate` over every element
 * (HashMatrix.cpp:1294-1332 inserts every (ni[i], nj[j]) pair) sorted by
 * Sortij/Buildp (HashMatrix.cpp:670-682, 992-1030).  elems = nel x nloc global
 * dof ids.  Returns nnz (count pass fills rowptr). */
int64_t ffo_pattern_rowptr(int n, int64_t nel, int nloc, const int32_t *elems,
                           int64_t *rowptr);
int ffo_pattern_colidx(int n, int64_t nel, int nloc, const int32_t *elems,
                       const int64_t *rowptr, int32_t *colidx);
/* All (i,j) pairs sorted + deduplicated — the brute-force cross-check. */
int64_t ffo_pattern_bruteforce(int n, int64_t nel, int nloc,
                               const int32_t *elems, int64_t *rowptr,
                               int32_t *colidx);

/* ---- P1 assembly -------------------------------------------------------- */
enum { FFO_COEF_CONST = 0, FFO_COEF_SIN3 = 1, FFO_COEF_P0 = 2, FFO_COEF_QP = 3 };
/* A P1 function (vertex values u) at the Tet_5 / T_5 points of every element,
 * FreeFEM's P1 evaluation order (P012_3d.cpp:98-106; FESpace.cpp:1078-1085):
 * out[e * nq + q]; returns nq. */
int ffo_p1_at_qp(int dim, int64_t nt, const int32_t *elems, const double *u, double *out);
/* Coefficient kappa(x,y,z): CONST -> c0; SIN3 -> 1+0.5*sin(2 pi x)sin(2 pi y)
 * sin(2 pi z) evaluated as the script expression; P0 -> kappa_e[element]. */
/* Element_Op full 3D (problem.cpp:6337-6437) for
 * int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v))), P1 (P012_3d.cpp:123-175)
 * with QF Tet_5. Ke row-major 4x4, row = test function. */
int ffo_p1_element_3d(const double *xyz, const int32_t *tet, int kind,
                      double c0, double kappa_e, double *Ke);
/* AssembleBilinearForm 3D element loop (problem.cpp:1398-1405) + HashMatrix +=
 * (element order, HashMatrix.cpp:1294-1332) into a CSR pattern. */
int ffo_assemble_p1_3d(int nv, int64_t nt, const double *xyz,
                       const int32_t *tets, int kind, double c0,
                       const double *kappa_e, const int64_t *rowptr,
                       const int32_t *colidx, double *vals);
/* 2D: Element_Op 2D (problem.cpp:6063) with P1 triangles, QF T_5. */
int ffo_p1_element_2d(const double *xy, const int32_t *tri, int kind,
                      double c0, double kappa_e, double *Ke);
int ffo_assemble_p1_2d(int nv, int64_t nt, const double *xy,
                       const int32_t *tris, int kind, double c0,
                       const double *kappa_e, const int64_t *rowptr,
                       const int32_t *colidx, double *vals);
/* int3d(Th)(kappa*(dx(u)*dx(v)+dy(u)*dy(v)+dz(u)*dz(v)) + m*u*v) (int2d in
 * 2D): the Element_Op loop with the op_id x op_id term after the derivative
 * terms at each quadrature point.  mkind < 0: no mass term (the functions
 * above); else m(x) like kappa (CONST m0, SIN3, P0 mass_e[element]). */
int ffo_p1_form_element_3d(const double *xyz, const int32_t *tet, int kind, double c0,
                           double kappa_e, int mkind, double m0, double mass_e, double *Ke);
int ffo_p1_form_element_2d(const double *xy, const int32_t *tri, int kind, double c0,
                           double kappa_e, int mkind, double m0, double mass_e, double *Ke);
int ffo_assemble_p1_form_3d(int nv, int64_t nt, const double *xyz, const int32_t *tets, int kind,
                            double c0, const double *kappa_e, int mkind, double m0,
                            const double *mass_e, const int64_t *rowptr, const int32_t *colidx,
                            double *vals);
int ffo_assemble_p1_form_2d(int nv, int64_t nt, const double *xy, const int32_t *tris, int kind,
                            double c0, const double *kappa_e, int mkind, double m0,
                            const double *mass_e, const int64_t *rowptr, const int32_t *colidx,
                            double *vals);
/* Element_rhs 3D (problem.cpp:7917-7985) for int3d(Th)(f*v), f constant,
 * accumulated into b in element order. */
int ffo_rhs_p1_3d(int nv, int64_t nt, const double *xyz, const int32_t *tets,
                  double f, double *b);
int ffo_rhs_p1_2d(int nv, int64_t nt, const double *xy, const int32_t *tris,
                  double f, double *b);

/* ---- adjoint building blocks (3D P1) ---------------------------------- */
/* mass varf int3d(Th)(u*v): Element_Op (problem.cpp:6384-6437) with op_id
 * terms, P1 values l_i = barycentrics, Tet_5; accumulated like
 * ffo_assemble_p1_3d. */
int ffo_assemble_p1_mass_3d(int nv, int64_t nt, const double *xyz,
                            const int32_t *tets, const int64_t *rowptr,
                            const int32_t *colidx, double *vals);
/* 1/2 int3d(Th)((u-ud)^2) with Tet_5 (the script expression's integral). */
double ffo_p1_misfit_3d(int64_t nt, const double *xyz, const int32_t *tets,
                        const double *u, const double *ud);
/* P0 linear form int3d(Th)(-(dx(u)*dx(l)+dy(u)*dy(l)+dz(u)*dz(l))*w):
 * Element_rhs (problem.cpp:7940-7978) with the P0 basis (w = 1 on K,
 * P012_3d.cpp:46-80), one value per element. */
int ffo_p0_grad_3d(int64_t nt, const double *xyz, const int32_t *tets,
                   const double *u, const double *lam, double *g);

/* ---- P2 and vector P2 (3D) --------------------------------------------- */
/* P2 nodes by first appearance over the elements: 4 vertex keys then the 6
 * edge keys (a,b) of Tet::nvedge {01,02,03,12,13,23} per element
 * (GenericMesh::BuildDFNumbering, GenericMesh.hpp:1880-1920).  elem_nodes =
 * nt x 10; returns the number of nodes. */
int64_t ffo_p2_nodes(int nv, int64_t nt, const int32_t *tets, int32_t *elem_nodes);
/* P2 basis at a reference point (P012_3d.cpp:199-270): f[10], df[10*3]
 * given the 4 P1 gradients G[12]. */
void ffo_p2_basis(const double *phat, const double *G, double *f, double *df);
/* Lame element matrix of int3d(lambda*div(u)div(v) + 2mu eps(u):eps(v)) with
 * Tet_5, vector [P2,P2,P2] with local dofs comp-blocked (c*10+i,
 * FESpacen.cpp:403-428): Ke 30x30 row-major (row = test function). */
int ffo_p2_elast_element(const double *xyz, const int32_t *tet, double lambda,
                         double mu, double *Ke);
/* element dofs of [P2,P2,P2]: dof(c*10+i) = 3*node_i + c (FESpacen.hpp:582) */
int ffo_p2v_elem_dofs(int64_t nt, const int32_t *elem_nodes, int32_t *elem_dofs);
/* assembly of the Lame matrix into a CSR pattern of the 30-dof elements */
int ffo_assemble_p2_elast(int64_t nt, const double *xyz, const int32_t *tets,
                          const int32_t *elem_dofs, double lambda, double mu,
                          int64_t ndof, const int64_t *rowptr,
                          const int32_t *colidx, double *vals);
/* int3d(f . v), constant body force f[3] (Element_rhs, problem.cpp:7940-7978) */
int ffo_rhs_p2_elast(int64_t nt, const double *xyz, const int32_t *tets,
                     const int32_t *elem_dofs, const double *f, int64_t ndof,
                     double *b);
/* Dirichlet dofs of boundary triangles with listed labels: their 3 vertex
 * nodes and 3 edge nodes, all components (AssembleBC, problem.cpp:10120-10164). */
int ffo_p2v_mark_bc(int nv, int64_t nt, const int32_t *tets,
                    const int32_t *elem_nodes, int64_t nbe, const int32_t *btri,
                    const int32_t *blab, int nlab, const int32_t *labs,
                    int64_t ndof, uint8_t *onbc);

/* ---- Dirichlet (tgv) ---------------------------------------------------- */
/* AssembleBC (problem.cpp:10039-10194): dofs of boundary faces whose label is
 * in labs[] are marked (P1: the face vertices). */
int ffo_mark_bc(int n, int64_t nbf, int nvf, const int32_t *bfaces,
                const int32_t *blab, int nlab, const int32_t *labs,
                uint8_t *onbc);
/* HashMatrix::SetBC with tgv>0 (HashMatrix.cpp:1194-1238): a_ii = tgv; and
 * the right-hand side B[ddf] = tgv*g (problem.cpp:10176). */
int ffo_set_bc(int n, const int64_t *rowptr, const int32_t *colidx,
               double *vals, const uint8_t *onbc, double tgv);
int ffo_rhs_bc(int n, const uint8_t *onbc, double tgv, double g, double *b);

/* ---- SpMV and CG -------------------------------------------------------- */
void ffo_spmv(int n, const int64_t *rowptr, const int32_t *colidx,
              const double *vals, const double *x, double *y);
/* y = A x of a half (lower-triangle) matrix, HashMatrix.cpp:1128-1136 */
void ffo_spmv_half(int n, const int64_t *rowptr, const int32_t *colidx,
                   const double *vals, const double *x, double *y);
void ffo_spmv_t(int n, const int64_t *rowptr, const int32_t *colidx,
                const double *vals, const double *x, double *y);
/* HashMatrix::gettgv (HashMatrix.cpp:1340-1371), ratio 1e6. */
double ffo_gettgv(int n, const int64_t *rowptr, const int32_t *colidx,
                  const double *vals, int *ntgv);
/* SolverCG::dosolver + HMatVirtPrecon (VirtualSolverCG.hpp:13-192) +
 * ConjugueGradient (CG.cpp:195-265).  x holds the initial guess on entry.
 * stats = {iterations, gCg0, gCg_final, converged(0/1/2)}.  trans != 0 solves
 * with A^T (HMatVirt(A,trans)).  Returns the reference's nret. */
int ffo_pcg(int n, const int64_t *rowptr, const int32_t *colidx,
            const double *vals, const double *b, double *x, double eps,
            int itmax, double tgv, int trans, double *stats);
/* LinearCG / AffineCG (lgfem.cpp:705-758) with the matrix product as the
 * operator, i.e. ConjuguedGradient2 (MatriceCreuse.hpp:180-240) on
 * g = A x - b (b NULL: AffineCG's zero right-hand side).  precon 0: identity
 * (MatriceIdentite), 1: Jacobi, diagonal inverse (1 where a_ii^2 < 1e-60).
 * *eps in: eps > 0 relative, < 0 absolute threshold on (Cg, g); out: the
 * threshold reps2 (unchanged when the start is already converged).  The loop
 * runs iter = 0..itmax.  stats = {iterations, g2_0, g2_final, nret}.
 * Returns ConjuguedGradient2's value (1 converged, 2 g2_0 < 1e-30, 0 not
 * converged) or -1 for its ExecError "Matrix is not defined (/0)" and -2 for
 * "ro is NaN". */
int ffo_cg2(int n, const int64_t *rowptr, const int32_t *colidx,
            const double *vals, const double *b, double *x, double *eps,
            int itmax, int precon, double *stats);
/* SolverGMRES::dosolver (VirtualSolverCG.hpp:236-255) + HMatVirtPrecon
 * (Jacobi, tgv rows, SetInitWithBC) + fgmres (CG.cpp:347-517): restarted
 * right-preconditioned GMRES with modified Gram-Schmidt, restart = NbSpace,
 * stop when |g_{it+1}| / ||b (tgv rows zeroed)|| < |eps|.  itmax <= 0 means n.
 * stats = {iterations (fgmres' nbitermx out), relres/normb, normb, restarts}.
 * Returns 1 converged, 0 not. */
int ffo_fgmres(int n, const int64_t *rowptr, const int32_t *colidx, const double *vals,
               const double *b, double *x, double eps, int itmax, int restart, double tgv,
               double *stats);

/* int3d(Th)(f*v) (int2d in 2D): Element_rhs with a coefficient f (CONST f0,
 * SIN3, P0 fcoef[element], QP fcoef[element * nq + q]); b zeroed first. */
int ffo_rhs_p1_coef(int dim, int nv, int64_t nt, const double *X, const int32_t *elems, int kind,
                    double f0, const double *fcoef, double *b);

/* ---- boundary integrals on P1 (ffo_bnd.c) ------------------------------ */
/* QF_GaussLegendre3 points (on [0,1]) and weights, reference order. */
int ffo_qf_gl3(double *x, double *w);
/* int2d(Th3, labels)(alpha*u*v) added to vals (when vals != NULL) and
 * int2d(Th3, labels)(g*v) added to b (when b != NULL), int1d on 2D meshes
 * (dim 2: X = xy, elems = triangles, belems = edges), over the boundary
 * elements whose label is listed.  Coefficient: CONST c0, SIN3, P0
 * coef[boundary element], QP coef[f * nq + q] at the points of boundary
 * element f's own rule (vertices in mesh order; nq = 7 on faces, 3 on edges).
 * Returns 0, -3 (slot missing), -4 (no element holds a boundary element),
 * -5 (QP point not found). */
int ffo_p1_boundary(int dim, int nv, int64_t nel, const double *X, const int32_t *elems,
                    int64_t nbe, const int32_t *belems, const int32_t *blab, int nlab,
                    const int32_t *labs, int kind, double c0, const double *coef,
                    const int64_t *rowptr, const int32_t *colidx, double *vals, double *b);

/* ---- scalar P2 (Vh(Th,P2), dof = node) --------------------------------- */
/* the coefficient of the P1/P2 forms at physical point P (CONST c0, SIN3, P0 ke) */
double ffo_coef_eval(int kind, double c0, double ke, const double *P, int dim);
/* Element_Op for int3d(Th)(kappa*grad(u)'*grad(v)), P2 FB, Tet_5; Ke 10x10
 * row-major, row = test function */
int ffo_p2_lap_element(const double *xyz, const int32_t *tet, int kind, double c0,
                       double kappa_e, double *Ke);
int ffo_assemble_p2_lap(int64_t nt, const double *xyz, const int32_t *tets,
                        const int32_t *elem_nodes, int kind, double c0, const double *kappa_e,
                        int64_t nn, const int64_t *rowptr, const int32_t *colidx, double *vals);
int ffo_rhs_p2(int64_t nt, const double *xyz, const int32_t *tets, const int32_t *elem_nodes,
               double fv, int64_t nn, double *b);
int ffo_p2_mark_bc(int nv, int64_t nt, const int32_t *tets, const int32_t *elem_nodes,
                   int64_t nbe, const int32_t *btri, const int32_t *blab, int nlab,
                   const int32_t *labs, int ncomp, int64_t ndof, uint8_t *onbc);

/* ---- all-core variant of the 3D P1 path (ffo_omp.c, libffo_omp.so) ----- */
struct ffo_omp_sys;
int ffo_omp_max_threads(void);
/* pattern + element-order values, bit-identical to ffo_pattern_* and
 * ffo_assemble_p1_3d; NULL on failure */
struct ffo_omp_sys *ffo_omp_assemble_p1_3d(int nv, int64_t nt, const double *xyz,
                                           const int32_t *tets, int kind, double c0,
                                           const double *kappa_e);
int64_t ffo_omp_nnz(const struct ffo_omp_sys *S);
void ffo_omp_arrays(const struct ffo_omp_sys *S, int64_t *rowptr, int32_t *colidx, double *vals);
void ffo_omp_setbc(struct ffo_omp_sys *S, const uint8_t *onbc, double tgv);
int ffo_omp_pcg(const struct ffo_omp_sys *S, const double *b, double *x, double eps, int itmax,
                double tgv, double *stats);
void ffo_omp_free(struct ffo_omp_sys *S);

#ifdef __cplusplus
}
#endif
#endif
