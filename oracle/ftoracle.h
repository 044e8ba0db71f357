/*
 * ftoracle.h -- CPU oracle for fluidtissue's level-set reinitialization.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark may load it; the product path
 * (libftreinit.so) never links or calls it.
 *
 * It restates, in plain C, the reference's CPU reinit:
 *   LevelSet::findInterfaceNodes_    src/fem/LevelSet.cpp:261-322
 *   LevelSet::reconstructInterface_  src/fem/LevelSet.cpp:326-425
 *   LevelSet::resetPhi_              src/fem/LevelSet.cpp:429-455
 *   EuclideanDistance::resetPhi (CPU branch)  src/fem/EuclideanDistance.cu:108-117
 *   distanceToInterface_ (host form)          src/fem/EuclideanDistance.cu:14-35
 *   post-reset blends                         src/fem/LevelSet.cpp:18,73-74,147-148
 * plus an FP64 nearest-triangle oracle (closed-form point-triangle distance)
 * for the triangle-distance mode, which has no counterpart in the reference.
 *
 * The whole file is compiled with -ffp-contract=off: every multiply and add
 * is rounded separately, the way Eigen's expression templates evaluate the
 * reference's vector arithmetic.
 *
 * Array conventions (the reference's data contract, src/fem/TetraMesh.hpp:250-256):
 *   p      n_nodes x 3 float, AoS (Vec3f)
 *   phi    n_nodes double (VecXd)
 *   tets   n_tets x 4 int32 (Vec4i)
 *   edges  n_edges x 2 int32 (Vec2i), orientation as given by the mesher
 *   t2e    n_tets x 6 int32 (Vec6i), local order {01,02,03,12,13,23}
 */
#ifndef FTORACLE_H
#define FTORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LevelSet::findInterfaceNodes_ (LevelSet.cpp:261-300).  Writes ifp (P x 3)
 * and ifp2e (P) in increasing edge order; buffers must hold n_edges entries.
 * Returns P. */
int64_t ora_find_interface_nodes(int64_t n_edges, const int32_t* edges,
                                 const float* p, const double* phi,
                                 float* ifp, int32_t* ifp2e);

/* LevelSet::reconstructInterface_ (LevelSet.cpp:326-425).  tris is F x 4
 * ({a,b,c,-1}), if_elems the crossing tets, ifp_area P floats.  Capacities:
 * tris 2*n_tets, if_elems n_tets.  Returns 0, or -1 for a tet whose t2e
 * yields a crossing-edge count other than 3 or 4 (LevelSet.cpp:400-405), or
 * -2 for a crossing tet whose four nodes all have |phi| <= FLT_MIN (the
 * reference reads past the tet there, LevelSet.cpp:340-341).  On error
 * *n_tri / *n_elems hold what was emitted before the failing tet and
 * *bad_tet its index. */
int ora_reconstruct_interface(int64_t n_tets, const int32_t* tets,
                              const int32_t* t2e, int64_t n_edges,
                              const int32_t* edges, const float* p,
                              const double* phi, int64_t n_ifp,
                              const float* ifp, const int32_t* ifp2e,
                              int32_t* tris, int32_t* if_elems,
                              float* ifp_area, int64_t* n_tri,
                              int64_t* n_elems, int64_t* bad_tet);

/* EuclideanDistance::resetPhi, CPU branch (EuclideanDistance.cu:108-117):
 * phi_n[i] = (phi_n[i] < 0 ? -1 : 1) * min_j |ifp_j - p_i| (float, first
 * minimum wins).  nearest (nullable) receives j, or -1 when n_ifp == 0. */
void ora_reset_phi_points(int64_t n_nodes, const float* p, const float* ifp,
                          int64_t n_ifp, double* phi_n, int32_t* nearest,
                          int nthreads);

/* Same distance for a subset of nodes: out_d[k] = unsigned float distance of
 * node nodes[k] (nodes == NULL: node k), out_j[k] the nearest point. */
void ora_point_distance(int64_t n_query, const int64_t* nodes, const float* p,
                        const float* ifp, int64_t n_ifp, float* out_d,
                        int32_t* out_j, int nthreads);

/* FP64 squared distance from x to the closed triangle (a, b, c): plane
 * projection when it falls inside, else the nearest edge segment; valid for
 * degenerate (zero-area) triangles too. */
double ora_point_triangle_dist2(const double* x, const double* a,
                                const double* b, const double* c);

/* FP64 nearest-triangle oracle by brute force over all triangles.  For each
 * queried node: out_d = unsigned FP64 distance, out_f = nearest triangle
 * (lowest index among exact FP64 ties), -1 when n_tri == 0. */
void ora_triangle_distance(int64_t n_query, const int64_t* nodes,
                           const float* p, const float* ifp,
                           const int32_t* tris, int64_t n_tri, double* out_d,
                           int32_t* out_f, int nthreads);

/* FP64 distance of node `node` to one triangle f. */
double ora_node_triangle_distance(int64_t node, const float* p,
                                  const float* ifp, const int32_t* tris,
                                  int64_t f);

/* LevelSet::resetPhi_ epilogue plus the caller's blend
 * (LevelSet.cpp:438-448, 73-74, 147-148):
 *   phi_ifn = signed reset distance (input, already computed)
 *   retained = phi_ifn, except phi_old at both nodes of every crossing edge
 *   blend == 0 (init):   phi_out = phi_ifn
 *   blend == 1 (update): phi_out = w_phi * retained + w_ifn * phi_ifn   */
void ora_finish_reset(int64_t n_nodes, const double* phi_old,
                      const double* phi_ifn, int64_t n_ifp,
                      const int32_t* ifp2e, const int32_t* edges, int blend,
                      double w_phi, double w_ifn, double* retained,
                      double* phi_out);

/* Mesh3D::p2t, node -> tetrahedra (TetraMesh.cpp:84-86, 115-131): the
 * (node, tet) pairs sorted, so each node's tets in increasing id order.
 * off: n_nodes+1 entries; tet_ids: 4*n_tets entries. */
void ora_node_tets(int64_t n_nodes, int64_t n_tets, const int32_t* tets,
                   int64_t* off, int32_t* tet_ids);

/* Assembler::grad (Assembler.cpp:11-42): per tet, the P1 gradient of u from
 * B = [p1-p0, p2-p0, p3-p0] (FP64 from the float nodes), its inverse
 * transposed by cofactors / determinant (Assembler.hpp:429-459), the basis
 * gradients (-1,-1,-1), e1, e2, e3 (Tetrahedron.hpp:36-50) mapped by it,
 * then g = u0 c0 + u1 c1 + u2 c2 + u3 c3, cast to float (3 per tet).
 * Returns the first tet with det B == 0 (the reference reports it and
 * returns -1), or -1. */
int64_t ora_tet_grad(int64_t n_tets, const int32_t* tets, const float* p,
                     const double* u, float* grad_t);

/* Assembler::grad_nodes (Assembler.cpp:46-71): per node, the float mean of
 * the gradients of its tets, summed in p2t order. */
void ora_grad_nodes(int64_t n_nodes, const int64_t* off, const int32_t* tet_ids,
                    const float* grad_t, float* grad_n);

/* LevelSet::getInterfaceCurvature (LevelSet.cpp:215-253): from the node
 * normals g (grad_nodes of phi_ifn), the unit-normal components as double
 * fields, their grad_nodes, k = gg_x.x + gg_y.y + gg_z.z per node, then at
 * each crossing point (crossing-edge order) t = float(phi0 / (phi0 - phi1))
 * and k_if = float((1.0 - t) k0 + float(t k1)).  off/tet_ids: node -> tets. */
void ora_interface_curvature(int64_t n_nodes, const float* p, int64_t n_tets,
                             const int32_t* tets, const int64_t* off,
                             const int32_t* tet_ids, const float* g,
                             const double* phi, int64_t n_ifp,
                             const int32_t* ifp2e, const int32_t* edges,
                             float* k_if);

/* Flow::surfaceTension_ (src/growth/Flow.cpp:307-341), serially in
 * increasing crossing point i (the reference's OpenMP loop races on nodes
 * that two points share; serial order is the defined result): on edge
 * (e0, e1), t = float(phi0 / (phi0 - phi1)), unit normal n = g / |g| in
 * float, then fx[e0] += float(((((1.0 - t) st) (-k_i)) area_i) rho) n.x
 * (a double chain, cast to float to scale the float normal) and fx[e1] +=
 * ((((t st) (-k_i)) area_i) rho) n.x in float; likewise y and z. */
void ora_surface_tension(const float* g, const double* phi, int64_t n_ifp,
                         const int32_t* ifp2e, const int32_t* edges, const float* k_if,
                         const float* area, float st, int32_t rho, double* fx, double* fy,
                         double* fz);

/* ---- The advection step of LevelSet::update (LevelSet.cpp:89-141) ---------
 * Crank-Nicolson on the P1 Galerkin form: (M + s A) phi_n = (M - s A) phi,
 * s = 0.5 dt, solved by BiCGSTAB with a Jacobi preconditioner from the guess
 * phi (Eigen::BiCGSTAB, LevelSet.hpp:145).
 *
 * Quadrature (Quadrature.hpp:23-36, Tetrahedron.hpp:30-50): points (a,b,b),
 * (b,a,b), (b,b,a), (b,b,b), a = (5 + 3 sqrt5)/20, b = (5 - sqrt5)/20, given
 * to the float shape functions (phi_0 = 1 - x - y - z, phi_1..3 = x, y, z,
 * rounded to float); weights 0.25 / 6.  phi[4 j + q]. */
void ora_fem_quad(double* phi, double* w);

/* Assembler::bilin_with_p2t (Assembler.hpp:131-226): row i sums, over the
 * tets of i in increasing id (Mesh3D::p2t) and their local nodes j, the
 * integral int f(u = phi_j, v = phi_i) into column t(j):
 *   form 0 (mass, LevelSet.cpp:37):       f = u v
 *   form 1 (advection, :107-110, data V): f = V1 (ux v) + V2 (uy v) + V3 (uz v),
 *     V_c(q) = sum_j V[c n + t(j)] phi_j(q) accumulated from 0 in j order;
 * each integral as ((f0 w + f2 w) + (f1 w + f3 w)) |det B|, basis gradients
 * from invBt as in ora_tet_grad.  A tet with det B = 0 contributes nothing
 * (the reference leaves its inverse unset).  The pattern row_off/cols is the
 * node and its tet neighbours in increasing order (the std::map of a row).
 * Returns the first such tet or -1. */
int64_t ora_fem_assemble(int form, int64_t n, const float* p, const int32_t* tets,
                         const int64_t* n2t_off, const int32_t* n2t_ids,
                         const int64_t* row_off, const int32_t* cols, const double* vel,
                         double* vals);

/* y = A x, each row summed in column order (Eigen's row-major sparse times
 * dense vector). */
void ora_spmv(int64_t n, const int64_t* row_off, const int32_t* cols, const double* vals,
              const double* x, double* y);

/* Eigen's BiCGSTAB (H. A. van der Vorst, SIAM J. Sci. Stat. Comput. 13,
 * 1992; Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, Eigen >= 3.3, not
 * vendored) with its DiagonalPreconditioner (1 / A_ii, 1 where A_ii = 0),
 * from the guess in x: r = b - A x, r0 = r, rho = alpha = w = 1, v = p = 0;
 * while |r|^2 > tol^2 |b|^2 and i < max_iters: rho_old = rho, rho = r0.r;
 * if |rho| < eps^2 |r0|^2 restart (r = b - A x, r0 = r, rho = |r0|^2 = |r|^2,
 * i = 0 on the first restart); beta = (rho/rho_old)(alpha/w);
 * p = r + beta (p - w v); y = P p; v = A y; alpha = rho / r0.v; s = r - alpha v;
 * z = P s; t = A z; w = t.s / t.t (0 if t.t = 0); x += alpha y + w z;
 * r = s - w t.  |b| = 0 gives x = 0.  Dot products here run left to right
 * (Eigen's order is its SIMD width's: parity with it is to the tolerance).
 * error = |r| / |b|.  Returns 1 if error <= tol. */
int ora_bicgstab(int64_t n, const int64_t* row_off, const int32_t* cols, const double* vals,
                 const double* b, double* x, double tol, int64_t max_iters, int64_t* iters,
                 double* error, int64_t* restarts);

#ifdef __cplusplus
}
#endif

#endif
