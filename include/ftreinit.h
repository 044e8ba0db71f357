/*
 * ftreinit.h -- C ABI of the MI355X (gfx950) level-set reinitialization.
 *
 * Replaces the reinit hot path of fluidtissue:
 *   EuclideanDistance::init / resetPhi     src/fem/EuclideanDistance.hpp:30-38,
 *                                          src/fem/EuclideanDistance.cu:66-118
 *   LevelSet::resetPhi_ and its helpers    src/fem/LevelSet.cpp:261-455
 *   the caller blends                      src/fem/LevelSet.cpp:18,73-74,147-148
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (FT_OK = 0, negative = error class) and never throws; ft_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Array conventions follow the reference's data contract
 * (src/fem/TetraMesh.hpp:250-256, src/fem/MathTypes.hpp:27-46):
 *   xyz    n_nodes x 3 float, AoS (Vec3f)
 *   phi    n_nodes double (VecXd)
 *   tets   n_tets x 4 int32 (Vec4i)
 *   edges  n_edges x 2 int32 (Vec2i), orientation as given by the mesher
 *   t2e    n_tets x 6 int32 (Vec6i), local edge order {01,02,03,12,13,23}
 * Interface outputs follow LevelSet.hpp:133-139:
 *   ifp        P x 3 float, ordered by crossing-edge id       (ifp_)
 *   ifp2e      P int32, the crossing edge of each point         (ifp2e_)
 *   tris       F x 4 int32 {a,b,c,-1}, in tet order              (ifMesh_.t)
 *   if_elems   crossing tets in increasing order                 (ifElements_)
 *   ifp_area   P float, summed triangle areas per point          (ifpArea_)
 *   p2t        CSR: offsets (P+1) and triangle ids, ascending    (ifMesh_.p2t)
 * e2ifp_ is the inverse of ifp2e (ifp2e is strictly increasing).
 */
#ifndef FTREINIT_H
#define FTREINIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FT_OK               0
#define FT_ERR_ARG         -1  /* bad argument (null pointer, size, option)   */
#define FT_ERR_NODEVICE    -2  /* no HIP device                               */
#define FT_ERR_HIP         -3  /* HIP runtime error                           */
#define FT_ERR_NOMEM       -4  /* device allocation failed                    */
#define FT_ERR_T2E         -5  /* crossing tet whose t2e edges do not cross 3
                                  or 4 times, or disagree with its node signs
                                  (LevelSet.cpp:400-405)                      */
#define FT_ERR_DEGENERATE  -6  /* crossing tet whose four nodes all have
                                  |phi| <= FLT_MIN: the reference reads past
                                  the tet there (LevelSet.cpp:338-341)        */
#define FT_ERR_STATE       -7  /* call order (e.g. reinit before set_mesh)    */
#define FT_ERR_BUDGET      -8  /* request exceeds the device memory budget    */
#define FT_ERR_SOLVER      -9  /* the advection solve did not converge        */
#define FT_ERR_INTERNAL   -10  /* a stage's own index check failed (a fault in
                                  an earlier stage): nothing was written past
                                  the check and phi is unchanged              */

/* ---- options ----------------------------------------------------------- */
#define FT_DIST_POINTS     0   /* reference semantics: nearest crossing point
                                  (EuclideanDistance.cu:14-35)                */
#define FT_DIST_TRIANGLES  1   /* nearest marching-tets triangle              */

#define FT_BLEND_INIT      0   /* LevelSet::init:   phi = phi_ifn (:74)       */
#define FT_BLEND_UPDATE    1   /* LevelSet::update: phi = w_phi*retained +
                                  w_ifn*phi_ifn (:148, weights :18)           */

typedef struct ft_opts {
    int32_t mode;        /* FT_DIST_*                                          */
    int32_t blend;       /* FT_BLEND_*                                         */
    double  w_phi;       /* 0.75 (LevelSet.cpp:18)                             */
    double  w_ifn;       /* 0.25                                               */
    double  band;        /* > 0: narrow band; nodes with no primitive within
                            band get +-band and are counted in n_clamped.
                            0: full domain                                      */
    float   bin_scale;   /* bin edge in units of the mean node spacing; 0 =
                            the mode's default (points 2.0, triangles 1.25)     */
    int32_t want_nearest;/* 1: keep the per-node nearest primitive index       */
    int32_t query_block; /* threads per block of the distance query: 64, 128,
                            256 (waves per workgroup 1, 2, 4); 0 = default     */
    int32_t want_normals;/* 1: also the level-set normals of phi_ifn, as
                            LevelSet::resetPhi_ ends (LevelSet.cpp:451-453,
                            Assembler::grad_nodes, Assembler.cpp:11-71)        */
    int32_t waves_per_simd; /* launch bound of the triangle-mode query: 5..8
                            waves per SIMD (register budget against
                            occupancy); 0 = 8                                  */
    int32_t schedule;    /* FT_SCHED_STATIC: one tile per wave, deep tiles on
                            the first workgroups; FT_SCHED_QUEUE: persistent
                            waves taking tiles from per-XCD counters;
                            FT_SCHED_BINS (narrow band, triangles only): no
                            tree walk, each tile tests the triangles of the
                            index bins within the band, staged in LDS; the
                            same results as the tree walk                    */
    int32_t reserved[2];
} ft_opts;

#define FT_SCHED_STATIC    0
#define FT_SCHED_QUEUE     1
#define FT_SCHED_BINS      2

typedef struct ft_counts {
    int64_t n_ifp;       /* P: crossing points                                 */
    int64_t n_tri;       /* F: interface triangles                             */
    int64_t n_if_elems;  /* crossing tets                                      */
    int64_t n_clamped;   /* nodes clamped to the band                          */
    int64_t bad_tet;     /* first offending tet on FT_ERR_T2E/_DEGENERATE (the
                            triangle or point on FT_ERR_INTERNAL)              */
} ft_counts;

typedef struct ft_ctx* ft_handle;

/* Defaults: points mode, init blend, weights (0.75, 0.25), full domain. */
void        ft_opts_default(ft_opts* o);
const char* ft_last_error(void);
int         ft_device_count(int32_t* n);

/* One handle per device and host thread (EuclideanDistance is not
 * thread-safe either, EuclideanDistance.hpp:13-52). */
int ft_create(int32_t device, ft_handle* out);
int ft_destroy(ft_handle h);

/* EuclideanDistance::init (EuclideanDistance.cu:66-86): uploads the node
 * coordinates once.  phi0 is not needed (the reference's copy of it is dead
 * data, re-uploaded by resetPhi). */
int ft_set_nodes(ft_handle h, const float* xyz, int64_t n_nodes);

/* Uploads nodes and connectivity once; they stay resident in HBM
 * (Mesh3D accessors, TetraMesh.hpp:49-62). */
int ft_set_mesh(ft_handle h, const float* xyz, int64_t n_nodes,
                const int32_t* tets, int64_t n_tets, const int32_t* edges,
                int64_t n_edges, const int32_t* t2e);

/* Batch of independent fields resident together (no counterpart in the
 * reference, which holds one mesh per LevelSet).  Field k owns nodes
 * [node_off[k], node_off[k+1]), tets [tet_off[k], ...) and edges
 * [edge_off[k], ...) of the concatenated arrays; its tets, edges and t2e use
 * field-local indices.  Offsets have n_fields+1 entries starting at 0.  A
 * reinit of the batch treats every field exactly as a reinit of that field
 * alone would (same results, bit for bit), in one pass of each kernel. */
int ft_set_mesh_batch(ft_handle h, int32_t n_fields, const int64_t* node_off,
                      const float* xyz, const int64_t* tet_off,
                      const int32_t* tets, const int64_t* edge_off,
                      const int32_t* edges, const int32_t* t2e);

/* Batch of Kuhn-grid fields generated on the device, identical to
 * ft_kuhn_mesh: dims 3 per field; origins 3 per field (NULL: 0); spacings one
 * per field (NULL: 1/(max dim - 1)); interior nodes jittered as in
 * ft_kuhn_mesh with the same seed for every field. */
int ft_set_kuhn_batch(ft_handle h, int32_t n_fields, const int32_t* dims,
                      const double* origins, const double* spacings,
                      double jitter, uint64_t seed);

/* Layout of the resident fields: n_fields and the global offsets (arrays of
 * n_fields+1; any pointer may be NULL). */
int ft_batch_info(ft_handle h, int32_t* n_fields, int64_t* node_off,
                  int64_t* tet_off, int64_t* edge_off);

/* Copies the resident mesh (global indices) to host memory; NULL skips. */
int ft_get_mesh(ft_handle h, float* xyz, int32_t* tets, int32_t* edges,
                int32_t* t2e);

/* Reinit of every resident field in one pass on device-resident phi (all
 * fields concatenated, node_off layout).  per_field (nullable, n_fields
 * entries) receives each field's counts; interface outputs are the
 * concatenation of the fields' interfaces in global numbering. */
int ft_reinit_batch(ft_handle h, double* d_phi, double* d_phi_ifn,
                    const ft_opts* o, ft_counts* per_field, void* stream);

/* ft_reinit_batch as one captured HIP graph (no counterpart in the
 * reference): every kernel of the reinit of all resident fields, with no
 * host round trip between them, replayed by one hipGraphLaunch per call.
 * The first call, or one whose d_phi, d_phi_ifn, options, fields or buffers
 * differ from the captured ones, runs a counting pass over phi (reads only),
 * captures with headroom over its stage sizes and launches.  A later call
 * whose interface outgrows the headroom is caught on the device before phi
 * is written, captured again with room and launched again.  Results are
 * those of ft_reinit_batch, bit for bit.  Returns after the graph has run. */
int ft_reinit_graph(ft_handle h, double* d_phi, double* d_phi_ifn,
                    const ft_opts* o, ft_counts* per_field, void* stream);

/* Level-set normals of the last reinit with want_normals = 1 (normals_,
 * LevelSet.hpp:131): per node the mean of its tets' P1 gradients of phi_ifn,
 * summed in increasing tet id (Assembler.cpp:46-71); n_nodes x 3 float.  A
 * tet with a singular map gets a zero gradient (the reference reports it,
 * Assembler.hpp:444-455, and leaves its inverse unset). */
int ft_get_normals(ft_handle h, float* normals);
/* Device address of the same (valid until the next call on h). */
int ft_normals_device(ft_handle h, const float** normals);

/* LevelSet::getInterfaceCurvature (LevelSet.cpp:215-253) after a reinit with
 * want_normals = 1: the divergence of the unit normal field (grad_nodes of
 * its three components, Assembler.cpp:46-71) at the mesh nodes, interpolated
 * along each crossing edge with t from the reinit's (blended) phi; k_if has
 * n_ifp entries in crossing-edge order.  The reinit's phi (host: the handle's
 * device copy; device: the caller's pointer) must still hold that phi. */
int ft_interface_curvature(ft_handle h, float* k_if, void* stream);

/* Flow::surfaceTension_ (src/growth/Flow.cpp:307-341) after a reinit with
 * want_normals = 1: adds to fx, fy, fz (n_nodes doubles each, read and
 * written) the surface-tension force of each crossing point i on edge
 * (e0, e1), t = phi[e0] / (phi[e0] - phi[e1]) from the reinit's phi:
 *   e0 += (1 - t) st (-k_i) ifpArea_i rho n(e0)/|n(e0)|
 *   e1 +=  t      st (-k_i) ifpArea_i rho n(e1)/|n(e1)|
 * with k_i the curvature of ft_interface_curvature.  Terms are added in
 * increasing i, the reference's serial order (its OpenMP loop races where
 * two points share a node).  Host arrays; the _device form takes device
 * arrays and a stream (null: the handle's). */
int ft_surface_tension(ft_handle h, float st, int32_t rho, double* fx, double* fy, double* fz);
int ft_surface_tension_device(ft_handle h, float st, int32_t rho, double* fx, double* fy,
                              double* fz, void* stream);

/* ---- Advection: the step before the reinit in LevelSet::update ----------
 * (LevelSet.cpp:89-141.)  Crank-Nicolson on the P1 Galerkin form of
 * phi_t + V.grad(phi) = 0 over the handle's mesh (one field):
 *   (M + s A) phi_n = (M - s A) phi,  s = 0.5 dt,
 * M the mass matrix (assembled once per mesh, LevelSet.cpp:35-39), A the
 * advection matrix of the velocity (Assembler::assemble_bilin with
 * bilin_with_p2t, Assembler.hpp:131-226, :107-117), solved by BiCGSTAB with a
 * Jacobi preconditioner from the guess phi (Eigen::BiCGSTAB, LevelSet.hpp:145).
 * The matrices are the reference's to the bit (row order, sums in column
 * order); the solver's dot products are reduced in a fixed tree, so phi_n
 * agrees with a CPU solve to the solver tolerance.
 * flow: 4 n_nodes doubles, velocity x, y, z then pressure (Flow's layout).
 * tol <= 0 and max_iters <= 0 take Eigen's defaults (machine epsilon, 2 n).
 * dt < FLT_MIN or |flow| < FLT_MIN: nothing is done (skipped = 1).  A NaN in
 * the velocity is FT_ERR_ARG.  If the solve stops above tol, FT_ERR_SOLVER
 * and phi is left as it was (the reference returns -1 before its swap). */
typedef struct ft_advect_info {
    int32_t iterations;   /* solver_.iterations()                          */
    int32_t restarts;     /* BiCGSTAB restarts (r0 too orthogonal to r)     */
    double  error;        /* solver_.error(): |b - L phi_n| / |b|           */
    int32_t converged;    /* error <= tol                                   */
    int32_t skipped;      /* dt or |flow| below FLT_MIN: phi untouched      */
} ft_advect_info;
int ft_advect(ft_handle h, double* phi, const double* flow, double dt, double tol,
              int32_t max_iters, ft_advect_info* info);
/* The same on device arrays, on stream (null: the handle's). */
int ft_advect_device(ft_handle h, double* d_phi, const double* d_flow, double dt, double tol,
                     int32_t max_iters, ft_advect_info* info, void* stream);
/* The assembled mass (form 0) or advection (form 1, vel = 3 n_nodes doubles,
 * host) matrix in CSR, rows in node order, columns increasing: row_off
 * (n_nodes + 1) is always written; with cols == NULL only the sizes are
 * (nnz = row_off[n_nodes]).  Assembler::assemble_bilin of LevelSet's forms. */
int ft_fem_assemble(ft_handle h, int32_t form, const double* vel, int64_t* row_off, int32_t* cols,
                    double* vals);

/* Totals of the last successful reinit on h (sizes for ft_get_interface). */
int ft_get_counts(ft_handle h, ft_counts* total);

/* Wavefront OBJ of the last reinit's interface: fname.obj with the vertices,
 * vertex normals (per point, the sum over its triangles of (p1-p0)x(p2-p0),
 * divided by its norm when unit_normals) and 1-based faces, and fname.mtl
 * with the materials table; the INTERFACE output of cfish::
 * Write_Wavefront_obj (src/leveque/WriteData.cpp:9-61, :89-118; raw normals
 * as src/growth/WriteData.cpp:48-82), text identical to its iostream output. */
int ft_write_interface_obj(ft_handle h, const char* fname, int32_t unit_normals);

/* The parts of cfish::Write_Wavefront_obj (src/leveque/WriteData.hpp:14,
 * the reference's INTERFACE and BOUNDARY flags). */
#define FT_OBJ_INTERFACE 0x02
#define FT_OBJ_BOUNDARY 0x04

/* Wavefront OBJ with the parts named in `parts` (src/leveque/WriteData.cpp:
 * 9-87): the interface as ft_write_interface_obj, then the domain boundary:
 * the nodes of b_tris (n_btris x 3 node ids of the resident mesh, the
 * reference's mesh.b_tris()) in increasing id, and the triangles in the
 * order given, indexed after the interface's points.  The boundary needs a
 * single resident mesh; the interface part needs a reinit first.  Then the
 * materials file, as ft_write_interface_obj. */
int ft_write_obj(ft_handle h, const char* fname, int32_t parts, int32_t unit_normals,
                 const int32_t* b_tris, int64_t n_btris);

/* Captures and launches of ft_reinit_graph on h so far (nullable). */
int ft_graph_stats(ft_handle h, int64_t* captures, int64_t* launches);

/* Resident meshes (no counterpart in the reference): a mesh uploaded or
 * generated once, node-sorted once, kept in HBM and shared by any number of
 * batches on its device.  Ids are process-wide.  The Kuhn form is generated
 * on the device exactly as ft_kuhn_mesh. */
int ft_mesh_create_kuhn(int32_t device, int32_t nx, int32_t ny, int32_t nz,
                        double jitter, uint64_t seed, int32_t* mesh_id);
int ft_mesh_create(int32_t device, const float* xyz, int64_t n_nodes,
                   const int32_t* tets, int64_t n_tets, const int32_t* edges,
                   int64_t n_edges, const int32_t* t2e, int32_t* mesh_id);
int ft_mesh_info(int32_t mesh_id, int64_t* n_nodes, int64_t* n_tets,
                 int64_t* n_edges, int64_t* bytes);
int ft_mesh_release(int32_t mesh_id);

/* Makes the handle's batch the fields of the given resident meshes, in order
 * (a mesh may appear more than once); device copies only, skipped when the
 * list names the very meshes of the current batch (a released id that now
 * names a new mesh is assembled afresh). */
int ft_set_batch_meshes(ft_handle h, int32_t n_fields, const int32_t* mesh_ids);

/* Per-field counts of the last reinit (n_fields entries). */
int ft_get_field_counts(ft_handle h, ft_counts* per_field);

/* EuclideanDistance::resetPhi (EuclideanDistance.cu:90-118), host pointers:
 * phi_n[i] = sign(phi_n[i] < 0) * min_j |ifp_j - p_i| in float, bit-equal to
 * the reference's CPU branch.  n_ifp == 0 gives +-FLT_MAX. */
int ft_reset_phi_points(ft_handle h, const float* ifp, int64_t n_ifp,
                        double* phi_n);

/* LevelSet::resetPhi_ plus the caller's blend, host phi (LevelSet.cpp:429-455,
 * 73-74, 147-148).  phi: in = phi_, out = blended phi_.  phi_ifn (nullable):
 * the full reset (phi_ifn_).  Interface outputs stay on the device; fetch
 * them with ft_get_interface. */
int ft_reinit(ft_handle h, double* phi, double* phi_ifn, const ft_opts* o,
              ft_counts* counts);

/* Same, on device-resident phi (device pointers on the handle's device).
 * stream: a hipStream_t, or NULL for the handle's own stream, which is
 * ordered after the caller's earlier work on the null stream (stream 0, the
 * same value as NULL: a caller that filled phi on the null stream may pass
 * NULL) and before its later work there. */
int ft_reinit_device(ft_handle h, double* d_phi, double* d_phi_ifn,
                     const ft_opts* o, ft_counts* counts, void* stream);

/* ft_reinit_device without allocating (no counterpart in the reference; for
 * callers that cannot allocate device memory: tight serving loops, code that
 * must not call hipMalloc).  Every per-call workspace (bitmaps, interface,
 * bin index, tree, scan scratch) is taken from one file-scope device array of
 * FT_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its load to
 * the end of the process; the handle's resident mesh is set up beforehand as
 * usual.  Results are those of ft_reinit_device, bit for bit.  A reinit whose
 * workspace exceeds the scratch fails with FT_ERR_BUDGET before phi is
 * written (up to about 128^3 nodes fits, depending on the interface).
 * want_nearest and want_normals must be 0, and ft_get_interface has nothing
 * after this call: those outputs would live in the scratch.  Calls that
 * overlap on one GPU (from several threads or streams) share that scratch and
 * race; the call returns after its work is done. */
#define FT_NOALLOC_SCRATCH_BYTES (128u << 20)
int ft_reinit_device_noalloc(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                             ft_counts* counts, void* stream);

/* Copies the interface of the last reinit to host memory.  Any pointer may
 * be NULL; sizes come from ft_counts (p2t_offsets: n_ifp+1, p2t_tris:
 * 3*n_tri). */
int ft_get_interface(ft_handle h, float* ifp, int32_t* ifp2e, int32_t* tris,
                     int32_t* if_elems, float* ifp_area, int32_t* p2t_offsets,
                     int32_t* p2t_tris);

/* Per-node nearest primitive of the last reinit with want_nearest = 1 (point
 * index in points mode, triangle index in triangles mode; -1 when none). */
int ft_get_nearest(ft_handle h, int32_t* nearest);

/* Device address of the last reinit's interface buffers (read-only, valid
 * until the next call on h). Any pointer may be NULL. */
int ft_interface_device(ft_handle h, const float** ifp, const int32_t** tris,
                        const int32_t** ifp2e);

/* Phase timing with HIP events on the reinit's stream (no counterpart in the
 * reference, whose only timing is tic/toc around each phase,
 * LevelSet.cpp:272-274,360-362,434-436).  on = 1 enables and clears it.
 * ft_get_timing: cumulative milliseconds of [0] marching tets, [1] bin
 * index build, [2] distance query kernel (with the FP64 settle of triangle
 * mode and the epilogue), [3] level-set normals when requested (else about
 * 0), over `calls` reinits; phase_ms has 4 entries. */
int ft_set_timing(ft_handle h, int32_t on);
int ft_get_timing(ft_handle h, double* phase_ms, int64_t* calls);

/* ---- synthetic inputs (no counterpart in the reference; the reference
 * meshes with CGAL, TetraMesh.cpp:23-165) ------------------------------- */

/* Counts of an nx*ny*nz-node grid cut into 6 Kuhn tets per cell. */
int ft_kuhn_counts(int32_t nx, int32_t ny, int32_t nz, int64_t* n_nodes,
                   int64_t* n_tets, int64_t* n_edges);

/* Kuhn mesh: node (i,j,k) -> index (i*ny + j)*nz + k (TetraMesh.cpp:199-211),
 * position origin + (i,j,k)*spacing in FP64 rounded to float; interior nodes
 * moved by U(-jitter, jitter)*spacing per axis from a splitmix64 stream of
 * `seed` (jitter = 0: none).  Each cell's 6 tets run 000 -> 111 along the 6
 * axis orders; edges are oriented low -> high node and numbered by
 * (low node, direction); t2e in local order {01,02,03,12,13,23}. */
int ft_kuhn_mesh(int32_t nx, int32_t ny, int32_t nz, const double* origin,
                 const double* spacing, double jitter, uint64_t seed,
                 float* xyz, int32_t* tets, int32_t* edges, int32_t* t2e);

/* Level-set fields of the benchmark workloads, FP64 from the float node
 * coordinates:
 *   FT_PHI_SPHERE_WAVE  (|x-c|-0.15)*(1+0.5 sin4pix sin4piy sin4piz), c=0.35
 *   FT_PHI_STAR         (|x-c|-r(theta,phi))*2.3, r=0.3+0.06cos5phi sin^3theta
 *   FT_PHI_SPHERES      1.5*min over 64 spheres (centres U[0.1,0.9]^3,
 *                       radii U[0.02,0.08]) of |x-c_i|-r_i
 *   FT_PHI_ELLIPSOIDS   min over 1..16 ellipsoids (axes U[0.03,0.2], random
 *                       rotations, centres U[0.25,0.75]^3)
 * Seeded with splitmix64. */
#define FT_PHI_SPHERE_WAVE  1
#define FT_PHI_STAR         2
#define FT_PHI_SPHERES      3
#define FT_PHI_ELLIPSOIDS   4
int ft_phi_field(int32_t kind, int64_t n_nodes, const float* xyz,
                 uint64_t seed, double* phi);

#ifdef __cplusplus
}
#endif

#endif /* FTREINIT_H */
