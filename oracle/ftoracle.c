/*
 * ftoracle.c -- CPU oracle for fluidtissue's level-set reinitialization.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.  See ftoracle.h
 * for the reference functions restated here.  Build with -ffp-contract=off
 * (oracle/Makefile does): each product and sum below is rounded on its own,
 * which is how the reference's Eigen expressions evaluate.
 *
 * Evaluation-order notes (the bits depend on them):
 *  - Eigen reduces a 3-vector sum (dot, squaredNorm) as x0 + (x1 + x2): its
 *    unrolled reducer splits a length-3 range into [0] and [1,2].
 *  - Eigen's cross product evaluates each component as a*b - c*d in one
 *    source expression.  Whether the reference's clang fused that into an FMA
 *    depends on its version and -march; the oracle (and the device code)
 *    define it unfused.  Parity with the original binary is unpinned on this
 *    point; the test suite pins everything else.
 *  - distanceToInterface_ (EuclideanDistance.cu:28) is plain C++:
 *    sqrt((dx*dx + dy*dy) + dz*dz) in float, left to right.
 */
#include "ftoracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int sb(double x) { return signbit(x) ? 1 : 0; }

/* ---------------------------------------------------------------------- */
/* LevelSet::findInterfaceNodes_  (src/fem/LevelSet.cpp:261-300)           */
/* ---------------------------------------------------------------------- */
int64_t ora_find_interface_nodes(int64_t n_edges, const int32_t* edges,
                                 const float* p, const double* phi,
                                 float* ifp, int32_t* ifp2e)
{
    int64_t np = 0;
    for (int64_t i = 0; i < n_edges; i++) {           /* :278 edge order */
        const int32_t a = edges[2 * i + 0];
        const int32_t b = edges[2 * i + 1];
        if (sb(phi[a]) == sb(phi[b]))                  /* :281 */
            continue;
        double t = 0.5;                                /* :287-290 */
        const double div = phi[a] - phi[b];
        if (fabs(div) > (double)FLT_MIN)
            t = phi[a] / div;
        for (int c = 0; c < 3; c++) {                  /* :292-294, 297 */
            const double p0 = (double)p[3 * (int64_t)a + c];
            const double p1 = (double)p[3 * (int64_t)b + c];
            const double v = (1.0 - t) * p0 + t * p1;
            ifp[3 * np + c] = (float)v;
        }
        ifp2e[np] = (int32_t)i;                        /* :298 */
        np++;                                          /* e2ifp_[i] = np-1 (:299) */
    }
    return np;
}

/* ---------------------------------------------------------------------- */
/* LevelSet::reconstructInterface_  (src/fem/LevelSet.cpp:326-425)         */
/* ---------------------------------------------------------------------- */

/* addTriangle lambda (LevelSet.cpp:330-357). Returns 0, or -2 when no tet
 * node has |phi| > FLT_MIN. */
static int add_triangle(const int32_t n[3], const int32_t* tet,
                        const float* p, const double* phi, const float* ifp,
                        int32_t* tris, int64_t* n_tri, float* ifp_area)
{
    const float* q0 = ifp + 3 * (int64_t)n[0];
    const float* q1 = ifp + 3 * (int64_t)n[1];
    const float* q2 = ifp + 3 * (int64_t)n[2];
    float e1[3], e2[3], nrm[3];
    for (int c = 0; c < 3; c++) {                      /* :332-333 */
        e1[c] = q1[c] - q0[c];
        e2[c] = q2[c] - q0[c];
    }
    nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];            /* :334 Eigen cross */
    nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];

    int k = 0;                                         /* :338-341 */
    while (k < 4 && fabs(phi[tet[k]]) <= (double)FLT_MIN)
        k++;
    if (k == 4)
        return -2;
    const int32_t pref = tet[k];

    float er[3];                                       /* :343 */
    for (int c = 0; c < 3; c++)
        er[c] = p[3 * (int64_t)pref + c] - q0[c];
    const float dot = nrm[0] * er[0] + (nrm[1] * er[1] + nrm[2] * er[2]);

    int32_t* t = tris + 4 * (*n_tri);
    t[0] = n[0];
    if ((signbit(dot) ? 1 : 0) != sb(phi[pref])) {     /* :344-347 */
        t[1] = n[2];
        t[2] = n[1];
    } else {
        t[1] = n[1];
        t[2] = n[2];
    }
    t[3] = -1;
    (*n_tri)++;

    const float sq = nrm[0] * nrm[0] + (nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const float norm = sqrtf(sq);                      /* normal.norm() */
    for (int j = 0; j < 3; j++)                        /* :356 float += double */
        ifp_area[n[j]] = (float)((double)ifp_area[n[j]] + 0.5 * (double)norm);
    return 0;
}

int ora_reconstruct_interface(int64_t n_tets, const int32_t* tets,
                              const int32_t* t2e, int64_t n_edges,
                              const int32_t* edges, const float* p,
                              const double* phi, int64_t n_ifp,
                              const float* ifp, const int32_t* ifp2e,
                              int32_t* tris, int32_t* if_elems,
                              float* ifp_area, int64_t* n_tri,
                              int64_t* n_elems, int64_t* bad_tet)
{
    /* e2ifp_ is a std::map edge -> ifp index; a dense table gives the same
     * lookups (every crossing edge was inserted by findInterfaceNodes_). */
    int32_t* e2ifp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_edges > 0 ? n_edges : 1));
    for (int64_t e = 0; e < n_edges; e++)
        e2ifp[e] = -1;
    for (int64_t i = 0; i < n_ifp; i++)
        e2ifp[ifp2e[i]] = (int32_t)i;
    for (int64_t i = 0; i < n_ifp; i++)                /* :366 */
        ifp_area[i] = 0.0f;

    *n_tri = 0;
    *n_elems = 0;
    *bad_tet = -1;
    int rc = 0;
    for (int64_t i = 0; i < n_tets; i++) {             /* :377 tet order */
        const int32_t* t = tets + 4 * i;
        const int s = sb(phi[t[0]]) + sb(phi[t[1]]) + sb(phi[t[2]]) + sb(phi[t[3]]);
        if (s == 4 || s == 0)                          /* :380-383 */
            continue;
        int32_t eif[6];
        int ne = 0;
        for (int j = 0; j < 6; j++) {                  /* :386-391 t2e order */
            const int32_t e = t2e[6 * i + j];
            const int32_t a = edges[2 * (int64_t)e + 0];
            const int32_t b = edges[2 * (int64_t)e + 1];
            if (sb(phi[a]) != sb(phi[b]))
                eif[ne++] = e;
        }
        if (ne == 3) {                                 /* :393-395 */
            const int32_t n[3] = {e2ifp[eif[0]], e2ifp[eif[1]], e2ifp[eif[2]]};
            rc = add_triangle(n, t, p, phi, ifp, tris, n_tri, ifp_area);
        } else if (ne == 4) {                          /* :396-399 */
            const int32_t na[3] = {e2ifp[eif[0]], e2ifp[eif[2]], e2ifp[eif[3]]};
            rc = add_triangle(na, t, p, phi, ifp, tris, n_tri, ifp_area);
            if (rc == 0) {
                const int32_t nb[3] = {e2ifp[eif[0]], e2ifp[eif[1]], e2ifp[eif[3]]};
                rc = add_triangle(nb, t, p, phi, ifp, tris, n_tri, ifp_area);
            }
        } else {                                       /* :400-405 */
            rc = -1;
        }
        if (rc != 0) {
            *bad_tet = i;
            break;
        }
        if_elems[(*n_elems)++] = (int32_t)i;           /* :407 */
    }
    free(e2ifp);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* distanceToInterface_ host form (EuclideanDistance.cu:14-35) and the     */
/* CPU branch of EuclideanDistance::resetPhi (:108-117)                    */
/* ---------------------------------------------------------------------- */
static inline float nearest_point(const float* x, const float* ifp,
                                  int64_t n_ifp, int32_t* jmin)
{
    float dmin = FLT_MAX;
    int32_t jm = -1;
    for (int64_t j = 0; j < n_ifp; j++) {
        const float dx = ifp[3 * j + 0] - x[0];
        const float dy = ifp[3 * j + 1] - x[1];
        const float dz = ifp[3 * j + 2] - x[2];
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        if (d < dmin) {                                /* :30 strict: first min */
            dmin = d;
            jm = (int32_t)j;
        }
    }
    *jmin = jm;
    return dmin;
}

void ora_reset_phi_points(int64_t n_nodes, const float* p, const float* ifp,
                          int64_t n_ifp, double* phi_n, int32_t* nearest,
                          int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n_nodes; i++) {
        int32_t j;
        float dmin = nearest_point(p + 3 * i, ifp, n_ifp, &j);
        if (phi_n[i] < 0.0f)                           /* :113 */
            dmin = -dmin;
        phi_n[i] = dmin;
        if (nearest)
            nearest[i] = j;
    }
}

void ora_point_distance(int64_t n_query, const int64_t* nodes, const float* p,
                        const float* ifp, int64_t n_ifp, float* out_d,
                        int32_t* out_j, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t k = 0; k < n_query; k++) {
        const int64_t i = nodes ? nodes[k] : k;
        int32_t j;
        const float d = nearest_point(p + 3 * i, ifp, n_ifp, &j);
        out_d[k] = d;
        if (out_j)
            out_j[k] = j;
    }
}

/* ---------------------------------------------------------------------- */
/* FP64 point-triangle distance.  The closest point of a triangle is the   */
/* orthogonal projection onto its plane when that falls inside (edge       */
/* functions n.(e_i x (x - v_i)) >= 0 for all three edges), else the       */
/* closest point of one of the three edge segments.  Triangles whose       */
/* height is below 1e-7 of their longest edge have no reliable normal in   */
/* FP64 and are measured by their edge segments alone (error <= that       */
/* height).                                                                */
/* ---------------------------------------------------------------------- */
static inline double dot3(const double* u, const double* v)
{
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

static inline void cross3(const double* u, const double* v, double* w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

static double seg_dist2(const double* x, const double* a, const double* b)
{
    double ab[3], ax[3];
    for (int c = 0; c < 3; c++) {
        ab[c] = b[c] - a[c];
        ax[c] = x[c] - a[c];
    }
    const double l2 = dot3(ab, ab);
    double t = 0.0;
    if (l2 > 0.0) {
        t = dot3(ax, ab) / l2;
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    double r[3];
    for (int c = 0; c < 3; c++)
        r[c] = ax[c] - t * ab[c];
    return dot3(r, r);
}

double ora_point_triangle_dist2(const double* x, const double* a,
                                const double* b, const double* c)
{
    double ab[3], bc[3], ca[3], n[3];
    for (int k = 0; k < 3; k++) {
        ab[k] = b[k] - a[k];
        bc[k] = c[k] - b[k];
        ca[k] = a[k] - c[k];
    }
    cross3(ab, bc, n);
    const double nn = dot3(n, n);
    double lmax = dot3(ab, ab);
    if (dot3(bc, bc) > lmax) lmax = dot3(bc, bc);
    if (dot3(ca, ca) > lmax) lmax = dot3(ca, ca);
    if (nn > 1e-14 * lmax * lmax) {
        double xa[3], xb[3], xc[3], t[3];
        for (int k = 0; k < 3; k++) {
            xa[k] = x[k] - a[k];
            xb[k] = x[k] - b[k];
            xc[k] = x[k] - c[k];
        }
        cross3(ab, xa, t);
        const double s0 = dot3(n, t);
        cross3(bc, xb, t);
        const double s1 = dot3(n, t);
        cross3(ca, xc, t);
        const double s2 = dot3(n, t);
        if (s0 >= 0.0 && s1 >= 0.0 && s2 >= 0.0) {
            const double h = dot3(n, xa);
            return h * h / nn;
        }
    }
    double d = seg_dist2(x, a, b);
    const double d2 = seg_dist2(x, b, c);
    const double d3 = seg_dist2(x, c, a);
    if (d2 < d) d = d2;
    if (d3 < d) d = d3;
    return d;
}

static inline void load_tri(const float* ifp, const int32_t* tris, int64_t f,
                            double* a, double* b, double* c)
{
    const int32_t* t = tris + 4 * f;
    for (int k = 0; k < 3; k++) {
        a[k] = (double)ifp[3 * (int64_t)t[0] + k];
        b[k] = (double)ifp[3 * (int64_t)t[1] + k];
        c[k] = (double)ifp[3 * (int64_t)t[2] + k];
    }
}

double ora_node_triangle_distance(int64_t node, const float* p,
                                  const float* ifp, const int32_t* tris,
                                  int64_t f)
{
    double x[3], a[3], b[3], c[3];
    for (int k = 0; k < 3; k++) x[k] = (double)p[3 * node + k];
    load_tri(ifp, tris, f, a, b, c);
    return sqrt(ora_point_triangle_dist2(x, a, b, c));
}

void ora_triangle_distance(int64_t n_query, const int64_t* nodes,
                           const float* p, const float* ifp,
                           const int32_t* tris, int64_t n_tri, double* out_d,
                           int32_t* out_f, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
    for (int64_t k = 0; k < n_query; k++) {
        const int64_t i = nodes ? nodes[k] : k;
        double x[3], a[3], b[3], c[3];
        for (int m = 0; m < 3; m++) x[m] = (double)p[3 * i + m];
        double best = INFINITY;
        int32_t bf = -1;
        for (int64_t f = 0; f < n_tri; f++) {
            load_tri(ifp, tris, f, a, b, c);
            const double d2 = ora_point_triangle_dist2(x, a, b, c);
            if (d2 < best) {
                best = d2;
                bf = (int32_t)f;
            }
        }
        out_d[k] = bf >= 0 ? sqrt(best) : INFINITY;
        out_f[k] = bf;
    }
}

/* ---------------------------------------------------------------------- */
/* LevelSet::resetPhi_ epilogue (LevelSet.cpp:438-448) and the blends of    */
/* LevelSet::init (:73-74) and LevelSet::update (:147-148, weights :18).    */
/* ---------------------------------------------------------------------- */
void ora_finish_reset(int64_t n_nodes, const double* phi_old,
                      const double* phi_ifn, int64_t n_ifp,
                      const int32_t* ifp2e, const int32_t* edges, int blend,
                      double w_phi, double w_ifn, double* retained,
                      double* phi_out)
{
    memcpy(retained, phi_ifn, sizeof(double) * (size_t)n_nodes);
    for (int64_t i = 0; i < n_ifp; i++) {              /* :443-447 */
        const int32_t e = ifp2e[i];
        const int32_t a = edges[2 * (int64_t)e + 0];
        const int32_t b = edges[2 * (int64_t)e + 1];
        retained[a] = phi_old[a];
        retained[b] = phi_old[b];
    }
    for (int64_t i = 0; i < n_nodes; i++) {
        if (blend == 0)
            phi_out[i] = phi_ifn[i];                   /* :74 */
        else
            phi_out[i] = w_phi * retained[i] + w_ifn * phi_ifn[i];  /* :148 */
    }
}

/* ---------------------------------------------------------------------- */
/* Level-set normals: Assembler::grad / grad_nodes (Assembler.cpp:11-71),   */
/* called at the end of LevelSet::resetPhi_ (LevelSet.cpp:451-453).         */
/* ---------------------------------------------------------------------- */
void ora_node_tets(int64_t n_nodes, int64_t n_tets, const int32_t* tets,
                   int64_t* off, int32_t* tet_ids)
{
    /* a stable counting sort of the (node, tet) pairs by node; tets are
     * visited in increasing id, so each node's list comes out sorted, as
     * std::sort of the pairs gives it (TetraMesh.cpp:117) */
    memset(off, 0, sizeof(int64_t) * (size_t)(n_nodes + 1));
    for (int64_t t = 0; t < n_tets; t++)
        for (int k = 0; k < 4; k++) off[tets[4 * t + k] + 1]++;
    for (int64_t i = 0; i < n_nodes; i++) off[i + 1] += off[i];
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    memcpy(cur, off, sizeof(int64_t) * (size_t)n_nodes);
    for (int64_t t = 0; t < n_tets; t++)
        for (int k = 0; k < 4; k++) {
            const int32_t v = tets[4 * t + k];
            tet_ids[cur[v]++] = (int32_t)t;
        }
    free(cur);
}

int64_t ora_tet_grad(int64_t n_tets, const int32_t* tets, const float* p,
                     const double* u, float* grad_t)
{
    int64_t bad = -1;
    for (int64_t t = 0; t < n_tets; t++) {
        const int32_t* q = tets + 4 * t;
        double x[4][3];
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) x[j][k] = (double)p[3 * (int64_t)q[j] + k];
        double B[3][3];                      /* column c = x[c+1] - x[0] */
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) B[r][c] = x[c + 1][r] - x[0][r];
        /* cofactors C(r,c) = B(r1,c1) B(r2,c2) - B(r1,c2) B(r2,c1), cyclic */
        double C[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                C[r][c] = B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1];
            }
        const double det = C[0][0] * B[0][0] + (C[1][0] * B[1][0] + C[2][0] * B[2][0]);
        if (det == 0.0) {
            if (bad < 0) bad = t;
            grad_t[3 * t] = grad_t[3 * t + 1] = grad_t[3 * t + 2] = 0.0f;
            continue;
        }
        const double inv = 1.0 / det;
        /* inverse = adjugate / det: inv(r,c) = C(c,r) / det; its transpose
         * invBt(r,c) = C(r,c) * inv */
        double Bt[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) Bt[r][c] = C[r][c] * inv;
        double g[3];
        for (int r = 0; r < 3; r++) {
            /* invBt times the basis gradients (a product with a dynamic-size
             * matrix, Quadrature.hpp:118-126: its inner sums run left to
             * right); the unit gradients select a column exactly */
            const double c0 = (Bt[r][0] * -1.0 + Bt[r][1] * -1.0) + Bt[r][2] * -1.0;
            const double c1 = Bt[r][0];
            const double c2 = Bt[r][1];
            const double c3 = Bt[r][2];
            g[r] = ((u[q[0]] * c0 + u[q[1]] * c1) + u[q[2]] * c2) + u[q[3]] * c3;
        }
        for (int k = 0; k < 3; k++) grad_t[3 * t + k] = (float)g[k];
    }
    return bad;
}

void ora_grad_nodes(int64_t n_nodes, const int64_t* off, const int32_t* tet_ids,
                    const float* grad_t, float* grad_n)
{
    for (int64_t i = 0; i < n_nodes; i++) {
        float n[3] = {0.0f, 0.0f, 0.0f};
        for (int64_t j = off[i]; j < off[i + 1]; j++)
            for (int k = 0; k < 3; k++) n[k] += grad_t[3 * (int64_t)tet_ids[j] + k];
        const float cnt = (float)(off[i + 1] - off[i]);
        for (int k = 0; k < 3; k++) grad_n[3 * i + k] = n[k] / cnt;
    }
}

void ora_interface_curvature(int64_t n_nodes, const float* p, int64_t n_tets,
                             const int32_t* tets, const int64_t* off,
                             const int32_t* tet_ids, const float* g,
                             const double* phi, int64_t n_ifp,
                             const int32_t* ifp2e, const int32_t* edges,
                             float* k_if)
{
    double* gc[3];
    for (int c = 0; c < 3; c++) gc[c] = (double*)malloc(sizeof(double) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    for (int64_t i = 0; i < n_nodes; i++) {                      /* :224-230 */
        const float x = g[3 * i], y = g[3 * i + 1], z = g[3 * i + 2];
        const float nrm = sqrtf(x * x + (y * y + z * z));
        gc[0][i] = (double)(x / nrm);
        gc[1][i] = (double)(y / nrm);
        gc[2][i] = (double)(z / nrm);
    }
    float* gt = (float*)malloc(sizeof(float) * 3 * (size_t)(n_tets > 0 ? n_tets : 1));
    float* gn = (float*)malloc(sizeof(float) * 3 * (size_t)(n_nodes > 0 ? n_nodes : 1));
    float* k = (float*)malloc(sizeof(float) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    for (int c = 0; c < 3; c++) {                                /* :233-239 */
        ora_tet_grad(n_tets, tets, p, gc[c], gt);
        ora_grad_nodes(n_nodes, off, tet_ids, gt, gn);
        for (int64_t i = 0; i < n_nodes; i++) k[i] = c == 0 ? gn[3 * i] : k[i] + gn[3 * i + c];
    }
    for (int64_t i = 0; i < n_ifp; i++) {                        /* :243-250 */
        const int32_t e = ifp2e[i];
        const int32_t a = edges[2 * (int64_t)e], b = edges[2 * (int64_t)e + 1];
        const float t = (float)(phi[a] / (phi[a] - phi[b]));
        const float tk1 = t * k[b];
        k_if[i] = (float)((1.0 - (double)t) * (double)k[a] + (double)tk1);
    }
    for (int c = 0; c < 3; c++) free(gc[c]);
    free(gt);
    free(gn);
    free(k);
}


/* ---------------------------------------------------------------------- */
/* Flow::surfaceTension_  (src/growth/Flow.cpp:307-341)                    */
/* ---------------------------------------------------------------------- */
static void unit_normal(const float* g, int32_t v, float* n)
{
    const float x = g[3 * (int64_t)v], y = g[3 * (int64_t)v + 1], z = g[3 * (int64_t)v + 2];
    const float nrm = sqrtf(x * x + (y * y + z * z));              /* :327, Vec3f::norm */
    n[0] = x / nrm;
    n[1] = y / nrm;
    n[2] = z / nrm;
}

void ora_surface_tension(const float* g, const double* phi, int64_t n_ifp,
                         const int32_t* ifp2e, const int32_t* edges, const float* k_if,
                         const float* area, float st, int32_t rho, double* fx, double* fy,
                         double* fz)
{
    for (int64_t i = 0; i < n_ifp; i++) {                          /* :313 */
        const int32_t e0 = edges[2 * (int64_t)ifp2e[i]], e1 = edges[2 * (int64_t)ifp2e[i] + 1];
        const float t = (float)(phi[e0] / (phi[e0] - phi[e1]));    /* :320 */
        float n[3];
        unit_normal(g, e0, n);                                     /* :327 */
        const double s0 = (1.0 - (double)t) * (double)st;          /* :328, left to right */
        const double s1 = s0 * (double)(-k_if[i]);
        const double s2 = s1 * (double)area[i];
        const float f0 = (float)(s2 * (double)rho);
        fx[e0] += (double)(f0 * n[0]);                             /* :329-331 */
        fy[e0] += (double)(f0 * n[1]);
        fz[e0] += (double)(f0 * n[2]);
        unit_normal(g, e1, n);                                     /* :333 */
        const float f1 = (((t * st) * (-k_if[i])) * area[i]) * (float)rho;   /* :334, float */
        fx[e1] += (double)(f1 * n[0]);                             /* :335-337 */
        fy[e1] += (double)(f1 * n[1]);
        fz[e1] += (double)(f1 * n[2]);
    }
}


/* ---------------------------------------------------------------------- */
/* LevelSet::update, advection (LevelSet.cpp:89-141; Assembler.hpp:131-226) */
/* ---------------------------------------------------------------------- */
void ora_fem_quad(double* phi, double* w)
{
    const double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
    const double qp[3][4] = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}};
    for (int q = 0; q < 4; q++) {
        const float x = (float)qp[0][q], y = (float)qp[1][q], z = (float)qp[2][q];
        phi[0 * 4 + q] = (double)(float)(((1.0 - (double)x) - (double)y) - (double)z);
        phi[1 * 4 + q] = (double)x;
        phi[2 * 4 + q] = (double)y;
        phi[3 * 4 + q] = (double)z;
    }
    *w = 0.25 / 6.0;
}

/* getElementInfo_ij_ (Assembler.hpp:429-459) as in ora_tet_grad, and the
 * mapped basis gradients G[j]; returns 0 for a singular map. */
static int fem_frame(const float* p, const int32_t* q, double G[4][3], double* detB)
{
    double x[4][3];
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) x[j][k] = (double)p[3 * (int64_t)q[j] + k];
    double B[3][3], C[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) B[r][c] = x[c + 1][r] - x[0][r];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            C[r][c] = B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1];
        }
    const double det = C[0][0] * B[0][0] + (C[1][0] * B[1][0] + C[2][0] * B[2][0]);
    if (det == 0.0) return 0;
    const double inv = 1.0 / det;
    for (int r = 0; r < 3; r++) {
        const double b0 = C[r][0] * inv, b1 = C[r][1] * inv, b2 = C[r][2] * inv;
        G[0][r] = (b0 * -1.0 + b1 * -1.0) + b2 * -1.0;
        G[1][r] = b0;
        G[2][r] = b1;
        G[3][r] = b2;
    }
    *detB = fabs(det);
    return 1;
}

int64_t ora_fem_assemble(int form, int64_t n, const float* p, const int32_t* tets,
                         const int64_t* n2t_off, const int32_t* n2t_ids,
                         const int64_t* row_off, const int32_t* cols, const double* vel,
                         double* vals)
{
    double phi[16], w;
    ora_fem_quad(phi, &w);
    int64_t bad = -1;
    for (int64_t i = 0; i < n; i++) {
        const int64_t rb = row_off[i], re = row_off[i + 1];
        for (int64_t c = rb; c < re; c++) vals[c] = 0.0;
        for (int64_t a = n2t_off[i]; a < n2t_off[i + 1]; a++) {
            const int32_t k = n2t_ids[a];
            const int32_t* q = tets + 4 * (int64_t)k;
            double G[4][3], detB;
            if (!fem_frame(p, q, G, &detB)) {
                if (bad < 0 || k < bad) bad = k;
                continue;
            }
            int li = 0;
            while (li < 3 && q[li] != (int32_t)i) li++;
            double V[3][4] = {{0.0}};
            if (form == 1)
                for (int c = 0; c < 3; c++)
                    for (int q4 = 0; q4 < 4; q4++)
                        for (int j = 0; j < 4; j++)
                            V[c][q4] = V[c][q4] + vel[c * n + q[j]] * phi[4 * j + q4];
            for (int j = 0; j < 4; j++) {
                double f[4];
                for (int q4 = 0; q4 < 4; q4++) {
                    const double v = phi[4 * li + q4];
                    if (form == 0)
                        f[q4] = phi[4 * j + q4] * v;
                    else
                        f[q4] = (V[0][q4] * (G[j][0] * v) + V[1][q4] * (G[j][1] * v)) +
                                V[2][q4] * (G[j][2] * v);
                }
                const double qint = ((f[0] * w + f[2] * w) + (f[1] * w + f[3] * w)) * detB;
                int64_t c = rb;
                while (c < re && cols[c] != q[j]) c++;
                if (c < re) vals[c] += qint;
            }
        }
    }
    return bad;
}

void ora_spmv(int64_t n, const int64_t* row_off, const int32_t* cols, const double* vals,
              const double* x, double* y)
{
    for (int64_t i = 0; i < n; i++) {
        double t = 0.0;
        for (int64_t c = row_off[i]; c < row_off[i + 1]; c++) t += vals[c] * x[cols[c]];
        y[i] = t;
    }
}

static double dotv(int64_t n, const double* a, const double* b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int ora_bicgstab(int64_t n, const int64_t* row_off, const int32_t* cols, const double* vals,
                 const double* b, double* x, double tol, int64_t max_iters, int64_t* iters,
                 double* error, int64_t* restarts)
{
    const size_t sz = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double* invd = (double*)malloc(sz);
    double *r = (double*)malloc(sz), *r0 = (double*)malloc(sz), *p = (double*)calloc(1, sz);
    double *v = (double*)calloc(1, sz), *y = (double*)malloc(sz), *z = (double*)malloc(sz);
    double *s = (double*)malloc(sz), *t = (double*)malloc(sz);
    for (int64_t i = 0; i < n; i++) {
        invd[i] = 1.0;
        for (int64_t c = row_off[i]; c < row_off[i + 1]; c++)
            if (cols[c] == i) { if (vals[c] != 0.0) invd[i] = 1.0 / vals[c]; break; }
    }
    ora_spmv(n, row_off, cols, vals, x, t);
    for (int64_t i = 0; i < n; i++) r[i] = r0[i] = b[i] - t[i];
    double r0_sq = dotv(n, r0, r0);
    const double b_sq = dotv(n, b, b);
    int64_t it = 0, rs = 0;
    int ok = 1;
    if (b_sq == 0.0) {
        for (int64_t i = 0; i < n; i++) x[i] = 0.0;
        *error = 0.0;
    } else {
        double rho = 1.0, alpha = 1.0, w = 1.0;
        const double tol2 = tol * tol * b_sq, eps2 = DBL_EPSILON * DBL_EPSILON;
        while (dotv(n, r, r) > tol2 && it < max_iters) {
            const double rho_old = rho;
            rho = dotv(n, r0, r);
            if (fabs(rho) < eps2 * r0_sq) {
                ora_spmv(n, row_off, cols, vals, x, t);
                for (int64_t i = 0; i < n; i++) r[i] = r0[i] = b[i] - t[i];
                rho = r0_sq = dotv(n, r, r);
                if (rs++ == 0) it = 0;
            }
            const double beta = (rho / rho_old) * (alpha / w);
            for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
            for (int64_t i = 0; i < n; i++) y[i] = invd[i] * p[i];
            ora_spmv(n, row_off, cols, vals, y, v);
            alpha = rho / dotv(n, r0, v);
            for (int64_t i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
            for (int64_t i = 0; i < n; i++) z[i] = invd[i] * s[i];
            ora_spmv(n, row_off, cols, vals, z, t);
            const double tt = dotv(n, t, t);
            w = tt > 0.0 ? dotv(n, t, s) / tt : 0.0;
            for (int64_t i = 0; i < n; i++) x[i] += alpha * y[i] + w * z[i];
            for (int64_t i = 0; i < n; i++) r[i] = s[i] - w * t[i];
            it++;
        }
        *error = sqrt(dotv(n, r, r) / b_sq);
        ok = *error <= tol;
    }
    *iters = it;
    *restarts = rs;
    free(invd); free(r); free(r0); free(p); free(v); free(y); free(z); free(s); free(t);
    return ok;
}
