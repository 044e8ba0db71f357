// ft_gen.cpp -- synthetic reinit inputs: Kuhn tetrahedral grids and the
// level-set fields of the benchmark workloads.
//
// The reference meshes a regular node grid with CGAL's Delaunay
// triangulation (src/fem/TetraMesh.cpp:23-165, nodes at :169-212).  CGAL is
// not available here, so grids are cut into Kuhn tets instead; the node
// numbering keeps the reference's x-major order (TetraMesh.cpp:199-211) and
// the local edge order of t2e is the reference's {01,02,03,12,13,23}
// (TetraMesh.cpp:103).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "ftreinit.h"
#include "ft_internal.hpp"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): the generator state advances by the
// golden-ratio increment and is mixed by two xor-shift-multiply rounds.
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() { s += kGolden; return mix64(s); }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Counter-based draw: the k-th value of the stream `seed`, so that threads
// can draw independently and the result does not depend on thread count.
inline double u01_at(uint64_t seed, uint64_t k)
{
    return (double)(mix64(seed + (k + 1) * kGolden) >> 11) * 0x1.0p-53;
}

template <class F>
void parallel_for(int64_t n, F&& f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536) nt = 1;
    const int64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &f] { for (int64_t i = b; i < e; i++) f(i); });
    }
    for (auto& x : th) x.join();
}

// The 7 edge directions leaving a node toward higher indices, as (dx,dy,dz).
constexpr int kDir[7][3] = {{1,0,0},{0,1,0},{0,0,1},{1,1,0},{1,0,1},{0,1,1},{1,1,1}};
inline int dir_id(int dx, int dy, int dz)
{
    for (int d = 0; d < 7; d++)
        if (kDir[d][0] == dx && kDir[d][1] == dy && kDir[d][2] == dz) return d;
    return -1;
}
// The 6 Kuhn tets of a cell: axis orders (0=x,1=y,2=z).
constexpr int kPerm[6][3] = {{0,1,2},{0,2,1},{1,0,2},{1,2,0},{2,0,1},{2,1,0}};

}  // namespace

extern "C" int ft_kuhn_counts(int32_t nx, int32_t ny, int32_t nz, int64_t* n_nodes,
                              int64_t* n_tets, int64_t* n_edges)
{
    if (nx < 2 || ny < 2 || nz < 2 || !n_nodes || !n_tets || !n_edges)
        return ft::set_error(FT_ERR_ARG, "ft_kuhn_counts: need nx,ny,nz >= 2 and outputs");
    const int64_t X = nx, Y = ny, Z = nz;
    *n_nodes = X * Y * Z;
    *n_tets = 6 * (X - 1) * (Y - 1) * (Z - 1);
    *n_edges = (X - 1) * Y * Z + X * (Y - 1) * Z + X * Y * (Z - 1) +
               (X - 1) * (Y - 1) * Z + (X - 1) * Y * (Z - 1) + X * (Y - 1) * (Z - 1) +
               (X - 1) * (Y - 1) * (Z - 1);
    if (*n_nodes > INT32_MAX || *n_edges > INT32_MAX || *n_tets > INT32_MAX)
        return ft::set_error(FT_ERR_ARG, "ft_kuhn_counts: grid exceeds int32 indices");
    return FT_OK;
}

extern "C" int ft_kuhn_mesh(int32_t nx, int32_t ny, int32_t nz, const double* origin,
                            const double* spacing, double jitter, uint64_t seed,
                            float* xyz, int32_t* tets, int32_t* edges, int32_t* t2e)
{
    int64_t N, T, E;
    int rc = ft_kuhn_counts(nx, ny, nz, &N, &T, &E);
    if (rc) return rc;
    if (!origin || !spacing || !xyz || !tets || !edges || !t2e)
        return ft::set_error(FT_ERR_ARG, "ft_kuhn_mesh: null pointer");
    if (!(jitter >= 0.0 && jitter < 0.5))
        return ft::set_error(FT_ERR_ARG, "ft_kuhn_mesh: jitter must be in [0, 0.5)");
    const int64_t X = nx, Y = ny, Z = nz;

    // Nodes (TetraMesh.cpp:199-211 order).
    parallel_for(N, [&](int64_t a) {
        const int64_t i = a / (Y * Z), j = (a / Z) % Y, k = a % Z;
        const bool interior = i > 0 && i < X - 1 && j > 0 && j < Y - 1 && k > 0 && k < Z - 1;
        const int64_t ijk[3] = {i, j, k};
        for (int c = 0; c < 3; c++) {
            double g = (double)ijk[c];
            if (jitter > 0.0 && interior)
                g += (2.0 * u01_at(seed, 3 * (uint64_t)a + c) - 1.0) * jitter;
            xyz[3 * a + c] = (float)(origin[c] + g * spacing[c]);
        }
    });

    // Edge numbering: by low node, then direction; offsets per node.
    auto valid = [&](int64_t i, int64_t j, int64_t k, int d) {
        return (kDir[d][0] == 0 || i < X - 1) && (kDir[d][1] == 0 || j < Y - 1) &&
               (kDir[d][2] == 0 || k < Z - 1);
    };
    // Offsets: count edges per (i,j)-row, then per node within a row.
    std::vector<int64_t> row_off(X * Y + 1, 0);
    for (int64_t r = 0; r < X * Y; r++) {
        const int64_t i = r / Y, j = r % Y;
        int64_t c = 0;
        for (int64_t k = 0; k < Z; k++)
            for (int d = 0; d < 7; d++) c += valid(i, j, k, d);
        row_off[r + 1] = row_off[r] + c;
    }
    if (row_off[X * Y] != E)
        return ft::set_error(FT_ERR_ARG, "ft_kuhn_mesh: internal edge count mismatch");
    auto node_edge_base = [&](int64_t i, int64_t j, int64_t k) {
        // Nodes (i,j,0..k-1) all have k' < Z-1, hence the same edge count.
        int64_t per = 0;
        for (int d = 0; d < 7; d++) per += valid(i, j, 0, d);
        return row_off[i * Y + j] + k * per;
    };
    auto edge_id = [&](int64_t i, int64_t j, int64_t k, int d) {
        int64_t base = node_edge_base(i, j, k);
        int r = 0;
        for (int q = 0; q < d; q++) r += valid(i, j, k, q);
        return (int32_t)(base + r);
    };
    parallel_for(X * Y, [&](int64_t r) {
        const int64_t i = r / Y, j = r % Y;
        int64_t e = row_off[r];
        for (int64_t k = 0; k < Z; k++) {
            const int64_t a = (i * Y + j) * Z + k;
            for (int d = 0; d < 7; d++) {
                if (!valid(i, j, k, d)) continue;
                const int64_t b = ((i + kDir[d][0]) * Y + (j + kDir[d][1])) * Z + (k + kDir[d][2]);
                edges[2 * e + 0] = (int32_t)a;
                edges[2 * e + 1] = (int32_t)b;
                e++;
            }
        }
    });

    // Tets: by cell (x-major over cells), then the 6 axis orders.
    const int64_t CX = X - 1, CY = Y - 1, CZ = Z - 1;
    parallel_for(CX * CY, [&](int64_t r) {
        const int64_t ci = r / CY, cj = r % CY;
        for (int64_t ck = 0; ck < CZ; ck++) {
            const int64_t cell = (ci * CY + cj) * CZ + ck;
            for (int p = 0; p < 6; p++) {
                int64_t v[4][3];
                v[0][0] = ci; v[0][1] = cj; v[0][2] = ck;
                for (int s = 1; s <= 3; s++) {
                    for (int c = 0; c < 3; c++) v[s][c] = v[s - 1][c];
                    v[s][kPerm[p][s - 1]] += 1;
                }
                const int64_t t = 6 * cell + p;
                for (int s = 0; s < 4; s++)
                    tets[4 * t + s] = (int32_t)((v[s][0] * Y + v[s][1]) * Z + v[s][2]);
                static const int kLoc[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
                for (int q = 0; q < 6; q++) {
                    const int64_t* lo = v[kLoc[q][0]];
                    const int64_t* hi = v[kLoc[q][1]];
                    const int d = dir_id((int)(hi[0] - lo[0]), (int)(hi[1] - lo[1]), (int)(hi[2] - lo[2]));
                    t2e[6 * t + q] = edge_id(lo[0], lo[1], lo[2], d);
                }
            }
        }
    });
    return FT_OK;
}

extern "C" int ft_phi_field(int32_t kind, int64_t n, const float* xyz, uint64_t seed,
                            double* phi)
{
    if (n < 0 || (n > 0 && (!xyz || !phi)))
        return ft::set_error(FT_ERR_ARG, "ft_phi_field: bad arguments");
    const double kPi = 3.14159265358979323846;
    switch (kind) {
    case FT_PHI_SPHERE_WAVE:
        parallel_for(n, [&](int64_t i) {
            const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            const double dx = x - 0.35, dy = y - 0.35, dz = z - 0.35;
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double w = 1.0 + 0.5 * std::sin(4 * kPi * x) * std::sin(4 * kPi * y) *
                                       std::sin(4 * kPi * z);
            phi[i] = (r - 0.15) * w;
        });
        return FT_OK;
    case FT_PHI_STAR:
        parallel_for(n, [&](int64_t i) {
            const double dx = xyz[3 * i] - 0.5, dy = xyz[3 * i + 1] - 0.5,
                         dz = xyz[3 * i + 2] - 0.5;
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double th = r > 0.0 ? std::acos(std::max(-1.0, std::min(1.0, dz / r))) : 0.0;
            const double ph = std::atan2(dy, dx);
            const double s = std::sin(th);
            const double R = 0.3 + 0.06 * std::cos(5.0 * ph) * s * s * s;
            phi[i] = (r - R) * 2.3;
        });
        return FT_OK;
    case FT_PHI_SPHERES: {
        SplitMix g(seed);
        double c[64][3], rad[64];
        for (int s = 0; s < 64; s++) {
            for (int k = 0; k < 3; k++) c[s][k] = 0.1 + 0.8 * g.u01();
            rad[s] = 0.02 + 0.06 * g.u01();
        }
        parallel_for(n, [&](int64_t i) {
            const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            double m = INFINITY;
            for (int s = 0; s < 64; s++) {
                const double dx = x - c[s][0], dy = y - c[s][1], dz = z - c[s][2];
                m = std::min(m, std::sqrt(dx * dx + dy * dy + dz * dz) - rad[s]);
            }
            phi[i] = 1.5 * m;
        });
        return FT_OK;
    }
    case FT_PHI_ELLIPSOIDS: {
        SplitMix g(seed);
        const int ne = 1 + (int)(16.0 * g.u01());
        struct Ell { double c[3], ax[3], R[3][3]; };
        std::vector<Ell> el(ne);
        for (auto& e : el) {
            for (int k = 0; k < 3; k++) e.c[k] = 0.25 + 0.5 * g.u01();
            for (int k = 0; k < 3; k++) e.ax[k] = 0.03 + 0.17 * g.u01();
            // Uniform random rotation from a unit quaternion (Shoemake 1992).
            const double u1 = g.u01(), u2 = g.u01(), u3 = g.u01();
            const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
            const double qw = a * std::sin(2 * kPi * u2), qx = a * std::cos(2 * kPi * u2),
                         qy = b * std::sin(2 * kPi * u3), qz = b * std::cos(2 * kPi * u3);
            e.R[0][0] = 1 - 2 * (qy * qy + qz * qz); e.R[0][1] = 2 * (qx * qy - qz * qw);
            e.R[0][2] = 2 * (qx * qz + qy * qw);     e.R[1][0] = 2 * (qx * qy + qz * qw);
            e.R[1][1] = 1 - 2 * (qx * qx + qz * qz); e.R[1][2] = 2 * (qy * qz - qx * qw);
            e.R[2][0] = 2 * (qx * qz - qy * qw);     e.R[2][1] = 2 * (qy * qz + qx * qw);
            e.R[2][2] = 1 - 2 * (qx * qx + qy * qy);
        }
        parallel_for(n, [&](int64_t i) {
            const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
            double m = INFINITY;
            for (const auto& e : el) {
                const double d[3] = {p[0] - e.c[0], p[1] - e.c[1], p[2] - e.c[2]};
                double s = 0.0;
                for (int r = 0; r < 3; r++) {  // body coordinates: R^T d, scaled
                    const double q = (e.R[0][r] * d[0] + e.R[1][r] * d[1] + e.R[2][r] * d[2]) / e.ax[r];
                    s += q * q;
                }
                const double amin = std::min(e.ax[0], std::min(e.ax[1], e.ax[2]));
                m = std::min(m, (std::sqrt(s) - 1.0) * amin);
            }
            phi[i] = m;
        });
        return FT_OK;
    }
    default:
        return ft::set_error(FT_ERR_ARG, "ft_phi_field: unknown kind %d", kind);
    }
}
