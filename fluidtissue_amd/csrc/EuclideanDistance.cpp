// EuclideanDistance.cpp -- the reference's EuclideanDistance class
// (src/fem/EuclideanDistance.cu:66-118) on libftreinit's C ABI.
//
// The class layout cannot change (LevelSet embeds it by value,
// src/fem/LevelSet.hpp:142), so the device handle of each object lives in a
// registry keyed by the object's address, beside the node array and count it
// was given.  Like the reference, objects have no destructor: a handle lives
// until the process ends.  A copied or moved object (the implicit copy, as in
// the reference, copies the node pointer and count) or a new object at the
// address of a dead one finds no entry, or one for other nodes, and uploads
// its own nodes (h_p, m_meshSize) on its first resetPhi.
//
// With no HIP device, init and resetPhi take the reference's CPU branch
// (EuclideanDistance.cu:82-85, 107-117): the node pointer is borrowed and
// each node's distance is the brute-force minimum over the crossing points,
// on m_threads host threads.  It says so once on stderr; FT_REQUIRE_DEVICE=1
// in the environment makes init throw instead.
#include "EuclideanDistance.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ftreinit.h"

namespace {

struct Entry {
    ft_handle h = nullptr;
    const float* nodes = nullptr;    // the node array uploaded to h
    int n = 0;
};

std::mutex g_mu;
std::unordered_map<const EuclideanDistance*, Entry>& registry()
{
    static std::unordered_map<const EuclideanDistance*, Entry> m;
    return m;
}

[[noreturn]] void fail(const char* where, int rc)
{
    throw std::runtime_error(std::string("EuclideanDistance::") + where + ": ftreinit error " +
                             std::to_string(rc) + ": " + ft_last_error());
}

// The handle of `self` with `nodes` (n of them) uploaded: created on the
// first call, nodes uploaded again when the object's nodes are not the ones
// its handle holds.
ft_handle handle_for(const EuclideanDistance* self, const float* nodes, int n, const char* where)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Entry& e = registry()[self];
    if (!e.h) {
        int rc = ft_create(0, &e.h);
        if (rc) {
            registry().erase(self);
            fail(where, rc);
        }
    }
    if (e.nodes != nodes || e.n != n) {
        int rc = ft_set_nodes(e.h, nodes, n);
        if (rc) {
            e.nodes = nullptr;
            fail(where, rc);
        }
        e.nodes = nodes;
        e.n = n;
    }
    return e.h;
}

// distanceToInterface_ of the host branch (EuclideanDistance.cu:14-35): the
// float minimum over the crossing points of sqrt(dx*dx + dy*dy + dz*dz),
// each product and sum rounded on its own, first minimum kept.
float point_distance(const float* p, const float* ifp, int n_ifp)
{
    float dmin = FLT_MAX;
    for (int j = 0; j < n_ifp; j++) {
        const float dx = ifp[3 * j + 0] - p[0];
        const float dy = ifp[3 * j + 1] - p[1];
        const float dz = ifp[3 * j + 2] - p[2];
        const float s2 = dx * dx + dy * dy + dz * dz;
        const float d = std::sqrt(s2);
        if (d < dmin) dmin = d;
    }
    return dmin;
}

// The host branch of resetPhi (EuclideanDistance.cu:107-117) over nodes
// [lo, hi): the sign from phi_n < 0, written back as double.
void host_range(const float* nodes, const float* ifp, int n_ifp, double* phi_n, int lo, int hi)
{
    for (int i = lo; i < hi; i++) {
        const float p[3] = {nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]};
        float dmin = point_distance(p, ifp, n_ifp);
        if (phi_n[i] < 0.0f) dmin = -dmin;
        phi_n[i] = dmin;
    }
}

bool device_required()
{
    const char* v = std::getenv("FT_REQUIRE_DEVICE");
    return v && std::strcmp(v, "0") != 0 && *v;
}

}  // namespace

void EuclideanDistance::init(const float* meshp, const double* phi0, const int n,
                             const int nThreads)
{
    (void)phi0;  // the reference uploads it but never reads it (resetPhi re-uploads phi)
    m_meshSize = n;
    m_threads = nThreads;
    h_p = meshp;
    int32_t count = 0;
    ft_device_count(&count);
    m_devCount = count;
    if (count <= 0) {
        if (device_required())
            throw std::runtime_error("EuclideanDistance::init: no HIP device and FT_REQUIRE_DEVICE is set");
        static std::once_flag note;
        std::call_once(note, [] {
            std::fprintf(stderr, "EuclideanDistance: no HIP device, running the CPU branch\n");
        });
        return;                                  // the CPU branch borrows meshp (:84)
    }
    handle_for(this, meshp, n, "init");
}

void EuclideanDistance::resetPhi(const float* ifp, const int n_ifp, double* phi_n)
{
    if (m_devCount <= 0) {
        const int n = m_meshSize;
        const int nt = std::max(1, std::min(m_threads, n));
        if (nt == 1 || n < 4096) {
            host_range(h_p, ifp, n_ifp, phi_n, 0, n);
            return;
        }
        std::vector<std::thread> th;
        const int chunk = (n + nt - 1) / nt;
        for (int t = 0; t < nt; t++) {
            const int lo = t * chunk, hi = std::min(n, lo + chunk);
            if (lo >= hi) break;
            th.emplace_back(host_range, h_p, ifp, n_ifp, phi_n, lo, hi);
        }
        for (std::thread& x : th) x.join();
        return;
    }
    ft_handle h = handle_for(this, h_p, m_meshSize, "resetPhi");
    int rc = ft_reset_phi_points(h, ifp, n_ifp, phi_n);
    if (rc) fail("resetPhi", rc);
}
