// EuclideanDistance.cpp -- the reference's EuclideanDistance class
// (src/fem/EuclideanDistance.cu:66-118) on libftreinit's C ABI.
//
// The class layout cannot change (LevelSet embeds it by value,
// src/fem/LevelSet.hpp:142), so the device handle of each object lives in a
// registry keyed by the object's address.  Like the reference, objects have
// no destructor; their handle lives until the process ends.
#include "EuclideanDistance.hpp"

#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "ftreinit.h"

namespace {

std::mutex g_mu;
std::unordered_map<const EuclideanDistance*, ft_handle>& registry()
{
    static std::unordered_map<const EuclideanDistance*, ft_handle> m;
    return m;
}

[[noreturn]] void fail(const char* where, int rc)
{
    throw std::runtime_error(std::string("EuclideanDistance::") + where + ": ftreinit error " +
                             std::to_string(rc) + ": " + ft_last_error());
}

ft_handle handle_of(const EuclideanDistance* self)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = registry().find(self);
    return it == registry().end() ? nullptr : it->second;
}

}  // namespace

void EuclideanDistance::init(const float* meshp, const double* phi0, const int n,
                             const int nThreads)
{
    (void)phi0;  // the reference uploads it but never reads it (resetPhi re-uploads phi)
    m_meshSize = n;
    m_threads = nThreads;
    int32_t count = 0;
    ft_device_count(&count);
    m_devCount = count;
    if (count <= 0)
        throw std::runtime_error("EuclideanDistance::init: no HIP device (this library has no CPU "
                                 "branch: the distance runs on the GPU only)");
    ft_handle h = handle_of(this);
    if (!h) {
        int rc = ft_create(0, &h);
        if (rc) fail("init", rc);
        std::lock_guard<std::mutex> lk(g_mu);
        registry()[this] = h;
    }
    h_p = meshp;
    int rc = ft_set_nodes(h, meshp, n);
    if (rc) fail("init", rc);
}

void EuclideanDistance::resetPhi(const float* ifp, const int n_ifp, double* phi_n)
{
    ft_handle h = handle_of(this);
    if (!h) fail("resetPhi", FT_ERR_STATE);
    int rc = ft_reset_phi_points(h, ifp, n_ifp, phi_n);
    if (rc) fail("resetPhi", rc);
}
