// Drop-in link test: a translation unit written like the reference's
// LevelSet.cpp (src/fem/LevelSet.cpp:68, 439) that includes the
// EuclideanDistance declaration and links libftreinit.so.  It calls
//   dist_.init(mesh_.p(0).data(), phi_.data(), n, nThreads_)
//   dist_.resetPhi(ifp_[0].data(), ifp_.size(), phi_n.data())
// on a small grid and prints its inputs and the resulting phi (hex floats)
// for the caller to compare against the oracle.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "EuclideanDistance.hpp"

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 12;
    const float h = 1.0f / (n - 1);
    std::vector<float> p;                       // mesh_.p(): Vec3f, x-major
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                p.push_back(i * h);
                p.push_back(j * h);
                p.push_back(k * h);
            }
    const int N = n * n * n;
    std::vector<double> phi(N);
    for (int a = 0; a < N; a++) {
        const double dx = p[3 * a] - 0.5, dy = p[3 * a + 1] - 0.5, dz = p[3 * a + 2] - 0.5;
        phi[a] = std::sqrt(dx * dx + dy * dy + dz * dz) - 0.3;
    }
    // a few interface points (ifp_): on the sphere of radius 0.3
    std::vector<float> ifp;
    for (int q = 0; q < 40; q++) {
        const double t = 0.7 * q, u = 0.13 * q;
        ifp.push_back((float)(0.5 + 0.3 * std::sin(t) * std::cos(u)));
        ifp.push_back((float)(0.5 + 0.3 * std::sin(t) * std::sin(u)));
        ifp.push_back((float)(0.5 + 0.3 * std::cos(t)));
    }
    EuclideanDistance dist;                     // embedded by value in LevelSet
    static_assert(sizeof(EuclideanDistance) == 40, "layout of the reference class");
    const int nThreads = std::thread::hardware_concurrency();
    dist.init(p.data(), phi.data(), N, nThreads);
    std::vector<double> phi_n = phi;            // LevelSet.cpp:438
    const int P = (int)(ifp.size() / 3);
    dist.resetPhi(ifp.data(), P, phi_n.data());
    const bool copies = argc > 2 && std::string(argv[2]) == "copy";
    // copies: a copy and a moved object of an initialised one (the implicit
    // copy of the reference class), then an object copied into the storage
    // of a destroyed one that had other nodes (the registry's address reused)
    std::vector<double> phi_c = phi, phi_m = phi, phi_r = phi;
    if (copies) {
        EuclideanDistance copy = dist;
        copy.resetPhi(ifp.data(), P, phi_c.data());
        EuclideanDistance moved = std::move(copy);
        moved.resetPhi(ifp.data(), P, phi_m.data());
        std::vector<float> shifted(p.size());
        for (size_t a = 0; a < p.size(); a++) shifted[a] = p[a] + 0.25f;
        alignas(EuclideanDistance) unsigned char store[sizeof(EuclideanDistance)];
        EuclideanDistance* other = new (store) EuclideanDistance();
        other->init(shifted.data(), phi.data(), N, nThreads);
        std::vector<double> scratch = phi;
        other->resetPhi(ifp.data(), P, scratch.data());
        other->~EuclideanDistance();
        EuclideanDistance* reused = new (store) EuclideanDistance(dist);
        reused->resetPhi(ifp.data(), P, phi_r.data());
    }
    std::printf("%d %d\n", N, P);
    for (float v : ifp) std::printf("%a\n", (double)v);
    for (int a = 0; a < N; a++) {
        if (copies)
            std::printf("%a %a %a %a %a\n", phi[a], phi_n[a], phi_c[a], phi_m[a], phi_r[a]);
        else
            std::printf("%a %a\n", phi[a], phi_n[a]);
    }
    return 0;
}
