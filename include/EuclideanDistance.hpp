//
// Drop-in declaration of the reference's EuclideanDistance class
// (src/fem/EuclideanDistance.hpp:13-52): same methods, same member layout
// (so sizeof(LevelSet) is unchanged), no CUDA or HIP include.  The methods are
// implemented in libftreinit.so on gfx950 HIP kernels; see include/ftreinit.h
// for the C ABI underneath.
//
// With no HIP device the methods run the reference's CPU branch
// (EuclideanDistance.cu:82-85, 107-117) in the library's own host loop, on
// nThreads threads; FT_REQUIRE_DEVICE=1 in the environment makes init throw
// instead.  On a device error the methods throw std::runtime_error where the
// reference fails silently.  Copies and moved objects work as the
// reference's do (each uploads its nodes to its own device handle on first
// use); like the reference, objects hold device memory until the process
// ends.
//

#pragma once


class EuclideanDistance
{

public:
    EuclideanDistance() :
        h_p(nullptr),
        d_p(nullptr),
        d_phi(nullptr),
        m_meshSize(0),
        m_devCount(0),
        m_threads(1)
    {}

    //
    // Upload the fixed mesh node position (meshp, 3xn) and initial level
    // set (phi0, n).
    //
    void        init( const float* meshp, const double* phi0, const int n,
                      const int nThreads );

    //
    // Given interface-mesh intersection points (ifp, size n_ifp), computes the
    // signed distance function (phi_n, size m_meshSize).
    // Note that phi_n must contain the old (pre-reset) distances.
    //
    void        resetPhi( const float* ifp, const int n_ifp, double* phi_n );


private:

    const float*    h_p;        // host mesh node positions
    float*          d_p;        // device mesh node positions (unused: held by the handle)
    double*         d_phi;      // device level set function (unused: held by the handle)

    int             m_meshSize; // number of mesh nodes.

    int             m_devCount; // number of HIP devices
    int             m_threads;  // max. number of threads for host-side operations

};
