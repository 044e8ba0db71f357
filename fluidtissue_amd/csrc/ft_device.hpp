// ft_device.hpp -- device-side building blocks shared by the reinit kernels:
// bit tests, wave64/block scans, Morton codes, box bounds and the
// point/triangle distance functions.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace ftd {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int sign_bit(double x)
{
    return (int)((unsigned long long)__double_as_longlong(x) >> 63);
}

__device__ __forceinline__ int get_bit(const uint64_t* bits, int64_t i)
{
    return (int)((bits[i >> 6] >> (i & 63)) & 1ull);
}

// Load of element i of a read-only array through the constant address space:
// at a wave-uniform address this compiles to scalar (SMEM) loads into SGPRs.
template <class T>
__device__ __forceinline__ T cload(const T* p, int64_t i)
{
    static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
    T v;
    unsigned* d = reinterpret_cast<unsigned*>(&v);
    const __attribute__((address_space(4))) unsigned* src =
        (const __attribute__((address_space(4))) unsigned*)(p + i);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) d[k] = src[k];
    return v;
}

// ---- scans ----------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int d, int width = kWave)
{
    const unsigned lo = __shfl_up((unsigned)(v & 0xffffffffu), d, width);
    const unsigned hi = __shfl_up((unsigned)(v >> 32), d, width);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_up_any(uint32_t v, int d, int width = kWave)
{
    return __shfl_up(v, d, width);
}
__device__ __forceinline__ uint64_t shfl_up_any(uint64_t v, int d, int width = kWave)
{
    return shfl_up_u64(v, d, width);
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const T t = shfl_up_any(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan across a block of BS threads; *total receives the block sum.
// lds must hold BS/64 + 1 elements.  Contains __syncthreads: call from all
// threads of the block.
template <class T, int BS>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T* total)
{
    constexpr int NW = BS / kWave;
    const T inc = wave_incl_scan(v);
    const int w = threadIdx.x / kWave, lane = lane_id();
    if (lane == kWave - 1) lds[w] = inc;
    __syncthreads();
    if (w == 0) {
        const T s = lane < NW ? lds[lane] : T(0);
        const T si = wave_incl_scan(s);
        if (lane < NW) lds[lane] = si - s;
        if (lane == NW - 1) lds[NW] = si;
    }
    __syncthreads();
    const T res = inc - v + lds[w];
    *total = lds[NW];
    __syncthreads();
    return res;
}

// ---- Morton codes (x in the highest bit of each triple) --------------------
__device__ __forceinline__ uint32_t part1by2(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return (part1by2(x) << 2) | (part1by2(y) << 1) | part1by2(z);
}

// Uniform grid description (device copy of the host's choice).
struct Grid {
    float lo[3];
    float inv_b;          // 1 / bin edge
    int dim[3];
    float h;              // the field's node spacing
};

__device__ __forceinline__ uint32_t grid_code(const Grid& g, float x, float y, float z)
{
    const float c[3] = {x, y, z};
    uint32_t b[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float f = (c[k] - g.lo[k]) * g.inv_b;
        int i = (f >= 0.0f) ? (int)f : 0;           // NaN -> 0
        i = i < g.dim[k] ? i : g.dim[k] - 1;
        b[k] = (uint32_t)i;
    }
    return morton3(b[0], b[1], b[2]);
}

// Field of element x given the exclusive offsets off[0..nf] (off[0] = 0):
// the largest k < nf with off[k] <= x.
__device__ __forceinline__ int field_of(const int64_t* off, int nf, int64_t x)
{
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Squared distance from x to an axis-aligned box (0 inside).
__device__ __forceinline__ float box_dist2(float x, float y, float z, const float* lo,
                                           const float* hi)
{
    const float dx = fmaxf(fmaxf(lo[0] - x, x - hi[0]), 0.0f);
    const float dy = fmaxf(fmaxf(lo[1] - y, y - hi[1]), 0.0f);
    const float dz = fmaxf(fmaxf(lo[2] - z, z - hi[2]), 0.0f);
    return __fmaf_rn(dz, dz, __fmaf_rn(dy, dy, dx * dx));
}

// ---- bounding volumes --------------------------------------------------------
// An axis-aligned box plus a cylinder (centre c, unit axis n, radius rho,
// half-thickness tau) that both contain every primitive below a tree node.
// A patch of a smooth surface lies in a thin cylinder around its normal, so
// for a node far away the cylinder bound is close to the true distance where
// the box bound is loose by about the patch size.  tau < 0: no cylinder.
struct __align__(16) Bound {
    float lo[3], hi[3];
    float c[3], n[3];
    float rho, tau;
    float pad[2];
};

// Square root for lower bounds only: the hardware v_sqrt_f32 (within 1 ulp),
// not the correctly rounded sequence the build flags give sqrtf.  Every bound
// that uses it subtracts a relative slack of 2^-20 or more afterwards, which
// covers the 2^-23 relative error; a flushed tiny input gives 0, still a
// lower bound.
__device__ __forceinline__ float sqrt_lb(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ float cyl_dist2(float x, float y, float z, const Bound& b)
{
    if (b.tau < 0.0f) return 0.0f;
    const float dx = x - b.c[0], dy = y - b.c[1], dz = z - b.c[2];
    const float hn = __fmaf_rn(dz, b.n[2], __fmaf_rn(dy, b.n[1], dx * b.n[0]));
    const float cx = __fmaf_rn(b.n[1], dz, -b.n[2] * dy);
    const float cy = __fmaf_rn(b.n[2], dx, -b.n[0] * dz);
    const float cz = __fmaf_rn(b.n[0], dy, -b.n[1] * dx);
    const float lat = sqrt_lb(__fmaf_rn(cz, cz, __fmaf_rn(cy, cy, cx * cx)));
    const float slack = 0x1.0p-20f * (fabsf(hn) + lat);     // rounding of hn and lat
    const float a = fmaxf(fabsf(hn) - b.tau - slack, 0.0f);
    const float r = fmaxf(lat - b.rho - slack, 0.0f);
    return __fmaf_rn(r, r, a * a);
}

// ---- FP64 vector helpers (index build, FP64 settle) -------------------------
__device__ inline double dot3d(const double* u, const double* v)
{
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}
__device__ inline void cross3d(const double* u, const double* v, double* w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

// ---- FP64 point-triangle distance (the search's final word) -----------------
// Squared distance from x to triangle abc (float points, arithmetic in
// FP64): the plane distance when the projection of x falls inside (edge
// functions n.(e x (x - v)) >= 0 for the three edges), else the nearest of
// the three edge segments; nearly collinear vertices (|n|^2 <= 1e-14 lmax^2)
// use the segments only.  Each difference (x - v, an edge) is formed from the
// float inputs where it is used, the same way each time, and each segment
// distance is taken beside its edge function, so few doubles are live at
// once: this runs at the end of the search kernel, whose occupancy depends on
// its register count.
__device__ inline void diff3d(const float* u, const float* v, double* w)
{
    for (int k = 0; k < 3; k++) w[k] = (double)u[k] - (double)v[k];
}

// Squared distance from the point at ax (relative to a) to the segment a,
// a + ab.
__device__ inline double seg_dist2_rel(const double* ax, const double* ab)
{
    const double l2 = dot3d(ab, ab);
    double t = l2 > 0.0 ? dot3d(ax, ab) / l2 : 0.0;
    t = fmin(fmax(t, 0.0), 1.0);
    double r = 0.0;
    for (int k = 0; k < 3; k++) { const double q = ax[k] - t * ab[k]; r += q * q; }
    return r;
}

__device__ inline double tri_dist2_f64(const float* x, const float* a, const float* b, const float* c)
{
    double n[3], nn, lmax;
    {
        double e0[3], e1[3], e2[3];
        diff3d(b, a, e0);
        diff3d(c, b, e1);
        diff3d(a, c, e2);
        cross3d(e0, e1, n);
        nn = dot3d(n, n);
        lmax = fmax(dot3d(e0, e0), fmax(dot3d(e1, e1), dot3d(e2, e2)));
    }
    bool inside = nn > 1e-14 * lmax * lmax;
    const float* v[3] = {a, b, c};
    const float* w[3] = {b, c, a};
    double hgt = 0.0, sd[3];
    for (int q = 0; q < 3; q++) {
        double xv[3], e[3], t[3];
        diff3d(x, v[q], xv);
        diff3d(w[q], v[q], e);
        if (q == 0) hgt = dot3d(n, xv);
        cross3d(e, xv, t);
        inside = inside && dot3d(n, t) >= 0.0;
        sd[q] = seg_dist2_rel(xv, e);
    }
    if (inside) return hgt * hgt / nn;
    return fmin(sd[0], fmin(sd[1], sd[2]));
}

// ---- FP32 point-triangle distance -------------------------------------------
// Triangle record (80 B), an orthonormal frame at vertex a, built in FP64 when
// the index is built, with u along ab (the longest edge) and n the unit
// normal; in it the triangle is (0,0), (B,0), (Cx,Cy), Cy >= 0:
//   F0 = (a, +-R)  F1 = (u, B)  F2 = (v, n.x)  F3 = (n.y, n.z, Cx, Cy)
//   F4 = (1/B, 1/|ac|^2, 1/|bc|^2, Cx - B)
// (inverse lengths are 0 for zero-length edges; R is the radius, rounded up,
// of the disk about the midpoint of ab that holds the triangle; its sign is
// negative for collinear vertices, which have no interior and whose distance
// is that of the edge segments).  The culling test reads F0-F3 only, one
// contiguous run of scalar loads.  The point is
// projected once (X, Y, Z); the in-plane distance is 0 inside the triangle,
// else the nearest clamped edge projection; d^2 = d2D^2 + Z^2.  With ab the
// longest edge the 2D problem stays well conditioned however thin the
// triangle is, so no case needs FP64 at query time.
struct TriRec {
    float4 F0, F1, F2, F3, F4;
};

// Projection of the point into the triangle's frame.
struct TriProj {
    float X, Y, Z;
};
__device__ __forceinline__ TriProj tri_project(float x, float y, float z, const TriRec& r)
{
    const float wx = x - r.F0.x, wy = y - r.F0.y, wz = z - r.F0.z;
    TriProj p;
    p.X = __fmaf_rn(wz, r.F1.z, __fmaf_rn(wy, r.F1.y, wx * r.F1.x));
    p.Y = __fmaf_rn(wz, r.F2.z, __fmaf_rn(wy, r.F2.y, wx * r.F2.x));
    p.Z = __fmaf_rn(wz, r.F3.y, __fmaf_rn(wy, r.F3.x, wx * r.F2.w));
    return p;
}

// Squared distance to the rectangle [0, B] x [0, Cy] of the triangle's frame,
// which holds it: ab is the longest edge, so C projects inside [0, B], and
// Cy >= 0 (0 for collinear vertices, a segment).  Tighter than a disk about
// the triangle for thin triangles, and with no square root.  As computed it
// may exceed the exact value by its rounding, within a relative 2^-18: the
// caller compares it with a bound raised by that much (rect_bound).
__device__ __forceinline__ float tri_rect_d2(const TriProj& p, const TriRec& r)
{
    const float dx = fmaxf(fmaxf(-p.X, p.X - r.F1.w), 0.0f);
    const float dy = fmaxf(fmaxf(-p.Y, p.Y - r.F3.w), 0.0f);
    return __fmaf_rn(dx, dx, __fmaf_rn(dy, dy, p.Z * p.Z));
}

// The bound to compare tri_rect_d2 with: b (1 + 2^-17), which after rounding
// is still at least b / (1 - 2^-18), so that d2 (1 - 2^-18) <= b implies
// d2 <= rect_bound(b).  Computed once per change of b, not per triangle.
__device__ __forceinline__ float rect_bound(float b) { return b * (1.0f + 0x1.0p-17f); }

// Larger of two non-negative floats (squared distances): ordered as their
// bit patterns, so one integer max gives fmaxf's value without the NaN
// canonicalisation fmaxf costs.
__device__ __forceinline__ float max_nonneg(float a, float b)
{
    return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

__device__ __forceinline__ float tri_dist2_from(const TriProj& p, const TriRec& r)
{
    const float X = p.X, Y = p.Y, Z = p.Z;
    const float B = r.F1.w, Cx = r.F3.z, Cy = r.F3.w, dCx = r.F4.w;
    const float xb = X - B;
    const bool inside = (int)(Y >= 0.0f) & (int)(__fmaf_rn(Cy, X, -Cx * Y) >= 0.0f) &
                        (int)(__fmaf_rn(dCx, Y, -Cy * xb) >= 0.0f) & (int)(r.F0.w > 0.0f);
    // edge ab
    float t = fminf(fmaxf(X * r.F4.x, 0.0f), 1.0f);
    float ex = __fmaf_rn(-t, B, X);
    float d = __fmaf_rn(Y, Y, ex * ex);
    // edge ac
    t = fminf(fmaxf(__fmaf_rn(Y, Cy, X * Cx) * r.F4.y, 0.0f), 1.0f);
    ex = __fmaf_rn(-t, Cx, X);
    float ey = __fmaf_rn(-t, Cy, Y);
    d = fminf(d, __fmaf_rn(ey, ey, ex * ex));
    // edge bc
    t = fminf(fmaxf(__fmaf_rn(Y, Cy, xb * dCx) * r.F4.z, 0.0f), 1.0f);
    ex = __fmaf_rn(-t, dCx, xb);
    ey = __fmaf_rn(-t, Cy, Y);
    d = fminf(d, __fmaf_rn(ey, ey, ex * ex));
    return __fmaf_rn(Z, Z, inside ? 0.0f : d);
}

}  // namespace ftd
