// ft_reinit.hip -- gfx950 level-set reinitialization behind the C ABI of
// include/ftreinit.h.
//
// Pipeline of one reinit (LevelSet::resetPhi_, src/fem/LevelSet.cpp:429-455),
// all on the device, mesh resident in HBM:
//   1 k_node_signs    phi -> sign bitmap (1 bit/node)
//   2 k_classify      one pass over all tets: crossing bitmap + per-block
//                     (crossing tets, triangles) counts      (LevelSet.cpp:377-383)
//   3 scan            block counts -> exclusive offsets
//   4 k_compact       crossing tets in increasing tet order + triangle offsets
//   5 k_xtet          crossing tets: t2e/edges -> crossing edges in t2e order,
//                     checked, marked in an edge bitmap       (LevelSet.cpp:385-405)
//   6 scan            edge-bitmap popcounts -> rank of each crossing edge, i.e.
//                     the ifp index in increasing edge order  (LevelSet.cpp:278-299)
//   7 k_points        FP64 crossing points, ifp2e, retained-node bitmap
//                                                            (LevelSet.cpp:281-298, 443-447)
//   8 k_tris          triangles in tet order with the reference's orientation
//                     rule, normal lengths                   (LevelSet.cpp:330-347, 393-399)
//   9 p2t / area      per-point triangle lists (CSR, ascending) and the
//                     ordered area sums                      (LevelSet.cpp:349-356)
//  10 index           uniform-bin counting sort of the primitives (points or
//                     triangles), leaves = non-empty bins in Morton order,
//                     binary radix tree over the leaves
//  11 k_query         per node (Morton-sorted order): exact nearest primitive,
//                     sign from phi < 0, retained values and the blend fused
//                     in the epilogue (EuclideanDistance.cu:45-58,
//                     LevelSet.cpp:438-448, 74, 148)
// The whole file is compiled with -ffp-contract=off: products and sums are
// rounded one at a time unless an fma is written out, so the bit-exact parts
// (interpolation, orientation, areas, point distances) match the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "ftreinit.h"
#include "ft_device.hpp"
#include "ft_internal.hpp"

using ftd::Grid;
using ftd::TriRec;
using ftd::Bound;

namespace {

constexpr int kBS = 256;                 // threads per block of the streaming kernels
constexpr int kScanTile = 2048;          // elements per scan block (256 x 8)
constexpr int kStack = 24;               // traversal stack depth (tree depth <= 24)
constexpr int kQueryBS = 256;
constexpr int kMaxBinsLog2PerAxis = 8;   // <= 256 bins per axis, 2^24 Morton cells

#define FT_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ft::set_error(FT_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define FT_TRY(...)                     \
    do {                                \
        int rc_ = (__VA_ARGS__);        \
        if (rc_ != FT_OK) return rc_;   \
    } while (0)

inline unsigned grid1d(int64_t n, int bs = kBS) { return (unsigned)((n + bs - 1) / bs); }

// ============================================================================
// Generic kernels: fill and exclusive scan
// ============================================================================
template <class T>
__global__ __launch_bounds__(kBS) void k_fill(T* p, int64_t n, T v)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i < n) p[i] = v;
}

struct LoadSame {
    template <class T> __device__ T operator()(T v) const { return v; }
};
struct LoadPopc {  // u64 bitmap word -> popcount
    __device__ uint32_t operator()(uint64_t v) const { return (uint32_t)__popcll(v); }
};
struct LoadNonEmpty {  // bin count -> (non-empty << 32) | count
    __device__ uint64_t operator()(uint32_t c) const { return ((uint64_t)(c > 0) << 32) | c; }
};

template <class TI, class TO, class Ld>
__global__ __launch_bounds__(kBS) void k_scan_reduce(const TI* in, int64_t n, TO* bsum, Ld ld)
{
    __shared__ TO lds[kBS / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    TO s = 0;
#pragma unroll
    for (int k = 0; k < kScanTile / kBS; k++) {
        const int64_t i = base + (int64_t)k * kBS + threadIdx.x;
        if (i < n) s += (TO)ld(in[i]);
    }
    TO tot;
    ftd::block_excl_scan<TO, kBS>(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <class TO>
__global__ __launch_bounds__(1024) void k_scan_top(TO* bsum, int64_t nb, TO* total)
{
    __shared__ TO lds[1024 / 64 + 1];
    TO carry = 0;
    for (int64_t c = 0; c < nb; c += 1024) {
        const int64_t i = c + threadIdx.x;
        const TO v = i < nb ? bsum[i] : TO(0);
        TO tot;
        const TO ex = ftd::block_excl_scan<TO, 1024>(v, lds, &tot);
        if (i < nb) bsum[i] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class TI, class TO, class Ld>
__global__ __launch_bounds__(kBS) void k_scan_down(const TI* in, TO* out, int64_t n,
                                                   const TO* bsum, Ld ld)
{
    __shared__ TO lds[kBS / 64 + 1];
    constexpr int IPT = kScanTile / kBS;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * IPT;
    TO v[IPT];
    TO s = 0;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int64_t i = base + k;
        v[k] = i < n ? (TO)ld(in[i]) : TO(0);
        s += v[k];
    }
    TO tot;
    TO ex = ftd::block_excl_scan<TO, kBS>(s, lds, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int64_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

// ============================================================================
// Mesh set-up: Morton order of the nodes (processing order of k_query)
// ============================================================================
__global__ __launch_bounds__(kBS) void k_node_bin(const float* xyz, int64_t n, Grid g,
                                                  uint32_t* code, uint32_t* count)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = ftd::grid_code(g, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    code[i] = c;
    atomicAdd(&count[c], 1u);
}

__global__ __launch_bounds__(kBS) void k_node_scatter(const float* xyz, int64_t n,
                                                      const uint32_t* code, uint32_t* count,
                                                      const uint32_t* offset, int32_t* perm,
                                                      float* sx, float* sy, float* sz)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = code[i];
    const uint32_t slot = offset[c] + atomicSub(&count[c], 1u) - 1u;
    perm[slot] = (int32_t)i;
    sx[slot] = xyz[3 * i];
    sy[slot] = xyz[3 * i + 1];
    sz[slot] = xyz[3 * i + 2];
}

// ============================================================================
// Marching tets (LevelSet::findInterfaceNodes_ / reconstructInterface_)
// ============================================================================
__global__ __launch_bounds__(kBS) void k_node_signs(const double* phi, int64_t n, uint64_t* bits)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    const int neg = (i < n) ? ftd::sign_bit(phi[i]) : 0;   // std::signbit (LevelSet.cpp:281,380)
    const uint64_t m = __ballot(neg);
    if (ftd::lane_id() == 0 && i < n) bits[i >> 6] = m;
}

// One tet per thread.  Crossing iff 1..3 of its nodes are negative
// (LevelSet.cpp:380-383); 2 negative nodes -> 4 crossing edges -> 2 triangles.
__global__ __launch_bounds__(kBS) void k_classify(const int4* tets, int64_t T,
                                                  const uint64_t* sbits, uint64_t* tbits,
                                                  uint64_t* blk)
{
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    int s = 0;
    if (t < T) {
        const int4 q = tets[t];
        s = ftd::get_bit(sbits, q.x) + ftd::get_bit(sbits, q.y) + ftd::get_bit(sbits, q.z) +
            ftd::get_bit(sbits, q.w);
    }
    const bool cross = (t < T) && s != 0 && s != 4;
    const uint64_t m = __ballot(cross);
    const uint64_t m2 = __ballot(cross && s == 2);
    if (ftd::lane_id() == 0 && t < T) tbits[t >> 6] = m;
    __shared__ uint64_t part[kBS / 64];
    if (ftd::lane_id() == 0)
        part[threadIdx.x / 64] = ((uint64_t)__popcll(m) << 32) | (uint64_t)(__popcll(m) + __popcll(m2));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s4 = 0;
        for (int w = 0; w < kBS / 64; w++) s4 += part[w];
        blk[blockIdx.x] = s4;
    }
}

// One thread per 64-tet word of the crossing bitmap.  Lanes 4k..4k+3 share a
// k_classify block, whose exclusive offset is blk_prefix[w/4].
__global__ __launch_bounds__(kBS) void k_compact(const int4* tets, const uint64_t* sbits,
                                                 const uint64_t* tbits, int64_t nwords,
                                                 const uint64_t* blk_prefix, int32_t* xtet,
                                                 int32_t* xtri_off)
{
    const int64_t w = (int64_t)blockIdx.x * kBS + threadIdx.x;
    const uint64_t word = w < nwords ? tbits[w] : 0ull;
    uint32_t ct = 0;
    uint64_t m = word;
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int4 q = tets[w * 64 + b];
        const int s = ftd::get_bit(sbits, q.x) + ftd::get_bit(sbits, q.y) +
                      ftd::get_bit(sbits, q.z) + ftd::get_bit(sbits, q.w);
        ct += (s == 2) ? 2u : 1u;
    }
    const uint64_t mine = ((uint64_t)__popcll(word) << 32) | ct;
    uint64_t inc = mine;
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        const uint64_t t = ftd::shfl_up_u64(inc, o, 4);
        if ((threadIdx.x & 3) >= o) inc += t;
    }
    if (w >= nwords || word == 0) return;
    const uint64_t base = blk_prefix[w >> 2] + (inc - mine);
    uint32_t bx = (uint32_t)(base >> 32), bt = (uint32_t)base;
    m = word;
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int64_t t = w * 64 + b;
        const int4 q = tets[t];
        const int s = ftd::get_bit(sbits, q.x) + ftd::get_bit(sbits, q.y) +
                      ftd::get_bit(sbits, q.z) + ftd::get_bit(sbits, q.w);
        xtet[bx] = (int32_t)t;
        xtri_off[bx] = (int32_t)bt;
        bx++;
        bt += (s == 2) ? 2u : 1u;
    }
}

// Error word: (code << 40) | tet, min-reduced so the lowest failing tet wins
// within a code; codes: 1 = t2e (FT_ERR_T2E), 2 = degenerate tet.
__device__ __forceinline__ void flag_error(unsigned long long* err, int code, int64_t t)
{
    atomicMin(err, ((unsigned long long)code << 40) | (unsigned long long)t);
}

// Crossing tets: collect crossing edges in t2e order (LevelSet.cpp:385-391),
// check the count (3 or 4, and consistent with the node signs used for the
// triangle offsets), mark them in the edge bitmap.
__global__ __launch_bounds__(kBS) void k_xtet(const int4* tets, const int2* t2e3,
                                              const int2* edges, const uint64_t* sbits,
                                              const int32_t* xtet, int64_t X,
                                              unsigned long long* ebits, int4* xedges,
                                              unsigned long long* err)
{
    const int64_t x = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (x >= X) return;
    const int64_t t = xtet[x];
    const int2 r0 = t2e3[3 * t], r1 = t2e3[3 * t + 1], r2 = t2e3[3 * t + 2];
    const int te[6] = {r0.x, r0.y, r1.x, r1.y, r2.x, r2.y};
    int ce[4] = {-1, -1, -1, -1};
    int n = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int2 ab = edges[te[j]];
        if (ftd::get_bit(sbits, ab.x) != ftd::get_bit(sbits, ab.y)) {
            if (n < 4) ce[n] = te[j];
            n++;
        }
    }
    const int4 q = tets[t];
    const int s = ftd::get_bit(sbits, q.x) + ftd::get_bit(sbits, q.y) + ftd::get_bit(sbits, q.z) +
                  ftd::get_bit(sbits, q.w);
    const int expect = (s == 2) ? 4 : 3;
    if (n != expect) {
        flag_error(err, 1, t);
        xedges[x] = make_int4(-1, -1, -1, -1);
        return;
    }
    for (int k = 0; k < n; k++)
        atomicOr(&ebits[ce[k] >> 6], 1ull << (ce[k] & 63));
    xedges[x] = make_int4(ce[0], ce[1], ce[2], ce[3]);
}

__device__ __forceinline__ int32_t edge_rank(const uint64_t* ebits, const uint32_t* eprefix,
                                             int32_t e)
{
    const uint64_t wv = ebits[e >> 6];
    const uint64_t below = (e & 63) ? (wv & ((1ull << (e & 63)) - 1ull)) : 0ull;
    return (int32_t)(eprefix[e >> 6] + (uint32_t)__popcll(below));
}

// Crossing points in increasing edge order (LevelSet.cpp:278-299): one thread
// per 64-edge word.  FP64 interpolation, no fma; marks both edge nodes as
// retained (LevelSet.cpp:443-447).
__global__ __launch_bounds__(kBS) void k_points(const uint64_t* ebits, const uint32_t* eprefix,
                                                int64_t nwords, const int2* edges,
                                                const float* xyz, const double* phi,
                                                float* ifp, int32_t* ifp2e,
                                                unsigned long long* retain)
{
    const int64_t w = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (w >= nwords) return;
    uint64_t m = ebits[w];
    uint32_t r = eprefix[w];
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int64_t e = w * 64 + b;
        const int2 ab = edges[e];
        const double pa = phi[ab.x], pb = phi[ab.y];
        double t = 0.5;
        const double div = pa - pb;
        if (fabs(div) > (double)FLT_MIN) t = pa / div;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double p0 = (double)xyz[3 * (int64_t)ab.x + c];
            const double p1 = (double)xyz[3 * (int64_t)ab.y + c];
            const double v = (1.0 - t) * p0 + t * p1;
            ifp[3 * (int64_t)r + c] = (float)v;
        }
        ifp2e[r] = (int32_t)e;
        atomicOr(&retain[ab.x >> 6], 1ull << (ab.x & 63));
        atomicOr(&retain[ab.y >> 6], 1ull << (ab.y & 63));
        r++;
    }
}

// addTriangle (LevelSet.cpp:330-357): float normal by Eigen's cross product,
// orientation toward phi > 0 through the first tet node with |phi| > FLT_MIN.
__device__ __forceinline__ void emit_tri(int32_t n0, int32_t n1, int32_t n2, int32_t pref,
                                         const float* xyz, const double* phi, const float* ifp,
                                         int32_t f, int4* tris, float* tri_norm)
{
    const float* q0 = ifp + 3 * (int64_t)n0;
    const float* q1 = ifp + 3 * (int64_t)n1;
    const float* q2 = ifp + 3 * (int64_t)n2;
    const float e1x = q1[0] - q0[0], e1y = q1[1] - q0[1], e1z = q1[2] - q0[2];
    const float e2x = q2[0] - q0[0], e2y = q2[1] - q0[1], e2z = q2[2] - q0[2];
    const float nx = e1y * e2z - e1z * e2y;
    const float ny = e1z * e2x - e1x * e2z;
    const float nz = e1x * e2y - e1y * e2x;
    const float* pr = xyz + 3 * (int64_t)pref;
    const float rx = pr[0] - q0[0], ry = pr[1] - q0[1], rz = pr[2] - q0[2];
    const float dot = nx * rx + (ny * ry + nz * rz);     // Eigen redux: x0 + (x1 + x2)
    const int sd = (int)(__float_as_uint(dot) >> 31);
    int4 out;
    out.x = n0;
    if (sd != ftd::sign_bit(phi[pref])) { out.y = n2; out.z = n1; }
    else { out.y = n1; out.z = n2; }
    out.w = -1;
    tris[f] = out;
    const float sq = nx * nx + (ny * ny + nz * nz);
    tri_norm[f] = sqrtf(sq);
}

__global__ __launch_bounds__(kBS) void k_tris(const int4* tets, const float* xyz,
                                              const double* phi, const int32_t* xtet,
                                              const int32_t* xtri_off, const int4* xedges,
                                              int64_t X, const uint64_t* ebits,
                                              const uint32_t* eprefix, const float* ifp,
                                              int4* tris, float* tri_norm, int32_t* if_elems,
                                              unsigned long long* err)
{
    const int64_t x = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (x >= X) return;
    const int64_t t = xtet[x];
    if_elems[x] = (int32_t)t;
    const int4 ce = xedges[x];
    if (ce.x < 0) return;                                  // flagged by k_xtet
    const int4 q = tets[t];
    const int nd[4] = {q.x, q.y, q.z, q.w};
    int k = 0;
    while (k < 4 && fabs(phi[nd[k]]) <= (double)FLT_MIN) k++;
    if (k == 4) { flag_error(err, 2, t); return; }
    const int32_t pref = nd[k];
    const int32_t v0 = edge_rank(ebits, eprefix, ce.x);
    const int32_t v1 = edge_rank(ebits, eprefix, ce.y);
    const int32_t v2 = edge_rank(ebits, eprefix, ce.z);
    const int32_t f = xtri_off[x];
    if (ce.w < 0) {
        emit_tri(v0, v1, v2, pref, xyz, phi, ifp, f, tris, tri_norm);          // :393-395
    } else {
        const int32_t v3 = edge_rank(ebits, eprefix, ce.w);
        emit_tri(v0, v2, v3, pref, xyz, phi, ifp, f, tris, tri_norm);          // :397
        emit_tri(v0, v1, v3, pref, xyz, phi, ifp, f + 1, tris, tri_norm);      // :398
    }
}

__global__ __launch_bounds__(kBS) void k_p2t_count(const int4* tris, int64_t F, uint32_t* deg)
{
    const int64_t f = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (f >= F) return;
    const int4 t = tris[f];
    atomicAdd(&deg[t.x], 1u);
    atomicAdd(&deg[t.y], 1u);
    atomicAdd(&deg[t.z], 1u);
}

__global__ __launch_bounds__(kBS) void k_p2t_fill(const int4* tris, int64_t F,
                                                  const uint32_t* off, uint32_t* cursor,
                                                  int32_t* p2t)
{
    const int64_t f = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (f >= F) return;
    const int4 t = tris[f];
    const int v[3] = {t.x, t.y, t.z};
    for (int k = 0; k < 3; k++) {
        const uint32_t slot = atomicAdd(&cursor[v[k]], 1u);
        p2t[off[v[k]] + slot] = (int32_t)f;
    }
}

// Per point: sort its triangle list ascending (= emission order, the order of
// p2t in the reference) and sum the areas in that order, each step as
// float(double(area) + 0.5 * double(norm)) (LevelSet.cpp:356).
__global__ __launch_bounds__(kBS) void k_point_area(const uint32_t* off, int64_t P,
                                                    int32_t* p2t, const float* tri_norm,
                                                    float* area)
{
    const int64_t v = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (v >= P) return;
    const uint32_t b = off[v], e = off[v + 1];
    for (uint32_t i = b + 1; i < e; i++) {
        const int32_t key = p2t[i];
        uint32_t j = i;
        while (j > b && p2t[j - 1] > key) { p2t[j] = p2t[j - 1]; j--; }
        p2t[j] = key;
    }
    float a = 0.0f;
    for (uint32_t i = b; i < e; i++)
        a = (float)((double)a + 0.5 * (double)tri_norm[p2t[i]]);
    area[v] = a;
}

// ============================================================================
// Spatial index: uniform bins by counting sort, binary radix tree on bins
// ============================================================================
struct __align__(16) BNode {  // child bounds, children, parent
    Bound l, r;
    int32_t left, right, parent, pad;
};

__device__ __forceinline__ void tri_vertices(const int4* tris, const float* ifp, int32_t f,
                                             float* a, float* b, float* c)
{
    const int4 t = tris[f];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a[k] = ifp[3 * (int64_t)t.x + k];
        b[k] = ifp[3 * (int64_t)t.y + k];
        c[k] = ifp[3 * (int64_t)t.z + k];
    }
}

// mode 0: points (ifp), mode 1: triangles (centroid).
__global__ __launch_bounds__(kBS) void k_prim_bin(int mode, int64_t n, const float* ifp,
                                                  const int4* tris, Grid g, uint32_t* code,
                                                  uint32_t* count)
{
    const int64_t j = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (j >= n) return;
    float c[3];
    if (mode == 0) {
        for (int k = 0; k < 3; k++) c[k] = ifp[3 * j + k];
    } else {
        float a[3], b[3], d[3];
        tri_vertices(tris, ifp, (int32_t)j, a, b, d);
        for (int k = 0; k < 3; k++) c[k] = (a[k] + b[k] + d[k]) * (1.0f / 3.0f);
    }
    const uint32_t cc = ftd::grid_code(g, c[0], c[1], c[2]);
    code[j] = cc;
    atomicAdd(&count[cc], 1u);
}

__global__ __launch_bounds__(kBS) void k_leaves(const uint32_t* count, const uint64_t* prefix,
                                                int64_t G, uint32_t* leaf_key,
                                                uint32_t* leaf_start, uint32_t* leaf_cnt)
{
    const int64_t c = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (c >= G) return;
    const uint32_t n = count[c];
    if (!n) return;
    const uint64_t p = prefix[c];
    const uint32_t l = (uint32_t)(p >> 32);
    leaf_key[l] = (uint32_t)c;
    leaf_start[l] = (uint32_t)p;
    leaf_cnt[l] = n;
}

// Scatter primitives into bin order and write their search records.
__global__ __launch_bounds__(kBS) void k_prim_scatter(int mode, int64_t n, const float* ifp,
                                                      const int4* tris, const uint32_t* code,
                                                      uint32_t* count, const uint64_t* prefix,
                                                      float4* prec, TriRec* trec,
                                                      int32_t* pidx)
{
    const int64_t j = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (j >= n) return;
    const uint32_t c = code[j];
    const uint32_t slot = (uint32_t)prefix[c] + atomicSub(&count[c], 1u) - 1u;
    pidx[slot] = (int32_t)j;
    if (mode == 0) {
        prec[slot] = make_float4(ifp[3 * j], ifp[3 * j + 1], ifp[3 * j + 2], 0.0f);
        return;
    }
    float v3[3][3];
    tri_vertices(tris, ifp, (int32_t)j, v3[0], v3[1], v3[2]);
    // Rotate (orientation kept) so that ab is the longest edge, then build the
    // frame in FP64: u along ab, n the unit normal, v = n x u.  Collinear or
    // coincident vertices get a frame along ab and no interior: the distance
    // is then that of the edge segments.
    double len[3];
    for (int q = 0; q < 3; q++) {
        const float* p0 = v3[q];
        const float* p1 = v3[(q + 1) % 3];
        double l = 0.0;
        for (int k = 0; k < 3; k++) { const double dd = (double)p1[k] - p0[k]; l += dd * dd; }
        len[q] = l;
    }
    const int r0 = (len[0] >= len[1] && len[0] >= len[2]) ? 0 : (len[1] >= len[2] ? 1 : 2);
    const float* a = v3[r0];
    const float* b = v3[(r0 + 1) % 3];
    const float* d = v3[(r0 + 2) % 3];
    double e0[3], e1[3];
    for (int k = 0; k < 3; k++) { e0[k] = (double)b[k] - a[k]; e1[k] = (double)d[k] - a[k]; }
    double nv[3];
    ftd::cross3d(e0, e1, nv);
    const double nn = ftd::dot3d(nv, nv);
    const double L0 = sqrt(ftd::dot3d(e0, e0));
    double u[3], v[3], nu[3];
    bool interior = nn > 0.0 && L0 > 0.0;
    if (L0 > 0.0) for (int k = 0; k < 3; k++) u[k] = e0[k] / L0;
    else { u[0] = 1.0; u[1] = 0.0; u[2] = 0.0; }
    if (interior) {
        const double inn = 1.0 / sqrt(nn);
        for (int k = 0; k < 3; k++) nu[k] = nv[k] * inn;
        ftd::cross3d(nu, u, v);
    } else {
        // any unit vector orthogonal to u
        const double t0[3] = {fabs(u[0]) < 0.9 ? 1.0 : 0.0, fabs(u[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        double w[3];
        ftd::cross3d(u, t0, w);
        const double wl = sqrt(ftd::dot3d(w, w));
        for (int k = 0; k < 3; k++) v[k] = w[k] / wl;
        ftd::cross3d(u, v, nu);
    }
    const double Cx = ftd::dot3d(e1, u), Cy = interior ? ftd::dot3d(e1, v) : 0.0;
    const double dCx = Cx - L0;
    const double lac = Cx * Cx + Cy * Cy, lbc = dCx * dCx + Cy * Cy;
    TriRec r;
    r.F0 = make_float4(a[0], a[1], a[2], (float)L0);
    r.F1 = make_float4((float)u[0], (float)u[1], (float)u[2], (float)Cx);
    r.F2 = make_float4((float)v[0], (float)v[1], (float)v[2], (float)Cy);
    r.F3 = make_float4((float)nu[0], (float)nu[1], (float)nu[2], L0 > 0.0 ? (float)(1.0 / L0) : 0.0f);
    r.F4 = make_float4(lac > 0.0 ? (float)(1.0 / lac) : 0.0f, lbc > 0.0 ? (float)(1.0 / lbc) : 0.0f,
                       (float)dCx, interior ? 1.0f : 0.0f);
    trec[slot] = r;
}

// Leaf bounds: box of the vertices and, for triangles, a cylinder about the
// area-weighted mean normal of the leaf's triangles.
__global__ __launch_bounds__(kBS) void k_leaf_bounds(int mode, int64_t L, const uint32_t* leaf_start,
                                                     const uint32_t* leaf_cnt, const int32_t* pidx,
                                                     const float* ifp, const int4* tris, Bound* leaf_b)
{
    const int64_t l = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (l >= L) return;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    double ns[3] = {0.0, 0.0, 0.0};
    const uint32_t b = leaf_start[l], e = b + leaf_cnt[l];
    for (uint32_t s = b; s < e; s++) {
        const int32_t j = pidx[s];
        if (mode == 0) {
            for (int k = 0; k < 3; k++) {
                const float v = ifp[3 * (int64_t)j + k];
                lo[k] = fminf(lo[k], v); hi[k] = fmaxf(hi[k], v);
            }
        } else {
            float va[3], vb[3], vc[3];
            tri_vertices(tris, ifp, j, va, vb, vc);
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], fminf(va[k], fminf(vb[k], vc[k])));
                hi[k] = fmaxf(hi[k], fmaxf(va[k], fmaxf(vb[k], vc[k])));
            }
            double e0[3], e1[3], nn[3];
            for (int k = 0; k < 3; k++) { e0[k] = (double)vb[k] - va[k]; e1[k] = (double)vc[k] - va[k]; }
            ftd::cross3d(e0, e1, nn);
            for (int k = 0; k < 3; k++) ns[k] += nn[k];
        }
    }
    Bound B;
    for (int k = 0; k < 3; k++) { B.lo[k] = lo[k]; B.hi[k] = hi[k]; B.c[k] = 0.5f * (lo[k] + hi[k]); }
    B.pad[0] = B.pad[1] = 0.0f;
    const double nl = sqrt(ftd::dot3d(ns, ns));
    if (mode == 0 || !(nl > 0.0)) {
        B.n[0] = 0.0f; B.n[1] = 0.0f; B.n[2] = 1.0f;
        B.rho = 0.0f; B.tau = -1.0f;                 // box only
    } else {
        for (int k = 0; k < 3; k++) B.n[k] = (float)(ns[k] / nl);
        double tau = 0.0, rho = 0.0;
        const double n[3] = {B.n[0], B.n[1], B.n[2]};
        for (uint32_t s = b; s < e; s++) {
            float v[3][3];
            tri_vertices(tris, ifp, pidx[s], v[0], v[1], v[2]);
            for (int q = 0; q < 3; q++) {
                double d[3], cr[3];
                for (int k = 0; k < 3; k++) d[k] = (double)v[q][k] - B.c[k];
                ftd::cross3d(n, d, cr);
                tau = fmax(tau, fabs(ftd::dot3d(n, d)));
                rho = fmax(rho, sqrt(ftd::dot3d(cr, cr)));
            }
        }
        B.tau = (float)(tau * (1.0 + 1e-6));
        B.rho = (float)(rho * (1.0 + 1e-6));
    }
    leaf_b[l] = B;
}

// Binary radix tree over L >= 2 sorted, distinct keys (T. Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG 2012).
// Children >= 0 are internal nodes, ~l is leaf l.
__device__ __forceinline__ int kdelta(const uint32_t* keys, int L, int i, int j)
{
    if (j < 0 || j >= L) return -1;
    return __clz(keys[i] ^ keys[j]);
}

__global__ __launch_bounds__(kBS) void k_karras(const uint32_t* keys, int L, BNode* nodes,
                                                int32_t* leaf_parent)
{
    const int i = blockIdx.x * kBS + threadIdx.x;
    if (i >= L - 1) return;
    const int d = (kdelta(keys, L, i, i + 1) - kdelta(keys, L, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(keys, L, i, i - d);
    int lmax = 2;
    while (kdelta(keys, L, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, L, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(keys, L, i, j);
    int s = 0, div = 2, t;
    do {
        t = (l + div - 1) / div;
        if (kdelta(keys, L, i, i + (s + t) * d) > dnode) s += t;
        div <<= 1;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int left = (min(i, j) == gamma) ? ~gamma : gamma;
    const int right = (max(i, j) == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    nodes[i].left = left;
    nodes[i].right = right;
    if (left >= 0) nodes[left].parent = i; else leaf_parent[~left] = i;
    if (right >= 0) nodes[right].parent = i; else leaf_parent[~right] = i;
    if (i == 0) nodes[0].parent = -1;
}

__device__ __forceinline__ void st_bound(Bound* dst, const Bound& v)
{
    const unsigned* src = reinterpret_cast<const unsigned*>(&v);
    unsigned* d = reinterpret_cast<unsigned*>(dst);
    for (int k = 0; k < 16; k++)
        __hip_atomic_store(&d[k], src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ Bound ld_bound(const Bound* srcp)
{
    Bound v;
    const unsigned* s = reinterpret_cast<const unsigned*>(srcp);
    unsigned* d = reinterpret_cast<unsigned*>(&v);
    for (int k = 0; k < 16; k++)
        d[k] = __hip_atomic_load(&s[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

// Parent bound of two children: box union and a cylinder that contains both
// child cylinders.  For a child cylinder (c_i, n_i, rho_i, tau_i) seen from
// the parent axis n at centre c, with cos = |n.n_i| and sin = sqrt(1-cos^2):
//   thickness  |n.(c_i-c)| + tau_i cos + rho_i sin
//   radius     |n x (c_i-c)| + rho_i + tau_i sin
__device__ inline Bound merge_bounds(const Bound& A, const Bound& Bb)
{
    Bound P;
    for (int k = 0; k < 3; k++) {
        P.lo[k] = fminf(A.lo[k], Bb.lo[k]);
        P.hi[k] = fmaxf(A.hi[k], Bb.hi[k]);
        P.c[k] = 0.5f * (P.lo[k] + P.hi[k]);
    }
    P.pad[0] = P.pad[1] = 0.0f;
    if (A.tau < 0.0f || Bb.tau < 0.0f) {
        P.n[0] = 0.0f; P.n[1] = 0.0f; P.n[2] = 1.0f;
        P.rho = 0.0f; P.tau = -1.0f;
        return P;
    }
    const double wa = (double)A.rho * A.rho + 1e-30, wb = (double)Bb.rho * Bb.rho + 1e-30;
    double n[3];
    for (int k = 0; k < 3; k++) n[k] = wa * A.n[k] + wb * Bb.n[k];
    double nl = sqrt(ftd::dot3d(n, n));
    if (!(nl > 1e-12 * (wa + wb))) { for (int k = 0; k < 3; k++) n[k] = A.n[k]; nl = 1.0; }
    for (int k = 0; k < 3; k++) n[k] /= nl;
    double tau = 0.0, rho = 0.0;
    const Bound* ch[2] = {&A, &Bb};
    for (int i = 0; i < 2; i++) {
        const Bound& C = *ch[i];
        double dc[3], ni[3], cr[3];
        for (int k = 0; k < 3; k++) { dc[k] = (double)C.c[k] - P.c[k]; ni[k] = C.n[k]; }
        const double cs = fmin(1.0, fabs(ftd::dot3d(n, ni)));
        const double sn = sqrt(fmax(0.0, 1.0 - cs * cs));
        ftd::cross3d(n, dc, cr);
        tau = fmax(tau, fabs(ftd::dot3d(n, dc)) + C.tau * cs + C.rho * sn);
        rho = fmax(rho, sqrt(ftd::dot3d(cr, cr)) + C.rho + C.tau * sn);
    }
    for (int k = 0; k < 3; k++) P.n[k] = (float)n[k];
    P.tau = (float)(tau * (1.0 + 1e-6));
    P.rho = (float)(rho * (1.0 + 1e-6));
    return P;
}

// Bottom-up fitting: the second child to arrive at a node merges both.
__global__ __launch_bounds__(kBS) void k_refit(int L, const int32_t* leaf_parent,
                                               const Bound* leaf_b, BNode* nodes, unsigned* flags)
{
    const int l = blockIdx.x * kBS + threadIdx.x;
    if (l >= L) return;
    Bound cur_b = leaf_b[l];
    int cur = ~l;
    int p = leaf_parent[l];
    for (int depth = 0; p >= 0 && depth < 64; depth++) {   // depth <= 24 for 24-bit keys
        BNode* nd = &nodes[p];
        const bool is_left = (nd->left == cur);
        st_bound(is_left ? &nd->l : &nd->r, cur_b);
        __threadfence();
        const unsigned old = atomicAdd(&flags[p], 1u);
        if (old == 0) return;                       // sibling not done yet
        __threadfence();
        const Bound sib = ld_bound(is_left ? &nd->r : &nd->l);
        cur_b = is_left ? merge_bounds(cur_b, sib) : merge_bounds(sib, cur_b);
        cur = p;
        p = __hip_atomic_load(&nd->parent, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ============================================================================
// Nearest-primitive search with the fused reinit epilogue
// ============================================================================
enum Epi : int { EPI_RAW = 0, EPI_INIT = 1, EPI_UPDATE = 2 };

struct QueryArgs {
    int64_t n;                       // nodes
    const float *sx, *sy, *sz;       // Morton-ordered coordinates
    const int32_t* perm;             // sorted slot -> node
    const float4* prec;              // point records (mode 0)
    const TriRec* trec;              // triangle records (mode 1)
    const int32_t* pidx;             // sorted slot -> primitive id
    const uint32_t* leaf_start;
    const uint32_t* leaf_cnt;
    const Bound* leaf_b;
    const BNode* nodes;
    int L;                           // leaves
    float band;                      // > 0: narrow band
    const float* ifp;                // FP64 fallback of triangles
    const int4* tris;
    const double* phi_old;           // sign + retained values
    double* phi_out;
    double* phi_ifn;                 // nullable
    const uint64_t* retain;          // nullable
    int32_t* nearest;                // nullable
    int epi;
    double w_phi, w_ifn;
    unsigned long long* n_clamped;
    unsigned* overflow;
    unsigned long long* stats;       // nullable: [tree nodes, leaves, primitives]
    int rep;                         // 1: seeding pass, one query per 64-node tile
    int32_t* seed_out;               // seeding pass: best slot of each tile's query
    const int32_t* seed_in;          // nullable: per-tile candidate tested first
    float h_est;                     // mean node spacing (statistics only)
    int probe;                       // wave kernel: greedy first candidate
};

// Margin of the box prune: a box is skipped only if its lower bound exceeds
// the best distance by more than the float rounding of both (relative) plus
// an absolute floor for near-zero distances.
__device__ __forceinline__ float prune_bound(float best2)
{
    return best2 * (1.0f + 0x1.0p-16f) + 1.0e-30f;
}

template <int MODE>
struct Searcher {
    float x, y, z;
    float best;      // mode 0: distance; mode 1: squared distance
    int32_t bj;      // best primitive (original id)
    const float4* prec;
    const TriRec* trec;
    const int32_t* pidx;
    const uint32_t* leaf_start;
    const uint32_t* leaf_cnt;
    const float* ifp;
    const int4* tris;

    __device__ __forceinline__ float best2() const { return MODE == 0 ? best * best : best; }

    __device__ __forceinline__ void consider(float d, uint32_t s)
    {
        if (d < best || (d == best && d < INFINITY)) {
            const int32_t j = pidx[s];
            if (d < best || bj < 0 || j < bj) { best = d; bj = j; bslot = (int32_t)s; }
        }
    }

    uint32_t n_leaves = 0, n_prims = 0;
    int32_t bslot = -1;

    __device__ void test_prim(uint32_t s)
    {
        if (MODE == 0) {
            const float4 p = prec[s];
            const float dx = p.x - x, dy = p.y - y, dz = p.z - z;
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);   // EuclideanDistance.cu:20-28
            consider(d, s);
        } else {
            const TriRec r = trec[s];
            const float d2 = ftd::tri_dist2_f32(x, y, z, r);
            consider(d2, s);
        }
    }

    // Leaf test at a wave-uniform leaf: records read through scalar loads.
    __device__ void test_leaf_uniform(int l)
    {
        const uint32_t b = ftd::cload(leaf_start, l), e = b + ftd::cload(leaf_cnt, l);
        n_leaves++;
        n_prims += e - b;
        for (uint32_t s = b; s < e; s++) {
            if (MODE == 0) {
                const float4 p = ftd::cload(prec, s);
                const float dx = p.x - x, dy = p.y - y, dz = p.z - z;
                const float d = sqrtf(dx * dx + dy * dy + dz * dz);   // EuclideanDistance.cu:20-28
                consider(d, s);
            } else {
                const TriRec r = ftd::cload(trec, s);
                const ftd::TriProj pj = ftd::tri_project(x, y, z, r);
                // wave-uniform early out when the disk bound rules it out for all lanes
                if (__ballot(ftd::tri_disk_lb2(pj, r) <= prune_bound(best2())) == 0ull) continue;
                const float d2 = ftd::tri_dist2_from(pj, r);
                consider(d2, s);
            }
        }
    }

    __device__ void test_leaf(int l)
    {
        const uint32_t b = leaf_start[l], e = b + leaf_cnt[l];
        n_leaves++;
        n_prims += e - b;
        for (uint32_t s = b; s < e; s++) test_prim(s);
    }
};

// Epilogue of one node (EuclideanDistance.cu:54-57, LevelSet.cpp:438-448, 74,
// 148): sign from phi_old < 0, retained value at crossing-edge nodes, blend.
template <int MODE>
__device__ __forceinline__ void finish_node(const QueryArgs& a, const Searcher<MODE>& S, int64_t i)
{
    float d;
    if (MODE == 0) d = S.best;
    else d = S.bj >= 0 ? sqrtf(S.best) : (a.band > 0.0f ? a.band : FLT_MAX);
    if (a.band > 0.0f && S.bj < 0) {
        d = a.band;
        atomicAdd(a.n_clamped, 1ull);
    }
    const int64_t node = a.perm[i];
    const double po = a.phi_old[node];
    const double dn = (po < 0.0) ? -(double)d : (double)d;
    if (a.phi_ifn) a.phi_ifn[node] = dn;
    if (a.nearest) a.nearest[node] = S.bj;
    double out = dn;
    if (a.epi == EPI_UPDATE) {
        const double kept = (a.retain && ftd::get_bit(a.retain, node)) ? po : dn;
        out = a.w_phi * kept + a.w_ifn * dn;
    }
    a.phi_out[node] = out;
}

template <int MODE>
__device__ __forceinline__ void init_searcher(Searcher<MODE>& S, const QueryArgs& a, float px,
                                              float py, float pz)
{
    S.x = px; S.y = py; S.z = pz;
    S.bj = -1;
    if (MODE == 0) S.best = a.band > 0.0f ? a.band : FLT_MAX;
    else S.best = a.band > 0.0f ? a.band * a.band : INFINITY;
    S.prec = a.prec; S.trec = a.trec; S.pidx = a.pidx;
    S.leaf_start = a.leaf_start; S.leaf_cnt = a.leaf_cnt;
    S.ifp = a.ifp; S.tris = a.tris;
}

template <int MODE>
__device__ __forceinline__ void record_stats(const QueryArgs& a, const Searcher<MODE>& S,
                                             int64_t i, uint32_t n_nodes)
{
    // buckets: distance < 2h, < 8h, < 32h, beyond; +4 when phi < 0
    const float dd = MODE == 0 ? S.best : sqrtf(S.best);
    const float hh = a.h_est;
    int bkt = dd < 2 * hh ? 0 : dd < 8 * hh ? 1 : dd < 32 * hh ? 2 : 3;
    if (a.phi_old[a.perm[i]] < 0.0) bkt += 4;
    unsigned long long* st = a.stats + 4 * bkt;
    atomicAdd(&st[0], 1ull);
    atomicAdd(&st[1], (unsigned long long)n_nodes);
    atomicAdd(&st[2], (unsigned long long)S.n_leaves);
    atomicAdd(&st[3], (unsigned long long)S.n_prims);
}

// Wave-coherent search: the 64 lanes of a wave hold 64 consecutive nodes in
// Morton order (a compact tile) and walk the tree together.  A child is
// entered when any lane's bound admits it; each lane keeps its own best.  All
// control flow is wave-uniform (ballots), node and primitive records are read
// at wave-uniform addresses, the stack is one per wave in LDS.
template <int MODE>
__global__ __launch_bounds__(kQueryBS) void k_query_wave(QueryArgs a)
{
    __shared__ int32_t wstack[kQueryBS / 64][kStack];
    const int w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kQueryBS + (threadIdx.x & ~63);
    if (base >= a.n) return;                          // whole wave out of range
    const int64_t gid = base + (threadIdx.x & 63);
    const int64_t i = gid < a.n ? gid : a.n - 1;      // tail lanes shadow the last node
    const float px = a.sx[i], py = a.sy[i], pz = a.sz[i];
    Searcher<MODE> S;
    init_searcher<MODE>(S, a, px, py, pz);
    if (a.seed_in) {
        const int32_t s0 = a.seed_in[base >> 6];
        if (s0 >= 0) S.test_prim((uint32_t)s0);
    }
    uint32_t n_nodes = 0;
    if (a.L == 1) {
        S.test_leaf_uniform(0);
    } else if (a.L > 1) {
        if (a.probe) {
            // Greedy descent for the tile's middle node to the leaf nearest to it
            // by bound: a first exact candidate for every lane.
            const float cx = __shfl(px, 32), cy = __shfl(py, 32), cz = __shfl(pz, 32);
            int cur = 0;
            for (int depth = 0; depth < kStack + 1 && cur >= 0; depth++) {
                const BNode nd = ftd::cload(a.nodes, cur);
                const float ll = ftd::bound_dist2(cx, cy, cz, nd.l);
                const float lr = ftd::bound_dist2(cx, cy, cz, nd.r);
                cur = __builtin_amdgcn_readfirstlane(ll <= lr ? nd.left : nd.right);
            }
            if (cur < 0) S.test_leaf_uniform(~cur);
        }
        int sp = 0;
        int cur = 0;
        int budget = 2 * a.L + 8;
        while (--budget > 0) {
            const BNode nd = ftd::cload(a.nodes, cur);    // scalar loads: cur is uniform
            const int left = nd.left, right = nd.right;
            const float bnd = prune_bound(S.best2());
            // box bound first; the cylinder only where some lane's box admits it
            float ll = ftd::box_dist2(px, py, pz, nd.l.lo, nd.l.hi);
            float lr = ftd::box_dist2(px, py, pz, nd.r.lo, nd.r.hi);
            if (__ballot(ll <= bnd)) ll = fmaxf(ll, ftd::cyl_dist2(px, py, pz, nd.l));
            if (__ballot(lr <= bnd)) lr = fmaxf(lr, ftd::cyl_dist2(px, py, pz, nd.r));
            n_nodes++;
            const uint64_t mL = __ballot(ll <= bnd), mR = __ballot(lr <= bnd);
            const uint64_t any = mL | mR;
            const bool left_first = 2 * __popcll(__ballot(ll <= lr) & any) >= __popcll(any);
            const int c0 = left_first ? left : right, c1 = left_first ? right : left;
            const uint64_t m0 = left_first ? mL : mR, m1 = left_first ? mR : mL;
            int next = -1;
            if (m0) {
                if (c0 < 0) S.test_leaf_uniform(~c0);
                else next = c0;
            }
            if (m1) {
                if (c1 < 0) S.test_leaf_uniform(~c1);
                else if (next < 0) next = c1;
                else if (sp < kStack) wstack[w][sp++] = c1;
                else atomicOr(a.overflow, 1u);
            }
            if (next < 0) {
                if (sp == 0) break;
                next = wstack[w][--sp];
            }
            cur = __builtin_amdgcn_readfirstlane(next);
        }
        if (budget <= 0) atomicOr(a.overflow, 2u);
    }
    if (a.stats && gid < a.n) record_stats<MODE>(a, S, i, n_nodes);
    if (gid < a.n) finish_node<MODE>(a, S, i);
}

template <int MODE>
__global__ __launch_bounds__(kQueryBS) void k_query(QueryArgs a)
{
    __shared__ int32_t stack_lds[kStack][kQueryBS];
    const int64_t gid = (int64_t)blockIdx.x * kQueryBS + threadIdx.x;
    // Seeding pass: one query per tile of 64 consecutive (Morton-ordered) nodes.
    const int64_t ntiles = (a.n + 63) >> 6;
    if (a.rep ? gid >= ntiles : gid >= a.n) return;
    const int64_t i = a.rep ? min(gid * 64 + 32, a.n - 1) : gid;
    const float px = a.sx[i], py = a.sy[i], pz = a.sz[i];
    Searcher<MODE> S;
    init_searcher<MODE>(S, a, px, py, pz);
    // The tile's seed is one exact candidate: a tight first bound for pruning.
    if (a.seed_in) {
        const int32_t s0 = a.seed_in[i >> 6];
        if (s0 >= 0) S.test_prim((uint32_t)s0);
    }

    if (a.L == 1) {
        const Bound lb = a.leaf_b[0];
        if (ftd::bound_dist2(px, py, pz, lb) <= prune_bound(S.best2())) S.test_leaf(0);
    } else if (a.L > 1) {
        int sp = 0;
        int cur = 0;
        // A binary tree over L leaves has 2L-1 nodes; no traversal visits more.
        int budget = 2 * a.L + 8;
        uint32_t n_nodes = 0;
        while (--budget > 0) {
            const BNode nd = a.nodes[cur];
            n_nodes++;
            const float ll = ftd::bound_dist2(px, py, pz, nd.l);
            const float lr = ftd::bound_dist2(px, py, pz, nd.r);
            int c0 = nd.left, c1 = nd.right;
            float l0 = ll, l1 = lr;
            if (lr < ll) { c0 = nd.right; c1 = nd.left; l0 = lr; l1 = ll; }
            int next = -1;
            // nearer child first; leaves are tested at once
            if (l0 <= prune_bound(S.best2())) {
                if (c0 < 0) S.test_leaf(~c0);
                else next = c0;
            }
            if (l1 <= prune_bound(S.best2())) {
                if (c1 < 0) S.test_leaf(~c1);
                else if (next < 0) next = c1;
                else if (sp < kStack) stack_lds[sp++][threadIdx.x] = c1;
                else atomicOr(a.overflow, 1u);
            }
            if (next < 0) {
                if (sp == 0) break;
                next = stack_lds[--sp][threadIdx.x];
            }
            cur = next;
        }
        if (budget <= 0) atomicOr(a.overflow, 2u);
        if (a.stats && !a.rep) record_stats<MODE>(a, S, i, n_nodes);
    }

    if (a.rep) {
        a.seed_out[gid] = S.bslot;
        return;
    }
    finish_node<MODE>(a, S, i);
}

// ============================================================================
// Host side
// ============================================================================
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct ft_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh
    int64_t N = 0, T = 0, E = 0;
    bool has_conn = false;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    float h_est = 0.0f;
    DevBuf xyz, tets, edges, t2e, perm, sx, sy, sz;
    // per reinit
    DevBuf phi_tmp, phi_ifn_tmp;
    DevBuf sbits, retain, tbits, blk, xtet, xtri_off, xedges, ebits, eprefix;
    DevBuf ifp, ifp2e, tris, tri_norm, if_elems, p2t_off, p2t_cursor, p2t, ifp_area;
    DevBuf nearest;
    // index
    DevBuf bin_count, bin_prefix, pcode, prec, trec, pidx, leaf_key, leaf_start, leaf_cnt;
    DevBuf leaf_b, leaf_parent, bnodes, bflags, seeds, stats;
    // scan scratch and small scalars
    DevBuf scan_bsum, scalars;
    uint64_t* h_scalars = nullptr;   // pinned mirror
    ft_counts last{};
    bool have_interface = false;
    bool have_nearest = false;
    // phase timing (ft_set_timing): events at the phase boundaries of a reinit
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double phase_ms[3] = {0, 0, 0};
    int64_t timed_calls = 0;
};

namespace {

int ensure(DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (bytes <= b.cap) return FT_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 8;
    want = (want + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return ft::set_error(FT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return FT_OK;
}

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

template <class T>
int fill(hipStream_t s, T* p, int64_t n, T v)
{
    if (n <= 0) return FT_OK;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid1d(n)), dim3(kBS), 0, s, p, n, v);
    FT_HIP(hipGetLastError());
    return FT_OK;
}

// Exclusive scan of n elements; the total lands in *d_total (device).
template <class TI, class TO, class Ld>
int scan(ft_ctx* c, hipStream_t s, const TI* in, TO* out, int64_t n, TO* d_total, Ld ld)
{
    const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    FT_TRY(ensure(c->scan_bsum, sizeof(TO) * (size_t)nb));
    TO* bs = c->scan_bsum.as<TO>();
    if (n > 0) {
        hipLaunchKernelGGL((k_scan_reduce<TI, TO, Ld>), dim3((unsigned)nb), dim3(kBS), 0, s, in, n, bs, ld);
    } else {
        FT_TRY(fill<TO>(s, bs, 1, TO(0)));
    }
    hipLaunchKernelGGL(k_scan_top<TO>, dim3(1), dim3(1024), 0, s, bs, nb, d_total);
    if (n > 0)
        hipLaunchKernelGGL((k_scan_down<TI, TO, Ld>), dim3((unsigned)nb), dim3(kBS), 0, s, in, out, n, bs, ld);
    FT_HIP(hipGetLastError());
    return FT_OK;
}

// Uniform cubic grid over the node box with edge >= `edge`, <= 256 per axis.
Grid make_grid(const ft_ctx* c, float edge, int64_t* G)
{
    Grid g{};
    float ext_max = 0.0f;
    for (int k = 0; k < 3; k++) ext_max = std::max(ext_max, c->hi[k] - c->lo[k]);
    float b = std::max(edge, ext_max / (float)(1 << kMaxBinsLog2PerAxis) * 1.0001f);
    if (!(b > 0.0f)) b = 1.0f;
    int lg = 0;
    for (int k = 0; k < 3; k++) {
        int d = (int)std::ceil((c->hi[k] - c->lo[k]) / b);
        d = std::max(1, std::min(d, 1 << kMaxBinsLog2PerAxis));
        g.dim[k] = d;
        g.lo[k] = c->lo[k];
        while ((1 << lg) < d) lg++;
    }
    g.inv_b = 1.0f / b;
    *G = (int64_t)1 << (3 * lg);
    return g;
}

int check_launch() { FT_HIP(hipGetLastError()); return FT_OK; }

// Copies `count` u64 device scalars (ctx->scalars) to the host and waits.
int read_scalars(ft_ctx* c, hipStream_t s, int count)
{
    FT_HIP(hipMemcpyAsync(c->h_scalars, c->scalars.p, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

void mark(ft_ctx* c, int k, hipStream_t s)
{
    if (c->timing) (void)hipEventRecord(c->ev[k], s);
}

// Scalar slots in ctx->scalars.
enum Slot { S_BLK = 0, S_EDGES = 1, S_ERR = 2, S_P2T = 3, S_BINS = 4, S_CLAMP = 5, S_OVF = 6, S_STATS = 8, S_N = 12 };

// Builds the bin index over `n` primitives (mode 0: points in ifp, mode 1:
// triangles tris/ifp) and runs the query with the epilogue.
int search(ft_ctx* c, hipStream_t s, int mode, int64_t n, const float* ifp, const int4* tris,
           float bin_scale, float band, const double* phi_old, double* phi_out, double* phi_ifn,
           const uint64_t* retain, int epi, double w_phi, double w_ifn, bool want_nearest,
           int64_t* n_clamped)
{
    uint64_t* sc = c->scalars.as<uint64_t>();
    int64_t L = 0;
    if (n > 0) {
        const float edge = (bin_scale > 0.0f ? bin_scale : 1.0f) * c->h_est;
        int64_t G = 0;
        const Grid g = make_grid(c, edge, &G);
        FT_TRY(ensure(c->bin_count, sizeof(uint32_t) * G));
        FT_TRY(ensure(c->bin_prefix, sizeof(uint64_t) * G));
        FT_TRY(ensure(c->pcode, sizeof(uint32_t) * n));
        FT_TRY(ensure(c->pidx, sizeof(int32_t) * n));
        if (mode == 0) FT_TRY(ensure(c->prec, sizeof(float4) * n));
        else FT_TRY(ensure(c->trec, sizeof(TriRec) * n));
        uint32_t* cnt = c->bin_count.as<uint32_t>();
        uint64_t* pre = c->bin_prefix.as<uint64_t>();
        FT_TRY(fill<uint32_t>(s, cnt, G, 0u));
        hipLaunchKernelGGL(k_prim_bin, dim3(grid1d(n)), dim3(kBS), 0, s, mode, n, ifp, tris, g,
                           c->pcode.as<uint32_t>(), cnt);
        FT_TRY(scan<uint32_t, uint64_t>(c, s, cnt, pre, G, sc + S_BINS, LoadNonEmpty{}));
        FT_TRY(read_scalars(c, s, S_N));
        L = (int64_t)(c->h_scalars[S_BINS] >> 32);
        FT_TRY(ensure(c->leaf_key, sizeof(uint32_t) * L));
        FT_TRY(ensure(c->leaf_start, sizeof(uint32_t) * L));
        FT_TRY(ensure(c->leaf_cnt, sizeof(uint32_t) * L));
        FT_TRY(ensure(c->leaf_b, sizeof(Bound) * L));
        FT_TRY(ensure(c->leaf_parent, sizeof(int32_t) * L));
        FT_TRY(ensure(c->bnodes, sizeof(BNode) * std::max<int64_t>(1, L - 1)));
        FT_TRY(ensure(c->bflags, sizeof(unsigned) * std::max<int64_t>(1, L - 1)));
        hipLaunchKernelGGL(k_leaves, dim3(grid1d(G)), dim3(kBS), 0, s, cnt, pre, G,
                           c->leaf_key.as<uint32_t>(), c->leaf_start.as<uint32_t>(),
                           c->leaf_cnt.as<uint32_t>());
        hipLaunchKernelGGL(k_prim_scatter, dim3(grid1d(n)), dim3(kBS), 0, s, mode, n, ifp, tris,
                           c->pcode.as<uint32_t>(), cnt, pre, c->prec.as<float4>(),
                           c->trec.as<TriRec>(), c->pidx.as<int32_t>());
        hipLaunchKernelGGL(k_leaf_bounds, dim3(grid1d(L)), dim3(kBS), 0, s, mode, L,
                           c->leaf_start.as<uint32_t>(), c->leaf_cnt.as<uint32_t>(),
                           c->pidx.as<int32_t>(), ifp, tris, c->leaf_b.as<Bound>());
        if (L > 1) {
            FT_TRY(fill<unsigned>(s, c->bflags.as<unsigned>(), L - 1, 0u));
            hipLaunchKernelGGL(k_karras, dim3(grid1d(L - 1)), dim3(kBS), 0, s,
                               c->leaf_key.as<uint32_t>(), (int)L, c->bnodes.as<BNode>(),
                               c->leaf_parent.as<int32_t>());
            hipLaunchKernelGGL(k_refit, dim3(grid1d(L)), dim3(kBS), 0, s, (int)L,
                               c->leaf_parent.as<int32_t>(), c->leaf_b.as<Bound>(),
                               c->bnodes.as<BNode>(),
                               c->bflags.as<unsigned>());
        }
        FT_TRY(check_launch());
    }
    FT_TRY(fill<uint64_t>(s, sc + S_CLAMP, 2, 0ull));
    QueryArgs q{};
    q.n = c->N;
    q.sx = c->sx.as<float>(); q.sy = c->sy.as<float>(); q.sz = c->sz.as<float>();
    q.perm = c->perm.as<int32_t>();
    q.prec = c->prec.as<float4>();
    q.trec = c->trec.as<TriRec>();
    q.pidx = c->pidx.as<int32_t>();
    q.leaf_start = c->leaf_start.as<uint32_t>();
    q.leaf_cnt = c->leaf_cnt.as<uint32_t>();
    q.leaf_b = c->leaf_b.as<Bound>();
    q.nodes = c->bnodes.as<BNode>();
    q.L = (int)L;
    q.band = band;
    q.ifp = ifp;
    q.tris = tris;
    q.phi_old = phi_old;
    q.phi_out = phi_out;
    q.phi_ifn = phi_ifn;
    q.retain = retain;
    if (want_nearest) {
        FT_TRY(ensure(c->nearest, sizeof(int32_t) * c->N));
        q.nearest = c->nearest.as<int32_t>();
    }
    q.epi = epi;
    q.w_phi = w_phi;
    q.w_ifn = w_ifn;
    q.n_clamped = (unsigned long long*)(sc + S_CLAMP);
    static const bool want_stats = getenv("FT_QUERY_STATS") != nullptr;
    q.h_est = c->h_est;
    if (want_stats) {
        FT_TRY(ensure(c->stats, 8 * 32));
        FT_TRY(fill<uint64_t>(s, c->stats.as<uint64_t>(), 32, 0ull));
        q.stats = c->stats.as<unsigned long long>();
    }
    q.overflow = (unsigned*)(sc + S_OVF);
    const int64_t ntiles = (c->N + 63) / 64;
    const bool seeded = L > 1 && c->N >= 4096;
    static const int kind = getenv("FT_QUERY_KERNEL") && !strcmp(getenv("FT_QUERY_KERNEL"), "lane") ? 0 : 1;
    static const bool use_seed = getenv("FT_SEED") && !strcmp(getenv("FT_SEED"), "1");
    q.probe = !(getenv("FT_PROBE") && !strcmp(getenv("FT_PROBE"), "0"));
    mark(c, 2, s);
    if (seeded && use_seed) {
        FT_TRY(ensure(c->seeds, sizeof(int32_t) * ntiles));
        QueryArgs qr = q;
        qr.rep = 1;
        qr.seed_out = c->seeds.as<int32_t>();
        qr.stats = nullptr;
        if (mode == 0)
            hipLaunchKernelGGL(k_query<0>, dim3(grid1d(ntiles, kQueryBS)), dim3(kQueryBS), 0, s, qr);
        else
            hipLaunchKernelGGL(k_query<1>, dim3(grid1d(ntiles, kQueryBS)), dim3(kQueryBS), 0, s, qr);
        q.seed_in = c->seeds.as<int32_t>();
    }
    if (c->N > 0) {
        if (kind == 1) {
            if (mode == 0)
                hipLaunchKernelGGL(k_query_wave<0>, dim3(grid1d(c->N, kQueryBS)), dim3(kQueryBS), 0, s, q);
            else
                hipLaunchKernelGGL(k_query_wave<1>, dim3(grid1d(c->N, kQueryBS)), dim3(kQueryBS), 0, s, q);
        } else {
            if (mode == 0)
                hipLaunchKernelGGL(k_query<0>, dim3(grid1d(c->N, kQueryBS)), dim3(kQueryBS), 0, s, q);
            else
                hipLaunchKernelGGL(k_query<1>, dim3(grid1d(c->N, kQueryBS)), dim3(kQueryBS), 0, s, q);
        }
        FT_TRY(check_launch());
    }
    mark(c, 3, s);
    FT_TRY(read_scalars(c, s, S_N));
    if (want_stats && c->N > 0) {
        uint64_t st[32];
        FT_HIP(hipMemcpy(st, c->stats.p, sizeof st, hipMemcpyDeviceToHost));
        static const char* names[8] = {"out <2h", "out 2-8h", "out 8-32h", "out >32h",
                                       "in <2h", "in 2-8h", "in 8-32h", "in >32h"};
        double tn = 0, tl = 0, tp = 0;
        for (int b = 0; b < 8; b++) { tn += st[4 * b + 1]; tl += st[4 * b + 2]; tp += st[4 * b + 3]; }
        fprintf(stderr, "[ftreinit] query mode %d: N=%lld L=%lld per node: tree nodes %.1f, "
                "leaves %.1f, primitives %.1f\n", mode, (long long)c->N, (long long)L,
                tn / c->N, tl / c->N, tp / c->N);
        for (int b = 0; b < 8; b++) {
            const double n = (double)st[4 * b];
            if (n > 0)
                fprintf(stderr, "[ftreinit]   %-9s %5.1f%% of nodes: tree %.1f leaves %.1f prims %.1f\n",
                        names[b], 100.0 * n / c->N, st[4 * b + 1] / n, st[4 * b + 2] / n,
                        st[4 * b + 3] / n);
        }
    }
    if (c->h_scalars[S_OVF])
        return ft::set_error(FT_ERR_HIP, "search: traversal exceeded its bounds (%llu)",
                             (unsigned long long)c->h_scalars[S_OVF]);
    *n_clamped = (int64_t)c->h_scalars[S_CLAMP];
    c->have_nearest = want_nearest;
    return FT_OK;
}

int check_handle(ft_handle h)
{
    if (!h) return ft::set_error(FT_ERR_ARG, "null handle");
    FT_HIP(hipSetDevice(h->device));
    return FT_OK;
}

int upload_nodes(ft_ctx* c, const float* xyz, int64_t n)
{
    c->N = n;
    FT_TRY(ensure(c->xyz, sizeof(float) * 3 * n));
    FT_HIP(hipMemcpyAsync(c->xyz.p, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
    // node box and spacing estimate
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], xyz[3 * i + k]);
            hi[k] = std::max(hi[k], xyz[3 * i + k]);
        }
    if (n == 0) for (int k = 0; k < 3; k++) { lo[k] = 0; hi[k] = 1; }
    double vol = 1.0, ext_max = 0.0;
    for (int k = 0; k < 3; k++) ext_max = std::max(ext_max, (double)hi[k] - lo[k]);
    if (!(ext_max > 0.0)) ext_max = 1.0;
    for (int k = 0; k < 3; k++) vol *= std::max((double)hi[k] - lo[k], ext_max * 1e-3);
    for (int k = 0; k < 3; k++) { c->lo[k] = lo[k]; c->hi[k] = hi[k]; }
    c->h_est = (float)std::cbrt(vol / (double)std::max<int64_t>(1, n));
    // Morton order of the nodes
    FT_TRY(ensure(c->perm, sizeof(int32_t) * n));
    FT_TRY(ensure(c->sx, sizeof(float) * n));
    FT_TRY(ensure(c->sy, sizeof(float) * n));
    FT_TRY(ensure(c->sz, sizeof(float) * n));
    if (n == 0) return FT_OK;
    int64_t G = 0;
    const Grid g = make_grid(c, c->h_est, &G);
    FT_TRY(ensure(c->bin_count, sizeof(uint32_t) * G));
    FT_TRY(ensure(c->bin_prefix, sizeof(uint64_t) * G));
    FT_TRY(ensure(c->pcode, sizeof(uint32_t) * n));
    uint32_t* cnt = c->bin_count.as<uint32_t>();
    uint32_t* off = reinterpret_cast<uint32_t*>(c->bin_prefix.p);
    FT_TRY(fill<uint32_t>(c->stream, cnt, G, 0u));
    hipLaunchKernelGGL(k_node_bin, dim3(grid1d(n)), dim3(kBS), 0, c->stream, c->xyz.as<float>(), n, g,
                       c->pcode.as<uint32_t>(), cnt);
    FT_TRY(scan<uint32_t, uint32_t>(c, c->stream, cnt, off, G,
                                    reinterpret_cast<uint32_t*>(c->scalars.as<uint64_t>() + S_BINS),
                                    LoadSame{}));
    hipLaunchKernelGGL(k_node_scatter, dim3(grid1d(n)), dim3(kBS), 0, c->stream, c->xyz.as<float>(), n,
                       c->pcode.as<uint32_t>(), cnt, off, c->perm.as<int32_t>(), c->sx.as<float>(),
                       c->sy.as<float>(), c->sz.as<float>());
    FT_TRY(check_launch());
    FT_HIP(hipStreamSynchronize(c->stream));
    return FT_OK;
}

// Marching tets on device: fills ifp, ifp2e, tris, if_elems, ifp_area, p2t
// and the retained-node bitmap.  Returns counts in `cnt`.
int marching_tets(ft_ctx* c, hipStream_t s, const double* phi, ft_counts* cnt)
{
    const int64_t N = c->N, T = c->T, E = c->E;
    uint64_t* sc = c->scalars.as<uint64_t>();
    const int64_t nw_nodes = (N + 63) / 64, nw_tets = (T + 63) / 64, nw_edges = (E + 63) / 64;
    const int64_t nblk = (T + kBS - 1) / kBS;
    FT_TRY(ensure(c->sbits, 8 * nw_nodes));
    FT_TRY(ensure(c->retain, 8 * nw_nodes));
    FT_TRY(ensure(c->tbits, 8 * nw_tets));
    FT_TRY(ensure(c->blk, 8 * std::max<int64_t>(1, nblk)));
    FT_TRY(ensure(c->ebits, 8 * nw_edges));
    FT_TRY(ensure(c->eprefix, 4 * nw_edges));
    uint64_t* sbits = c->sbits.as<uint64_t>();
    mark(c, 0, s);
    hipLaunchKernelGGL(k_node_signs, dim3(grid1d(N)), dim3(kBS), 0, s, phi, N, sbits);
    FT_TRY(fill<uint64_t>(s, c->retain.as<uint64_t>(), nw_nodes, 0ull));
    FT_TRY(fill<uint64_t>(s, c->ebits.as<uint64_t>(), nw_edges, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_ERR, 1, ~0ull));
    if (T > 0)
        hipLaunchKernelGGL(k_classify, dim3((unsigned)nblk), dim3(kBS), 0, s, c->tets.as<int4>(), T,
                           sbits, c->tbits.as<uint64_t>(), c->blk.as<uint64_t>());
    FT_TRY(scan<uint64_t, uint64_t>(c, s, c->blk.as<uint64_t>(), c->blk.as<uint64_t>(), nblk,
                                    sc + S_BLK, LoadSame{}));
    FT_TRY(read_scalars(c, s, S_N));
    const int64_t X = (int64_t)(c->h_scalars[S_BLK] >> 32);
    const int64_t F = (int64_t)(c->h_scalars[S_BLK] & 0xffffffffull);
    FT_TRY(ensure(c->xtet, 4 * std::max<int64_t>(1, X)));
    FT_TRY(ensure(c->xtri_off, 4 * std::max<int64_t>(1, X)));
    FT_TRY(ensure(c->xedges, 16 * std::max<int64_t>(1, X)));
    FT_TRY(ensure(c->if_elems, 4 * std::max<int64_t>(1, X)));
    FT_TRY(ensure(c->tris, 16 * std::max<int64_t>(1, F)));
    FT_TRY(ensure(c->tri_norm, 4 * std::max<int64_t>(1, F)));
    if (nw_tets > 0)
        hipLaunchKernelGGL(k_compact, dim3(grid1d(nw_tets)), dim3(kBS), 0, s, c->tets.as<int4>(), sbits,
                           c->tbits.as<uint64_t>(), nw_tets, c->blk.as<uint64_t>(),
                           c->xtet.as<int32_t>(), c->xtri_off.as<int32_t>());
    if (X > 0)
        hipLaunchKernelGGL(k_xtet, dim3(grid1d(X)), dim3(kBS), 0, s, c->tets.as<int4>(),
                           c->t2e.as<int2>(), c->edges.as<int2>(), sbits, c->xtet.as<int32_t>(), X,
                           c->ebits.as<unsigned long long>(), c->xedges.as<int4>(),
                           (unsigned long long*)(sc + S_ERR));
    FT_TRY(fill<uint64_t>(s, sc + S_EDGES, 1, 0ull));   // the scan writes the low 32 bits
    FT_TRY(scan<uint64_t, uint32_t>(c, s, c->ebits.as<uint64_t>(), c->eprefix.as<uint32_t>(),
                                    nw_edges, reinterpret_cast<uint32_t*>(sc + S_EDGES), LoadPopc{}));
    FT_TRY(read_scalars(c, s, S_N));
    const int64_t P = (int64_t)(c->h_scalars[S_EDGES] & 0xffffffffull);
    cnt->n_ifp = P;
    cnt->n_tri = F;
    cnt->n_if_elems = X;
    cnt->bad_tet = -1;
    if (c->h_scalars[S_ERR] != ~0ull) {
        const uint64_t ev = c->h_scalars[S_ERR];
        cnt->bad_tet = (int64_t)(ev & ((1ull << 40) - 1));
        return ft::set_error(FT_ERR_T2E,
                             "tet %lld: crossing edges from t2e do not match its node signs "
                             "(LevelSet.cpp:400-405)", (long long)cnt->bad_tet);
    }
    FT_TRY(ensure(c->ifp, 12 * std::max<int64_t>(1, P)));
    FT_TRY(ensure(c->ifp2e, 4 * std::max<int64_t>(1, P)));
    FT_TRY(ensure(c->ifp_area, 4 * std::max<int64_t>(1, P)));
    FT_TRY(ensure(c->p2t_off, 4 * (P + 1)));
    FT_TRY(ensure(c->p2t_cursor, 4 * std::max<int64_t>(1, P)));
    FT_TRY(ensure(c->p2t, 4 * std::max<int64_t>(1, 3 * F)));
    if (nw_edges > 0)
        hipLaunchKernelGGL(k_points, dim3(grid1d(nw_edges)), dim3(kBS), 0, s, c->ebits.as<uint64_t>(),
                           c->eprefix.as<uint32_t>(), nw_edges, c->edges.as<int2>(), c->xyz.as<float>(),
                           phi, c->ifp.as<float>(), c->ifp2e.as<int32_t>(),
                           c->retain.as<unsigned long long>());
    if (X > 0)
        hipLaunchKernelGGL(k_tris, dim3(grid1d(X)), dim3(kBS), 0, s, c->tets.as<int4>(), c->xyz.as<float>(),
                           phi, c->xtet.as<int32_t>(), c->xtri_off.as<int32_t>(), c->xedges.as<int4>(), X,
                           c->ebits.as<uint64_t>(), c->eprefix.as<uint32_t>(), c->ifp.as<float>(),
                           c->tris.as<int4>(), c->tri_norm.as<float>(), c->if_elems.as<int32_t>(),
                           (unsigned long long*)(sc + S_ERR));
    // p2t and areas
    uint32_t* deg = c->p2t_cursor.as<uint32_t>();
    FT_TRY(fill<uint32_t>(s, deg, P, 0u));
    if (F > 0)
        hipLaunchKernelGGL(k_p2t_count, dim3(grid1d(F)), dim3(kBS), 0, s, c->tris.as<int4>(), F, deg);
    FT_TRY(scan<uint32_t, uint32_t>(c, s, deg, c->p2t_off.as<uint32_t>(), P,
                                    c->p2t_off.as<uint32_t>() + P, LoadSame{}));
    FT_TRY(fill<uint32_t>(s, deg, P, 0u));
    if (F > 0)
        hipLaunchKernelGGL(k_p2t_fill, dim3(grid1d(F)), dim3(kBS), 0, s, c->tris.as<int4>(), F,
                           c->p2t_off.as<uint32_t>(), deg, c->p2t.as<int32_t>());
    if (P > 0)
        hipLaunchKernelGGL(k_point_area, dim3(grid1d(P)), dim3(kBS), 0, s, c->p2t_off.as<uint32_t>(), P,
                           c->p2t.as<int32_t>(), c->tri_norm.as<float>(), c->ifp_area.as<float>());
    FT_TRY(check_launch());
    mark(c, 1, s);
    FT_TRY(read_scalars(c, s, S_N));
    if (c->h_scalars[S_ERR] != ~0ull) {
        const uint64_t ev = c->h_scalars[S_ERR];
        cnt->bad_tet = (int64_t)(ev & ((1ull << 40) - 1));
        return ft::set_error(FT_ERR_DEGENERATE,
                             "tet %lld crosses the interface but all its nodes have |phi| <= "
                             "FLT_MIN (LevelSet.cpp:338-341)", (long long)cnt->bad_tet);
    }
    c->have_interface = true;
    return FT_OK;
}

int validate_opts(const ft_opts* o)
{
    if (!o) return ft::set_error(FT_ERR_ARG, "null options");
    if (o->mode != FT_DIST_POINTS && o->mode != FT_DIST_TRIANGLES)
        return ft::set_error(FT_ERR_ARG, "unknown distance mode %d", o->mode);
    if (o->blend != FT_BLEND_INIT && o->blend != FT_BLEND_UPDATE)
        return ft::set_error(FT_ERR_ARG, "unknown blend %d", o->blend);
    if (!(o->band >= 0.0) || o->band > 1e30)
        return ft::set_error(FT_ERR_ARG, "band must be >= 0");
    if (!(o->bin_scale >= 0.0f) || o->bin_scale > 1e6f)
        return ft::set_error(FT_ERR_ARG, "bin_scale must be >= 0");
    return FT_OK;
}

int reinit_device(ft_ctx* c, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                  ft_counts* counts, hipStream_t s)
{
    FT_TRY(validate_opts(o));
    if (!c->has_conn) return ft::set_error(FT_ERR_STATE, "reinit before ft_set_mesh");
    if (!d_phi) return ft::set_error(FT_ERR_ARG, "null phi");
    c->have_interface = false;
    ft_counts cnt{};
    int rc = marching_tets(c, s, d_phi, &cnt);
    if (counts) *counts = cnt;
    c->last = cnt;
    if (rc) return rc;
    // the query reads phi_old and writes phi_out in place per node
    const int epi = (o->blend == FT_BLEND_INIT) ? EPI_INIT : EPI_UPDATE;
    int64_t clamped = 0;
    if (o->mode == FT_DIST_POINTS)
        rc = search(c, s, 0, cnt.n_ifp, c->ifp.as<float>(), nullptr, o->bin_scale, (float)o->band,
                    d_phi, d_phi, d_phi_ifn, c->retain.as<uint64_t>(), epi, o->w_phi, o->w_ifn,
                    o->want_nearest != 0, &clamped);
    else
        rc = search(c, s, 1, cnt.n_tri, c->ifp.as<float>(), c->tris.as<int4>(), o->bin_scale,
                    (float)o->band, d_phi, d_phi, d_phi_ifn, c->retain.as<uint64_t>(), epi, o->w_phi,
                    o->w_ifn, o->want_nearest != 0, &clamped);
    cnt.n_clamped = clamped;
    if (counts) *counts = cnt;
    c->last = cnt;
    if (rc == FT_OK && c->timing) {
        float ms[3] = {0, 0, 0};
        (void)hipEventSynchronize(c->ev[3]);
        for (int k = 0; k < 3; k++) (void)hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]);
        for (int k = 0; k < 3; k++) c->phase_ms[k] += ms[k];
        c->timed_calls++;
    }
    return rc;
}

std::mutex g_create_mu;

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void ft_opts_default(ft_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->mode = FT_DIST_POINTS;
    o->blend = FT_BLEND_INIT;
    o->w_phi = 0.75;
    o->w_ifn = 0.25;
    o->band = 0.0;
    o->bin_scale = 0.0f;
    o->want_nearest = 0;
}

const char* ft_last_error(void) { return ft::error_slot().c_str(); }

int ft_device_count(int32_t* n)
{
    if (!n) return ft::set_error(FT_ERR_ARG, "null output");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = (e == hipSuccess) ? c : 0;
    return FT_OK;
}

int ft_create(int32_t device, ft_handle* out)
{
    if (!out) return ft::set_error(FT_ERR_ARG, "null output");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(g_create_mu);
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
        return ft::set_error(FT_ERR_NODEVICE, "no HIP device");
    if (device < 0 || device >= nd) return ft::set_error(FT_ERR_ARG, "device %d out of range", device);
    FT_HIP(hipSetDevice(device));
    ft_ctx* c = new ft_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ft::set_error(FT_ERR_HIP, "hipStreamCreate failed");
    }
    if (ensure(c->scalars, 8 * S_N) != FT_OK ||
        hipHostMalloc((void**)&c->h_scalars, 8 * S_N, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        release(c->scalars);
        delete c;
        return ft::set_error(FT_ERR_NOMEM, "scalar buffers");
    }
    *out = c;
    return FT_OK;
}

int ft_destroy(ft_handle h)
{
    if (!h) return FT_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    DevBuf* bufs[] = {&h->xyz, &h->tets, &h->edges, &h->t2e, &h->perm, &h->sx, &h->sy, &h->sz,
                      &h->phi_tmp, &h->phi_ifn_tmp, &h->sbits, &h->retain, &h->tbits, &h->blk,
                      &h->xtet, &h->xtri_off, &h->xedges, &h->ebits, &h->eprefix, &h->ifp,
                      &h->ifp2e, &h->tris, &h->tri_norm, &h->if_elems, &h->p2t_off,
                      &h->p2t_cursor, &h->p2t, &h->ifp_area, &h->nearest, &h->bin_count,
                      &h->bin_prefix, &h->pcode, &h->prec, &h->trec, &h->pidx, &h->leaf_key,
                      &h->leaf_start, &h->leaf_cnt, &h->leaf_b, &h->leaf_parent,
                      &h->bnodes, &h->bflags, &h->seeds, &h->stats, &h->scan_bsum, &h->scalars};
    for (DevBuf* b : bufs) release(*b);
    if (h->h_scalars) (void)hipHostFree(h->h_scalars);
    for (int k = 0; k < 4; k++)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return FT_OK;
}

int ft_set_nodes(ft_handle h, const float* xyz, int64_t n_nodes)
{
    FT_TRY(check_handle(h));
    if (n_nodes < 0 || n_nodes > INT32_MAX || (n_nodes > 0 && !xyz))
        return ft::set_error(FT_ERR_ARG, "ft_set_nodes: bad arguments");
    h->has_conn = false;
    h->have_interface = false;
    h->T = h->E = 0;
    return upload_nodes(h, xyz, n_nodes);
}

int ft_set_mesh(ft_handle h, const float* xyz, int64_t n_nodes, const int32_t* tets,
                int64_t n_tets, const int32_t* edges, int64_t n_edges, const int32_t* t2e)
{
    FT_TRY(check_handle(h));
    if (n_nodes <= 0 || n_nodes > INT32_MAX || n_tets < 0 || n_tets > INT32_MAX || n_edges < 0 ||
        n_edges > INT32_MAX || !xyz || (n_tets > 0 && (!tets || !t2e)) || (n_edges > 0 && !edges))
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh: bad arguments");
    // Index ranges are validated once here so that no kernel reads out of bounds.
    for (int64_t i = 0; i < 4 * n_tets; i++)
        if (tets[i] < 0 || tets[i] >= n_nodes)
            return ft::set_error(FT_ERR_ARG, "ft_set_mesh: tet %lld has node %d out of range",
                                 (long long)(i / 4), tets[i]);
    for (int64_t i = 0; i < 2 * n_edges; i++)
        if (edges[i] < 0 || edges[i] >= n_nodes)
            return ft::set_error(FT_ERR_ARG, "ft_set_mesh: edge %lld has node %d out of range",
                                 (long long)(i / 2), edges[i]);
    for (int64_t i = 0; i < 6 * n_tets; i++)
        if (t2e[i] < 0 || t2e[i] >= n_edges)
            return ft::set_error(FT_ERR_ARG, "ft_set_mesh: t2e of tet %lld has edge %d out of range",
                                 (long long)(i / 6), t2e[i]);
    h->has_conn = false;
    h->have_interface = false;
    FT_TRY(upload_nodes(h, xyz, n_nodes));
    h->T = n_tets;
    h->E = n_edges;
    FT_TRY(ensure(h->tets, 16 * std::max<int64_t>(1, n_tets)));
    FT_TRY(ensure(h->edges, 8 * std::max<int64_t>(1, n_edges)));
    FT_TRY(ensure(h->t2e, 24 * std::max<int64_t>(1, n_tets)));
    if (n_tets > 0) {
        FT_HIP(hipMemcpyAsync(h->tets.p, tets, 16 * n_tets, hipMemcpyHostToDevice, h->stream));
        FT_HIP(hipMemcpyAsync(h->t2e.p, t2e, 24 * n_tets, hipMemcpyHostToDevice, h->stream));
    }
    if (n_edges > 0)
        FT_HIP(hipMemcpyAsync(h->edges.p, edges, 8 * n_edges, hipMemcpyHostToDevice, h->stream));
    FT_HIP(hipStreamSynchronize(h->stream));
    h->has_conn = true;
    return FT_OK;
}

int ft_reset_phi_points(ft_handle h, const float* ifp, int64_t n_ifp, double* phi_n)
{
    FT_TRY(check_handle(h));
    if (n_ifp < 0 || (n_ifp > 0 && !ifp) || (h->N > 0 && !phi_n))
        return ft::set_error(FT_ERR_ARG, "ft_reset_phi_points: bad arguments");
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->phi_tmp, 8 * std::max<int64_t>(1, h->N)));
    FT_TRY(ensure(h->ifp, 12 * std::max<int64_t>(1, n_ifp)));
    if (n_ifp > 0) FT_HIP(hipMemcpyAsync(h->ifp.p, ifp, 12 * n_ifp, hipMemcpyHostToDevice, s));
    if (h->N > 0) FT_HIP(hipMemcpyAsync(h->phi_tmp.p, phi_n, 8 * h->N, hipMemcpyHostToDevice, s));
    int64_t clamped = 0;
    h->have_interface = false;
    FT_TRY(search(h, s, 0, n_ifp, h->ifp.as<float>(), nullptr, 0.0f, 0.0f, h->phi_tmp.as<double>(),
                  h->phi_tmp.as<double>(), nullptr, nullptr, EPI_RAW, 0.0, 0.0, true, &clamped));
    if (h->N > 0) FT_HIP(hipMemcpyAsync(phi_n, h->phi_tmp.p, 8 * h->N, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_reinit_device(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                     ft_counts* counts, void* stream)
{
    FT_TRY(check_handle(h));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return reinit_device(h, d_phi, d_phi_ifn, o, counts, s);
}

int ft_reinit(ft_handle h, double* phi, double* phi_ifn, const ft_opts* o, ft_counts* counts)
{
    FT_TRY(check_handle(h));
    if (!phi) return ft::set_error(FT_ERR_ARG, "null phi");
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->phi_tmp, 8 * std::max<int64_t>(1, h->N)));
    if (phi_ifn) FT_TRY(ensure(h->phi_ifn_tmp, 8 * std::max<int64_t>(1, h->N)));
    FT_HIP(hipMemcpyAsync(h->phi_tmp.p, phi, 8 * h->N, hipMemcpyHostToDevice, s));
    FT_TRY(reinit_device(h, h->phi_tmp.as<double>(), phi_ifn ? h->phi_ifn_tmp.as<double>() : nullptr,
                         o, counts, s));
    FT_HIP(hipMemcpyAsync(phi, h->phi_tmp.p, 8 * h->N, hipMemcpyDeviceToHost, s));
    if (phi_ifn) FT_HIP(hipMemcpyAsync(phi_ifn, h->phi_ifn_tmp.p, 8 * h->N, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_get_interface(ft_handle h, float* ifp, int32_t* ifp2e, int32_t* tris, int32_t* if_elems,
                     float* ifp_area, int32_t* p2t_offsets, int32_t* p2t_tris)
{
    FT_TRY(check_handle(h));
    if (!h->have_interface) return ft::set_error(FT_ERR_STATE, "no interface from a reinit");
    const int64_t P = h->last.n_ifp, F = h->last.n_tri, X = h->last.n_if_elems;
    hipStream_t s = h->stream;
    if (ifp && P) FT_HIP(hipMemcpyAsync(ifp, h->ifp.p, 12 * P, hipMemcpyDeviceToHost, s));
    if (ifp2e && P) FT_HIP(hipMemcpyAsync(ifp2e, h->ifp2e.p, 4 * P, hipMemcpyDeviceToHost, s));
    if (tris && F) FT_HIP(hipMemcpyAsync(tris, h->tris.p, 16 * F, hipMemcpyDeviceToHost, s));
    if (if_elems && X) FT_HIP(hipMemcpyAsync(if_elems, h->if_elems.p, 4 * X, hipMemcpyDeviceToHost, s));
    if (ifp_area && P) FT_HIP(hipMemcpyAsync(ifp_area, h->ifp_area.p, 4 * P, hipMemcpyDeviceToHost, s));
    if (p2t_offsets) FT_HIP(hipMemcpyAsync(p2t_offsets, h->p2t_off.p, 4 * (P + 1), hipMemcpyDeviceToHost, s));
    if (p2t_tris && F) FT_HIP(hipMemcpyAsync(p2t_tris, h->p2t.p, 12 * F, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_get_nearest(ft_handle h, int32_t* nearest)
{
    FT_TRY(check_handle(h));
    if (!nearest) return ft::set_error(FT_ERR_ARG, "null output");
    if (!h->have_nearest) return ft::set_error(FT_ERR_STATE, "last call kept no nearest indices");
    FT_HIP(hipMemcpyAsync(nearest, h->nearest.p, 4 * h->N, hipMemcpyDeviceToHost, h->stream));
    FT_HIP(hipStreamSynchronize(h->stream));
    return FT_OK;
}

int ft_set_timing(ft_handle h, int32_t on)
{
    FT_TRY(check_handle(h));
    if (on && !h->ev[0])
        for (int k = 0; k < 4; k++) FT_HIP(hipEventCreate(&h->ev[k]));
    h->timing = on != 0;
    for (int k = 0; k < 3; k++) h->phase_ms[k] = 0.0;
    h->timed_calls = 0;
    return FT_OK;
}

int ft_get_timing(ft_handle h, double* phase_ms, int64_t* calls)
{
    FT_TRY(check_handle(h));
    if (!phase_ms || !calls) return ft::set_error(FT_ERR_ARG, "null output");
    for (int k = 0; k < 3; k++) phase_ms[k] = h->phase_ms[k];
    *calls = h->timed_calls;
    return FT_OK;
}

int ft_interface_device(ft_handle h, const float** ifp, const int32_t** tris, const int32_t** ifp2e)
{
    FT_TRY(check_handle(h));
    if (!h->have_interface) return ft::set_error(FT_ERR_STATE, "no interface from a reinit");
    if (ifp) *ifp = h->ifp.as<float>();
    if (tris) *tris = h->tris.as<int32_t>();
    if (ifp2e) *ifp2e = h->ifp2e.as<int32_t>();
    return FT_OK;
}

}  // extern "C"
