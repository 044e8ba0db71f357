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
//  11 schedule        64-node tiles deep inside the surface listed first
//                     (k_tile_deep, scan, k_tile_list; full domain only)
//  12 k_query         per node (Morton-sorted order): exact nearest primitive
//                     (triangles: FP32 search, FP64 settle of the best two),
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
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ftreinit.h"
#include "ft_device.hpp"
#include "ft_internal.hpp"
#include "ft_p2t.hpp"

using ftd::Grid;
using ftd::TriRec;
using ftd::Bound;

namespace {

constexpr int kBS = 256;                 // threads per block of the streaming kernels
constexpr int kScanTile = 2048;          // elements per scan block (256 x 8)
constexpr int kStack = 24;               // traversal stack depth (tree depth <= 24)
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

// Scalar slots of a handle (a small device u64 array): stage totals written
// by the scans, the error word, capacity and traversal overflow flags.
//   S_BLK    (crossing tets << 32) | triangles
//   S_EDGES  crossing points (low 32 bits)
//   S_ERR    ~0, or (code << 40) | first failing tet
//   S_BINS   (leaves << 32) | primitives
//   S_DEEP   query tiles scheduled first (low 32 bits), of which the very deep
//            ones, split over several waves (high 32 bits)
//   S_OVF    traversal stack/budget overflow of the query
//   S_CAP    a stage total exceeded the capacity its buffers were sized for
//   S_Q      the query's work-queue counters (9 x u32: deep tiles, then one
//            per XCD), zeroed per launch
//   S_MARGIN the bins schedule: largest distance from a triangle's centroid to
//            its vertices (float bits, low 32)
enum Slot { S_BLK = 0, S_EDGES = 1, S_ERR = 2, S_P2T = 3, S_BINS = 4, S_DEEP = 5, S_OVF = 6, S_CAP = 7,
            S_Q = 8, S_MARGIN = 13, S_N = 14 };

// Stage totals of the running reinit as the kernels see them.  Kernels after
// a scan read the total from the device, so a launch sized for a capacity
// (a replayed graph) processes exactly this call's elements.
__device__ __forceinline__ int64_t cnt_xtets(const uint64_t* sc) { return (int64_t)(sc[S_BLK] >> 32); }
__device__ __forceinline__ int64_t cnt_tris(const uint64_t* sc) { return (int64_t)(sc[S_BLK] & 0xffffffffull); }
__device__ __forceinline__ int64_t cnt_points(const uint64_t* sc) { return (int64_t)(sc[S_EDGES] & 0xffffffffull); }
__device__ __forceinline__ int64_t cnt_leaves(const uint64_t* sc) { return (int64_t)(sc[S_BINS] >> 32); }
// Set once a stage failed (t2e or degenerate tet) or outgrew its buffers:
// every later kernel of the reinit then does nothing and phi stays as it was.
__device__ __forceinline__ bool halted(const uint64_t* sc)
{
    return sc[S_CAP] != 0ull || sc[S_ERR] != ~0ull;
}

// Connectivity of the resident fields as the marching-tets kernels see it:
// one concatenated array per kind (global indices), or, for a batch of
// resident meshes, each field's own arrays read in place (field-local
// indices, to which the field's node or edge offset is added), so that a
// batch is assembled without copying its tets, t2e and edges.
struct MeshView {
    const int4* tets;                // concatenated (null: per field)
    const int2* t2e3;                // t2e rows as 3 int2
    const int2* edges;
    const int4* const* f_tets;       // per field
    const int2* const* f_t2e3;
    const int2* const* f_edges;
    const int64_t* node_off;         // device offsets, nf + 1 each
    const int64_t* tet_off;
    const int64_t* edge_off;
    int nf;
};

// Field of tet t (t < T) and of edge e: binary search over the offsets.
__device__ __forceinline__ int mv_tet_field(const MeshView& m, int64_t t)
{
    return ftd::field_of(m.tet_off, m.nf, t);
}
__device__ __forceinline__ int mv_edge_field(const MeshView& m, int64_t e)
{
    return ftd::field_of(m.edge_off, m.nf, e);
}
// Tet t of field k (k = its field) with global node ids.
__device__ __forceinline__ int4 mv_tet(const MeshView& m, int64_t t, int k)
{
    if (m.tets) return m.tets[t];
    int4 q = m.f_tets[k][t - m.tet_off[k]];
    const int32_t a = (int32_t)m.node_off[k];
    q.x += a; q.y += a; q.z += a; q.w += a;
    return q;
}
// Row j (0..2) of tet t's t2e, as two global edge ids.
__device__ __forceinline__ int2 mv_t2e(const MeshView& m, int64_t t, int j, int k)
{
    if (m.t2e3) return m.t2e3[3 * t + j];
    int2 r = m.f_t2e3[k][3 * (t - m.tet_off[k]) + j];
    const int32_t a = (int32_t)m.edge_off[k];
    r.x += a; r.y += a;
    return r;
}
// Edge e of field k with global node ids.
__device__ __forceinline__ int2 mv_edge(const MeshView& m, int64_t e, int k)
{
    if (m.edges) return m.edges[e];
    int2 ab = m.f_edges[k][e - m.edge_off[k]];
    const int32_t a = (int32_t)m.node_off[k];
    ab.x += a; ab.y += a;
    return ab;
}

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
struct LoadTileClass {  // tile class 0 / 1 deep / 2 very deep -> (very deep << 32) | deep
    __device__ uint64_t operator()(uint32_t c) const { return ((uint64_t)(c == 2) << 32) | (uint64_t)(c >= 1); }
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

// K consecutive elements at a 16-byte aligned address, as 16-byte accesses.
template <class T, int K>
__device__ __forceinline__ void load_run16(const T* p, T (&v)[K])
{
    static_assert((sizeof(T) * K) % 16 == 0, "whole 16-byte words");
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) * K / 16); i++) {
        const uint4 u = q[i];
        memcpy(reinterpret_cast<char*>(v) + 16 * i, &u, 16);
    }
}

template <class T, int K>
__device__ __forceinline__ void store_run16(T* p, const T (&v)[K])
{
    static_assert((sizeof(T) * K) % 16 == 0, "whole 16-byte words");
    uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) * K / 16); i++) {
        uint4 u;
        memcpy(&u, reinterpret_cast<const char*>(v) + 16 * i, 16);
        q[i] = u;
    }
}

// Each thread scans IPT consecutive elements: read and written as 16-byte
// accesses when the run is whole and both arrays are 16-byte aligned (the
// run starts at a multiple of IPT elements, at least 32 bytes).
template <class TI, class TO, class Ld>
__global__ __launch_bounds__(kBS) void k_scan_down(const TI* in, TO* out, int64_t n,
                                                   const TO* bsum, Ld ld)
{
    __shared__ TO lds[kBS / 64 + 1];
    constexpr int IPT = kScanTile / kBS;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * IPT;
    const bool vec = base + IPT <= n &&
                     ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    TO v[IPT];
    TO s = 0;
    if (vec) {
        TI r[IPT];
        load_run16(in + base, r);
#pragma unroll
        for (int k = 0; k < IPT; k++) v[k] = (TO)ld(r[k]);
    } else {
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            const int64_t i = base + k;
            v[k] = i < n ? (TO)ld(in[i]) : TO(0);
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) s += v[k];
    TO tot;
    TO ex = ftd::block_excl_scan<TO, kBS>(s, lds, &tot) + bsum[blockIdx.x];
    if (vec) {
        TO o[IPT];
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            o[k] = ex;
            ex += v[k];
        }
        store_run16(out + base, o);
    } else {
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            const int64_t i = base + k;
            if (i < n) out[i] = ex;
            ex += v[k];
        }
    }
}

// ============================================================================
// Mesh set-up: Morton order of the nodes (processing order of k_query)
// ============================================================================
// Fields: a handle holds one or more independent fields (a plain reinit has
// one).  Field k owns nodes [node_off[k], node_off[k+1]) and likewise tets and
// edges; its index bins and query tiles never mix with another field's.
struct FieldGeom {
    const int64_t* node_off;         // nf + 1
    const int64_t* cnt_off;          // unused by the slot padding
    const Grid* grids;               // unused by the slot padding
    int nf;
};

// Node sort of one field (nodes node_base.. of xyz, local index i).
__global__ __launch_bounds__(kBS) void k_node_bin(const float* xyz, int64_t n, Grid g,
                                                  uint32_t* code, uint32_t* count)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = ftd::grid_code(g, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    code[i] = c;
    atomicAdd(&count[c], 1u);
}

// Sorted position p of local node i goes to query slot slot_base + p.
__global__ __launch_bounds__(kBS) void k_node_scatter(const float* xyz, int64_t n, int64_t node_base,
                                                      const uint32_t* code, uint32_t* count,
                                                      const uint32_t* offset, int32_t* perm,
                                                      float* sx, float* sy, float* sz)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = code[i];
    const int64_t slot = (int64_t)offset[c] + atomicSub(&count[c], 1u) - 1u;
    perm[slot] = (int32_t)(node_base + i);
    sx[slot] = xyz[3 * i];
    sy[slot] = xyz[3 * i + 1];
    sz[slot] = xyz[3 * i + 2];
}

// Pad slots of each field's last wave: perm -1, coordinates of the field's
// first slot (lanes that only shadow a real node); tile -> field table.
__global__ __launch_bounds__(kBS) void k_pad_slots(int64_t n_slots, FieldGeom fg,
                                                   const int64_t* slot_off, int32_t* perm,
                                                   float* sx, float* sy, float* sz,
                                                   int32_t* tile_field)
{
    const int64_t s = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (s >= n_slots) return;
    const int k = ftd::field_of(slot_off, fg.nf, s);
    if ((s & 63) == 0) tile_field[s >> 6] = k;
    const int64_t nk = fg.node_off[k + 1] - fg.node_off[k];
    if (s - slot_off[k] < nk) return;
    const int64_t f0 = slot_off[k];
    perm[s] = -1;
    sx[s] = sx[f0];
    sy[s] = sy[f0];
    sz[s] = sz[f0];
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

// Crossing iff 1..3 of a tet's nodes are negative (LevelSet.cpp:380-383); 2
// negative nodes -> 4 crossing edges -> 2 triangles.  A workgroup handles
// kClassifyGroups consecutive groups of kBS tets, one tet of each per
// thread, so that each thread has that many tet loads in flight; per group
// it writes the words of two bitmaps, crossing tets (tbits) and those with
// two triangles (t2bits, a subset), and blk[group] = (crossing tets << 32) |
// triangles.
constexpr int kClassifyGroups = 2;
__global__ __launch_bounds__(kBS) void k_classify(MeshView mv, int64_t T,
                                                  const uint64_t* sbits, uint64_t* tbits,
                                                  uint64_t* t2bits, uint64_t* blk)
{
    __shared__ uint64_t part[kClassifyGroups][kBS / 64];
    __shared__ int kf0;
    int4 q[kClassifyGroups];
    int64_t t[kClassifyGroups];
    // per-field arrays: the block's first field by binary search, then each
    // thread steps forward over the few field ends inside the block
    int k = 0;
    if (!mv.tets) {
        if (threadIdx.x == 0) kf0 = mv_tet_field(mv, std::min<int64_t>(T - 1, (int64_t)blockIdx.x * kClassifyGroups * kBS));
        __syncthreads();
        k = kf0;
    }
#pragma unroll
    for (int g = 0; g < kClassifyGroups; g++) {
        t[g] = ((int64_t)blockIdx.x * kClassifyGroups + g) * kBS + threadIdx.x;
        if (t[g] < T) {
            if (!mv.tets)
                while (t[g] >= mv.tet_off[k + 1]) k++;
            q[g] = mv_tet(mv, t[g], k);
        }
    }
#pragma unroll
    for (int g = 0; g < kClassifyGroups; g++) {
        int sgn = 0;
        if (t[g] < T)
            sgn = ftd::get_bit(sbits, q[g].x) + ftd::get_bit(sbits, q[g].y) + ftd::get_bit(sbits, q[g].z) +
                  ftd::get_bit(sbits, q[g].w);
        const bool cross = (t[g] < T) && sgn != 0 && sgn != 4;
        const uint64_t m = __ballot(cross);
        const uint64_t m2 = __ballot(cross && sgn == 2);
        if (ftd::lane_id() == 0 && t[g] < T) {
            tbits[t[g] >> 6] = m;
            t2bits[t[g] >> 6] = m2;
        }
        if (ftd::lane_id() == 0)
            part[g][threadIdx.x / 64] = ((uint64_t)__popcll(m) << 32) | (uint64_t)(__popcll(m) + __popcll(m2));
    }
    __syncthreads();
    if (threadIdx.x < kClassifyGroups) {
        const int64_t grp = (int64_t)blockIdx.x * kClassifyGroups + threadIdx.x;
        if (grp * kBS < T) {
            uint64_t s4 = 0;
            for (int w = 0; w < kBS / 64; w++) s4 += part[threadIdx.x][w];
            blk[grp] = s4;
        }
    }
}

// One thread per 64-tet word of the crossing bitmap.  Lanes 4k..4k+3 share a
// k_classify block, whose exclusive offset is blk_prefix[w/4].
__global__ __launch_bounds__(kBS) void k_compact(const uint64_t* tbits, const uint64_t* t2bits,
                                                 int64_t nwords, const uint64_t* blk_prefix,
                                                 int32_t* xtet, int32_t* xtri_off, const uint64_t* sc)
{
    if (halted(sc)) return;
    const int64_t w = (int64_t)blockIdx.x * kBS + threadIdx.x;
    const uint64_t word = w < nwords ? tbits[w] : 0ull;
    const uint64_t two = w < nwords ? t2bits[w] : 0ull;
    const uint32_t ct = (uint32_t)(__popcll(word) + __popcll(two));
    const uint64_t mine = ((uint64_t)__popcll(word) << 32) | ct;
    uint64_t inc = mine;
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        const uint64_t t = ftd::shfl_up_u64(inc, o, 4);
        if ((int)(threadIdx.x & 3) >= o) inc += t;
    }
    if (w >= nwords || word == 0) return;
    const uint64_t base = blk_prefix[w >> 2] + (inc - mine);
    uint32_t bx = (uint32_t)(base >> 32), bt = (uint32_t)base;
    uint64_t m = word;
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        xtet[bx] = (int32_t)(w * 64 + b);
        xtri_off[bx] = (int32_t)bt;
        bx++;
        bt += 1u + (uint32_t)((two >> b) & 1ull);
    }
}

// Capacity check of a reinit whose buffers were sized in advance (graph
// replay): raises S_CAP when a stage total exceeds what was allocated, which
// halts every later kernel.  One thread.
__global__ void k_check_caps(uint64_t* sc, int64_t x_cap, int64_t f_cap, int64_t p_cap)
{
    uint64_t v = 0;
    if (x_cap >= 0 && (cnt_xtets(sc) > x_cap || cnt_tris(sc) > f_cap)) v |= 1ull;
    if (p_cap >= 0 && cnt_points(sc) > p_cap) v |= 2ull;
    if (v) sc[S_CAP] |= v;
}

// Error word: (code << 40) | tet, min-reduced so the lowest failing tet wins
// within a code; codes: 1 = t2e (FT_ERR_T2E), 2 = degenerate tet.
__device__ __forceinline__ void flag_error(unsigned long long* err, int code, int64_t t)
{
    atomicMin(err, ((unsigned long long)code << 40) | (unsigned long long)t);
}

// Crossing tets: collect crossing edges in t2e order (LevelSet.cpp:385-391),
// check the count (3 or 4, and consistent with the node signs used for the
// triangle offsets: 4 exactly for the tets of the two-triangle bitmap t2bits),
// mark them in the edge bitmap.
__global__ __launch_bounds__(kBS) void k_xtet(const uint64_t* t2bits, MeshView mv,
                                              const uint64_t* sbits,
                                              const int32_t* xtet, uint64_t* sc,
                                              unsigned long long* ebits, int4* xedges)
{
    if (halted(sc)) return;
    const int64_t x = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (x >= cnt_xtets(sc)) return;
    unsigned long long* err = (unsigned long long*)(sc + S_ERR);
    const int64_t t = xtet[x];
    const int k = mv.t2e3 ? 0 : mv_tet_field(mv, t);
    const int2 r0 = mv_t2e(mv, t, 0, k), r1 = mv_t2e(mv, t, 1, k), r2 = mv_t2e(mv, t, 2, k);
    const int te[6] = {r0.x, r0.y, r1.x, r1.y, r2.x, r2.y};
    int ce[4] = {-1, -1, -1, -1};
    int n = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int2 ab = mv_edge(mv, te[j], k);          // a tet's edges are its field's
        if (ftd::get_bit(sbits, ab.x) != ftd::get_bit(sbits, ab.y)) {
            if (n < 4) ce[n] = te[j];
            n++;
        }
    }
    const int expect = ftd::get_bit(t2bits, t) ? 4 : 3;
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

// Position of the n-th (from 0) set bit of w; n < popcount(w).
__device__ __forceinline__ int nth_set_bit(uint64_t w, uint32_t n)
{
    int pos = 0;
#pragma unroll
    for (int sh = 32; sh; sh >>= 1) {
        const uint32_t c = (uint32_t)__popcll(w & ((1ull << sh) - 1ull));
        const bool skip = n >= c;
        n -= skip ? c : 0u;
        w = skip ? (w >> sh) : w;
        pos += skip ? sh : 0;
    }
    return pos;
}

// Crossing points in increasing edge order (LevelSet.cpp:278-299).  A wave
// takes 64 consecutive words of the edge bitmap and their crossing edges 64
// at a time, one per lane: the lane finds the word holding its edge by a
// binary search over the wave's prefix counts, then the bit.  FP64
// interpolation, no fma; marks both edge nodes as retained
// (LevelSet.cpp:443-447).
__global__ __launch_bounds__(kBS) void k_points(const uint64_t* ebits, const uint32_t* eprefix,
                                                int64_t nwords, MeshView mv,
                                                const float* xyz, const double* phi,
                                                float* ifp, int32_t* ifp2e,
                                                unsigned long long* retain, const uint64_t* sc)
{
    if (halted(sc)) return;
    const int lane = ftd::lane_id();
    const int64_t w = (int64_t)blockIdx.x * kBS + threadIdx.x;
    const int64_t w0 = w - lane;                       // the wave's first word
    if (w0 >= nwords) return;                          // whole wave out of range
    const uint64_t word = w < nwords ? ebits[w] : 0ull;
    const uint32_t cnt = (uint32_t)__popcll(word);
    const uint32_t inc = ftd::wave_incl_scan(cnt);
    const uint32_t total = __shfl(inc, 63);
    const uint32_t r0 = eprefix[w0];                   // rank of the wave's first crossing edge
    for (uint32_t k0 = 0; k0 < total; k0 += 64) {
        const uint32_t k = k0 + lane;
        int L = 0;                                     // lanes whose words end at or before k
#pragma unroll
        for (int step = 32; step; step >>= 1) {
            const uint32_t v = __shfl(inc, L + step - 1);
            L += (v <= k) ? step : 0;
        }
        const uint64_t wl = __shfl(word, L);
        const uint32_t ex = __shfl(inc, L) - __shfl(cnt, L);
        if (k >= total) continue;
        const int64_t e = (w0 + L) * 64 + nth_set_bit(wl, k - ex);
        const uint32_t r = r0 + k;
        const int2 ab = mv_edge(mv, e, mv.edges ? 0 : mv_edge_field(mv, e));
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

__global__ __launch_bounds__(kBS) void k_tris(MeshView mv, const float* xyz,
                                              const double* phi, const int32_t* xtet,
                                              const int32_t* xtri_off, const int4* xedges,
                                              const uint64_t* ebits,
                                              const uint32_t* eprefix, const float* ifp,
                                              int4* tris, float* tri_norm, int32_t* if_elems,
                                              uint64_t* sc)
{
    if (halted(sc)) return;
    const int64_t x = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (x >= cnt_xtets(sc)) return;
    unsigned long long* err = (unsigned long long*)(sc + S_ERR);
    const int64_t t = xtet[x];
    if_elems[x] = (int32_t)t;
    const int4 ce = xedges[x];
    if (ce.x < 0) return;                                  // flagged by k_xtet
    const int4 q = mv_tet(mv, t, mv.tets ? 0 : mv_tet_field(mv, t));
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

// Point -> triangle lists and point areas; the per-element steps, with their
// index checks, are in ft_p2t.hpp.
struct AtomicInc {
    __device__ uint32_t operator()(uint32_t* p) const { return atomicAdd(p, 1u); }
};

__global__ __launch_bounds__(kBS) void k_p2t_count(const int4* tris, uint64_t* sc, uint32_t* deg)
{
    if (halted(sc)) return;
    const int64_t f = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (f >= cnt_tris(sc)) return;
    const int4 t = tris[f];
    const int32_t v[3] = {t.x, t.y, t.z};
    if (!ftp::count_tri(v, cnt_points(sc), deg, AtomicInc{}))
        flag_error((unsigned long long*)(sc + S_ERR), ftp::kErrIndex, f);
}

__global__ __launch_bounds__(kBS) void k_p2t_fill(const int4* tris, uint64_t* sc,
                                                  const uint32_t* off, uint32_t* cursor,
                                                  int32_t* p2t)
{
    if (halted(sc)) return;
    const int64_t f = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (f >= cnt_tris(sc)) return;
    const int4 t = tris[f];
    const int32_t v[3] = {t.x, t.y, t.z};
    if (!ftp::fill_tri(v, (int32_t)f, cnt_points(sc), off, cursor, p2t, AtomicInc{}))
        flag_error((unsigned long long*)(sc + S_ERR), ftp::kErrIndex, f);
}

__global__ __launch_bounds__(kBS) void k_point_area(const uint32_t* off, uint64_t* sc,
                                                    int32_t* p2t, const float* tri_norm,
                                                    float* area)
{
    if (halted(sc)) return;
    const int64_t v = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (v >= cnt_points(sc)) return;
    if (!ftp::point_area(v, off, p2t, cnt_tris(sc), tri_norm, area))
        flag_error((unsigned long long*)(sc + S_ERR), ftp::kErrIndex, v);
}

// ============================================================================
// Spatial index: uniform bins by counting sort, binary radix tree on bins
// ============================================================================
// Leaves per block of the local refit pass: a node whose leaf range stays
// inside one block is fitted by that block's workgroup; the rest ("top"
// nodes) by the single-workgroup pass after it.
constexpr int kRefitBlock = 256;

struct __align__(16) BNode {  // child bounds, children, parent
    Bound l, r;
    int32_t left, right, parent;
    int32_t top;                 // 1: the leaf range spans refit blocks
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

// Index geometry of one reinit: per-field grids and counter offsets, the
// field of a crossing point through its edge (edge_off), and the key layout
// (field << code_bits) | Morton code.
struct IndexGeom {
    const int64_t* edge_off;         // nf + 1
    const int64_t* cnt_off;          // nf + 1
    const Grid* grids;               // nf
    int nf;
    int code_bits;
};

// mode 0: points (ifp), mode 1: triangles (centroid).
__device__ __forceinline__ int64_t cnt_prims(int mode, const uint64_t* sc)
{
    return mode == 0 ? cnt_points(sc) : cnt_tris(sc);
}

__global__ __launch_bounds__(kBS) void k_prim_bin(int mode, const uint64_t* sc, const float* ifp,
                                                  const int4* tris, const int32_t* ifp2e,
                                                  IndexGeom ig, uint32_t* code, uint32_t* count)
{
    if (halted(sc)) return;
    const int64_t j = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (j >= cnt_prims(mode, sc)) return;
    float c[3];
    int32_t v0;
    if (mode == 0) {
        for (int k = 0; k < 3; k++) c[k] = ifp[3 * j + k];
        v0 = (int32_t)j;
    } else {
        float a[3], b[3], d[3];
        tri_vertices(tris, ifp, (int32_t)j, a, b, d);
        for (int k = 0; k < 3; k++) c[k] = (a[k] + b[k] + d[k]) * (1.0f / 3.0f);
        v0 = tris[j].x;
    }
    const int f = ifp2e ? ftd::field_of(ig.edge_off, ig.nf, ifp2e[v0]) : 0;
    const uint32_t cc = (uint32_t)ig.cnt_off[f] + ftd::grid_code(ig.grids[f], c[0], c[1], c[2]);
    code[j] = cc;
    atomicAdd(&count[cc], 1u);
}

__global__ __launch_bounds__(kBS) void k_leaves(const uint32_t* count, const uint64_t* prefix,
                                                int64_t G, IndexGeom ig, uint32_t* leaf_key,
                                                uint32_t* leaf_start, uint32_t* leaf_cnt)
{
    const int64_t c = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (c >= G) return;
    const uint32_t n = count[c];
    if (!n) return;
    const uint64_t p = prefix[c];
    const uint32_t l = (uint32_t)(p >> 32);
    const int f = ftd::field_of(ig.cnt_off, ig.nf, c);
    leaf_key[l] = ((uint32_t)f << ig.code_bits) | (uint32_t)(c - ig.cnt_off[f]);
    leaf_start[l] = (uint32_t)p;
    leaf_cnt[l] = n;
}

// Leaf range [leaf_off[f], leaf_off[f+1]) of each field and its tree root:
// none (INT_MIN) without leaves, the leaf itself when alone; fields with two
// or more leaves get the internal node covering exactly their range from
// k_karras.
__global__ void k_field_roots(int nf, const int64_t* cnt_off, const uint64_t* prefix,
                              const uint64_t* total, int64_t G, int32_t* leaf_off, int32_t* roots)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > nf) return;
    const uint32_t L = (uint32_t)(*total >> 32);
    const int64_t c = cnt_off[f];
    const uint32_t lf = c < G ? (uint32_t)(prefix[c] >> 32) : L;
    leaf_off[f] = (int32_t)lf;
    if (f == nf) return;
    const int64_t c1 = cnt_off[f + 1];
    const uint32_t ll = c1 < G ? (uint32_t)(prefix[c1] >> 32) : L;
    roots[f] = (ll == lf) ? INT32_MIN : (ll == lf + 1 ? ~(int32_t)lf : -1);
}

// Scatter primitives into bin order and write their search records.
__global__ __launch_bounds__(kBS) void k_prim_scatter(int mode, const uint64_t* sc, const float* ifp,
                                                      const int4* tris, const uint32_t* code,
                                                      uint32_t* count, const uint64_t* prefix,
                                                      float4* prec, TriRec* trec,
                                                      int32_t* pidx)
{
    if (halted(sc)) return;
    const int64_t j = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (j >= cnt_prims(mode, sc)) return;
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
    // disk about the midpoint of ab holding the triangle (ab is the longest
    // edge), radius rounded up; its sign carries the interior flag
    const double Rm = fmax(0.5 * L0, sqrt((Cx - 0.5 * L0) * (Cx - 0.5 * L0) + Cy * Cy)) * (1.0 + 1e-6);
    r.F0 = make_float4(a[0], a[1], a[2], interior ? (float)Rm : -(float)Rm);
    r.F1 = make_float4((float)u[0], (float)u[1], (float)u[2], (float)L0);
    r.F2 = make_float4((float)v[0], (float)v[1], (float)v[2], (float)nu[0]);
    r.F3 = make_float4((float)nu[1], (float)nu[2], (float)Cx, (float)Cy);
    r.F4 = make_float4(L0 > 0.0 ? (float)(1.0 / L0) : 0.0f, lac > 0.0 ? (float)(1.0 / lac) : 0.0f,
                       lbc > 0.0 ? (float)(1.0 / lbc) : 0.0f, (float)dCx);
    trec[slot] = r;
}

// Unit eigenvector of the smallest eigenvalue of the symmetric 3x3 matrix C
// (c00, c11, c22, c01, c02, c12): the eigenvalue in closed form (O. K. Smith,
// "Eigenvalues of a symmetric 3 x 3 matrix", CACM 4(4), 1961), the vector as
// the longest cross product of two rows of C - lambda I.  False when the
// points are (nearly) isotropic or there is no clear smallest direction.
__device__ inline bool smallest_eigvec(const double c[6], double n[3])
{
    const double p1 = c[3] * c[3] + c[4] * c[4] + c[5] * c[5];
    const double q = (c[0] + c[1] + c[2]) / 3.0;
    const double a = c[0] - q, b = c[1] - q, d = c[2] - q;
    const double p2 = a * a + b * b + d * d + 2.0 * p1;
    if (!(p2 > 0.0)) return false;
    const double p = sqrt(p2 / 6.0);
    const double B[6] = {a / p, b / p, d / p, c[3] / p, c[4] / p, c[5] / p};
    const double det = B[0] * (B[1] * B[2] - B[5] * B[5]) - B[3] * (B[3] * B[2] - B[5] * B[4]) +
                       B[4] * (B[3] * B[5] - B[1] * B[4]);
    const double r = fmin(1.0, fmax(-1.0, det / 2.0));
    const double lam = q + 2.0 * p * cos(acos(r) / 3.0 + 2.0943951023931957);   // smallest
    const double r0[3] = {c[0] - lam, c[3], c[4]}, r1[3] = {c[3], c[1] - lam, c[5]},
                 r2[3] = {c[4], c[5], c[2] - lam};
    double x[3][3];
    ftd::cross3d(r0, r1, x[0]);
    ftd::cross3d(r0, r2, x[1]);
    ftd::cross3d(r1, r2, x[2]);
    int best = 0;
    double bl = ftd::dot3d(x[0], x[0]);
    for (int k = 1; k < 3; k++) {
        const double lk = ftd::dot3d(x[k], x[k]);
        if (lk > bl) { bl = lk; best = k; }
    }
    if (!(bl > 1e-24 * p2 * p2)) return false;
    const double inv = 1.0 / sqrt(bl);
    for (int k = 0; k < 3; k++) n[k] = x[best][k] * inv;
    return true;
}

// The vertices a leaf's bound must hold: the triangles' three vertices, or
// the crossing points themselves.
template <int MODE>
struct LeafVerts {
    static constexpr int kPerPrim = MODE == 0 ? 1 : 3;
    const int32_t* pidx;
    const float* ifp;
    const int4* tris;
    // the vertices of primitive s into v[0..kPerPrim)
    __device__ void get(uint32_t s, float v[3][3]) const
    {
        const int32_t j = pidx[s];
        if (MODE == 0) {
            for (int k = 0; k < 3; k++) v[0][k] = ifp[3 * (int64_t)j + k];
        } else {
            tri_vertices(tris, ifp, j, v[0], v[1], v[2]);
        }
    }
};

// Leaf bounds: box of the vertices and a cylinder about an axis: for
// triangles the area-weighted mean normal, for crossing points the normal of
// their least-squares plane (three or more points that are not collinear).
template <int MODE>
__global__ __launch_bounds__(kBS) void k_leaf_bounds(const uint64_t* sc, const uint32_t* leaf_start,
                                                     const uint32_t* leaf_cnt, const int32_t* pidx,
                                                     const float* ifp, const int4* tris, Bound* leaf_b)
{
    if (halted(sc)) return;
    const int64_t l = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (l >= cnt_leaves(sc)) return;
    using LV = LeafVerts<MODE>;
    const LV lv{pidx, ifp, tris};
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    double ns[3] = {0.0, 0.0, 0.0}, m[3] = {0.0, 0.0, 0.0};
    const uint32_t b = leaf_start[l], e = b + leaf_cnt[l];
    for (uint32_t s = b; s < e; s++) {
        float v[3][3];
        lv.get(s, v);
        for (int q = 0; q < LV::kPerPrim; q++) {
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(lo[k], v[q][k]);
                hi[k] = fmaxf(hi[k], v[q][k]);
                if (MODE == 0) m[k] += v[q][k];
            }
        }
        if (MODE == 1) {
            double e0[3], e1[3], nn[3];
            for (int k = 0; k < 3; k++) { e0[k] = (double)v[1][k] - v[0][k]; e1[k] = (double)v[2][k] - v[0][k]; }
            ftd::cross3d(e0, e1, nn);
            for (int k = 0; k < 3; k++) ns[k] += nn[k];
        }
    }
    Bound B;
    for (int k = 0; k < 3; k++) { B.lo[k] = lo[k]; B.hi[k] = hi[k]; B.c[k] = 0.5f * (lo[k] + hi[k]); }
    B.pad[0] = B.pad[1] = 0.0f;
    double n[3];
    bool axis = false;
    if (MODE == 1) {
        const double nl = sqrt(ftd::dot3d(ns, ns));
        if (nl > 0.0) {
            for (int k = 0; k < 3; k++) n[k] = ns[k] / nl;
            axis = true;
        }
    } else if (e - b >= 3) {
        double cov[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const double cnt = (double)(e - b);
        for (int k = 0; k < 3; k++) m[k] /= cnt;
        for (uint32_t s = b; s < e; s++) {
            float v[3][3];
            lv.get(s, v);
            const double d0 = v[0][0] - m[0], d1 = v[0][1] - m[1], d2 = v[0][2] - m[2];
            cov[0] += d0 * d0; cov[1] += d1 * d1; cov[2] += d2 * d2;
            cov[3] += d0 * d1; cov[4] += d0 * d2; cov[5] += d1 * d2;
        }
        axis = smallest_eigvec(cov, n);
    }
    if (!axis) {
        B.n[0] = 0.0f; B.n[1] = 0.0f; B.n[2] = 1.0f;
        B.rho = 0.0f; B.tau = -1.0f;                 // box only
    } else {
        for (int k = 0; k < 3; k++) { B.n[k] = (float)n[k]; n[k] = B.n[k]; }
        // One pass over the vertices, relative to the box centre c: their
        // extent [hlo, hhi] along n and their largest distance rho0 from the
        // axis through c.  The slab is then centred on that extent, not on the
        // box centre: a surface patch can pass well off its box's centre, and
        // a slab centred off the patch is thicker by that offset.  Moving the
        // centre along n leaves the axis distances as they are, up to the
        // float rounding of the moved centre, which is added to rho.
        double hlo = INFINITY, hhi = -INFINITY, rho0 = 0.0;
        for (uint32_t s = b; s < e; s++) {
            float v[3][3];
            lv.get(s, v);
            for (int q = 0; q < LV::kPerPrim; q++) {
                double d[3], cr[3];
                for (int k = 0; k < 3; k++) d[k] = (double)v[q][k] - B.c[k];
                const double hv = ftd::dot3d(n, d);
                hlo = fmin(hlo, hv);
                hhi = fmax(hhi, hv);
                ftd::cross3d(n, d, cr);
                rho0 = fmax(rho0, ftd::dot3d(cr, cr));
            }
        }
        const double mid = 0.5 * (hlo + hhi);
        double sh[3], cr[3];
        for (int k = 0; k < 3; k++) {
            const float ck = (float)(B.c[k] + mid * n[k]);
            sh[k] = (double)ck - B.c[k];              // the centre's actual move
            B.c[k] = ck;
        }
        const double m = ftd::dot3d(n, sh);
        ftd::cross3d(n, sh, cr);
        const double tau = fmax(hhi - m, m - hlo);
        const double rho = sqrt(rho0) + sqrt(ftd::dot3d(cr, cr));
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

__global__ __launch_bounds__(kBS) void k_karras(const uint32_t* keys, const uint64_t* sc, BNode* nodes,
                                                int32_t* leaf_parent, int code_bits,
                                                const int32_t* leaf_off, int32_t* roots)
{
    if (halted(sc)) return;
    const int L = (int)cnt_leaves(sc);
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
    // the node whose range is exactly one field's leaves is that field's root
    const int lo = min(i, j), hi = max(i, j);
    nodes[i].top = (lo / kRefitBlock) != (hi / kRefitBlock);
    const int f = (int)(keys[lo] >> code_bits);
    if (leaf_off[f] == lo && leaf_off[f + 1] == hi + 1) roots[f] = i;
}

// Parent bound of two children: box union and a cylinder that contains both
// child cylinders.  For a child cylinder (c_i, n_i, rho_i, tau_i) seen from
// the parent axis n at centre c, with cos = |n.n_i| and sin = |n x n_i|:
//   thickness  |n.(c_i-c)| + tau_i cos + rho_i sin
//   radius     |n x (c_i-c)| + rho_i + tau_i sin
// In FP32 with hardware square roots (the merge is on the refit's critical
// path, one per tree level): every term is a sum of non-negative parts with a
// relative error under 1e-6, the axis is unit to 3e-7, and the results are
// rounded up by 2^-12 relative plus 2^-18 of the box extent, so the cylinder
// still contains both children.
__device__ inline Bound merge_bounds(const Bound& A, const Bound& Bb)
{
    Bound P;
    float ext = 0.0f;
    for (int k = 0; k < 3; k++) {
        P.lo[k] = fminf(A.lo[k], Bb.lo[k]);
        P.hi[k] = fmaxf(A.hi[k], Bb.hi[k]);
        P.c[k] = 0.5f * (P.lo[k] + P.hi[k]);
        ext = fmaxf(ext, P.hi[k] - P.lo[k]);
    }
    P.pad[0] = P.pad[1] = 0.0f;
    if (A.tau < 0.0f || Bb.tau < 0.0f) {
        P.n[0] = 0.0f; P.n[1] = 0.0f; P.n[2] = 1.0f;
        P.rho = 0.0f; P.tau = -1.0f;
        return P;
    }
    const float wa = A.rho * A.rho + 1e-30f, wb = Bb.rho * Bb.rho + 1e-30f;
    float n[3];
    for (int k = 0; k < 3; k++) n[k] = wa * A.n[k] + wb * Bb.n[k];
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (nn > 1e-20f * (wa + wb) * (wa + wb) && nn > 1e-30f) {
        const float inv = __builtin_amdgcn_rsqf(nn);
        for (int k = 0; k < 3; k++) n[k] *= inv;
    } else {
        for (int k = 0; k < 3; k++) n[k] = A.n[k];
    }
    // The centre may be any point (thickness and radius below are taken
    // about the rounded one); it is moved along n to centre the slab on the
    // children's extents along n.
    const Bound* ch[2] = {&A, &Bb};
    float cs[2], sn[2], hlo = INFINITY, hhi = -INFINITY;
    for (int i = 0; i < 2; i++) {
        const Bound& C = *ch[i];
        cs[i] = fminf(1.0f, fabsf(n[0] * C.n[0] + n[1] * C.n[1] + n[2] * C.n[2]));
        const float sx = n[1] * C.n[2] - n[2] * C.n[1], sy = n[2] * C.n[0] - n[0] * C.n[2],
                    sz = n[0] * C.n[1] - n[1] * C.n[0];
        sn[i] = __builtin_amdgcn_sqrtf(sx * sx + sy * sy + sz * sz);
        const float hc = n[0] * (C.c[0] - P.c[0]) + n[1] * (C.c[1] - P.c[1]) + n[2] * (C.c[2] - P.c[2]);
        const float ec = C.tau * cs[i] + C.rho * sn[i];
        hlo = fminf(hlo, hc - ec);
        hhi = fmaxf(hhi, hc + ec);
    }
    const float mid = 0.5f * (hlo + hhi);
    for (int k = 0; k < 3; k++) P.c[k] += mid * n[k];
    float tau = 0.0f, rho = 0.0f;
    for (int i = 0; i < 2; i++) {
        const Bound& C = *ch[i];
        const float dc[3] = {C.c[0] - P.c[0], C.c[1] - P.c[1], C.c[2] - P.c[2]};
        const float ax = fabsf(n[0] * dc[0] + n[1] * dc[1] + n[2] * dc[2]);
        const float rx = n[1] * dc[2] - n[2] * dc[1], ry = n[2] * dc[0] - n[0] * dc[2],
                    rz = n[0] * dc[1] - n[1] * dc[0];
        const float rad = __builtin_amdgcn_sqrtf(rx * rx + ry * ry + rz * rz);
        tau = fmaxf(tau, ax + C.tau * cs[i] + C.rho * sn[i]);
        rho = fmaxf(rho, rad + C.rho + C.tau * sn[i]);
    }
    for (int k = 0; k < 3; k++) P.n[k] = n[k];
    P.tau = tau * (1.0f + 0x1.0p-12f) + 0x1.0p-18f * ext;
    P.rho = rho * (1.0f + 0x1.0p-12f) + 0x1.0p-18f * ext;
    return P;
}

// Bottom-up fitting: the second child to arrive at a node merges both.
// A cylinder whose volume is not below alpha times its box's prunes little
// beyond the box test and costs more: such a bound is kept as a box only.
__device__ __forceinline__ void drop_loose_cylinder(Bound& b, float alpha)
{
    if (alpha <= 0.0f || b.tau < 0.0f) return;
    const float vb = (b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
    const float vc = 3.14159265f * b.rho * b.rho * 2.0f * b.tau;
    if (vc >= alpha * vb) { b.rho = 0.0f; b.tau = -1.0f; }
}

// Bottom-up fitting in two passes, with no device-wide fence: with one L2 per
// XCD, an agent-scope release or acquire writes back or invalidates that L2,
// and a fence per tree level made the one-pass form the index's slowest
// kernel.  Karras node i spans leaves [min(i,j), max(i,j)] with i one end, so
// a node inside refit block b has its index in that block.
//
// Pass 1, one workgroup per block of kRefitBlock leaves: a thread per leaf
// climbs while the parent is inside the block.  Arrivals are counted in LDS;
// the second child to arrive merges both (workgroup-scope fences order the
// bound stores, which the workgroup's waves see through their CU).  At a top
// node it stores its bound and counts the arrival in flags[p]; the second
// arrival there appends p to the ready list (both children done).
__global__ __launch_bounds__(kRefitBlock) void k_refit_local(const uint64_t* sc, const int32_t* leaf_parent,
                                                             const Bound* leaf_b, BNode* nodes,
                                                             unsigned* flags, unsigned* ready,
                                                             float alpha)
{
    __shared__ unsigned arrived[kRefitBlock];
    if (halted(sc)) return;
    const int L = (int)cnt_leaves(sc);
    const int b0 = blockIdx.x * kRefitBlock;
    if (L < 2 || b0 >= L) return;                 // uniform over the block
    arrived[threadIdx.x] = 0u;
    __syncthreads();
    const int l = b0 + threadIdx.x;
    if (l >= L) return;
    Bound cur_b = leaf_b[l];
    int cur = ~l;
    int p = leaf_parent[l];
    for (int depth = 0; p >= 0 && depth < 64; depth++) {   // depth <= 32 for 32-bit keys
        BNode* nd = &nodes[p];
        const bool is_left = (nd->left == cur);
        if (is_left) nd->l = cur_b; else nd->r = cur_b;
        if (nd->top) {
            if (atomicAdd(&flags[p], 1u) == 1u) ready[1 + atomicAdd(&ready[0], 1u)] = (unsigned)p;
            return;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (atomicAdd(&arrived[p - b0], 1u) == 0u) return;   // sibling not done yet
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const Bound sib = is_left ? nd->r : nd->l;
        cur_b = is_left ? merge_bounds(cur_b, sib) : merge_bounds(sib, cur_b);
        drop_loose_cylinder(cur_b, alpha);
        cur = p;
        p = nd->parent;
    }
}

// Pass 2, one workgroup: from each ready top node up.  flags[p] already counts
// the children that pass 1 finished; the thread whose arrival completes a node
// carries on with it.  All threads share one workgroup (one CU), so the
// bound stores need only workgroup-scope fences; pass 1's stores are visible
// across the kernel boundary.
constexpr int kRefitTopThreads = 1024;
__global__ __launch_bounds__(kRefitTopThreads) void k_refit_top(const uint64_t* sc, BNode* nodes,
                                                                unsigned* flags, const unsigned* ready,
                                                                float alpha)
{
    if (halted(sc)) return;
    const unsigned nr = ready[0];
    for (unsigned k = threadIdx.x; k < nr; k += kRefitTopThreads) {
        int cur = (int)ready[1 + k];
        Bound cur_b = merge_bounds(nodes[cur].l, nodes[cur].r);
        drop_loose_cylinder(cur_b, alpha);
        int p = nodes[cur].parent;
        for (int depth = 0; p >= 0 && depth < 64; depth++) {
            BNode* nd = &nodes[p];
            const bool is_left = (nd->left == cur);
            if (is_left) nd->l = cur_b; else nd->r = cur_b;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (atomicAdd(&flags[p], 1u) == 0u) break;   // the other child is still on its way
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const Bound sib = is_left ? nd->r : nd->l;
            cur_b = is_left ? merge_bounds(cur_b, sib) : merge_bounds(sib, cur_b);
            drop_loose_cylinder(cur_b, alpha);
            cur = p;
            p = nd->parent;
        }
    }
}

// ============================================================================
// Nearest-primitive search with the fused reinit epilogue
// ============================================================================
enum Epi : int { EPI_RAW = 0, EPI_INIT = 1, EPI_UPDATE = 2 };

struct QueryArgs {
    int64_t n_slots;                 // query slots (per-field Morton order, 64-padded)
    const float *sx, *sy, *sz;       // slot coordinates
    const int32_t* perm;             // slot -> node (-1: padding)
    const int32_t* tile_field;       // 64-slot tile -> field
    const int32_t* tile_list;        // launch order of the tiles, deep ones first (null: slot order)
    const uint64_t* n_deep;          // deep tiles at the head of tile_list (low 32 bits)
    const int32_t* roots;            // field -> tree root (>= 0 node, < 0 ~leaf, INT_MIN none)
    const float4* prec;              // point records (mode 0)
    const TriRec* trec;              // triangle records (mode 1)
    const int32_t* pidx;             // sorted slot -> primitive id
    const uint32_t* leaf_start;
    const uint32_t* leaf_cnt;
    const Bound* leaf_b;
    const BNode* nodes;
    int L;                           // leaves
    float band;                      // > 0: narrow band
    const float* ifp;
    const int4* tris;
    const double* phi_old;           // sign + retained values
    double* phi_out;
    double* phi_ifn;                 // nullable
    const uint64_t* retain;          // nullable
    int32_t* nearest;                // nullable
    int epi;
    double w_phi, w_ifn;
    unsigned long long* n_clamped;   // per field
    unsigned* overflow;
    unsigned long long* stats;       // nullable: per distance bucket
    float h_est;                     // node spacing (statistics only)
    int probe;                       // greedy first candidate
    const uint64_t* sc;              // stage totals and halt flags
    unsigned xcd_chunk;              // tiles per XCD chunk (0: dispatch order)
    uint32_t* queue;                 // work-queue schedule: deep counter, then one per XCD (null: static)
    ulonglong2* split_out;           // best and second key per lane of each part of a split tile (null: none split)
};

// The dispatcher hands workgroup b to XCD b % 8, each XCD with its own L2.
// Remapped, XCD x processes runs of `chunk` consecutive tiles (in Morton
// order, a compact region whose triangle and tree records its L2 keeps)
// instead of every 8th tile; the runs are dealt to the XCDs in turn, so costly
// regions still spread over all of them (one contiguous run per XCD was
// slower: the XCD holding the costliest region finished last).  A bijection
// on [0, nb).
constexpr unsigned kXcds = 8;
// Workgroups per run, measured at 256^3 (triangles, deep tiles first): with
// 256-thread workgroups 512 (2048 tiles; 1024 is 4% slower), with 64-thread
// workgroups 1024 (1024 tiles; 512 is 1.3% slower, 2048 0.5%).
constexpr unsigned kXcdChunk = 512;
constexpr unsigned kXcdChunk64 = 1024;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, unsigned chunk)
{
    if (chunk == 0) return b;
    // chunks of `chunk` consecutive tiles, dealt to the XCDs in turn
    const unsigned span = kXcds * chunk;
    const unsigned full = nb - nb % span;          // blocks in whole spans
    if (b >= full) return b;                       // the ragged tail keeps its order
    const unsigned x = b % kXcds, i = b / kXcds;
    return (i / chunk) * span + x * chunk + i % chunk;
}

// Margin of the box prune: a box is skipped only if its lower bound exceeds
// the best distance by more than the float rounding of both (relative) plus
// an absolute floor for near-zero distances.
__device__ __forceinline__ float prune_bound(float best2)
{
    return __fmaf_rn(best2, 1.0f + 0x1.0p-16f, 1.0e-30f);
}

// A candidate as one 64-bit key: the distance's bits above the primitive's
// id.  Distances are non-negative, so their bit patterns order as their
// values, and key order is the order (distance, then lower id); no candidate
// is id -1, the largest id as unsigned.
__device__ __forceinline__ uint64_t cand_key(float d, int32_t j)
{
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)j;
}

template <int MODE>
struct Searcher {
    float x, y, z;
    uint64_t bkey;   // best: mode 0 distance, mode 1 squared distance; primitive (original id)
    const float4* prec;
    const TriRec* trec;
    const int32_t* pidx;
    const uint32_t* leaf_start;
    const uint32_t* leaf_cnt;
    const float* ifp;
    const int4* tris;

    __device__ __forceinline__ float best() const { return __uint_as_float((uint32_t)(bkey >> 32)); }
    __device__ __forceinline__ int32_t bj() const { return (int32_t)(uint32_t)bkey; }
    __device__ __forceinline__ float sec() const { return __uint_as_float((uint32_t)(skey >> 32)); }
    __device__ __forceinline__ int32_t sj() const { return (int32_t)(uint32_t)skey; }
    __device__ __forceinline__ float best2() const { return MODE == 0 ? best() * best() : best(); }
    float bnd;       // prune_bound(best2()), refreshed whenever best changes
    float bnd_r;     // mode 1: rect_bound(bnd), the triangle cull's bound
    // no candidate yet, distance d (the band, or no limit)
    __device__ __forceinline__ void init_best(float d)
    {
        bkey = cand_key(d, -1);
        skey = bkey;
        bnd = prune_bound(best2());
        bnd_r = ftd::rect_bound(bnd);
    }

    // Mode 0 keeps the nearest point (ties: lowest id, the reference's strict
    // < over ascending ids).  Mode 1 keeps the best two triangles by FP32
    // distance, so that the epilogue can settle near-ties in FP64.  j is the
    // primitive's original id (wave-uniform).  Written with selects, not
    // branches: a lane-divergent branch costs exec-mask bookkeeping on every
    // call, which is most of the work here.
    // d is finite and non-negative.  A primitive seen again has the same key
    // as before, so it is neither a new best nor a new second.
    __device__ __forceinline__ void consider(float d, int32_t j)
    {
        // (bitwise & and |: short-circuit operators compile to branches)
        const uint64_t key = cand_key(d, j);
        const bool nb = key < bkey;
        if (MODE == 1) {
            // the old best becomes second (no candidate yet: it equals the
            // initial second); else a key between the two is the new second
            const bool ns = !nb & (key != bkey) & (key < skey);
            skey = nb ? bkey : (ns ? key : skey);
        }
        bkey = nb ? key : bkey;
        const float nbnd = prune_bound(MODE == 0 ? d * d : d);
        bnd = nb ? nbnd : bnd;
        if (MODE == 1) bnd_r = nb ? ftd::rect_bound(nbnd) : bnd_r;
    }

    uint32_t n_leaves = 0, n_prims = 0, n_pairs = 0;   // statistics (FT_QUERY_STATS)
    uint64_t skey;   // mode 1: second best squared distance and its triangle
    int parts = 1, part = 0;   // a split tile's part: only leaves l with l % parts == part are tested

    // One triangle for all lanes: a wave-uniform early out when the bounding
    // rectangle of its frame rules it out for every lane, else the exact FP32
    // distance.
    __device__ __forceinline__ void test_tri(const TriRec& r, int32_t j)
    {
        const ftd::TriProj pj = ftd::tri_project(x, y, z, r);
        if (__ballot(ftd::tri_rect_d2(pj, r) <= bnd_r) == 0ull) return;
        consider(ftd::tri_dist2_from(pj, r), j);
    }

    // Leaf test at a wave-uniform leaf: records read through scalar loads.
    __device__ void test_leaf_uniform(int l, uint64_t admit = ~0ull)
    {
        if (parts > 1 && l % parts != part) return;
        const uint32_t b = ftd::cload(leaf_start, l), e = b + ftd::cload(leaf_cnt, l);
        n_leaves++;
        n_prims += e - b;
        n_pairs += (uint32_t)__popcll(admit) * (e - b);   // lane-triangle pairs whose lane admits the leaf
        if (MODE == 0) {
            for (uint32_t s = b; s < e; s++) {
                const float4 p = ftd::cload(prec, s);
                const int32_t j = ftd::cload(pidx, s);
                const float dx = p.x - x, dy = p.y - y, dz = p.z - z;
                const float s2 = dx * dx + dy * dy + dz * dz;
                // bnd = best^2 (1 + 2^-16): beyond it sqrtf(s2) > best, which the
                // strict compare would reject anyway, so the point is skipped
                // when no lane is within it
                if (__ballot(s2 <= bnd) == 0ull) continue;
                const float d = sqrtf(s2);                            // EuclideanDistance.cu:20-28
                consider(d, j);
            }
        } else {
            // two records per step: their scalar loads are in flight together
            // (three or four need more scalar registers than full occupancy
            // leaves: 106 SGPRs, occupancy 7, spills)
            uint32_t s = b;
            for (; s + 1 < e; s += 2) {
                const TriRec r0 = ftd::cload(trec, s);
                const TriRec r1 = ftd::cload(trec, s + 1);
                const int32_t j0 = ftd::cload(pidx, s), j1 = ftd::cload(pidx, s + 1);
                test_tri(r0, j0);
                test_tri(r1, j1);
            }
            if (s < e) test_tri(ftd::cload(trec, s), ftd::cload(pidx, s));
        }
    }

};

// Epilogue of one node (EuclideanDistance.cu:54-57, LevelSet.cpp:438-448, 74,
// 148): sign from phi_old < 0, retained value at crossing-edge nodes, blend.
__device__ __forceinline__ void node_epilogue(const QueryArgs& a, int64_t node, double d, int32_t near)
{
    const double po = a.phi_old[node];
    const double dn = (po < 0.0) ? -d : d;
    if (a.phi_ifn) a.phi_ifn[node] = dn;
    if (a.nearest) a.nearest[node] = near;
    double out = dn;
    if (a.epi == EPI_UPDATE) {
        const double kept = (a.retain && ftd::get_bit(a.retain, node)) ? po : dn;
        out = a.w_phi * kept + a.w_ifn * dn;
    }
    a.phi_out[node] = out;
}

// FP64 squared distance from point (x, y, z) to triangle f of the interface.
__device__ inline double node_tri_dist2(const QueryArgs& a, float x, float y, float z, int32_t f)
{
    const int4 t = a.tris[f];
    const float p[3] = {x, y, z};
    float va[3], vb[3], vc[3];
    for (int k = 0; k < 3; k++) {
        va[k] = a.ifp[3 * (int64_t)t.x + k];
        vb[k] = a.ifp[3 * (int64_t)t.y + k];
        vc[k] = a.ifp[3 * (int64_t)t.z + k];
    }
    return ftd::tri_dist2_f64(p, va, vb, vc);
}

// End of the search of one node, then its epilogue.  Points: the float
// distance, as the reference.  Triangles: the FP32 search knows the nearest
// triangle up to its rounding (at 256^3 float rounding alone nears 1e-5 h),
// so the best and, when that close, the runner-up are evaluated in FP64; the
// lower wins, ties to the lower id.  No candidate: +-band in a narrow band,
// else +-FLT_MAX.
template <int MODE>
__device__ __forceinline__ void finish_node(const QueryArgs& a, const Searcher<MODE>& S, int64_t node)
{
    if (MODE == 0) {
        const float d = (a.band > 0.0f && S.bj() < 0) ? a.band : S.best();
        node_epilogue(a, node, (double)d, S.bj());
    } else {
        double d;
        int32_t near = S.bj();
        if (near >= 0) {
            double d2 = node_tri_dist2(a, S.x, S.y, S.z, near);
            const int32_t sj = S.sj();
            const bool close = sj >= 0 && S.sec() <= S.best() * (1.0f + 0x1.0p-16f);
            if (close) {
                const double e2 = node_tri_dist2(a, S.x, S.y, S.z, sj);
                if (e2 < d2 || (e2 == d2 && sj < near)) { d2 = e2; near = sj; }
            }
            d = sqrt(d2);
        } else {
            d = (double)(a.band > 0.0f ? a.band : FLT_MAX);
        }
        node_epilogue(a, node, d, near);
    }
}

template <int MODE>
__device__ __forceinline__ void record_stats(const QueryArgs& a, const Searcher<MODE>& S,
                                             int64_t node, uint32_t n_nodes)
{
    // buckets: distance < 2h, < 8h, < 32h, beyond; +4 when phi < 0
    const float dd = MODE == 0 ? S.best() : sqrtf(S.best());
    const float hh = a.h_est;
    int bkt = dd < 2 * hh ? 0 : dd < 8 * hh ? 1 : dd < 32 * hh ? 2 : 3;
    if (a.phi_old[node] < 0.0) bkt += 4;
    unsigned long long* st = a.stats + 5 * bkt;
    atomicAdd(&st[0], 1ull);
    atomicAdd(&st[1], (unsigned long long)n_nodes);
    atomicAdd(&st[2], (unsigned long long)S.n_leaves);
    atomicAdd(&st[3], (unsigned long long)S.n_prims);
    atomicAdd(&st[4], (unsigned long long)S.n_pairs);
}

// Query schedule.  A tile that lies deep inside the surface costs 10-20x the
// median tile (near the middle of a closed surface every triangle is nearly
// equidistant, and the exact search must certify them all); started late,
// such tiles set the kernel's end.  A tile is deep when one of its nodes has
// phi_old <= -kDeepTileH h of its field.  Deep tiles are listed first and
// take the lowest workgroups, which the dispatcher starts first and deals
// over all XCDs; the rest keep their slot (Morton) order.  Measured at 256^3
// (64 spheres, full domain): step 15.3 -> 12.6 ms; thresholds of 6-10 h do
// about the same there, 10 h is best on the 128^3 field, 3 h and 20 h lose.
constexpr float kDeepTileH = 10.0f;

// The deepest tiles still set the kernel's end when they start first: at
// 256^3 the tiles at the centres of the largest spheres (phi_old near -29 h)
// test some 25k triangles each, and one wave's chain of dependent record
// loads took 9.3-9.6 ms of a 9.6 ms kernel while the other tiles had ended by
// about 9.0 ms (per-tile timestamps, a probe build).  A tile is very deep when
// one of its nodes has phi_old <= -kSplitTileH h: it is searched by
// kSplitParts waves, each with all 64 nodes but testing only the leaves l with
// l % kSplitParts equal to its part, and k_split_merge combines the parts.
// Each part repeats the walk of the inner tree nodes, so splitting costs work:
// at 256^3 (query ms, this field and its mirror image across x = z) 2 parts
// from 25 h gave 9.05 / 8.98, 4 parts 9.07-9.25 / 9.02, 4 parts from 20 h
// 9.17 / 9.25, 8 parts 9.36, from 30 h the deepest tiles ran whole again
// (10.9), and the unsplit query took 9.6-11.2 / 8.9-9.4 depending on the
// order of the tiles.  Four parts keep more margin for fields deeper than
// this one.  At most split_tiles_cap tiles per launch are split (the parts'
// results take kSplitParts x 1 KiB per tile); the rest run whole.  The split
// only changes which wave tests a leaf: each node's result is the same.
constexpr float kSplitTileH = 25.0f;
constexpr int kSplitParts = 4;
constexpr int64_t kSplitCap = 1024;

__host__ __device__ inline int64_t split_tiles_cap(int64_t n_tiles)
{
    const int64_t cap = n_tiles / 64 > kSplitCap ? n_tiles / 64 : kSplitCap;
    return n_tiles < cap ? n_tiles : cap;
}

// One wave per tile: cls[t] = 2 when some node of tile t is very deep, 1 when
// deep, else 0.
__global__ __launch_bounds__(kBS) void k_tile_deep(int64_t n_tiles, const int32_t* perm, const double* phi,
                                                   const int32_t* tile_field, const Grid* grids,
                                                   uint32_t* cls, const uint64_t* sc)
{
    if (halted(sc)) return;
    const int64_t t = ((int64_t)blockIdx.x * kBS + threadIdx.x) >> 6;
    if (t >= n_tiles) return;                          // whole wave
    const int32_t node = perm[t * 64 + ftd::lane_id()];
    const float h = grids[tile_field[t]].h;
    const double p = node >= 0 ? phi[node] : 0.0;
    const uint64_t m1 = __ballot(p <= -(double)(kDeepTileH * h));
    const uint64_t m2 = __ballot(p <= -(double)(kSplitTileH * h));
    if (ftd::lane_id() == 0) cls[t] = m2 ? 2u : (m1 ? 1u : 0u);
}

// Stable partition from the exclusive scan ex of LoadTileClass(cls) (K deep
// tiles, K2 of them very deep): the very deep tiles to list[0, K2), the other
// deep ones to list[K2, K), the rest after them, each in tile order.
__global__ __launch_bounds__(kBS) void k_tile_list(int64_t n_tiles, const uint32_t* cls, const uint64_t* ex,
                                                   const uint64_t* sc, int32_t* list)
{
    if (halted(sc)) return;
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (t >= n_tiles) return;
    const uint64_t tot = sc[S_DEEP];
    const int64_t K = (int64_t)(uint32_t)tot, K2 = (int64_t)(tot >> 32);
    const int64_t e1 = (int64_t)(uint32_t)ex[t], e2 = (int64_t)(ex[t] >> 32);
    const uint32_t c = cls[t];
    const int64_t at = c == 2 ? e2 : (c == 1 ? K2 + (e1 - e2) : K + (t - e1));
    list[at] = (int32_t)t;
}

// Wave-coherent search: the 64 lanes of a wave hold 64 consecutive slots of
// one field in Morton order (a compact tile) and walk that field's tree
// together.  A child is entered when any lane's bound admits it; each lane
// keeps its own best.  All control flow is wave-uniform (ballots), node and
// primitive records are read at wave-uniform addresses through scalar loads,
// the stack is one per wave in LDS.
// split >= 0: part split % kSplitParts of split tile split / kSplitParts,
// whose best and second keys go to a.split_out instead of the epilogue.
template <int MODE>
__device__ __forceinline__ void search_tile(const QueryArgs& a, int64_t base, int32_t* stack, int split = -1)
{
    const int64_t slot = base + (threadIdx.x & 63);
    const int32_t node = a.perm[slot];                 // -1: padding lane
    const float px = a.sx[slot], py = a.sy[slot], pz = a.sz[slot];
    const int field = __builtin_amdgcn_readfirstlane(a.tile_field[base >> 6]);
    const int32_t root = __builtin_amdgcn_readfirstlane(ftd::cload(a.roots, field));
    Searcher<MODE> S;
    S.x = px; S.y = py; S.z = pz;
    if (MODE == 0) S.init_best(a.band > 0.0f ? a.band : FLT_MAX);
    else S.init_best(a.band > 0.0f ? a.band * a.band : INFINITY);
    S.prec = a.prec; S.trec = a.trec; S.pidx = a.pidx;
    S.leaf_start = a.leaf_start; S.leaf_cnt = a.leaf_cnt;
    S.ifp = a.ifp; S.tris = a.tris;
    if (split >= 0) {
        S.parts = kSplitParts;
        S.part = __builtin_amdgcn_readfirstlane(split % kSplitParts);
    }
    uint32_t n_nodes = 0;
    if (root != INT32_MIN && root < 0) {
        S.test_leaf_uniform(~root);
    } else if (root >= 0) {
        if (a.probe) {
            // Greedy descent for the tile's middle slot to a leaf near it: a
            // first exact candidate for every lane.  The child whose bound
            // centre is nearer is taken (cheaper than the bounds and, in a
            // model of the traversal, a seed after which less is tested).
            const float cx = __shfl(px, 32), cy = __shfl(py, 32), cz = __shfl(pz, 32);
            int cur = root;
            for (int depth = 0; depth < kStack + 1 && cur >= 0; depth++) {
                const BNode nd = ftd::cload(a.nodes, cur);
                const float lx = cx - nd.l.c[0], ly = cy - nd.l.c[1], lz = cz - nd.l.c[2];
                const float rx = cx - nd.r.c[0], ry = cy - nd.r.c[1], rz = cz - nd.r.c[2];
                const float ll = __fmaf_rn(lz, lz, __fmaf_rn(ly, ly, lx * lx));
                const float lr = __fmaf_rn(rz, rz, __fmaf_rn(ry, ry, rx * rx));
                cur = __builtin_amdgcn_readfirstlane(ll <= lr ? nd.left : nd.right);
            }
            if (cur < 0) S.test_leaf_uniform(~cur);
        }
        int sp = 0;
        int cur = root;
        int budget = 2 * a.L + 8;
        while (--budget > 0) {
            const BNode nd = ftd::cload(a.nodes, cur);    // scalar loads: cur is uniform
            const int left = nd.left, right = nd.right;
            const float bnd = S.bnd;
            // box bound first; the cylinder only where some lane's box admits it
            float ll = ftd::box_dist2(px, py, pz, nd.l.lo, nd.l.hi);
            float lr = ftd::box_dist2(px, py, pz, nd.r.lo, nd.r.hi);
            if (__ballot(ll <= bnd)) ll = ftd::max_nonneg(ll, ftd::cyl_dist2(px, py, pz, nd.l));
            if (__ballot(lr <= bnd)) lr = ftd::max_nonneg(lr, ftd::cyl_dist2(px, py, pz, nd.r));
            n_nodes++;
            const uint64_t mL = __ballot(ll <= bnd), mR = __ballot(lr <= bnd);
            const uint64_t any = mL | mR;
            const bool left_first = 2 * __popcll(__ballot(ll <= lr) & any) >= __popcll(any);
            const int c0 = left_first ? left : right, c1 = left_first ? right : left;
            const uint64_t m0 = left_first ? mL : mR, m1 = left_first ? mR : mL;
            int next = -1;
            if (m0) {
                if (c0 < 0) S.test_leaf_uniform(~c0, m0);
                else next = c0;
            }
            if (m1) {
                if (c1 < 0) S.test_leaf_uniform(~c1, m1);
                else if (next < 0) next = c1;
                else if (sp < kStack) stack[sp++] = c1;
                else atomicOr(a.overflow, 1u);
            }
            if (next < 0) {
                if (sp == 0) break;
                next = stack[--sp];
            }
            cur = __builtin_amdgcn_readfirstlane(next);
        }
        if (budget <= 0) atomicOr(a.overflow, 2u);
    }
    if (split >= 0) {                                  // full domain: nothing is clamped
        a.split_out[(int64_t)split * 64 + (threadIdx.x & 63)] = make_ulonglong2(S.bkey, S.skey);
        return;
    }
    const bool clamped = node >= 0 && a.band > 0.0f && S.bj() < 0;
    const uint64_t cm = __ballot(clamped);
    if ((threadIdx.x & 63) == 0 && cm) atomicAdd(&a.n_clamped[field], (unsigned long long)__popcll(cm));
    if (node < 0) return;
    if (a.stats) record_stats<MODE>(a, S, node, n_nodes);
    finish_node<MODE>(a, S, node);
}

// Query schedule of a launch: deep tiles first (kDeepTileH), the rest in
// XCD-chunked Morton order (xcd_block).  Static: one tile per wave, the
// deep tiles on the first workgroups (which the dispatcher deals over all
// XCDs; a very deep tile on kSplitParts of them, a.split_out), the others
// after them.  Work queue (a.queue; every tile whole): a persistent grid
// whose waves take tiles from counters -- the deep tiles from one shared
// counter, then the rest from the counter of their XCD (workgroup b runs on
// XCD b % 8), each XCD's share being its chunks of the static order; a wave
// whose XCD has none left takes from the other XCDs' counters in turn.
// Every wave leaves once all nine counters are past their ends.
// Launch bound: WPS waves per SIMD (8: at most 64 VGPRs).  Unbounded, the
// FP64 settle of the epilogue takes 84 VGPRs and the occupancy drops to 5
// (measured 25% slower at 256^3); bounded at 8, it spills a few registers
// that are touched once per node.
template <int MODE, int BS, int WPS>
__global__ __launch_bounds__(BS, WPS) void k_query_wave(QueryArgs a)
{
    __shared__ int32_t wstack[BS / 64][kStack];
    if (halted(a.sc)) return;                         // an earlier stage failed: phi untouched
    const int w = threadIdx.x >> 6;
    if (a.queue) {
        // with no tile list (narrow band): the tiles in slot order, none deep
        const int64_t n_tiles = a.n_slots / 64;
        const uint32_t K = a.tile_list ? (uint32_t)ftd::cload(a.n_deep, 0) : 0u;
        const uint32_t nn = (uint32_t)(n_tiles - K);
        const uint32_t chunk = a.xcd_chunk;
        const unsigned xcd = blockIdx.x % kXcds;
        unsigned src = 0;                              // 0: deep counter; 1..8: XCD counters
        for (;;) {
            uint32_t got = 0;
            if (ftd::lane_id() == 0) got = atomicAdd(&a.queue[src == 0 ? 0 : 1 + (xcd + src - 1) % kXcds], 1u);
            got = (uint32_t)__shfl(got, 0);
            int64_t r = -1;
            if (src == 0) {
                if (got < K) r = got;
            } else {
                // the got-th tile of XCD x: its chunks are x, x + 8, x + 16, ...
                // (increasing in got, so once past nn it stays past)
                const unsigned x = (xcd + src - 1) % kXcds;
                const uint64_t t = (uint64_t)(got / chunk) * kXcds * chunk + (uint64_t)x * chunk + got % chunk;
                if (t < nn) r = (int64_t)K + (int64_t)t;
            }
            if (r >= 0) {
                search_tile<MODE>(a, (a.tile_list ? (int64_t)a.tile_list[r] : r) * 64, wstack[w]);
                continue;
            }
            if (++src > kXcds) break;                  // every counter exhausted
        }
        return;
    }
    int64_t base;
    if (a.tile_list) {
        // deep tiles on the first Kb workgroups (Kb a multiple of the 8 XCDs):
        // first kSplitParts waves for each split tile, then one per other deep
        // tile; the rest on the workgroups after them in XCD-chunked order
        constexpr unsigned wpb = BS / 64;
        const int64_t n_tiles = a.n_slots / 64;
        const uint64_t nd = ftd::cload(a.n_deep, 0);
        const unsigned K = (unsigned)nd;
        // (explicit compares: min() of two 64-bit integers resolves to a
        // floating-point overload here and rounds them)
        const uint64_t cap = (uint64_t)split_tiles_cap(n_tiles);
        const unsigned K2 = a.split_out ? (unsigned)((nd >> 32) < cap ? (nd >> 32) : cap) : 0u;
        const unsigned D = K2 * kSplitParts + (K - K2);          // deep waves
        const unsigned Kb = ((D + wpb - 1) / wpb + 7u) & ~7u;
        const unsigned b = blockIdx.x;
        int64_t r;
        if (b < Kb) {
            const unsigned rr = b * wpb + w;
            if (rr >= D) return;                       // whole wave: padding of the deep workgroups
            if (rr < K2 * kSplitParts) {
                search_tile<MODE>(a, (int64_t)a.tile_list[rr / kSplitParts] * 64, wstack[w], (int)rr);
                return;
            }
            r = (int64_t)K2 + (rr - K2 * kSplitParts);
        } else {
            const int64_t nn = n_tiles - K;
            const unsigned nbn = (unsigned)((nn + wpb - 1) / wpb);
            if (b - Kb >= nbn) return;                 // whole wave: past the last tile
            r = (int64_t)xcd_block(b - Kb, nbn, a.xcd_chunk) * wpb + w;
            if (r >= nn) return;
            r += K;
        }
        base = (int64_t)a.tile_list[r] * 64;
    } else {
        base = (int64_t)xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk) * BS + (threadIdx.x & ~63);
        if (base >= a.n_slots) return;                // whole wave out of range
    }
    search_tile<MODE>(a, base, wstack[w]);
}

// The parts of each split tile combined, one wave per tile: per lane the
// least key and the least key other than it over the parts' best and second
// keys (a part's two differ, and a triangle is tested by one part only, the
// one owning its leaf), then the node's epilogue.  Every triangle that can be
// the best or a second within the FP64 settle's margin is tested by its part:
// a part prunes with its own best, which is no better than the tile's.
template <int MODE>
__global__ __launch_bounds__(kBS) void k_split_merge(QueryArgs a)
{
    if (halted(a.sc)) return;
    const int64_t n_tiles = a.n_slots / 64;
    const int64_t k = ((int64_t)blockIdx.x * kBS + threadIdx.x) >> 6;
    const int64_t cap = split_tiles_cap(n_tiles), nd2 = (int64_t)(ftd::cload(a.n_deep, 0) >> 32);
    const int64_t K2 = nd2 < cap ? nd2 : cap;
    if (k >= K2) return;                               // whole wave
    const int lane = threadIdx.x & 63;
    const int64_t slot = (int64_t)a.tile_list[k] * 64 + lane;
    const int32_t node = a.perm[slot];
    const ulonglong2 v = a.split_out[k * kSplitParts * 64 + lane];
    uint64_t best = v.x, second = v.y;
    for (int p = 1; p < kSplitParts; p++) {
        const ulonglong2 u = a.split_out[(k * kSplitParts + p) * 64 + lane];
        // explicit compares: min() of two 64-bit keys would round them
        const uint64_t nb = u.x < best ? u.x : best;
        const uint64_t k1 = best != nb ? best : second, k2 = u.x != nb ? u.x : u.y;
        second = k2 < k1 ? k2 : k1;
        best = nb;
    }
    if (node < 0) return;
    Searcher<MODE> S;
    S.x = a.sx[slot]; S.y = a.sy[slot]; S.z = a.sz[slot];
    S.bkey = best;
    S.skey = second;
    S.ifp = a.ifp; S.tris = a.tris;
    finish_node<MODE>(a, S, node);
}

// ----------------------------------------------------------------------------
// Narrow band without the tree (schedule FT_SCHED_BINS, triangles): each wave
// takes its tile's box, widened by the band and by the largest distance from
// a triangle's centroid to its vertices (the primitives are binned by
// centroid), and walks the index bins in that box in their Morton order.  The
// triangles of each non-empty bin are staged in LDS, 64 at a time, by the
// wave's lanes, and every lane then tests each staged triangle (the frame
// rectangle cull against its own bound, then the exact distance).  Each lane
// keeps its own best two, as the tree walk does, so no cross-lane reduction
// is needed.  Every triangle within the band of a lane has its centroid in
// the widened box, and the lane's first bound is the band: the results are
// the tree walk's, bit for bit.
// ----------------------------------------------------------------------------
struct BinQuery {
    int64_t G;                       // bins over all fields (cnt_off[f] + code)
    const uint64_t* prefix;          // (bins with triangles before << 32) | triangles before
                                     // (the index's scan; the scatter consumed the counts)
    const Grid* grids;               // per field
    const int64_t* cnt_off;          // per field: its first bin
};

// Largest distance from a triangle's centroid (as k_prim_bin computes it) to
// its vertices, as float bits (non-negative: integer max orders them).  A
// grid-stride loop and one atomic per workgroup (one per wave contended).
constexpr int kMarginBlocks = 512;
__global__ __launch_bounds__(kBS) void k_tri_margin(const uint64_t* sc, const float* ifp, const int4* tris,
                                                    uint32_t* out)
{
    __shared__ float part[kBS / 64];
    if (halted(sc)) return;
    const int64_t F = cnt_tris(sc);
    float r = 0.0f;
    for (int64_t j = (int64_t)blockIdx.x * kBS + threadIdx.x; j < F; j += (int64_t)gridDim.x * kBS) {
        float a[3], b[3], d[3], c[3];
        tri_vertices(tris, ifp, (int32_t)j, a, b, d);
        for (int k = 0; k < 3; k++) c[k] = (a[k] + b[k] + d[k]) * (1.0f / 3.0f);
        float m = 0.0f;
        const float* v[3] = {a, b, d};
        for (int q = 0; q < 3; q++) {
            const float dx = v[q][0] - c[0], dy = v[q][1] - c[1], dz = v[q][2] - c[2];
            m = fmaxf(m, sqrtf(dx * dx + dy * dy + dz * dz));
        }
        r = fmaxf(r, m);
    }
    for (int o = 32; o >= 1; o >>= 1) r = fmaxf(r, __shfl_xor(r, o));
    if (ftd::lane_id() == 0) part[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = part[0];
        for (int k = 1; k < kBS / 64; k++) m = fmaxf(m, part[k]);
        if (m > 0.0f) atomicMax(out, __float_as_uint(m));
    }
}

// Orders the wave's LDS stores before its loads (a wave runs its lanes
// together: a wave-scope fence waits for its own LDS accesses).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kBinStage = 64;        // triangles staged per pass (one per lane)

template <int BS>
__global__ __launch_bounds__(BS, 8) void k_query_bins(QueryArgs a, BinQuery bq)
{
    __shared__ TriRec srec[BS / 64][kBinStage];
    __shared__ int32_t sid[BS / 64][kBinStage];
    if (halted(a.sc)) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t base = (int64_t)xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk) * BS + (threadIdx.x & ~63);
    if (base >= a.n_slots) return;                    // whole wave out of range
    const int64_t slot = base + lane;
    const int32_t node = a.perm[slot];
    const float px = a.sx[slot], py = a.sy[slot], pz = a.sz[slot];
    const int field = __builtin_amdgcn_readfirstlane(a.tile_field[base >> 6]);
    Searcher<1> S;
    S.x = px; S.y = py; S.z = pz;
    S.init_best(a.band * a.band);
    S.ifp = a.ifp; S.tris = a.tris;
    // the tile's box over its real nodes, widened
    float lo[3] = {node >= 0 ? px : INFINITY, node >= 0 ? py : INFINITY, node >= 0 ? pz : INFINITY};
    float hi[3] = {node >= 0 ? px : -INFINITY, node >= 0 ? py : -INFINITY, node >= 0 ? pz : -INFINITY};
    for (int o = 32; o >= 1; o >>= 1)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
        }
    const Grid g = bq.grids[field];
    const float wide = a.band + __uint_as_float((uint32_t)ftd::cload(a.sc, S_MARGIN));
    int b0[3], n[3];
    for (int k = 0; k < 3; k++) {
        // one more bin each side: the centroid's bin and this range round alike
        // only up to the last ulp
        // (clamped to the grid before the conversion: a wide band overflows int)
        const float dmax = (float)g.dim[k];
        const float f0 = fminf((lo[k] - wide - g.lo[k]) * g.inv_b, dmax);
        const float f1 = fminf((hi[k] + wide - g.lo[k]) * g.inv_b, dmax);
        const int i0 = f0 >= 1.0f ? (int)f0 - 1 : 0;
        int i1 = f1 >= 0.0f ? (int)f1 + 1 : 0;
        i1 = i1 < g.dim[k] - 1 ? i1 : g.dim[k] - 1;
        b0[k] = __builtin_amdgcn_readfirstlane(i0 < g.dim[k] ? i0 : g.dim[k] - 1);
        n[k] = __builtin_amdgcn_readfirstlane(i1 >= b0[k] ? i1 - b0[k] + 1 : 0);
    }
    // The widened range is walked in blocks of 4 x 4 x 4 bins, 64 consecutive
    // Morton codes, whose triangle count is a difference of two prefix entries:
    // a tile far from the surface finds every block empty in one pass.
    const uint32_t off = (uint32_t)bq.cnt_off[field], fend = (uint32_t)bq.cnt_off[field + 1];
    const uint32_t total = (uint32_t)cnt_tris(a.sc);
    auto before = [&](uint32_t x) -> uint32_t { return (int64_t)x < bq.G ? (uint32_t)bq.prefix[x] : total; };
    int B0[3], nB[3];
    for (int k = 0; k < 3; k++) {
        B0[k] = b0[k] >> 2;
        nB[k] = n[k] > 0 ? ((b0[k] + n[k] - 1) >> 2) - B0[k] + 1 : 0;
    }
    const int64_t nblk = (int64_t)nB[0] * nB[1] * nB[2];
    // this lane's place in a block (Morton bits: z0 y0 x0 z1 y1 x1)
    const int lz = (lane & 1) | ((lane >> 2) & 2), ly = ((lane >> 1) & 1) | ((lane >> 3) & 2),
              lx = ((lane >> 2) & 1) | ((lane >> 4) & 2);
    for (int64_t q0 = 0; q0 < nblk; q0 += 64) {
        const int64_t q = q0 + lane;
        uint32_t first = 0, bc = 0;
        int bxq = 0, byq = 0, bzq = 0;
        if (q < nblk) {
            bzq = B0[2] + (int)(q % nB[2]);
            byq = B0[1] + (int)((q / nB[2]) % nB[1]);
            bxq = B0[0] + (int)(q / ((int64_t)nB[2] * nB[1]));
            first = off + (ftd::morton3((uint32_t)bxq, (uint32_t)byq, (uint32_t)bzq) << 6);
            const uint32_t last = first + 64 < fend ? first + 64 : fend;
            bc = first < fend ? before(last) - before(first) : 0u;
        }
        // non-empty blocks nearest the tile first (the nodes' bounds then
        // tighten early and reject more of the later bins): distance from the
        // tile's box to the block's, only to order them
        float bd = INFINITY;
        if (bc > 0) {
            const float e = 4.0f / g.inv_b;
            const int bq3[3] = {bxq, byq, bzq};
            float d2 = 0.0f;
            for (int k = 0; k < 3; k++) {
                const float l0 = g.lo[k] + (float)bq3[k] * e;
                const float gap = fmaxf(fmaxf(l0 - hi[k], lo[k] - (l0 + e)), 0.0f);
                d2 = __fmaf_rn(gap, gap, d2);
            }
            bd = fminf(d2, FLT_MAX);                  // finite for every block in `blocks`
        }
        uint64_t blocks = __ballot(bc > 0);
        // each pass takes out one block (a lane in `blocks` has the least key:
        // finite, against infinity for the others), so at most 64 passes
        for (int pass = 0; pass < 64 && blocks; pass++) {
            // the lane holding the nearest remaining block
            float mv = bd;
            int ml = lane;
            for (int o = 32; o >= 1; o >>= 1) {
                const float ov = __shfl_xor(mv, o);
                const int ol = __shfl_xor(ml, o);
                if (ov < mv || (ov == mv && ol < ml)) { mv = ov; ml = ol; }
            }
            const int sb = __builtin_amdgcn_readfirstlane(ml);
            blocks &= ~(1ull << sb);
            if (lane == sb) bd = INFINITY;
            const uint32_t c = (uint32_t)__shfl((int)first, sb) + (uint32_t)lane;
            const int bx = 4 * __shfl(bxq, sb) + lx, by = 4 * __shfl(byq, sb) + ly, bz = 4 * __shfl(bzq, sb) + lz;
            uint32_t cnt = 0;
            if (bx >= b0[0] && bx < b0[0] + n[0] && by >= b0[1] && by < b0[1] + n[1] && bz >= b0[2] &&
                bz < b0[2] + n[2]) {
                const uint64_t pc = bq.prefix[c];
                cnt = before(c + 1) - (uint32_t)pc;
                if (cnt) {
                    // the bin's leaf box against the tile's box: beyond the band
                    // of every node, skipped (the box gap is a lower bound on
                    // each node's distance to the leaf)
                    const Bound& lb = a.leaf_b[(uint32_t)(pc >> 32)];
                    float d2 = 0.0f;
                    for (int k = 0; k < 3; k++) {
                        const float gap = fmaxf(fmaxf(lb.lo[k] - hi[k], lo[k] - lb.hi[k]), 0.0f);
                        d2 = __fmaf_rn(gap, gap, d2);
                    }
                    if (d2 > prune_bound(a.band * a.band)) cnt = 0;
                }
            }
            uint64_t has = __ballot(cnt > 0);
            while (has) {
                const int src = __ffsll((unsigned long long)has) - 1;
                has &= has - 1;
                const uint32_t cc = (uint32_t)__shfl((int)c, src), nt = (uint32_t)__shfl((int)cnt, src);
                const uint64_t pcc = bq.prefix[cc];
                const uint32_t start = (uint32_t)pcc;
                // each node against the bin's leaf box, with its bound so far
                const Bound& lb = a.leaf_b[(uint32_t)(pcc >> 32)];
                if (__ballot(ftd::box_dist2(px, py, pz, lb.lo, lb.hi) <= S.bnd) == 0ull) continue;
                for (uint32_t s0 = 0; s0 < nt; s0 += kBinStage) {
                    const uint32_t k = nt - s0 < (uint32_t)kBinStage ? nt - s0 : (uint32_t)kBinStage;
                    if ((uint32_t)lane < k) {
                        srec[w][lane] = a.trec[start + s0 + lane];
                        sid[w][lane] = a.pidx[start + s0 + lane];
                    }
                    wave_lds_sync();
                    for (uint32_t t = 0; t < k; t++) S.test_tri(srec[w][t], sid[w][t]);
                    wave_lds_sync();                   // before the next pass overwrites them
                }
            }
        }
    }
    const bool clamped = node >= 0 && S.bj() < 0;
    const uint64_t cm = __ballot(clamped);
    if (lane == 0 && cm) atomicAdd(&a.n_clamped[field], (unsigned long long)__popcll(cm));
    if (node < 0) return;
    finish_node<1>(a, S, node);
}

// ============================================================================
// Level-set normals: Assembler::grad / grad_nodes (Assembler.cpp:11-71), the
// last step of LevelSet::resetPhi_ (LevelSet.cpp:451-453)
// ============================================================================
// Node -> tets lists (Mesh3D::p2t, TetraMesh.cpp:84-86, 115-131): counts,
// scan, scatter, then each list sorted to increasing tet id.
// Node -> tets counts and fills with the references of a workgroup's tets
// gathered first in an LDS table keyed by node (consecutive tets share most
// of their nodes): one memory atomic per distinct node of the workgroup
// instead of one per reference.  A reference that finds no slot within
// kN2tProbe probes goes to memory on its own.  The order inside a node's
// list is arbitrary here; k_n2t_sort makes it the tet order.
constexpr int kN2tHash = 1024;            // LDS slots per workgroup (power of 2)
constexpr int kN2tProbe = 16;
__device__ __forceinline__ int n2t_slot(int32_t* key, int32_t v)
{
    uint32_t h = ((uint32_t)v * 2654435761u) >> (32 - 10);
    for (int p = 0; p < kN2tProbe; p++) {
        const int32_t prev = atomicCAS(&key[h], -1, v);
        if (prev == -1 || prev == v) return (int)h;
        h = (h + 1) & (kN2tHash - 1);
    }
    return -1;
}

__global__ __launch_bounds__(kBS) void k_n2t_count(const int4* tets, int64_t T, uint32_t* cnt)
{
    static_assert(kN2tHash == 1 << 10, "the slot hash takes 10 bits");
    __shared__ int32_t key[kN2tHash];
    __shared__ uint32_t num[kN2tHash];
    for (int k = threadIdx.x; k < kN2tHash; k += kBS) {
        key[k] = -1;
        num[k] = 0u;
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (t < T) {
        const int4 q = tets[t];
        const int32_t v[4] = {q.x, q.y, q.z, q.w};
        for (int j = 0; j < 4; j++) {
            const int sl = n2t_slot(key, v[j]);
            if (sl >= 0) atomicAdd(&num[sl], 1u);
            else atomicAdd(&cnt[v[j]], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kN2tHash; k += kBS)
        if (key[k] >= 0) atomicAdd(&cnt[key[k]], num[k]);
}

__global__ __launch_bounds__(kBS) void k_n2t_fill(const int4* tets, int64_t T, const uint64_t* off,
                                                  uint32_t* cursor, int32_t* ids)
{
    __shared__ int32_t key[kN2tHash];
    __shared__ uint32_t num[kN2tHash];
    __shared__ uint32_t base[kN2tHash];
    for (int k = threadIdx.x; k < kN2tHash; k += kBS) {
        key[k] = -1;
        num[k] = 0u;
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    int32_t v[4] = {-1, -1, -1, -1};
    int sl[4] = {-1, -1, -1, -1};
    if (t < T) {
        const int4 q = tets[t];
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        for (int j = 0; j < 4; j++) {
            sl[j] = n2t_slot(key, v[j]);
            if (sl[j] >= 0) atomicAdd(&num[sl[j]], 1u);
        }
    }
    __syncthreads();
    // one reservation per distinct node of the workgroup
    for (int k = threadIdx.x; k < kN2tHash; k += kBS) {
        if (key[k] >= 0) base[k] = atomicAdd(&cursor[key[k]], num[k]);
        num[k] = 0u;                                    // now the slot's local cursor
    }
    __syncthreads();
    if (t < T) {
        for (int j = 0; j < 4; j++) {
            const uint32_t r = sl[j] >= 0 ? base[sl[j]] + atomicAdd(&num[sl[j]], 1u) : atomicAdd(&cursor[v[j]], 1u);
            ids[off[v[j]] + r] = (int32_t)t;
        }
    }
}

// Each node's list sorted to increasing tet id: lists of up to kN2tReg ids in
// registers by a bitonic network (padding sorts last), longer ones in place by
// insertion.
constexpr int kN2tReg = 32;
__global__ __launch_bounds__(kBS) void k_n2t_sort(const uint64_t* off, int64_t N, int32_t* ids)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const uint64_t b = off[i], e = off[i + 1];
    const int n = (int)(e - b);
    if (n <= kN2tReg) {
        int32_t v[kN2tReg];
#pragma unroll
        for (int a = 0; a < kN2tReg; a++) v[a] = a < n ? ids[b + a] : INT32_MAX;
#pragma unroll
        for (int k = 2; k <= kN2tReg; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (int a = 0; a < kN2tReg; a++) {
                    const int l = a ^ j;
                    if (l > a) {
                        const int32_t x = v[a], y = v[l];
                        const bool sw = (a & k) == 0 ? x > y : x < y;
                        v[a] = sw ? y : x;
                        v[l] = sw ? x : y;
                    }
                }
            }
        }
#pragma unroll
        for (int a = 0; a < kN2tReg; a++)
            if (a < n) ids[b + a] = v[a];
        return;
    }
    for (uint64_t a = b + 1; a < e; a++) {
        const int32_t key = ids[a];
        uint64_t j = a;
        while (j > b && ids[j - 1] > key) { ids[j] = ids[j - 1]; j--; }
        ids[j] = key;
    }
}

// Assembler::grad: per tet, B = [p1-p0, p2-p0, p3-p0] in FP64 from the float
// nodes; its inverse transposed is cofactor(r,c) / det, det = sum of
// cofactor(r,0) B(r,0) as x0 + (x1 + x2) (getElementInfo_ij_,
// Assembler.hpp:429-459); the basis gradients (-1,-1,-1), e1, e2, e3
// (Tetrahedron.hpp:36-50) mapped by it, the first summed left to right; then
// g = ((u0 c0 + u1 c1) + u2 c2) + u3 c3, cast to float.  A singular tet (det
// 0, which the reference only reports) gets a zero gradient.
__global__ __launch_bounds__(kBS) void k_tet_grad(const int4* tets, int64_t T, const float* xyz,
                                                  const double* u, float* grad_t, const uint64_t* sc)
{
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (t >= T || halted(sc)) return;
    const int4 q = tets[t];
    const int qi[4] = {q.x, q.y, q.z, q.w};
    double x[4][3];
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) x[j][k] = (double)xyz[3 * (int64_t)qi[j] + k];
    double B[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) B[r][c] = x[c + 1][r] - x[0][r];
    double C[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            C[r][c] = B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1];
        }
    const double det = C[0][0] * B[0][0] + (C[1][0] * B[1][0] + C[2][0] * B[2][0]);
    float g[3] = {0.0f, 0.0f, 0.0f};
    if (det != 0.0) {
        const double inv = 1.0 / det;
        const double u0 = u[qi[0]], u1 = u[qi[1]], u2 = u[qi[2]], u3 = u[qi[3]];
        for (int r = 0; r < 3; r++) {
            const double b0 = C[r][0] * inv, b1 = C[r][1] * inv, b2 = C[r][2] * inv;
            const double c0 = (b0 * -1.0 + b1 * -1.0) + b2 * -1.0;
            g[r] = (float)(((u0 * c0 + u1 * b0) + u2 * b1) + u3 * b2);
        }
    }
    grad_t[3 * t] = g[0];
    grad_t[3 * t + 1] = g[1];
    grad_t[3 * t + 2] = g[2];
}

// Assembler::grad_nodes: per node, the float mean of its tets' gradients,
// summed in increasing tet id.
constexpr int kNodeGather = 8;

__global__ __launch_bounds__(kBS) void k_node_normals(const uint64_t* off, const int32_t* ids,
                                                      const float* grad_t, int64_t N, float* out,
                                                      const uint64_t* sc)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N || halted(sc)) return;
    const uint64_t b = off[i], e = off[i + 1];
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
    // in list order; the list entries, then their gradients, kNodeGather at a
    // time (loads in flight together instead of one dependent pair per tet)
    for (uint64_t j0 = b; j0 < e; j0 += kNodeGather) {
        int64_t t[kNodeGather];
#pragma unroll
        for (int u = 0; u < kNodeGather; u++) t[u] = j0 + u < e ? (int64_t)ids[j0 + u] : -1;
        float g[kNodeGather][3];
#pragma unroll
        for (int u = 0; u < kNodeGather; u++)
            for (int c = 0; c < 3; c++) g[u][c] = t[u] >= 0 ? grad_t[3 * t[u] + c] : 0.0f;
#pragma unroll
        for (int u = 0; u < kNodeGather; u++)
            if (j0 + u < e) {
                n0 += g[u][0];
                n1 += g[u][1];
                n2 += g[u][2];
            }
    }
    const float cnt = (float)(e - b);
    out[3 * i] = n0 / cnt;
    out[3 * i + 1] = n1 / cnt;
    out[3 * i + 2] = n2 / cnt;
}

// ---- interface curvature (LevelSet::getInterfaceCurvature, LevelSet.cpp:215-253)
// g / |g| per node in float (|g| = sqrt(x^2 + (y^2 + z^2))), kept as three
// double fields (g_x, g_y, g_z, :224-230).
__global__ __launch_bounds__(kBS) void k_unit_split(const float* g, int64_t N, double* gx, double* gy,
                                                    double* gz)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const float x = g[3 * i], y = g[3 * i + 1], z = g[3 * i + 2];
    const float nrm = sqrtf(x * x + (y * y + z * z));
    gx[i] = (double)(x / nrm);
    gy[i] = (double)(y / nrm);
    gz[i] = (double)(z / nrm);
}

// One component of grad_nodes (the float mean over the node's tets of that
// component) added into k: k = mean on the first call, k += mean after, so
// k = gg_x.x + gg_y.y + gg_z.z in that order (:237-238).
// The curvature's gradients (LevelSet.cpp:233-235, the entries :238 uses):
// for c = x, y, z the c-th entry of
// the P1 gradient of g_c, in k_tet_grad's frame and operation order, with one
// frame per tet for the three fields instead of one pass per field.
__global__ __launch_bounds__(kBS) void k_tet_grad_diag(const int4* tets, int64_t T, const float* xyz,
                                                       const double* gx, const double* gy, const double* gz,
                                                       float* out, const uint64_t* sc)
{
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (t >= T || halted(sc)) return;
    const int4 q = tets[t];
    const int qi[4] = {q.x, q.y, q.z, q.w};
    double x[4][3];
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) x[j][k] = (double)xyz[3 * (int64_t)qi[j] + k];
    double B[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) B[r][c] = x[c + 1][r] - x[0][r];
    double C[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            C[r][c] = B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1];
        }
    const double det = C[0][0] * B[0][0] + (C[1][0] * B[1][0] + C[2][0] * B[2][0]);
    float g[3] = {0.0f, 0.0f, 0.0f};
    if (det != 0.0) {
        const double inv = 1.0 / det;
        const double* uf[3] = {gx, gy, gz};
        for (int r = 0; r < 3; r++) {
            const double u0 = uf[r][qi[0]], u1 = uf[r][qi[1]], u2 = uf[r][qi[2]], u3 = uf[r][qi[3]];
            const double b0 = C[r][0] * inv, b1 = C[r][1] * inv, b2 = C[r][2] * inv;
            const double c0 = (b0 * -1.0 + b1 * -1.0) + b2 * -1.0;
            g[r] = (float)(((u0 * c0 + u1 * b0) + u2 * b1) + u3 * b2);
        }
    }
    out[3 * t] = g[0];
    out[3 * t + 1] = g[1];
    out[3 * t + 2] = g[2];
}

// Node curvature (LevelSet.cpp:237-239): per component the float mean over the node's
// tets in increasing id (grad_nodes), then the sum x + y + z in that order.
__global__ __launch_bounds__(kBS) void k_node_div(const uint64_t* off, const int32_t* ids, const float* gd,
                                                  int64_t N, float* k)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const uint64_t b = off[i], e = off[i + 1];
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (uint64_t j0 = b; j0 < e; j0 += kNodeGather) {      // in list order, as k_node_normals
        int64_t t[kNodeGather];
#pragma unroll
        for (int u = 0; u < kNodeGather; u++) t[u] = j0 + u < e ? (int64_t)ids[j0 + u] : -1;
        float g[kNodeGather][3];
#pragma unroll
        for (int u = 0; u < kNodeGather; u++)
            for (int c = 0; c < 3; c++) g[u][c] = t[u] >= 0 ? gd[3 * t[u] + c] : 0.0f;
#pragma unroll
        for (int u = 0; u < kNodeGather; u++)
            if (j0 + u < e) {
                s0 += g[u][0];
                s1 += g[u][1];
                s2 += g[u][2];
            }
    }
    const float cnt = (float)(e - b);
    const float m0 = s0 / cnt, m1 = s1 / cnt, m2 = s2 / cnt;
    k[i] = (m0 + m1) + m2;
}

// Curvature at the crossing points (:243-250): t = float(phi0 / (phi0 - phi1))
// from the blended phi, then (1.0 - t) k0 in double plus t k1 in float, the
// sum in double rounded to float; in crossing-edge order (e2ifp_).
__global__ __launch_bounds__(kBS) void k_curv_points(const int32_t* ifp2e, int64_t P, const int2* edges,
                                                     const double* phi, const float* k, float* k_if)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= P) return;
    const int2 e = edges[ifp2e[i]];
    const float t = (float)(phi[e.x] / (phi[e.x] - phi[e.y]));
    const float tk1 = t * k[e.y];
    k_if[i] = (float)((1.0 - (double)t) * (double)k[e.x] + (double)tk1);
}

// ---- Flow::surfaceTension_ (src/growth/Flow.cpp:307-341) -------------------
// Crossing point i (on edge e = (e0, e1), t = phi[e0] / (phi[e0] - phi[e1])
// as float) adds to node e0 the force (1 - t) st (-k_i) area_i rho n0 and to
// e1 t st (-k_i) area_i rho n1, with n the node's normal over its norm.  The
// reference's loop is an OpenMP loop whose += on a node shared by two points
// races; the result defined here is the serial one, terms added in ascending
// i.  Each node gathers its terms (ids 2i + side, sorted), so the order is
// exact with no atomics on the forces.
__global__ __launch_bounds__(kBS) void k_st_count(const int32_t* ifp2e, int64_t P, const int2* edges,
                                                  uint32_t* cnt)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= P) return;
    const int2 e = edges[ifp2e[i]];
    atomicAdd(&cnt[e.x], 1u);
    atomicAdd(&cnt[e.y], 1u);
}

__global__ __launch_bounds__(kBS) void k_st_fill(const int32_t* ifp2e, int64_t P, const int2* edges,
                                                 const uint64_t* off, uint32_t* cursor, int32_t* ids)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= P) return;
    const int2 e = edges[ifp2e[i]];
    ids[off[e.x] + atomicAdd(&cursor[e.x], 1u)] = (int32_t)(2 * i);
    ids[off[e.y] + atomicAdd(&cursor[e.y], 1u)] = (int32_t)(2 * i + 1);
}

// Types as the reference evaluates them: (1.0 - t) makes the e0 factor a
// double product, cast to float when it scales the float normal; t * st
// keeps the e1 factor in float; fx is double, each term added as a float.
__global__ __launch_bounds__(kBS) void k_st_apply(const uint64_t* off, const int32_t* ids, int64_t N,
                                                  const int32_t* ifp2e, const int2* edges,
                                                  const double* phi, const float* k, const float* area,
                                                  const float* normals, float st, int32_t rho,
                                                  double* fx, double* fy, double* fz)
{
    const int64_t v = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (v >= N) return;
    const uint64_t b = off[v], e = off[v + 1];
    if (b == e) return;
    const float x = normals[3 * v], y = normals[3 * v + 1], z = normals[3 * v + 2];
    const float nrm = sqrtf(x * x + (y * y + z * z));
    const float nx = x / nrm, ny = y / nrm, nz = z / nrm;
    double ax = fx[v], ay = fy[v], az = fz[v];
    for (uint64_t q = b; q < e; q++) {
        const int32_t id = ids[q];
        const int32_t i = id >> 1;
        const int2 ed = edges[ifp2e[i]];
        const float t = (float)(phi[ed.x] / (phi[ed.x] - phi[ed.y]));
        float sc;
        if ((id & 1) == 0)
            sc = (float)((((1.0 - (double)t) * (double)st) * (double)(-k[i])) * (double)area[i] * (double)rho);
        else
            sc = (((t * st) * (-k[i])) * area[i]) * (float)rho;
        ax += (double)(sc * nx);
        ay += (double)(sc * ny);
        az += (double)(sc * nz);
    }
    fx[v] = ax;
    fy[v] = ay;
    fz[v] = az;
}

// ============================================================================
// Advection (LevelSet::update, LevelSet.cpp:89-141): Crank-Nicolson on the P1
// Galerkin form, (M + s A) phi_n = (M - s A) phi with s = 0.5 dt, solved by
// BiCGSTAB with a Jacobi preconditioner.
//
// Matrices are stored in sliced ELL: slices of kFemSlice rows, each slice
// column-major over its width (the longest row in it), so that a wave reads
// k-th entries of 64 rows in one coalesced access.  Rows keep the reference's
// column order (increasing; Assembler::bilin_with_p2t fills a std::map), and
// every row sum runs over it in that order.
// ============================================================================
constexpr int kFemSlice = 64;          // rows per ELL slice (one wave)
constexpr int kFemMaxCols = 48;        // columns per row (the node and its tet neighbours)

struct FemQuad {
    double phi[16];                    // phi_j at quadrature point q: [4 j + q]
    double w;                          // weight (0.25 / 6)
};

struct FemEll {
    const int32_t* len;                // row lengths
    const uint64_t* soff;              // slice offsets in units of kFemSlice entries
    const int32_t* cols;               // ELL columns (padding: the row itself)
};
__device__ __forceinline__ int64_t ell_at(const FemEll& e, int64_t i, int k)
{
    return ((int64_t)e.soff[i / kFemSlice] + k) * kFemSlice + (i % kFemSlice);
}

// getElementInfo_ij_ (Assembler.hpp:429-459) and the mapped basis gradients
// G[j] = invBt grad(phi_j) (the constant gradients (-1,-1,-1), e1, e2, e3,
// Tetrahedron.hpp:36-50; sums left to right, as in k_tet_grad).  False for a
// singular map, whose tet the reference leaves with an unset inverse.
// x: the tet's four nodes, in its local order.
__device__ __forceinline__ bool fem_frame(const double x[4][3], double G[4][3], double& detB)
{
    double B[3][3], C[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) B[r][c] = x[c + 1][r] - x[0][r];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            C[r][c] = B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1];
        }
    const double det = C[0][0] * B[0][0] + (C[1][0] * B[1][0] + C[2][0] * B[2][0]);
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    for (int r = 0; r < 3; r++) {
        const double b0 = C[r][0] * inv, b1 = C[r][1] * inv, b2 = C[r][2] * inv;
        G[0][r] = (b0 * -1.0 + b1 * -1.0) + b2 * -1.0;
        G[1][r] = b0;
        G[2][r] = b1;
        G[3][r] = b2;
    }
    detB = fabs(det);
    return true;
}

// First position in the sorted lst[0, cnt) whose entry is not below v
// (binary search: most of a row's candidates are already in its list).
__device__ __forceinline__ int fem_lower_bound(const int32_t* lst, int cnt, int32_t v)
{
    int lo = 0, hi = cnt;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lst[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The tets a0 + u (u < kFemGather, those below a1) of a row's list, their
// list entries and then the tets loaded together: kFemGather of each load in
// flight per lane instead of one dependent pair at a time (at 256^3
// k_fem_rowfill 19.7 ms with one, 9.6 with four, 8.8 with eight).
constexpr int kFemGather = 8;
__device__ __forceinline__ void fem_gather_tets(const int32_t* n2t_ids, const int4* tets, uint64_t a0,
                                                uint64_t a1, int4 (&q)[kFemGather])
{
    int32_t id[kFemGather];
#pragma unroll
    for (int u = 0; u < kFemGather; u++) id[u] = a0 + u < a1 ? n2t_ids[a0 + u] : -1;
#pragma unroll
    for (int u = 0; u < kFemGather; u++) q[u] = id[u] >= 0 ? tets[id[u]] : make_int4(0, 0, 0, 0);
}

// Sorted unique columns of row i (the node and every node of its tets) into
// lst; returns the count, or -1 past kFemMaxCols.
__device__ int fem_row_cols(const uint64_t* n2t_off, const int32_t* n2t_ids, const int4* tets, int64_t i,
                            int32_t* lst)
{
    int cnt = 0;
    const uint64_t a1 = n2t_off[i + 1];
    for (uint64_t a0 = n2t_off[i]; a0 < a1; a0 += kFemGather) {
        int4 q[kFemGather];
        fem_gather_tets(n2t_ids, tets, a0, a1, q);
        for (int u = 0; u < kFemGather && a0 + u < a1; u++) {
            const int v4[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
            for (int j = 0; j < 4; j++) {
                const int32_t v = v4[j];
                const int pos = fem_lower_bound(lst, cnt, v);
                if (pos < cnt && lst[pos] == v) continue;
                if (cnt == kFemMaxCols) return -1;
                for (int k = cnt; k > pos; k--) lst[k] = lst[k - 1];
                lst[pos] = v;
                cnt++;
            }
        }
    }
    return cnt;
}

// Pass 1 of the pattern: row lengths and slice widths.  One wave per slice.
__global__ __launch_bounds__(kFemSlice) void k_fem_rowlen(const uint64_t* n2t_off, const int32_t* n2t_ids,
                                                          const int4* tets, int64_t N, int32_t* len,
                                                          uint32_t* width, unsigned* flags)
{
    __shared__ int32_t lst[kFemSlice][kFemMaxCols + 1];
    const int64_t i = (int64_t)blockIdx.x * kFemSlice + threadIdx.x;
    int cnt = 0;
    if (i < N) {
        cnt = fem_row_cols(n2t_off, n2t_ids, tets, i, lst[threadIdx.x]);
        if (cnt < 0) { atomicOr(&flags[0], 1u); cnt = 0; }
        len[i] = cnt;
    }
    int wmax = cnt;
    for (int o = 32; o >= 1; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
    if (threadIdx.x == 0) {
        width[blockIdx.x] = (uint32_t)wmax;
        atomicMax(&flags[1], (unsigned)wmax);                // longest row of the mesh
    }
}

// Pass 2: the ELL columns (padding entries name the row itself) and the
// scatter map: for the a-th tet of row i (n2t order), the row positions of
// its four nodes, one byte each, in fmap[n2t_off[i] + a].
__global__ __launch_bounds__(kFemSlice) void k_fem_rowfill(const uint64_t* n2t_off, const int32_t* n2t_ids,
                                                           const int4* tets, int64_t N, FemEll e,
                                                           const uint32_t* width, int32_t* cols,
                                                           uint32_t* fmap)
{
    __shared__ int32_t lst[kFemSlice][kFemMaxCols + 1];
    const int64_t i = (int64_t)blockIdx.x * kFemSlice + threadIdx.x;
    if (i >= N) return;
    int32_t* l = lst[threadIdx.x];
    const int cnt = fem_row_cols(n2t_off, n2t_ids, tets, i, l);
    const int W = (int)width[blockIdx.x];
    for (int k = 0; k < W; k++) cols[ell_at(e, i, k)] = (k < cnt) ? l[k] : (int32_t)i;
    const uint64_t a1 = n2t_off[i + 1];
    for (uint64_t a0 = n2t_off[i]; a0 < a1; a0 += kFemGather) {
        int4 q[kFemGather];
        fem_gather_tets(n2t_ids, tets, a0, a1, q);
        for (int u = 0; u < kFemGather && a0 + u < a1; u++) {
            const int v4[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
            uint32_t m = 0;
            for (int j = 0; j < 4; j++) {
                // every node of the row's tets is in its list (cnt >= 1 here)
                const int k = fem_lower_bound(l, cnt, v4[j]);
                m |= (uint32_t)(k < cnt ? k : cnt - 1) << (8 * j);
            }
            fmap[a0 + u] = m;
        }
    }
}

// Rows of the mass (FORM 0) or advection (FORM 1) form of
// Assembler::bilin_with_p2t (Assembler.hpp:131-226): over the tets of row i
// in increasing id and their local nodes j, the integral
// ((f0 w + f2 w) + (f1 w + f3 w)) |det B| of f(u = phi_j, v = phi_i) is
// added to column t(j):  mass  f = u v;  advection  f = V1 (ux v) + V2 (uy v)
// + V3 (uz v) with V_c(q) = sum_j V_c[t(j)] phi_j(q) accumulated from 0 in
// j order (vel: the velocity interleaved per node).
//
// kFemLanes lanes share a row and take its tets in rounds.  First the row's
// columns are gathered once into LDS: each column node's coordinates (and,
// for the advection form, its velocity), so that a tet's nodes are read
// there through its column map (fmap: the row slot of each of its four
// nodes) instead of through the tet list.  Then in each round lane L
// computes the four contributions of tet kFemLanes r + L into LDS, and the
// row's first lane adds that round's tets into the row accumulator in tet
// order, so every entry is summed exactly as the reference's std::map does.
// A workgroup is one wave of kFemRowsPerWave rows.
constexpr int kFemLanes = 4;
constexpr int kFemRowsPerWave = 64 / kFemLanes;
constexpr int kFemQb = 4 * kFemLanes + 1;            // round buffer per row, padded against bank conflicts
constexpr int kFemQm = kFemLanes + 1;                // round column maps per row, padded likewise

// LDS per wave, in this order: kFemRowsPerWave accumulators of `stride`
// doubles; the round buffers (kFemQb doubles per row); for the advection
// form the column velocities (3 doubles per column slot); the column
// coordinates (3 floats per slot); the round maps (kFemQm words per row);
// the row's own slot (one word per row).  The velocity region is free after
// the rounds: k_fem_cn reuses it.
__host__ __device__ inline size_t fem_lds_bytes(int stride, int form)
{
    const size_t R = kFemRowsPerWave, S = (size_t)stride;
    return 8 * R * (S + kFemQb) + (form == 1 ? 24 * R * S : 0) + 12 * R * S + 4 * R * kFemQm + 4 * R;
}

template <int FORM>
struct FemLds {
    double* acc;       // the row's accumulator
    double* qb;        // the row's round buffer
    double* vel;       // the row's column velocities (FORM 1)
    float* xyz;        // the row's column coordinates
    uint32_t* qm;      // the row's round maps
    int* own;          // the row's own slot
    __device__ FemLds(double* lds, int stride, int rl)
    {
        const size_t R = kFemRowsPerWave, S = (size_t)stride;
        acc = lds + rl * S;
        qb = lds + R * S + rl * kFemQb;
        double* v0 = lds + R * (S + kFemQb);
        vel = v0 + rl * 3 * S;
        float* x0 = reinterpret_cast<float*>(v0 + (FORM == 1 ? 3 * R * S : 0));
        xyz = x0 + rl * 3 * S;
        uint32_t* m0 = reinterpret_cast<uint32_t*>(x0 + 3 * R * S);
        qm = m0 + rl * kFemQm;
        own = reinterpret_cast<int*>(m0 + R * kFemQm) + rl;
    }
};

template <int FORM>
__device__ void fem_rows(const uint64_t* __restrict__ n2t_off, const uint32_t* __restrict__ fmap,
                         const float* __restrict__ xyz, const double* __restrict__ vel, int64_t N,
                         const FemQuad& qd, const FemEll& e, int stride, double* lds, int64_t& row, double*& acc,
                         int& len)
{
    const int lane = threadIdx.x % kFemLanes, rl = threadIdx.x / kFemLanes;
    const int64_t i = (int64_t)blockIdx.x * kFemRowsPerWave + rl;
    const FemLds<FORM> L(lds, stride, rl);
    acc = L.acc;
    row = i;
    len = 0;
    uint64_t a0 = 0;
    int na = 0;
    if (i < N) {
        a0 = n2t_off[i];
        na = (int)(n2t_off[i + 1] - a0);
        len = e.len[i];
        for (int k = lane; k < len; k += kFemLanes) {
            const int32_t c = e.cols[ell_at(e, i, k)];
            for (int d = 0; d < 3; d++) L.xyz[3 * k + d] = xyz[3 * (int64_t)c + d];
            if (FORM == 1)
                for (int d = 0; d < 3; d++) L.vel[3 * k + d] = vel[3 * (int64_t)c + d];
            if (c == (int32_t)i) *L.own = k;
            acc[k] = 0.0;
        }
    }
    int na_max = na;                                   // rounds are uniform over the wave
    for (int o = 32; o >= 1; o >>= 1) na_max = max(na_max, __shfl_xor(na_max, o));
    uint32_t mn = lane < na ? fmap[a0 + lane] : 0u;    // maps one round ahead
    __syncthreads();
    const int own = i < N ? *L.own : 0;
    for (int r0 = 0; r0 < na_max; r0 += kFemLanes) {
        const int a = r0 + lane;
        const uint32_t mp = mn;
        if (a + kFemLanes < na) mn = fmap[a0 + a + kFemLanes];
        if (a < na) {
            L.qm[lane] = mp;
            int sl[4];
            double x[4][3];
            for (int j = 0; j < 4; j++) {
                sl[j] = (int)((mp >> (8 * j)) & 0xffu);
                for (int d = 0; d < 3; d++) x[j][d] = (double)L.xyz[3 * sl[j] + d];
            }
            double G[4][3], detB;
            if (!fem_frame(x, G, detB)) {
                for (int j = 0; j < 4; j++) L.qb[4 * lane + j] = 0.0;   // no contribution
            } else {
                const int li = sl[0] == own ? 0 : sl[1] == own ? 1 : sl[2] == own ? 2 : 3;
                // quadrature points in the order 0, 2, 1, 3: s02 = f0 w + f2 w and
                // s13 = f1 w + f3 w per basis function, then (s02 + s13) |det B|
                const double w = qd.w;
                double s02[4], s13[4];
                for (int qq = 0; qq < 4; qq++) {
                    const int q4 = qq == 1 ? 2 : qq == 2 ? 1 : qq;
                    const double v = qd.phi[4 * li + q4];
                    double V0 = 0.0, V1 = 0.0, V2 = 0.0;
                    if (FORM == 1) {
                        for (int j = 0; j < 4; j++) {
                            const double* vj = L.vel + 3 * sl[j];
                            const double pj = qd.phi[4 * j + q4];
                            V0 = V0 + vj[0] * pj;
                            V1 = V1 + vj[1] * pj;
                            V2 = V2 + vj[2] * pj;
                        }
                    }
                    for (int j = 0; j < 4; j++) {
                        const double f = FORM == 0 ? qd.phi[4 * j + q4] * v
                                                   : (V0 * (G[j][0] * v) + V1 * (G[j][1] * v)) + V2 * (G[j][2] * v);
                        const double pw = f * w;
                        if (qq == 0) s02[j] = pw;
                        else if (qq == 1) s02[j] = s02[j] + pw;
                        else if (qq == 2) s13[j] = pw;
                        else s13[j] = s13[j] + pw;
                    }
                }
                for (int j = 0; j < 4; j++) L.qb[4 * lane + j] = (s02[j] + s13[j]) * detB;
            }
        }
        __syncthreads();
        if (lane == 0) {
            // a tet's four nodes are four distinct columns: its four adds are
            // independent, so they are issued together; tets stay in order
            const int nt = min(kFemLanes, na - r0);
            for (int t = 0; t < nt; t++) {
                const uint32_t pos = L.qm[t];
                int c[4];
                double s4[4];
                for (int j = 0; j < 4; j++) {
                    c[j] = (int)((pos >> (8 * j)) & 0xffu);
                    s4[j] = acc[c[j]] + L.qb[4 * t + j];
                }
                for (int j = 0; j < 4; j++) acc[c[j]] = s4[j];
            }
        }
        __syncthreads();
    }
}

// The velocity interleaved per node (vx, vy, vz), so that a tet's gathers
// touch one 24-byte run per node instead of three arrays.
__global__ __launch_bounds__(kBS) void k_vel_interleave(const double* flow, int64_t N, double* vel3)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    vel3[3 * i] = flow[i];
    vel3[3 * i + 1] = flow[N + i];
    vel3[3 * i + 2] = flow[2 * N + i];
}

// The mass matrix, once per mesh (LevelSet::init, LevelSet.cpp:35-39).
__global__ __launch_bounds__(64) void k_fem_mass(const uint64_t* n2t_off, const uint32_t* fmap,
                                                 const float* xyz, int64_t N, FemQuad qd, FemEll e, int stride, double* mvals)
{
    extern __shared__ double fem_lds[];
    int64_t i;
    double* acc;
    int len;
    fem_rows<0>(n2t_off, fmap, xyz, nullptr, N, qd, e, stride, fem_lds, i, acc, len);
    for (int k = threadIdx.x % kFemLanes; k < len; k += kFemLanes) mvals[ell_at(e, i, k)] = acc[k];
}

// The advection matrix alone (ft_fem_assemble), into vals in ELL layout.
__global__ __launch_bounds__(64) void k_fem_adv(const uint64_t* n2t_off, const uint32_t* fmap,
                                                const float* xyz, const double* vel, int64_t N, FemQuad qd, FemEll e, int stride,
                                                double* vals)
{
    extern __shared__ double fem_lds[];
    int64_t i;
    double* acc;
    int len;
    fem_rows<1>(n2t_off, fmap, xyz, vel, N, qd, e, stride, fem_lds, i, acc, len);
    for (int k = threadIdx.x % kFemLanes; k < len; k += kFemLanes) vals[ell_at(e, i, k)] = acc[k];
}

// ELL rows to CSR at the row offsets off (a scan of the lengths).
__global__ __launch_bounds__(kBS) void k_ell_to_csr(int64_t N, FemEll e, const double* vals, const uint64_t* off,
                                                    int32_t* cols, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i >= N) return;
    const int len = e.len[i];
    for (int k = 0; k < len; k++) {
        const int64_t at = ell_at(e, i, k);
        cols[off[i] + k] = e.cols[at];
        out[off[i] + k] = vals[at];
    }
}

// Deterministic block sum of up to two values per thread into part[2 b], part[2 b + 1].
template <int BS>
__device__ __forceinline__ void block_sum2(double a, double b, double* part)
{
    __shared__ double red[2][BS / 64];
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = a; red[1][threadIdx.x >> 6] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < BS / 64; w++) { sa += red[0][w]; sb += red[1][w]; }
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = sb;
    }
}

// Per call: the advection row, then L = M + s A and R = M - s A entrywise
// (LevelSet.cpp:126-127), invdiag = 1 / L_ii (1 where 0: Eigen's
// DiagonalPreconditioner), b = R phi and r = b - L phi (BiCGSTAB's first
// residual from the guess phi), each summed in column order; r0 = r, x = phi,
// p = v = 0.
// At least 5 waves per SIMD: 94 VGPRs, no spills (measured faster than 118 at 4).
__global__ __launch_bounds__(64, 3) void k_fem_cn(const uint64_t* __restrict__ n2t_off,
                                               const uint32_t* __restrict__ fmap, const float* __restrict__ xyz,
                                               const double* __restrict__ vel, int64_t N, FemQuad qd, FemEll e,
                                               int stride, const double* __restrict__ mvals, double s,
                                               const double* __restrict__ phi, double* __restrict__ lvals,
                                               double* __restrict__ invdiag, double* __restrict__ b,
                                               double* __restrict__ x, double* __restrict__ r,
                                               double* __restrict__ r0, double* __restrict__ p,
                                               double* __restrict__ v)
{
    extern __shared__ double fem_lds[];
    int64_t i;
    double* A;
    int len;
    fem_rows<1>(n2t_off, fmap, xyz, vel, N, qd, e, stride, fem_lds, i, A, len);
    // the row's entries split over its lanes: L = M + s A into lvals, the
    // products R phi_c (in place of A) and L phi_c into LDS; then the first
    // lane sums both in column order, as a sequential row loop would
    const int lane = threadIdx.x % kFemLanes;
    double* LP = FemLds<1>(fem_lds, stride, threadIdx.x / kFemLanes).vel;   // free after the rounds
    double id = 0.0;
    int found = 0;
    for (int k = lane; k < len; k += kFemLanes) {
        const int64_t at = ell_at(e, i, k);
        const int32_t c = e.cols[at];
        const double m = mvals[at], sa = s * A[k];
        const double L = m + sa, R = m - sa;
        const double pc = phi[c];
        lvals[at] = L;
        A[k] = R * pc;
        LP[k] = L * pc;
        if (c == (int32_t)i && L != 0.0) {
            id = 1.0 / L;
            found = 1;
        }
    }
    for (int o = 1; o < kFemLanes; o <<= 1) {          // the diagonal's lane to all
        const double vo = __shfl_xor(id, o);
        const int fo = __shfl_xor(found, o);
        if (!found && fo) {
            id = vo;
            found = 1;
        }
    }
    __syncthreads();
    if (i < N && lane == 0) {
        double bi = 0.0, lx = 0.0;
        for (int k = 0; k < len; k++) {
            bi += A[k];
            lx += LP[k];
        }
        const double ri = bi - lx;
        invdiag[i] = found ? id : 1.0;
        b[i] = bi;
        x[i] = phi[i];
        r[i] = ri;
        r0[i] = ri;
        p[i] = 0.0;
        v[i] = 0.0;
    }
}

// ---- BiCGSTAB (Eigen's bicgstab; H. A. van der Vorst 1992) ----------------
// Scalars live on the device; each finishing kernel reduces the block
// partials (in block order) and updates them.
struct BcgScalars {
    double rho, rho_old, alpha, w;
    double rr, bb, r0sq;
    double r0v, tt, ts;
};

constexpr int kBcgBS = 256;
constexpr int kBcgGrid = 2048;         // workgroups of the solver kernels (block partials)

// Grid-stride loops: row i is handled by workgroup (i / kBcgBS) % gridDim.x;
// each thread sums its rows in order, so the partials are fixed by N and the
// grid size alone.
#define BCG_ROWS(i) for (int64_t i = (int64_t)blockIdx.x * kBcgBS + threadIdx.x; i < N; \
                         i += (int64_t)gridDim.x * kBcgBS)

// The preconditioned vectors y = P p and z = P s (P = invdiag) are formed
// once per row and stored, so that each SpMV entry gathers one value instead
// of forming them from two or three gathers; each is the same product as
// when formed where read.

// p = r + beta (p - w v), beta = (rho / rho_old) (alpha / w); y = invdiag p.
__global__ __launch_bounds__(kBcgBS) void k_bcg_p(int64_t N, const BcgScalars* sc, const double* r,
                                                  const double* v, const double* invdiag, double* p, double* y)
{
    const double beta = (sc->rho / sc->rho_old) * (sc->alpha / sc->w), w = sc->w;
    BCG_ROWS(i) {
        const double pi = r[i] + beta * (p[i] - w * v[i]);
        p[i] = pi;
        y[i] = invdiag[i] * pi;
    }
}

// v = L y (y = P p, then v = A y); partial r0.v.
__global__ __launch_bounds__(kBcgBS) void k_bcg_v(int64_t N, FemEll e, const double* lvals,
                                                  const double* y, const double* r0, double* v, double* part)
{
    double d = 0.0;
    BCG_ROWS(i) {
        const int len = e.len[i];
        double t = 0.0;
        for (int k = 0; k < len; k++) {
            const int64_t at = ell_at(e, i, k);
            t += lvals[at] * y[e.cols[at]];
        }
        v[i] = t;
        d += r0[i] * t;
    }
    block_sum2<kBcgBS>(d, 0.0, part);
}

// s = r - alpha v; z = invdiag s.
__global__ __launch_bounds__(kBcgBS) void k_bcg_s(int64_t N, const BcgScalars* sc, const double* invdiag,
                                                  const double* r, const double* v, double* s_, double* z)
{
    const double alpha = sc->alpha;
    BCG_ROWS(i) {
        const double si = r[i] - alpha * v[i];
        s_[i] = si;
        z[i] = invdiag[i] * si;
    }
}

// t = L z; partials t.t and t.s.
__global__ __launch_bounds__(kBcgBS) void k_bcg_t(int64_t N, FemEll e, const double* lvals,
                                                  const double* z, const double* s_, double* t, double* part)
{
    double tt = 0.0, ts = 0.0;
    BCG_ROWS(i) {
        const int len = e.len[i];
        double acc = 0.0;
        for (int k = 0; k < len; k++) {
            const int64_t at = ell_at(e, i, k);
            acc += lvals[at] * z[e.cols[at]];
        }
        t[i] = acc;
        tt += acc * acc;
        ts += acc * s_[i];
    }
    block_sum2<kBcgBS>(tt, ts, part);
}

// x += alpha y + w z;  r = s - w t;  partials r.r and r0.r.
__global__ __launch_bounds__(kBcgBS) void k_bcg_x(int64_t N, const double* y, const double* s_,
                                                  const double* z, const double* t, const double* r0,
                                                  const BcgScalars* sc, double* x, double* r, double* part)
{
    const double alpha = sc->alpha, w = sc->w;
    double rr = 0.0, r0r = 0.0;
    BCG_ROWS(i) {
        x[i] = x[i] + (alpha * y[i] + w * z[i]);
        const double ri = s_[i] - w * t[i];
        r[i] = ri;
        rr += ri * ri;
        r0r += r0[i] * ri;
    }
    block_sum2<kBcgBS>(rr, r0r, part);
}

// Restart (a residual too orthogonal to r0): r = b - L x, r0 = r; partial r.r.
__global__ __launch_bounds__(kBcgBS) void k_bcg_restart(int64_t N, FemEll e, const double* lvals,
                                                        const double* b, const double* x, double* r,
                                                        double* r0, double* part)
{
    double rr = 0.0;
    BCG_ROWS(i) {
        const int len = e.len[i];
        double t = 0.0;
        for (int k = 0; k < len; k++) {
            const int64_t at = ell_at(e, i, k);
            t += lvals[at] * x[e.cols[at]];
        }
        const double ri = b[i] - t;
        r[i] = ri;
        r0[i] = ri;
        rr += ri * ri;
    }
    block_sum2<kBcgBS>(rr, 0.0, part);
}

// Partials of a.b and c.d.
__global__ __launch_bounds__(kBcgBS) void k_dots2(int64_t N, const double* a, const double* b,
                                                  const double* c, const double* d, double* part)
{
    double ab = 0.0, cd = 0.0;
    BCG_ROWS(i) {
        ab += a[i] * b[i];
        cd += c[i] * d[i];
    }
    block_sum2<kBcgBS>(ab, cd, part);
}

enum BcgStep : int { BCG_INIT = 0, BCG_ALPHA = 1, BCG_W = 2, BCG_END = 3, BCG_RESTART = 4, BCG_SUMS = 5 };

// Sums nb block partial pairs in block order (one workgroup: thread k sums a
// strided subset in order, then the 256 results in order) and updates the
// scalars for the step.
__global__ __launch_bounds__(256) void k_bcg_fin(const double* part, int64_t nb, int step, BcgScalars* sc)
{
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int64_t k = threadIdx.x; k < nb; k += 256) { a += part[2 * k]; b += part[2 * k + 1]; }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double A = 0.0, B = 0.0;
    for (int k = 0; k < 256; k++) { A += red[0][k]; B += red[1][k]; }
    switch (step) {
    case BCG_INIT:                       // r.r (= r0.r0 = r0.r), b.b; rho = alpha = w = 1
        sc->rr = A; sc->r0sq = A; sc->bb = B;
        sc->alpha = 1.0; sc->w = 1.0;
        sc->rho_old = 1.0; sc->rho = A;  // the first iteration's rho_old = rho, rho = r0.r
        break;
    case BCG_ALPHA:
        sc->r0v = A;
        sc->alpha = sc->rho / A;
        break;
    case BCG_W:
        sc->tt = A; sc->ts = B;
        sc->w = A > 0.0 ? B / A : 0.0;
        break;
    case BCG_END:                        // r.r, and the next iteration's rho_old, rho
        sc->rr = A;
        sc->rho_old = sc->rho;
        sc->rho = B;
        break;
    case BCG_RESTART:                    // rho = |r0|^2 = |r|^2 (rho_old kept)
        sc->rr = A; sc->r0sq = A; sc->rho = A;
        break;
    case BCG_SUMS:                       // the two sums as they are
        sc->tt = A; sc->ts = B;
        break;
    }
}

// |flow|^2 over 4 N entries and whether the velocity (first 3 N) has a NaN.
__global__ __launch_bounds__(kBcgBS) void k_flow_check(const double* flow, int64_t N, double* part)
{
    double sq = 0.0, nan = 0.0;
    BCG_ROWS(i) {
        for (int c = 0; c < 4; c++) {
            const double f = flow[c * N + i];
            sq += f * f;
            if (c < 3 && f != f) nan = 1.0;
        }
    }
    block_sum2<kBcgBS>(sq, nan, part);
}

// ============================================================================
// Per-field interface counts (batches): ranks at the field boundaries
// ============================================================================
__device__ __forceinline__ uint64_t low_mask(int b) { return b ? ((1ull << b) - 1ull) : 0ull; }

// crossing tets before tet t (t <= T): block prefix + words of the block
__device__ inline uint32_t tet_rank(int64_t t, int64_t T, const uint64_t* tbits,
                                    const uint64_t* blk_prefix, uint32_t X)
{
    if (t >= T) return X;
    uint32_t r = (uint32_t)(blk_prefix[t >> 8] >> 32);
    for (int64_t w = (t >> 8) << 2; w < (t >> 6); w++) r += __popcll(tbits[w]);
    return r + __popcll(tbits[t >> 6] & low_mask((int)(t & 63)));
}

__device__ inline uint32_t edge_rank_at(int64_t e, int64_t E, const uint64_t* ebits,
                                        const uint32_t* eprefix, uint32_t P)
{
    if (e >= E) return P;
    return eprefix[e >> 6] + (uint32_t)__popcll(ebits[e >> 6] & low_mask((int)(e & 63)));
}

// out[3f..3f+2] = (crossing points, triangles, crossing tets) of field f
__global__ void k_field_counts(int nf, const int64_t* tet_off, const int64_t* edge_off, int64_t T,
                               int64_t E, const uint64_t* tbits, const uint64_t* blk_prefix,
                               const int32_t* xtri_off, const uint64_t* ebits,
                               const uint32_t* eprefix, const uint64_t* sc, int64_t* out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf || halted(sc)) return;
    const uint32_t X = (uint32_t)cnt_xtets(sc), F = (uint32_t)cnt_tris(sc), P = (uint32_t)cnt_points(sc);
    const uint32_t x0 = tet_rank(tet_off[f], T, tbits, blk_prefix, X);
    const uint32_t x1 = tet_rank(tet_off[f + 1], T, tbits, blk_prefix, X);
    const uint32_t f0 = x0 < X ? (uint32_t)xtri_off[x0] : F;
    const uint32_t f1 = x1 < X ? (uint32_t)xtri_off[x1] : F;
    const uint32_t p0 = edge_rank_at(edge_off[f], E, ebits, eprefix, P);
    const uint32_t p1 = edge_rank_at(edge_off[f + 1], E, ebits, eprefix, P);
    out[3 * f + 0] = p1 - p0;
    out[3 * f + 1] = f1 - f0;
    out[3 * f + 2] = x1 - x0;
}

// ============================================================================
// Kuhn grids generated on the device, numbered exactly as ft_kuhn_mesh
// ============================================================================
struct KuhnField {
    int nx, ny, nz, pad;
    double origin[3], spacing[3];
    double jitter;
    uint64_t seed;
    int64_t node_off, tet_off, edge_off;
};

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__constant__ int c_kdir[7][3] = {{1,0,0},{0,1,0},{0,0,1},{1,1,0},{1,0,1},{0,1,1},{1,1,1}};
__constant__ int c_kperm[6][3] = {{0,1,2},{0,2,1},{1,0,2},{1,2,0},{2,0,1},{2,1,0}};

// Edges numbered by low node (x-major), then direction x,y,z,xy,xz,yz,xyz; a
// node has (1+I)(1+J)(1+K)-1 of them (I = not on the upper x face, ...), so the
// count before node (i,j,k) factorises over the axes.
__device__ inline int64_t kuhn_edges_before(int64_t i, int64_t j, int64_t k, int64_t X, int64_t Y,
                                            int64_t Z)
{
    auto S = [](int64_t q, int64_t Q) { return q + (q < Q - 1 ? q : Q - 1); };
    const int64_t I = i < X - 1, J = j < Y - 1;
    return (S(i, X) * S(Y, Y) * S(Z, Z) - i * Y * Z) + ((1 + I) * S(j, Y) * S(Z, Z) - j * Z) +
           ((1 + I) * (1 + J) * S(k, Z) - k);
}

__device__ inline bool kuhn_valid(int64_t i, int64_t j, int64_t k, int d, int64_t X, int64_t Y,
                                  int64_t Z)
{
    return (c_kdir[d][0] == 0 || i < X - 1) && (c_kdir[d][1] == 0 || j < Y - 1) &&
           (c_kdir[d][2] == 0 || k < Z - 1);
}

__device__ inline int32_t kuhn_edge_id(int64_t i, int64_t j, int64_t k, int d, int64_t X, int64_t Y,
                                       int64_t Z)
{
    int r = 0;
    for (int q = 0; q < d; q++) r += kuhn_valid(i, j, k, q, X, Y, Z);
    return (int32_t)(kuhn_edges_before(i, j, k, X, Y, Z) + r);
}

__global__ __launch_bounds__(kBS) void k_kuhn_nodes(KuhnField f, float* xyz)
{
    const int64_t X = f.nx, Y = f.ny, Z = f.nz;
    const int64_t a = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (a >= X * Y * Z) return;
    const int64_t i = a / (Y * Z), j = (a / Z) % Y, k = a % Z;
    const bool interior = i > 0 && i < X - 1 && j > 0 && j < Y - 1 && k > 0 && k < Z - 1;
    const int64_t ijk[3] = {i, j, k};
    float* out = xyz + 3 * (f.node_off + a);
    for (int c = 0; c < 3; c++) {
        double g = (double)ijk[c];
        if (f.jitter > 0.0 && interior) {
            const uint64_t z = mix64(f.seed + (3 * (uint64_t)a + c + 1) * 0x9E3779B97F4A7C15ULL);
            const double u = (double)(z >> 11) * 0x1.0p-53;
            g += (2.0 * u - 1.0) * f.jitter;
        }
        out[c] = (float)(f.origin[c] + g * f.spacing[c]);
    }
}

__global__ __launch_bounds__(kBS) void k_kuhn_edges(KuhnField f, int2* edges)
{
    const int64_t X = f.nx, Y = f.ny, Z = f.nz;
    const int64_t a = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (a >= X * Y * Z) return;
    const int64_t i = a / (Y * Z), j = (a / Z) % Y, k = a % Z;
    int64_t e = f.edge_off + kuhn_edges_before(i, j, k, X, Y, Z);
    for (int d = 0; d < 7; d++) {
        if (!kuhn_valid(i, j, k, d, X, Y, Z)) continue;
        const int64_t b = ((i + c_kdir[d][0]) * Y + (j + c_kdir[d][1])) * Z + (k + c_kdir[d][2]);
        edges[e++] = make_int2((int32_t)(f.node_off + a), (int32_t)(f.node_off + b));
    }
}

__global__ __launch_bounds__(kBS) void k_kuhn_tets(KuhnField f, int4* tets, int32_t* t2e)
{
    const int64_t X = f.nx, Y = f.ny, Z = f.nz;
    const int64_t CX = X - 1, CY = Y - 1, CZ = Z - 1;
    const int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (t >= 6 * CX * CY * CZ) return;
    const int64_t cell = t / 6;
    const int p = (int)(t % 6);
    int64_t v[4][3];
    v[0][0] = cell / (CY * CZ); v[0][1] = (cell / CZ) % CY; v[0][2] = cell % CZ;
    for (int q = 1; q <= 3; q++) {
        for (int c = 0; c < 3; c++) v[q][c] = v[q - 1][c];
        v[q][c_kperm[p][q - 1]] += 1;
    }
    int32_t nd[4];
    for (int q = 0; q < 4; q++) nd[q] = (int32_t)(f.node_off + (v[q][0] * Y + v[q][1]) * Z + v[q][2]);
    tets[f.tet_off + t] = make_int4(nd[0], nd[1], nd[2], nd[3]);
    const int loc[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
    for (int q = 0; q < 6; q++) {
        const int64_t* lo = v[loc[q][0]];
        const int64_t* hi = v[loc[q][1]];
        const int dx = (int)(hi[0] - lo[0]), dy = (int)(hi[1] - lo[1]), dz = (int)(hi[2] - lo[2]);
        int d = 0;
        for (int r = 0; r < 7; r++)
            if (c_kdir[r][0] == dx && c_kdir[r][1] == dy && c_kdir[r][2] == dz) d = r;
        t2e[6 * (f.tet_off + t) + q] = (int32_t)(f.edge_off + kuhn_edge_id(lo[0], lo[1], lo[2], d, X, Y, Z));
    }
}

// Localised copies of index arrays (local -> global): mesh uploads and batch
// assembly from resident meshes.  Negative entries (slot padding) stay.
__global__ __launch_bounds__(kBS) void k_add_offset(const int32_t* src, int64_t n, int64_t add,
                                                    int32_t* dst)
{
    const int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x;
    if (i < n) dst[i] = (int32_t)(src[i] + add);
}

// Batch assembly: the concatenation of per-field pieces into one array in
// one launch.  Piece p fills dst[off[p], off[p+1]) from src[p] (element
// e - off[p]) plus add[p], or with the constant add[p] when src[p] is null;
// with keep_neg, negative source values (padding slots) are kept as they
// are.  A block copies kGatherChunk consecutive elements; it finds its first
// piece by binary search and its threads step to the next piece as they
// cross a boundary.  32-bit words: float arrays copy bit for bit (add 0).
constexpr int kGatherChunk = kBS * 8;
__global__ __launch_bounds__(kBS) void k_gather(const int32_t* const* src, const int64_t* off,
                                                const int64_t* add, int np, int keep_neg, int32_t* dst)
{
    __shared__ int p0;
    const int64_t total = off[np];
    const int64_t c0 = (int64_t)blockIdx.x * kGatherChunk;
    if (c0 >= total) return;
    if (threadIdx.x == 0) p0 = ftd::field_of(off, np, c0);
    __syncthreads();
    int p = p0;
    for (int j = 0; j < kGatherChunk / kBS; j++) {
        const int64_t e = c0 + (int64_t)j * kBS + threadIdx.x;
        if (e >= total) break;
        while (e >= off[p + 1]) p++;
        const int32_t* sp = src[p];
        if (!sp) { dst[e] = (int32_t)add[p]; continue; }
        const int32_t v = sp[e - off[p]];
        dst[e] = (keep_neg && v < 0) ? v : (int32_t)(v + add[p]);
    }
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

struct FieldHost {
    float lo[3], hi[3];
    float h;                         // mean node spacing estimate
};

struct Caps {
    int64_t X = 0;   // crossing tets
    int64_t F = 0;   // triangles
    int64_t P = 0;   // crossing points
};

// A captured reinit and what it was captured for.
struct GraphState {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double* phi = nullptr;
    double* phi_ifn = nullptr;
    ft_opts opts{};
    uint64_t layout = 0;
    Caps caps;
    std::vector<void*> bufs;         // device buffers the graph's kernels address
    int64_t captures = 0, launches = 0;
};

struct ft_ctx {
    int device = 0;
    int cu_count = 256;              // compute units of the device
    hipStream_t stream = nullptr;
    // A second stream for the stages the query does not read (the interface's
    // point-to-triangle lists and areas) and the query's tile schedule: they
    // run beside the index build.  side_fork/side_join order it against the
    // calling stream; side_open: forked and not joined yet.
    hipStream_t side = nullptr;
    hipEvent_t side_fork_ev = nullptr, side_join_ev = nullptr;
    bool side_open = false;
    // fields and mesh (field k: nodes [node_off[k], node_off[k+1]), ...)
    int nf = 0;
    std::vector<int64_t> node_off, tet_off, edge_off, slot_off;
    std::vector<FieldHost> fh;
    int64_t N = 0, T = 0, E = 0, n_slots = 0;
    uint64_t layout_gen = 0;         // bumped whenever the fields change
    bool has_conn = false;
    float h_est = 0.0f;
    DevBuf xyz, tets, edges, t2e, perm, sx, sy, sz, tile_field;
    DevBuf tile_sched;               // per query tile: class, launch order, class scan
    DevBuf deep_part;                // the parts' keys of the split tiles
    DevBuf gather_tab;               // piece tables of the batch assembly (ft_set_batch_meshes)
    // A batch of resident meshes reads their tets, t2e and edges in place
    // (virt): field_tab holds the per-field pointers, field_meshes keeps the
    // meshes alive while the batch refers to them.  Operations that need the
    // concatenated arrays copy them once (materialize_conn).
    bool virt = false;
    DevBuf field_tab;
    std::vector<std::shared_ptr<ft_ctx>> field_meshes;
    DevBuf d_node_off, d_tet_off, d_edge_off, d_slot_off;
    // per reinit
    DevBuf phi_tmp, phi_ifn_tmp;
    DevBuf sbits, retain, tbits, blk, xtet, xtri_off, xedges, ebits, eprefix;
    DevBuf ifp, ifp2e, tris, tri_norm, if_elems, p2t_off, p2t_cursor, p2t, ifp_area;
    DevBuf nearest, fcounts, fclamp;
    // normals: node -> tets lists (per field layout), tet gradients, node normals
    DevBuf n2t_off, n2t_ids, grad_t, normals, phi_ifn_int;
    DevBuf gsplit, kmesh, k_if;      // curvature: unit-normal components, node and point curvature
    DevBuf st_cnt, st_off, st_ids;   // surface tension: node -> crossing-point terms
    DevBuf st_f;                     // surface tension: device copy of host force arrays
    // advection: sliced-ELL pattern and mass matrix (per layout), the
    // Crank-Nicolson matrix, BiCGSTAB vectors, block partials and scalars
    DevBuf fem_len, fem_w, fem_soff, fem_cols, fem_m, fem_l, fem_flag, fem_map;
    DevBuf bcg_vec, bcg_part, bcg_sc, adv_phi, adv_flow, adv_vel3, fem_csr_off;
    uint64_t fem_layout = ~0ull;
    int64_t fem_nnz = 0;             // ELL entries with padding
    int fem_stride = 1;              // LDS doubles per row accumulator (longest row, odd)
    void* h_bcg = nullptr;           // pinned mirror of the solver scalars
    const double* last_phi = nullptr;   // device phi of the last reinit (blended)
    uint64_t n2t_layout = ~0ull;
    bool have_normals = false;
    // index
    DevBuf bin_count, bin_prefix, pcode, prec, trec, pidx, leaf_key, leaf_start, leaf_cnt;
    DevBuf leaf_b, leaf_parent, bnodes, bflags, stats;
    DevBuf d_grids, d_cnt_off, leaf_off, roots;
    uint64_t idx_layout = ~0ull;     // layout and bin scale the grids were made for
    float idx_bin_scale = 0.0f;
    int64_t idx_G = 0;               // bins over all fields
    int idx_code_bits = 0;
    GraphState graph;
    // capacities for the next eager call on the same fields and mode
    Caps spec_caps;
    uint64_t spec_layout = ~0ull;
    int spec_mode = -1;
    int64_t spec_hits = 0;
    // scan scratch and small scalars
    DevBuf scan_bsum, scalars;
    DevBuf scan_side;                // scan scratch of the side stream
    uint64_t* h_scalars = nullptr;   // pinned mirror
    ft_counts last{};
    std::vector<ft_counts> field_counts;
    std::vector<uint64_t> batch_meshes;  // serials of the resident meshes of the current batch
    bool have_interface = false;
    bool have_nearest = false;
    // phase timing (ft_set_timing): events at the phase boundaries of a reinit
    bool timing = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // host <-> device staging of large host arrays: two pinned chunks
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    double phase_ms[4] = {0, 0, 0, 0};
    int64_t timed_calls = 0;
};

namespace {

void drop_graph(ft_ctx* c);
int materialize_conn(ft_ctx* h);

std::vector<DevBuf*> all_buffers(ft_ctx* h)
{
    return {&h->xyz, &h->tets, &h->edges, &h->t2e, &h->perm, &h->sx, &h->sy, &h->sz,
            &h->tile_field, &h->tile_sched, &h->gather_tab, &h->field_tab, &h->d_node_off, &h->d_tet_off, &h->d_edge_off, &h->d_slot_off,
            &h->fcounts, &h->fclamp, &h->d_grids, &h->d_cnt_off, &h->leaf_off, &h->roots,
            &h->phi_tmp, &h->phi_ifn_tmp, &h->sbits, &h->retain, &h->tbits, &h->blk,
            &h->xtet, &h->xtri_off, &h->xedges, &h->ebits, &h->eprefix, &h->ifp,
            &h->ifp2e, &h->tris, &h->tri_norm, &h->if_elems, &h->p2t_off,
            &h->p2t_cursor, &h->p2t, &h->ifp_area, &h->nearest, &h->bin_count,
            &h->bin_prefix, &h->pcode, &h->prec, &h->trec, &h->pidx, &h->leaf_key,
            &h->leaf_start, &h->leaf_cnt, &h->leaf_b, &h->leaf_parent,
            &h->bnodes, &h->bflags, &h->stats, &h->scan_bsum, &h->scan_side, &h->deep_part, &h->scalars,
            &h->n2t_off, &h->n2t_ids, &h->grad_t, &h->normals, &h->phi_ifn_int,
            &h->gsplit, &h->kmesh, &h->k_if, &h->st_cnt, &h->st_off, &h->st_ids, &h->st_f,
            &h->fem_len, &h->fem_w, &h->fem_soff, &h->fem_cols, &h->fem_m, &h->fem_l, &h->fem_flag, &h->fem_map,
            &h->bcg_vec, &h->bcg_part, &h->bcg_sc, &h->adv_phi, &h->adv_flow, &h->adv_vel3, &h->fem_csr_off};
}

std::vector<void*> buffer_addresses(ft_ctx* c)
{
    std::vector<void*> a;
    for (DevBuf* b : all_buffers(c)) a.push_back(b->p);
    return a;
}

// Workspace of an allocation-free call (ft_reinit_device_noalloc): while it
// is set on the calling thread, ensure() hands out consecutive pieces of it
// instead of allocating, and release() drops buffers without freeing.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
};
thread_local Arena* t_arena = nullptr;

int ensure(DevBuf& b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (bytes <= b.cap) return FT_OK;
    if (t_arena) {
        const size_t want = (bytes + 255) & ~(size_t)255;
        if (t_arena->used + want > t_arena->size)
            return ft::set_error(FT_ERR_BUDGET,
                                 "the workspace of this reinit exceeds the %zu-byte scratch of "
                                 "ft_reinit_device_noalloc (%zu bytes taken, %zu more asked); use "
                                 "ft_reinit_device", t_arena->size, t_arena->used, want);
        b.p = t_arena->base + t_arena->used;
        b.cap = want;
        t_arena->used += want;
        return FT_OK;
    }
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
    if (b.p && !t_arena) (void)hipFree(b.p);
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
// ws: the block-sum scratch (default: the handle's; each stream needs its own).
template <class TI, class TO, class Ld>
int scan(ft_ctx* c, hipStream_t s, const TI* in, TO* out, int64_t n, TO* d_total, Ld ld, DevBuf* ws = nullptr)
{
    DevBuf& w = ws ? *ws : c->scan_bsum;
    const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    FT_TRY(ensure(w, sizeof(TO) * (size_t)nb));
    TO* bs = w.as<TO>();
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

// Uniform cubic grid over a field's box with edge >= `edge` and at most
// 2^max_lg bins per axis; *lg receives log2 of the Morton side.
Grid make_grid(const FieldHost& f, float edge, int max_lg, int* lg_out)
{
    Grid g{};
    float ext_max = 0.0f;
    for (int k = 0; k < 3; k++) ext_max = std::max(ext_max, f.hi[k] - f.lo[k]);
    float b = std::max(edge, ext_max / (float)(1 << max_lg) * 1.0001f);
    if (!(b > 0.0f)) b = 1.0f;
    int lg = 0;
    for (int k = 0; k < 3; k++) {
        int d = (int)std::ceil((f.hi[k] - f.lo[k]) / b);
        d = std::max(1, std::min(d, 1 << max_lg));
        g.dim[k] = d;
        g.lo[k] = f.lo[k];
        while ((1 << lg) < d) lg++;
    }
    g.inv_b = 1.0f / b;
    g.h = f.h;
    *lg_out = lg;
    return g;
}

int field_bits(int nf)
{
    int b = 0;
    while ((1 << b) < nf) b++;
    return b;
}

int check_launch() { FT_HIP(hipGetLastError()); return FT_OK; }

// Copies `count` u64 device scalars (ctx->scalars) to the host and waits.
int read_scalars(ft_ctx* c, hipStream_t s, int count)
{
    FT_HIP(hipMemcpyAsync(c->h_scalars, c->scalars.p, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int check_handle(ft_handle h)
{
    if (!h) return ft::set_error(FT_ERR_ARG, "null handle");
    FT_HIP(hipSetDevice(h->device));
    return FT_OK;
}

// ---- Staged host <-> device copies -----------------------------------------
// A pageable copy runs at 17 GB/s (to host) to 30 GB/s (to device) on this
// platform, a pinned one at about 56 GB/s both ways.  Large host arrays go
// through two pinned chunks instead: host threads copy one chunk between
// the caller's array and pinned memory while the DMA engine moves the other.
constexpr size_t kStageChunk = size_t(32) << 20;
constexpr int kStageThreads = 4;
constexpr size_t kStageMin = size_t(8) << 20;          // smaller: one plain copy

void par_memcpy(void* dst, const void* src, size_t n)
{
    if (n < (size_t(4) << 20)) { std::memcpy(dst, src, n); return; }
    const size_t part = ((n + kStageThreads - 1) / kStageThreads + 63) & ~size_t(63);
    std::thread th[kStageThreads - 1];
    int started = 0;
    try {
        for (int t = 1; t < kStageThreads; t++) {
            const size_t lo = std::min(n, part * t), hi = std::min(n, part * (t + 1));
            if (lo >= hi) break;
            th[t - 1] = std::thread([=] { std::memcpy((char*)dst + lo, (const char*)src + lo, hi - lo); });
            started = t;
        }
    } catch (...) {                                     // no thread: copy the rest here
        const size_t lo = std::min(n, part * (started + 1));
        std::memcpy((char*)dst + lo, (const char*)src + lo, n - lo);
    }
    std::memcpy(dst, src, std::min(n, part));
    for (int t = 0; t < started; t++) th[t].join();
}

int stage_setup(ft_ctx* h)
{
    for (int k = 0; k < 2; k++) {
        if (!h->stage[k]) FT_HIP(hipHostMalloc(&h->stage[k], kStageChunk, hipHostMallocDefault));
        if (!h->stage_ev[k]) FT_HIP(hipEventCreateWithFlags(&h->stage_ev[k], hipEventDisableTiming));
    }
    return FT_OK;
}

// Host to device, enqueued on s; returns once the host array may be reused.
int copy_h2d(ft_ctx* h, void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes < kStageMin) {
        if (bytes) FT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        FT_HIP(hipStreamSynchronize(s));
        return FT_OK;
    }
    FT_TRY(stage_setup(h));
    size_t off = 0;
    for (int k = 0; off < bytes; k++, off += kStageChunk) {
        const int b = k & 1;
        const size_t len = std::min(kStageChunk, bytes - off);
        if (k >= 2) FT_HIP(hipEventSynchronize(h->stage_ev[b]));   // chunk k-2 has left this buffer
        par_memcpy(h->stage[b], (const char*)src + off, len);
        FT_HIP(hipMemcpyAsync((char*)dst + off, h->stage[b], len, hipMemcpyHostToDevice, s));
        FT_HIP(hipEventRecord(h->stage_ev[b], s));
    }
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

// Device to host after the work already on s, for a list of (host, device,
// bytes) segments moved as one sequence of chunks.
struct D2H {
    void* host;
    const void* dev;
    size_t bytes;
};
int copy_d2h(ft_ctx* h, const std::vector<D2H>& segs, hipStream_t s)
{
    size_t total = 0;
    for (const D2H& g : segs) total += g.bytes;
    if (total < kStageMin) {
        for (const D2H& g : segs)
            if (g.bytes) FT_HIP(hipMemcpyAsync(g.host, g.dev, g.bytes, hipMemcpyDeviceToHost, s));
        FT_HIP(hipStreamSynchronize(s));
        return FT_OK;
    }
    FT_TRY(stage_setup(h));
    struct Chunk { char* host; const char* dev; size_t len; };
    std::vector<Chunk> ch;
    for (const D2H& g : segs)
        for (size_t off = 0; off < g.bytes; off += kStageChunk)
            ch.push_back({(char*)g.host + off, (const char*)g.dev + off, std::min(kStageChunk, g.bytes - off)});
    for (size_t k = 0; k < ch.size() && k < 2; k++) {
        FT_HIP(hipMemcpyAsync(h->stage[k], ch[k].dev, ch[k].len, hipMemcpyDeviceToHost, s));
        FT_HIP(hipEventRecord(h->stage_ev[k], s));
    }
    for (size_t k = 0; k < ch.size(); k++) {
        const int b = (int)(k & 1);
        FT_HIP(hipEventSynchronize(h->stage_ev[b]));
        par_memcpy(ch[k].host, h->stage[b], ch[k].len);
        if (k + 2 < ch.size()) {
            FT_HIP(hipMemcpyAsync(h->stage[b], ch[k + 2].dev, ch[k + 2].len, hipMemcpyDeviceToHost, s));
            FT_HIP(hipEventRecord(h->stage_ev[b], s));
        }
    }
    return FT_OK;
}


// Field layout: offsets on host and device, slot layout (each field padded to
// whole waves).
int set_fields(ft_ctx* c, int nf, const int64_t* node_off, const int64_t* tet_off,
               const int64_t* edge_off)
{
    c->batch_meshes.clear();
    c->virt = false;
    c->field_meshes.clear();
    c->layout_gen++;
    c->nf = nf;
    c->node_off.assign(node_off, node_off + nf + 1);
    c->tet_off.assign(tet_off, tet_off + nf + 1);
    c->edge_off.assign(edge_off, edge_off + nf + 1);
    c->slot_off.assign(nf + 1, 0);
    for (int k = 0; k < nf; k++)
        c->slot_off[k + 1] = c->slot_off[k] + ((node_off[k + 1] - node_off[k] + 63) / 64) * 64;
    c->N = node_off[nf];
    c->T = tet_off[nf];
    c->E = edge_off[nf];
    c->n_slots = c->slot_off[nf];
    c->fh.assign(nf, FieldHost{});
    c->field_counts.assign(nf, ft_counts{});
    const size_t ob = sizeof(int64_t) * (nf + 1);
    FT_TRY(ensure(c->d_node_off, ob));
    FT_TRY(ensure(c->d_tet_off, ob));
    FT_TRY(ensure(c->d_edge_off, ob));
    FT_TRY(ensure(c->d_slot_off, ob));
    FT_HIP(hipMemcpyAsync(c->d_node_off.p, c->node_off.data(), ob, hipMemcpyHostToDevice, c->stream));
    FT_HIP(hipMemcpyAsync(c->d_tet_off.p, c->tet_off.data(), ob, hipMemcpyHostToDevice, c->stream));
    FT_HIP(hipMemcpyAsync(c->d_edge_off.p, c->edge_off.data(), ob, hipMemcpyHostToDevice, c->stream));
    FT_HIP(hipMemcpyAsync(c->d_slot_off.p, c->slot_off.data(), ob, hipMemcpyHostToDevice, c->stream));
    FT_HIP(hipStreamSynchronize(c->stream));
    return FT_OK;
}

void finish_field_box(FieldHost& f, int64_t n)
{
    if (n == 0) for (int k = 0; k < 3; k++) { f.lo[k] = 0; f.hi[k] = 1; }
    double vol = 1.0, ext_max = 0.0;
    for (int k = 0; k < 3; k++) ext_max = std::max(ext_max, (double)f.hi[k] - f.lo[k]);
    if (!(ext_max > 0.0)) ext_max = 1.0;
    for (int k = 0; k < 3; k++) vol *= std::max((double)f.hi[k] - f.lo[k], ext_max * 1e-3);
    f.h = (float)std::cbrt(vol / (double)std::max<int64_t>(1, n));
}

// Field boxes from host coordinates.
void boxes_from_host(ft_ctx* c, const float* xyz)
{
    for (int k = 0; k < c->nf; k++) {
        FieldHost& f = c->fh[k];
        for (int q = 0; q < 3; q++) { f.lo[q] = FLT_MAX; f.hi[q] = -FLT_MAX; }
        for (int64_t i = c->node_off[k]; i < c->node_off[k + 1]; i++)
            for (int q = 0; q < 3; q++) {
                f.lo[q] = std::min(f.lo[q], xyz[3 * i + q]);
                f.hi[q] = std::max(f.hi[q], xyz[3 * i + q]);
            }
        finish_field_box(f, c->node_off[k + 1] - c->node_off[k]);
    }
    c->h_est = c->nf ? c->fh[0].h : 0.0f;
}

// Grid of the node sort.  A query wave takes 64 consecutive slots, so the
// tiles are compact only if the bins follow the nodes.  When a field's node
// count and box fit a structured grid -- n = nx ny nz nodes with nx - 1,
// ny - 1, nz - 1 spacings s along the box's edges, as a box mesher's (or
// the Kuhn generator's) grid, jittered by less than s / 2 -- the bins are
// kf s wide and start s / 2 before the box: every bin then holds the
// kf^3 nodes of one block, and 64 consecutive slots in Morton order are an
// aligned 4 x 4 x 4 block of nodes (kf = 1 or 2).  Bins of the estimated
// spacing from the box corner cut those blocks where jittered nodes cross a
// bin edge; at 256^3 aligned blocks made the query 4.6% faster.  Any other
// field keeps the grid of about one node per bin.  The order only affects
// speed: each node's result is the same in any slot.
Grid node_sort_grid(const FieldHost& f, int64_t n, int max_lg, int* lg_out)
{
    const Grid fallback = make_grid(f, f.h, max_lg, lg_out);
    double ext[3], ext_max = 0.0;
    for (int k = 0; k < 3; k++) {
        ext[k] = (double)f.hi[k] - f.lo[k];
        if (!(ext[k] >= 0.0)) return fallback;
        ext_max = std::max(ext_max, ext[k]);
    }
    if (n < 8 || !(ext_max > 0.0)) return fallback;
    // prod_k (ext_k / s + 1) decreases in s: bisection for the s that gives n
    auto count = [&](double s) {
        double p = 1.0;
        for (int k = 0; k < 3; k++) p *= ext[k] / s + 1.0;
        return p;
    };
    double lo_s = ext_max / (double)n, hi_s = ext_max;   // count(lo_s) > n >= 8 >= count(hi_s)
    if (!(count(hi_s) <= (double)n)) return fallback;
    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo_s + hi_s);
        (count(mid) > (double)n ? lo_s : hi_s) = mid;
    }
    const double s0 = 0.5 * (lo_s + hi_s);
    int64_t nk[3], prod = 1, cells = 1;
    double s_sum = 0.0;
    int s_cnt = 0;
    for (int k = 0; k < 3; k++) {
        const double q = ext[k] / s0 + 1.0;
        nk[k] = std::llround(q);
        if (nk[k] < 1 || std::fabs(q - (double)nk[k]) > 0.05) return fallback;
        prod *= nk[k];
        cells = std::max(cells, nk[k]);
        if (nk[k] > 1) { s_sum += ext[k] / (double)(nk[k] - 1); s_cnt++; }
    }
    if (prod != n || s_cnt == 0) return fallback;
    const double s = s_sum / s_cnt;
    for (int k = 0; k < 3; k++)
        if (nk[k] > 1 && std::fabs(ext[k] / (double)(nk[k] - 1) - s) > 1e-4 * s) return fallback;
    const int64_t kf = (cells + (1ll << max_lg) - 1) >> max_lg;   // spacings per bin
    if (kf > 2) return fallback;                                   // blocks would not be 4 x 4 x 4
    Grid g{};
    int lg = 0;
    for (int k = 0; k < 3; k++) {
        g.lo[k] = (float)((double)f.lo[k] - 0.5 * s);
        // node i of the axis lands in bin floor((i + 1/2) / kf)
        g.dim[k] = (int)(((nk[k] - 1) * 2 + 1) / (2 * kf)) + 1;
        while ((1 << lg) < g.dim[k]) lg++;
    }
    g.inv_b = (float)(1.0 / ((double)kf * s));
    g.h = f.h;
    *lg_out = lg;
    return g;
}

// Morton order of each field's nodes into the query slots (device xyz):
// one counting sort per field on node_sort_grid.
int sort_nodes(ft_ctx* c)
{
    const int nf = c->nf;
    FT_TRY(ensure(c->perm, sizeof(int32_t) * std::max<int64_t>(1, c->n_slots)));
    FT_TRY(ensure(c->sx, sizeof(float) * std::max<int64_t>(1, c->n_slots)));
    FT_TRY(ensure(c->sy, sizeof(float) * std::max<int64_t>(1, c->n_slots)));
    FT_TRY(ensure(c->sz, sizeof(float) * std::max<int64_t>(1, c->n_slots)));
    FT_TRY(ensure(c->tile_field, sizeof(int32_t) * std::max<int64_t>(1, c->n_slots / 64)));
    FT_TRY(ensure(c->tile_sched, 4 * sizeof(int32_t) * std::max<int64_t>(1, c->n_slots / 64)));
    if (c->N == 0) return FT_OK;
    hipStream_t s = c->stream;
    for (int k = 0; k < nf; k++) {
        const int64_t n = c->node_off[k + 1] - c->node_off[k];
        if (n == 0) continue;
        int lg = 0;
        const Grid g = node_sort_grid(c->fh[k], n, kMaxBinsLog2PerAxis, &lg);
        const int64_t G = (int64_t)1 << (3 * lg);
        FT_TRY(ensure(c->bin_count, sizeof(uint32_t) * G));
        FT_TRY(ensure(c->bin_prefix, sizeof(uint64_t) * G));
        FT_TRY(ensure(c->pcode, sizeof(uint32_t) * n));
        uint32_t* cnt = c->bin_count.as<uint32_t>();
        uint32_t* off = reinterpret_cast<uint32_t*>(c->bin_prefix.p);
        const float* xk = c->xyz.as<float>() + 3 * c->node_off[k];
        const int64_t sb = c->slot_off[k];
        FT_TRY(fill<uint32_t>(s, cnt, G, 0u));
        hipLaunchKernelGGL(k_node_bin, dim3(grid1d(n)), dim3(kBS), 0, s, xk, n, g, c->pcode.as<uint32_t>(), cnt);
        FT_TRY(scan<uint32_t, uint32_t>(c, s, cnt, off, G,
                                        reinterpret_cast<uint32_t*>(c->scalars.as<uint64_t>() + S_BINS),
                                        LoadSame{}));
        hipLaunchKernelGGL(k_node_scatter, dim3(grid1d(n)), dim3(kBS), 0, s, xk, n, c->node_off[k],
                           c->pcode.as<uint32_t>(), cnt, off, c->perm.as<int32_t>() + sb,
                           c->sx.as<float>() + sb, c->sy.as<float>() + sb, c->sz.as<float>() + sb);
    }
    FieldGeom fg{c->d_node_off.as<int64_t>(), nullptr, nullptr, nf};
    hipLaunchKernelGGL(k_pad_slots, dim3(grid1d(c->n_slots)), dim3(kBS), 0, s, c->n_slots, fg,
                       c->d_slot_off.as<int64_t>(), c->perm.as<int32_t>(), c->sx.as<float>(),
                       c->sy.as<float>(), c->sz.as<float>(), c->tile_field.as<int32_t>());
    FT_TRY(check_launch());
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

// ----------------------------------------------------------------------------
// One reinit: stage buffers, the kernel sequence, and the readback
// ----------------------------------------------------------------------------
// Buffer capacities of one reinit, in elements.  An eager reinit sizes them
// to the exact stage totals it reads back between stages.  A captured graph
// fixes them when it is captured; the device then raises S_CAP when a total
// outgrows them, every later kernel does nothing, and the host captures again
// with more room (phi is only written by the last kernel, so nothing is lost).
// Default bin edge, in node spacings, per distance mode (0 points, 1
// triangles).  A bin's primitives form one leaf.  Points are cheap to test and
// their least-squares-plane cylinder tightens with more points, so their
// optimum is larger: measured on 32^3..256^3 fields, points 2.0-2.5, triangles
// 1.25.
constexpr float kDefaultBinScale[2] = {2.0f, 1.25f};

// Per-field bin grids of the index for one bin_scale (0: the mode's default),
// on the device; cached per field layout.
int index_setup(ft_ctx* c, hipStream_t s, float bin_scale, int mode)
{
    const float bs = bin_scale > 0.0f ? bin_scale : kDefaultBinScale[mode];
    if (c->idx_layout == c->layout_gen && c->idx_bin_scale == bs) return FT_OK;
    const int nf = c->nf;
    const int fb = field_bits(nf);
    // keys (field << code_bits) | code stay below 2^30
    const int max_lg = std::min(kMaxBinsLog2PerAxis, (30 - fb) / 3);
    std::vector<Grid> grids(nf);
    std::vector<int64_t> cnt_off(nf + 1, 0);
    int lg_max = 0;
    for (int k = 0; k < nf; k++) {
        int lg = 0;
        grids[k] = make_grid(c->fh[k], bs * c->fh[k].h, max_lg, &lg);
        lg_max = std::max(lg_max, lg);
        cnt_off[k + 1] = cnt_off[k] + ((int64_t)1 << (3 * lg));
    }
    const int64_t G = cnt_off[nf];
    if (G > ((int64_t)1 << 32))
        return ft::set_error(FT_ERR_ARG, "index: %lld bins exceed 2^32; raise bin_scale", (long long)G);
    FT_TRY(ensure(c->d_grids, sizeof(Grid) * nf));
    FT_TRY(ensure(c->d_cnt_off, sizeof(int64_t) * (nf + 1)));
    FT_HIP(hipMemcpyAsync(c->d_grids.p, grids.data(), sizeof(Grid) * nf, hipMemcpyHostToDevice, s));
    FT_HIP(hipMemcpyAsync(c->d_cnt_off.p, cnt_off.data(), sizeof(int64_t) * (nf + 1),
                          hipMemcpyHostToDevice, s));
    FT_HIP(hipStreamSynchronize(s));
    c->idx_G = G;
    c->idx_code_bits = 3 * lg_max;
    c->idx_layout = c->layout_gen;
    c->idx_bin_scale = bs;
    return FT_OK;
}

inline int64_t at_least1(int64_t v) { return std::max<int64_t>(1, v); }

// Cylinder kept only while its volume is below its box's (measured at 256^3:
// 0 -> 21.2 ms, 0.5 -> 20.9, 1 -> 20.8, 2 -> 21.1, 4 -> 21.2 ms query).
constexpr float kCylAlpha = 1.0f;
float cyl_alpha() { return kCylAlpha; }
inline int64_t scan_blocks(int64_t n) { return std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile); }

// Device buffers of the stages of one reinit.  Stage 0: sized by the mesh;
// 1: by crossing tets and triangles; 2: by crossing points; 3: the index over
// n primitives (leaves <= min(bins, n)).
int ensure_stage(ft_ctx* c, int stage, const Caps& cp, int mode, bool want_nearest,
                 bool want_normals = false)
{
    const int64_t N = c->N, T = c->T, E = c->E, nf = c->nf;
    const int64_t nw_nodes = (N + 63) / 64, nw_tets = (T + 63) / 64, nw_edges = (E + 63) / 64;
    const int64_t nblk = (T + kBS - 1) / kBS;
    if (stage == 0) {
        FT_TRY(ensure(c->sbits, 8 * at_least1(nw_nodes)));
        FT_TRY(ensure(c->retain, 8 * at_least1(nw_nodes)));
        FT_TRY(ensure(c->tbits, 8 * at_least1(2 * nw_tets)));   // crossing, then two-triangle
        FT_TRY(ensure(c->blk, 8 * at_least1(nblk)));
        FT_TRY(ensure(c->ebits, 8 * at_least1(nw_edges)));
        FT_TRY(ensure(c->eprefix, 4 * at_least1(nw_edges)));
        FT_TRY(ensure(c->fcounts, sizeof(int64_t) * 3 * nf));
        FT_TRY(ensure(c->fclamp, sizeof(uint64_t) * nf));
        FT_TRY(ensure(c->roots, sizeof(int32_t) * nf));
        FT_TRY(ensure(c->leaf_off, sizeof(int32_t) * (nf + 1)));
        if (want_nearest) FT_TRY(ensure(c->nearest, sizeof(int32_t) * at_least1(N)));
        if (want_normals) {
            FT_TRY(ensure(c->normals, 12 * at_least1(N)));
            FT_TRY(ensure(c->grad_t, 12 * at_least1(T)));
            FT_TRY(ensure(c->phi_ifn_int, 8 * at_least1(N)));
        }
    } else if (stage == 1) {
        FT_TRY(ensure(c->xtet, 4 * at_least1(cp.X)));
        FT_TRY(ensure(c->xtri_off, 4 * at_least1(cp.X)));
        FT_TRY(ensure(c->xedges, 16 * at_least1(cp.X)));
        FT_TRY(ensure(c->if_elems, 4 * at_least1(cp.X)));
        FT_TRY(ensure(c->tris, 16 * at_least1(cp.F)));
        FT_TRY(ensure(c->tri_norm, 4 * at_least1(cp.F)));
    } else if (stage == 2) {
        FT_TRY(ensure(c->ifp, 12 * at_least1(cp.P)));
        FT_TRY(ensure(c->ifp2e, 4 * at_least1(cp.P)));
        FT_TRY(ensure(c->ifp_area, 4 * at_least1(cp.P)));
        FT_TRY(ensure(c->p2t_off, 4 * (cp.P + 1)));
        FT_TRY(ensure(c->p2t_cursor, 4 * at_least1(cp.P)));
        FT_TRY(ensure(c->p2t, 4 * at_least1(3 * cp.F)));
        // the side stream's scans: the point lists, the tile schedule
        FT_TRY(ensure(c->scan_side, sizeof(uint64_t) * std::max(scan_blocks(cp.P), scan_blocks(c->n_slots / 64))));
    } else {
        const int64_t n = mode == 0 ? cp.P : cp.F;
        const int64_t G = c->idx_G;
        const int64_t L = std::min(G, n);
        FT_TRY(ensure(c->bin_count, sizeof(uint32_t) * at_least1(G)));
        FT_TRY(ensure(c->bin_prefix, sizeof(uint64_t) * at_least1(G)));
        FT_TRY(ensure(c->pcode, sizeof(uint32_t) * at_least1(n)));
        FT_TRY(ensure(c->pidx, sizeof(int32_t) * at_least1(n)));
        if (mode == 0) FT_TRY(ensure(c->prec, sizeof(float4) * at_least1(n)));
        else FT_TRY(ensure(c->trec, sizeof(TriRec) * at_least1(n)));
        FT_TRY(ensure(c->leaf_key, sizeof(uint32_t) * at_least1(L)));
        FT_TRY(ensure(c->leaf_start, sizeof(uint32_t) * at_least1(L)));
        FT_TRY(ensure(c->leaf_cnt, sizeof(uint32_t) * at_least1(L)));
        FT_TRY(ensure(c->leaf_b, sizeof(Bound) * at_least1(L)));
        FT_TRY(ensure(c->leaf_parent, sizeof(int32_t) * at_least1(L)));
        FT_TRY(ensure(c->bnodes, sizeof(BNode) * at_least1(L - 1)));
        // arrival counters (L - 1), then the ready count and list (1 + L - 1)
        FT_TRY(ensure(c->bflags, sizeof(unsigned) * at_least1(2 * L)));
        // scan scratch of the largest scan of a reinit
        const int64_t nb = std::max({scan_blocks(nblk), scan_blocks(nw_edges), scan_blocks(cp.P),
                                     scan_blocks(G)});
        FT_TRY(ensure(c->scan_bsum, sizeof(uint64_t) * nb));
        FT_TRY(ensure(c->deep_part, sizeof(ulonglong2) * 64 * kSplitParts * at_least1(split_tiles_cap(c->n_slots / 64))));
    }
    return FT_OK;
}

// The error word of the reinit (S_ERR) as a status.
int stage_error(uint64_t ev, ft_counts* cnt)
{
    if (ev == ~0ull) return FT_OK;
    cnt->bad_tet = (int64_t)(ev & ((1ull << 40) - 1));
    if ((ev >> 40) == (uint64_t)ftp::kErrIndex)
        return ft::set_error(FT_ERR_INTERNAL,
                             "interface index out of range at triangle or point %lld: the "
                             "point -> triangle stage stopped before writing",
                             (long long)cnt->bad_tet);
    if ((ev >> 40) == 1)
        return ft::set_error(FT_ERR_T2E,
                             "tet %lld: crossing edges from t2e do not match its node signs "
                             "(LevelSet.cpp:400-405)", (long long)cnt->bad_tet);
    return ft::set_error(FT_ERR_DEGENERATE,
                         "tet %lld crosses the interface but all its nodes have |phi| <= "
                         "FLT_MIN (LevelSet.cpp:338-341)", (long long)cnt->bad_tet);
}

// Node -> tets lists of the current fields (once per field layout; not in a
// captured graph: it synchronises).
int build_node_tets(ft_ctx* c, hipStream_t s)
{
    FT_TRY(materialize_conn(c));
    if (c->n2t_layout == c->layout_gen) return FT_OK;
    const int64_t N = c->N, T = c->T;
    FT_TRY(ensure(c->n2t_off, sizeof(uint64_t) * (N + 1)));
    FT_TRY(ensure(c->n2t_ids, sizeof(int32_t) * at_least1(4 * T)));
    FT_TRY(ensure(c->pcode, sizeof(uint32_t) * at_least1(N)));          // counters
    uint32_t* cnt = c->pcode.as<uint32_t>();
    uint64_t* off = c->n2t_off.as<uint64_t>();
    FT_TRY(fill<uint32_t>(s, cnt, N, 0u));
    if (T > 0) hipLaunchKernelGGL(k_n2t_count, dim3(grid1d(T)), dim3(kBS), 0, s, c->tets.as<int4>(), T, cnt);
    FT_TRY(scan<uint32_t, uint64_t>(c, s, cnt, off, N, off + N, LoadSame{}));
    FT_TRY(fill<uint32_t>(s, cnt, N, 0u));
    if (T > 0)
        hipLaunchKernelGGL(k_n2t_fill, dim3(grid1d(T)), dim3(kBS), 0, s, c->tets.as<int4>(), T, off, cnt,
                           c->n2t_ids.as<int32_t>());
    if (N > 0) hipLaunchKernelGGL(k_n2t_sort, dim3(grid1d(N)), dim3(kBS), 0, s, off, N, c->n2t_ids.as<int32_t>());
    FT_TRY(check_launch());
    FT_HIP(hipStreamSynchronize(s));
    c->n2t_layout = c->layout_gen;
    return FT_OK;
}

// What the search of a reinit reads and writes.
struct SearchSpec {
    int mode = 0;                    // 0: crossing points, 1: triangles
    float band = 0.0f;
    const double* phi_old = nullptr;
    double* phi_out = nullptr;
    double* phi_ifn = nullptr;
    bool use_retain = false;         // keep phi_old at crossing-edge nodes (update blend)
    int epi = EPI_RAW;
    double w_phi = 0.0, w_ifn = 0.0;
    bool want_nearest = false;
    bool want_normals = false;       // Assembler::grad_nodes of phi_ifn after the epilogue
    bool int_ifn = false;            // phi_ifn into the handle's buffer (normals need it)
    int query_block = 0;
    int waves_per_simd = 8;          // launch bound of the query (triangles: 5..8)
    bool work_queue = false;         // persistent query over tile counters
    bool bins = false;               // narrow band: bins staged in LDS instead of the tree walk
};

// Copies of the resident meshes ms into the handle's concatenated arrays,
// one gather launch per array over all fields: the node coordinates and
// query slots (perm with padding kept at -1, sx, sy, sz as bit copies, the
// tile -> field map as constants), or with conn the tets, t2e and edges
// (offset to global ids).  Their piece tables go up in one copy.
int gather_meshes(ft_ctx* h, const std::vector<std::shared_ptr<ft_ctx>>& ms, hipStream_t s, bool conn)
{
    enum { G_XYZ, G_PERM, G_SX, G_SY, G_SZ, G_TILE, G_TET, G_T2E, G_EDGE, G_N };
    const int g0 = conn ? G_TET : G_XYZ, g1 = conn ? G_N : G_TET;
    const int nf = (int)ms.size();
    const size_t tab = (size_t)(3 * nf + 1);             // src pointers (nf), offsets (nf+1), adds (nf)
    std::vector<int64_t> host((size_t)(g1 - g0) * tab);
    for (int g = g0; g < g1; g++) {
        int64_t* t = host.data() + (g - g0) * tab;
        const int32_t** srcp = reinterpret_cast<const int32_t**>(t);
        int64_t* off = t + nf;
        int64_t* add = t + 2 * nf + 1;
        off[0] = 0;
        for (int k = 0; k < nf; k++) {
            const ft_ctx* m = ms[k].get();
            int64_t cnt = 0, a = 0;
            const void* src = nullptr;
            switch (g) {
            case G_XYZ: src = m->xyz.p; cnt = 3 * m->N; break;
            case G_PERM: src = m->perm.p; cnt = m->n_slots; a = h->node_off[k]; break;
            case G_SX: src = m->sx.p; cnt = m->n_slots; break;
            case G_SY: src = m->sy.p; cnt = m->n_slots; break;
            case G_SZ: src = m->sz.p; cnt = m->n_slots; break;
            case G_TILE: src = nullptr; cnt = m->n_slots / 64; a = k; break;
            case G_TET: src = m->tets.p; cnt = 4 * m->T; a = h->node_off[k]; break;
            case G_T2E: src = m->t2e.p; cnt = 6 * m->T; a = h->edge_off[k]; break;
            default: src = m->edges.p; cnt = 2 * m->E; a = h->node_off[k]; break;
            }
            srcp[k] = static_cast<const int32_t*>(src);
            off[k + 1] = off[k] + cnt;
            add[k] = a;
        }
    }
    FT_TRY(ensure(h->gather_tab, sizeof(int64_t) * host.size()));
    FT_HIP(hipMemcpyAsync(h->gather_tab.p, host.data(), sizeof(int64_t) * host.size(), hipMemcpyHostToDevice, s));
    void* dsts[G_N] = {h->xyz.p, h->perm.p, h->sx.p, h->sy.p, h->sz.p, h->tile_field.p, h->tets.p, h->t2e.p,
                       h->edges.p};
    for (int g = g0; g < g1; g++) {
        const int64_t* t = h->gather_tab.as<int64_t>() + (g - g0) * tab;
        const int64_t total = host[(g - g0) * tab + 2 * nf];
        if (total == 0) continue;
        hipLaunchKernelGGL(k_gather, dim3(grid1d(total, kGatherChunk)), dim3(kBS), 0, s,
                           reinterpret_cast<const int32_t* const*>(t), t + nf, t + 2 * nf + 1, nf,
                           g == G_PERM ? 1 : 0, static_cast<int32_t*>(dsts[g]));
    }
    FT_TRY(check_launch());
    // the host tables must outlive their (possibly staged) copy
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

// The concatenated tets, t2e and edges of a batch of resident meshes, for the
// operations that index them globally (normals, curvature, surface tension,
// advection, ft_get_mesh); a no-op for any other handle.
int materialize_conn(ft_ctx* h)
{
    if (!h->virt) return FT_OK;
    drop_graph(h);                   // captured with the per-field view
    FT_TRY(ensure(h->tets, 16 * std::max<int64_t>(1, h->T)));
    FT_TRY(ensure(h->edges, 8 * std::max<int64_t>(1, h->E)));
    FT_TRY(ensure(h->t2e, 24 * std::max<int64_t>(1, h->T)));
    FT_TRY(gather_meshes(h, h->field_meshes, h->stream, true));
    h->virt = false;
    h->field_meshes.clear();
    return FT_OK;
}

MeshView mesh_view(ft_ctx* c)
{
    MeshView m{};
    m.nf = c->nf;
    m.node_off = c->d_node_off.as<int64_t>();
    m.tet_off = c->d_tet_off.as<int64_t>();
    m.edge_off = c->d_edge_off.as<int64_t>();
    if (!c->virt) {
        m.tets = c->tets.as<int4>();
        m.t2e3 = c->t2e.as<int2>();
        m.edges = c->edges.as<int2>();
    } else {
        const void* const* tab = static_cast<const void* const*>(c->field_tab.p);
        m.f_tets = reinterpret_cast<const int4* const*>(tab);
        m.f_t2e3 = reinterpret_cast<const int2* const*>(tab + c->nf);
        m.f_edges = reinterpret_cast<const int2* const*>(tab + 2 * c->nf);
    }
    return m;
}

// Enqueues the marching tets of one reinit on stream s.  eager: reads the
// stage totals back after the two counting scans (sizing the buffers exactly,
// and stopping early on a failed stage).  !eager (graph capture): no host
// reads and no allocation; the buffers were sized for `cp` beforehand and
// k_check_caps guards them.  counts_only: stops once the totals are known
// (phi is not written).
// Work enqueued on c->side from now on runs after everything enqueued on s
// so far.
int side_fork(ft_ctx* c, hipStream_t s)
{
    FT_HIP(hipEventRecord(c->side_fork_ev, s));
    FT_HIP(hipStreamWaitEvent(c->side, c->side_fork_ev, 0));
    c->side_open = true;
    return FT_OK;
}

// Work enqueued on s from now on runs after everything enqueued on c->side.
int side_join(ft_ctx* c, hipStream_t s)
{
    if (!c->side_open) return FT_OK;
    c->side_open = false;
    FT_HIP(hipEventRecord(c->side_join_ev, c->side));
    FT_HIP(hipStreamWaitEvent(s, c->side_join_ev, 0));
    return FT_OK;
}

// Joins the side stream into s when it leaves scope (early returns included).
struct SideJoin {
    ft_ctx* c;
    hipStream_t s;
    ~SideJoin() { (void)side_join(c, s); }
};

int enqueue_marching(ft_ctx* c, hipStream_t s, double* phi, int mode, bool want_nearest,
                     bool want_normals, Caps& cp, bool eager, bool counts_only, ft_counts* cnt,
                     bool timed)
{
    const int64_t N = c->N, T = c->T, E = c->E;
    uint64_t* sc = c->scalars.as<uint64_t>();
    const int64_t nw_nodes = (N + 63) / 64, nw_tets = (T + 63) / 64, nw_edges = (E + 63) / 64;
    const int64_t nblk = (T + kBS - 1) / kBS;
    timed = timed && !counts_only;
    FT_TRY(ensure_stage(c, 0, cp, mode, want_nearest, want_normals));
    uint64_t* sbits = c->sbits.as<uint64_t>();
    if (timed) (void)hipEventRecord(c->ev[0], s);
    hipLaunchKernelGGL(k_node_signs, dim3(grid1d(N)), dim3(kBS), 0, s, phi, N, sbits);
    FT_TRY(fill<uint64_t>(s, c->retain.as<uint64_t>(), nw_nodes, 0ull));
    FT_TRY(fill<uint64_t>(s, c->ebits.as<uint64_t>(), nw_edges, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_ERR, 1, ~0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_CAP, 1, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_OVF, 1, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_EDGES, 1, 0ull));   // the edge scan writes the low 32 bits
    if (T > 0)
        hipLaunchKernelGGL(k_classify, dim3((unsigned)((nblk + kClassifyGroups - 1) / kClassifyGroups)),
                           dim3(kBS), 0, s, mesh_view(c), T, sbits, c->tbits.as<uint64_t>(),
                           c->tbits.as<uint64_t>() + nw_tets, c->blk.as<uint64_t>());
    FT_TRY(scan<uint64_t, uint64_t>(c, s, c->blk.as<uint64_t>(), c->blk.as<uint64_t>(), nblk,
                                    sc + S_BLK, LoadSame{}));
    if (eager) {
        FT_TRY(read_scalars(c, s, S_N));
        cp.X = (int64_t)(c->h_scalars[S_BLK] >> 32);
        cp.F = (int64_t)(c->h_scalars[S_BLK] & 0xffffffffull);
        cnt->n_if_elems = cp.X;
        cnt->n_tri = cp.F;
    } else {
        hipLaunchKernelGGL(k_check_caps, dim3(1), dim3(1), 0, s, sc, cp.X, cp.F, (int64_t)-1);
    }
    FT_TRY(ensure_stage(c, 1, cp, mode, want_nearest));
    if (nw_tets > 0)
        hipLaunchKernelGGL(k_compact, dim3(grid1d(nw_tets)), dim3(kBS), 0, s, c->tbits.as<uint64_t>(),
                           c->tbits.as<uint64_t>() + nw_tets, nw_tets, c->blk.as<uint64_t>(),
                           c->xtet.as<int32_t>(), c->xtri_off.as<int32_t>(), sc);
    if (cp.X > 0)
        hipLaunchKernelGGL(k_xtet, dim3(grid1d(cp.X)), dim3(kBS), 0, s, c->tbits.as<uint64_t>() + nw_tets,
                           mesh_view(c), sbits, c->xtet.as<int32_t>(), sc,
                           c->ebits.as<unsigned long long>(), c->xedges.as<int4>());
    FT_TRY(scan<uint64_t, uint32_t>(c, s, c->ebits.as<uint64_t>(), c->eprefix.as<uint32_t>(),
                                    nw_edges, reinterpret_cast<uint32_t*>(sc + S_EDGES), LoadPopc{}));
    if (eager) {
        FT_TRY(read_scalars(c, s, S_N));
        cp.P = (int64_t)(c->h_scalars[S_EDGES] & 0xffffffffull);
        cnt->n_ifp = cp.P;
        FT_TRY(stage_error(c->h_scalars[S_ERR], cnt));
        if (counts_only) return FT_OK;
    } else {
        hipLaunchKernelGGL(k_check_caps, dim3(1), dim3(1), 0, s, sc, (int64_t)-1, (int64_t)-1, cp.P);
    }
    FT_TRY(ensure_stage(c, 2, cp, mode, want_nearest));
    if (nw_edges > 0)
        hipLaunchKernelGGL(k_points, dim3(grid1d(nw_edges)), dim3(kBS), 0, s, c->ebits.as<uint64_t>(),
                           c->eprefix.as<uint32_t>(), nw_edges, mesh_view(c), c->xyz.as<float>(),
                           phi, c->ifp.as<float>(), c->ifp2e.as<int32_t>(),
                           c->retain.as<unsigned long long>(), sc);
    if (cp.X > 0)
        hipLaunchKernelGGL(k_tris, dim3(grid1d(cp.X)), dim3(kBS), 0, s, mesh_view(c), c->xyz.as<float>(),
                           phi, c->xtet.as<int32_t>(), c->xtri_off.as<int32_t>(), c->xedges.as<int4>(),
                           c->ebits.as<uint64_t>(), c->eprefix.as<uint32_t>(), c->ifp.as<float>(),
                           c->tris.as<int4>(), c->tri_norm.as<float>(), c->if_elems.as<int32_t>(), sc);
    // p2t and areas; entries past the point count stay empty.  The query
    // reads neither: they run on the side stream, beside the index build,
    // and the query's stream joins it before the query.
    FT_TRY(side_fork(c, s));
    {
        hipStream_t ss = c->side;
        uint32_t* deg = c->p2t_cursor.as<uint32_t>();
        FT_TRY(fill<uint32_t>(ss, deg, cp.P, 0u));
        if (cp.F > 0)
            hipLaunchKernelGGL(k_p2t_count, dim3(grid1d(cp.F)), dim3(kBS), 0, ss, c->tris.as<int4>(), sc, deg);
        FT_TRY(scan<uint32_t, uint32_t>(c, ss, deg, c->p2t_off.as<uint32_t>(), cp.P,
                                        c->p2t_off.as<uint32_t>() + cp.P, LoadSame{}, &c->scan_side));
        FT_TRY(fill<uint32_t>(ss, deg, cp.P, 0u));
        if (cp.F > 0)
            hipLaunchKernelGGL(k_p2t_fill, dim3(grid1d(cp.F)), dim3(kBS), 0, ss, c->tris.as<int4>(), sc,
                               c->p2t_off.as<uint32_t>(), deg, c->p2t.as<int32_t>());
        if (cp.P > 0)
            hipLaunchKernelGGL(k_point_area, dim3(grid1d(cp.P)), dim3(kBS), 0, ss, c->p2t_off.as<uint32_t>(),
                               sc, c->p2t.as<int32_t>(), c->tri_norm.as<float>(), c->ifp_area.as<float>());
    }
    if (c->nf > 1)
        hipLaunchKernelGGL(k_field_counts, dim3((c->nf + 63) / 64), dim3(64), 0, s, c->nf,
                           c->d_tet_off.as<int64_t>(), c->d_edge_off.as<int64_t>(), T, E,
                           c->tbits.as<uint64_t>(), c->blk.as<uint64_t>(), c->xtri_off.as<int32_t>(),
                           c->ebits.as<uint64_t>(), c->eprefix.as<uint32_t>(), sc,
                           c->fcounts.as<int64_t>());
    FT_TRY(check_launch());
    if (timed) (void)hipEventRecord(c->ev[1], s);
    // a degenerate tet found by k_tris halts the rest on the device; the final
    // readback reports it
    return FT_OK;
}

// The query kernel for (mode, threads per workgroup, waves per SIMD); points
// mode has the 8-wave bound only (its epilogue has no FP64 settle to spill).
template <int MODE, int BS>
int launch_query_bs(int wps, dim3 g, dim3 b, hipStream_t s, const QueryArgs& q)
{
    switch (MODE == 0 ? 8 : wps) {
    case 5: hipLaunchKernelGGL((k_query_wave<MODE, BS, 5>), g, b, 0, s, q); break;
    case 6: hipLaunchKernelGGL((k_query_wave<MODE, BS, 6>), g, b, 0, s, q); break;
    case 7: hipLaunchKernelGGL((k_query_wave<MODE, BS, 7>), g, b, 0, s, q); break;
    default: hipLaunchKernelGGL((k_query_wave<MODE, BS, 8>), g, b, 0, s, q); break;
    }
    return check_launch();
}

int launch_query(int mode, int qb, int wps, dim3 g, dim3 b, hipStream_t s, const QueryArgs& q)
{
    if (mode == 0) {
        if (qb == 64) return launch_query_bs<0, 64>(wps, g, b, s, q);
        if (qb == 128) return launch_query_bs<0, 128>(wps, g, b, s, q);
        return launch_query_bs<0, 256>(wps, g, b, s, q);
    }
    if (qb == 64) return launch_query_bs<1, 64>(wps, g, b, s, q);
    if (qb == 128) return launch_query_bs<1, 128>(wps, g, b, s, q);
    return launch_query_bs<1, 256>(wps, g, b, s, q);
}

// Enqueues the bin index over the primitives (crossing points c->ifp, or
// triangles c->tris) and the distance query with its epilogue.  The number of
// primitives is read on the device (S_EDGES or S_BLK); cp bounds it.
int enqueue_search(ft_ctx* c, hipStream_t s, const SearchSpec& sp, const Caps& cp, bool eager,
                   bool timed)
{
    const int mode = sp.mode;
    uint64_t* sc = c->scalars.as<uint64_t>();
    (void)eager;
    const int64_t n_cap = mode == 0 ? cp.P : cp.F;
    const float* ifp = c->ifp.as<float>();
    const int4* tris = mode == 0 ? nullptr : c->tris.as<int4>();
    const int nf = c->nf;
    if (c->idx_layout != c->layout_gen)
        return ft::set_error(FT_ERR_STATE, "index grids not set up for this field layout");
    FT_TRY(ensure_stage(c, 3, cp, mode, sp.want_nearest));
    // Deep tiles first, in a full-domain search (in a narrow band they end
    // early, clamped).  The schedule reads only phi and the slots: it is
    // built on the side stream while the index is built here.
    const int64_t n_tiles = c->n_slots / 64;
    int32_t* tile_list = nullptr;
    if (sp.band <= 0.0f && n_tiles > 0) {
        FT_TRY(side_fork(c, s));
        hipStream_t ss = c->side;
        // tile_sched: classes (u32), the launch order (i32), the scan (u64)
        uint32_t* cls = c->tile_sched.as<uint32_t>();
        tile_list = reinterpret_cast<int32_t*>(cls + n_tiles);
        uint64_t* ex = reinterpret_cast<uint64_t*>(cls + 2 * n_tiles);
        hipLaunchKernelGGL(k_tile_deep, dim3(grid1d(64 * n_tiles)), dim3(kBS), 0, ss, n_tiles,
                           c->perm.as<int32_t>(), sp.phi_old, c->tile_field.as<int32_t>(),
                           c->d_grids.as<Grid>(), cls, sc);
        FT_TRY(scan<uint32_t, uint64_t>(c, ss, cls, ex, n_tiles, sc + S_DEEP, LoadTileClass{}, &c->scan_side));
        hipLaunchKernelGGL(k_tile_list, dim3(grid1d(n_tiles)), dim3(kBS), 0, ss, n_tiles, cls, ex, sc,
                           tile_list);
    }
    const int64_t G = c->idx_G;
    const int64_t L_cap = std::min(G, n_cap);
    IndexGeom ig{c->d_edge_off.as<int64_t>(), c->d_cnt_off.as<int64_t>(), c->d_grids.as<Grid>(), nf,
                 c->idx_code_bits};
    uint32_t* bcnt = c->bin_count.as<uint32_t>();
    uint64_t* pre = c->bin_prefix.as<uint64_t>();
    FT_TRY(fill<uint32_t>(s, bcnt, G, 0u));
    if (n_cap > 0)
        hipLaunchKernelGGL(k_prim_bin, dim3(grid1d(n_cap)), dim3(kBS), 0, s, mode, sc, ifp, tris,
                           nf > 1 ? c->ifp2e.as<int32_t>() : nullptr, ig, c->pcode.as<uint32_t>(), bcnt);
    FT_TRY(scan<uint32_t, uint64_t>(c, s, bcnt, pre, G, sc + S_BINS, LoadNonEmpty{}));
    hipLaunchKernelGGL(k_field_roots, dim3((nf + 1 + 63) / 64), dim3(64), 0, s, nf,
                       c->d_cnt_off.as<int64_t>(), pre, sc + S_BINS, G,
                       c->leaf_off.as<int32_t>(), c->roots.as<int32_t>());
    if (n_cap > 0) {
        hipLaunchKernelGGL(k_leaves, dim3(grid1d(G)), dim3(kBS), 0, s, bcnt, pre, G, ig,
                           c->leaf_key.as<uint32_t>(), c->leaf_start.as<uint32_t>(),
                           c->leaf_cnt.as<uint32_t>());
        hipLaunchKernelGGL(k_prim_scatter, dim3(grid1d(n_cap)), dim3(kBS), 0, s, mode, sc, ifp, tris,
                           c->pcode.as<uint32_t>(), bcnt, pre, c->prec.as<float4>(),
                           c->trec.as<TriRec>(), c->pidx.as<int32_t>());
        hipLaunchKernelGGL(mode == 0 ? k_leaf_bounds<0> : k_leaf_bounds<1>, dim3(grid1d(L_cap)), dim3(kBS), 0, s, sc,
                           c->leaf_start.as<uint32_t>(), c->leaf_cnt.as<uint32_t>(),
                           c->pidx.as<int32_t>(), ifp, tris, c->leaf_b.as<Bound>());
    }
    if (L_cap > 1) {
        // arrival counters and the ready count
        FT_TRY(fill<unsigned>(s, c->bflags.as<unsigned>(), L_cap, 0u));
        hipLaunchKernelGGL(k_karras, dim3(grid1d(L_cap - 1)), dim3(kBS), 0, s,
                           c->leaf_key.as<uint32_t>(), sc, c->bnodes.as<BNode>(),
                           c->leaf_parent.as<int32_t>(), c->idx_code_bits,
                           c->leaf_off.as<int32_t>(), c->roots.as<int32_t>());
        unsigned* ready = c->bflags.as<unsigned>() + (L_cap - 1);
        hipLaunchKernelGGL(k_refit_local, dim3((unsigned)((L_cap + kRefitBlock - 1) / kRefitBlock)),
                           dim3(kRefitBlock), 0, s, sc, c->leaf_parent.as<int32_t>(),
                           c->leaf_b.as<Bound>(), c->bnodes.as<BNode>(), c->bflags.as<unsigned>(),
                           ready, cyl_alpha());
        hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(kRefitTopThreads), 0, s, sc,
                           c->bnodes.as<BNode>(), c->bflags.as<unsigned>(), ready, cyl_alpha());
    }
    FT_TRY(check_launch());

    // ---- distance query with the fused epilogue --------------------------------
    FT_TRY(fill<uint64_t>(s, c->fclamp.as<uint64_t>(), nf, 0ull));
    QueryArgs q{};
    q.n_slots = c->n_slots;
    q.sx = c->sx.as<float>(); q.sy = c->sy.as<float>(); q.sz = c->sz.as<float>();
    q.perm = c->perm.as<int32_t>();
    q.tile_field = c->tile_field.as<int32_t>();
    q.roots = c->roots.as<int32_t>();
    q.prec = c->prec.as<float4>();
    q.trec = c->trec.as<TriRec>();
    q.pidx = c->pidx.as<int32_t>();
    q.leaf_start = c->leaf_start.as<uint32_t>();
    q.leaf_cnt = c->leaf_cnt.as<uint32_t>();
    q.leaf_b = c->leaf_b.as<Bound>();
    q.nodes = c->bnodes.as<BNode>();
    q.L = (int)L_cap;
    q.band = sp.band;
    q.ifp = ifp;
    q.tris = tris;
    q.phi_old = sp.phi_old;
    q.phi_out = sp.phi_out;
    q.phi_ifn = sp.phi_ifn ? sp.phi_ifn : (sp.int_ifn ? c->phi_ifn_int.as<double>() : nullptr);
    q.retain = sp.use_retain ? c->retain.as<uint64_t>() : nullptr;   // resolved after allocation
    q.nearest = sp.want_nearest ? c->nearest.as<int32_t>() : nullptr;
    q.epi = sp.epi;
    q.w_phi = sp.w_phi;
    q.w_ifn = sp.w_ifn;
    q.n_clamped = c->fclamp.as<unsigned long long>();
    static const bool want_stats = getenv("FT_QUERY_STATS") != nullptr;
    q.h_est = c->h_est;
    if (want_stats && eager) {
        FT_TRY(ensure(c->stats, 8 * 40));
        FT_TRY(fill<uint64_t>(s, c->stats.as<uint64_t>(), 40, 0ull));
        q.stats = c->stats.as<unsigned long long>();
    }
    q.overflow = (unsigned*)(sc + S_OVF);
    // The greedy probe seeds each full-domain search with a candidate (256^3:
    // query 11.3 -> 10.4 ms).  In a narrow band the band itself bounds the
    // search and most tiles end clamped, for which the probe is wasted work
    // (128^3 star, 6h band: 0.696 -> 0.675 ms without it).
    q.probe = sp.band <= 0.0f && !(getenv("FT_PROBE") && !strcmp(getenv("FT_PROBE"), "0"));
    q.sc = sc;
    q.xcd_chunk = kXcdChunk;
    if (timed) (void)hipEventRecord(c->ev[2], s);
    FT_TRY(side_join(c, s));                          // the schedule and the point lists
    if (tile_list) {
        q.tile_list = tile_list;
        q.n_deep = sc + S_DEEP;
    }
    if (c->n_slots > 0) {
        const int qb = sp.query_block == 64 || sp.query_block == 128 ? sp.query_block : 256;
        q.xcd_chunk = qb == 64 ? kXcdChunk64 : kXcdChunk;
        const int wps = mode == 0 ? 8 : sp.waves_per_simd;
        if (sp.bins) {
            // narrow band from the bins, staged in LDS (no tree walk)
            FT_TRY(fill<uint64_t>(s, sc + S_MARGIN, 1, 0ull));
            if (n_cap > 0)
                hipLaunchKernelGGL(k_tri_margin, dim3((unsigned)std::min<int64_t>(grid1d(n_cap), kMarginBlocks)),
                                   dim3(kBS), 0, s, sc, ifp, tris, reinterpret_cast<uint32_t*>(sc + S_MARGIN));
            const BinQuery bq{G, c->bin_prefix.as<uint64_t>(), c->d_grids.as<Grid>(), c->d_cnt_off.as<int64_t>()};
            const dim3 gb(grid1d(c->n_slots, qb)), bb(qb);
            if (qb == 64) hipLaunchKernelGGL(k_query_bins<64>, gb, bb, 0, s, q, bq);
            else if (qb == 128) hipLaunchKernelGGL(k_query_bins<128>, gb, bb, 0, s, q, bq);
            else hipLaunchKernelGGL(k_query_bins<256>, gb, bb, 0, s, q, bq);
            FT_TRY(check_launch());
        } else {
        // with a tile list: up to 7 padding workgroups after the deep ones,
        // one more for the split of the remainder, and the extra waves of
        // the split tiles (static schedule)
        if (q.tile_list && !sp.work_queue) q.split_out = c->deep_part.as<ulonglong2>();
        const int64_t split_waves = q.split_out ? split_tiles_cap(n_tiles) * (kSplitParts - 1) : 0;
        dim3 g(grid1d(c->n_slots, qb) + (q.tile_list ? 8u : 0u) + (unsigned)((split_waves + qb / 64 - 1) / (qb / 64))),
            b(qb);
        if (sp.work_queue) {
            // persistent: as many waves as the device holds at this launch
            // bound (4 SIMDs per CU), whole workgroups, at least one per XCD
            const int64_t waves = (int64_t)c->cu_count * 4 * wps;
            g = dim3((unsigned)std::max<int64_t>(kXcds, waves / (qb / 64)));
            q.queue = reinterpret_cast<uint32_t*>(sc + S_Q);
            FT_TRY(fill<uint32_t>(s, q.queue, 9, 0u));
        }
        FT_TRY(launch_query(mode, qb, wps, g, b, s, q));
        if (q.split_out) {
            const dim3 gm(grid1d(64 * split_tiles_cap(n_tiles)));
            if (mode == 0) hipLaunchKernelGGL(k_split_merge<0>, gm, dim3(kBS), 0, s, q);
            else hipLaunchKernelGGL(k_split_merge<1>, gm, dim3(kBS), 0, s, q);
            FT_TRY(check_launch());
        }
        }
    }
    if (timed) (void)hipEventRecord(c->ev[3], s);
    if (sp.want_normals && c->N > 0) {
        if (c->n2t_layout != c->layout_gen)
            return ft::set_error(FT_ERR_STATE, "node -> tets lists not built for this field layout");
        if (c->T > 0)
            hipLaunchKernelGGL(k_tet_grad, dim3(grid1d(c->T)), dim3(kBS), 0, s, c->tets.as<int4>(), c->T,
                               c->xyz.as<float>(), q.phi_ifn, c->grad_t.as<float>(), sc);
        hipLaunchKernelGGL(k_node_normals, dim3(grid1d(c->N)), dim3(kBS), 0, s, c->n2t_off.as<uint64_t>(),
                           c->n2t_ids.as<int32_t>(), c->grad_t.as<float>(), c->N, c->normals.as<float>(), sc);
        FT_TRY(check_launch());
    }
    if (timed) (void)hipEventRecord(c->ev[4], s);
    return FT_OK;
}

// After the reinit's kernels: totals, errors and per-field counts from the
// device (one readback).  Returns FT_ERR_BUDGET internally when a stage
// outgrew the capacities of a captured graph.
int collect(ft_ctx* c, hipStream_t s, bool want_nearest, ft_counts* cnt)
{
    const int nf = c->nf;
    FT_HIP(hipMemcpyAsync(c->h_scalars, c->scalars.p, sizeof(uint64_t) * S_N, hipMemcpyDeviceToHost, s));
    std::vector<uint64_t> clamp(nf);
    std::vector<int64_t> fc(nf > 1 ? 3 * nf : 0);
    FT_HIP(hipMemcpyAsync(clamp.data(), c->fclamp.p, 8 * nf, hipMemcpyDeviceToHost, s));
    if (nf > 1)
        FT_HIP(hipMemcpyAsync(fc.data(), c->fcounts.p, sizeof(int64_t) * 3 * nf, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    const uint64_t* hs = c->h_scalars;
    *cnt = ft_counts{};
    cnt->n_if_elems = (int64_t)(hs[S_BLK] >> 32);
    cnt->n_tri = (int64_t)(hs[S_BLK] & 0xffffffffull);
    cnt->n_ifp = (int64_t)(hs[S_EDGES] & 0xffffffffull);
    cnt->bad_tet = -1;
    if (hs[S_CAP]) return FT_ERR_BUDGET;
    FT_TRY(stage_error(hs[S_ERR], cnt));
    c->have_interface = true;
    if (hs[S_OVF])
        return ft::set_error(FT_ERR_HIP, "search: traversal exceeded its bounds (%llu)",
                             (unsigned long long)hs[S_OVF]);
    c->field_counts.assign(nf, ft_counts{});
    for (int k = 0; k < nf; k++) {
        ft_counts& f = c->field_counts[k];
        if (nf > 1) {
            f.n_ifp = fc[3 * k];
            f.n_tri = fc[3 * k + 1];
            f.n_if_elems = fc[3 * k + 2];
        } else {
            f = *cnt;
        }
        f.bad_tet = -1;
        f.n_clamped = (int64_t)clamp[k];
        cnt->n_clamped += f.n_clamped;
    }
    c->have_nearest = want_nearest;
    return FT_OK;
}

void print_query_stats(ft_ctx* c, int mode)
{
    uint64_t st[40];
    if (hipMemcpy(st, c->stats.p, sizeof st, hipMemcpyDeviceToHost) != hipSuccess) return;
    static const char* names[8] = {"out <2h", "out 2-8h", "out 8-32h", "out >32h",
                                   "in <2h", "in 2-8h", "in 8-32h", "in >32h"};
    double tn = 0, tl = 0, tp = 0, tq = 0;
    for (int b = 0; b < 8; b++) {
        tn += st[5 * b + 1]; tl += st[5 * b + 2]; tp += st[5 * b + 3]; tq += st[5 * b + 4];
    }
    fprintf(stderr, "[ftreinit] query mode %d: N=%lld per node: tree nodes %.1f, leaves %.1f, "
            "primitives %.1f, lanes admitting a tested leaf %.1f%%\n", mode, (long long)c->N,
            tn / c->N, tl / c->N, tp / c->N, tp > 0 ? 100.0 * tq / (64.0 * tp) : 0.0);
    for (int b = 0; b < 8; b++) {
        const double nn = (double)st[5 * b];
        if (nn > 0)
            fprintf(stderr, "[ftreinit]   %-9s %5.1f%% of nodes: tree %.1f leaves %.1f prims %.1f "
                    "admitting lanes %.1f%%\n", names[b], 100.0 * nn / c->N, st[5 * b + 1] / nn,
                    st[5 * b + 2] / nn, st[5 * b + 3] / nn,
                    st[5 * b + 3] > 0 ? 100.0 * st[5 * b + 4] / (64.0 * st[5 * b + 3]) : 0.0);
    }
}

SearchSpec reinit_search_spec(double* phi, double* phi_ifn, const ft_opts* o)
{
    SearchSpec sp;
    sp.mode = o->mode == FT_DIST_POINTS ? 0 : 1;
    sp.band = (float)o->band;
    sp.phi_old = phi;
    sp.phi_out = phi;                // read and written per node in place
    sp.phi_ifn = phi_ifn;
    sp.use_retain = true;
    sp.epi = (o->blend == FT_BLEND_INIT) ? EPI_INIT : EPI_UPDATE;
    sp.w_phi = o->w_phi;
    sp.w_ifn = o->w_ifn;
    sp.want_nearest = o->want_nearest != 0;
    sp.want_normals = o->want_normals != 0;
    sp.int_ifn = sp.want_normals && !phi_ifn;
    sp.query_block = o->query_block;
    sp.waves_per_simd = o->waves_per_simd > 0 ? o->waves_per_simd : 8;
    sp.work_queue = o->schedule == FT_SCHED_QUEUE;
    sp.bins = o->schedule == FT_SCHED_BINS;
    return sp;
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
    if (o->query_block != 0 && o->query_block != 64 && o->query_block != 128 && o->query_block != 256)
        return ft::set_error(FT_ERR_ARG, "query_block must be 0, 64, 128 or 256");
    if (o->waves_per_simd != 0 && (o->waves_per_simd < 5 || o->waves_per_simd > 8))
        return ft::set_error(FT_ERR_ARG, "waves_per_simd must be 0 or 5..8");
    if (o->schedule != FT_SCHED_STATIC && o->schedule != FT_SCHED_QUEUE && o->schedule != FT_SCHED_BINS)
        return ft::set_error(FT_ERR_ARG, "unknown query schedule %d", o->schedule);
    if (o->schedule == FT_SCHED_BINS && !(o->band > 0.0 && o->mode == FT_DIST_TRIANGLES))
        return ft::set_error(FT_ERR_ARG, "the bins schedule is for a narrow band in triangles mode");
    return FT_OK;
}

// Capacity with headroom over a measured stage total.
inline int64_t with_room(int64_t v) { return v + v / 4 + 256; }

int reinit_device(ft_ctx* c, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                  ft_counts* counts, hipStream_t s, bool allow_spec = true)
{
    FT_TRY(validate_opts(o));
    if (!c->has_conn) return ft::set_error(FT_ERR_STATE, "reinit before a mesh was set");
    if (!d_phi) return ft::set_error(FT_ERR_ARG, "null phi");
    c->have_interface = false;
    c->have_nearest = false;
    c->have_normals = false;
    ft_counts cnt{};
    cnt.bad_tet = -1;
    Caps cp;
    const SearchSpec sp = reinit_search_spec(d_phi, d_phi_ifn, o);
    SideJoin join{c, s};
    int rc = index_setup(c, s, o->bin_scale, sp.mode);
    if (rc == FT_OK && sp.want_normals) rc = build_node_tets(c, s);
    // Speculative pass: the capacities of the last call on these fields, no
    // readback between stages (the device guards them, and halts before phi is
    // written if a stage outgrows them); else, or then, the exact pass, which
    // reads the stage totals back to size the buffers.
    const bool spec = allow_spec && rc == FT_OK && c->spec_layout == c->layout_gen && c->spec_mode == sp.mode;
    if (spec) {
        cp = c->spec_caps;
        rc = enqueue_marching(c, s, d_phi, sp.mode, sp.want_nearest, sp.want_normals, cp, false, false,
                              &cnt, c->timing);
        if (rc == FT_OK) rc = enqueue_search(c, s, sp, cp, false, c->timing);
        if (rc == FT_OK) rc = collect(c, s, sp.want_nearest, &cnt);
        c->spec_hits += rc != FT_ERR_BUDGET;
    }
    if (!spec || rc == FT_ERR_BUDGET) {
        cp = Caps{};
        cnt = ft_counts{};
        cnt.bad_tet = -1;
        rc = index_setup(c, s, o->bin_scale, sp.mode);
        if (rc == FT_OK)
            rc = enqueue_marching(c, s, d_phi, sp.mode, sp.want_nearest, sp.want_normals, cp, true, false,
                                  &cnt, c->timing);
        if (rc == FT_OK) rc = enqueue_search(c, s, sp, cp, true, c->timing);
        if (rc == FT_OK) rc = collect(c, s, sp.want_nearest, &cnt);
        if (rc == FT_OK) {
            c->spec_caps.X = with_room(cp.X);
            c->spec_caps.F = with_room(cp.F);
            c->spec_caps.P = with_room(cp.P);
            c->spec_layout = c->layout_gen;
            c->spec_mode = sp.mode;
        }
    }
    c->have_normals = rc == FT_OK && sp.want_normals;
    c->last_phi = d_phi;
    if (counts) *counts = cnt;
    c->last = cnt;
    static const bool want_stats = getenv("FT_QUERY_STATS") != nullptr;
    if (rc == FT_OK && want_stats && c->N > 0) print_query_stats(c, sp.mode);
    if (rc == FT_OK && c->timing) {
        float ms[4] = {0, 0, 0, 0};
        (void)hipEventSynchronize(c->ev[4]);
        for (int k = 0; k < 4; k++) (void)hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]);
        for (int k = 0; k < 4; k++) c->phase_ms[k] += ms[k];
        c->timed_calls++;
    }
    return rc;
}

// ----------------------------------------------------------------------------
// Graph form: the whole reinit of the resident fields as one captured graph
// ----------------------------------------------------------------------------
bool same_opts(const ft_opts& a, const ft_opts& b)
{
    return a.mode == b.mode && a.blend == b.blend && a.w_phi == b.w_phi && a.w_ifn == b.w_ifn &&
           a.band == b.band && a.bin_scale == b.bin_scale && a.want_nearest == b.want_nearest &&
           a.query_block == b.query_block && a.want_normals == b.want_normals &&
           a.waves_per_simd == b.waves_per_simd && a.schedule == b.schedule;
}

void drop_graph(ft_ctx* c)
{
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    if (c->graph.graph) (void)hipGraphDestroy(c->graph.graph);
    c->graph.exec = nullptr;
    c->graph.graph = nullptr;
}

// Captures the reinit of (d_phi, d_phi_ifn, o) with capacities cp on the
// handle's stream.  Everything it touches is allocated before the capture.
int capture_reinit(ft_ctx* c, double* d_phi, double* d_phi_ifn, const ft_opts* o, const Caps& cp)
{
    drop_graph(c);
    const SearchSpec sp = reinit_search_spec(d_phi, d_phi_ifn, o);
    if (sp.want_normals) FT_TRY(build_node_tets(c, c->stream));
    for (int st = 0; st < 4; st++) FT_TRY(ensure_stage(c, st, cp, sp.mode, sp.want_nearest, sp.want_normals));
    FT_HIP(hipStreamSynchronize(c->stream));
    FT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    Caps cc = cp;
    ft_counts unused{};
    int rc = enqueue_marching(c, c->stream, d_phi, sp.mode, sp.want_nearest, sp.want_normals, cc, false,
                              false, &unused, false);
    if (rc == FT_OK) rc = enqueue_search(c, c->stream, sp, cc, false, false);
    const int rj = side_join(c, c->stream);          // the capture must end joined
    if (rc == FT_OK) rc = rj;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != FT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess)
        return ft::set_error(FT_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiateWithFlags(&exec, graph, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return ft::set_error(FT_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    GraphState& g = c->graph;
    g.graph = graph;
    g.exec = exec;
    g.phi = d_phi;
    g.phi_ifn = d_phi_ifn;
    g.opts = *o;
    g.layout = c->layout_gen;
    g.caps = cp;
    g.bufs = buffer_addresses(c);
    g.captures++;
    return FT_OK;
}

// Reinit of the resident fields through a captured graph: the first call (or
// a call whose pointers, options, fields or buffer sizes differ from the
// captured ones) sizes the stages with a counting pass over phi, captures
// with headroom and launches; later calls launch the same graph.  A call
// whose interface outgrows the headroom is detected on the device (nothing
// is written), captured again with room for it, and launched again.
int reinit_graph(ft_ctx* c, double* d_phi, double* d_phi_ifn, const ft_opts* o, ft_counts* counts,
                 hipStream_t s)
{
    FT_TRY(validate_opts(o));
    if (!c->has_conn) return ft::set_error(FT_ERR_STATE, "reinit before a mesh was set");
    if (!d_phi) return ft::set_error(FT_ERR_ARG, "null phi");
    c->have_interface = false;
    c->have_nearest = false;
    c->have_normals = false;
    GraphState& g = c->graph;
    ft_counts cnt{};
    cnt.bad_tet = -1;
    FT_TRY(index_setup(c, s, o->bin_scale, o->mode == FT_DIST_POINTS ? 0 : 1));
    const bool same = g.exec && g.phi == d_phi && g.phi_ifn == d_phi_ifn && same_opts(g.opts, *o) &&
                      g.layout == c->layout_gen;
    Caps cp = g.caps;
    bool fresh = !same;
    if (fresh) {
        Caps exact;
        const int mode = o->mode == FT_DIST_POINTS ? 0 : 1;
        const int rc = enqueue_marching(c, s, d_phi, mode, o->want_nearest != 0, o->want_normals != 0,
                                        exact, true, true, &cnt, false);
        if (rc != FT_OK) {
            if (counts) *counts = cnt;
            c->last = cnt;
            return rc;
        }
        cp.X = with_room(exact.X);
        cp.F = with_room(exact.F);
        cp.P = with_room(exact.P);
    }
    for (int attempt = 0; attempt < 4; attempt++) {
        if (fresh || buffer_addresses(c) != g.bufs) FT_TRY(capture_reinit(c, d_phi, d_phi_ifn, o, cp));
        FT_HIP(hipGraphLaunch(g.exec, s));
        g.launches++;
        const int rc = collect(c, s, o->want_nearest != 0, &cnt);
        c->have_normals = rc == FT_OK && o->want_normals != 0;
        c->last_phi = d_phi;
        if (rc != FT_ERR_BUDGET) {
            if (counts) *counts = cnt;
            c->last = cnt;
            return rc;
        }
        // Outgrown: the totals of the stages that ran are this call's.
        const uint64_t why = c->h_scalars[S_CAP];
        cp.X = std::max(cp.X, with_room(cnt.n_if_elems));
        cp.F = std::max(cp.F, with_room(cnt.n_tri));
        if (why & 2ull) cp.P = std::max(cp.P, with_room(cnt.n_ifp));
        fresh = true;
    }
    return ft::set_error(FT_ERR_HIP, "graph reinit: stage capacities did not settle");
}

std::mutex g_create_mu;

// ---- allocation-free reinit ---------------------------------------------------
// Per-call workspaces of ft_reinit_device_noalloc: one file-scope device array
// per GPU, held by every process that loads the library until it ends.  Its
// device address comes from hipGetSymbolAddress, once per device (the host
// name `scratch` is only a stand-in, not an address a kernel may use).
constexpr size_t kScratchBytes = size_t(FT_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;    // per device; null until looked up

int scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // device is current
        if (e != hipSuccess || !p)
            return ft::set_error(FT_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                                 hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return FT_OK;
}

// The handle's per-call buffers (everything but the resident mesh, its query
// slots and tile lists, and the small scalars).
std::vector<DevBuf*> call_buffers(ft_ctx* h)
{
    return {&h->phi_tmp, &h->phi_ifn_tmp, &h->sbits, &h->retain, &h->tbits, &h->blk, &h->xtet,
            &h->xtri_off, &h->xedges, &h->ebits, &h->eprefix, &h->ifp, &h->ifp2e, &h->tris,
            &h->tri_norm, &h->if_elems, &h->p2t_off, &h->p2t_cursor, &h->p2t, &h->ifp_area,
            &h->nearest, &h->fcounts, &h->fclamp, &h->normals, &h->grad_t, &h->phi_ifn_int,
            &h->bin_count, &h->bin_prefix, &h->pcode, &h->prec, &h->trec, &h->pidx, &h->leaf_key,
            &h->leaf_start, &h->leaf_cnt, &h->leaf_b, &h->leaf_parent, &h->bnodes, &h->bflags,
            &h->stats, &h->d_grids, &h->d_cnt_off, &h->leaf_off, &h->roots, &h->scan_bsum,
            &h->scan_side, &h->deep_part};
}

// reinit_device with every per-call buffer taken from the device's scratch:
// the handle's own buffers are set aside for the call and put back after it,
// and the interface, nearest indices and normals of the call are not kept.
int reinit_noalloc(ft_ctx* c, double* d_phi, double* d_phi_ifn, const ft_opts* o, ft_counts* counts,
                   hipStream_t s)
{
    FT_TRY(validate_opts(o));
    if (o->want_nearest || o->want_normals)
        return ft::set_error(FT_ERR_ARG, "ft_reinit_device_noalloc keeps no nearest indices or normals "
                                         "(they would live in the shared scratch)");
    void* base = nullptr;
    FT_TRY(scratch_address(c->device, &base));
    DevBuf saved[64];
    const std::vector<DevBuf*> bufs = call_buffers(c);
    if (bufs.size() > 64) return ft::set_error(FT_ERR_STATE, "too many call buffers");
    for (size_t i = 0; i < bufs.size(); i++) { saved[i] = *bufs[i]; *bufs[i] = DevBuf{}; }
    const uint64_t idx_layout = c->idx_layout;
    const float idx_bs = c->idx_bin_scale;
    const int64_t idx_G = c->idx_G;
    const int idx_bits = c->idx_code_bits;
    c->idx_layout = ~0ull;                           // grids are uploaded into the scratch
    Arena arena;
    arena.base = static_cast<char*>(base);
    arena.size = kScratchBytes;
    t_arena = &arena;
    const int rc = reinit_device(c, d_phi, d_phi_ifn, o, counts, s, false);
    t_arena = nullptr;
    // a call stopped part way (e.g. out of scratch) leaves kernels in flight
    // that write into the scratch: let them end before returning
    const hipError_t es = hipStreamSynchronize(s);
    for (size_t i = 0; i < bufs.size(); i++) *bufs[i] = saved[i];
    c->idx_layout = idx_layout;
    c->idx_bin_scale = idx_bs;
    c->idx_G = idx_G;
    c->idx_code_bits = idx_bits;
    c->have_interface = c->have_nearest = c->have_normals = false;
    // Every buffer a reinit sizes must be in call_buffers; one that is not
    // would keep pointing into the shared scratch after the call.  Such a
    // buffer is dropped (not freed) and the call reports it.
    const char* lo = static_cast<const char*>(base);
    bool escaped = false;
    for (DevBuf* b : all_buffers(c)) {
        const char* p = static_cast<const char*>(b->p);
        if (p && p >= lo && p < lo + kScratchBytes) {
            *b = DevBuf{};
            escaped = true;
        }
    }
    if (escaped) return ft::set_error(FT_ERR_INTERNAL, "a per-call buffer was left in the noalloc scratch");
    if (rc == FT_OK && es != hipSuccess)
        return ft::set_error(FT_ERR_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(es));
    return rc;
}

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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cu_count = cus;
    // A blocking stream: a call that passes no stream (NULL, which is also
    // the address of the null stream) runs on this one, and it must still see
    // the caller's work on the null stream before it (a copy into phi) and
    // come before the caller's null-stream work after it.
    if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess) {
        delete c;
        return ft::set_error(FT_ERR_HIP, "hipStreamCreate failed");
    }
    // the side stream is ordered by events alone (default priority, as the other)
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_join_ev, hipEventDisableTiming) != hipSuccess) {
        if (c->side_fork_ev) (void)hipEventDestroy(c->side_fork_ev);
        if (c->side) (void)hipStreamDestroy(c->side);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return ft::set_error(FT_ERR_HIP, "side stream set-up failed");
    }
    if (ensure(c->scalars, 8 * S_N) != FT_OK ||
        hipHostMalloc((void**)&c->h_scalars, 8 * S_N, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->side);
        (void)hipEventDestroy(c->side_fork_ev);
        (void)hipEventDestroy(c->side_join_ev);
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
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (DevBuf* b : all_buffers(h)) release(*b);
    if (h->h_scalars) (void)hipHostFree(h->h_scalars);
    if (h->h_bcg) (void)hipHostFree(h->h_bcg);
    for (int k = 0; k < 2; k++) {
        if (h->stage[k]) (void)hipHostFree(h->stage[k]);
        if (h->stage_ev[k]) (void)hipEventDestroy(h->stage_ev[k]);
    }
    for (int k = 0; k < 5; k++)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    (void)hipEventDestroy(h->side_fork_ev);
    (void)hipEventDestroy(h->side_join_ev);
    (void)hipStreamDestroy(h->side);
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
    const int64_t no[2] = {0, n_nodes}, zo[2] = {0, 0};
    FT_TRY(set_fields(h, 1, no, zo, zo));
    FT_TRY(ensure(h->xyz, sizeof(float) * 3 * std::max<int64_t>(1, n_nodes)));
    if (n_nodes > 0)
        FT_HIP(hipMemcpyAsync(h->xyz.p, xyz, sizeof(float) * 3 * n_nodes, hipMemcpyHostToDevice, h->stream));
    boxes_from_host(h, xyz);
    return sort_nodes(h);
}

int ft_set_mesh_batch(ft_handle h, int32_t n_fields, const int64_t* node_off, const float* xyz,
                      const int64_t* tet_off, const int32_t* tets, const int64_t* edge_off,
                      const int32_t* edges, const int32_t* t2e)
{
    FT_TRY(check_handle(h));
    if (n_fields < 1 || !node_off || !tet_off || !edge_off || !xyz)
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh_batch: bad arguments");
    if (node_off[0] != 0 || tet_off[0] != 0 || edge_off[0] != 0)
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh_batch: offsets must start at 0");
    for (int k = 0; k < n_fields; k++)
        if (node_off[k + 1] <= node_off[k] || tet_off[k + 1] < tet_off[k] || edge_off[k + 1] < edge_off[k])
            return ft::set_error(FT_ERR_ARG, "ft_set_mesh_batch: field %d has bad offsets", k);
    const int64_t N = node_off[n_fields], T = tet_off[n_fields], E = edge_off[n_fields];
    if (N > INT32_MAX || T > INT32_MAX || E > INT32_MAX)
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh_batch: sizes exceed int32 indices");
    if ((T > 0 && (!tets || !t2e)) || (E > 0 && !edges))
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh_batch: null connectivity");
    // Field-local indices are validated here so that no kernel reads out of bounds.
    for (int k = 0; k < n_fields; k++) {
        const int64_t nk = node_off[k + 1] - node_off[k], ek = edge_off[k + 1] - edge_off[k];
        for (int64_t i = 4 * tet_off[k]; i < 4 * tet_off[k + 1]; i++)
            if (tets[i] < 0 || tets[i] >= nk)
                return ft::set_error(FT_ERR_ARG, "tet %lld (field %d) has node %d out of range",
                                     (long long)(i / 4), k, tets[i]);
        for (int64_t i = 2 * edge_off[k]; i < 2 * edge_off[k + 1]; i++)
            if (edges[i] < 0 || edges[i] >= nk)
                return ft::set_error(FT_ERR_ARG, "edge %lld (field %d) has node %d out of range",
                                     (long long)(i / 2), k, edges[i]);
        for (int64_t i = 6 * tet_off[k]; i < 6 * tet_off[k + 1]; i++)
            if (t2e[i] < 0 || t2e[i] >= ek)
                return ft::set_error(FT_ERR_ARG, "t2e of tet %lld (field %d) has edge %d out of range",
                                     (long long)(i / 6), k, t2e[i]);
    }
    h->has_conn = false;
    h->have_interface = false;
    FT_TRY(set_fields(h, n_fields, node_off, tet_off, edge_off));
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->xyz, sizeof(float) * 3 * N));
    FT_TRY(ensure(h->tets, 16 * std::max<int64_t>(1, T)));
    FT_TRY(ensure(h->edges, 8 * std::max<int64_t>(1, E)));
    FT_TRY(ensure(h->t2e, 24 * std::max<int64_t>(1, T)));
    FT_HIP(hipMemcpyAsync(h->xyz.p, xyz, sizeof(float) * 3 * N, hipMemcpyHostToDevice, s));
    if (T > 0) {
        FT_HIP(hipMemcpyAsync(h->tets.p, tets, 16 * T, hipMemcpyHostToDevice, s));
        FT_HIP(hipMemcpyAsync(h->t2e.p, t2e, 24 * T, hipMemcpyHostToDevice, s));
    }
    if (E > 0) FT_HIP(hipMemcpyAsync(h->edges.p, edges, 8 * E, hipMemcpyHostToDevice, s));
    // field-local -> global indices, in place
    for (int k = 0; k < n_fields; k++) {
        const int64_t nt = tet_off[k + 1] - tet_off[k], ne = edge_off[k + 1] - edge_off[k];
        if (node_off[k] != 0) {
            int32_t* tp = h->tets.as<int32_t>() + 4 * tet_off[k];
            int32_t* ep = h->edges.as<int32_t>() + 2 * edge_off[k];
            if (nt) hipLaunchKernelGGL(k_add_offset, dim3(grid1d(4 * nt)), dim3(kBS), 0, s, tp, 4 * nt, node_off[k], tp);
            if (ne) hipLaunchKernelGGL(k_add_offset, dim3(grid1d(2 * ne)), dim3(kBS), 0, s, ep, 2 * ne, node_off[k], ep);
        }
        if (edge_off[k] != 0 && nt) {
            int32_t* qp = h->t2e.as<int32_t>() + 6 * tet_off[k];
            hipLaunchKernelGGL(k_add_offset, dim3(grid1d(6 * nt)), dim3(kBS), 0, s, qp, 6 * nt, edge_off[k], qp);
        }
    }
    FT_TRY(check_launch());
    boxes_from_host(h, xyz);
    FT_TRY(sort_nodes(h));
    h->has_conn = true;
    return FT_OK;
}

int ft_set_mesh(ft_handle h, const float* xyz, int64_t n_nodes, const int32_t* tets,
                int64_t n_tets, const int32_t* edges, int64_t n_edges, const int32_t* t2e)
{
    FT_TRY(check_handle(h));
    if (n_nodes <= 0 || n_tets < 0 || n_edges < 0)
        return ft::set_error(FT_ERR_ARG, "ft_set_mesh: bad arguments");
    const int64_t no[2] = {0, n_nodes}, to[2] = {0, n_tets}, eo[2] = {0, n_edges};
    return ft_set_mesh_batch(h, 1, no, xyz, to, tets, eo, edges, t2e);
}

int ft_set_kuhn_batch(ft_handle h, int32_t n_fields, const int32_t* dims, const double* origins,
                      const double* spacings, double jitter, uint64_t seed)
{
    FT_TRY(check_handle(h));
    if (n_fields < 1 || !dims || !(jitter >= 0.0 && jitter < 0.5))
        return ft::set_error(FT_ERR_ARG, "ft_set_kuhn_batch: bad arguments");
    std::vector<int64_t> no(n_fields + 1, 0), to(n_fields + 1, 0), eo(n_fields + 1, 0);
    for (int k = 0; k < n_fields; k++) {
        int64_t nn, nt, ne;
        FT_TRY(ft_kuhn_counts(dims[3 * k], dims[3 * k + 1], dims[3 * k + 2], &nn, &nt, &ne));
        no[k + 1] = no[k] + nn;
        to[k + 1] = to[k] + nt;
        eo[k + 1] = eo[k] + ne;
    }
    if (no[n_fields] > INT32_MAX || to[n_fields] > INT32_MAX || eo[n_fields] > INT32_MAX)
        return ft::set_error(FT_ERR_ARG, "ft_set_kuhn_batch: batch exceeds int32 indices");
    h->has_conn = false;
    h->have_interface = false;
    FT_TRY(set_fields(h, n_fields, no.data(), to.data(), eo.data()));
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->xyz, sizeof(float) * 3 * no[n_fields]));
    FT_TRY(ensure(h->tets, 16 * std::max<int64_t>(1, to[n_fields])));
    FT_TRY(ensure(h->edges, 8 * std::max<int64_t>(1, eo[n_fields])));
    FT_TRY(ensure(h->t2e, 24 * std::max<int64_t>(1, to[n_fields])));
    for (int k = 0; k < n_fields; k++) {
        KuhnField f{};
        f.nx = dims[3 * k]; f.ny = dims[3 * k + 1]; f.nz = dims[3 * k + 2];
        const double sp = 1.0 / (std::max(f.nx, std::max(f.ny, f.nz)) - 1);
        for (int q = 0; q < 3; q++) {
            f.origin[q] = origins ? origins[3 * k + q] : 0.0;
            f.spacing[q] = spacings ? spacings[k] : sp;
        }
        f.jitter = jitter;
        f.seed = seed;
        f.node_off = no[k]; f.tet_off = to[k]; f.edge_off = eo[k];
        const int64_t nn = no[k + 1] - no[k], nt = to[k + 1] - to[k];
        hipLaunchKernelGGL(k_kuhn_nodes, dim3(grid1d(nn)), dim3(kBS), 0, s, f, h->xyz.as<float>());
        hipLaunchKernelGGL(k_kuhn_edges, dim3(grid1d(nn)), dim3(kBS), 0, s, f, h->edges.as<int2>());
        hipLaunchKernelGGL(k_kuhn_tets, dim3(grid1d(nt)), dim3(kBS), 0, s, f, h->tets.as<int4>(),
                           h->t2e.as<int32_t>());
        // box: the grid's extent (jitter moves interior nodes only)
        FieldHost& fb = h->fh[k];
        const int d3[3] = {f.nx, f.ny, f.nz};
        for (int q = 0; q < 3; q++) {
            fb.lo[q] = (float)f.origin[q];
            fb.hi[q] = (float)(f.origin[q] + (d3[q] - 1) * f.spacing[q]);
        }
        finish_field_box(fb, nn);
    }
    h->h_est = h->fh[0].h;
    FT_TRY(check_launch());
    FT_TRY(sort_nodes(h));
    h->has_conn = true;
    return FT_OK;
}

int ft_batch_info(ft_handle h, int32_t* n_fields, int64_t* node_off, int64_t* tet_off,
                  int64_t* edge_off)
{
    FT_TRY(check_handle(h));
    if (n_fields) *n_fields = h->nf;
    for (int k = 0; k <= h->nf; k++) {
        if (node_off) node_off[k] = h->node_off[k];
        if (tet_off) tet_off[k] = h->tet_off[k];
        if (edge_off) edge_off[k] = h->edge_off[k];
    }
    return FT_OK;
}

int ft_get_mesh(ft_handle h, float* xyz, int32_t* tets, int32_t* edges, int32_t* t2e)
{
    FT_TRY(check_handle(h));
    if (tets || edges || t2e) FT_TRY(materialize_conn(h));
    hipStream_t s = h->stream;
    if (xyz && h->N) FT_HIP(hipMemcpyAsync(xyz, h->xyz.p, 12 * h->N, hipMemcpyDeviceToHost, s));
    if (tets && h->T) FT_HIP(hipMemcpyAsync(tets, h->tets.p, 16 * h->T, hipMemcpyDeviceToHost, s));
    if (edges && h->E) FT_HIP(hipMemcpyAsync(edges, h->edges.p, 8 * h->E, hipMemcpyDeviceToHost, s));
    if (t2e && h->T) FT_HIP(hipMemcpyAsync(t2e, h->t2e.p, 24 * h->T, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

// ---- resident meshes --------------------------------------------------------
// A resident mesh is a one-field handle kept only for its mesh and its query
// slots.  A batch of resident meshes reads their tets, t2e and edges in place
// and copies only their node coordinates and slots; it holds a reference to
// each, so a mesh released while a batch still refers to it is freed when
// that batch moves on (its id is free at once).
}  // extern "C" (the registry below is C++)

namespace {
// Ids of released meshes are reused; the serial is not (it grows with every
// mesh created in the process), so (id, serial) names one mesh for good.
struct Resident {
    std::shared_ptr<ft_ctx> c;
    int64_t bytes = 0;
    uint64_t serial = 0;
};
std::mutex g_res_mu;
std::vector<Resident> g_res;
uint64_t g_res_serial = 0;

int resident_finish(ft_ctx* c, int32_t* id)
{
    // drop the set-up workspaces, keep mesh + slots
    DevBuf* tmp[] = {&c->bin_count, &c->bin_prefix, &c->pcode, &c->scan_bsum};
    for (DevBuf* b : tmp) release(*b);
    Resident r;
    r.c = std::shared_ptr<ft_ctx>(c, [](ft_ctx* p) { ft_destroy(p); });
    r.bytes = (int64_t)(c->xyz.cap + c->tets.cap + c->edges.cap + c->t2e.cap + c->perm.cap + c->sx.cap +
                        c->sy.cap + c->sz.cap + c->tile_field.cap + c->tile_sched.cap);
    std::lock_guard<std::mutex> lk(g_res_mu);
    r.serial = ++g_res_serial;
    for (size_t i = 0; i < g_res.size(); i++)
        if (!g_res[i].c) { g_res[i] = r; *id = (int32_t)i; return FT_OK; }
    g_res.push_back(r);
    *id = (int32_t)g_res.size() - 1;
    return FT_OK;
}

std::shared_ptr<ft_ctx> resident_get(int32_t id, uint64_t* serial = nullptr)
{
    std::lock_guard<std::mutex> lk(g_res_mu);
    if (id < 0 || id >= (int32_t)g_res.size()) return nullptr;
    if (serial) *serial = g_res[id].serial;
    return g_res[id].c;
}
}  // namespace

extern "C" {

int ft_mesh_create_kuhn(int32_t device, int32_t nx, int32_t ny, int32_t nz, double jitter,
                        uint64_t seed, int32_t* id)
{
    if (!id) return ft::set_error(FT_ERR_ARG, "null output");
    ft_handle h = nullptr;
    FT_TRY(ft_create(device, &h));
    const int32_t dims[3] = {nx, ny, nz};
    int rc = ft_set_kuhn_batch(h, 1, dims, nullptr, nullptr, jitter, seed);
    if (rc) { ft_destroy(h); return rc; }
    return resident_finish(h, id);
}

int ft_mesh_create(int32_t device, const float* xyz, int64_t n_nodes, const int32_t* tets,
                   int64_t n_tets, const int32_t* edges, int64_t n_edges, const int32_t* t2e,
                   int32_t* id)
{
    if (!id) return ft::set_error(FT_ERR_ARG, "null output");
    ft_handle h = nullptr;
    FT_TRY(ft_create(device, &h));
    int rc = ft_set_mesh(h, xyz, n_nodes, tets, n_tets, edges, n_edges, t2e);
    if (rc) { ft_destroy(h); return rc; }
    return resident_finish(h, id);
}

int ft_mesh_info(int32_t id, int64_t* n_nodes, int64_t* n_tets, int64_t* n_edges, int64_t* bytes)
{
    std::lock_guard<std::mutex> lk(g_res_mu);
    if (id < 0 || id >= (int32_t)g_res.size() || !g_res[id].c)
        return ft::set_error(FT_ERR_ARG, "no resident mesh %d", id);
    const ft_ctx* c = g_res[id].c.get();
    if (n_nodes) *n_nodes = c->N;
    if (n_tets) *n_tets = c->T;
    if (n_edges) *n_edges = c->E;
    if (bytes) *bytes = g_res[id].bytes;
    return FT_OK;
}

int ft_mesh_release(int32_t id)
{
    std::shared_ptr<ft_ctx> c;
    {
        std::lock_guard<std::mutex> lk(g_res_mu);
        if (id < 0 || id >= (int32_t)g_res.size() || !g_res[id].c)
            return ft::set_error(FT_ERR_ARG, "no resident mesh %d", id);
        c.swap(g_res[id].c);
        g_res[id] = Resident{};
    }
    return FT_OK;                     // destroyed here unless a batch still refers to it
}

int ft_set_batch_meshes(ft_handle h, int32_t n_fields, const int32_t* mesh_ids)
{
    FT_TRY(check_handle(h));
    if (n_fields < 1 || !mesh_ids) return ft::set_error(FT_ERR_ARG, "ft_set_batch_meshes: bad arguments");
    std::vector<std::shared_ptr<ft_ctx>> ms(n_fields);
    std::vector<uint64_t> serials(n_fields);
    for (int k = 0; k < n_fields; k++) {
        ms[k] = resident_get(mesh_ids[k], &serials[k]);
        if (!ms[k]) return ft::set_error(FT_ERR_ARG, "no resident mesh %d", mesh_ids[k]);
        if (ms[k]->device != h->device)
            return ft::set_error(FT_ERR_ARG, "resident mesh %d lives on another device", mesh_ids[k]);
    }
    // already set from these very meshes: an id alone is not enough, a
    // released id may since name a new mesh of other dimensions
    if (h->has_conn && h->batch_meshes == serials) return FT_OK;
    std::vector<int64_t> no(n_fields + 1, 0), to(n_fields + 1, 0), eo(n_fields + 1, 0);
    for (int k = 0; k < n_fields; k++) {
        no[k + 1] = no[k] + ms[k]->N;
        to[k + 1] = to[k] + ms[k]->T;
        eo[k + 1] = eo[k] + ms[k]->E;
    }
    if (no[n_fields] > INT32_MAX || to[n_fields] > INT32_MAX || eo[n_fields] > INT32_MAX)
        return ft::set_error(FT_ERR_ARG, "batch exceeds int32 indices");
    h->has_conn = false;
    h->have_interface = false;
    FT_TRY(set_fields(h, n_fields, no.data(), to.data(), eo.data()));
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->xyz, 12 * no[n_fields]));
    FT_TRY(ensure(h->perm, 4 * std::max<int64_t>(1, h->n_slots)));
    FT_TRY(ensure(h->sx, 4 * std::max<int64_t>(1, h->n_slots)));
    FT_TRY(ensure(h->sy, 4 * std::max<int64_t>(1, h->n_slots)));
    FT_TRY(ensure(h->sz, 4 * std::max<int64_t>(1, h->n_slots)));
    FT_TRY(ensure(h->tile_field, 4 * std::max<int64_t>(1, h->n_slots / 64)));
    FT_TRY(ensure(h->tile_sched, 16 * std::max<int64_t>(1, h->n_slots / 64)));
    // per-field connectivity pointers: tets, t2e, edges of each mesh
    std::vector<const void*> tab(3 * (size_t)n_fields);
    for (int k = 0; k < n_fields; k++) {
        tab[k] = ms[k]->tets.p;
        tab[n_fields + k] = ms[k]->t2e.p;
        tab[2 * n_fields + k] = ms[k]->edges.p;
    }
    FT_TRY(ensure(h->field_tab, sizeof(void*) * tab.size()));
    FT_HIP(hipMemcpyAsync(h->field_tab.p, tab.data(), sizeof(void*) * tab.size(), hipMemcpyHostToDevice, s));
    FT_TRY(gather_meshes(h, ms, s, false));
    for (int k = 0; k < n_fields; k++) h->fh[k] = ms[k]->fh[0];
    h->h_est = h->fh[0].h;
    FT_HIP(hipStreamSynchronize(s));
    h->field_meshes = std::move(ms);
    h->virt = true;
    h->batch_meshes = serials;
    h->has_conn = true;
    return FT_OK;
}

int ft_get_field_counts(ft_handle h, ft_counts* per_field)
{
    FT_TRY(check_handle(h));
    if (!per_field) return ft::set_error(FT_ERR_ARG, "null output");
    for (int k = 0; k < h->nf; k++)
        per_field[k] = k < (int)h->field_counts.size() ? h->field_counts[k] : ft_counts{};
    return FT_OK;
}

int ft_reinit_batch(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                    ft_counts* per_field, void* stream)
{
    FT_TRY(check_handle(h));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    ft_counts tot{};
    const int rc = reinit_device(h, d_phi, d_phi_ifn, o, &tot, s);
    if (per_field)
        for (int k = 0; k < h->nf && k < (int)h->field_counts.size(); k++) per_field[k] = h->field_counts[k];
    return rc;
}

int ft_reinit_graph(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                    ft_counts* per_field, void* stream)
{
    FT_TRY(check_handle(h));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    ft_counts tot{};
    const int rc = reinit_graph(h, d_phi, d_phi_ifn, o, &tot, s);
    if (per_field)
        for (int k = 0; k < h->nf && k < (int)h->field_counts.size(); k++) per_field[k] = h->field_counts[k];
    return rc;
}

int ft_get_counts(ft_handle h, ft_counts* total)
{
    FT_TRY(check_handle(h));
    if (!total) return ft::set_error(FT_ERR_ARG, "null output");
    if (!h->have_interface) return ft::set_error(FT_ERR_STATE, "no interface: reinit first");
    *total = h->last;
    return FT_OK;
}

int ft_get_normals(ft_handle h, float* normals)
{
    FT_TRY(check_handle(h));
    if (!normals) return ft::set_error(FT_ERR_ARG, "null output");
    if (!h->have_normals) return ft::set_error(FT_ERR_STATE, "no normals: reinit with want_normals first");
    if (h->N > 0) FT_HIP(hipMemcpyAsync(normals, h->normals.p, 12 * h->N, hipMemcpyDeviceToHost, h->stream));
    FT_HIP(hipStreamSynchronize(h->stream));
    return FT_OK;
}

int ft_normals_device(ft_handle h, const float** normals)
{
    FT_TRY(check_handle(h));
    if (!normals) return ft::set_error(FT_ERR_ARG, "null output");
    if (!h->have_normals) return ft::set_error(FT_ERR_STATE, "no normals: reinit with want_normals first");
    *normals = h->normals.as<float>();
    return FT_OK;
}

// LevelSet::getInterfaceCurvature (LevelSet.cpp:215-253) of the last reinit
// into the handle's k_if, on stream s.
int enqueue_curvature(ft_ctx* h, hipStream_t s)
{
    FT_TRY(materialize_conn(h));
    if (!h->have_normals || !h->have_interface || !h->last_phi)
        return ft::set_error(FT_ERR_STATE, "curvature needs a reinit with want_normals first");
    const int64_t N = h->N, T = h->T, P = h->last.n_ifp;
    FT_TRY(ensure(h->gsplit, 24 * at_least1(N)));
    FT_TRY(ensure(h->kmesh, 4 * at_least1(N)));
    FT_TRY(ensure(h->k_if, 4 * at_least1(P)));
    if (N > 0) {
        double* g = h->gsplit.as<double>();
        const uint64_t* sc = h->scalars.as<uint64_t>();
        hipLaunchKernelGGL(k_unit_split, dim3(grid1d(N)), dim3(kBS), 0, s, h->normals.as<float>(), N, g,
                           g + N, g + 2 * N);
        if (T > 0)
            hipLaunchKernelGGL(k_tet_grad_diag, dim3(grid1d(T)), dim3(kBS), 0, s, h->tets.as<int4>(), T,
                               h->xyz.as<float>(), g, g + N, g + 2 * N, h->grad_t.as<float>(), sc);
        hipLaunchKernelGGL(k_node_div, dim3(grid1d(N)), dim3(kBS), 0, s, h->n2t_off.as<uint64_t>(),
                           h->n2t_ids.as<int32_t>(), h->grad_t.as<float>(), N, h->kmesh.as<float>());
        if (P > 0)
            hipLaunchKernelGGL(k_curv_points, dim3(grid1d(P)), dim3(kBS), 0, s, h->ifp2e.as<int32_t>(), P,
                               h->edges.as<int2>(), h->last_phi, h->kmesh.as<float>(), h->k_if.as<float>());
        FT_TRY(check_launch());
    }
    return FT_OK;
}

int ft_interface_curvature(ft_handle h, float* k_if, void* stream)
{
    FT_TRY(check_handle(h));
    if (!k_if) return ft::set_error(FT_ERR_ARG, "null output");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    FT_TRY(enqueue_curvature(h, s));
    const int64_t P = h->last.n_ifp;
    if (P > 0) FT_HIP(hipMemcpyAsync(k_if, h->k_if.p, 4 * P, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

// Flow::surfaceTension_ (src/growth/Flow.cpp:307-341) on device arrays.
int surface_tension(ft_ctx* h, float st, int32_t rho, double* fx, double* fy, double* fz, hipStream_t s)
{
    FT_TRY(materialize_conn(h));
    FT_TRY(enqueue_curvature(h, s));
    const int64_t N = h->N, P = h->last.n_ifp;
    if (N == 0 || P == 0) return FT_OK;
    FT_TRY(ensure(h->st_cnt, 4 * N));
    FT_TRY(ensure(h->st_off, 8 * (N + 1)));
    FT_TRY(ensure(h->st_ids, 4 * 2 * P));
    uint32_t* cnt = h->st_cnt.as<uint32_t>();
    uint64_t* off = h->st_off.as<uint64_t>();
    int32_t* ids = h->st_ids.as<int32_t>();
    const int32_t* ifp2e = h->ifp2e.as<int32_t>();
    const int2* edges = h->edges.as<int2>();
    FT_TRY(fill<uint32_t>(s, cnt, N, 0u));
    hipLaunchKernelGGL(k_st_count, dim3(grid1d(P)), dim3(kBS), 0, s, ifp2e, P, edges, cnt);
    FT_TRY(scan<uint32_t, uint64_t>(h, s, cnt, off, N, off + N, LoadSame{}));
    FT_TRY(fill<uint32_t>(s, cnt, N, 0u));
    hipLaunchKernelGGL(k_st_fill, dim3(grid1d(P)), dim3(kBS), 0, s, ifp2e, P, edges, off, cnt, ids);
    hipLaunchKernelGGL(k_n2t_sort, dim3(grid1d(N)), dim3(kBS), 0, s, off, N, ids);
    hipLaunchKernelGGL(k_st_apply, dim3(grid1d(N)), dim3(kBS), 0, s, off, ids, N, ifp2e, edges,
                       h->last_phi, h->k_if.as<float>(), h->ifp_area.as<float>(), h->normals.as<float>(),
                       st, rho, fx, fy, fz);
    FT_TRY(check_launch());
    return FT_OK;
}

int ft_surface_tension_device(ft_handle h, float st, int32_t rho, double* fx, double* fy, double* fz,
                              void* stream)
{
    FT_TRY(check_handle(h));
    if (h->N > 0 && (!fx || !fy || !fz)) return ft::set_error(FT_ERR_ARG, "null force array");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    FT_TRY(surface_tension(h, st, rho, fx, fy, fz, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_surface_tension(ft_handle h, float st, int32_t rho, double* fx, double* fy, double* fz)
{
    FT_TRY(check_handle(h));
    const int64_t N = h->N;
    if (N > 0 && (!fx || !fy || !fz)) return ft::set_error(FT_ERR_ARG, "null force array");
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->st_f, 24 * at_least1(N)));
    double* d = h->st_f.as<double>();
    double* host[3] = {fx, fy, fz};
    for (int c = 0; c < 3; c++) FT_TRY(copy_h2d(h, d + c * N, host[c], 8 * (size_t)N, s));
    FT_TRY(surface_tension(h, st, rho, d, d + N, d + 2 * N, s));
    return copy_d2h(h, {{fx, d, 8 * (size_t)N}, {fy, d + N, 8 * (size_t)N}, {fz, d + 2 * N, 8 * (size_t)N}}, s);
}

// ---- Advection (LevelSet::update, LevelSet.cpp:89-141) ---------------------
// Quadrature(2, P1) (Quadrature.hpp:23-36) at the float shape functions
// (Tetrahedron.hpp:30-34): the point coordinates go in as float and phi_0 =
// 1.0 - x - y - z comes back rounded to float.
FemQuad fem_quad()
{
    FemQuad q;
    const double a = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, b = (5.0 - std::sqrt(5.0)) / 20.0;
    const double qp[3][4] = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}};
    for (int k = 0; k < 4; k++) {
        const float x = (float)qp[0][k], y = (float)qp[1][k], z = (float)qp[2][k];
        q.phi[0 * 4 + k] = (double)(float)(((1.0 - (double)x) - (double)y) - (double)z);
        q.phi[1 * 4 + k] = (double)x;
        q.phi[2 * 4 + k] = (double)y;
        q.phi[3 * 4 + k] = (double)z;
    }
    q.w = 0.25 / 6.0;
    return q;
}

size_t fem_lds(const ft_ctx* h, int form) { return fem_lds_bytes(h->fem_stride, form); }
unsigned fem_grid(int64_t N) { return (unsigned)((N + kFemRowsPerWave - 1) / kFemRowsPerWave); }

FemEll fem_ell(ft_ctx* h)
{
    return FemEll{h->fem_len.as<int32_t>(), h->fem_soff.as<uint64_t>(), h->fem_cols.as<int32_t>()};
}

// The sliced-ELL pattern and the mass matrix (LevelSet::init,
// LevelSet.cpp:35-39), once per mesh layout.
int fem_setup(ft_ctx* h, hipStream_t s)
{
    FT_TRY(materialize_conn(h));
    if (h->fem_layout == h->layout_gen) return FT_OK;
    if (!h->has_conn) return ft::set_error(FT_ERR_STATE, "advection needs a mesh (ft_set_mesh)");
    FT_TRY(build_node_tets(h, s));
    const int64_t N = h->N, ns = (N + kFemSlice - 1) / kFemSlice;
    FT_TRY(ensure(h->fem_len, 4 * at_least1(N)));
    FT_TRY(ensure(h->fem_w, 4 * at_least1(ns)));
    FT_TRY(ensure(h->fem_soff, 8 * (ns + 1)));
    FT_TRY(ensure(h->fem_flag, 16));
    uint32_t* w = h->fem_w.as<uint32_t>();
    uint64_t* soff = h->fem_soff.as<uint64_t>();
    FT_TRY(fill<unsigned>(s, h->fem_flag.as<unsigned>(), 2, 0u));   // overflow, longest row
    FT_TRY(fill<uint64_t>(s, soff + ns, 1, 0ull));
    if (ns > 0) {
        hipLaunchKernelGGL(k_fem_rowlen, dim3((unsigned)ns), dim3(kFemSlice), 0, s, h->n2t_off.as<uint64_t>(),
                           h->n2t_ids.as<int32_t>(), h->tets.as<int4>(), N, h->fem_len.as<int32_t>(), w,
                           h->fem_flag.as<unsigned>());
        FT_TRY(scan<uint32_t, uint64_t>(h, s, w, soff, ns, soff + ns, LoadSame{}));
    }
    uint64_t total = 0;
    unsigned flag[2] = {0, 0};
    FT_HIP(hipMemcpyAsync(&total, soff + ns, 8, hipMemcpyDeviceToHost, s));
    FT_HIP(hipMemcpyAsync(flag, h->fem_flag.p, 8, hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    if (flag[0])
        return ft::set_error(FT_ERR_ARG, "a node shares tets with more than %d nodes", kFemMaxCols - 1);
    const int64_t nnz = (int64_t)total * kFemSlice;
    FT_TRY(ensure(h->fem_cols, 4 * at_least1(nnz)));
    FT_TRY(ensure(h->fem_m, 8 * at_least1(nnz)));
    FT_TRY(ensure(h->fem_l, 8 * at_least1(nnz)));
    FT_TRY(ensure(h->fem_map, 4 * at_least1(4 * h->T)));             // one word per n2t entry
    h->fem_stride = (int)flag[1] | 1;
    const FemEll e = fem_ell(h);
    if (ns > 0) {
        hipLaunchKernelGGL(k_fem_rowfill, dim3((unsigned)ns), dim3(kFemSlice), 0, s, h->n2t_off.as<uint64_t>(),
                           h->n2t_ids.as<int32_t>(), h->tets.as<int4>(), N, e, w, h->fem_cols.as<int32_t>(),
                           h->fem_map.as<uint32_t>());
        hipLaunchKernelGGL(k_fem_mass, dim3(fem_grid(N)), dim3(64), fem_lds(h, 0), s,
                           h->n2t_off.as<uint64_t>(), h->fem_map.as<uint32_t>(), h->xyz.as<float>(), N, fem_quad(), e, h->fem_stride,
                           h->fem_m.as<double>());
    }
    FT_TRY(check_launch());
    FT_HIP(hipStreamSynchronize(s));
    h->fem_nnz = nnz;
    h->fem_layout = h->layout_gen;
    return FT_OK;
}

int bcg_fin(ft_ctx* h, hipStream_t s, int64_t nb, int step, bool read)
{
    hipLaunchKernelGGL(k_bcg_fin, dim3(1), dim3(256), 0, s, h->bcg_part.as<double>(), nb, step,
                       h->bcg_sc.as<BcgScalars>());
    if (read) {
        FT_HIP(hipMemcpyAsync(h->h_bcg, h->bcg_sc.p, sizeof(BcgScalars), hipMemcpyDeviceToHost, s));
        FT_HIP(hipStreamSynchronize(s));
    }
    return FT_OK;
}

// Crank-Nicolson step of LevelSet::update on device phi (in/out) and flow
// (4 N: velocity x, y, z, pressure), BiCGSTAB to tol in at most max_iters.
int advect(ft_ctx* h, double* d_phi, const double* d_flow, double dt, double tol, int32_t max_iters,
           ft_advect_info* info, hipStream_t s)
{
    *info = ft_advect_info{};
    if (h->nf != 1) return ft::set_error(FT_ERR_STATE, "advection needs a single field");
    if (!h->has_conn) return ft::set_error(FT_ERR_STATE, "advection needs a mesh (ft_set_mesh)");
    const int64_t N = h->N;
    if (dt < FLT_MIN || N == 0) { info->skipped = 1; info->converged = 1; return FT_OK; }   // :92-93
    const int64_t nb = std::min<int64_t>((N + kBcgBS - 1) / kBcgBS, kBcgGrid);
    FT_TRY(ensure(h->bcg_part, 16 * (size_t)nb));
    FT_TRY(ensure(h->bcg_sc, sizeof(BcgScalars)));
    if (!h->h_bcg) FT_HIP(hipHostMalloc(&h->h_bcg, sizeof(BcgScalars), hipHostMallocDefault));
    BcgScalars* hs = static_cast<BcgScalars*>(h->h_bcg);
    double* part = h->bcg_part.as<double>();
    BcgScalars* sc = h->bcg_sc.as<BcgScalars>();
    // flow.norm() < FLT_MIN: nothing to do (:93); NaN velocity: error (:112-115)
    hipLaunchKernelGGL(k_flow_check, dim3((unsigned)nb), dim3(kBcgBS), 0, s, d_flow, N, part);
    FT_TRY(bcg_fin(h, s, nb, BCG_SUMS, true));
    if (std::sqrt(hs->tt) < FLT_MIN) { info->skipped = 1; info->converged = 1; return FT_OK; }
    if (hs->ts > 0.0) return ft::set_error(FT_ERR_ARG, "the velocity has NaN (LevelSet.cpp:112-115)");
    FT_TRY(fem_setup(h, s));
    FT_TRY(ensure(h->bcg_vec, 11 * 8 * (size_t)N));
    double* vec = h->bcg_vec.as<double>();
    double *x = vec, *r = vec + N, *r0 = vec + 2 * N, *p = vec + 3 * N, *v = vec + 4 * N, *t = vec + 5 * N,
           *invdiag = vec + 6 * N, *b = vec + 7 * N, *y = vec + 8 * N, *sv = vec + 9 * N, *z = vec + 10 * N;
    const FemEll e = fem_ell(h);
    double* lv = h->fem_l.as<double>();
    FT_TRY(ensure(h->adv_vel3, 24 * (size_t)N));
    hipLaunchKernelGGL(k_vel_interleave, dim3(grid1d(N)), dim3(kBS), 0, s, d_flow, N, h->adv_vel3.as<double>());
    hipLaunchKernelGGL(k_fem_cn, dim3(fem_grid(N)), dim3(64), fem_lds(h, 1), s, h->n2t_off.as<uint64_t>(),
                       h->fem_map.as<uint32_t>(), h->xyz.as<float>(), h->adv_vel3.as<double>(), N, fem_quad(), e, h->fem_stride,
                       h->fem_m.as<double>(), 0.5 * dt, d_phi, lv, invdiag, b, x, r, r0, p, v);
    hipLaunchKernelGGL(k_dots2, dim3((unsigned)nb), dim3(kBcgBS), 0, s, N, r, r, b, b, part);
    FT_TRY(check_launch());
    FT_TRY(bcg_fin(h, s, nb, BCG_INIT, true));
    const double tol_ = tol > 0.0 ? tol : DBL_EPSILON;                 // Eigen's defaults
    const int64_t max_it = max_iters > 0 ? max_iters : 2 * N;
    if (hs->bb == 0.0) {                                               // zero right-hand side: x = 0
        FT_TRY(fill<double>(s, d_phi, N, 0.0));
        FT_HIP(hipStreamSynchronize(s));
        info->iterations = (int32_t)std::min<int64_t>(max_it, INT32_MAX);
        info->error = tol_;
        info->converged = 1;
        return FT_OK;
    }
    const double tol2 = tol_ * tol_ * hs->bb, eps2 = DBL_EPSILON * DBL_EPSILON;
    int64_t it = 0, restarts = 0;
    const dim3 g((unsigned)nb), bs(kBcgBS);
    while (hs->rr > tol2 && it < max_it) {
        if (std::fabs(hs->rho) < eps2 * hs->r0sq) {
            hipLaunchKernelGGL(k_bcg_restart, g, bs, 0, s, N, e, lv, b, x, r, r0, part);
            FT_TRY(bcg_fin(h, s, nb, BCG_RESTART, false));
            if (restarts++ == 0) it = 0;
        }
        hipLaunchKernelGGL(k_bcg_p, g, bs, 0, s, N, sc, r, v, invdiag, p, y);
        hipLaunchKernelGGL(k_bcg_v, g, bs, 0, s, N, e, lv, y, r0, v, part);
        FT_TRY(bcg_fin(h, s, nb, BCG_ALPHA, false));
        hipLaunchKernelGGL(k_bcg_s, g, bs, 0, s, N, sc, invdiag, r, v, sv, z);
        hipLaunchKernelGGL(k_bcg_t, g, bs, 0, s, N, e, lv, z, sv, t, part);
        FT_TRY(bcg_fin(h, s, nb, BCG_W, false));
        hipLaunchKernelGGL(k_bcg_x, g, bs, 0, s, N, y, sv, z, t, r0, sc, x, r, part);
        FT_TRY(bcg_fin(h, s, nb, BCG_END, true));
        it++;
    }
    FT_TRY(check_launch());
    info->iterations = (int32_t)std::min<int64_t>(it, INT32_MAX);
    info->restarts = (int32_t)restarts;
    info->error = std::sqrt(hs->rr / hs->bb);
    info->converged = info->error <= tol_;
    if (!info->converged)                                              // :135-139, phi kept
        return ft::set_error(FT_ERR_SOLVER, "BiCGSTAB stopped at |r|/|b| = %g > %g after %lld iterations",
                             info->error, tol_, (long long)it);
    FT_HIP(hipMemcpyAsync(d_phi, x, 8 * N, hipMemcpyDeviceToDevice, s));
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_advect_device(ft_handle h, double* d_phi, const double* d_flow, double dt, double tol,
                     int32_t max_iters, ft_advect_info* info, void* stream)
{
    FT_TRY(check_handle(h));
    if (h->N > 0 && (!d_phi || !d_flow)) return ft::set_error(FT_ERR_ARG, "null phi or flow");
    ft_advect_info tmp;
    return advect(h, d_phi, d_flow, dt, tol, max_iters, info ? info : &tmp,
                  stream ? (hipStream_t)stream : h->stream);
}

int ft_advect(ft_handle h, double* phi, const double* flow, double dt, double tol, int32_t max_iters,
              ft_advect_info* info)
{
    FT_TRY(check_handle(h));
    const int64_t N = h->N;
    if (N > 0 && (!phi || !flow)) return ft::set_error(FT_ERR_ARG, "null phi or flow");
    ft_advect_info tmp;
    if (!info) info = &tmp;
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->adv_phi, 8 * at_least1(N)));
    FT_TRY(ensure(h->adv_flow, 32 * at_least1(N)));
    FT_TRY(copy_h2d(h, h->adv_phi.p, phi, 8 * (size_t)N, s));
    FT_TRY(copy_h2d(h, h->adv_flow.p, flow, 32 * (size_t)N, s));
    FT_TRY(advect(h, h->adv_phi.as<double>(), h->adv_flow.as<double>(), dt, tol, max_iters, info, s));
    if (!info->skipped) FT_TRY(copy_d2h(h, {{phi, h->adv_phi.p, 8 * (size_t)N}}, s));
    return FT_OK;
}

int ft_fem_assemble(ft_handle h, int32_t form, const double* vel, int64_t* row_off, int32_t* cols,
                    double* vals)
{
    FT_TRY(check_handle(h));
    if ((form != 0 && form != 1) || !row_off || (cols && !vals) || (form == 1 && cols && !vel))
        return ft::set_error(FT_ERR_ARG, "ft_fem_assemble: bad arguments");
    if (h->nf != 1) return ft::set_error(FT_ERR_STATE, "needs a single field");
    hipStream_t s = h->stream;
    FT_TRY(fem_setup(h, s));
    const int64_t N = h->N;
    FT_TRY(ensure(h->fem_csr_off, 8 * (N + 1)));
    uint64_t* off = h->fem_csr_off.as<uint64_t>();
    FT_TRY(fill<uint64_t>(s, off + N, 1, 0ull));
    if (N > 0) FT_TRY(scan<int32_t, uint64_t>(h, s, h->fem_len.as<int32_t>(), off, N, off + N, LoadSame{}));
    FT_HIP(hipMemcpyAsync(row_off, off, 8 * (N + 1), hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    if (!cols) return FT_OK;                                           // sizes only
    const int64_t nnz = row_off[N];
    const FemEll e = fem_ell(h);
    const double* src = h->fem_m.as<double>();
    if (form == 1) {
        FT_TRY(ensure(h->adv_flow, 32 * at_least1(N)));
        FT_HIP(hipMemcpyAsync(h->adv_flow.p, vel, 24 * N, hipMemcpyHostToDevice, s));
        FT_TRY(ensure(h->adv_vel3, 24 * (size_t)N));
        hipLaunchKernelGGL(k_vel_interleave, dim3(grid1d(N)), dim3(kBS), 0, s, h->adv_flow.as<double>(), N,
                           h->adv_vel3.as<double>());
        hipLaunchKernelGGL(k_fem_adv, dim3(fem_grid(N)), dim3(64), fem_lds(h, 1), s, h->n2t_off.as<uint64_t>(),
                           h->fem_map.as<uint32_t>(), h->xyz.as<float>(), h->adv_vel3.as<double>(), N, fem_quad(), e, h->fem_stride,
                           h->fem_l.as<double>());
        src = h->fem_l.as<double>();
    }
    const int64_t cbytes = (4 * at_least1(nnz) + 255) & ~(int64_t)255;   // CSR staging: cols, vals
    FT_TRY(ensure(h->bcg_vec, cbytes + 8 * at_least1(nnz)));
    int32_t* dc = h->bcg_vec.as<int32_t>();
    double* dv = reinterpret_cast<double*>(h->bcg_vec.as<char>() + cbytes);
    if (N > 0)
        hipLaunchKernelGGL(k_ell_to_csr, dim3(grid1d(N)), dim3(kBS), 0, s, N, e, src, off, dc, dv);
    FT_TRY(check_launch());
    if (nnz > 0) {
        FT_HIP(hipMemcpyAsync(cols, dc, 4 * nnz, hipMemcpyDeviceToHost, s));
        FT_HIP(hipMemcpyAsync(vals, dv, 8 * nnz, hipMemcpyDeviceToHost, s));
    }
    FT_HIP(hipStreamSynchronize(s));
    return FT_OK;
}

int ft_graph_stats(ft_handle h, int64_t* captures, int64_t* launches)
{
    FT_TRY(check_handle(h));
    if (captures) *captures = h->graph.captures;
    if (launches) *launches = h->graph.launches;
    return FT_OK;
}

int ft_reset_phi_points(ft_handle h, const float* ifp, int64_t n_ifp, double* phi_n)
{
    FT_TRY(check_handle(h));
    if (n_ifp < 0 || n_ifp > INT32_MAX || (n_ifp > 0 && !ifp) || (h->N > 0 && !phi_n))
        return ft::set_error(FT_ERR_ARG, "ft_reset_phi_points: bad arguments");
    if (h->nf != 1) return ft::set_error(FT_ERR_STATE, "ft_reset_phi_points needs a single field");
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->phi_tmp, 8 * std::max<int64_t>(1, h->N)));
    FT_TRY(ensure(h->ifp, 12 * std::max<int64_t>(1, n_ifp)));
    FT_TRY(copy_h2d(h, h->ifp.p, ifp, 12 * (size_t)n_ifp, s));
    FT_TRY(copy_h2d(h, h->phi_tmp.p, phi_n, 8 * (size_t)h->N, s));
    h->have_interface = false;
    // the crossing points are given: their count goes where the edge scan
    // would have put it, and the search runs as in a reinit
    Caps cp;
    cp.P = n_ifp;
    SearchSpec sp;
    sp.mode = 0;
    sp.phi_old = h->phi_tmp.as<double>();
    sp.phi_out = h->phi_tmp.as<double>();
    sp.epi = EPI_RAW;
    sp.want_nearest = true;
    uint64_t* sc = h->scalars.as<uint64_t>();
    FT_TRY(index_setup(h, s, 0.0f, sp.mode));
    FT_TRY(ensure_stage(h, 0, cp, 0, true));
    FT_TRY(fill<uint64_t>(s, sc + S_BLK, 1, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_EDGES, 1, (uint64_t)n_ifp));
    FT_TRY(fill<uint64_t>(s, sc + S_ERR, 1, ~0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_CAP, 1, 0ull));
    FT_TRY(fill<uint64_t>(s, sc + S_OVF, 1, 0ull));
    SideJoin join{h, s};
    FT_TRY(enqueue_search(h, s, sp, cp, true, false));
    ft_counts cnt{};
    FT_TRY(collect(h, s, true, &cnt));
    h->have_interface = false;        // points only: no triangles, p2t or areas
    return copy_d2h(h, {{phi_n, h->phi_tmp.p, 8 * (size_t)h->N}}, s);
}

int ft_reinit_device(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                     ft_counts* counts, void* stream)
{
    FT_TRY(check_handle(h));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return reinit_device(h, d_phi, d_phi_ifn, o, counts, s);
}

int ft_reinit_device_noalloc(ft_handle h, double* d_phi, double* d_phi_ifn, const ft_opts* o,
                             ft_counts* counts, void* stream)
{
    FT_TRY(check_handle(h));
    if (!d_phi) return ft::set_error(FT_ERR_ARG, "null phi");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return reinit_noalloc(h, d_phi, d_phi_ifn, o, counts, s);
}

int ft_reinit(ft_handle h, double* phi, double* phi_ifn, const ft_opts* o, ft_counts* counts)
{
    FT_TRY(check_handle(h));
    if (!phi) return ft::set_error(FT_ERR_ARG, "null phi");
    hipStream_t s = h->stream;
    FT_TRY(ensure(h->phi_tmp, 8 * std::max<int64_t>(1, h->N)));
    if (phi_ifn) FT_TRY(ensure(h->phi_ifn_tmp, 8 * std::max<int64_t>(1, h->N)));
    FT_TRY(copy_h2d(h, h->phi_tmp.p, phi, 8 * h->N, s));
    FT_TRY(reinit_device(h, h->phi_tmp.as<double>(), phi_ifn ? h->phi_ifn_tmp.as<double>() : nullptr,
                         o, counts, s));
    std::vector<D2H> out{{phi, h->phi_tmp.p, 8 * (size_t)h->N}};
    if (phi_ifn) out.push_back({phi_ifn, h->phi_ifn_tmp.p, 8 * (size_t)h->N});
    return copy_d2h(h, out, s);
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
        for (int k = 0; k < 5; k++) FT_HIP(hipEventCreate(&h->ev[k]));
    h->timing = on != 0;
    for (int k = 0; k < 4; k++) h->phase_ms[k] = 0.0;
    h->timed_calls = 0;
    return FT_OK;
}

int ft_get_timing(ft_handle h, double* phase_ms, int64_t* calls)
{
    FT_TRY(check_handle(h));
    if (!phase_ms || !calls) return ft::set_error(FT_ERR_ARG, "null output");
    for (int k = 0; k < 4; k++) phase_ms[k] = h->phase_ms[k];
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
