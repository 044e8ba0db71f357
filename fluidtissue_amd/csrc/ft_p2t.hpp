// ft_p2t.hpp -- the per-element steps of the point -> triangle lists and the
// point areas (ifMesh_.p2t and ifpArea_, src/fem/LevelSet.cpp:349-356), as
// plain functions that compile for the device (called by k_p2t_count,
// k_p2t_fill and k_point_area) and for the host (a CPU test drives them over
// a whole interface, planted bad indices included).
//
// Every index that addresses an array is checked first.  A triangle whose
// vertex is not a crossing point of this reinit, a list that would overflow
// its counted length, or a list entry that is not a triangle of this reinit
// sets the error word (code 3, the triangle or point index) and writes
// nothing; the kernels after a set error word do nothing, so phi stays as it
// was and the call returns FT_ERR_INTERNAL.  None of these can happen when
// the stages before are correct; the checks keep a fault in them from
// becoming an out-of-bounds write here.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define FT_HD __host__ __device__ __forceinline__
#else
#define FT_HD inline
#endif

namespace ftp {

// Error code of the stage in the reinit's error word ((code << 40) | index).
constexpr int kErrIndex = 3;

// The three vertices of a triangle ({a, b, c, -1}) are crossing points.
FT_HD bool vertices_ok(const int32_t* tri, int64_t n_points)
{
    return (uint64_t)(int64_t)tri[0] < (uint64_t)n_points && (uint64_t)(int64_t)tri[1] < (uint64_t)n_points &&
           (uint64_t)(int64_t)tri[2] < (uint64_t)n_points;
}

// Degree count of triangle f.  False: a bad vertex, nothing counted.  inc(p)
// adds 1 to *p and returns the old value: an atomic on the device, a plain
// add on the host.
template <class Inc>
FT_HD bool count_tri(const int32_t* tri, int64_t n_points, uint32_t* deg, Inc inc)
{
    if (!vertices_ok(tri, n_points)) return false;
    inc(&deg[tri[0]]);
    inc(&deg[tri[1]]);
    inc(&deg[tri[2]]);
    return true;
}

// Triangle f into the lists of its three vertices at off[v] + cursor[v]++.
// False: a bad vertex, or a list already full (nothing more written).
template <class Inc>
FT_HD bool fill_tri(const int32_t* tri, int32_t f, int64_t n_points, const uint32_t* off,
                    uint32_t* cursor, int32_t* p2t, Inc inc)
{
    if (!vertices_ok(tri, n_points)) return false;
    for (int k = 0; k < 3; k++) {
        const int32_t v = tri[k];
        const uint32_t slot = inc(&cursor[v]);
        if (slot >= off[v + 1] - off[v]) return false;
        p2t[off[v] + slot] = f;
    }
    return true;
}

// Point v: sorts its list ascending (the emission order, the order of p2t in
// the reference) and sums the areas in that order, each step as
// float(double(area) + 0.5 * double(norm)) (LevelSet.cpp:356).  False: an
// entry that is not a triangle of this reinit (area left unwritten).
FT_HD bool point_area(int64_t v, const uint32_t* off, int32_t* p2t, int64_t n_tris,
                      const float* tri_norm, float* area)
{
    const uint32_t b = off[v], e = off[v + 1];
    if (e < b || e - b > (uint64_t)3 * (uint64_t)n_tris) return false;
    for (uint32_t i = b; i < e; i++)
        if ((uint64_t)(int64_t)p2t[i] >= (uint64_t)n_tris) return false;
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
    return true;
}

}  // namespace ftp
