// CPU build of the point -> triangle stage (fluidtissue_amd/csrc/ft_p2t.hpp,
// the steps k_p2t_count, k_p2t_fill and k_point_area run per element).
// Driven over whole interfaces on the host, built with AddressSanitizer by
// the test that runs it, so any write out of bounds aborts the run:
//   1. a valid interface: lists equal the ascending triangle ids of each
//      point, areas the float sums in that order;
//   2. a triangle with a vertex out of range (P, -1, INT32_MAX), planted in
//      the count and, separately, in the fill only: the step reports it and
//      writes nothing;
//   3. a list that is already full: the fill reports it and stops;
//   4. a list entry that is not a triangle: the area step reports it and
//      leaves the area unwritten.
// Prints "ok" and exits 0 when every check holds.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ft_p2t.hpp"

namespace {

struct Inc {
    uint32_t operator()(uint32_t* p) const { return (*p)++; }
};

int fails = 0;
#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            fails++;                                              \
        }                                                         \
    } while (0)

struct Itf {
    int64_t P = 0;
    std::vector<int32_t> tris;   // F x 4 {a, b, c, -1}
    std::vector<float> norm;     // F
    int64_t F() const { return (int64_t)tris.size() / 4; }
};

Itf random_interface(int64_t P, int64_t F, unsigned seed)
{
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int32_t> pick(0, (int32_t)P - 1);
    std::uniform_real_distribution<float> len(0.0f, 1.0f);
    Itf it;
    it.P = P;
    for (int64_t f = 0; f < F; f++) {
        int32_t a = pick(rng), b = pick(rng), c = pick(rng);
        while (b == a) b = pick(rng);
        while (c == a || c == b) c = pick(rng);
        it.tris.insert(it.tris.end(), {a, b, c, -1});
        it.norm.push_back(len(rng));
    }
    return it;
}

// The device pipeline on the host: count, exclusive scan, fill, area.  The
// first failing step stops the pipeline, as the error word halts the
// kernels after it.  Returns false when a step reported a bad index.
bool run(const Itf& it, std::vector<uint32_t>& off, std::vector<int32_t>& p2t, std::vector<float>& area)
{
    const int64_t P = it.P, F = it.F();
    std::vector<uint32_t> deg(P, 0u), cursor(P, 0u);
    for (int64_t f = 0; f < F; f++)
        if (!ftp::count_tri(&it.tris[4 * f], P, deg.data(), Inc{})) return false;
    off.assign(P + 1, 0u);
    for (int64_t v = 0; v < P; v++) off[v + 1] = off[v] + deg[v];
    p2t.assign(off[P], -7);
    for (int64_t f = 0; f < F; f++)
        if (!ftp::fill_tri(&it.tris[4 * f], (int32_t)f, P, off.data(), cursor.data(), p2t.data(), Inc{}))
            return false;
    area.assign(P, -1.0f);
    for (int64_t v = 0; v < P; v++)
        if (!ftp::point_area(v, off.data(), p2t.data(), F, it.norm.data(), area.data())) return false;
    return true;
}

}  // namespace

int main()
{
    // 1. valid interfaces
    for (unsigned seed = 1; seed <= 5; seed++) {
        const Itf it = random_interface(40 + seed * 13, 200 + seed * 37, seed);
        std::vector<uint32_t> off;
        std::vector<int32_t> p2t;
        std::vector<float> area;
        CHECK(run(it, off, p2t, area));
        for (int64_t v = 0; v < it.P; v++) {
            std::vector<int32_t> want;
            for (int64_t f = 0; f < it.F(); f++)
                for (int k = 0; k < 3; k++)
                    if (it.tris[4 * f + k] == v) want.push_back((int32_t)f);
            CHECK(std::vector<int32_t>(p2t.begin() + off[v], p2t.begin() + off[v + 1]) == want);
            float a = 0.0f;
            for (int32_t f : want) a = (float)((double)a + 0.5 * (double)it.norm[f]);
            CHECK(area[v] == a);
        }
    }
    // 2a. bad vertex seen by the count: nothing counted for it, pipeline stops
    for (int32_t bad : {40, -1, INT32_MAX}) {
        Itf it = random_interface(40, 120, 9);
        it.tris[4 * 77 + 1] = bad;
        std::vector<uint32_t> off;
        std::vector<int32_t> p2t;
        std::vector<float> area;
        CHECK(!run(it, off, p2t, area));
        std::vector<uint32_t> deg(it.P, 0u);
        CHECK(!ftp::count_tri(&it.tris[4 * 77], it.P, deg.data(), Inc{}));
        CHECK(std::all_of(deg.begin(), deg.end(), [](uint32_t d) { return d == 0u; }));
    }
    // 2b. bad vertex seen only by the fill (the count was of a good interface)
    {
        const Itf good = random_interface(40, 120, 11);
        std::vector<uint32_t> off;
        std::vector<int32_t> p2t;
        std::vector<float> area;
        CHECK(run(good, off, p2t, area));
        std::vector<uint32_t> cursor(good.P, 0u);
        std::vector<int32_t> out(p2t.size(), -7);
        const int32_t tri[4] = {3, 40, 5, -1};
        CHECK(!ftp::fill_tri(tri, 0, good.P, off.data(), cursor.data(), out.data(), Inc{}));
        CHECK(std::all_of(out.begin(), out.end(), [](int32_t x) { return x == -7; }));
    }
    // 3. a full list: the fill stops at it
    {
        const Itf it = random_interface(30, 90, 13);
        std::vector<uint32_t> off;
        std::vector<int32_t> p2t;
        std::vector<float> area;
        CHECK(run(it, off, p2t, area));
        std::vector<uint32_t> cursor(it.P, 0u);
        const int32_t v = it.tris[0];
        cursor[v] = off[v + 1] - off[v];              // v's list already holds its count
        std::vector<int32_t> out(p2t.size(), -7);
        CHECK(!ftp::fill_tri(&it.tris[0], 0, it.P, off.data(), cursor.data(), out.data(), Inc{}));
        for (uint32_t i = off[v]; i < off[v + 1]; i++) CHECK(out[i] == -7);
    }
    // 4. a list entry that is not a triangle
    {
        const Itf it = random_interface(30, 90, 17);
        std::vector<uint32_t> off;
        std::vector<int32_t> p2t;
        std::vector<float> area;
        CHECK(run(it, off, p2t, area));
        int64_t v = 0;
        while (off[v + 1] == off[v]) v++;
        for (int32_t bad : {(int32_t)it.F(), -1}) {
            p2t[off[v]] = bad;
            area[v] = -1.0f;
            CHECK(!ftp::point_area(v, off.data(), p2t.data(), it.F(), it.norm.data(), area.data()));
            CHECK(area[v] == -1.0f);
        }
    }
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
