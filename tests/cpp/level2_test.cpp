// A host program that drives the whole reinit through the C ABI the way the
// Level-2 integration of LevelSet::resetPhi_ does (src/fem/LevelSet.cpp:
// 429-455 with the blends of :73-74 and :147-148): a mesh uploaded once, a
// reinit with the init blend, the interface read back, the OBJ written, then a
// reinit with the update blend on the result.  Only ftreinit.h and the C++
// standard library; the mesh and phi come from the library's own generators.
//
//   level2_test <n> <out_dir> <points|triangles>
//
// Writes out_dir/{xyz,tets,edges,t2e,phi0,phi1,ifn1,phi2,ifn2,ifp,ifp2e,tris,
// elems,area,p2t_off,p2t}.bin (raw arrays) and out_dir/shape.{obj,mtl}, and
// prints "N T E P F X" on stdout.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ftreinit.h"

namespace {

int die(const char* what)
{
    std::fprintf(stderr, "%s: %s\n", what, ft_last_error());
    return 1;
}

template <class T>
bool dump(const std::string& dir, const char* name, const std::vector<T>& v)
{
    const std::string path = dir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: level2_test <n> <out_dir> <points|triangles>\n");
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const std::string dir = argv[2];
    const bool tri = std::string(argv[3]) == "triangles";

    int64_t N = 0, T = 0, E = 0;
    if (ft_kuhn_counts(n, n, n, &N, &T, &E)) return die("ft_kuhn_counts");
    std::vector<float> xyz(3 * N);
    std::vector<int32_t> tets(4 * T), edges(2 * E), t2e(6 * T);
    const double origin[3] = {0.0, 0.0, 0.0};
    const double spacing[3] = {1.0 / (n - 1), 1.0 / (n - 1), 1.0 / (n - 1)};   // the unit cube
    if (ft_kuhn_mesh(n, n, n, origin, spacing, 0.2, 5, xyz.data(), tets.data(), edges.data(), t2e.data()))
        return die("ft_kuhn_mesh");
    std::vector<double> phi(N);
    if (ft_phi_field(FT_PHI_SPHERE_WAVE, N, xyz.data(), 0, phi.data())) return die("ft_phi_field");
    const std::vector<double> phi0 = phi;

    // LevelSet::init, after the mesh is built
    ft_handle ft = nullptr;
    if (ft_create(0, &ft)) return die("ft_create");
    if (ft_set_mesh(ft, xyz.data(), N, tets.data(), T, edges.data(), E, t2e.data())) return die("ft_set_mesh");

    // LevelSet::resetPhi_ with the init blend (LevelSet::init, :73-74)
    ft_opts o;
    ft_opts_default(&o);
    o.mode = tri ? FT_DIST_TRIANGLES : FT_DIST_POINTS;
    o.blend = FT_BLEND_INIT;
    std::vector<double> ifn1(N);
    ft_counts c;
    if (ft_reinit(ft, phi.data(), ifn1.data(), &o, &c)) return die("ft_reinit (init)");
    const std::vector<double> phi1 = phi;

    // the interface members LevelSet keeps
    std::vector<float> ifp(3 * c.n_ifp), area(c.n_ifp);
    std::vector<int32_t> ifp2e(c.n_ifp), trisv(4 * c.n_tri), elems(c.n_if_elems), off(c.n_ifp + 1),
        p2t(3 * c.n_tri);
    if (ft_get_interface(ft, ifp.data(), ifp2e.data(), trisv.data(), elems.data(), area.data(), off.data(),
                         p2t.data()))
        return die("ft_get_interface");
    if (ft_write_interface_obj(ft, (dir + "/shape").c_str(), 1)) return die("ft_write_interface_obj");

    // LevelSet::update's reinit with the update blend (:147-148), on phi as it
    // now is
    o.blend = FT_BLEND_UPDATE;
    std::vector<double> ifn2(N);
    ft_counts c2;
    if (ft_reinit(ft, phi.data(), ifn2.data(), &o, &c2)) return die("ft_reinit (update)");
    ft_destroy(ft);

    const bool ok = dump(dir, "xyz", xyz) && dump(dir, "tets", tets) && dump(dir, "edges", edges) &&
                    dump(dir, "t2e", t2e) && dump(dir, "phi0", phi0) && dump(dir, "phi1", phi1) &&
                    dump(dir, "ifn1", ifn1) && dump(dir, "phi2", phi) && dump(dir, "ifn2", ifn2) &&
                    dump(dir, "ifp", ifp) && dump(dir, "ifp2e", ifp2e) && dump(dir, "tris", trisv) &&
                    dump(dir, "elems", elems) && dump(dir, "area", area) && dump(dir, "p2t_off", off) &&
                    dump(dir, "p2t", p2t);
    if (!ok) {
        std::fprintf(stderr, "cannot write into %s\n", dir.c_str());
        return 1;
    }
    std::printf("%lld %lld %lld %lld %lld %lld\n", (long long)N, (long long)T, (long long)E, (long long)c.n_ifp,
                (long long)c.n_tri, (long long)c.n_if_elems);
    return 0;
}
