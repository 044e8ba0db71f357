// ft_obj.cpp -- Wavefront OBJ export: the INTERFACE and BOUNDARY parts of
// cfish::Write_Wavefront_obj (src/leveque/WriteData.cpp:9-87; the interface
// with unit vertex normals, or raw ones as src/growth/WriteData.cpp:48-82)
// and its materials file (src/leveque/WriteData.cpp:89-118).  The text is the
// reference's iostream output (default float formatting), so files compare
// byte for byte.  The boundary triangles are the caller's (the reference's
// mesh.b_tris(), which its mesh builder takes from the Delaunay hull in
// CGAL's order): they are written as given.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "ftreinit.h"
#include "ft_internal.hpp"

extern "C" int ft_write_obj(ft_handle h, const char* fname, int32_t parts, int32_t unit_normals,
                            const int32_t* b_tris, int64_t n_btris)
{
    if (!h || !fname || (parts & ~(FT_OBJ_INTERFACE | FT_OBJ_BOUNDARY)) || n_btris < 0 ||
        (n_btris > 0 && !b_tris))
        return ft::set_error(FT_ERR_ARG, "ft_write_obj: bad arguments");
    // the boundary: its nodes in increasing id (unique_nodes, src/common/
    // Utils.hpp:53-70) and each node's position in that list
    std::vector<float> mesh_xyz;
    std::vector<int32_t> bnodes;
    std::unordered_map<int32_t, int64_t> bpos;
    if (parts & FT_OBJ_BOUNDARY) {
        int64_t noff[2] = {0, 0};
        int32_t nf = 0;
        int rc = ft_batch_info(h, &nf, nullptr, nullptr, nullptr);
        if (rc != FT_OK) return rc;
        if (nf != 1) return ft::set_error(FT_ERR_STATE, "ft_write_obj: the boundary needs a single resident mesh");
        rc = ft_batch_info(h, &nf, noff, nullptr, nullptr);
        if (rc != FT_OK) return rc;
        const int64_t N = noff[1];
        bnodes.assign(b_tris, b_tris + 3 * n_btris);
        for (int32_t v : bnodes)
            if (v < 0 || v >= N) return ft::set_error(FT_ERR_ARG, "ft_write_obj: boundary node %d out of range", v);
        std::sort(bnodes.begin(), bnodes.end());
        bnodes.erase(std::unique(bnodes.begin(), bnodes.end()), bnodes.end());
        for (size_t k = 0; k < bnodes.size(); k++) bpos[bnodes[k]] = (int64_t)k;
        mesh_xyz.resize(3 * (size_t)N);
        rc = ft_get_mesh(h, mesh_xyz.data(), nullptr, nullptr, nullptr);
        if (rc != FT_OK) return rc;
    }
    int64_t P = 0, F = 0;
    std::vector<float> p;
    std::vector<int32_t> t, off, p2t;
    if (parts & FT_OBJ_INTERFACE) {
        ft_counts c{};
        int rc = ft_get_counts(h, &c);
        if (rc != FT_OK) return rc;
        P = c.n_ifp;
        F = c.n_tri;
        p.resize(3 * P);
        t.resize(4 * F);
        off.resize(P + 1);
        p2t.resize(3 * F);
        rc = ft_get_interface(h, p.data(), nullptr, t.data(), nullptr, nullptr, off.data(), p2t.data());
        if (rc != FT_OK) return rc;
    }
    const std::string base(fname);
    std::ofstream out(base + ".obj");
    if (!out.good()) return ft::set_error(FT_ERR_ARG, "cannot write %s.obj", fname);
    out << "mtllib " << base << ".mtl" << "\n\n";
    int64_t offset = 1;                        // OBJ indices start from 1
    if (parts & FT_OBJ_INTERFACE) {
        out << "#\n";
        out << "# shape interface\n";
        out << "#\n\n";
        for (int64_t i = 0; i < P; i++)
            out << "v " << p[3 * i] << " " << p[3 * i + 1] << " " << p[3 * i + 2] << "\n";
        out << "\n";
        // vertex normal: sum over the point's triangles (p2t order) of
        // (p1 - p0) x (p2 - p0) in float, optionally divided by its norm
        for (int64_t i = 0; i < P; i++) {
            float n[3] = {0.0f, 0.0f, 0.0f};
            for (int32_t j = off[i]; j < off[i + 1]; j++) {
                const int32_t* q = &t[4 * (int64_t)p2t[j]];
                const float* a = &p[3 * (int64_t)q[0]];
                const float* b = &p[3 * (int64_t)q[1]];
                const float* d = &p[3 * (int64_t)q[2]];
                const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const float e2[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
                n[0] += e1[1] * e2[2] - e1[2] * e2[1];
                n[1] += e1[2] * e2[0] - e1[0] * e2[2];
                n[2] += e1[0] * e2[1] - e1[1] * e2[0];
            }
            if (unit_normals) {
                const float nrm = std::sqrt(n[0] * n[0] + (n[1] * n[1] + n[2] * n[2]));
                for (int k = 0; k < 3; k++) n[k] /= nrm;
            }
            out << "vn " << n[0] << " " << n[1] << " " << n[2] << "\n";
        }
        out << "\n";
        out << "g interface\n";
        out << "usemtl 0\n";
        for (int64_t f = 0; f < F; f++) {
            const int64_t a = t[4 * f] + offset, b = t[4 * f + 1] + offset, d = t[4 * f + 2] + offset;
            out << "f " << a << "//" << a << " " << b << "//" << b << " " << d << "//" << d << "\n";
        }
        offset += P;
    }
    if (parts & FT_OBJ_BOUNDARY) {
        out << "#\n";
        out << "# domain boundary\n";
        out << "#\n\n";
        for (int32_t v : bnodes)
            out << "v " << mesh_xyz[3 * (size_t)v] << " " << mesh_xyz[3 * (size_t)v + 1] << " "
                << mesh_xyz[3 * (size_t)v + 2] << "\n";
        out << "\n";
        out << "g boundary\n";
        out << "usemtl 1\n";
        for (int64_t f = 0; f < n_btris; f++) {
            const int32_t* q = b_tris + 3 * f;
            out << "f " << offset + bpos[q[0]] << " " << offset + bpos[q[1]] << " " << offset + bpos[q[2]]
                << "\n";
        }
    }
    out.close();
    std::ofstream mtl(base + ".mtl");
    if (!mtl.good()) return ft::set_error(FT_ERR_ARG, "cannot write %s.mtl", fname);
    const char* materials[][6] = {
        {"newmtl 0", "Ka 0 0 0", "Kd 0.6 0.6 0.6", "d 1", "Ks 0 0 0", "Ns 0 interface"},
        {"newmtl 1", "Ka 0 0 0", "Kd 0.0 0.0 1.0", "d 0.2", "Ks 0 0 0", "Ns 0 boundary"},
        {"newmtl 2", "Ka 0 0 0", "Kd 1.0 0.0 0.0", "d 1", "Ks 0 0 0", "Ns 0 knots"}};
    for (int m = 0; m < 3; m++) {
        for (int k = 0; k < 6; k++) mtl << materials[m][k] << "\n";
        if (m == 0) mtl << "\n";
    }
    return mtl.good() ? FT_OK : ft::set_error(FT_ERR_ARG, "cannot write %s.mtl", fname);
}

extern "C" int ft_write_interface_obj(ft_handle h, const char* fname, int32_t unit_normals)
{
    return ft_write_obj(h, fname, FT_OBJ_INTERFACE, unit_normals, nullptr, 0);
}
