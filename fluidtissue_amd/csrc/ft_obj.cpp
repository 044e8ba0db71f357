// ft_obj.cpp -- Wavefront OBJ export of the reconstructed interface, the
// INTERFACE part of cfish::Write_Wavefront_obj (src/leveque/WriteData.cpp:9-61
// with unit vertex normals, src/growth/WriteData.cpp:48-82 with raw ones) and
// its materials file (src/leveque/WriteData.cpp:89-118).  The text is the
// reference's iostream output (default float formatting), so files compare
// byte for byte.
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "ftreinit.h"
#include "ft_internal.hpp"

extern "C" int ft_write_interface_obj(ft_handle h, const char* fname, int32_t unit_normals)
{
    if (!h || !fname) return ft::set_error(FT_ERR_ARG, "ft_write_interface_obj: bad arguments");
    ft_counts c{};
    int rc = ft_get_counts(h, &c);
    if (rc != FT_OK) return rc;
    const int64_t P = c.n_ifp, F = c.n_tri;
    std::vector<float> p(3 * P);
    std::vector<int32_t> t(4 * F), off(P + 1), p2t(3 * F);
    rc = ft_get_interface(h, p.data(), nullptr, t.data(), nullptr, nullptr, off.data(), p2t.data());
    if (rc != FT_OK) return rc;
    const std::string base(fname);
    std::ofstream out(base + ".obj");
    if (!out.good()) return ft::set_error(FT_ERR_ARG, "cannot write %s.obj", fname);
    out << "mtllib " << base << ".mtl" << "\n\n";
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
    for (int64_t f = 0; f < F; f++) {          // OBJ indices start from 1
        const int32_t a = t[4 * f] + 1, b = t[4 * f + 1] + 1, d = t[4 * f + 2] + 1;
        out << "f " << a << "//" << a << " " << b << "//" << b << " " << d << "//" << d << "\n";
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
