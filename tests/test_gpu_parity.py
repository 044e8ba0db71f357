"""GPU parity: the HIP reinit through the C ABI against the CPU oracle."""
import ctypes
import os

import numpy as np
import pytest

import fluidtissue_amd as fa
import oracle
from fluidtissue_amd import _lib
from tests.helpers import (assert_interface_bit_exact, oracle_interface, triangle_errors,
                           workload)

pytestmark = pytest.mark.gpu

FLT_MAX = float(np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def rein():
    r = fa.Reinit(0)
    yield r
    r.close()


def _run(rein, name, **kw):
    m, phi, h = workload(name)
    if getattr(rein, "_mesh_name", None) != name:
        rein.set_mesh(m)
        rein._mesh_name = name
    out, ifn, cnt = rein.reset_phi(phi, **kw)
    return m, phi, h, out, ifn, cnt


@pytest.mark.parametrize("name", ["sphere_wave_32", "star_64", "spheres_96", "ellipsoids", "reference_domain_40",
                                  "slab_48x32x16", "shuffled_32"])
def test_marching_tets_bit_exact(rein, name):
    m, phi, h, out, ifn, cnt = _run(rein, name)
    ref = oracle_interface(name)
    print(f"{name}: P={cnt.n_ifp} F={cnt.n_tri} Tx={cnt.n_if_elems}", flush=True)
    assert (cnt.n_ifp, cnt.n_tri, cnt.n_if_elems) == (len(ref["ifp"]), len(ref["tris"]),
                                                      len(ref["if_elems"]))
    assert_interface_bit_exact(rein.interface(), ref)


def test_marching_tets_bit_exact_star_128(rein):
    m, phi, h, out, ifn, cnt = _run(rein, "star_128")
    assert_interface_bit_exact(rein.interface(), oracle_interface("star_128"))


@pytest.mark.parametrize("name", ["sphere_wave_32", "ellipsoids", "reference_domain_40", "slab_48x32x16",
                                  "shuffled_32"])
def test_points_mode_bit_exact_full(rein, name):
    """Reference semantics (EuclideanDistance.cu:14-35,108-117) on every node."""
    m, phi, h, out, ifn, cnt = _run(rein, name, mode="points", blend="init", want_nearest=True)
    ref = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=0)
    assert np.array_equal(ifn.view(np.uint64), ref["phi_ifn"].view(np.uint64))
    assert np.array_equal(out.view(np.uint64), ref["phi"].view(np.uint64))
    _, near = oracle.reset_phi_points(m.xyz, ref["ifp"], phi, want_nearest=True)
    assert np.array_equal(rein.nearest(), near)
    # LevelSet::update blend (LevelSet.cpp:147-148)
    out_u, ifn_u, _ = rein.reset_phi(phi, mode="points", blend="update")
    ref_u = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=1)
    assert np.array_equal(ifn_u.view(np.uint64), ref_u["phi_ifn"].view(np.uint64))
    assert np.array_equal(out_u.view(np.uint64), ref_u["phi"].view(np.uint64))


@pytest.mark.parametrize("name", ["sphere_wave_32", "ellipsoids"])
def test_update_blend_on_first_and_grown_calls(name):
    """The retained values of the update blend (LevelSet.cpp:443-447, 148) on
    the first call of a fresh handle, and after its buffers were regrown by
    a larger mesh, bitwise against the oracle."""
    small = fa.kuhn_mesh(9)
    m, phi, h = workload(name)
    ref = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=1)
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, _ = r.reset_phi(phi, mode="points", blend="update")
    assert np.array_equal(out.view(np.uint64), ref["phi"].view(np.uint64))
    r.set_mesh(small)
    r.reset_phi(fa.phi_field("star", small.xyz), mode="points", blend="update")
    r.set_mesh(m)
    out2, _, _ = r.reset_phi(phi, mode="points", blend="update")
    assert np.array_equal(out2.view(np.uint64), ref["phi"].view(np.uint64))
    r.close()


@pytest.mark.parametrize("name", ["star_128", "spheres_96"])
def test_points_mode_bit_exact_sampled(rein, name):
    m, phi, h, out, ifn, cnt = _run(rein, name, mode="points", want_nearest=True)
    ref = oracle_interface(name)
    rng = np.random.default_rng(11)
    nodes = rng.choice(m.n_nodes, 20000, replace=False)
    d, j = oracle.point_distance(m.xyz, ref["ifp"], nodes)
    got = ifn[nodes]
    assert np.array_equal(np.abs(got).astype(np.float32).view(np.uint32), d.view(np.uint32))
    assert np.array_equal(np.abs(got), d.astype(np.float64))
    assert np.array_equal(np.signbit(got), phi[nodes] < 0.0)
    assert np.array_equal(rein.nearest()[nodes], j)


@pytest.mark.parametrize("name", ["sphere_wave_32", "ellipsoids", "reference_domain_40", "slab_48x32x16",
                                  "shuffled_32"])
def test_triangle_mode_full(rein, name):
    m, phi, h, out, ifn, cnt = _run(rein, name, mode="triangles", want_nearest=True)
    nodes = np.arange(m.n_nodes)
    emax, erms, ties, mism = triangle_errors(name, np.abs(ifn), rein.nearest(), nodes, h)
    print(f"{name} triangles: max err {emax:.3e} ({emax / h:.2e} h) rms {erms:.3e}, "
          f"eps-ties {ties}, mismatches {mism}", flush=True)
    assert emax <= 1e-5 * h
    assert mism == 0
    assert np.array_equal(np.signbit(ifn), phi < 0.0)
    # |d_tri| <= |d_pt| with the same interface
    _, ifn_p, _ = rein.reset_phi(phi, mode="points")
    assert np.all(np.abs(ifn) <= np.abs(ifn_p) * (1 + 1e-6))


def test_checkerboard_every_tet_crossing(rein):
    """phi = +-1 by node parity: every Kuhn tet has two negative nodes (four
    crossing edges, two triangles) and every axis edge and body diagonal
    crosses, so the edge bitmap's words are dense and each wave deals out
    many more crossing points than it has lanes."""
    n = 20
    m = fa.kuhn_mesh(n)
    h = 1.0 / (n - 1)
    ijk = np.rint(m.xyz.astype(np.float64) / h).astype(np.int64)
    phi = np.where(ijk.sum(axis=1) % 2 == 0, 1.0, -1.0)
    rein.set_mesh(m)
    rein._mesh_name = "checkerboard_20"
    out, ifn, cnt = rein.reset_phi(phi, mode="points", blend="init", want_nearest=True)
    ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
    tris, elems, area = oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
    n_tets = len(m.tets)
    print(f"checkerboard {n}^3: P={cnt.n_ifp} F={cnt.n_tri} Tx={cnt.n_if_elems} of {n_tets} tets",
          flush=True)
    assert cnt.n_if_elems == n_tets and cnt.n_tri == 2 * n_tets
    assert_interface_bit_exact(rein.interface(),
                               dict(ifp=ifp, ifp2e=ifp2e, tris=tris, if_elems=elems, ifp_area=area))
    ref = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=0)
    assert np.array_equal(ifn.view(np.uint64), ref["phi_ifn"].view(np.uint64))
    assert np.array_equal(out.view(np.uint64), ref["phi"].view(np.uint64))
    _, ifn_t, _ = rein.reset_phi(phi, mode="triangles")
    nodes = np.random.default_rng(5).choice(len(m.xyz), 1000, replace=False)
    d64, _ = oracle.triangle_distance(m.xyz, ifp, tris, nodes)
    assert np.max(np.abs(np.abs(ifn_t[nodes]) - d64)) <= 1e-5 * h
    assert np.array_equal(np.signbit(ifn_t), phi < 0.0)


@pytest.mark.parametrize("name", ["star_64", "spheres_96"])
def test_triangle_mode_sampled(rein, name):
    m, phi, h, out, ifn, cnt = _run(rein, name, mode="triangles", want_nearest=True)
    rng = np.random.default_rng(12)
    nodes = rng.choice(m.n_nodes, 3000, replace=False)
    emax, erms, ties, mism = triangle_errors(name, np.abs(ifn[nodes]), rein.nearest()[nodes],
                                             nodes, h)
    print(f"{name} triangles: max err {emax:.3e} ({emax / h:.2e} h) rms {erms:.3e}, "
          f"eps-ties {ties}, mismatches {mism}", flush=True)
    assert emax <= 1e-5 * h
    assert mism == 0


@pytest.mark.parametrize("name,band_h,n_sample", [("spheres_256", 0.0, 2000), ("star_128", 6.0, 3000)])
def test_full_size_workloads(name, band_h, n_sample):
    """The benchmark configurations at full size: 256^3 jittered, 64 spheres,
    full domain; 128^3 star, narrow band 6h.  Interface bit-exact against the
    oracle; distances on a seeded node sample within 1e-5 h of FP64, nearest
    triangle equal outside eps-ties."""
    m, phi, h = workload(name)
    band = band_h * h
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, cnt = r.reset_phi(phi, mode="triangles", band=band, want_nearest=True)
    ref = oracle_interface(name)
    assert_interface_bit_exact(r.interface(), ref)
    rng = np.random.default_rng(14)
    nodes = rng.choice(m.n_nodes, n_sample, replace=False)
    near = r.nearest()[nodes]
    got = np.abs(ifn[nodes])
    r.close()
    if band > 0:
        d64, _ = oracle.triangle_distance(m.xyz, ref["ifp"], ref["tris"], nodes)
        outb = d64 > band * (1 + 1e-6)
        assert np.all(near[outb] == -1) and np.all(got[outb] == np.float64(np.float32(band)))
        keep = d64 < band * (1 - 1e-6)
        nodes, near, got = nodes[keep], near[keep], got[keep]
    emax, erms, ties, mism = triangle_errors(name, got, near, nodes, h, ref=ref)
    print(f"{name} full size triangles: {len(nodes)} nodes checked, max err {emax:.3e} "
          f"({emax / h:.2e} h) rms {erms:.3e} ({erms / h:.2e} h), eps-ties {ties}, "
          f"mismatches {mism}", flush=True)
    assert emax <= 1e-5 * h
    assert mism == 0


def test_points_mode_full_size_sampled():
    """Points mode (the reference's own distance, EuclideanDistance.cu:14-35)
    on the 256^3 workload: a seeded node sample bit-exact against the
    oracle's brute force over all 601k crossing points, nearest index
    included (its leaves use the plane-fit cylinder bounds)."""
    m, phi, h = workload("spheres_256")
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, cnt = r.reset_phi(phi, mode="points", blend="init", want_nearest=True)
    ifp = r.interface().ifp
    near = r.nearest()
    r.close()
    nodes = np.random.default_rng(15).choice(m.n_nodes, 2000, replace=False)
    ref, ref_near = oracle.reset_phi_points(m.xyz[nodes], ifp, phi[nodes], want_nearest=True)
    print(f"spheres_256 points: P={cnt.n_ifp}, {len(nodes)} nodes checked", flush=True)
    assert np.array_equal(ifn[nodes].view(np.uint64), ref.view(np.uint64))
    assert np.array_equal(near[nodes], ref_near)


def test_triangle_mode_band(rein):
    """Narrow band 6h: exact inside, clamped to +-band outside."""
    name = "star_64"
    m, phi, h, _, _, _ = _run(rein, name)
    band = 6 * h
    out, ifn, cnt = rein.reset_phi(phi, mode="triangles", band=band, want_nearest=True)
    rng = np.random.default_rng(13)
    nodes = rng.choice(m.n_nodes, 4000, replace=False)
    itf = oracle_interface(name)
    d64, f64 = oracle.triangle_distance(m.xyz, itf["ifp"], itf["tris"], nodes)
    near = rein.nearest()[nodes]
    got = np.abs(ifn[nodes])
    inb = d64 < band * (1 - 1e-6)
    outb = d64 > band * (1 + 1e-6)
    assert np.all(near[outb] == -1) and np.all(got[outb] == np.float64(np.float32(band)))
    assert np.max(np.abs(got[inb] - d64[inb])) <= 1e-5 * h
    assert cnt.n_clamped == int((rein.nearest() == -1).sum())
    assert cnt.n_clamped > 0


def test_euclidean_distance_abi(rein):
    """EuclideanDistance::init/resetPhi drop-in (host pointers)."""
    m, phi, h = workload("sphere_wave_32")
    ref = oracle_interface("sphere_wave_32")
    ed = fa.EuclideanDistance(0)
    ed.init(m.xyz, phi, m.n_nodes, 8)
    phi_n = phi.copy()
    ed.reset_phi(ref["ifp"], len(ref["ifp"]), phi_n)
    ora, _ = oracle.reset_phi_points(m.xyz, ref["ifp"], phi)
    assert np.array_equal(phi_n.view(np.uint64), ora.view(np.uint64))
    # no interface at all: +-FLT_MAX like the reference's loop over zero points
    phi_n = phi.copy()
    ed.reset_phi(ref["ifp"][:0], 0, phi_n)
    assert np.all(np.abs(phi_n) == FLT_MAX) and np.array_equal(np.signbit(phi_n), phi < 0)
    ed.close()


def test_no_crossing_gives_flt_max(rein):
    m, phi, h = workload("sphere_wave_32")
    rein.set_mesh(m)
    rein._mesh_name = "sphere_wave_32"
    pos = np.abs(phi) + 1.0
    out, ifn, cnt = rein.reset_phi(pos, mode="points")
    assert cnt.n_ifp == 0 and cnt.n_tri == 0
    assert np.all(ifn == FLT_MAX)
    out, ifn, cnt = rein.reset_phi(-pos, mode="triangles")
    assert np.all(ifn == -FLT_MAX)


def test_malformed_t2e_error(rein):
    m = fa.kuhn_mesh(8)
    phi = m.xyz[:, 0].astype(np.float64) - 0.51
    sb = np.signbit(phi)[m.tets].sum(1)
    bad = int(np.nonzero((sb > 0) & (sb < 4))[0][5])
    t2e = m.t2e.copy()
    t2e[bad] = t2e[bad][0]
    r = fa.Reinit(0)
    r.set_mesh(fa.Mesh(m.xyz, m.tets, m.edges, t2e))
    with pytest.raises(_lib.FtError) as ei:
        r.reset_phi(phi)
    assert ei.value.code == _lib.FT_ERR_T2E
    assert r.counts.bad_tet == bad
    r.close()


def test_degenerate_tet_error(rein):
    m = fa.kuhn_mesh(5)
    phi = np.ones(m.n_nodes)
    t = m.tets[17]
    phi[t] = [0.0, -0.0, 0.0, 0.0]
    r = fa.Reinit(0)
    r.set_mesh(m)
    with pytest.raises(_lib.FtError) as ei:
        r.reset_phi(phi)
    assert ei.value.code == _lib.FT_ERR_DEGENERATE
    r.close()


def test_set_mesh_rejects_out_of_range(rein):
    m = fa.kuhn_mesh(4)
    tets = m.tets.copy()
    tets[3, 2] = m.n_nodes
    r = fa.Reinit(0)
    with pytest.raises(_lib.FtError) as ei:
        r.set_mesh(fa.Mesh(m.xyz, tets, m.edges, m.t2e))
    assert ei.value.code == _lib.FT_ERR_ARG
    r.close()


def test_device_resident_matches_host(rein):
    import torch
    m, phi, h, out, ifn, cnt = _run(rein, "star_64", mode="triangles")
    d_phi = torch.from_numpy(phi.copy()).cuda()
    d_ifn = torch.empty_like(d_phi)
    torch.cuda.synchronize()
    for _ in range(2):     # repeated calls reuse the buffers
        d_phi.copy_(torch.from_numpy(phi))
        torch.cuda.synchronize()
        rein.reset_phi_device(d_phi.data_ptr(), d_ifn.data_ptr(), mode="triangles")
        assert np.array_equal(d_phi.cpu().numpy().view(np.uint64), out.view(np.uint64))
        assert np.array_equal(d_ifn.cpu().numpy().view(np.uint64), ifn.view(np.uint64))


def test_triangle_mode_degenerate_triangles(rein):
    """phi = 0 exactly on a grid layer: crossing points coincide with nodes and
    half of the triangles have zero area.  The surface is the plane x = x_k, so
    every node's distance is |x - x_k| in closed form."""
    n = 24
    m = fa.kuhn_mesh(n)
    xk = m.xyz[11 * n * n, 0]                      # node layer i = 11, exact float
    phi = m.xyz[:, 0].astype(np.float64) - float(xk)
    assert (phi == 0.0).sum() == n * n
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, cnt = r.reset_phi(phi, mode="triangles", want_nearest=True)
    itf = r.interface()
    q = itf.ifp[itf.tris[:, :3]].astype(np.float64)
    area2 = np.linalg.norm(np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0]), axis=1)
    assert (area2 == 0).sum() > 100                 # many degenerate triangles
    h = 1.0 / (n - 1)
    exact = np.abs(m.xyz[:, 0].astype(np.float64) - float(xk))
    assert np.max(np.abs(np.abs(ifn) - exact)) <= 1e-5 * h
    d64, f64 = oracle.triangle_distance(m.xyz, itf.ifp, itf.tris)
    assert np.max(np.abs(np.abs(ifn) - d64)) <= 1e-5 * h
    r.close()


def test_cpp_drop_in_link(tmp_path):
    """A reference-style C++ TU links the drop-in class and gets the oracle's
    bits (EuclideanDistance.cu:108-117 semantics) on the device, from the
    object, a copy, a moved object and a copy placed at a reused address."""
    from tests.drop_in import run_link_test
    got, ref = run_link_test(tmp_path, copies=True)
    assert "no HIP device" not in got["stderr"]
    for col in got["phi_cols"]:
        assert np.array_equal(col.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("mode", ["points", "triangles"])
def test_cpp_level2_integration(tmp_path, mode):
    """A C++ program drives LevelSet::resetPhi_'s Level-2 replacement through
    the C ABI alone (mesh once, reinit with the init blend, interface and OBJ
    read back, reinit with the update blend on the result), as documented for
    integrators.  Its interface is the oracle's bit for bit; points mode is
    bitwise the oracle's reinit for both blends; triangle mode is within
    1e-5 h of the FP64 brute force, with the blends and retained values
    bitwise the oracle's epilogue of those distances."""
    from tests.drop_in import run_level2
    from tests.helpers import expected_interface_obj
    n = 24
    r = run_level2(tmp_path, n, mode)
    h = 1.0 / (n - 1)
    xyz, tets, edges, t2e = r["xyz"], r["tets"], r["edges"], r["t2e"]
    ref = oracle.reinit(xyz, tets, edges, t2e, r["phi0"], blend=0)
    assert r["P"] == len(ref["ifp"]) > 0
    assert np.array_equal(r["ifp"].view(np.uint32), ref["ifp"].view(np.uint32))
    assert np.array_equal(r["ifp2e"], ref["ifp2e"]) and np.array_equal(r["tris"], ref["tris"])
    assert np.array_equal(r["elems"], ref["if_elems"])
    assert np.array_equal(r["area"].view(np.uint32), ref["ifp_area"].view(np.uint32))
    off, p2t = oracle.p2t_lists(ref["tris"], len(ref["ifp"]))
    assert np.array_equal(r["p2t_off"], off) and np.array_equal(r["p2t"], p2t)
    base = os.path.join(r["dir"], "shape")
    assert open(base + ".obj").read() == expected_interface_obj(ref["ifp"], ref["tris"], base, True)
    if mode == "points":
        assert np.array_equal(r["ifn1"].view(np.uint64), ref["phi_ifn"].view(np.uint64))
        assert np.array_equal(r["phi1"].view(np.uint64), ref["phi"].view(np.uint64))
        ref2 = oracle.reinit(xyz, tets, edges, t2e, r["phi1"], blend=1)
        assert np.array_equal(r["ifn2"].view(np.uint64), ref2["phi_ifn"].view(np.uint64))
        assert np.array_equal(r["phi2"].view(np.uint64), ref2["phi"].view(np.uint64))
    else:
        d64, _ = oracle.triangle_distance(xyz, ref["ifp"], ref["tris"])
        assert np.max(np.abs(np.abs(r["ifn1"]) - d64)) <= 1e-5 * h
        assert np.array_equal(np.signbit(r["ifn1"]), r["phi0"] < 0.0)
        _, out1 = oracle.finish_reset(r["phi0"], r["ifn1"], ref["ifp2e"], edges, 0)
        assert np.array_equal(r["phi1"].view(np.uint64), out1.view(np.uint64))
        ifp1, ifp2e1 = oracle.find_interface_nodes(edges, xyz, r["phi1"])
        _, out2 = oracle.finish_reset(r["phi1"], r["ifn2"], ifp2e1, edges, 1)
        assert np.array_equal(r["phi2"].view(np.uint64), out2.view(np.uint64))


def test_repeated_calls_outgrowing_capacities():
    """An eager call sized from the previous call's totals that finds a larger
    interface halts on the device, leaves phi alone and reruns exactly: every
    call equals the same call on a fresh handle."""
    m = fa.kuhn_mesh(30, jitter=0.2, seed=8)
    x = m.xyz.astype(np.float64)
    small = np.linalg.norm(x - 0.5, axis=1) - 0.12
    big = fa.phi_field("spheres", m.xyz, seed=3)
    r = fa.Reinit(0)
    r.set_mesh(m)
    for k, phi in enumerate([small, small, big, big, small, big]):
        for mode in ("triangles", "points"):
            out, ifn, cnt = r.reset_phi(phi, mode=mode, blend="update", want_nearest=True)
            f = fa.Reinit(0)
            f.set_mesh(m)
            o2, i2, c2 = f.reset_phi(phi, mode=mode, blend="update", want_nearest=True)
            assert (cnt.n_ifp, cnt.n_tri) == (c2.n_ifp, c2.n_tri), (k, mode)
            assert np.array_equal(out.view(np.uint64), o2.view(np.uint64)), (k, mode)
            assert np.array_equal(ifn.view(np.uint64), i2.view(np.uint64)), (k, mode)
            assert np.array_equal(r.nearest(), f.nearest()), (k, mode)
            f.close()
    r.close()
