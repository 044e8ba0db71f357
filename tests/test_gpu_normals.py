"""GPU: level-set normals (LevelSet::resetPhi_ ends with Assembler::grad_nodes,
LevelSet.cpp:451-453, Assembler.cpp:11-71) against the oracle, bit for bit."""
import numpy as np
import pytest

import fluidtissue_amd as fa
import oracle
from tests.helpers import workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,mode", [("sphere_wave_32", "points"), ("sphere_wave_32", "triangles"),
                                       ("ellipsoids", "triangles"), ("spheres_96", "triangles")])
def test_normals_match_oracle(name, mode):
    m, phi, h = workload(name)
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, _ = r.reset_phi(phi, mode=mode, blend="update", want_normals=True)
    got = r.normals()
    r.close()
    ref = oracle.grad_nodes(m.tets, m.xyz, ifn)           # from the same phi_ifn
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # a distance function has unit gradients away from the interface's kinks
    norm = np.linalg.norm(got.astype(np.float64), axis=1)
    assert np.median(np.abs(norm - 1.0)) < 0.05


def test_normals_without_phi_ifn_buffer():
    """want_normals with no phi_ifn output: the handle keeps phi_ifn itself."""
    import torch
    m, phi, h = workload("ellipsoids")
    r = fa.Reinit(0)
    r.set_mesh(m)
    _, ifn, _ = r.reset_phi(phi, mode="triangles", want_normals=True)
    ref = r.normals()
    d = torch.from_numpy(phi.copy()).cuda()
    r.reset_phi_device(d.data_ptr(), None, mode="triangles", want_normals=True)
    torch.cuda.synchronize()
    assert np.array_equal(r.normals().view(np.uint32), ref.view(np.uint32))
    r.close()


def test_normals_batch_equals_alone_and_graph():
    import torch
    sizes = [20, 27, (24, 18, 30)]
    meshes = [fa.kuhn_mesh(n, jitter=0.2, seed=2) for n in sizes]
    kinds = ["star", "spheres", "ellipsoids"]
    phis = [fa.phi_field(kinds[k], m.xyz, seed=k) for k, m in enumerate(meshes)]
    b = fa.Reinit(0)
    b.set_batch(meshes)
    no, _, _ = b.batch_info()
    b.reset_phi(np.concatenate(phis), mode="triangles", want_normals=True)
    nb = b.normals()
    for k, (m, phi) in enumerate(zip(meshes, phis)):
        a = fa.Reinit(0)
        a.set_mesh(m)
        a.reset_phi(phi, mode="triangles", want_normals=True)
        assert np.array_equal(nb[no[k]:no[k + 1]].view(np.uint32), a.normals().view(np.uint32))
        a.close()
    d = torch.from_numpy(np.concatenate(phis)).cuda()
    src = d.clone()
    for rep in range(3):
        d.copy_(src)
        b.reset_phi_graph_device(d.data_ptr(), None, mode="triangles", want_normals=True)
        torch.cuda.synchronize()
        assert np.array_equal(b.normals().view(np.uint32), nb.view(np.uint32)), rep
    assert b.graph_stats()[0] == 1
    b.close()


def test_normals_need_a_reinit_with_them():
    m, phi, h = workload("sphere_wave_32")
    r = fa.Reinit(0)
    r.set_mesh(m)
    r.reset_phi(phi)
    with pytest.raises(fa._lib.FtError):
        r.normals()
    r.close()


@pytest.mark.parametrize("name,mode,blend", [("sphere_wave_32", "points", "update"), ("ellipsoids", "triangles", "init"),
                                             ("star_64", "triangles", "update")])
def test_interface_curvature_matches_oracle(name, mode, blend):
    """LevelSet::getInterfaceCurvature (LevelSet.cpp:215-253) on the device,
    bit for bit against the oracle from the same reinit outputs."""
    m, phi, h = workload(name)
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, cnt = r.reset_phi(phi, mode=mode, blend=blend, want_normals=True)
    g = r.normals()
    k = r.curvature()
    itf = r.interface()
    r.close()
    ref = oracle.interface_curvature(m.tets, m.xyz, g, out, itf.ifp2e, m.edges)
    assert len(k) == cnt.n_ifp
    assert np.array_equal(k.view(np.uint32), ref.view(np.uint32))


def test_interface_curvature_sphere_on_device():
    m = fa.kuhn_mesh(48, jitter=0.1, seed=5)
    R = 0.27
    phi = np.linalg.norm(m.xyz.astype(np.float64) - 0.5, axis=1) - R
    r = fa.Reinit(0)
    r.set_mesh(m)
    r.reset_phi(phi, mode="triangles", want_normals=True)
    k = r.curvature()
    r.close()
    assert abs(np.median(k) - 2 / R) < 0.05 * 2 / R


@pytest.mark.parametrize("name,mode,blend", [("sphere_wave_32", "points", "update"), ("ellipsoids", "triangles", "init"),
                                             ("star_64", "triangles", "update")])
def test_surface_tension_matches_oracle(name, mode, blend):
    """Flow::surfaceTension_ (src/growth/Flow.cpp:307-341) on the device, added
    into given force arrays, bit for bit against the oracle's serial loop; the
    host and device entry points agree."""
    import ctypes as C
    import torch
    m, phi, h = workload(name)
    r = fa.Reinit(0)
    r.set_mesh(m)
    out, ifn, cnt = r.reset_phi(phi, mode=mode, blend=blend, want_normals=True)
    g, k, itf = r.normals(), r.curvature(), r.interface()
    rng = np.random.default_rng(3)
    f0 = [rng.standard_normal(len(m.xyz)) for _ in range(3)]
    got = r.surface_tension(0.7, 3, *[a.copy() for a in f0])
    d = [torch.from_numpy(a.copy()).cuda() for a in f0]
    fa._lib.check(fa._lib.lib.ft_surface_tension_device(r._h, C.c_float(0.7), C.c_int32(3),
                                                        *[C.c_void_p(t.data_ptr()) for t in d], None))
    r.close()
    ref = oracle.surface_tension(g, out, itf.ifp2e, m.edges, k, itf.ifp_area, 0.7, 3, *f0)
    for a in range(3):
        assert np.array_equal(got[a].view(np.uint64), ref[a].view(np.uint64)), a
        assert np.array_equal(d[a].cpu().numpy().view(np.uint64), ref[a].view(np.uint64)), a
    assert np.any(got[0] != f0[0])


def test_surface_tension_needs_a_reinit_with_normals():
    m, phi, h = workload("sphere_wave_32")
    r = fa.Reinit(0)
    r.set_mesh(m)
    with pytest.raises(fa._lib.FtError) as e:
        r.surface_tension(1.0, 1)
    assert e.value.code == fa._lib.FT_ERR_STATE
    r.reset_phi(phi)                                      # no normals kept
    with pytest.raises(fa._lib.FtError) as e:
        r.surface_tension(1.0, 1)
    assert e.value.code == fa._lib.FT_ERR_STATE
    r.close()


def test_surface_tension_host_arrays_equal_device_arrays():
    """At 128^3 the force arrays (16.8 MB each) go through the pinned staging
    of the host entry point; the result is the device entry point's."""
    import ctypes as C
    import torch
    m, phi, h = workload("star_128")
    r = fa.Reinit(0)
    r.set_mesh(m)
    r.reset_phi(phi, mode="triangles", blend="update", want_normals=True)
    rng = np.random.default_rng(8)
    f0 = [rng.standard_normal(len(m.xyz)) for _ in range(3)]
    got = r.surface_tension(1.3, 2, *[a.copy() for a in f0])
    d = [torch.from_numpy(a.copy()).cuda() for a in f0]
    fa._lib.check(fa._lib.lib.ft_surface_tension_device(r._h, C.c_float(1.3), C.c_int32(2),
                                                        *[C.c_void_p(t.data_ptr()) for t in d], None))
    r.close()
    for a in range(3):
        assert np.array_equal(got[a].view(np.uint64), d[a].cpu().numpy().view(np.uint64))
    assert np.count_nonzero(got[0] != f0[0]) > 1000


@pytest.mark.parametrize("unit", [True, False])
def test_interface_obj_matches_reference_text(tmp_path, unit):
    """The OBJ export of the interface, byte for byte the text the reference's
    writer prints for the oracle's interface."""
    from tests.helpers import EXPECTED_MTL, expected_interface_obj, oracle_interface
    m, phi, h = workload("sphere_wave_32")
    r = fa.Reinit(0)
    r.set_mesh(m)
    r.reset_phi(phi, mode="triangles")
    base = str(tmp_path / "shape")
    r.write_obj(base, unit_normals=unit)
    r.close()
    ref = oracle_interface("sphere_wave_32")
    assert open(base + ".obj").read() == expected_interface_obj(ref["ifp"], ref["tris"], base, unit)
    assert open(base + ".mtl").read() == EXPECTED_MTL


def test_obj_boundary_section_matches_reference_text(tmp_path):
    """INTERFACE | BOUNDARY: the interface text, then the boundary nodes in
    increasing id and the given triangles indexed after the interface's
    points (WriteData.cpp:63-87); BOUNDARY alone indexes from 1; bad node
    ids are refused."""
    from tests.helpers import (EXPECTED_MTL, boundary_faces, expected_boundary_obj, expected_interface_obj,
                               oracle_interface)
    m, phi, h = workload("sphere_wave_32")
    bt = boundary_faces(m.tets)
    assert len(bt) == 12 * 31 * 31                     # two triangles per boundary square
    r = fa.Reinit(0)
    r.set_mesh(m)
    base = str(tmp_path / "only_boundary")
    r.write_obj(base, boundary=bt, interface=False)
    head = f"mtllib {base}.mtl\n\n"
    assert open(base + ".obj").read() == head + expected_boundary_obj(m.xyz, bt, 1)
    assert open(base + ".mtl").read() == EXPECTED_MTL
    r.reset_phi(phi, mode="triangles")
    base = str(tmp_path / "both")
    r.write_obj(base, boundary=bt)
    ref = oracle_interface("sphere_wave_32")
    want = expected_interface_obj(ref["ifp"], ref["tris"], base) + expected_boundary_obj(m.xyz, bt,
                                                                                       1 + len(ref["ifp"]))
    assert open(base + ".obj").read() == want
    bad = bt.copy()
    bad[5, 1] = m.xyz.shape[0]
    with pytest.raises(fa._lib.FtError):
        r.write_obj(str(tmp_path / "bad"), boundary=bad)
    r.close()
