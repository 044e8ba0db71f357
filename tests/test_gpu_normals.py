"""GPU: level-set normals (LevelSet::resetPhi_ ends with Assembler::grad_nodes,
LevelSet.cpp:451-453, Assembler.cpp:11-71) against the oracle, bit for bit."""
import numpy as np
import pytest

import fluidtissue_amd as fa
import oracle
from tests.helpers import workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,mode", [("cfg1", "points"), ("cfg1", "triangles"),
                                       ("ellipsoids", "triangles"), ("cfg3s", "triangles")])
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
    m, phi, h = workload("cfg1")
    r = fa.Reinit(0)
    r.set_mesh(m)
    r.reset_phi(phi)
    with pytest.raises(fa._lib.FtError):
        r.normals()
    r.close()
