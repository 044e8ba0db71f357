"""GPU: the reinit of a batch as one captured HIP graph (ft_reinit_graph).

Every replay is checked bit for bit against the eager reinit of the same
input, with other eager GPU work and synchronisations between the replays.
"""
import numpy as np
import pytest

import fluidtissue_amd as fa
from fluidtissue_amd import _lib

pytestmark = pytest.mark.gpu

DIMS = [20, 33, (28, 24, 30), 17, 26]


def _batch_phi(r, kind_seed):
    m = r.mesh()
    no, _, _ = r.batch_info()
    kinds = ["ellipsoids", "star", "spheres", "sphere_wave"]
    return np.concatenate([fa.phi_field(kinds[(k + kind_seed) % 4], m.xyz[no[k]:no[k + 1]],
                                        seed=kind_seed * 100 + k) for k in range(len(no) - 1)])


def _eager(r, phi, mode, blend, want_nearest=False):
    import torch
    d = torch.from_numpy(phi.copy()).cuda()
    di = torch.empty_like(d)
    cnts = r.reset_phi_batch_device(d.data_ptr(), di.data_ptr(), mode=mode, blend=blend,
                                    want_nearest=want_nearest)
    torch.cuda.synchronize()
    itf = r.interface()
    near = r.nearest() if want_nearest else None
    return d.cpu().numpy(), di.cpu().numpy(), [(c.n_ifp, c.n_tri, c.n_if_elems) for c in cnts], itf, near


def _same_interface(a, b):
    assert np.array_equal(a.ifp.view(np.uint32), b.ifp.view(np.uint32))
    assert np.array_equal(a.ifp2e, b.ifp2e)
    assert np.array_equal(a.tris, b.tris)
    assert np.array_equal(a.if_elems, b.if_elems)
    assert np.array_equal(a.ifp_area.view(np.uint32), b.ifp_area.view(np.uint32))
    assert np.array_equal(a.p2t_offsets, b.p2t_offsets)
    assert np.array_equal(a.p2t, b.p2t)


@pytest.mark.parametrize("mode,blend", [("triangles", "update"), ("points", "init")])
def test_graph_replays_equal_eager(mode, blend):
    import torch
    r = fa.Reinit(0)
    r.set_kuhn_batch(DIMS, jitter=0.15, seed=4)
    phi = _batch_phi(r, 1)
    ref_out, ref_ifn, ref_cnt, ref_itf, ref_near = _eager(r, phi, mode, blend, want_nearest=True)
    other = fa.Reinit(0)                       # eager work between the replays
    other.set_kuhn_batch([24, 18])
    other_phi = _batch_phi(other, 2)
    d = torch.empty(len(phi), dtype=torch.float64, device="cuda")
    di = torch.empty_like(d)
    src = torch.from_numpy(phi).cuda()
    for rep in range(4):
        d.copy_(src)
        di.fill_(7.0)
        cnts = r.reset_phi_graph_device(d.data_ptr(), di.data_ptr(), mode=mode, blend=blend,
                                        want_nearest=True)
        torch.cuda.synchronize()
        assert [(c.n_ifp, c.n_tri, c.n_if_elems) for c in cnts] == ref_cnt, rep
        assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64)), rep
        assert np.array_equal(di.cpu().numpy().view(np.uint64), ref_ifn.view(np.uint64)), rep
        _same_interface(r.interface(), ref_itf)
        assert np.array_equal(r.nearest(), ref_near)
        other.reset_phi(other_phi, mode="triangles")
        torch.cuda.synchronize()
    captures, launches = r.graph_stats()
    assert captures == 1 and launches == 4
    other.close()
    r.close()


@pytest.mark.parametrize("mode", ["triangles", "points"])
def test_graph_split_deep_tiles(mode):
    """A batch holding a field with very deep tiles (the 64^3 star reaches about
    -50 h, where tiles are searched by several waves and merged): the eager
    call equals the work-queue schedule, which searches every tile whole, and
    the captured graph replays it bit for bit."""
    import torch
    r = fa.Reinit(0)
    r.set_kuhn_batch([64, 20, 33])
    m = r.mesh()
    no, _, _ = r.batch_info()
    phi = np.concatenate([fa.phi_field("star", m.xyz[no[0]:no[1]]),
                          fa.phi_field("sphere_wave", m.xyz[no[1]:no[2]]),
                          fa.phi_field("spheres", m.xyz[no[2]:no[3]], seed=3)])
    assert phi[:no[1]].min() <= -25.0 / 63
    ref_out, ref_ifn, ref_cnt, ref_itf, ref_near = _eager(r, phi, mode, "update", want_nearest=True)
    d = torch.from_numpy(phi.copy()).cuda()
    di = torch.empty_like(d)
    r.reset_phi_batch_device(d.data_ptr(), di.data_ptr(), mode=mode, blend="update", want_nearest=True,
                             schedule="queue")
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64))
    assert np.array_equal(di.cpu().numpy().view(np.uint64), ref_ifn.view(np.uint64))
    assert np.array_equal(r.nearest(), ref_near)
    other = fa.Reinit(0)                       # eager work between the replays
    other.set_kuhn_batch([24, 18])
    other_phi = _batch_phi(other, 2)
    src = torch.from_numpy(phi).cuda()
    for rep in range(3):
        d.copy_(src)
        di.fill_(7.0)
        cnts = r.reset_phi_graph_device(d.data_ptr(), di.data_ptr(), mode=mode, blend="update",
                                        want_nearest=True)
        torch.cuda.synchronize()
        assert [(c.n_ifp, c.n_tri, c.n_if_elems) for c in cnts] == ref_cnt, rep
        assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64)), rep
        assert np.array_equal(di.cpu().numpy().view(np.uint64), ref_ifn.view(np.uint64)), rep
        assert np.array_equal(r.nearest(), ref_near), rep
        other.reset_phi(other_phi, mode="triangles")
        torch.cuda.synchronize()
    other.close()
    r.close()


@pytest.mark.parametrize("mode", ["triangles", "points"])
def test_graph_deep_tiles_first(mode):
    """Fields with nodes deep inside the surface (a sphere of radius 0.4 on
    40^3 and 33^3 grids: nodes down to ~15h inside), so the query's schedule
    lists deep tiles first: the captured graph, replayed, equals the eager
    reinit bit for bit, and the batch equals each field alone."""
    import torch
    dims = [40, 33]
    r = fa.Reinit(0)
    r.set_kuhn_batch(dims)
    m = r.mesh()
    no, _, _ = r.batch_info()
    phi = np.concatenate([np.linalg.norm(m.xyz[no[k]:no[k + 1]].astype(np.float64) - 0.5, axis=1) - 0.4
                          for k in range(len(dims))])
    assert np.min(phi[:no[1]]) <= -10.0 / (dims[0] - 1)            # deep tiles exist
    ref_out, ref_ifn, ref_cnt, ref_itf, ref_near = _eager(r, phi, mode, "init", want_nearest=True)
    if mode == "points":                       # every node written: against the oracle, bitwise
        import oracle
        for k in range(len(dims)):
            sl = slice(no[k], no[k + 1])
            n = dims[k]
            mk = fa.kuhn_mesh(n)
            o = oracle.reinit(mk.xyz, mk.tets, mk.edges, mk.t2e, phi[sl], blend=0)
            assert np.array_equal(ref_ifn[sl].view(np.uint64), o["phi_ifn"].view(np.uint64)), k
    d = torch.empty(len(phi), dtype=torch.float64, device="cuda")
    di = torch.empty_like(d)
    src = torch.from_numpy(phi).cuda()
    for rep in range(4):
        d.copy_(src)
        di.fill_(7.0)
        cnts = r.reset_phi_graph_device(d.data_ptr(), di.data_ptr(), mode=mode, blend="init",
                                        want_nearest=True)
        torch.cuda.synchronize()
        assert [(c.n_ifp, c.n_tri, c.n_if_elems) for c in cnts] == ref_cnt, rep
        assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64)), rep
        assert np.array_equal(di.cpu().numpy().view(np.uint64), ref_ifn.view(np.uint64)), rep
        assert np.array_equal(r.nearest(), ref_near)
    r.close()
    for k, n in enumerate(dims):                                     # each field alone
        one = fa.Reinit(0)
        one.set_kuhn_batch([n])
        assert np.array_equal(one.mesh().xyz, m.xyz[no[k]:no[k + 1]])
        out1, ifn1, _, _, _ = _eager(one, phi[no[k]:no[k + 1]], mode, "init")
        assert np.array_equal(out1.view(np.uint64), ref_out[no[k]:no[k + 1]].view(np.uint64))
        one.close()


def test_graph_regrows_and_matches_eager():
    """A replay whose interface outgrows the captured headroom is captured
    again and still gives the eager result; a smaller one reuses the graph."""
    import torch
    r = fa.Reinit(0)
    r.set_kuhn_batch([30, 30, 30])
    m = r.mesh()
    c = np.array([0.5, 0.5, 0.5])
    small = np.linalg.norm(m.xyz.astype(np.float64) - c, axis=1) - 0.1        # one small sphere
    big = _batch_phi(r, 2)                                                   # many features
    d = torch.empty(len(small), dtype=torch.float64, device="cuda")
    for k, phi in enumerate([small, big, small, big]):
        ref_out, ref_ifn, ref_cnt, ref_itf, _ = _eager(r, phi, "triangles", "update")
        d.copy_(torch.from_numpy(phi))
        cnts = r.reset_phi_graph_device(d.data_ptr(), None, mode="triangles", blend="update")
        torch.cuda.synchronize()
        assert [(x.n_ifp, x.n_tri, x.n_if_elems) for x in cnts] == ref_cnt, k
        assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64)), k
        _same_interface(r.interface(), ref_itf)
    captures, launches = r.graph_stats()
    # eager calls between the replays do not move the buffers the graph uses
    # except when they grow them; the big interface forces at least one capture
    assert captures >= 2 and launches >= 4
    r.close()


def test_graph_degenerate_tet_leaves_phi():
    """A replay that hits a degenerate crossing tet reports it and writes
    nothing: the query is halted on the device."""
    import torch
    r = fa.Reinit(0)
    r.set_kuhn_batch([16, 16])
    m = r.mesh()
    phi = _batch_phi(r, 0)
    d = torch.from_numpy(phi.copy()).cuda()
    r.reset_phi_graph_device(d.data_ptr(), None, mode="triangles")
    bad = phi.copy()
    t = m.tets[40]
    bad[t] = [0.0, -0.0, 0.0, 0.0]
    d.copy_(torch.from_numpy(bad))
    with pytest.raises(_lib.FtError) as ei:
        r.reset_phi_graph_device(d.data_ptr(), None, mode="triangles")
    assert ei.value.code == _lib.FT_ERR_DEGENERATE
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint64), bad.view(np.uint64))
    # and the graph still works afterwards
    d.copy_(torch.from_numpy(phi))
    ref_out = _eager(r, phi, "triangles", "init")[0]
    r.reset_phi_graph_device(d.data_ptr(), None, mode="triangles")
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint64), ref_out.view(np.uint64))
    r.close()
