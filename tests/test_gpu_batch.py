"""GPU: device mesh generation and batches of independent fields."""
import numpy as np
import pytest

import fluidtissue_amd as fa
import oracle

pytestmark = pytest.mark.gpu

KINDS = ["sphere_wave", "ellipsoids", "star", "spheres"]


def _local_interface(itf, cnts, k, node_off, edge_off):
    """Field k's part of a batched interface, renumbered field-locally."""
    p0 = sum(c.n_ifp for c in cnts[:k])
    f0 = sum(c.n_tri for c in cnts[:k])
    x0 = sum(c.n_if_elems for c in cnts[:k])
    P, F, X = cnts[k].n_ifp, cnts[k].n_tri, cnts[k].n_if_elems
    tris = itf.tris[f0:f0 + F].copy()
    tris[:, :3] -= p0
    return dict(ifp=itf.ifp[p0:p0 + P], ifp2e=itf.ifp2e[p0:p0 + P] - edge_off[k], tris=tris,
                ifp_area=itf.ifp_area[p0:p0 + P], if_elems_n=X)


def test_device_kuhn_matches_host():
    dims = [(9, 7, 5), (12, 12, 12), (6, 11, 8)]
    r = fa.Reinit(0)
    r.set_kuhn_batch(dims, jitter=0.2, seed=3)
    got = r.mesh()
    no, to, eo = r.batch_info()
    for k, d in enumerate(dims):
        ref = fa.kuhn_mesh(d, jitter=0.2, seed=3)
        assert np.array_equal(got.xyz[no[k]:no[k + 1]].view(np.uint32), ref.xyz.view(np.uint32))
        assert np.array_equal(got.tets[to[k]:to[k + 1]] - no[k], ref.tets)
        assert np.array_equal(got.edges[eo[k]:eo[k + 1]] - no[k], ref.edges)
        assert np.array_equal(got.t2e[to[k]:to[k + 1]] - eo[k], ref.t2e)
    r.close()


@pytest.mark.parametrize("mode", ["points", "triangles"])
def test_batch_equals_alone(mode):
    """Every field of a batch gets exactly what a reinit of it alone gives."""
    sizes = [20, 28, (33, 30, 26), 24, 17]
    meshes = [fa.kuhn_mesh(n) for n in sizes]
    phis = [fa.phi_field(KINDS[k % 4], m.xyz, seed=k + 1) for k, m in enumerate(meshes)]
    b = fa.Reinit(0)
    b.set_batch(meshes)
    no, to, eo = b.batch_info()
    out, ifn, _ = b.reset_phi(np.concatenate(phis), mode=mode, blend="update", want_nearest=True)
    cnts = b.field_counts()
    itf = b.interface()
    near = b.nearest()
    for k, (m, phi) in enumerate(zip(meshes, phis)):
        a = fa.Reinit(0)
        a.set_mesh(m)
        o1, i1, c1 = a.reset_phi(phi, mode=mode, blend="update", want_nearest=True)
        assert (cnts[k].n_ifp, cnts[k].n_tri, cnts[k].n_if_elems) == (c1.n_ifp, c1.n_tri, c1.n_if_elems)
        sl = slice(no[k], no[k + 1])
        assert np.array_equal(out[sl].view(np.uint64), o1.view(np.uint64))
        assert np.array_equal(ifn[sl].view(np.uint64), i1.view(np.uint64))
        loc = _local_interface(itf, cnts, k, no, eo)
        ia = a.interface()
        assert np.array_equal(loc["ifp"].view(np.uint32), ia.ifp.view(np.uint32))
        assert np.array_equal(loc["ifp2e"], ia.ifp2e)
        assert np.array_equal(loc["tris"], ia.tris)
        assert np.array_equal(loc["ifp_area"].view(np.uint32), ia.ifp_area.view(np.uint32))
        # nearest primitive, renumbered locally
        off = sum(c.n_ifp if mode == "points" else c.n_tri for c in cnts[:k])
        nk = near[sl].copy()
        nk[nk >= 0] -= off
        assert np.array_equal(nk, a.nearest())
        a.close()
    b.close()


def test_batch_many_fields_against_oracle():
    """64 small fields in one pass; points mode against the oracle per field."""
    rng = np.random.default_rng(5)
    sizes = rng.integers(8, 21, size=64)
    b = fa.Reinit(0)
    b.set_kuhn_batch([int(n) for n in sizes])
    m = b.mesh()
    no, to, eo = b.batch_info()
    phis = [fa.phi_field("ellipsoids", m.xyz[no[k]:no[k + 1]], seed=100 + k) for k in range(64)]
    out, ifn, _ = b.reset_phi(np.concatenate(phis), mode="points")
    cnts = b.field_counts()
    for k in rng.choice(64, 12, replace=False):
        mk = fa.kuhn_mesh(int(sizes[k]))
        ref = oracle.reinit(mk.xyz, mk.tets, mk.edges, mk.t2e, phis[k], blend=0)
        assert cnts[k].n_ifp == len(ref["ifp"]) and cnts[k].n_tri == len(ref["tris"])
        assert np.array_equal(ifn[no[k]:no[k + 1]].view(np.uint64), ref["phi_ifn"].view(np.uint64))
    b.close()


def test_served_results_equal_alone():
    """Jobs served concurrently in mixed batches get exactly their solo result."""
    from fluidtissue_amd.serve import Server
    rng = np.random.default_rng(9)
    dims = [int(n) for n in rng.integers(12, 30, size=14)]
    kinds = ["star", "spheres", "ellipsoids"]
    inputs = []
    for k, n in enumerate(dims):
        m = fa.kuhn_mesh(n)
        inputs.append((n, m, fa.phi_field(kinds[k % 3], m.xyz, seed=k)))
    jobs = []
    with Server(device=0, lanes=2, max_fields=5, mode="triangles", blend="update") as srv:
        with srv.cv:                                       # one burst: queue all, then release
            for n, m, phi in inputs:
                jobs.append((m, phi, srv.submit((n, n, n), phi)))
        for _, _, j in jobs:
            j.wait(120)
        assert max(b[0] for b in srv.batches) > 1        # some batches mixed jobs
    for m, phi, j in jobs:
        a = fa.Reinit(0)
        a.set_mesh(m)
        o1, i1, c1 = a.reset_phi(phi, mode="triangles", blend="update")
        assert np.array_equal(j.out.view(np.uint64), o1.view(np.uint64))
        assert np.array_equal(j.ifn.view(np.uint64), i1.view(np.uint64))
        assert (j.counts.n_ifp, j.counts.n_tri) == (c1.n_ifp, c1.n_tri)
        a.close()


def test_served_inplace_pinned_and_device_inputs():
    """phi given as a numpy array, a pinned host tensor or a device tensor,
    reinitialised in place or into new storage: all equal the solo result."""
    import torch
    from fluidtissue_amd.serve import Server
    sizes = [14, 22, 18, 14, 26, 20]
    inputs = []
    for k, n in enumerate(sizes):
        m = fa.kuhn_mesh(n)
        inputs.append((n, m, fa.phi_field("spheres" if k % 2 else "star", m.xyz, seed=k)))
    with Server(device=0, lanes=2, max_fields=4, mode="triangles", blend="update",
                want_ifn=False) as srv:
        jobs = []
        with srv.cv:
            for k, (n, m, phi) in enumerate(inputs):
                kind = k % 3
                if kind == 0:
                    src, inplace = phi.copy(), True
                elif kind == 1:
                    src, inplace = torch.from_numpy(phi).pin_memory(), True
                else:
                    src, inplace = torch.from_numpy(phi).cuda(), False
                jobs.append((m, phi, src, inplace, srv.submit((n, n, n), src, inplace=inplace)))
        for *_, j in jobs:
            j.wait(120)
    for m, phi, src, inplace, j in jobs:
        a = fa.Reinit(0)
        a.set_mesh(m)
        o1, _, _ = a.reset_phi(phi, mode="triangles", blend="update")
        a.close()
        got = j.out.cpu().numpy() if isinstance(j.out, torch.Tensor) else j.out
        assert np.array_equal(np.asarray(got).view(np.uint64), o1.view(np.uint64))
        assert j.ifn is None
        if inplace:
            back = src.numpy() if isinstance(src, torch.Tensor) else src
            assert np.array_equal(back.view(np.uint64), o1.view(np.uint64))


def test_resident_mesh_batch_matches_set_mesh():
    ids = []
    import ctypes as C
    from fluidtissue_amd._lib import lib
    for n in (10, 14, 10):
        mid = C.c_int32()
        assert lib.ft_mesh_create_kuhn(0, n, n, n, 0.0, 0, C.byref(mid)) == 0
        ids.append(mid.value)
    b = fa.Reinit(0)
    b.set_batch_meshes(ids)
    ref = fa.Reinit(0)
    ref.set_batch([fa.kuhn_mesh(n) for n in (10, 14, 10)])
    mb, mr = b.mesh(), ref.mesh()
    for x, y in zip((mb.xyz, mb.tets, mb.edges, mb.t2e), (mr.xyz, mr.tets, mr.edges, mr.t2e)):
        assert np.array_equal(x, y)
    phi = fa.phi_field("star", mb.xyz)
    o1, i1, _ = b.reset_phi(phi, mode="triangles")
    o2, i2, _ = ref.reset_phi(phi, mode="triangles")
    assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
    b.close()
    ref.close()
    for mid in ids:
        assert lib.ft_mesh_release(mid) == 0
