"""GPU: the batched and served configurations, the reference's edge cases and
the narrow band, through the C ABI against the oracle.

  batched   512 fields N ~ U{24..96}, unions of 1-16 ellipsoids, triangle
            mode, through the captured-graph entry point (ft_reinit_graph):
            every field bitwise equal to its own plain reinit, and 16 seeded
            fields (the largest among them) against the oracle: interface
            bit-exact, sampled distances within 1e-5 h of FP64, nearest
            triangle equal outside eps-ties.
  served    a stream of jobs of 32^3..192^3 (log-uniform) arriving while
            others run, two lanes: each job bitwise equal to its plain
            reinit; four jobs (the largest among them) against the oracle.
  edges     +0.0 / -0.0 nodes and edges with |dphi| <= FLT_MIN
            (LevelSet.cpp:281, 287-290, 380; EuclideanDistance.cu:55, 113).
  band      the 128^3 star with a 6h band: 2500 nodes inside the band.
  reuse     a released resident mesh id that names a new mesh of other
            dimensions is assembled afresh by ft_set_batch_meshes.
"""
import ctypes as C
import time

import numpy as np
import pytest

import fluidtissue_amd as fa
import oracle
from fluidtissue_amd._lib import check, lib
from tests.helpers import assert_interface_bit_exact, triangle_errors, workload

pytestmark = pytest.mark.gpu

FLT_MIN = float(np.finfo(np.float32).tiny)


def log(*a):
    print(*a, flush=True)


def _resident(n):
    mid = C.c_int32()
    check(lib.ft_mesh_create_kuhn(0, n, n, n, 0.0, 0, C.byref(mid)))
    return mid.value


def _grid_xyz(n):
    """The Kuhn generator's node coordinates (x-major, float of i*h in double)."""
    h = 1.0 / (n - 1)
    g = (np.arange(n, dtype=np.float64) * h).astype(np.float32)
    xyz = np.empty((n, n, n, 3), np.float32)
    xyz[..., 0] = g[:, None, None]
    xyz[..., 1] = g[None, :, None]
    xyz[..., 2] = g[None, None, :]
    return xyz.reshape(-1, 3)


def _oracle_check(tag, m, phi, ifn, near, h, itf=None, counts=None, n_sample=500, seed=0):
    """Interface bit-exact (when given), signs, and a seeded node sample's
    triangle distances and nearest triangles against the FP64 oracle."""
    ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
    tris, elems, area = oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
    ref = dict(ifp=ifp, ifp2e=ifp2e, tris=tris, if_elems=elems, ifp_area=area)
    if counts is not None:
        assert (counts.n_ifp, counts.n_tri, counts.n_if_elems) == (len(ifp), len(tris), len(elems))
    if itf is not None:
        assert np.array_equal(itf["ifp"].view(np.uint32), ifp.view(np.uint32))
        assert np.array_equal(itf["ifp2e"], ifp2e)
        assert np.array_equal(itf["tris"], tris)
        assert np.array_equal(itf["ifp_area"].view(np.uint32), area.view(np.uint32))
    assert np.array_equal(np.signbit(ifn), phi < 0.0)          # EuclideanDistance.cu:55
    rng = np.random.default_rng(seed)
    nodes = rng.choice(len(phi), min(n_sample, len(phi)), replace=False)
    d64, f64 = oracle.triangle_distance(m.xyz, ifp, tris, nodes)
    got = np.abs(ifn[nodes])
    err = np.abs(got - d64)
    ties = mism = 0
    for k in np.nonzero(near[nodes] != f64)[0]:
        dg = oracle.node_triangle_distance(nodes[k], m.xyz, ifp, tris, near[nodes][k])
        if abs(dg - d64[k]) <= 1e-5 * h:
            ties += 1
        else:
            mism += 1
    log(f"{tag}: P={len(ifp)} F={len(tris)}; {len(nodes)} nodes, max err {err.max() / h:.2e} h, "
        f"rms {np.sqrt((err ** 2).mean()) / h:.2e} h, eps-ties {ties}, mismatches {mism}")
    assert err.max() <= 1e-5 * h
    assert mism == 0
    return err.max() / h


def _local_interface(itf, cnts, k, edge_off):
    """Field k's part of a batched interface, renumbered field-locally."""
    p0 = sum(c.n_ifp for c in cnts[:k])
    f0 = sum(c.n_tri for c in cnts[:k])
    P, F = cnts[k].n_ifp, cnts[k].n_tri
    tris = itf.tris[f0:f0 + F].copy()
    tris[:, :3] -= p0
    return dict(ifp=itf.ifp[p0:p0 + P], ifp2e=itf.ifp2e[p0:p0 + P] - edge_off[k], tris=tris,
                ifp_area=itf.ifp_area[p0:p0 + P]), f0


@pytest.mark.timeout(400)
def test_batched_512_fields_graph_against_oracle():
    import torch
    t0 = time.perf_counter()
    rng = np.random.default_rng(7)                  # the batched workload's seeding
    sizes = [int(n) for n in rng.integers(24, 97, size=512)]
    ids = {n: _resident(n) for n in sorted(set(sizes))}
    xyz = {n: _grid_xyz(n) for n in ids}
    phis = [fa.phi_field("ellipsoids", xyz[n], seed=1000 + k) for k, n in enumerate(sizes)]
    N = sum(n ** 3 for n in sizes)
    log(f"batched: 512 fields, {N} nodes, {len(ids)} sizes ({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda:0")
    b = fa.Reinit(0)
    b.set_batch_meshes([ids[n] for n in sizes])
    no, to, eo = b.batch_info()
    d_phi0 = torch.from_numpy(np.concatenate(phis)).to(dev)
    d_phi = d_phi0.clone()
    d_ifn = torch.empty_like(d_phi)
    stream = torch.cuda.current_stream().cuda_stream
    cnts = b.reset_phi_graph_device(d_phi.data_ptr(), d_ifn.data_ptr(), mode="triangles",
                                    want_nearest=True, stream=stream)
    d_phi.copy_(d_phi0)                             # a replay of the captured graph
    cnts = b.reset_phi_graph_device(d_phi.data_ptr(), d_ifn.data_ptr(), mode="triangles",
                                    want_nearest=True, stream=stream)
    torch.cuda.synchronize()
    captures, launches = b.graph_stats()
    assert captures >= 1 and launches >= 2
    out = d_phi.cpu().numpy()
    ifn = d_ifn.cpu().numpy()
    near = b.nearest()
    itf = b.interface()
    log(f"batched graph reinit done: {captures} capture(s), {launches} launches "
        f"({time.perf_counter() - t0:.1f} s)")
    # every field: bitwise its plain reinit (one handle per size, device phi)
    singles = {}
    for n, mid in ids.items():
        h = fa.Reinit(0)
        h.set_batch_meshes([mid])
        singles[n] = h
    d_work = torch.empty(max(n ** 3 for n in sizes), dtype=torch.float64, device=dev)
    d_wifn = torch.empty_like(d_work)
    solo_near = {}
    check_fields = set(int(k) for k in rng.choice(512, 15, replace=False))
    check_fields.add(int(np.argmax(sizes)))
    for k, n in enumerate(sizes):
        sl = slice(int(no[k]), int(no[k + 1]))
        w, wi = d_work[:n ** 3], d_wifn[:n ** 3]
        w.copy_(d_phi0[sl])
        c1 = singles[n].reset_phi_device(w.data_ptr(), wi.data_ptr(), mode="triangles",
                                         want_nearest=k in check_fields, stream=stream)
        assert (cnts[k].n_ifp, cnts[k].n_tri) == (c1.n_ifp, c1.n_tri)
        assert np.array_equal(out[sl].view(np.uint64), w.cpu().numpy().view(np.uint64))
        assert np.array_equal(ifn[sl].view(np.uint64), wi.cpu().numpy().view(np.uint64))
        if k in check_fields:
            solo_near[k] = singles[n].nearest()
    log(f"batched: all 512 fields equal their plain reinit ({time.perf_counter() - t0:.1f} s)")
    for h in singles.values():
        h.close()
    # 16 fields against the oracle
    worst = 0.0
    for k in sorted(check_fields):
        n = sizes[k]
        m = fa.kuhn_mesh(n)
        sl = slice(int(no[k]), int(no[k + 1]))
        loc, f0 = _local_interface(itf, cnts, k, eo)
        nk = near[sl].copy()
        nk[nk >= 0] -= f0
        assert np.array_equal(nk, solo_near[k])
        worst = max(worst, _oracle_check(f"batched field {k} ({n}^3)", m, phis[k], ifn[sl], nk,
                                         1.0 / (n - 1), itf=loc, counts=cnts[k], seed=k))
    log(f"batched: 16 fields against the oracle, worst max err {worst:.2e} h "
        f"({time.perf_counter() - t0:.1f} s)")
    b.close()
    for mid in ids.values():
        check(lib.ft_mesh_release(mid))


@pytest.mark.timeout(400)
def test_served_stream_32_to_192_against_alone_and_oracle():
    import threading
    import torch
    from fluidtissue_amd.serve import Server
    t0 = time.perf_counter()
    rng = np.random.default_rng(23)
    raw = np.exp(rng.uniform(np.log(32), np.log(192), size=22))
    dims = [int(min(192, max(32, 8 * round(x / 8)))) for x in raw] + [192, 32]
    kinds = ["star" if k % 2 else "spheres" for k in range(len(dims))]
    xyz = {n: _grid_xyz(n) for n in set(dims)}
    phis = [fa.phi_field(kinds[k], xyz[n], seed=0) for k, n in enumerate(dims)]
    log(f"served: {len(dims)} jobs, sizes {sorted(set(dims))} ({time.perf_counter() - t0:.1f} s)")
    jobs = [None] * len(dims)
    with Server(device=0, lanes=2, max_fields=8, mode="triangles", blend="update") as srv:
        def submit_all():
            # a burst of ten (queued while the lanes are busy), then arrivals
            # spaced while earlier jobs run
            with srv.cv:
                for k in range(10):
                    jobs[k] = srv.submit((dims[k],) * 3, phis[k])
            for k in range(10, len(dims)):
                jobs[k] = srv.submit((dims[k],) * 3, phis[k])
                time.sleep(float(rng.exponential(0.002)))
        th = threading.Thread(target=submit_all)
        th.start()
        th.join()
        for j in jobs:
            j.wait(300)
        nb = [b[0] for b in srv.batches]
    log(f"served: {len(nb)} batches, largest {max(nb)} jobs ({time.perf_counter() - t0:.1f} s)")
    assert max(nb) > 1                              # some jobs shared a batch
    # each job bitwise its plain reinit
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    singles = {}
    for n in sorted(set(dims)):
        mid = _resident(n)
        h = fa.Reinit(0)
        h.set_batch_meshes([mid])
        singles[n] = (h, mid)
    oracle_jobs = {int(np.argmax(dims))} | set(int(k) for k in rng.choice(len(dims), 3, replace=False))
    for k, n in enumerate(dims):
        h, _ = singles[n]
        d = torch.from_numpy(phis[k]).to(dev)
        di = torch.empty_like(d)
        h.reset_phi_device(d.data_ptr(), di.data_ptr(), mode="triangles", blend="update",
                           want_nearest=k in oracle_jobs, stream=stream)
        o1, i1 = d.cpu().numpy(), di.cpu().numpy()
        assert np.array_equal(jobs[k].out.view(np.uint64), o1.view(np.uint64)), k
        assert np.array_equal(jobs[k].ifn.view(np.uint64), i1.view(np.uint64)), k
        if k in oracle_jobs:
            m = fa.kuhn_mesh(n)
            _oracle_check(f"served job {k} ({n}^3, {kinds[k]})", m, phis[k], i1, h.nearest(),
                          1.0 / (n - 1), counts=jobs[k].counts, seed=k)
    for h, mid in singles.values():
        h.close()
        check(lib.ft_mesh_release(mid))
    log(f"served: all jobs equal alone ({time.perf_counter() - t0:.1f} s)")


def _signed_zero_phi(n, seed):
    """A plane phi = x - 0.5 + small wave whose nodes are, at random, replaced
    by +0.0, -0.0, +-FLT_MIN/4, +-1e-300 or +-FLT_MIN: crossings with -0.0 on
    one side, edges with |dphi| <= FLT_MIN (midpoint rule) and tiny-valued
    orientation references.  No crossing tet has all four nodes at
    |phi| <= FLT_MIN (the reference reads past such a tet)."""
    m = fa.kuhn_mesh(n)
    x = m.xyz.astype(np.float64)
    phi = (x[:, 0] - 0.5) + 0.05 * np.sin(7 * x[:, 1]) * np.cos(5 * x[:, 2])
    rng = np.random.default_rng(seed)
    pick = rng.random(len(phi)) < 0.3
    vals = np.array([0.0, -0.0, FLT_MIN / 4, -FLT_MIN / 4, 1e-300, -1e-300, FLT_MIN, -FLT_MIN])
    phi[pick] = vals[rng.integers(0, len(vals), pick.sum())]
    orig = (x[:, 0] - 0.5) + 0.05 * np.sin(7 * x[:, 1]) * np.cos(5 * x[:, 2])
    while True:
        tiny = np.abs(phi[m.tets]) <= FLT_MIN
        sb = np.signbit(phi[m.tets])
        bad = tiny.all(axis=1) & (sb.any(axis=1) & ~sb.all(axis=1))
        if not bad.any():
            break
        phi[m.tets[bad, 0]] = orig[m.tets[bad, 0]]
    return m, phi


@pytest.mark.parametrize("seed", [1, 2])
def test_signed_zero_and_flt_min_edges(seed):
    n = 14
    h = 1.0 / (n - 1)
    m, phi = _signed_zero_phi(n, seed)
    neg0 = np.signbit(phi) & (phi == 0.0)
    e0, e1 = phi[m.edges[:, 0]], phi[m.edges[:, 1]]
    mid = (np.signbit(e0) != np.signbit(e1)) & (np.abs(e0 - e1) <= FLT_MIN)
    log(f"seed {seed}: {neg0.sum()} nodes at -0.0, {mid.sum()} crossing edges on the midpoint rule")
    assert neg0.sum() > 50 and mid.sum() > 20
    r = fa.Reinit(0)
    r.set_mesh(m)
    ref0 = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=0)
    ref1 = oracle.reinit(m.xyz, m.tets, m.edges, m.t2e, phi, blend=1)
    for blend, ref in (("init", ref0), ("update", ref1)):
        out, ifn, cnt = r.reset_phi(phi, mode="points", blend=blend)
        itf = r.interface()
        ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
        tris, elems, area = oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
        assert_interface_bit_exact(itf, dict(ifp=ifp, ifp2e=ifp2e, tris=tris, if_elems=elems,
                                             ifp_area=area))
        assert np.array_equal(ifn.view(np.uint64), ref["phi_ifn"].view(np.uint64))
        assert np.array_equal(out.view(np.uint64), ref["phi"].view(np.uint64))
    # -0.0 is positive for the distance sign (phi < 0.0 is false)
    assert not np.signbit(ifn[neg0]).any()
    _, ifn_t, cnt_t = r.reset_phi(phi, mode="triangles", want_nearest=True)
    _oracle_check(f"signed zeros seed {seed} triangles", m, phi, ifn_t, r.nearest(), h,
                  n_sample=len(phi), seed=seed)
    assert not np.signbit(ifn_t[neg0]).any()
    r.close()


def test_star_128_band_nodes_inside_the_band():
    """The 128^3 star with a 6h band, checked on 2500 nodes that lie inside
    the band: |phi|/2.3 bounds the distance to the analytic surface from
    above (phi is 2.3 times the radial gap), so nodes with |phi|/2.3 < 5.5h
    are within the band of the triangulated surface as well."""
    m, phi, h = workload("star_128")
    band = 6.0 * h
    r = fa.Reinit(0)
    r.set_mesh(m)
    _, ifn, cnt = r.reset_phi(phi, mode="triangles", band=band, want_nearest=True)
    near = r.nearest()
    r.close()
    cand = np.nonzero(np.abs(phi) / 2.3 < 5.5 * h)[0]
    rng = np.random.default_rng(31)
    nodes = rng.choice(cand, 2500, replace=False)
    ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
    tris, _, _ = oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
    ref = dict(ifp=ifp, tris=tris)
    d64, _ = oracle.triangle_distance(m.xyz, ifp, tris, nodes)
    assert np.all(d64 < band * (1 - 1e-6))
    assert np.all(near[nodes] >= 0)
    emax, erms, ties, mism = triangle_errors("star_128", np.abs(ifn[nodes]), near[nodes], nodes, h,
                                             ref=ref)
    log(f"star_128 band 6h: {len(nodes)} in-band nodes, max err {emax / h:.2e} h, "
        f"rms {erms / h:.2e} h, eps-ties {ties}, mismatches {mism}; clamped {cnt.n_clamped}")
    assert emax <= 1e-5 * h and mism == 0
    assert np.array_equal(np.signbit(ifn[nodes]), phi[nodes] < 0.0)


def test_released_mesh_id_reused_by_a_new_mesh():
    """ft_set_batch_meshes([id]) after id was released and reissued to a mesh
    of other dimensions assembles the new mesh (not the handle's stale copy)."""
    a = _resident(12)
    b = fa.Reinit(0)
    b.set_batch_meshes([a])
    m12 = fa.kuhn_mesh(12)
    b.reset_phi(fa.phi_field("star", m12.xyz), mode="triangles")
    check(lib.ft_mesh_release(a))
    a2 = _resident(14)
    assert a2 == a                                   # the id was reissued
    b.set_batch_meshes([a2])
    no, to, eo = b.batch_info()
    m14 = fa.kuhn_mesh(14)
    assert (no[-1], to[-1], eo[-1]) == (len(m14.xyz), len(m14.tets), len(m14.edges))
    phi = fa.phi_field("star", m14.xyz)
    o1, i1, c1 = b.reset_phi(phi, mode="triangles", blend="update")
    s = fa.Reinit(0)
    s.set_mesh(m14)
    o2, i2, c2 = s.reset_phi(phi, mode="triangles", blend="update")
    assert (c1.n_ifp, c1.n_tri, c1.n_if_elems) == (c2.n_ifp, c2.n_tri, c2.n_if_elems)
    assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
    assert np.array_equal(i1.view(np.uint64), i2.view(np.uint64))
    s.close()
    b.close()
    check(lib.ft_mesh_release(a2))


@pytest.mark.parametrize("mode", ["triangles", "points"])
def test_launch_bounds_and_work_queue_give_the_same_bits(mode):
    """Every launch bound (5..8 waves per SIMD) and the work-queue schedule
    give the default launch's results bit for bit: a single jittered field
    (deep tiles first) and a batch through the captured graph, replayed."""
    import torch
    m, phi, h = workload("spheres_96")
    r = fa.Reinit(0)
    r.set_mesh(m)
    ref = r.reset_phi(phi, mode=mode, blend="update", want_nearest=True)
    ref_near = r.nearest()
    for wps in (5, 6, 7, 8):
        for sch in ("static", "queue"):
            out, ifn, cnt = r.reset_phi(phi, mode=mode, blend="update", want_nearest=True,
                                        waves_per_simd=wps, schedule=sch, query_block=64 if wps % 2 else 256)
            assert np.array_equal(out.view(np.uint64), ref[0].view(np.uint64)), (wps, sch)
            assert np.array_equal(ifn.view(np.uint64), ref[1].view(np.uint64)), (wps, sch)
            assert np.array_equal(r.nearest(), ref_near), (wps, sch)
    r.close()
    sizes = [20, 33, 27, 41, 18]
    b = fa.Reinit(0)
    b.set_kuhn_batch(sizes, jitter=0.2, seed=4)
    mb = b.mesh()
    phib = fa.phi_field("spheres", mb.xyz, seed=5)
    d0 = torch.from_numpy(phib).cuda()
    outs = []
    for sch, reps in (("static", 1), ("queue", 3)):
        for _ in range(reps):
            d = d0.clone()
            torch.cuda.synchronize()
            b.reset_phi_graph_device(d.data_ptr(), None, mode=mode, schedule=sch, waves_per_simd=6)
            torch.cuda.synchronize()
            outs.append(d.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))
    b.close()


@pytest.mark.parametrize("mode", ["triangles", "points"])
def test_split_deep_tiles_give_the_same_bits(mode):
    """Tiles deep inside the surface (a node with phi_old <= -25 h) are searched
    by several waves, each testing the leaves it owns, and merged afterwards.
    Their nodes get the whole-tile search's results bit for bit, nearest
    indices included: the work-queue schedule searches every tile whole.  The
    64^3 star reaches about -50 h, so it has such tiles; the triangle results
    are also checked against the oracle on its deepest nodes."""
    m, phi, h = workload("star_64")
    assert phi.min() <= -25 * h
    r = fa.Reinit(0)
    r.set_mesh(m)
    split = r.reset_phi(phi, mode=mode, blend="update", want_nearest=True)
    near_split = r.nearest()
    whole = r.reset_phi(phi, mode=mode, blend="update", want_nearest=True, schedule="queue")
    near_whole = r.nearest()
    assert np.array_equal(split[0].view(np.uint64), whole[0].view(np.uint64))
    assert np.array_equal(split[1].view(np.uint64), whole[1].view(np.uint64))
    assert np.array_equal(near_split, near_whole)
    if mode == "triangles":
        deep = np.argsort(phi)[:300]
        emax, _, _, mism = triangle_errors("star_64", np.abs(split[1][deep]), near_split[deep], deep, h)
        assert emax <= 1e-5 * h
        assert mism == 0
    r.close()


@pytest.mark.parametrize("name,band_h", [("star_64", 6.0), ("spheres_96", 6.0), ("spheres_96", 2.0)])
def test_bins_schedule_equals_tree_walk(name, band_h):
    """The narrow band without the tree (each tile tests the triangles of the
    index bins within the band of its box, staged in LDS) gives the tree
    walk's results bit for bit: phi, phi_ifn, nearest triangles and the
    clamped count."""
    m, phi, h = workload(name)
    r = fa.Reinit(0)
    r.set_mesh(m)
    walk = r.reset_phi(phi, mode="triangles", blend="update", band=band_h * h, want_nearest=True)
    near_walk = r.nearest()
    for qb in (64, 256):
        bins = r.reset_phi(phi, mode="triangles", blend="update", band=band_h * h, want_nearest=True,
                           schedule="bins", query_block=qb)
        assert np.array_equal(bins[0].view(np.uint64), walk[0].view(np.uint64)), qb
        assert np.array_equal(bins[1].view(np.uint64), walk[1].view(np.uint64)), qb
        assert np.array_equal(r.nearest(), near_walk), qb
        assert bins[2].n_clamped == walk[2].n_clamped > 0
    with pytest.raises(fa._lib.FtError):
        r.reset_phi(phi, mode="points", band=band_h * h, schedule="bins")
    with pytest.raises(fa._lib.FtError):
        r.reset_phi(phi, mode="triangles", schedule="bins")
    r.close()


def test_bins_schedule_on_a_batch():
    """The bins kernel on a batch: each tile reads its own field's bin grid
    and bin range, and the results equal the tree walk's bit for bit."""
    import torch
    r = fa.Reinit(0)
    r.set_kuhn_batch([24, 40, (33, 28, 30), 17], jitter=0.15, seed=3)
    m = r.mesh()
    no, _, _ = r.batch_info()
    kinds = ["star", "spheres", "ellipsoids", "sphere_wave"]
    phi = np.concatenate([fa.phi_field(kinds[k], m.xyz[no[k]:no[k + 1]], seed=k + 7) for k in range(len(no) - 1)])
    outs = []
    for sch in ("static", "bins"):
        d = torch.from_numpy(phi.copy()).cuda()
        di = torch.empty_like(d)
        cnts = r.reset_phi_batch_device(d.data_ptr(), di.data_ptr(), mode="triangles", blend="update",
                                        band=0.1, want_nearest=True, schedule=sch)
        torch.cuda.synchronize()
        outs.append((d.cpu().numpy(), di.cpu().numpy(), r.nearest(), [c.n_clamped for c in cnts]))
    (o1, i1, n1, c1), (o2, i2, n2, c2) = outs
    assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
    assert np.array_equal(i1.view(np.uint64), i2.view(np.uint64))
    assert np.array_equal(n1, n2)
    assert c1 == c2 and sum(c1) > 0
    r.close()
