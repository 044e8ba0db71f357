"""Host logic of the served form (no GPU): mesh cache budget/LRU/pins and
batch admission."""
import collections
import types

import pytest

from fluidtissue_amd.serve import WORK_BYTES_PER_NODE, MeshCache, admit


def _fake_cache(budget):
    made, freed = [], []

    def create(device, dims):
        made.append(tuple(dims))
        return len(made), MeshCache.estimate_bytes(dims)

    return MeshCache(0, budget, create=create, release=freed.append), made, freed


def test_cache_reuses_and_evicts_least_recently_used():
    b20 = MeshCache.estimate_bytes((20, 20, 20))
    c, made, freed = _fake_cache(2.5 * b20)
    m1 = c.acquire((20, 20, 20)); c.release((20, 20, 20))
    m2 = c.acquire((21, 20, 20)); c.release((21, 20, 20))
    assert c.acquire((20, 20, 20)) == m1            # resident: no new mesh
    c.release((20, 20, 20))
    assert made == [(20, 20, 20), (21, 20, 20)]
    c.acquire((22, 20, 20))                         # needs room: evicts (21,...), the LRU one
    assert freed == [m2] and (20, 20, 20) in c.meshes and (21, 20, 20) not in c.meshes
    assert c.bytes == sum(e.bytes for e in c.meshes.values()) <= c.budget


def test_cache_never_evicts_pinned_meshes():
    b = MeshCache.estimate_bytes((20, 20, 20))
    c, made, freed = _fake_cache(1.5 * b)
    c.acquire((20, 20, 20))                         # pinned (not released)
    assert not c.can_admit((20, 21, 20), 0)
    with pytest.raises(Exception):
        c.acquire((20, 21, 20))
    assert freed == []
    c.release((20, 20, 20))
    assert c.can_admit((20, 21, 20), 0)
    c.acquire((20, 21, 20))
    assert freed == [1]


def _job(n):
    return types.SimpleNamespace(dims=(n, n, n), n=n ** 3)


def test_admit_fifo_counts_nodes_and_budget():
    q = collections.deque(_job(n) for n in (10, 12, 8, 30, 9))
    always = lambda dims, extra: ("entry", False)
    jobs, res, extra = admit(q, max_fields=3, max_nodes=10 ** 9, inflight_bytes=0, reserve=always)
    assert [j.n for j in jobs] == [1000, 1728, 512] and len(q) == 2 and len(res) == 3
    assert extra == WORK_BYTES_PER_NODE * (1000 + 1728 + 512)
    # node cap: the 30^3 job does not join a batch that would exceed it, but
    # is admitted alone next time
    jobs, _, _ = admit(q, max_fields=8, max_nodes=20000, inflight_bytes=0, reserve=always)
    assert [j.n for j in jobs] == [27000]
    # budget: nothing is taken past a job that does not fit (no overtaking)
    q = collections.deque(_job(n) for n in (10, 40, 8))
    seen = []

    def fits(dims, extra):
        seen.append((dims, extra))
        return ("entry", False) if dims[0] != 40 else None
    jobs, _, extra = admit(q, 8, 10 ** 9, inflight_bytes=5, reserve=fits)
    assert [j.dims[0] for j in jobs] == [10] and [j.dims[0] for j in q] == [40, 8]
    assert seen[0][1] == 5 + WORK_BYTES_PER_NODE * 1000    # in-flight work counted


def test_admission_reserves_meshes_so_two_lanes_cannot_share_free_bytes():
    """Two lanes each admitting a job whose mesh is not resident, with a budget
    that holds one mesh: the first reservation books the bytes, so the second
    lane is refused (it waits) instead of both passing a check and the second
    failing at creation."""
    b = MeshCache.estimate_bytes((20, 20, 20))
    c, made, freed = _fake_cache(1.5 * b)
    work = WORK_BYTES_PER_NODE * 8000
    lane1 = admit(collections.deque([_job(20)]), 8, 10 ** 9, 0, c.reserve)
    assert len(lane1[0]) == 1 and lane1[1][0][1] is True       # booked, to be made by lane 1
    q2 = collections.deque([types.SimpleNamespace(dims=(20, 21, 20), n=8400)])
    lane2 = admit(q2, 8, 10 ** 9, lane1[2], c.reserve)          # before lane 1 made its mesh
    assert lane2[0] == [] and len(q2) == 1 and made == []
    entry, make = lane1[1][0]
    assert c.materialize((20, 20, 20), entry, make) == 1 and made == [(20, 20, 20)]
    c.unpin(entry)                                              # lane 1's batch ends
    lane2 = admit(q2, 8, 10 ** 9, 0, c.reserve)                 # now the pinned-free mesh is evicted
    assert len(lane2[0]) == 1 and freed == [1]
    assert c.bytes <= c.budget and work > 0


def test_admission_shares_a_mesh_being_made():
    """A second job for the same dims pins the booked entry and waits for the
    lane that makes it, and a failed creation reaches both jobs."""
    import threading
    b = MeshCache.estimate_bytes((16, 16, 16))
    gate = threading.Event()

    def slow_fail(device, dims):
        gate.wait(5)
        raise RuntimeError("no memory")
    c = MeshCache(0, 4 * b, create=slow_fail, release=lambda mid: None)
    e1, make1 = c.reserve((16, 16, 16), 0)
    e2, make2 = c.reserve((16, 16, 16), 0)
    assert e1 is e2 and make1 and not make2 and e1.pins == 2
    errs = []

    def waiter():
        try:
            c.materialize((16, 16, 16), e2, make2)
        except RuntimeError as err:
            errs.append(err)
    t = threading.Thread(target=waiter)
    t.start()
    gate.set()
    with pytest.raises(RuntimeError):
        c.materialize((16, 16, 16), e1, make1)
    t.join(5)
    assert len(errs) == 1 and c.bytes == 0 and (16, 16, 16) not in c.meshes
    c.unpin(e1); c.unpin(e2)
    assert c.reserve((16, 16, 16), 0)[1] is True                # booked afresh


def test_autotune_picks_fastest_and_caches():
    from fluidtissue_amd import autotune
    clock = [0.0]
    calls = []
    cost = {(1.0, 64): 3.0, (1.0, 256): 2.0, (1.5, 64): 1.0, (1.5, 256): 4.0}
    wcost = {5: 0.5, 6: 0.2, 7: 0.4, 8: 0.3}

    class FakeHandle:
        def reset_phi_batch_device(self, phi, ifn, mode, band, bin_scale, query_block, waves_per_simd,
                                   schedule, stream=None):
            calls.append((bin_scale, query_block, waves_per_simd, schedule))
            clock[0] += cost[(bin_scale, query_block)] + wcost[waves_per_simd] + (0.05 if schedule == "queue" else 0)

    restores = []
    orig = autotune.time.perf_counter
    autotune.time.perf_counter = lambda: clock[0]
    try:
        args = dict(sync=lambda: None, copy=lambda d, s, n: restores.append((d, s, n)),
                    signature=("test", 1), bin_scales=(1.0, 1.5), query_blocks=(64, 256), reps=2)
        best, table = autotune.tune(FakeHandle(), 11, 22, 33, **args)
        assert (best["bin_scale"], best["query_block"]) == (1.5, 64)
        assert best["waves_per_simd"] == 6 and best["schedule"] == "static"
        # stage 1: bin x block at 8 waves; stage 2: waves 5..7 (8 already run); stage 3: the queue
        assert [(b, q, w, s) for b, q, w, s, _ in table] == [
            (1.0, 64, 8, "static"), (1.0, 256, 8, "static"), (1.5, 64, 8, "static"), (1.5, 256, 8, "static"),
            (1.5, 64, 5, "static"), (1.5, 64, 6, "static"), (1.5, 64, 7, "static"), (1.5, 64, 6, "queue")]
        assert abs(table[0][4] - 3300.0) < 1e-6 and abs(best["ms"] - 1200.0) < 1e-6
        assert len(calls) == 8 * 3 and restores == [(11, 22, 33)] * 24   # input restored each run
        assert autotune.tune(FakeHandle(), 11, 22, 33, **args)[0] == best
        assert len(calls) == 24                                        # cached per signature
    finally:
        autotune.time.perf_counter = orig
        autotune._cache.pop(("triangles", 0.0, ("test", 1)), None)


def test_autotune_default_candidates_follow_the_mode():
    """Without bin_scales the tuner tries the neighbourhood of each mode's
    measured optimum (points 2.0-2.5, triangles 1.25 node spacings), and
    launch bounds only in triangles mode."""
    from fluidtissue_amd import autotune
    tried = []

    class FakeHandle:
        def reset_phi_batch_device(self, phi, ifn, mode, band, bin_scale, query_block, waves_per_simd,
                                   schedule, stream=None):
            tried.append((mode, bin_scale, waves_per_simd))

    for mode in ("points", "triangles"):
        autotune.tune(FakeHandle(), 1, 2, 3, mode=mode, reps=1, sync=lambda: None,
                      copy=lambda d, s, n: None, query_blocks=(64,))
    pts = sorted({b for m, b, _ in tried if m == "points"})
    tri = sorted({b for m, b, _ in tried if m == "triangles"})
    assert pts == sorted(autotune.MODE_BIN_SCALES["points"]) and 2.0 in pts and 2.5 in pts
    assert tri == sorted(autotune.MODE_BIN_SCALES["triangles"]) and 1.25 in tri
    assert {w for m, _, w in tried if m == "points"} == {8}
    assert {w for m, _, w in tried if m == "triangles"} == {5, 6, 7, 8}
