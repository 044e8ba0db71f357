"""Benchmark of the level-set reinit on one MI355X.

One step = one full reinit of the 256^3 workload with phi resident in HBM:
marching tets (crossing points, triangles, areas, p2t), the uniform-bin
index of the triangles, and the full-domain nearest-triangle signed distance
of every node with the blend fused in (LevelSet::resetPhi_ + LevelSet::init's
reset, src/fem/LevelSet.cpp:73-74, 429-455).  Each step first restores the
same input phi with a device copy (included in the timed region).

Other workloads (--workload): star_128 and spheres_256_band6h (narrow
bands), batched (512 fields in one captured graph), serving (the served
stream against serial calls), advect (LevelSet::update's advection step) and
normals (the reinit with the level-set normals, then the interface
curvature).

Prints one JSON line (rank 0).  --gpus N > 1 is not supported: the workload
is one field on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "sphere_wave_32": dict(n=32, jitter=0.0, phi="sphere_wave", desc="32^3 Kuhn, sphere+wave"),
    "star_128": dict(n=128, jitter=0.0, phi="star", band_h=6.0, desc="128^3 Kuhn, star surface"),
    "spheres_256": dict(n=256, jitter=0.2, phi="spheres",
                 desc="256^3 Kuhn, interior nodes jittered +-0.2h, 64 spheres"),
    "spheres_256_band6h": dict(n=256, jitter=0.2, phi="spheres", band_h=6.0,
                               desc="256^3 Kuhn, interior nodes jittered +-0.2h, 64 spheres, 6h band"),
    "spheres_128": dict(n=128, jitter=0.2, phi="spheres",
                     desc="128^3 Kuhn, interior nodes jittered +-0.2h, 64 spheres (the 256^3 field, coarser)"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(m, phi, nthreads, sample_div=1024, seed=0):
    """The oracle (C restatement of the reference's CPU reinit) on the host:
    findInterfaceNodes_ + reconstructInterface_ in full (serial, as the
    reference), brute-force point distance (EuclideanDistance.cu:108-117,
    OpenMP) on a seeded 1/sample_div node sample, extrapolated linearly."""
    import oracle
    t0 = time.perf_counter()
    ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
    oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
    t_mt = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    n = m.n_nodes
    ns = max(1, n // sample_div)
    nodes = rng.choice(n, ns, replace=False)
    t0 = time.perf_counter()
    oracle.point_distance(m.xyz, ifp, nodes, nthreads=nthreads)
    t_d = (time.perf_counter() - t0) * (n / ns)
    total = t_mt + t_d
    return dict(value=n / total / 1e6, unit="Mnodes/s", cores=nthreads, kind="port",
                sample=(f"interface extraction timed in full ({t_mt:.2f} s, 1 thread); "
                        f"point distance on {ns} of {n} nodes ({nthreads} threads), "
                        f"extrapolated x{n / ns:.0f} to {t_d:.1f} s"),
                seconds=total)


def grid_xyz(n):
    """Node coordinates of an n^3 grid on the unit cube, as the Kuhn generators
    make them (x-major; float of i*h computed in double)."""
    h = 1.0 / (n - 1)
    g = (np.arange(n, dtype=np.float64) * h).astype(np.float32)
    xyz = np.empty((n, n, n, 3), np.float32)
    xyz[..., 0] = g[:, None, None]
    xyz[..., 1] = g[None, :, None]
    xyz[..., 2] = g[None, None, :]
    return xyz.reshape(-1, 3)


def advect_flow(xyz):
    """Velocity of the advection workload: a rotation about the vertical axis
    through the domain centre plus a vertical shear, |V| <= 0.9; pressure 0."""
    x = xyz.astype(np.float64)
    n = len(x)
    v = np.concatenate([-(x[:, 1] - 0.5), x[:, 0] - 0.5, 0.2 * np.sin(2 * np.pi * x[:, 0])])
    return np.concatenate([v, np.zeros(n)])


def run_advect(args):
    """One step = the advection of LevelSet::update (LevelSet.cpp:89-141) on
    the 256^3 mesh: assembly of the advection matrix, the Crank-Nicolson
    system and BiCGSTAB to |r| <= 1e-10 |b|, phi restored from a device copy
    before each step (timed)."""
    import torch
    import fluidtissue_amd as fa
    from fluidtissue_amd import _lib
    import ctypes as C
    n = args.advect_n
    m = fa.kuhn_mesh(n, jitter=0.2, seed=1)
    phi = fa.phi_field("spheres", m.xyz, seed=0)
    flow = advect_flow(m.xyz)
    h = 1.0 / (n - 1)
    dt = 0.5 * h / 0.9                           # CFL 0.5
    tol, max_iters = 1e-10, 500
    dev = torch.device("cuda:0")
    r = fa.Reinit(0)
    r.set_mesh(m)
    d_phi0 = torch.from_numpy(phi).to(dev)
    d_phi = torch.empty_like(d_phi0)
    d_flow = torch.from_numpy(flow).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    info = _lib.FtAdvectInfo()

    def step():
        d_phi.copy_(d_phi0)
        _lib.check(_lib.lib.ft_advect_device(r._h, C.c_void_p(d_phi.data_ptr()), C.c_void_p(d_flow.data_ptr()),
                                             dt, tol, max_iters, C.byref(info), C.c_void_p(stream)))
        return info.iterations

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    N = m.n_nodes
    row_off = np.empty(N + 1, np.int64)
    _lib.check(_lib.lib.ft_fem_assemble(r._h, 0, None, row_off.ctypes.data_as(C.c_void_p), None, None))
    nnz = int(row_off[-1])
    log(f"advect {n}^3: {ms:.3f} ms per step, {np.mean(its):.1f} BiCGSTAB iterations, "
        f"error {info.error:.3g}, nnz {nnz}")
    # per iteration: two SpMVs (8 B value + 4 B column per entry) and the
    # vector passes (p 32, v 32, t 32, x/r 72 bytes per row)
    it_bytes = 24 * nnz + 168 * N
    cpu = None
    if not args.no_cpu:
        import oracle
        nc = 64
        mc = fa.kuhn_mesh(nc, jitter=0.2, seed=1)
        pc = fa.phi_field("spheres", mc.xyz, seed=0)
        t1 = time.perf_counter()
        _, ci = oracle.advect(mc.xyz, mc.tets, pc, advect_flow(mc.xyz), 0.5 / (nc - 1) / 0.9, tol, max_iters)
        tc = time.perf_counter() - t1
        cpu = dict(value=mc.n_nodes / tc / 1e6, unit="Mnodes/s", cores=1, kind="port",
                   sample=f"oracle advection of the {nc}^3 mesh (same field, velocity and CFL): "
                          f"{tc:.2f} s, {ci['iterations']} iterations, 1 thread")
    out = {
        "metric": f"Mnodes/s advection step (Crank-Nicolson + BiCGSTAB to 1e-10) at {n}^3",
        "value": N / (ms * 1e-3) / 1e6, "unit": "Mnodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"advect: {n}^3 Kuhn (jittered 0.2h), 64 spheres phi, rotation + shear "
                               f"velocity, CFL 0.5", "nodes": N, "nnz": nnz, "dt": dt,
                   "iterations": float(np.mean(its)), "error": info.error},
        "iteration_bytes": it_bytes,
        "cpu_baseline": cpu,
    }
    r.close()
    return out


def run_normals(args):
    """One step = the reinit of the 256^3 workload with the level-set normals
    (LevelSet::resetPhi_ to its end, LevelSet.cpp:429-455 with
    Assembler::grad_nodes), then the interface curvature
    (LevelSet::getInterfaceCurvature, LevelSet.cpp:215-253), phi restored from
    a device copy before each step; both timed, each also on its own.  One
    surface-tension call (Flow::surfaceTension_) follows, untimed: its force
    arrays are the caller's host arrays."""
    import torch
    import fluidtissue_amd as fa
    wl = WORKLOADS["spheres_256"]
    m = fa.kuhn_mesh(wl["n"], jitter=wl["jitter"], seed=1)
    phi = fa.phi_field(wl["phi"], m.xyz, seed=0)
    r = fa.Reinit(0)
    r.set_mesh(m)
    d0 = torch.from_numpy(phi).cuda()
    d = torch.empty_like(d0)
    kw = dict(mode="triangles", bin_scale=1.25, query_block=64, want_normals=True)

    def reinit():
        d.copy_(d0)
        r.reset_phi_device(d.data_ptr(), None, **kw)

    for _ in range(args.warmup):
        reinit()
        r.curvature()
    torch.cuda.synchronize()
    t_re = t_cu = 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        reinit()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        k = r.curvature()
        t2 = time.perf_counter()
        t_re += t1 - t0
        t_cu += t2 - t1
    fx, fy, fz = r.surface_tension(0.1, 1)
    ms_re, ms_cu = t_re * 1e3 / args.steps, t_cu * 1e3 / args.steps
    N = m.n_nodes
    log(f"normals 256^3: reinit with normals {ms_re:.3f} ms, curvature {ms_cu:.3f} ms "
        f"({k.shape[0]} crossing points)")
    out = {
        "metric": "Mnodes/s reinit with normals + interface curvature at 256^3",
        "value": N / ((ms_re + ms_cu) * 1e-3) / 1e6, "unit": "Mnodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_re + ms_cu, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 search; f64 gradients", "data": "synthetic",
        "config": {"workload": "normals: the 256^3 reinit workload with want_normals, then the curvature",
                   "nodes": N, "crossing_points": int(k.shape[0])},
        "reinit_with_normals_ms": ms_re, "curvature_ms": ms_cu,
        "surface_tension_force_max": float(max(np.abs(fx).max(), np.abs(fy).max(), np.abs(fz).max())),
    }
    r.close()
    return out


def run_batched(args):
    """512 fields, N ~ U{24..96}, each a union of 1-16 ellipsoids: one batched
    reinit (autotuned bin size and workgroup size) against 512 serial plain
    reinits of the same fields on resident meshes."""
    import ctypes as C
    import torch
    import fluidtissue_amd as fa
    from fluidtissue_amd import autotune
    from fluidtissue_amd._lib import check, lib

    rng = np.random.default_rng(7)
    nfields = args.fields
    sizes = rng.integers(args.min_n, args.max_n + 1, size=nfields)
    t0 = time.perf_counter()
    ids = {}
    for n in sorted(set(int(x) for x in sizes)):
        mid = C.c_int32()
        check(lib.ft_mesh_create_kuhn(0, n, n, n, 0.0, 0, C.byref(mid)))
        ids[n] = mid.value
    xyz_cache = {}
    phis = []
    for k, n in enumerate(sizes):
        n = int(n)
        if n not in xyz_cache:
            xyz_cache[n] = grid_xyz(n)
        phis.append(fa.phi_field("ellipsoids", xyz_cache[n], seed=1000 + k))
    N = int(sum(int(n) ** 3 for n in sizes))
    log(f"batched inputs: {nfields} fields, {N} nodes, {len(ids)} resident meshes "
        f"({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    b = fa.Reinit(0)
    b.set_batch_meshes([ids[int(n)] for n in sizes])
    d_phi0 = torch.from_numpy(np.concatenate(phis)).to(dev)
    d_phi = torch.empty_like(d_phi0)
    nbytes = d_phi0.numel() * 8

    def copy(dst, src, nb):
        d_phi.copy_(d_phi0)

    best, table = autotune.tune(b, d_phi.data_ptr(), d_phi0.data_ptr(), nbytes, mode=args.mode,
                                reps=1, sync=torch.cuda.synchronize, copy=copy, log=log)
    log(f"autotune picked {best}")
    L = {k: best[k] for k in ("bin_scale", "query_block", "waves_per_simd", "schedule")}
    for _ in range(args.warmup):
        d_phi.copy_(d_phi0)
        b.reset_phi_batch_device(d_phi.data_ptr(), None, mode=args.mode, stream=stream, **L)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d_phi.copy_(d_phi0)
        b.reset_phi_batch_device(d_phi.data_ptr(), None, mode=args.mode, stream=stream, **L)
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t0) * 1e3 / args.steps
    eager_out = d_phi.cpu().numpy()
    # the same batch as one captured HIP graph per call (captured on the first)
    for _ in range(max(1, args.warmup)):
        d_phi.copy_(d_phi0)
        b.reset_phi_graph_device(d_phi.data_ptr(), None, mode=args.mode, stream=stream, **L)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d_phi.copy_(d_phi0)
        b.reset_phi_graph_device(d_phi.data_ptr(), None, mode=args.mode, stream=stream, **L)
    torch.cuda.synchronize()
    ms_batched = (time.perf_counter() - t0) * 1e3 / args.steps
    batched_out = d_phi.cpu().numpy()
    graph_same = bool(np.array_equal(batched_out.view(np.uint64), eager_out.view(np.uint64)))
    captures, launches = b.graph_stats()
    log(f"batched eager {ms_eager:.2f} ms, graph {ms_batched:.2f} ms "
        f"({captures} capture(s), {launches} launches), graph == eager: {graph_same}")
    # serial plain: one reinit per field on a single-field handle of its size
    singles = {}
    for n, mid in ids.items():
        h = fa.Reinit(0)
        h.set_batch_meshes([mid])
        singles[n] = h
    d_fields = [torch.from_numpy(p).to(dev) for p in phis]
    d_work = [torch.empty_like(p) for p in d_fields]

    def serial():
        for k, n in enumerate(sizes):
            d_work[k].copy_(d_fields[k])
            singles[int(n)].reset_phi_device(d_work[k].data_ptr(), None, mode=args.mode, stream=stream, **L)
    serial()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    serial()
    torch.cuda.synchronize()
    ms_serial = (time.perf_counter() - t0) * 1e3
    serial_out = torch.cat(d_work).cpu().numpy()
    same = graph_same and bool(np.array_equal(batched_out.view(np.uint64), serial_out.view(np.uint64)))
    log(f"batched {ms_batched:.2f} ms, serial plain {ms_serial:.2f} ms, "
        f"speedup {ms_serial / ms_batched:.2f}x, identical results: {same}")
    for h in singles.values():
        h.close()
    b.close()
    for mid in ids.values():
        lib.ft_mesh_release(mid)
    return {
        "metric": f"Mnodes/s batched reinit ({nfields} fields, one captured graph) vs serial plain calls",
        "value": N / (ms_batched * 1e-3) / 1e6, "unit": "Mnodes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_batched,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 distance search; f64 phi and crossing points", "data": "synthetic",
        "config": {"workload": f"batched: {nfields} fields N~U{{{args.min_n}..{args.max_n}}}, 1-16 ellipsoids, "
                               f"{args.mode} distance", "nodes": N, "fields": nfields},
        "serial_plain": {"ms": ms_serial, "Mnodes_per_s": N / (ms_serial * 1e-3) / 1e6},
        "eager_batched": {"ms": ms_eager, "Mnodes_per_s": N / (ms_eager * 1e-3) / 1e6},
        "graph": {"captures": captures, "launches": launches, "identical_to_eager": graph_same},
        "speedup_vs_serial": ms_serial / ms_batched, "identical_to_serial": same,
        "autotune": dict(best, table_ms=[[a, q, w, sch, round(t, 3)] for a, q, w, sch, t in table],
                         table_columns=["bin_scale", "query_block", "waves_per_simd", "schedule", "ms"]),
    }


def run_serving(args):
    """Poisson stream of reinit jobs (sizes log-uniform 32^3..192^3, snapped to
    multiples of 8; phi from the star and 64-sphere generators) served with
    continuous batching against serial plain calls, at each rate.

    Each job's phi is the simulation's own buffer, reinitialised in place:
    in pinned host memory (--phi-on host, the default), or in HBM
    (--phi-on device: a simulation whose state stays on the GPU, as with the
    device advection step).  Serial plain: one reinit call per job, in
    arrival order, on a handle holding that job's resident mesh, with the
    same transfers (none for device phi: the call works on the job's
    buffer)."""
    import ctypes as C
    import torch
    import fluidtissue_amd as fa
    from fluidtissue_amd.serve import Server, latency_stats
    from fluidtissue_amd._lib import check, lib

    rng = np.random.default_rng(11)
    raw = np.exp(rng.uniform(np.log(32), np.log(192), size=args.jobs))
    dims = [int(min(192, max(32, 8 * round(x / 8)))) for x in raw]
    gens = rng.integers(0, 2, size=args.jobs)
    t0 = time.perf_counter()
    fields = {}
    for n in sorted(set(dims)):
        xyz = grid_xyz(n)
        fields[(n, 0)] = torch.from_numpy(fa.phi_field("star", xyz))
        fields[(n, 1)] = torch.from_numpy(fa.phi_field("spheres", xyz, seed=0))
    sizes = [n ** 3 for n in dims]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    on_dev = args.phi_on == "device"
    dev = torch.device("cuda:0")
    if on_dev:
        store = torch.empty(int(offs[-1]), dtype=torch.float64, device=dev)
    else:
        store = torch.empty(int(offs[-1]), dtype=torch.float64, pin_memory=True)
    bufs = [store[int(offs[k]):int(offs[k + 1])] for k in range(args.jobs)]

    def refill():
        for k in range(args.jobs):
            bufs[k].copy_(fields[(dims[k], int(gens[k]))])

    refill()
    log(f"serving inputs: {args.jobs} jobs over {len(set(dims))} sizes, "
        f"{store.numel() * 8 / 1e9:.1f} GB {'in HBM' if on_dev else 'pinned'} "
        f"({time.perf_counter() - t0:.1f} s)")
    kw = dict(mode=args.mode, blend="init")

    # serial plain: a handle per size with its resident mesh, a device buffer per size
    singles = {}
    for n in sorted(set(dims)):
        mid = C.c_int32()
        check(lib.ft_mesh_create_kuhn(0, n, n, n, 0.0, 0, C.byref(mid)))
        h = fa.Reinit(0)
        h.set_batch_meshes([mid.value])
        singles[n] = (h, mid.value, torch.empty(n ** 3, dtype=torch.float64, device=dev))
    sstream = torch.cuda.Stream(device=dev)

    def serial_job(k):
        h, _, d = singles[dims[k]]
        if on_dev:
            h.reset_phi_batch_device(bufs[k].data_ptr(), None, stream=sstream.cuda_stream, **kw)
            sstream.synchronize()
            return
        with torch.cuda.stream(sstream):
            d.copy_(bufs[k], non_blocking=True)
        h.reset_phi_batch_device(d.data_ptr(), None, stream=sstream.cuda_stream, **kw)
        with torch.cuda.stream(sstream):
            bufs[k].copy_(d, non_blocking=True)
        sstream.synchronize()

    for n in sorted(set(dims)):                      # warm the kernels (untimed)
        serial_job(dims.index(n))
    call_ms = {}
    for n in sorted(set(dims)):                      # one plain call per size, median of 5
        k = dims.index(n)
        ts = []
        for _ in range(5):
            t1 = time.perf_counter()
            serial_job(k)
            ts.append((time.perf_counter() - t1) * 1e3)
        call_ms[n] = float(np.median(ts))
    log("serial plain call ms by size: " + ", ".join(f"{n}^3 {t:.2f}" for n, t in call_ms.items()))
    refill()

    def run_serial(arrivals):
        count = len(arrivals)
        start = time.perf_counter()
        lat = []
        t_end = start
        for k in range(count):
            ta = start + arrivals[k]
            now = time.perf_counter()
            if ta > now:
                time.sleep(ta - now)
            serial_job(k)
            t_end = time.perf_counter()
            lat.append(t_end - ta)
        lat = np.array(lat)
        return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                    mean_ms=float(lat.mean() * 1e3), jobs=count, jobs_per_s=count / (t_end - start),
                    Mnodes_per_s=sum(sizes[:count]) / (t_end - start) / 1e6, mean_batch=1.0)

    def run_served(srv, arrivals, count):
        srv.batches.clear()
        start = time.perf_counter()
        jobs = []
        if arrivals is None:                          # saturated: the whole stream at once
            with srv.cv:
                jobs = [srv.submit((n, n, n), bufs[k], t_arrival=start, inplace=True)
                        for k, n in enumerate(dims[:count])]
        else:
            for k in range(count):
                ta = start + arrivals[k]
                now = time.perf_counter()
                if ta > now:
                    time.sleep(ta - now)
                n = dims[k]
                jobs.append(srv.submit((n, n, n), bufs[k], t_arrival=ta, inplace=True))
        for j in jobs:
            j.wait(1200)
        end = max(j.t_done for j in jobs)
        st = latency_stats(jobs)
        st.update(jobs=count, jobs_per_s=count / (end - start),
                  Mnodes_per_s=sum(sizes[:count]) / (end - start) / 1e6,
                  mean_batch=float(np.mean([b[0] for b in srv.batches])))
        return st

    results = {}
    with Server(device=0, lanes=args.lanes, max_fields=64, budget_bytes=200e9, want_ifn=False, **kw) as srv:
        warm = [srv.submit((n, n, n), fields[(n, 0)].clone()) for n in sorted(set(dims))]
        for j in warm:
            j.wait(600)
        for rate in args.rates:
            count = min(args.jobs, getattr(args, "jobs_at_rate", {}).get(rate, args.jobs))
            arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
            res = {}
            refill()
            res["served"] = run_served(srv, arrivals, count)
            log(f"rate {rate}/s served: {res['served']}")
            refill()
            res["serial_plain"] = run_serial(arrivals)
            log(f"rate {rate}/s serial_plain: {res['serial_plain']}")
            res["throughput_ratio"] = res["served"]["Mnodes_per_s"] / res["serial_plain"]["Mnodes_per_s"]
            results[str(rate)] = res
        # saturated capacity: the whole stream queued at once
        sat = {}
        refill()
        sat["served"] = run_served(srv, None, args.jobs)
        log(f"saturated served: {sat['served']}")
        refill()
        sat["serial_plain"] = run_serial(np.zeros(args.jobs))
        log(f"saturated serial_plain: {sat['serial_plain']}")
        sat["throughput_ratio"] = sat["served"]["Mnodes_per_s"] / sat["serial_plain"]["Mnodes_per_s"]
        results["saturated"] = sat
    for h, mid, _ in singles.values():
        h.close()
        lib.ft_mesh_release(mid)
    return {
        "metric": "served reinit p50/p99 latency and throughput vs serial plain (Poisson stream)",
        "value": results["saturated"]["served"]["Mnodes_per_s"], "unit": "Mnodes/s",
        "n_gpus": 1, "steps": args.jobs, "warmup": 0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 distance search; f64 phi and crossing points",
        "data": "synthetic",
        "config": {"workload": f"serving: {args.jobs} jobs, sizes log-uniform 32^3-192^3 (multiples "
                               f"of 8), star/64-sphere phi, {args.mode} distance, phi in "
                               f"{'HBM' if on_dev else 'pinned host memory'} reinitialised in place",
                   "phi_on": args.phi_on, "lanes": args.lanes,
                   "rates_per_s": args.rates},
        "rates": results,
        "serial_call_ms_by_size": {str(n): t for n, t in call_ms.items()},
    }


def host_info():
    """The host's CPU as nproc / lscpu report it (the CPU baseline's threads)."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket",
                             "Socket(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except Exception as e:                      # noqa: BLE001 (informational only)
        info["lscpu"] = f"unavailable: {e}"
    share = info["nproc"]
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            info["cgroup cpu.max"] = f.read().strip()
        quota, period = info["cgroup cpu.max"].split()[:2]
        if quota != "max":
            share = min(share, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        info["OMP_NUM_THREADS"] = int(os.environ["OMP_NUM_THREADS"])
        share = min(share, info["OMP_NUM_THREADS"])
    # the threads the CPU baseline runs on: every hardware thread this
    # process may use (LevelSet.hpp:38-39 takes hardware_concurrency()),
    # within the affinity mask, the cgroup CPU quota and OMP_NUM_THREADS
    info["threads_used"] = share
    return info


def gpu_sample(r, phi, mode, band, stream, launch):
    """One more reinit (untimed) keeping the nearest primitive of every node:
    the GPU side of the accuracy check (phi_ifn, nearest index, interface)."""
    import torch
    d = torch.from_numpy(phi).cuda()
    r.reset_phi_device(d.data_ptr(), None, mode=mode, blend="init", band=band, want_nearest=True,
                       stream=stream, **launch)
    torch.cuda.synchronize()
    return dict(ifn=d.cpu().numpy(), near=r.nearest(), itf=r.interface())


def accuracy_check(m, phi, h, mode, band, gpu, n_sample=2000, seed=3, nodes=None):
    """The oracle as the checker of the GPU sample (run in the CPU leg): the
    interface against the oracle's marching tets (bit-exact), every sign,
    and a seeded node sample's distances against the oracle's FP64 brute
    force over all primitives; in triangle mode also against a second,
    independent checker (oracle/tri_qp.py, long double) on the triangles the
    GPU and the oracle chose."""
    import oracle
    from oracle import tri_qp
    t0 = time.perf_counter()
    ifn, near, itf = gpu["ifn"], gpu["near"], gpu["itf"]
    ifp, ifp2e = oracle.find_interface_nodes(m.edges, m.xyz, phi)
    tris, elems, area = oracle.reconstruct_interface(m.tets, m.t2e, m.edges, m.xyz, phi, ifp, ifp2e)
    exact = bool(np.array_equal(itf.ifp.view(np.uint32), ifp.view(np.uint32)) and
                 np.array_equal(itf.ifp2e, ifp2e) and np.array_equal(itf.tris, tris) and
                 np.array_equal(itf.if_elems, elems) and
                 np.array_equal(itf.ifp_area.view(np.uint32), area.view(np.uint32)))
    if nodes is None:
        rng = np.random.default_rng(seed)
        nodes = rng.choice(m.n_nodes, n_sample, replace=False)
    n_sample = len(nodes)
    got = np.abs(ifn[nodes])
    signs = bool(np.array_equal(np.signbit(ifn), phi < 0.0))
    if mode == "points":
        d64, j64 = oracle.point_distance(m.xyz, ifp, nodes)
        d64 = d64.astype(np.float64)
    else:
        d64, j64 = oracle.triangle_distance(m.xyz, ifp, tris, nodes)
    inb = np.ones(len(nodes), bool)
    clamped_ok = True
    if band > 0:
        outb = d64 > band * (1 + 1e-6)
        clamped_ok = bool(np.all(near[nodes][outb] == -1) and
                          np.all(got[outb] == np.float64(np.float32(band))))
        inb = d64 < band * (1 - 1e-6)
    err = np.abs(got[inb] - d64[inb])
    ties = mism = 0
    eps = 1e-5 * h
    for k in np.nonzero(inb & (near[nodes] != j64))[0]:
        if mode == "points":
            dg = float(np.sqrt(np.sum((m.xyz[nodes[k]].astype(np.float64) -
                                       ifp[near[nodes[k]]].astype(np.float64)) ** 2)))
        else:
            dg = oracle.node_triangle_distance(nodes[k], m.xyz, ifp, tris, near[nodes[k]])
        if abs(dg - d64[k]) <= eps:
            ties += 1
        else:
            mism += 1
    out = {"nodes_checked": int(inb.sum()), "sample": n_sample, "seed": seed,
           "max_err": float(err.max()) if err.size else 0.0,
           "max_err_h": float(err.max() / h) if err.size else 0.0,
           "rms_err_h": float(np.sqrt((err ** 2).mean()) / h) if err.size else 0.0,
           "tolerance_h": 1e-5, "eps_ties": ties, "index_mismatches": mism,
           "signs_bit_exact_all_nodes": signs, "interface_bit_exact": exact,
           "reference": "FP64 brute force over the oracle's interface (all primitives)"}
    if mode == "triangles":
        sel = nodes[inb]
        fg = near[sel]
        ok = fg >= 0
        q_gpu = tri_qp.node_distance(m.xyz, ifp, tris, sel[ok], fg[ok])
        q_ora = tri_qp.node_distance(m.xyz, ifp, tris, sel, j64[inb])
        e_ind = np.abs(got[inb][ok].astype(np.longdouble) - q_gpu).astype(np.float64)
        out["independent"] = {
            "max_err_h": float(e_ind.max() / h) if e_ind.size else 0.0,
            "rms_err_h": float(np.sqrt((e_ind ** 2).mean()) / h) if e_ind.size else 0.0,
            "oracle_vs_checker_max_h": float(np.max(np.abs(q_ora.astype(np.float64) - d64[inb])) / h),
            "reference": "the GPU's distance against its chosen triangle's distance from a second "
                         "checker (constrained quadratic minimization, long double)"}
    if band > 0:
        out["outside_band_clamped_exactly"] = clamped_ok
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def batched_fields(k_count, seed=7, lo=24, hi=96):
    """The first k_count fields of the batched workload: sizes N ~ U{lo..hi}
    (seeded as run_batched draws them) and each field's ellipsoids."""
    import fluidtissue_amd as fa
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=512)[:k_count]
    out = []
    for k, n in enumerate(sizes):
        m = fa.kuhn_mesh(int(n))
        out.append((int(n), m, fa.phi_field("ellipsoids", m.xyz, seed=1000 + k)))
    return out


def config_samples(mode):
    """GPU side of the per-config accuracy (no oracle here): one reinit of
    each of the other configs' fields, keeping the nearest primitive.
      sphere_wave_32: the 32^3 field, full domain, every node checked later;
      star_128 band: the 6h band, 1500 nodes the GPU puts inside the band
        and 500 anywhere;
      batched: 4 fields of the 512-field batch, reinitialised as one batch
        through the captured graph and each also alone; the batch's phi must
        equal the solo runs' bit for bit, and the solo runs carry the
        interface and the nearest index;
      served sizes: a 96^3 jittered field of the 64 spheres (the serving
        stream's sizes are 32^3-192^3 from these generators)."""
    import torch
    import fluidtissue_amd as fa
    out = {}
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream      # the stream torch copies phi on

    def solo(m, phi, band):
        r = fa.Reinit(0)
        r.set_mesh(m)
        g = gpu_sample(r, phi, mode, band, st, {})
        r.close()
        return g

    m = fa.kuhn_mesh(32)
    phi = fa.phi_field("sphere_wave", m.xyz)
    out["sphere_wave_32"] = dict(m=m, phi=phi, h=1.0 / 31, band=0.0, gpu=solo(m, phi, 0.0), all_nodes=True)
    m = fa.kuhn_mesh(128)
    phi = fa.phi_field("star", m.xyz)
    h = 1.0 / 127
    g = solo(m, phi, 6.0 * h)
    rng = np.random.default_rng(11)
    inside = np.nonzero(np.abs(g["ifn"]) < 6.0 * h)[0]
    nodes = np.unique(np.concatenate([rng.choice(inside, min(1500, len(inside)), replace=False),
                                      rng.choice(m.n_nodes, 500, replace=False)]))
    out["star_128_band6h"] = dict(m=m, phi=phi, h=h, band=6.0 * h, gpu=g, nodes=nodes)
    fields = batched_fields(4)
    b = fa.Reinit(0)
    b.set_batch([m for _, m, _ in fields])
    d = torch.from_numpy(np.concatenate([p for _, _, p in fields])).to(dev)
    b.reset_phi_graph_device(d.data_ptr(), None, mode=mode, stream=st)
    torch.cuda.synchronize()
    batch_phi = d.cpu().numpy()
    b.close()
    off = 0
    for k, (n, m, phi) in enumerate(fields):
        g = solo(m, phi, 0.0)
        same = bool(np.array_equal(batch_phi[off:off + m.n_nodes].view(np.uint64), g["ifn"].view(np.uint64)))
        off += m.n_nodes
        out[f"batched_field{k}_{n}^3"] = dict(m=m, phi=phi, h=1.0 / (n - 1), band=0.0, gpu=g,
                                             batch_equals_solo=same, n_sample=1000)
    m = fa.kuhn_mesh(96, jitter=0.2, seed=1)
    phi = fa.phi_field("spheres", m.xyz, seed=0)
    out["served_size_spheres_96"] = dict(m=m, phi=phi, h=1.0 / 95, band=0.0, gpu=solo(m, phi, 0.0))
    return out


def config_accuracy(samples, mode):
    """CPU leg: the oracle checks each config's GPU sample."""
    res = {}
    for name, c in samples.items():
        m = c["m"]
        nodes = np.arange(m.n_nodes) if c.get("all_nodes") else c.get("nodes")
        a = accuracy_check(m, c["phi"], c["h"], mode, c["band"], c["gpu"], n_sample=c.get("n_sample", 2000),
                           nodes=nodes)
        a["nodes_total"] = int(m.n_nodes)
        if "batch_equals_solo" in c:
            a["batch_equals_solo_bitwise"] = c["batch_equals_solo"]
        res[name] = a
    return res


def _read_counters(dirpath):
    """{(kernel, dispatch): {counter: value}} from a rocprofv3 --pmc csv run."""
    import csv
    import glob
    rows = {}
    for fn in glob.glob(os.path.join(dirpath, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for row in csv.DictReader(f):
                key = (row.get("Kernel_Name", ""), row.get("Dispatch_Id", ""))
                d = rows.setdefault(key, {})
                name = row.get("Counter_Name", "")
                d[name] = d.get(name, 0.0) + float(row.get("Counter_Value", 0.0))
                try:                            # the dispatch's own duration, ns
                    d["duration_ns"] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
                except (KeyError, ValueError):
                    pass
    return rows


def live_counters(args, N, q_ms, out_root):
    """rocprofv3 counter passes of one reinit of this workload (a child run of
    this script per pass, `--profile-child`): FETCH_SIZE, WRITE_SIZE and the
    instruction counts of the query kernel, scaled by what the same counters
    read for k_node_signs, whose traffic is known (it reads 8 B of phi per
    node and writes 1 bit per node)."""
    import shutil
    import subprocess
    prof = shutil.which("rocprofv3")
    if prof is None:
        return {"status": "rocprofv3 not found"}
    child = [sys.executable, os.path.abspath(__file__), "--profile-child", "--workload", args.workload,
             "--mode", args.mode, "--bin-scale", str(args.bin_scale), "--query-block", str(args.query_block),
             "--waves-per-simd", str(args.waves_per_simd), "--schedule", args.schedule]
    passes = {"fetch": ["FETCH_SIZE"], "write": ["WRITE_SIZE"],
              "sq": ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_SMEM"]}
    env = dict(os.environ, TMPDIR="/tmp")
    got = {}
    cmds = []
    for name, ctrs in passes.items():
        d = os.path.join(out_root, name)
        shutil.rmtree(d, ignore_errors=True)
        cmd = [prof, "--pmc"] + ctrs + ["--output-format", "csv", "-d", d, "-o", name, "--"] + child
        cmds.append(" ".join(["rocprofv3", "--pmc"] + ctrs + ["--output-format", "csv", "-d", f"<out>/{name}",
                                                             "-o", name, "--", "python3", "bench.py"] + child[2:]))
        t0 = time.perf_counter()
        try:
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=180,
                           cwd="/tmp", env=env)
        except subprocess.TimeoutExpired:
            return {"status": f"counter pass {name} timed out"}
        got[name] = _read_counters(d)
        log(f"counter pass {name}: {len(got[name])} dispatches ({time.perf_counter() - t0:.1f} s)")

    def last(rows, kname, ctr):
        ks = [k for k in rows if kname in k[0] and ctr in rows[k]]
        if not ks:
            return None
        k = max(ks, key=lambda k: int(k[1]) if k[1].isdigit() else 0)
        return rows[k][ctr]

    qk = "k_query_wave"
    fetch_raw = last(got["fetch"], qk, "FETCH_SIZE")
    write_raw = last(got["write"], qk, "WRITE_SIZE")
    sig_f = last(got["fetch"], "k_node_signs", "FETCH_SIZE")
    sig_w = last(got["write"], "k_node_signs", "WRITE_SIZE")
    if None in (fetch_raw, write_raw, sig_f, sig_w):
        return {"status": "counters missing from the rocprofv3 output", "commands": cmds}
    f_scale = 8.0 * N / (sig_f * 1024.0)
    w_scale = (N / 8.0) / (sig_w * 1024.0)
    traffic = fetch_raw * 1024.0 * f_scale + write_raw * 1024.0 * w_scale
    res = {"status": "ok", "traffic_bytes": traffic, "fetch_kib_raw": fetch_raw, "write_kib_raw": write_raw,
           "fetch_scale": f_scale, "write_scale": w_scale,
           "calibration": "k_node_signs: reads 8 B/node of phi, writes 1 bit/node", "commands": cmds}
    waves = last(got["sq"], qk, "SQ_WAVES")
    valu = last(got["sq"], qk, "SQ_INSTS_VALU")
    if waves and valu:
        res["waves"] = waves
        res["valu_per_wave"] = valu / waves
        res["salu_per_wave"] = (last(got["sq"], qk, "SQ_INSTS_SALU") or 0.0) / waves
        res["smem_per_wave"] = (last(got["sq"], qk, "SQ_INSTS_SMEM") or 0.0) / waves
        # each wave64 VALU instruction issues over 4 cycles of a 16-lane SIMD;
        # 256 CUs x 4 SIMDs at 2.4 GHz, over the profiled dispatch's own
        # duration (its timestamps in the same counter pass)
        dur = last(got["sq"], qk, "duration_ns")
        if dur:
            res["query_dispatch_ms_profiled"] = dur * 1e-6
            res["valu_issue_frac"] = valu * 4.0 / (1024 * 2.4e9 * dur * 1e-9)

    return res


def profile_child(args):
    """One reinit of the workload (after one warm-up), for the counter passes."""
    import torch
    import fluidtissue_amd as fa
    wl = WORKLOADS[args.workload]
    m = fa.kuhn_mesh(wl["n"], jitter=wl["jitter"], seed=1)
    phi = fa.phi_field(wl["phi"], m.xyz, seed=0)
    band = wl.get("band_h", 0.0) / (wl["n"] - 1)
    r = fa.Reinit(0)
    r.set_mesh(m)
    d0 = torch.from_numpy(phi).cuda()
    for _ in range(2):
        d = d0.clone()
        torch.cuda.synchronize()
        r.reset_phi_device(d.data_ptr(), None, mode=args.mode, blend="init", band=band,
                           bin_scale=args.bin_scale, query_block=args.query_block,
                           waves_per_simd=args.waves_per_simd, schedule=args.schedule)
        torch.cuda.synchronize()
    r.close()


def served_block(args):
    """The served form at the default workload's side: the serving stream
    (sizes log-uniform 32^3-192^3) at each rate and saturated, served against
    serial plain calls; p50/p99 and throughput ratios."""
    sa = argparse.Namespace(**vars(args))
    sa.jobs = args.served_jobs
    sa.rates = args.served_rates
    sa.jobs_at_rate = {50.0: min(args.served_jobs, 500)}
    res = run_serving(sa)
    out = {"jobs": sa.jobs, "lanes": res["config"]["lanes"], "phi_on": res["config"]["phi_on"],
           "note": "jobs per rate: " + ", ".join(f"{r:g}/s {sa.jobs_at_rate.get(r, sa.jobs)}" for r in sa.rates)
                   + f"; saturated {sa.jobs}"}
    for key, v in res["rates"].items():
        out[key] = {"served_p50_ms": v["served"]["p50_ms"], "served_p99_ms": v["served"]["p99_ms"],
                    "serial_p50_ms": v["serial_plain"]["p50_ms"], "serial_p99_ms": v["serial_plain"]["p99_ms"],
                    "served_Mnodes_per_s": v["served"]["Mnodes_per_s"],
                    "serial_Mnodes_per_s": v["serial_plain"]["Mnodes_per_s"],
                    "throughput_ratio": v["throughput_ratio"], "mean_batch": v["served"]["mean_batch"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="spheres_256", choices=sorted(WORKLOADS) + ["batched", "serving", "advect", "normals"])
    ap.add_argument("--advect-n", type=int, default=256)
    ap.add_argument("--fields", type=int, default=512)
    ap.add_argument("--min-n", type=int, default=24, help="batched: smallest field side")
    ap.add_argument("--max-n", type=int, default=96, help="batched: largest field side")
    ap.add_argument("--jobs", type=int, default=2000)
    ap.add_argument("--rates", type=float, nargs="+", default=[50.0, 200.0, 500.0])
    ap.add_argument("--lanes", type=int, default=2, help="serving: worker lanes (handles and streams)")
    ap.add_argument("--phi-on", choices=["host", "device"], default="host",
                    help="serving: where each job's phi lives")
    ap.add_argument("--mode", default="triangles", choices=["triangles", "points"])
    ap.add_argument("--bin-scale", type=float, default=0.0)
    ap.add_argument("--query-block", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=0)
    ap.add_argument("--schedule", default="static", choices=["static", "queue", "bins"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-served", action="store_true", help="skip the served block")
    ap.add_argument("--served-jobs", type=int, default=2000)
    ap.add_argument("--served-rates", type=float, nargs="+", default=[50.0, 200.0, 500.0])
    ap.add_argument("--profile-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one field on one GPU; --gpus must be 1 "
              "(sharding one field over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    import fluidtissue_amd as fa

    if args.profile_child:
        profile_child(args)
        return
    if args.workload == "batched":
        print(json.dumps(run_batched(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(run_serving(args)), flush=True)
        return
    if args.workload == "advect":
        print(json.dumps(run_advect(args)), flush=True)
        return
    if args.workload == "normals":
        print(json.dumps(run_normals(args)), flush=True)
        return
    wl = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    m = fa.kuhn_mesh(wl["n"], jitter=wl["jitter"], seed=1)
    phi = fa.phi_field(wl["phi"], m.xyz, seed=0)
    h = 1.0 / (wl["n"] - 1)
    band = wl.get("band_h", 0.0) * h
    log(f"inputs: {m.n_nodes} nodes {len(m.tets)} tets {len(m.edges)} edges "
        f"({time.perf_counter() - t0:.1f} s)")

    dev = torch.device("cuda:0")
    # one stream for the input copy and the reinit, so each step's reinit
    # starts after its copy of phi
    bstream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(bstream)
    stream = bstream.cuda_stream
    r = fa.Reinit(0)
    r.set_mesh(m)
    d_phi0 = torch.from_numpy(phi).to(dev)
    d_phi = torch.empty_like(d_phi0)

    def step():
        d_phi.copy_(d_phi0)
        return r.reset_phi_device(d_phi.data_ptr(), None, mode=args.mode, blend="init",
                                  band=band, bin_scale=args.bin_scale, stream=stream,
                                  query_block=args.query_block, waves_per_simd=args.waves_per_simd,
                                  schedule=args.schedule)

    tuned = None
    if not args.no_tune:
        from fluidtissue_amd import autotune

        def copy(dst, src, nb):
            d_phi.copy_(d_phi0)
        best, table = autotune.tune(r, d_phi.data_ptr(), d_phi0.data_ptr(), d_phi0.numel() * 8,
                                    mode=args.mode, reps=3, sync=torch.cuda.synchronize, copy=copy,
                                    bin_scales=autotune.MODE_BIN_SCALES[args.mode], query_blocks=(64, 256),
                                    band=band, stream=stream)
        args.bin_scale, args.query_block = best["bin_scale"], best["query_block"]
        args.waves_per_simd, args.schedule = best["waves_per_simd"], best["schedule"]
        tuned = dict(best, table_ms=[[a, q, w, sch, round(t, 3)] for a, q, w, sch, t in table],
                     table_columns=["bin_scale", "query_block", "waves_per_simd", "schedule", "ms"])
        log(f"autotune picked {best}")
    if args.bin_scale == 0.0:
        args.bin_scale = 1.25 if args.mode == "triangles" else 2.0
    if args.query_block == 0:
        args.query_block = 256
    if args.waves_per_simd == 0:
        args.waves_per_simd = 8
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    r.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cnt = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phase_ms, calls = r.timing()
    r.set_timing(False)
    log(f"interface: P={cnt.n_ifp} F={cnt.n_tri} Tx={cnt.n_if_elems}")
    ms = (t1 - t0) * 1e3 / args.steps
    N, T = m.n_nodes, len(m.tets)
    P, F, X = cnt.n_ifp, cnt.n_tri, cnt.n_if_elems
    ph = [p / max(calls, 1) for p in phase_ms]
    log(f"step {ms:.3f} ms; phases: marching tets {ph[0]:.3f} ms, index {ph[1]:.3f} ms, "
        f"query {ph[2]:.3f} ms, normals {ph[3]:.3f} ms")

    # Algorithmic bytes per reinit: B = 28N + 16T + 20P + 24Tx + 16F (coords
    # 12 + phi in 8 + phi out 8 per node; tet connectivity once; crossing-edge
    # endpoints and ifp per point; t2e per crossing tet; triangle indices).
    # The query kernel's own share is 28N.
    b_path = 28 * N + 16 * T + 20 * P + 24 * X + 16 * F
    b_query = 28 * N
    t_query = ph[2] * 1e-3
    peak = 8000.0

    sample = None
    if not args.no_accuracy and not args.no_cpu:
        sample = gpu_sample(r, phi, args.mode, band, stream,
                            dict(bin_scale=args.bin_scale, query_block=args.query_block,
                                 waves_per_simd=args.waves_per_simd, schedule=args.schedule))

    cfg_samples = None
    if not args.no_accuracy and not args.no_cpu and args.workload == "spheres_256":
        t_c = time.perf_counter()
        cfg_samples = config_samples(args.mode)
        log(f"per-config GPU samples: {len(cfg_samples)} ({time.perf_counter() - t_c:.1f} s)")

    # the drop-in path with host phi (ft_reinit on the caller's own arrays, as
    # LevelSet::resetPhi_ would call it: H2D of phi, the reinit, D2H of phi and
    # phi_ifn); reported beside the device-resident rate, never as `value`
    import ctypes as C
    from fluidtissue_amd import _lib
    from fluidtissue_amd.reinit import make_opts
    h_phi = phi.copy()
    h_ifn = np.zeros_like(phi)
    o = make_opts(args.mode, "init", band, args.bin_scale, False, query_block=args.query_block,
                  waves_per_simd=args.waves_per_simd, schedule=args.schedule)
    hc = _lib.FtCounts()
    ms_host = []
    for _ in range(3):
        np.copyto(h_phi, phi)
        t_h = time.perf_counter()
        _lib.check(_lib.lib.ft_reinit(r._h, h_phi.ctypes.data_as(C.c_void_p), h_ifn.ctypes.data_as(C.c_void_p),
                                      C.byref(o), C.byref(hc)))
        ms_host.append((time.perf_counter() - t_h) * 1e3)
    ms_host = float(np.median(ms_host))
    log(f"host-phi (PCIe-inclusive) reinit: {ms_host:.3f} ms")
    r.close()
    del d_phi, d_phi0
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.set_stream(torch.cuda.default_stream(dev))

    counters = None
    if not args.no_counters:
        # raw counter files: FT_BENCH_PROFILE_DIR if set, else a temporary
        # folder (the JSON line carries what they say)
        import tempfile
        out_root = os.environ.get("FT_BENCH_PROFILE_DIR") or tempfile.mkdtemp(prefix="ft_bench_counters_")
        counters = live_counters(args, N, ph[2], out_root)
        log(f"counters: { {k: v for k, v in counters.items() if k != 'commands'} }")

    # The CPU leg: the oracle on the host, as the reference CPU path's timed
    # baseline and as the checker of the GPU sample above.
    cpu = accuracy = accuracy_cfg = None
    hinfo = host_info()
    if not args.no_cpu:
        cpu = cpu_baseline(m, phi, hinfo["threads_used"])
        log(f"cpu baseline: {cpu['value']:.3f} Mnodes/s ({cpu['sample']})")
        cpu.pop("seconds")
        cpu["host"] = hinfo
        if sample is not None:
            accuracy = accuracy_check(m, phi, h, args.mode, band, sample)
            log(f"accuracy: {accuracy}")
        if cfg_samples is not None:
            t_c = time.perf_counter()
            accuracy_cfg = config_accuracy(cfg_samples, args.mode)
            for k, v in accuracy_cfg.items():
                log(f"accuracy {k}: max {v['max_err_h']:.3g} h, rms {v['rms_err_h']:.3g} h, "
                    f"mismatches {v['index_mismatches']}, ties {v['eps_ties']}, "
                    f"interface exact {v['interface_bit_exact']}, signs exact {v['signs_bit_exact_all_nodes']}")
            log(f"per-config accuracy: {time.perf_counter() - t_c:.1f} s")

    served = None
    if not args.no_served:
        served = served_block(args)
        log(f"served: {served}")

    traffic = counters.get("traffic_bytes") if counters and counters.get("status") == "ok" else None
    valu_frac = counters.get("valu_issue_frac") if counters and counters.get("status") == "ok" else None
    achieved = b_path / t_query / 1e9
    out = {
        "metric": "Mnodes/s reinit at 256^3 + served p50/p99 vs serial plain, <=1e-5*h err",
        "value": N / (ms * 1e-3) / 1e6,
        "unit": "Mnodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 distance search; f64 phi, crossing points and thin triangles",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {wl['desc']}, {args.mode} distance, "
                               + (f"narrow band {wl['band_h']:g}h" if band > 0 else "full domain"),
                   "nodes": N, "tets": T, "edges": len(m.edges), "crossing_points": P,
                   "triangles": F, "crossing_tets": X, "h": h, "band": band,
                   "clamped_nodes": cnt.n_clamped, "bin_scale": args.bin_scale,
                   "query_block": args.query_block, "waves_per_simd": args.waves_per_simd,
                   "schedule": args.schedule},
        "phases_ms": {"marching_tets": ph[0], "index_build": ph[1], "query": ph[2],
                      "normals": ph[3]},
        "roofline": {"kernel": f"k_query_wave<{1 if args.mode == 'triangles' else 0},"
                               f"{args.query_block},{args.waves_per_simd if args.mode == 'triangles' else 8}> "
                               f"(nearest-primitive search, {args.schedule} schedule)",
                     "bound": "valu", "achieved": achieved, "peak": peak, "unit": "GB/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "basis": "achieved = B / query-kernel time, B = 28N + 16T + 20P + 24Tx + 16F "
                              "bytes per reinit; the kernel is bound by vector instruction issue",
                     "valu_issue_frac": valu_frac,
                     "query_kernel_ms": ph[2],
                     "query_own_bytes": b_query,
                     "query_own_frac": b_query / t_query / 1e9 / peak,
                     "step_frac": b_path / (ms * 1e-3) / 1e9 / peak},
        "path_roofline": {"algorithmic_bytes": b_path,
                          "achieved_GBps": b_path / (ms * 1e-3) / 1e9,
                          "frac": b_path / (ms * 1e-3) / 1e9 / peak},
        "accuracy": accuracy,
        "accuracy_by_config": accuracy_cfg,
        "counters": counters,
        "served": served,
        "cpu_baseline": cpu,
        "host_phi_reinit": {"ms": ms_host, "Mnodes_per_s": N / (ms_host * 1e-3) / 1e6,
                            "note": "ft_reinit on the caller's pageable host arrays (median of 3): H2D of "
                                    "phi + reinit + D2H of phi and phi_ifn, staged through pinned chunks"},
        "autotune": tuned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
