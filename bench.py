"""Benchmark of the level-set reinit on one MI355X.

One step = one full reinit of the 256^3 workload with phi resident in HBM:
marching tets (crossing points, triangles, areas, p2t), the uniform-bin
index of the triangles, and the full-domain nearest-triangle signed distance
of every node with the blend fused in (LevelSet::resetPhi_ + LevelSet::init's
reset, src/fem/LevelSet.cpp:73-74, 429-455).  Each step first restores the
same input phi with a device copy (included in the timed region).

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
    "cfg1": dict(n=32, jitter=0.0, phi="sphere_wave", desc="32^3 Kuhn, sphere+wave"),
    "cfg2": dict(n=128, jitter=0.0, phi="star", desc="128^3 Kuhn, star surface"),
    "cfg3": dict(n=256, jitter=0.2, phi="spheres",
                 desc="256^3 Kuhn, interior nodes jittered +-0.2h, 64 spheres"),
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg3", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="triangles", choices=["triangles", "points"])
    ap.add_argument("--bin-scale", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one field on one GPU; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    import fluidtissue_amd as fa

    wl = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    m = fa.kuhn_mesh(wl["n"], jitter=wl["jitter"], seed=1)
    phi = fa.phi_field(wl["phi"], m.xyz, seed=0)
    h = 1.0 / (wl["n"] - 1)
    log(f"inputs: {m.n_nodes} nodes {len(m.tets)} tets {len(m.edges)} edges "
        f"({time.perf_counter() - t0:.1f} s)")

    dev = torch.device("cuda:0")
    r = fa.Reinit(0)
    r.set_mesh(m)
    d_phi0 = torch.from_numpy(phi).to(dev)
    d_phi = torch.empty_like(d_phi0)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        d_phi.copy_(d_phi0)
        return r.reset_phi_device(d_phi.data_ptr(), None, mode=args.mode, blend="init",
                                  bin_scale=args.bin_scale, stream=stream)

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
    log(f"interface: P={cnt.n_ifp} F={cnt.n_tri} Tx={cnt.n_if_elems}")
    ms = (t1 - t0) * 1e3 / args.steps
    N, T = m.n_nodes, len(m.tets)
    P, F, X = cnt.n_ifp, cnt.n_tri, cnt.n_if_elems
    ph = [p / max(calls, 1) for p in phase_ms]
    log(f"step {ms:.3f} ms; phases: marching tets {ph[0]:.3f} ms, index {ph[1]:.3f} ms, "
        f"query {ph[2]:.3f} ms")

    # Algorithmic bytes (per reinit): 28N + 16T + 20P + 24Tx + 16F; the query
    # kernel's share is 28N (coords 12 + phi_old 8 + phi_out 8).
    b_path = 28 * N + 16 * T + 20 * P + 24 * X + 16 * F
    b_query = 28 * N
    t_query = ph[2] * 1e-3
    peak = 8000.0
    achieved = b_query / t_query / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.mode}.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("query_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        nthreads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(m, phi, nthreads)
        log(f"cpu baseline: {cpu['value']:.3f} Mnodes/s ({cpu['sample']})")
        cpu.pop("seconds")

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
        "config": {"workload": f"{args.workload}: {wl['desc']}, {args.mode} distance, full domain",
                   "nodes": N, "tets": T, "edges": len(m.edges), "crossing_points": P,
                   "triangles": F, "crossing_tets": X, "h": h},
        "phases_ms": {"marching_tets": ph[0], "index_build": ph[1], "query": ph[2]},
        "roofline": {"kernel": "k_query (nearest-triangle search + epilogue)", "bound": "hbm",
                     "achieved": achieved, "peak": peak, "unit": "GB/s",
                     "frac": achieved / peak, "traffic": traffic},
        "path_roofline": {"algorithmic_bytes": b_path,
                          "achieved_GBps": b_path / (ms * 1e-3) / 1e9,
                          "frac": b_path / (ms * 1e-3) / 1e9 / peak},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
