"""Autotuner for the reinit's index and query launch shape.

Tunes, on the caller's own resident batch and level set:
  bin_scale       the bin edge of the uniform-bin index, in units of each
                  field's node spacing (leaf size: the balance between tree
                  depth and primitives per leaf);
  query_block     threads per workgroup of the query (1, 2 or 4 waves: LDS
                  footprint and how workgroups spread over the XCDs);
  waves_per_simd  the query's launch bound, 5..8 waves per SIMD (registers
                  against occupancy; triangles mode only);
  schedule        one tile per wave in a fixed deep-first order ("static"),
                  or persistent waves taking tiles from per-XCD counters
                  ("queue"); in a narrow band in triangles mode also the
                  bins staged in LDS instead of the tree walk ("bins").
Each candidate runs the full reinit; the fastest median wins.  The search
is staged: bin_scale x query_block at 8 waves, static; then waves_per_simd
at the best of those; then the schedule.  Results are cached per (mode,
band, batch signature).
"""
import time

import numpy as np

# bin edges tried per mode: the neighbourhood of each mode's measured optimum
# (points 2.0-2.5, triangles 1.25 node spacings on 32^3..256^3 fields)
MODE_BIN_SCALES = {"points": (1.5, 2.0, 2.5, 3.0), "triangles": (1.0, 1.25, 1.5, 2.0)}
QUERY_BLOCKS = (64, 128, 256)
WAVES_PER_SIMD = (5, 6, 7, 8)
SCHEDULES = ("static", "queue")

_cache = {}


def tune(handle, phi_ptr, phi_backup_ptr, nbytes, mode="triangles", reps=2, sync=None, copy=None,
         signature=None, bin_scales=None, query_blocks=QUERY_BLOCKS, waves=None, schedules=SCHEDULES,
         log=None, band=0.0, stream=None):
    """handle: a Reinit with its batch set; phi_ptr: device phi (overwritten);
    phi_backup_ptr: device copy of the input, restored before every run by
    copy(dst, src, nbytes); sync(): waits for the device; stream: the HIP
    stream of the runs (None: the handle's).  waves: launch bounds to try
    (default 5..8 in triangles mode, 8 in points mode).  Returns (best,
    table): best = dict(bin_scale, query_block, waves_per_simd, schedule, ms);
    table rows (bin_scale, query_block, waves_per_simd, schedule, ms)."""
    if bin_scales is None:
        bin_scales = MODE_BIN_SCALES[mode]
    if waves is None:
        waves = WAVES_PER_SIMD if mode == "triangles" else (8,)
    key = (mode, band, signature)
    if signature is not None and key in _cache:
        return _cache[key]
    table = []
    seen = {}

    def run(bs, qb, wps, sch):
        if (bs, qb, wps, sch) in seen:
            return seen[(bs, qb, wps, sch)]
        ts = []
        for _ in range(reps + 1):
            copy(phi_ptr, phi_backup_ptr, nbytes)
            sync()
            t0 = time.perf_counter()
            handle.reset_phi_batch_device(phi_ptr, None, mode=mode, band=band, bin_scale=bs,
                                          query_block=qb, waves_per_simd=wps, schedule=sch,
                                          stream=stream)
            sync()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts[1:]) * 1e3)
        seen[(bs, qb, wps, sch)] = ms
        table.append((bs, qb, wps, sch, ms))
        if log:
            log(f"autotune: bin_scale {bs} query_block {qb} waves/SIMD {wps} {sch}: {ms:.3f} ms")
        return ms

    w0 = 8 if 8 in waves else waves[-1]
    s0 = schedules[0]
    if band > 0.0 and mode == "triangles" and "bins" not in schedules:
        schedules = tuple(schedules) + ("bins",)
    best = min(((run(bs, qb, w0, s0), bs, qb) for bs in bin_scales for qb in query_blocks))
    _, bs, qb = best
    best_w = min(((run(bs, qb, w, s0), w) for w in waves))[1]
    best_s = min(((run(bs, qb, best_w, sch), sch) for sch in schedules))[1]
    res = (dict(bin_scale=bs, query_block=qb, waves_per_simd=best_w, schedule=best_s,
                ms=seen[(bs, qb, best_w, best_s)]), table)
    if signature is not None:
        _cache[key] = res
    return res
