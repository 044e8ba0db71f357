"""Autotuner for the reinit's index and query launch shape.

Tunes, on the caller's own resident batch and level set, the bin edge of the
uniform-bin index (in units of each field's node spacing: leaf size, hence
the balance between tree depth and primitives per leaf) and the query's
workgroup size (1, 2 or 4 waves per workgroup: waves per CU and LDS
footprint).  Each candidate runs the full reinit; the fastest median wins.
Results are cached per (mode, batch signature).
"""
import time

import numpy as np

# bin edges tried per mode: the neighbourhood of each mode's measured optimum
# (points 2.0-2.5, triangles 1.25 node spacings on 32^3..256^3 fields)
MODE_BIN_SCALES = {"points": (1.5, 2.0, 2.5, 3.0), "triangles": (1.0, 1.25, 1.5, 2.0)}
QUERY_BLOCKS = (64, 128, 256)

_cache = {}


def tune(handle, phi_ptr, phi_backup_ptr, nbytes, mode="triangles", reps=2, sync=None, copy=None,
         signature=None, bin_scales=None, query_blocks=QUERY_BLOCKS, log=None, band=0.0, stream=None):
    """handle: a Reinit with its batch set; phi_ptr: device phi (overwritten);
    phi_backup_ptr: device copy of the input, restored before every run by
    copy(dst, src, nbytes); sync(): waits for the device; stream: the HIP
    stream of the runs (None: the handle's).  Returns (bin_scale,
    query_block, table of (bin_scale, query_block, ms))."""
    if bin_scales is None:
        bin_scales = MODE_BIN_SCALES[mode]
    key = (mode, band, signature)
    if signature is not None and key in _cache:
        return _cache[key]
    table = []
    best = None
    for bs in bin_scales:
        for qb in query_blocks:
            ts = []
            for _ in range(reps + 1):
                copy(phi_ptr, phi_backup_ptr, nbytes)
                sync()
                t0 = time.perf_counter()
                handle.reset_phi_batch_device(phi_ptr, None, mode=mode, band=band, bin_scale=bs,
                                              query_block=qb, stream=stream)
                sync()
                ts.append(time.perf_counter() - t0)
            ms = float(np.median(ts[1:]) * 1e3)
            table.append((bs, qb, ms))
            if log:
                log(f"autotune: bin_scale {bs} query_block {qb}: {ms:.3f} ms")
            if best is None or ms < best[2]:
                best = (bs, qb, ms)
    res = (best[0], best[1], table)
    if signature is not None:
        _cache[key] = res
    return res
