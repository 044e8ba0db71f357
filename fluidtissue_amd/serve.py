"""Served reinit: continuous batching of reinit jobs over resident meshes.

Concurrent simulations submit reinit jobs (a level set on a grid of some
resolution).  Meshes stay resident in HBM, keyed by grid dimensions, and
are evicted least-recently-used only to stay within a byte budget.  Each
worker lane owns a handle and a HIP stream; whenever a lane is free it admits
the queued jobs that fit (by count, nodes and memory budget) as one batch and
reinitialises all of them in one pass (ft_reinit_batch) over their resident
meshes, read in place.  Jobs that arrive while batches are on the GPU wait
for the next free lane, so two lanes keep two batches in flight: one lane's
transfers run on the copy engines while the other lane's kernels run.  The
lanes take turns at the kernels (exclusive_kernels): two batches' kernels
side by side only compete for the compute units and their caches (measured
on the 2000-job stream: 1067 against 1032 Mnodes/s).

A job's phi is a numpy array or a torch tensor.  Host phi is copied straight
into the lane's device buffer at the job's offset (asynchronously when it is
in pinned memory) and the result comes back into pinned host memory; phi
already on the device is copied device to device and the result stays there.
With inplace=True the result is written back into the job's own phi (as
LevelSet::resetPhi_ updates phi_ in place, src/fem/LevelSet.cpp:448).

A batch gives every job exactly the result the job would get alone (the
batched kernels keep each field's data apart); tests check it bit for bit.
"""
import collections
import contextlib
import ctypes as C
import threading
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .reinit import Reinit

# Device bytes per node of one reinit beyond the resident mesh: phi, phi_ifn,
# bitmaps, interface and index workspaces (upper estimate).
WORK_BYTES_PER_NODE = 96


def _create_kuhn(device, dims):
    """Resident Kuhn mesh on the device: (mesh id, bytes)."""
    mid = C.c_int32()
    check(lib.ft_mesh_create_kuhn(device, dims[0], dims[1], dims[2], 0.0, 0, C.byref(mid)))
    b = C.c_int64()
    check(lib.ft_mesh_info(mid.value, None, None, None, C.byref(b)))
    return mid.value, b.value


def _release(mid):
    check(lib.ft_mesh_release(mid))


class _Entry:
    """A resident mesh of the cache: id (None while it is being created),
    bytes, pins; `ready` is set once the creator has made it (or failed,
    then `error` holds why)."""
    __slots__ = ("id", "bytes", "pins", "ready", "error")

    def __init__(self, mid, nbytes, pins):
        self.id, self.bytes, self.pins = mid, nbytes, pins
        self.ready = threading.Event()
        self.error = None


class MeshCache:
    """Resident Kuhn meshes keyed by dims, within a byte budget (LRU).

    create(device, dims) -> (id, bytes) and release(id) default to the
    library's resident meshes.

    Admission reserves: reserve(dims, extra) pins the job's mesh and, when it
    is not resident, books its estimated bytes at once (evicting unpinned
    meshes least-recently-used), all under one lock, so that two lanes can
    never both be admitted against the same free bytes.  The mesh itself is
    made afterwards, outside the lock, by the lane that booked it
    (materialize); a lane that pins a mesh another lane is still making waits
    for it."""

    def __init__(self, device=0, budget_bytes=200e9, create=_create_kuhn, release=_release):
        self.device = device
        self.budget = float(budget_bytes)
        self.meshes = collections.OrderedDict()   # dims -> _Entry
        self.bytes = 0                            # booked: made meshes + meshes being made
        self.lock = threading.Lock()
        self._create = create
        self._release = release

    @staticmethod
    def estimate_bytes(dims):
        nx, ny, nz = dims
        n = nx * ny * nz
        t = 6 * (nx - 1) * (ny - 1) * (nz - 1)
        return 12 * n + 40 * t + 8 * 7 * n + 24 * n

    def _evict_for(self, need, keep=None):
        """Releases unpinned made meshes, least recently used first, until
        `need` more bytes fit; True when they do."""
        for key in list(self.meshes):
            if self.bytes + need <= self.budget:
                return True
            e = self.meshes[key]
            if key != keep and e.pins == 0 and e.id is not None:
                self._release(e.id)
                self.bytes -= e.bytes
                del self.meshes[key]
        return self.bytes + need <= self.budget

    def can_admit(self, dims, extra):
        """Would reserve(dims, extra) succeed now?  (No side effects.)"""
        with self.lock:
            e = self.meshes.get(tuple(dims))
            need = extra + (0 if e is not None else self.estimate_bytes(dims))
            free = sum(x.bytes for k, x in self.meshes.items()
                       if x.pins == 0 and x.id is not None and k != tuple(dims))
            return self.bytes - free + need <= self.budget

    def reserve(self, dims, extra=0):
        """Pins the mesh of `dims` and books `extra` bytes of work beside all
        resident meshes: returns (entry, must_create) or None when it does
        not fit (nothing changes then)."""
        key = tuple(dims)
        with self.lock:
            e = self.meshes.get(key)
            if e is not None:
                if not self._evict_for(extra, keep=key):
                    return None
                self.meshes.move_to_end(key)
                e.pins += 1
                return e, False
            est = self.estimate_bytes(key)
            if not self._evict_for(est + extra):
                return None
            e = _Entry(None, est, 1)
            self.meshes[key] = e
            self.bytes += est
            return e, True

    def materialize(self, dims, entry, must_create):
        """Makes a booked mesh (must_create) or waits until the lane that
        booked it has; returns its id."""
        if must_create:
            try:
                mid, b = self._create(self.device, tuple(dims))
            except Exception as err:
                with self.lock:
                    entry.error = err
                    self.bytes -= entry.bytes
                    if self.meshes.get(tuple(dims)) is entry:
                        del self.meshes[tuple(dims)]
                entry.ready.set()
                raise
            with self.lock:
                self.bytes += b - entry.bytes
                entry.id, entry.bytes = mid, b
            entry.ready.set()
        else:
            entry.ready.wait()
            if entry.error is not None:
                raise entry.error
        return entry.id

    def acquire(self, dims):
        """reserve + materialize for one mesh; FtError(FT_ERR_BUDGET) when it
        does not fit."""
        r = self.reserve(dims, 0)
        if r is None:
            raise _lib.FtError(_lib.FT_ERR_BUDGET, f"mesh {tuple(dims)} does not fit the budget")
        return self.materialize(dims, *r)

    def release(self, dims):
        with self.lock:
            e = self.meshes.get(tuple(dims))
            if e is not None:
                e.pins -= 1

    def unpin(self, entry):
        """Drops one pin of a reserved entry (a failed creation may already
        have removed it, and a new entry may hold its dims since)."""
        with self.lock:
            entry.pins -= 1

    def close(self):
        with self.lock:
            for e in self.meshes.values():
                if e.id is not None:
                    self._release(e.id)
            self.meshes.clear()
            self.bytes = 0


class Job:
    __slots__ = ("dims", "phi", "t_arrival", "t_start", "t_done", "out", "ifn", "counts",
                 "error", "done", "n", "on_device", "inplace", "numpy_in")

    def __init__(self, dims, phi, t_arrival=None, inplace=False):
        import torch
        self.dims = tuple(int(d) for d in dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.inplace = bool(inplace)
        self.numpy_in = not isinstance(phi, torch.Tensor)
        if self.inplace and self.numpy_in and not (isinstance(phi, np.ndarray) and phi.dtype == np.float64
                                                   and phi.flags.c_contiguous and phi.flags.writeable):
            raise ValueError("inplace needs a writeable contiguous float64 array")
        if isinstance(phi, torch.Tensor):
            if phi.dtype != torch.float64 or not phi.is_contiguous():
                raise ValueError("phi must be a contiguous float64 tensor")
            self.phi = phi
        else:
            self.phi = torch.from_numpy(np.ascontiguousarray(phi, np.float64))
        self.on_device = self.phi.is_cuda
        if tuple(self.phi.shape) != (self.n,):
            raise ValueError("phi does not match dims")
        self.t_arrival = time.perf_counter() if t_arrival is None else t_arrival
        self.t_start = self.t_done = None
        self.out = self.ifn = self.counts = self.error = None
        self.done = threading.Event()

    @property
    def latency(self):
        return self.t_done - self.t_arrival

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError("reinit job not done")
        if self.error:
            raise self.error
        return self.out


def admit(queue, max_fields, max_nodes, inflight_bytes, reserve):
    """Takes from the front of `queue` the jobs that fit one batch: FIFO with no
    overtaking, at most max_fields jobs and max_nodes nodes (a single larger
    job is still admitted alone), and only while reserve(dims, bytes) can pin
    the job's mesh with the work of the in-flight and admitted jobs beside
    the resident meshes.  Returns (jobs, reservations, work bytes of the
    admitted jobs); the caller holds the reservations until the batch ends."""
    jobs, res, nodes, extra = [], [], 0, 0
    while queue and len(jobs) < max_fields:
        j = queue[0]
        work = WORK_BYTES_PER_NODE * j.n
        if jobs and nodes + j.n > max_nodes:
            break
        r = reserve(j.dims, inflight_bytes + extra + work)
        if r is None:
            break
        queue.popleft()
        jobs.append(j)
        res.append(r)
        nodes += j.n
        extra += work
    return jobs, res, extra


class Server:
    """Continuous-batching reinit server (see module docstring)."""

    def __init__(self, device=0, lanes=2, budget_bytes=200e9, max_fields=64, max_nodes=64_000_000,
                 mode="triangles", blend="init", band=0.0, bin_scale=0.0, query_block=0,
                 want_ifn=True, exclusive_kernels=True):
        self.device = device
        self.want_ifn = bool(want_ifn)
        self.cache = MeshCache(device, budget_bytes)
        self.budget = float(budget_bytes)
        self.max_fields = int(max_fields)
        self.max_nodes = int(max_nodes)
        self.kw = dict(mode=mode, blend=blend, band=band, bin_scale=bin_scale,
                       query_block=query_block)
        self.queue = collections.deque()
        self.cv = threading.Condition(threading.RLock())
        self.gpu_turn = threading.Lock() if exclusive_kernels else contextlib.nullcontext()
        self.stop = False
        self.inflight_bytes = 0
        self.batches = []              # (n_jobs, n_nodes, seconds) per batch
        self.handles = [Reinit(device) for _ in range(lanes)]
        self.threads = [threading.Thread(target=self._lane, args=(h,), daemon=True)
                        for h in self.handles]
        for t in self.threads:
            t.start()

    def submit(self, dims, phi, t_arrival=None, inplace=False):
        job = Job(dims, phi, t_arrival, inplace)
        with self.cv:                  # re-entrant: a caller may hold cv to submit a burst
            self.queue.append(job)
            self.cv.notify()
        return job

    def _admit(self):
        """Under cv: the next batch with its meshes pinned and booked.  A job
        that cannot fit even with nothing in flight (every mesh evictable)
        never will: it fails with FT_ERR_BUDGET instead of waiting."""
        jobs, res, extra = admit(self.queue, self.max_fields, self.max_nodes, self.inflight_bytes,
                                 self.cache.reserve)
        if not jobs and self.queue and self.inflight_bytes == 0:
            j = self.queue.popleft()
            j.error = _lib.FtError(_lib.FT_ERR_BUDGET,
                                   f"job of {j.dims} does not fit the {self.budget:.3g}-byte budget")
            j.t_done = time.perf_counter()
            j.done.set()
        self.inflight_bytes += extra
        return jobs, res, extra

    def _lane(self, h):
        import torch
        dev = torch.device("cuda", self.device)
        stream = torch.cuda.Stream(device=dev)      # default priority
        d_phi = d_ifn = None
        while True:
            with self.cv:
                while not self.queue and not self.stop:
                    self.cv.wait()
                if self.stop and not self.queue:
                    return
                jobs, res, extra = self._admit()
                if not jobs:              # memory is held by the other lane
                    self.cv.wait(0.001)
                    continue
            t0 = time.perf_counter()
            ids = []
            try:
                for j, (entry, make) in zip(jobs, res):
                    j.t_start = t0
                    ids.append(self.cache.materialize(j.dims, entry, make))
                h.set_batch_meshes(ids)
                total = sum(j.n for j in jobs)
                with torch.cuda.stream(stream):
                    if d_phi is None or d_phi.numel() < total:
                        d_phi = d_ifn = None
                        d_phi = torch.empty(total, dtype=torch.float64, device=dev)
                        if self.want_ifn:
                            d_ifn = torch.empty_like(d_phi)
                    off = 0
                    for j in jobs:
                        d_phi[off:off + j.n].copy_(j.phi, non_blocking=True)
                        off += j.n
                # one lane's kernels at a time (two batches' kernels side by
                # side only compete for the CUs and their caches); the other
                # lanes' copies overlap them
                with self.gpu_turn:
                    counts = h.reset_phi_batch_device(
                        d_phi.data_ptr(), d_ifn.data_ptr() if d_ifn is not None else None,
                        stream=stream.cuda_stream, **self.kw)
                with torch.cuda.stream(stream):
                    off = 0
                    for j in jobs:
                        for name, src in (("out", d_phi), ("ifn", d_ifn)):
                            if src is None:
                                continue
                            part = src[off:off + j.n]
                            if name == "out" and j.inplace:
                                dst = j.phi
                                dst.copy_(part, non_blocking=True)
                            elif j.on_device:
                                dst = part.clone()
                            else:
                                dst = torch.empty(j.n, dtype=torch.float64, pin_memory=True)
                                dst.copy_(part, non_blocking=True)
                            setattr(j, name, dst)
                        off += j.n
                stream.synchronize()
                t1 = time.perf_counter()
                for j, c in zip(jobs, counts):
                    if j.numpy_in:
                        j.out = j.out.numpy()
                        if j.ifn is not None:
                            j.ifn = j.ifn.numpy()
                    j.counts = c
                    j.t_done = t1
                self.batches.append((len(jobs), total, t1 - t0))
            except Exception as e:        # reported on every job of the batch
                for j in jobs:
                    j.error = e
                    j.t_done = time.perf_counter()
            finally:
                for entry, _ in res:      # every admitted job holds a pin
                    self.cache.unpin(entry)
                with self.cv:
                    self.inflight_bytes -= extra
                    self.cv.notify_all()
                for j in jobs:
                    j.done.set()

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        for h in self.handles:
            h.close()
        self.cache.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def latency_stats(jobs):
    lat = np.array([j.latency for j in jobs])
    return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))
