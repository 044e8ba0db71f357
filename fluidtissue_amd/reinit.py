"""Host-side mirror of fluidtissue's reinit interface over libftreinit.so.

  EuclideanDistance   src/fem/EuclideanDistance.hpp:13-52 (init / resetPhi)
  Reinit              LevelSet::resetPhi_ and the blends of LevelSet::init /
                      LevelSet::update (src/fem/LevelSet.cpp:73-74, 147-148,
                      429-455) on a mesh resident in HBM, with the interface
                      members of LevelSet.hpp:133-139 as outputs
  kuhn_mesh, phi_field
                      synthetic inputs (x-major nodes as TetraMesh.cpp:199-211)
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib

MODES = {"points": _lib.FT_DIST_POINTS, "triangles": _lib.FT_DIST_TRIANGLES}
BLENDS = {"init": _lib.FT_BLEND_INIT, "update": _lib.FT_BLEND_UPDATE}
PHI_KINDS = {
    "sphere_wave": _lib.FT_PHI_SPHERE_WAVE,
    "star": _lib.FT_PHI_STAR,
    "spheres": _lib.FT_PHI_SPHERES,
    "ellipsoids": _lib.FT_PHI_ELLIPSOIDS,
}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


@dataclass
class Mesh:
    """Tet mesh in the reference's layout (TetraMesh.hpp:250-256)."""
    xyz: np.ndarray     # (N, 3) float32
    tets: np.ndarray    # (T, 4) int32
    edges: np.ndarray   # (E, 2) int32
    t2e: np.ndarray     # (T, 6) int32
    shape: tuple = ()
    spacing: float = 0.0

    @property
    def n_nodes(self):
        return self.xyz.shape[0]


@dataclass
class Interface:
    """LevelSet.hpp:133-139: ifp_, ifp2e_, ifMesh_ (t, p2t), ifElements_, ifpArea_."""
    ifp: np.ndarray
    ifp2e: np.ndarray
    tris: np.ndarray
    if_elems: np.ndarray
    ifp_area: np.ndarray
    p2t_offsets: np.ndarray
    p2t: np.ndarray


def device_count():
    n = C.c_int32(0)
    check(lib.ft_device_count(C.byref(n)))
    return n.value


def kuhn_mesh(n, origin=(0.0, 0.0, 0.0), spacing=None, jitter=0.0, seed=0):
    """n or (nx, ny, nz) nodes; default spacing puts the grid on the unit cube."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
    if spacing is None:
        spacing = 1.0 / (max(nx, ny, nz) - 1)
    sp = (spacing,) * 3 if np.isscalar(spacing) else tuple(spacing)
    N, T, E = C.c_int64(), C.c_int64(), C.c_int64()
    check(lib.ft_kuhn_counts(nx, ny, nz, C.byref(N), C.byref(T), C.byref(E)))
    xyz = np.empty((N.value, 3), np.float32)
    tets = np.empty((T.value, 4), np.int32)
    edges = np.empty((E.value, 2), np.int32)
    t2e = np.empty((T.value, 6), np.int32)
    o = np.array(origin, np.float64)
    s = np.array(sp, np.float64)
    check(lib.ft_kuhn_mesh(nx, ny, nz, _ptr(o), _ptr(s), float(jitter), int(seed) & (2**64 - 1),
                           _ptr(xyz), _ptr(tets), _ptr(edges), _ptr(t2e)))
    return Mesh(xyz, tets, edges, t2e, (nx, ny, nz), float(sp[0]))


def phi_field(kind, xyz, seed=0):
    xyz = _c(xyz, np.float32)
    phi = np.empty(xyz.shape[0], np.float64)
    check(lib.ft_phi_field(PHI_KINDS[kind], xyz.shape[0], _ptr(xyz), int(seed) & (2**64 - 1),
                           _ptr(phi)))
    return phi


SCHEDULES = {"static": _lib.FT_SCHED_STATIC, "queue": _lib.FT_SCHED_QUEUE, "bins": _lib.FT_SCHED_BINS}


def make_opts(mode="points", blend="init", band=0.0, bin_scale=0.0, want_nearest=False,
              w_phi=0.75, w_ifn=0.25, query_block=0, want_normals=False, waves_per_simd=0,
              schedule="static"):
    o = _lib.FtOpts()
    lib.ft_opts_default(C.byref(o))
    o.mode = MODES[mode]
    o.blend = BLENDS[blend]
    o.band = float(band)
    o.bin_scale = float(bin_scale)
    o.want_nearest = 1 if want_nearest else 0
    o.w_phi = float(w_phi)
    o.w_ifn = float(w_ifn)
    o.query_block = int(query_block)
    o.want_normals = 1 if want_normals else 0
    o.waves_per_simd = int(waves_per_simd)
    o.schedule = SCHEDULES[schedule]
    return o


class _Handle:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib.ft_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib.ft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class EuclideanDistance(_Handle):
    """EuclideanDistance::init / resetPhi (EuclideanDistance.hpp:30-38)."""

    def init(self, meshp, phi0=None, n=None, n_threads=None):
        xyz = _c(meshp, np.float32).reshape(-1, 3)
        if n is not None and n != xyz.shape[0]:
            raise ValueError("n does not match meshp")
        self._xyz = xyz
        self.n = xyz.shape[0]
        check(lib.ft_set_nodes(self._h, _ptr(xyz), self.n))

    def reset_phi(self, ifp, n_ifp, phi_n):
        """In place on phi_n (float64, length n), as the reference's void resetPhi."""
        ifp = _c(ifp, np.float32).reshape(-1, 3)
        if n_ifp != ifp.shape[0]:
            raise ValueError("n_ifp does not match ifp")
        if phi_n.dtype != np.float64 or not phi_n.flags.c_contiguous or phi_n.shape[0] != self.n:
            raise ValueError("phi_n must be a contiguous float64 array of length n")
        check(lib.ft_reset_phi_points(self._h, _ptr(ifp) if n_ifp else None, n_ifp, _ptr(phi_n)))


class Reinit(_Handle):
    """LevelSet::resetPhi_ + blend on a resident mesh."""

    def set_mesh(self, mesh):
        self.mesh = mesh
        xyz = _c(mesh.xyz, np.float32)
        tets = _c(mesh.tets, np.int32)
        edges = _c(mesh.edges, np.int32)
        t2e = _c(mesh.t2e, np.int32)
        check(lib.ft_set_mesh(self._h, _ptr(xyz), xyz.shape[0], _ptr(tets), tets.shape[0],
                              _ptr(edges), edges.shape[0], _ptr(t2e)))
        self.n = xyz.shape[0]

    def reset_phi(self, phi, mode="points", blend="init", band=0.0, bin_scale=0.0,
                  want_ifn=True, want_nearest=False, query_block=0, want_normals=False,
                  waves_per_simd=0, schedule="static"):
        """Returns (phi_out, phi_ifn, counts); phi is not modified."""
        out = np.array(phi, dtype=np.float64, copy=True)
        ifn = np.empty_like(out) if want_ifn else None
        o = make_opts(mode, blend, band, bin_scale, want_nearest, query_block=query_block,
                      want_normals=want_normals, waves_per_simd=waves_per_simd, schedule=schedule)
        cnt = _lib.FtCounts()
        self.counts = cnt
        check(lib.ft_reinit(self._h, _ptr(out), _ptr(ifn), C.byref(o), C.byref(cnt)))
        return out, ifn, cnt

    def reset_phi_device(self, phi_ptr, ifn_ptr=None, mode="points", blend="init", band=0.0,
                         bin_scale=0.0, want_nearest=False, stream=None, query_block=0,
                         want_normals=False, waves_per_simd=0, schedule="static"):
        """Device-resident phi (raw device addresses, e.g. tensor.data_ptr())."""
        o = make_opts(mode, blend, band, bin_scale, want_nearest, query_block=query_block,
                      want_normals=want_normals, waves_per_simd=waves_per_simd, schedule=schedule)
        cnt = _lib.FtCounts()
        self.counts = cnt
        check(lib.ft_reinit_device(self._h, C.c_void_p(phi_ptr),
                                   C.c_void_p(ifn_ptr) if ifn_ptr else None, C.byref(o),
                                   C.byref(cnt), C.c_void_p(stream) if stream else None))
        return cnt

    def reset_phi_device_noalloc(self, phi_ptr, ifn_ptr=None, mode="points", blend="init", band=0.0,
                                 bin_scale=0.0, stream=None, query_block=0, waves_per_simd=0,
                                 schedule="static"):
        """reset_phi_device with no device allocation: every workspace comes
        from the library's fixed scratch (FT_NOALLOC_SCRATCH_BYTES per GPU).
        No nearest indices, normals or interface are kept."""
        o = make_opts(mode, blend, band, bin_scale, False, query_block=query_block,
                      waves_per_simd=waves_per_simd, schedule=schedule)
        cnt = _lib.FtCounts()
        self.counts = cnt
        check(lib.ft_reinit_device_noalloc(self._h, C.c_void_p(phi_ptr),
                                           C.c_void_p(ifn_ptr) if ifn_ptr else None, C.byref(o),
                                           C.byref(cnt), C.c_void_p(stream) if stream else None))
        return cnt

    # ---- batches of independent fields ------------------------------------
    def set_batch(self, meshes):
        """Resident batch of independent fields from host meshes."""
        nf = len(meshes)
        no = np.zeros(nf + 1, np.int64)
        to = np.zeros(nf + 1, np.int64)
        eo = np.zeros(nf + 1, np.int64)
        for k, m in enumerate(meshes):
            no[k + 1] = no[k] + m.xyz.shape[0]
            to[k + 1] = to[k] + m.tets.shape[0]
            eo[k + 1] = eo[k] + m.edges.shape[0]
        xyz = np.ascontiguousarray(np.concatenate([m.xyz for m in meshes]), np.float32)
        tets = np.ascontiguousarray(np.concatenate([m.tets for m in meshes]), np.int32)
        edges = np.ascontiguousarray(np.concatenate([m.edges for m in meshes]), np.int32)
        t2e = np.ascontiguousarray(np.concatenate([m.t2e for m in meshes]), np.int32)
        check(lib.ft_set_mesh_batch(self._h, nf, _ptr(no), _ptr(xyz), _ptr(to), _ptr(tets),
                                    _ptr(eo), _ptr(edges), _ptr(t2e)))
        self.n = int(no[-1])

    def set_kuhn_batch(self, dims, jitter=0.0, seed=0, origins=None, spacings=None):
        """Resident batch of Kuhn grids generated on the device; dims: list of
        n or (nx, ny, nz)."""
        d = np.array([(x, x, x) if np.isscalar(x) else tuple(x) for x in dims], np.int32)
        o = None if origins is None else np.ascontiguousarray(origins, np.float64)
        sp = None if spacings is None else np.ascontiguousarray(spacings, np.float64)
        check(lib.ft_set_kuhn_batch(self._h, d.shape[0], _ptr(d), _ptr(o), _ptr(sp), float(jitter),
                                    int(seed) & (2**64 - 1)))
        self.n = int(self.batch_info()[0][-1])

    def batch_info(self):
        """(node_off, tet_off, edge_off) of the resident fields."""
        nf = C.c_int32()
        check(lib.ft_batch_info(self._h, C.byref(nf), None, None, None))
        no = np.empty(nf.value + 1, np.int64)
        to = np.empty(nf.value + 1, np.int64)
        eo = np.empty(nf.value + 1, np.int64)
        check(lib.ft_batch_info(self._h, C.byref(nf), _ptr(no), _ptr(to), _ptr(eo)))
        return no, to, eo

    def mesh(self):
        """The resident mesh (global indices)."""
        no, to, eo = self.batch_info()
        xyz = np.empty((no[-1], 3), np.float32)
        tets = np.empty((to[-1], 4), np.int32)
        edges = np.empty((eo[-1], 2), np.int32)
        t2e = np.empty((to[-1], 6), np.int32)
        check(lib.ft_get_mesh(self._h, _ptr(xyz), _ptr(tets), _ptr(edges), _ptr(t2e)))
        return Mesh(xyz, tets, edges, t2e)

    def set_batch_meshes(self, mesh_ids):
        ids = np.ascontiguousarray(mesh_ids, np.int32)
        check(lib.ft_set_batch_meshes(self._h, ids.shape[0], _ptr(ids)))
        self.n = int(self.batch_info()[0][-1])

    def field_counts(self):
        nf = len(self.batch_info()[0]) - 1
        arr = (_lib.FtCounts * nf)()
        check(lib.ft_get_field_counts(self._h, arr))
        return list(arr)

    def reset_phi_batch_device(self, phi_ptr, ifn_ptr=None, mode="points", blend="init", band=0.0,
                               bin_scale=0.0, want_nearest=False, stream=None, query_block=0,
                               want_normals=False, waves_per_simd=0, schedule="static"):
        o = make_opts(mode, blend, band, bin_scale, want_nearest, query_block=query_block,
                      want_normals=want_normals, waves_per_simd=waves_per_simd, schedule=schedule)
        rc = lib.ft_reinit_batch(self._h, C.c_void_p(phi_ptr),
                                 C.c_void_p(ifn_ptr) if ifn_ptr else None, C.byref(o), None,
                                 C.c_void_p(stream) if stream else None)
        return self._batch_done(rc)

    def _batch_done(self, rc):
        """Per-field counts of a batch call; the totals become self.counts."""
        check(rc)
        per = self.field_counts()
        tot = _lib.FtCounts()
        for c in per:
            tot.n_ifp += c.n_ifp
            tot.n_tri += c.n_tri
            tot.n_if_elems += c.n_if_elems
            tot.n_clamped += c.n_clamped
        tot.bad_tet = -1
        self.counts = tot
        return per

    def reset_phi_graph_device(self, phi_ptr, ifn_ptr=None, mode="points", blend="init", band=0.0,
                               bin_scale=0.0, want_nearest=False, stream=None, query_block=0,
                               want_normals=False, waves_per_simd=0, schedule="static"):
        """reset_phi_batch_device through one captured HIP graph per call
        (captured on the first call for these buffers and options)."""
        o = make_opts(mode, blend, band, bin_scale, want_nearest, query_block=query_block,
                      want_normals=want_normals, waves_per_simd=waves_per_simd, schedule=schedule)
        rc = lib.ft_reinit_graph(self._h, C.c_void_p(phi_ptr),
                                 C.c_void_p(ifn_ptr) if ifn_ptr else None, C.byref(o), None,
                                 C.c_void_p(stream) if stream else None)
        return self._batch_done(rc)

    def graph_stats(self):
        """(captures, launches) of the graph form on this handle."""
        a, b = C.c_int64(), C.c_int64()
        check(lib.ft_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_timing(self, on=True):
        check(lib.ft_set_timing(self._h, 1 if on else 0))

    def timing(self):
        """(marching tets, index build, query, normals) cumulative ms, number of calls."""
        ms = (C.c_double * 4)()
        n = C.c_int64()
        check(lib.ft_get_timing(self._h, ms, C.byref(n)))
        return list(ms), n.value

    def interface(self):
        c = self.counts
        P, F, X = c.n_ifp, c.n_tri, c.n_if_elems
        ifp = np.empty((P, 3), np.float32)
        ifp2e = np.empty(P, np.int32)
        tris = np.empty((F, 4), np.int32)
        elems = np.empty(X, np.int32)
        area = np.empty(P, np.float32)
        off = np.empty(P + 1, np.int32)
        p2t = np.empty(3 * F, np.int32)
        check(lib.ft_get_interface(self._h, _ptr(ifp), _ptr(ifp2e), _ptr(tris), _ptr(elems),
                                   _ptr(area), _ptr(off), _ptr(p2t)))
        return Interface(ifp, ifp2e, tris, elems, area, off, p2t)

    def normals(self):
        """Level-set normals of the last reinit with want_normals (n x 3 float)."""
        out = np.empty((self.n, 3), np.float32)
        check(lib.ft_get_normals(self._h, _ptr(out)))
        return out

    def curvature(self, stream=None):
        """LevelSet::getInterfaceCurvature after a reinit with want_normals:
        one float per crossing point, in crossing-edge order."""
        out = np.empty(self.counts.n_ifp, np.float32)
        check(lib.ft_interface_curvature(self._h, _ptr(out), C.c_void_p(stream) if stream else None))
        return out

    def fem_matrix(self, form, vel=None):
        """Assembled mass (form 0) or advection (form 1, velocity vel: 3n
        floats, x then y then z) matrix in CSR: (row_off, cols, vals)."""
        row_off = np.empty(self.n + 1, np.int64)
        check(lib.ft_fem_assemble(self._h, int(form), None, _ptr(row_off), None, None))
        nnz = int(row_off[-1])
        cols = np.empty(max(nnz, 1), np.int32)
        vals = np.empty(max(nnz, 1), np.float64)
        v = None if vel is None else _c(vel, np.float64)
        if v is not None and v.shape[0] < 3 * self.n:
            raise ValueError("velocity: 3 values per node")
        check(lib.ft_fem_assemble(self._h, int(form), _ptr(v), _ptr(row_off), _ptr(cols), _ptr(vals)))
        return row_off, cols[:nnz], vals[:nnz]

    def advect(self, phi, flow, dt, tol=1e-10, max_iters=1000):
        """The advection step of LevelSet::update: Crank-Nicolson with
        BiCGSTAB.  flow: 4n floats (velocity x, y, z, then pressure).
        Returns (phi_n, info dict); FtError(FT_ERR_SOLVER) if the solve stops
        above tol."""
        phi_n = np.array(phi, np.float64, copy=True, order="C")
        flow = _c(flow, np.float64)
        if phi_n.shape != (self.n,) or flow.shape != (4 * self.n,):
            raise ValueError("phi: one value per node; flow: four per node")
        info = _lib.FtAdvectInfo()
        check(lib.ft_advect(self._h, _ptr(phi_n), _ptr(flow), float(dt), float(tol), int(max_iters),
                            C.byref(info)))
        return phi_n, dict(iterations=info.iterations, restarts=info.restarts, error=info.error,
                           converged=bool(info.converged), skipped=bool(info.skipped))

    def update(self, phi, flow, dt, tol=1e-10, max_iters=1000, mode="points", **kw):
        """LevelSet::update (LevelSet.cpp:89-152): the advection, then the
        reinit with the update blend.  A skipped advection (dt or |flow| below
        FLT_MIN) does nothing, as the reference returns before the reinit.
        Returns (phi_out, phi_ifn, counts, advection info)."""
        phi_a, info = self.advect(phi, flow, dt, tol, max_iters)
        if info["skipped"]:
            return phi_a, None, None, info
        out, ifn, cnt = self.reset_phi(phi_a, mode=mode, blend="update", **kw)
        return out, ifn, cnt, info

    def surface_tension(self, st, rho, fx=None, fy=None, fz=None):
        """Flow::surfaceTension_ after a reinit with want_normals: adds the
        surface-tension force into fx, fy, fz (float64, one per node; zeros
        when not given) and returns them."""
        f = []
        for a in (fx, fy, fz):
            if a is None:
                a = np.zeros(self.n, np.float64)
            elif a.dtype != np.float64 or a.shape != (self.n,) or not a.flags.c_contiguous:
                raise ValueError("force arrays: contiguous float64 of one entry per node")
            f.append(a)
        check(lib.ft_surface_tension(self._h, C.c_float(st), C.c_int32(rho), _ptr(f[0]), _ptr(f[1]),
                                     _ptr(f[2])))
        return tuple(f)

    def write_obj(self, fname, unit_normals=True, boundary=None, interface=True):
        """fname.obj / fname.mtl (cfish::Write_Wavefront_obj): the last
        reinit's interface when `interface`, then, when `boundary` is given
        (n x 3 node ids: the mesh's exterior boundary triangles), the domain
        boundary section."""
        if boundary is None:
            if not interface:
                raise ValueError("write_obj: nothing to write")
            check(lib.ft_write_interface_obj(self._h, str(fname).encode(), 1 if unit_normals else 0))
            return
        b = np.ascontiguousarray(boundary, np.int32)
        if b.ndim != 2 or b.shape[1] != 3:
            raise ValueError("boundary: an n x 3 array of node ids")
        parts = _lib.FT_OBJ_BOUNDARY | (_lib.FT_OBJ_INTERFACE if interface else 0)
        check(lib.ft_write_obj(self._h, str(fname).encode(), parts, 1 if unit_normals else 0, _ptr(b),
                               len(b)))

    def nearest(self):
        out = np.empty(self.n, np.int32)
        check(lib.ft_get_nearest(self._h, _ptr(out)))
        return out
