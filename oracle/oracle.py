"""ctypes binding of oracle/libftoracle.so -- TEST INFRASTRUCTURE ONLY.

The C functions restate the reference's CPU reinit (see ftoracle.h for the
file:line of each); this module only marshals numpy arrays.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftoracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
_o = C.CDLL(LIB_PATH)

_P = C.c_void_p
_i64 = C.c_int64
_o.ora_find_interface_nodes.restype = _i64
_o.ora_find_interface_nodes.argtypes = [_i64, _P, _P, _P, _P, _P]
_o.ora_reconstruct_interface.restype = C.c_int
_o.ora_reconstruct_interface.argtypes = [_i64, _P, _P, _i64, _P, _P, _P, _i64, _P, _P, _P, _P, _P,
                                         C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]
_o.ora_reset_phi_points.restype = None
_o.ora_reset_phi_points.argtypes = [_i64, _P, _P, _i64, _P, _P, C.c_int]
_o.ora_point_distance.restype = None
_o.ora_point_distance.argtypes = [_i64, _P, _P, _P, _i64, _P, _P, C.c_int]
_o.ora_point_triangle_dist2.restype = C.c_double
_o.ora_point_triangle_dist2.argtypes = [_P, _P, _P, _P]
_o.ora_triangle_distance.restype = None
_o.ora_triangle_distance.argtypes = [_i64, _P, _P, _P, _P, _i64, _P, _P, C.c_int]
_o.ora_node_triangle_distance.restype = C.c_double
_o.ora_node_triangle_distance.argtypes = [_i64, _P, _P, _P, _i64]
_o.ora_finish_reset.restype = None
_o.ora_finish_reset.argtypes = [_i64, _P, _P, _i64, _P, _P, C.c_int, C.c_double, C.c_double, _P, _P]

_o.ora_node_tets.restype = None
_o.ora_node_tets.argtypes = [_i64, _i64, _P, _P, _P]
_o.ora_tet_grad.restype = _i64
_o.ora_tet_grad.argtypes = [_i64, _P, _P, _P, _P]
_o.ora_grad_nodes.restype = None
_o.ora_grad_nodes.argtypes = [_i64, _P, _P, _P, _P]

_o.ora_interface_curvature.restype = None
_o.ora_interface_curvature.argtypes = [_i64, _P, _i64, _P, _P, _P, _P, _P, _i64, _P, _P, _P]
_o.ora_fem_quad.restype = None
_o.ora_fem_quad.argtypes = [_P, _P]
_o.ora_fem_assemble.restype = _i64
_o.ora_fem_assemble.argtypes = [C.c_int, _i64, _P, _P, _P, _P, _P, _P, _P, _P]
_o.ora_spmv.restype = None
_o.ora_spmv.argtypes = [_i64, _P, _P, _P, _P, _P]
_o.ora_bicgstab.restype = C.c_int
_o.ora_bicgstab.argtypes = [_i64, _P, _P, _P, _P, _P, C.c_double, _i64, _P, _P, _P]
_o.ora_surface_tension.restype = None
_o.ora_surface_tension.argtypes = [_P, _P, _i64, _P, _P, _P, _P, C.c_float, C.c_int32, _P, _P, _P]

NTHREADS = max(1, min(16, os.cpu_count() or 1))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def find_interface_nodes(edges, xyz, phi):
    """LevelSet::findInterfaceNodes_ -> (ifp (P,3) f32, ifp2e (P,) i32)."""
    edges, xyz, phi = _c(edges, np.int32), _c(xyz, np.float32), _c(phi, np.float64)
    E = edges.shape[0]
    ifp = np.empty((max(E, 1), 3), np.float32)
    ifp2e = np.empty(max(E, 1), np.int32)
    P = _o.ora_find_interface_nodes(E, _p(edges), _p(xyz), _p(phi), _p(ifp), _p(ifp2e))
    return ifp[:P].copy(), ifp2e[:P].copy()


class ReconstructError(RuntimeError):
    def __init__(self, code, tet):
        super().__init__(f"reconstructInterface_ failed with {code} at tet {tet}")
        self.code = code
        self.tet = tet


def reconstruct_interface(tets, t2e, edges, xyz, phi, ifp, ifp2e):
    """LevelSet::reconstructInterface_ -> (tris (F,4), if_elems, ifp_area)."""
    tets, t2e, edges = _c(tets, np.int32), _c(t2e, np.int32), _c(edges, np.int32)
    xyz, phi = _c(xyz, np.float32), _c(phi, np.float64)
    ifp, ifp2e = _c(ifp, np.float32), _c(ifp2e, np.int32)
    T = tets.shape[0]
    tris = np.empty((max(2 * T, 1), 4), np.int32)
    elems = np.empty(max(T, 1), np.int32)
    area = np.empty(max(ifp.shape[0], 1), np.float32)
    nt, ne, bad = _i64(), _i64(), _i64()
    rc = _o.ora_reconstruct_interface(T, _p(tets), _p(t2e), edges.shape[0], _p(edges), _p(xyz),
                                      _p(phi), ifp.shape[0], _p(ifp), _p(ifp2e), _p(tris),
                                      _p(elems), _p(area), C.byref(nt), C.byref(ne), C.byref(bad))
    if rc != 0:
        raise ReconstructError(rc, bad.value)
    return tris[:nt.value].copy(), elems[:ne.value].copy(), area[:ifp.shape[0]].copy()


def p2t_lists(tris, n_ifp):
    """ifMesh_.p2t as CSR (ascending triangle ids per point)."""
    F = tris.shape[0]
    v = tris[:, :3].reshape(-1).astype(np.int64)
    f = np.repeat(np.arange(F, dtype=np.int64), 3)
    order = np.lexsort((f, v))
    counts = np.bincount(v, minlength=n_ifp)
    off = np.zeros(n_ifp + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    return off.astype(np.int32), f[order].astype(np.int32)


def reset_phi_points(xyz, ifp, phi_n, nthreads=NTHREADS, want_nearest=False):
    """EuclideanDistance::resetPhi CPU branch; returns (phi, nearest)."""
    xyz, ifp = _c(xyz, np.float32), _c(ifp, np.float32)
    out = np.array(phi_n, dtype=np.float64, copy=True)
    near = np.empty(xyz.shape[0], np.int32) if want_nearest else None
    _o.ora_reset_phi_points(xyz.shape[0], _p(xyz), _p(ifp), ifp.shape[0], _p(out), _p(near),
                            nthreads)
    return out, near


def point_distance(xyz, ifp, nodes=None, nthreads=NTHREADS):
    """Unsigned float distance to the nearest crossing point (and its index)."""
    xyz, ifp = _c(xyz, np.float32), _c(ifp, np.float32)
    nodes = None if nodes is None else _c(nodes, np.int64)
    n = xyz.shape[0] if nodes is None else nodes.shape[0]
    d = np.empty(n, np.float32)
    j = np.empty(n, np.int32)
    _o.ora_point_distance(n, _p(nodes), _p(xyz), _p(ifp), ifp.shape[0], _p(d), _p(j), nthreads)
    return d, j


def point_triangle_dist2(x, a, b, c):
    x, a, b, c = (_c(v, np.float64) for v in (x, a, b, c))
    return _o.ora_point_triangle_dist2(_p(x), _p(a), _p(b), _p(c))


def triangle_distance(xyz, ifp, tris, nodes=None, nthreads=NTHREADS):
    """FP64 nearest triangle by brute force: (distance, triangle id)."""
    xyz, ifp, tris = _c(xyz, np.float32), _c(ifp, np.float32), _c(tris, np.int32)
    nodes = None if nodes is None else _c(nodes, np.int64)
    n = xyz.shape[0] if nodes is None else nodes.shape[0]
    d = np.empty(n, np.float64)
    f = np.empty(n, np.int32)
    _o.ora_triangle_distance(n, _p(nodes), _p(xyz), _p(ifp), _p(tris), tris.shape[0], _p(d), _p(f),
                             nthreads)
    return d, f


def node_triangle_distance(node, xyz, ifp, tris, f):
    xyz, ifp, tris = _c(xyz, np.float32), _c(ifp, np.float32), _c(tris, np.int32)
    return _o.ora_node_triangle_distance(int(node), _p(xyz), _p(ifp), _p(tris), int(f))


def finish_reset(phi_old, phi_ifn, ifp2e, edges, blend, w_phi=0.75, w_ifn=0.25):
    """LevelSet::resetPhi_ epilogue + blend -> (retained, phi_out)."""
    phi_old, phi_ifn = _c(phi_old, np.float64), _c(phi_ifn, np.float64)
    ifp2e, edges = _c(ifp2e, np.int32), _c(edges, np.int32)
    n = phi_old.shape[0]
    ret = np.empty(n, np.float64)
    out = np.empty(n, np.float64)
    _o.ora_finish_reset(n, _p(phi_old), _p(phi_ifn), ifp2e.shape[0], _p(ifp2e), _p(edges),
                        int(blend), w_phi, w_ifn, _p(ret), _p(out))
    return ret, out


def reinit(mesh_xyz, tets, edges, t2e, phi, blend=0, nthreads=NTHREADS):
    """Whole reference reinit (points mode): LevelSet::resetPhi_ + blend."""
    ifp, ifp2e = find_interface_nodes(edges, mesh_xyz, phi)
    tris, elems, area = reconstruct_interface(tets, t2e, edges, mesh_xyz, phi, ifp, ifp2e)
    ifn, _ = reset_phi_points(mesh_xyz, ifp, phi, nthreads)
    ret, out = finish_reset(phi, ifn, ifp2e, edges, blend)
    return dict(ifp=ifp, ifp2e=ifp2e, tris=tris, if_elems=elems, ifp_area=area, phi_ifn=ifn,
                retained=ret, phi=out)


def node_tets(n_nodes, tets):
    """Mesh3D::p2t as CSR: (offsets n_nodes+1, tet ids in increasing order)."""
    tets = _c(tets, np.int32)
    off = np.empty(n_nodes + 1, np.int64)
    ids = np.empty(4 * len(tets), np.int32)
    _o.ora_node_tets(n_nodes, len(tets), _p(tets), _p(off), _p(ids))
    return off, ids


def tet_grad(tets, xyz, u):
    """Assembler::grad: float gradient per tet; also the first singular tet or -1."""
    tets = _c(tets, np.int32)
    xyz = _c(xyz, np.float32)
    u = _c(u, np.float64)
    g = np.empty((len(tets), 3), np.float32)
    bad = _o.ora_tet_grad(len(tets), _p(tets), _p(xyz), _p(u), _p(g))
    return g, int(bad)


def grad_nodes(tets, xyz, u):
    """Assembler::grad_nodes: per-node float mean of the tet gradients."""
    n = len(xyz)
    off, ids = node_tets(n, tets)
    g, _ = tet_grad(tets, xyz, u)
    out = np.empty((n, 3), np.float32)
    _o.ora_grad_nodes(n, _p(off), _p(ids), _p(g), _p(out))
    return out


def interface_curvature(tets, xyz, normals, phi, ifp2e, edges):
    """LevelSet::getInterfaceCurvature from the node normals (grad_nodes of
    phi_ifn) and the blended phi; one value per crossing point."""
    tets = _c(tets, np.int32)
    xyz = _c(xyz, np.float32)
    n = len(xyz)
    off, ids = node_tets(n, tets)
    g = _c(normals, np.float32)
    phi = _c(phi, np.float64)
    ifp2e = _c(ifp2e, np.int32)
    edges = _c(edges, np.int32)
    out = np.empty(len(ifp2e), np.float32)
    _o.ora_interface_curvature(n, _p(xyz), len(tets), _p(tets), _p(off), _p(ids), _p(g), _p(phi),
                               len(ifp2e), _p(ifp2e), _p(edges), _p(out))
    return out


def surface_tension(normals, phi, ifp2e, edges, k_if, area, st, rho, fx=None, fy=None, fz=None):
    """Flow::surfaceTension_ in serial crossing-point order; adds into fx, fy,
    fz (float64 per node, zeros when not given) and returns them."""
    g = _c(normals, np.float32)
    n = len(g)
    out = []
    for a in (fx, fy, fz):
        out.append(np.zeros(n, np.float64) if a is None else np.array(a, np.float64, copy=True))
    phi, ifp2e, edges = _c(phi, np.float64), _c(ifp2e, np.int32), _c(edges, np.int32)
    k_if, area = _c(k_if, np.float32), _c(area, np.float32)
    _o.ora_surface_tension(_p(g), _p(phi), len(ifp2e), _p(ifp2e), _p(edges), _p(k_if), _p(area),
                           float(st), int(rho), _p(out[0]), _p(out[1]), _p(out[2]))
    return tuple(out)


FLT_MIN = float(np.finfo(np.float32).tiny)


def fem_quad():
    """P1 basis values at the 4 quadrature points (float-rounded) and the weight."""
    phi = np.empty(16, np.float64)
    w = np.empty(1, np.float64)
    _o.ora_fem_quad(_p(phi), _p(w))
    return phi.reshape(4, 4), float(w[0])


def fem_pattern(n_nodes, tets):
    """Sparsity of Assembler::bilin_with_p2t: row i holds i and every node
    that shares a tet with it, in increasing order (CSR: row_off, cols)."""
    t = np.asarray(tets, np.int64)
    r = np.repeat(t, 4, axis=1).ravel()
    c = np.tile(t, (1, 4)).ravel()
    key = np.unique(r * n_nodes + c)
    rows = key // n_nodes
    row_off = np.zeros(n_nodes + 1, np.int64)
    row_off[1:] = np.cumsum(np.bincount(rows, minlength=n_nodes))
    return row_off, (key % n_nodes).astype(np.int32)


def fem_assemble(form, xyz, tets, row_off, cols, vel=None):
    """Assembler::assemble_bilin of the mass (form 0) or advection (form 1,
    velocity vel = [vx, vy, vz] per node, 3n) bilinear form on the pattern."""
    xyz, tets = _c(xyz, np.float32), _c(tets, np.int32)
    n = len(xyz)
    off, ids = node_tets(n, tets)
    vals = np.empty(len(cols), np.float64)
    v = None if vel is None else _c(vel, np.float64)
    bad = _o.ora_fem_assemble(int(form), n, _p(xyz), _p(tets), _p(off), _p(ids), _p(_c(row_off, np.int64)),
                              _p(_c(cols, np.int32)), _p(v), _p(vals))
    return vals, int(bad)


def spmv(row_off, cols, vals, x):
    x = _c(x, np.float64)
    y = np.empty(len(row_off) - 1, np.float64)
    _o.ora_spmv(len(y), _p(_c(row_off, np.int64)), _p(_c(cols, np.int32)), _p(_c(vals, np.float64)),
                _p(x), _p(y))
    return y


def bicgstab(row_off, cols, vals, b, x0, tol, max_iters):
    """Eigen's BiCGSTAB with the Jacobi preconditioner from guess x0:
    (x, converged, iterations, error, restarts)."""
    x = np.array(x0, np.float64, copy=True)
    it, err, rs = np.zeros(1, np.int64), np.zeros(1, np.float64), np.zeros(1, np.int64)
    ok = _o.ora_bicgstab(len(x), _p(_c(row_off, np.int64)), _p(_c(cols, np.int32)),
                         _p(_c(vals, np.float64)), _p(_c(b, np.float64)), _p(x), float(tol),
                         int(max_iters), _p(it), _p(err), _p(rs))
    return x, bool(ok), int(it[0]), float(err[0]), int(rs[0])


def advect_system(xyz, tets, phi, flow, dt):
    """The Crank-Nicolson system of LevelSet::update (LevelSet.cpp:99-129):
    (row_off, cols, L = M + s A, b = (M - s A) phi), s = 0.5 dt."""
    n = len(xyz)
    row_off, cols = fem_pattern(n, tets)
    M, _ = fem_assemble(0, xyz, tets, row_off, cols)
    A, _ = fem_assemble(1, xyz, tets, row_off, cols, np.asarray(flow, np.float64)[:3 * n])
    s = 0.5 * dt
    L = M + s * A
    R = M - s * A
    return row_off, cols, L, spmv(row_off, cols, R, phi)


def advect(xyz, tets, phi, flow, dt, tol=0.0, max_iters=0):
    """The advection of LevelSet::update (LevelSet.cpp:89-141): phi_n and
    dict(iterations, error, restarts, converged, skipped)."""
    n = len(xyz)
    flow = np.asarray(flow, np.float64)
    if dt < FLT_MIN or np.sqrt(np.sum(flow * flow)) < FLT_MIN:
        return np.array(phi, np.float64, copy=True), dict(skipped=True, converged=True, iterations=0,
                                                         error=0.0, restarts=0)
    if np.isnan(flow[:3 * n]).any():
        raise ValueError("velocity has NaN")
    row_off, cols, L, b = advect_system(xyz, tets, phi, flow, dt)
    tol = tol if tol > 0 else float(np.finfo(np.float64).eps)
    max_iters = max_iters if max_iters > 0 else 2 * n
    x, ok, it, err, rs = bicgstab(row_off, cols, L, b, phi, tol, max_iters)
    return x, dict(skipped=False, converged=ok, iterations=it, error=err, restarts=rs)
