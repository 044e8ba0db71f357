"""Point-triangle distance as a constrained quadratic minimization -- TEST
INFRASTRUCTURE ONLY (a second, independent checker of triangle distances).

The oracle (ftoracle.c) computes a point-triangle distance geometrically:
the plane distance when the point projects inside (edge functions of the
normal), else the nearest of the three edge segments.  The device's FP64
settle does the same.  This module reaches the same number another way, so
that an error figure does not rest on one formula: with the triangle
parametrized as T(s, t) = V0 + s E0 + t E1 over s, t >= 0, s + t <= 1, the
squared distance from P is the quadratic

    Q(s, t) = a s^2 + 2 b s t + c t^2 + 2 d s + 2 e t + f,
    a = E0.E0, b = E0.E1, c = E1.E1, d = E0.D, e = E1.D, f = D.D, D = V0 - P,

a convex function, whose minimum over the triangle is either its stationary
point (when that lies inside; solved by Cramer's rule) or the minimum over
one of the three sides, each a one-dimensional quadratic minimized in
closed form and clamped to its segment.  Evaluated in numpy's long double
(80-bit extended on x86-64), from the float coordinates promoted exactly.
"""
import numpy as np

LD = np.longdouble


def _side_min(q0, q1, q2):
    """Minimum over u in [0, 1] of q0 + 2 q1 u + q2 u^2 (q2 >= 0)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(q2 > 0, -q1 / np.where(q2 > 0, q2, 1), 0)
    u = np.clip(u, 0, 1)
    return q0 + 2 * q1 * u + q2 * u * u


def dist2(p, v0, v1, v2):
    """Squared distance from points p (n, 3) to triangles (v0, v1, v2), each
    (n, 3); long double."""
    p, v0, v1, v2 = (np.asarray(x, np.float64).astype(LD) for x in (p, v0, v1, v2))
    e0, e1, dd = v1 - v0, v2 - v0, v0 - p
    a = np.sum(e0 * e0, axis=1)
    b = np.sum(e0 * e1, axis=1)
    c = np.sum(e1 * e1, axis=1)
    d = np.sum(e0 * dd, axis=1)
    e = np.sum(e1 * dd, axis=1)
    f = np.sum(dd * dd, axis=1)
    det = a * c - b * b
    # stationary point of Q: [a b; b c] (s, t) = -(d, e)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(det > 0, (b * e - c * d) / np.where(det > 0, det, 1), -1)
        t = np.where(det > 0, (b * d - a * e) / np.where(det > 0, det, 1), -1)
    inside = (det > 0) & (s >= 0) & (t >= 0) & (s + t <= 1)
    q_in = a * s * s + 2 * b * s * t + c * t * t + 2 * d * s + 2 * e * t + f
    # side t = 0 (s in [0, 1]): f + 2 d s + a s^2
    m0 = _side_min(f, d, a)
    # side s = 0 (t in [0, 1]): f + 2 e t + c t^2
    m1 = _side_min(f, e, c)
    # side s + t = 1, s = 1 - u, t = u: Q(1 - u, u)
    #   = (a + 2 d + f) + 2 (b - a + e - d) u + (a - 2 b + c) u^2
    m2 = _side_min(a + 2 * d + f, b - a + e - d, a - 2 * b + c)
    side = np.minimum(m0, np.minimum(m1, m2))
    return np.maximum(np.where(inside, np.minimum(q_in, side), side), 0)


def node_distance(xyz, ifp, tris, nodes, faces):
    """Distances (long double) from nodes[k] to triangle faces[k] of the
    interface (ifp points, tris {a, b, c, ...})."""
    xyz = np.asarray(xyz)
    ifp = np.asarray(ifp)
    t = np.asarray(tris)[np.asarray(faces)]
    return np.sqrt(dist2(xyz[nodes], ifp[t[:, 0]], ifp[t[:, 1]], ifp[t[:, 2]]))
