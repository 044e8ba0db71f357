"""Oracle of the advection step of LevelSet::update (LevelSet.cpp:89-141):
the P1 mass and advection forms (Assembler::bilin_with_p2t) and Eigen's
BiCGSTAB, pinned by closed forms and a direct solve."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import fluidtissue_amd as fa
import oracle


def _mesh(n=9, jitter=0.2, seed=3):
    return fa.kuhn_mesh(n, jitter=jitter, seed=seed)


def _tet_vols(m):
    x = m.xyz.astype(np.float64)[m.tets]
    return np.abs(np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]))) / 6


def _dense(row_off, cols, vals, n):
    return sp.csr_matrix((vals, cols, row_off), shape=(n, n)).toarray()


def test_quadrature_values():
    """Degree-2 rule: the float-rounded P1 basis at the 4 points sums to 1 and
    integrates u v exactly (1/10 on the diagonal, 1/20 off it, times 6|T|)."""
    phi, w = oracle.fem_quad()
    assert w == 0.25 / 6.0
    assert np.allclose(phi.sum(0), 1.0, atol=1e-7)
    mass = (phi[:, None, :] * phi[None, :, :] * w).sum(-1) * 6
    assert np.allclose(mass, np.where(np.eye(4) > 0, 0.1, 0.05), atol=1e-7)


def test_mass_matrix_closed_form():
    """P1 mass: M_ii = sum of |T|/10 over the tets of i, M_ij = sum of |T|/20
    over the tets holding both (exact for the degree-2 rule up to the float
    basis values); the total is the domain volume."""
    m = _mesh()
    n = len(m.xyz)
    row_off, cols = oracle.fem_pattern(n, m.tets)
    M, bad = oracle.fem_assemble(0, m.xyz, m.tets, row_off, cols)
    assert bad == -1
    vol = _tet_vols(m)
    ref = np.zeros((n, n))
    for t, v in zip(m.tets, vol):
        for a in t:
            for b in t:
                ref[a, b] += v / 10 if a == b else v / 20
    D = _dense(row_off, cols, M, n)
    assert np.allclose(D, ref, rtol=1e-6, atol=1e-12)
    assert abs(D.sum() - 1.0) < 1e-6                      # unit cube
    assert np.array_equal(D, D.T)


def test_advection_matrix_closed_forms():
    """A_ij = int (V . grad phi_j) phi_i: rows sum to 0 (the basis sums to 1);
    for constant V and linear u = c.x, A u = (V.c) M 1; for a linear V the
    row sums of the transpose weigh div V."""
    m = _mesh()
    n = len(m.xyz)
    row_off, cols = oracle.fem_pattern(n, m.tets)
    M, _ = oracle.fem_assemble(0, m.xyz, m.tets, row_off, cols)
    rng = np.random.default_rng(1)
    V = rng.standard_normal(3 * n)
    A, _ = oracle.fem_assemble(1, m.xyz, m.tets, row_off, cols, V)
    Ad = _dense(row_off, cols, A, n)
    assert np.max(np.abs(Ad.sum(1))) < 1e-12 * np.abs(Ad).sum(1).max()
    Vc = np.array([0.3, -1.1, 0.7])
    Ac, _ = oracle.fem_assemble(1, m.xyz, m.tets, row_off, cols, np.repeat(Vc, n))
    c = np.array([0.5, 0.25, -2.0])
    u = m.xyz.astype(np.float64) @ c
    lhs = oracle.spmv(row_off, cols, Ac, u)
    rhs = (Vc @ c) * oracle.spmv(row_off, cols, M, np.ones(n))
    assert np.allclose(lhs, rhs, rtol=1e-6, atol=1e-9 * np.abs(rhs).max())


def test_bicgstab_matches_direct_solve():
    m = _mesh(8)
    n = len(m.xyz)
    x = m.xyz.astype(np.float64)
    V = np.concatenate([-(x[:, 1] - 0.5), x[:, 0] - 0.5, 0.2 * np.ones(n)])     # rotation + drift
    phi = np.linalg.norm(x - 0.45, axis=1) - 0.3
    row_off, cols, L, b = oracle.advect_system(m.xyz, m.tets, phi, np.concatenate([V, np.zeros(n)]), 0.05)
    xs, ok, it, err, rs = oracle.bicgstab(row_off, cols, L, b, phi, 1e-12, 1000)
    assert ok and err <= 1e-12 and 0 < it < 100
    direct = spl.spsolve(sp.csr_matrix((L, cols, row_off), shape=(n, n)).tocsc(), b)
    assert np.max(np.abs(xs - direct)) < 1e-10 * np.abs(direct).max()
    r = b - oracle.spmv(row_off, cols, L, xs)
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(b) * 1.0001


def test_crank_nicolson_translates_a_plane_exactly():
    """phi = x - a under constant V = (v, 0, 0): the Galerkin CN step is exact
    for linear fields (A 1 = 0 and A phi = v M 1), phi_n = phi - v dt."""
    m = _mesh(10)
    n = len(m.xyz)
    x = m.xyz.astype(np.float64)
    phi = x[:, 0] - 0.43
    v, dt = 0.8, 0.02
    flow = np.concatenate([np.full(n, v), np.zeros(n), np.zeros(n), np.ones(n)])
    out, info = oracle.advect(m.xyz, m.tets, phi, flow, dt, tol=1e-13, max_iters=500)
    assert info["converged"] and not info["skipped"]
    assert np.max(np.abs(out - (phi - v * dt))) < 1e-10


def test_advect_edge_cases():
    m = _mesh(5)
    n = len(m.xyz)
    phi = m.xyz[:, 0].astype(np.float64) - 0.5
    out, info = oracle.advect(m.xyz, m.tets, phi, np.zeros(4 * n), 0.1)
    assert info["skipped"] and np.array_equal(out, phi)                 # |flow| < FLT_MIN
    out, info = oracle.advect(m.xyz, m.tets, phi, np.ones(4 * n), 0.0)
    assert info["skipped"]                                              # dt < FLT_MIN
    bad = np.ones(4 * n)
    bad[5] = np.nan
    try:
        oracle.advect(m.xyz, m.tets, phi, bad, 0.1)
        assert False
    except ValueError:
        pass
