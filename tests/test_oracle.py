"""Pins of the C oracle (oracle/symx_oracle.c) against mathematics and numpy/scipy,
independent of the code under test. CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from scene_np import tet4_block


def nh_energy_np(x12, X12, lam, mu):
    """Neo-Hookean Tet4 energy written directly from examples/elasticity_potentials.h:7-13."""
    x = x12.reshape(4, 3)
    X = X12.reshape(4, 3)
    Ds = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], 1)
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], 1)
    F = Ds @ np.linalg.inv(Dm)
    J = np.linalg.det(F)
    vol = np.linalg.det(Dm) / 6.0
    return vol * (0.5 * mu * (np.trace(F.T @ F) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2)


def random_tet(rng, scale=0.15):
    X = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float) * rng.uniform(0.5, 1.5)
    X = X + rng.uniform(-0.1, 0.1, 12)
    x = X + rng.uniform(-scale, scale, 12)
    return x, X


def test_nh_tet4_gradient_hessian_match_finite_differences(oracle):
    rng = np.random.default_rng(0)
    lam, mu = 142857.14285714287, 35714.28571428572
    for _ in range(6):
        x, X = random_tet(rng)
        E, g, H = oracle.nh_tet4(x, X, lam, mu)
        assert np.isclose(E, nh_energy_np(x, X, lam, mu), rtol=1e-12)
        h = 1e-6
        gfd = np.array([(nh_energy_np(x + h * e, X, lam, mu) - nh_energy_np(x - h * e, X, lam, mu)) / (2 * h) for e in np.eye(12)])
        assert np.abs(g - gfd).max() <= 1e-6 * np.abs(g).max()
        Hfd = np.stack([(oracle.nh_tet4(x + h * e, X, lam, mu)[1] - oracle.nh_tet4(x - h * e, X, lam, mu)[1]) / (2 * h) for e in np.eye(12)], 1)
        assert np.abs(H - Hfd).max() <= 1e-6 * np.abs(H).max()
        # invariants of a translation-invariant energy: gradient sums to zero, H t = 0, symmetric
        assert np.abs(g.reshape(4, 3).sum(0)).max() <= 1e-9 * np.abs(g).max()
        assert np.abs(H @ np.tile([1.0, 0, 0], 4)).max() <= 1e-9 * np.abs(H).max()
        assert np.array_equal(H, H.T) or np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


def test_nh_tet4_rest_state_closed_form(oracle):
    # At F = I: E = 0, gradient = 0 (a dropped or sign-flipped term breaks one of these).
    X = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float)
    E, g, H = oracle.nh_tet4(X, X, 2.0, 3.0)
    assert abs(E) < 1e-15 and np.abs(g).max() < 1e-14
    # For a pure dilation F = s I of the unit tet: E = vol [ mu/2 (3 s^2 - 3) - 3 mu ln s + lambda/2 (3 ln s)^2 ]
    s = 1.1
    E, g, H = oracle.nh_tet4(s * X, X, 2.0, 3.0)
    ref = (1 / 6) * (0.5 * 3.0 * (3 * s * s - 3) - 3.0 * 3 * np.log(s) + 0.5 * 2.0 * (3 * np.log(s)) ** 2)
    assert np.isclose(E, ref, rtol=1e-14)


def eigh_project(A, eps, mirror):
    w, V = np.linalg.eigh(A)
    if not np.any(w < eps):
        return A.copy(), False
    w = np.where(w < eps, -w if mirror else eps, w)
    return (V * w) @ V.T, True


@pytest.mark.parametrize("n", [3, 6, 12, 15])
@pytest.mark.parametrize("mirror", [False, True])
def test_projection_matches_eigh(oracle, n, mirror):
    rng = np.random.default_rng(n + 10 * mirror)
    M = rng.normal(size=(40, n, n))
    A = M + np.transpose(M, (0, 2, 1))
    P, changed = oracle.project(A, eps=1e-10, mirror=mirror)
    for i in range(A.shape[0]):
        R, ch = eigh_project(A[i], 1e-10, mirror)
        assert ch == changed[i]
        assert np.abs(P[i] - R).max() <= 1e-12 * np.abs(A[i]).max()


def test_projection_invariants(oracle):
    rng = np.random.default_rng(3)
    M = rng.normal(size=(30, 12, 12))
    A = M + np.transpose(M, (0, 2, 1))
    P, changed = oracle.project(A, eps=1e-3, mirror=False)
    assert changed.all()
    for i in range(A.shape[0]):
        assert np.linalg.eigvalsh(P[i]).min() >= 1e-3 * (1 - 1e-9) - 1e-12
    # idempotent up to round-off, and a PD matrix is returned bitwise unchanged
    P2, ch2 = oracle.project(P, eps=1e-3 * (1 - 1e-6), mirror=False)
    assert not ch2.any() and np.array_equal(P2, P)
    # mirroring maps lambda -> |lambda|: result equals the matrix absolute value
    Pm, _ = oracle.project(A, eps=1e-10, mirror=True)
    for i in range(A.shape[0]):
        w, V = np.linalg.eigh(A[i])
        assert np.abs(Pm[i] - (V * np.abs(w)) @ V.T).max() <= 1e-12 * np.abs(A[i]).max()


def spd_block_matrix(nb, rng):
    """Random SPD block tridiagonal-ish matrix in block CSR (ascending columns)."""
    rows, cols, vals = [], [], []
    blocks = {}
    for r in range(nb):
        for c in {max(r - 1, 0), r, min(r + 1, nb - 1), (r * 7 + 3) % nb}:
            if (c, r) in blocks:
                blocks[(r, c)] = blocks[(c, r)].T.copy()
            elif (r, c) not in blocks:
                blocks[(r, c)] = rng.normal(size=(3, 3)) * 0.3
    for r in range(nb):
        B = blocks[(r, r)]
        blocks[(r, r)] = B + B.T + 6.0 * np.eye(3)
    rp = [0]
    for r in range(nb):
        cs = sorted(c for (rr, c) in blocks if rr == r)
        for c in cs:
            cols.append(c)
            vals.append(blocks[(r, c)])
        rp.append(len(cols))
    rp = np.array(rp, np.int64)
    cols = np.array(cols, np.int32)
    vals = np.array(vals)
    dense = np.zeros((3 * nb, 3 * nb))
    for r in range(nb):
        for k in range(rp[r], rp[r + 1]):
            dense[3 * r:3 * r + 3, 3 * cols[k]:3 * cols[k] + 3] = vals[k]
    return rp, cols, vals, dense


def textbook_pcg(A, b, Dinv, abs_tol, rel_tol, max_iter):
    """Preconditioned CG (Saad, Iterative Methods, Alg. 9.1) with the stopping rules of solve_pcg.h:119-210."""
    x = np.zeros_like(b)
    bb = b @ b
    if bb < abs_tol ** 2:
        return x, 0
    r = b - A @ x
    err0 = np.sqrt(r @ r / bb)
    z = Dinv @ r
    p = z.copy()
    rz = r @ z
    for it in range(1, max_iter + 1):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        err = np.sqrt(r @ r / bb)
        if err < abs_tol or err / err0 < rel_tol:
            return x, it
        z = Dinv @ r
        rz_new = r @ z
        p = z + rz_new / rz * p
        rz = rz_new
    return x, max_iter


def test_pcg_matches_textbook_and_direct_solve(oracle):
    rng = np.random.default_rng(7)
    nb = 60
    rp, cols, vals, A = spd_block_matrix(nb, rng)
    b = rng.normal(size=3 * nb)
    Dinv = np.zeros_like(A)
    for r in range(nb):
        Dinv[3 * r:3 * r + 3, 3 * r:3 * r + 3] = np.linalg.inv(A[3 * r:3 * r + 3, 3 * r:3 * r + 3])
    for rel in [1e-3, 1e-8]:
        x, info = oracle.pcg(rp, cols, vals, b, abs_tol=1e-14, rel_tol=rel)
        xt, it = textbook_pcg(A, b, Dinv, 1e-14, rel, 1000)
        assert info["converged"] and info["n_iterations"] == it
        assert np.abs(x - xt).max() <= 1e-10 * np.abs(xt).max()
    x, info = oracle.pcg(rp, cols, vals, b, abs_tol=1e-14, rel_tol=1e-13)
    xd = spla.spsolve(sp.csr_matrix(A), b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    # zero right-hand side: immediate exit with x = 0
    x0, info0 = oracle.pcg(rp, cols, vals, np.zeros_like(b), abs_tol=1e-12, rel_tol=1e-3)
    assert info0["converged"] and info0["n_iterations"] == 0 and not x0.any()
    # indefinite matrix: stop on p.Ap <= 0
    valsn = vals.copy()
    for r in range(nb):
        for k in range(rp[r], rp[r + 1]):
            if cols[k] == r:
                valsn[k] -= 20.0 * np.eye(3)
    _, infon = oracle.pcg(rp, cols, valsn, b, abs_tol=1e-14, rel_tol=1e-10)
    assert infon["found_indefiniteness"] and not infon["converged"]


def test_oracle_newton_converges_to_stationary_point(oracle):
    from symx_amd import NewtonParams

    S = tet4_block(cells=(3, 3, 3), noise=0.2 / 3)
    np_ = NewtonParams(0, 1, 0, 100, 0.0, 1e-8, 1e-6, 1e-12, 10000, 1e-10, 0.0)
    E0, g0 = oracle.energy(S)
    res = oracle.newton(S, np_)
    assert res["result"] == 0
    E1, g1 = oracle.energy(S)
    assert E1 < E0
    assert np.abs(g1).max() <= 1e-8 * np.abs(g0).max()
    assert np.all(np.diff(res["energies"]) <= 1e-13 * abs(res["energies"][0]))  # monotone up to round-off
