"""Contact scene: a Neo-Hookean ball dropped at 2 m/s on the plane z = 0, with an IPC
log-barrier per surface vertex (conditional potential, active below dhat) and the
plane step filter in the line search.

CPU tests pin the mesher (numpy restatement), the oracle's barrier (closed form vs
central differences) and non-penetration of the oracle's trajectory. GPU tests run the
HIP path through the C ABI against the oracle step by step, mixed precision against
FP64, and the device step filter against numpy. Reference parity: SymX has no barrier
or step filter code (its example contact is a cubic penalty), so these rows are pinned by
the oracle and the invariants only.
"""
import numpy as np
import pytest

import symx_amd
from oracle_py import SceneData, sphere_arrays
from scene_np import sphere_mesh


def small_params(cells=8, **kw):
    # center high enough that the jitter of the lowest vertices keeps them above the plane
    return symx_amd.default_sphere_params(cells=cells, center=[0.0, 0.0, 0.51 + 0.1 * 1.0 / cells], **kw)


def oracle_scene(p):
    X, T, C = symx_amd.host_sphere_mesh(p)
    A = X[T[:, 0]]
    vol = np.abs(np.linalg.det(np.stack([X[T[:, 1]] - A, X[T[:, 2]] - A, X[T[:, 3]] - A], 2))) / 6
    mass = np.zeros(len(X))
    np.add.at(mass, T.ravel(), np.repeat(vol * p.density / 4, 4))
    E, nu = p.youngs_modulus, p.poissons_ratio
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    v0 = np.tile(np.array(list(p.velocity)), (len(X), 1))
    return SceneData(T, X, X.copy(), X.copy(), v0, mass, np.zeros(0, np.int32), np.zeros((0, 3)), lam, mu, p.dt, 0.0, list(p.gravity),
                     contact=C, kappa=p.contact_stiffness, dhat=p.dhat, step_eta=p.step_filter)


def oracle_step(oracle, S, newton):
    r = oracle.newton(S, newton)
    x = S.x.copy()
    S.v0[...] = (x - S.x0) / S.dt
    S.x0[...] = x
    return r


# ---- CPU ------------------------------------------------------------------------------
@pytest.mark.parametrize("cells", [6, 8, 13])
def test_sphere_mesh_matches_numpy_restatement(cells):
    p = symx_amd.default_sphere_params(cells=cells)
    X, T, C = symx_amd.host_sphere_mesh(p)
    Xn, Tn, Cn = sphere_mesh(radius=p.radius, cells=cells, jitter=p.jitter, seed=p.seed, center=list(p.center))
    assert np.array_equal(T, Tn) and np.array_equal(C, Cn)
    assert np.array_equal(X, Xn)


def test_sphere_mesh_full_size():
    p = symx_amd.default_sphere_params()
    X, T, C = symx_amd.host_sphere_mesh(p)
    assert 380_000 <= T.shape[0] <= 420_000
    A = X[T[:, 0]]
    vol = np.linalg.det(np.stack([X[T[:, 1]] - A, X[T[:, 2]] - A, X[T[:, 3]] - A], 2)) / 6
    assert vol.min() > 0.0
    assert abs(vol.sum() - 4.0 / 3.0 * np.pi * p.radius ** 3) < 0.01 * vol.sum()
    assert X[:, 2].min() > 0.0


def test_oracle_barrier_matches_finite_differences(oracle):
    """Energy of a single tet with two vertices inside the barrier band: gradient and
    (through the assembled block diagonal) the barrier Hessian against differences."""
    X = np.array([[0, 0, 5e-4], [0.1, 0, 8e-4], [0, 0.1, 3e-3], [0, 0, 0.1]], float)
    T = np.array([[0, 1, 2, 3]], np.int32)
    S = SceneData(T, X, X.copy(), X.copy(), np.zeros_like(X), np.ones(4), np.zeros(0, np.int32), np.zeros((0, 3)), 0.0, 0.0, 0.0, 0.0,
                  [0, 0, 0], contact=np.array([0, 1, 2], np.int32), kappa=1e6, dhat=1e-3)
    x = X.reshape(-1).copy()
    E, g = oracle.energy(S, x)
    h = 1e-9
    for k in [2, 5, 8]:
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (oracle.energy(S, xp, False) - oracle.energy(S, xm, False)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))
    # closed form: b(d) = -kappa (d - dhat)^2 ln(d / dhat)
    d = 5e-4
    assert abs(g[2] - (-1e6 * (2 * (d - 1e-3) * np.log(d / 1e-3) + (d - 1e-3) ** 2 / d))) <= 1e-9
    assert g[8] == 0.0  # vertex 2 is outside the band
    xin = x.copy()
    xin[2] = -1e-4
    assert oracle.energy(S, xin, False) == np.inf


def test_oracle_contact_keeps_vertices_above_the_plane(oracle):
    p = small_params(cells=6)
    S = oracle_scene(p)
    newton = symx_amd.default_newton_params(precision=symx_amd.FP64)
    for _ in range(12):
        r = oracle_step(oracle, S, newton)
        assert r["result"] == 0
        assert S.x[:, 2].min() > 0.0
    assert S.x[:, 2].min() < p.dhat  # contact reached
    assert S.v0[:, 2].mean() > -2.0  # and it decelerates the ball


# ---- GPU ------------------------------------------------------------------------------
@pytest.mark.gpu
def test_sphere_scene_inputs_match_oracle_arrays(gpu):
    p = small_params()
    s = symx_amd.Sphere(p)
    X, T, C = symx_amd.host_sphere_mesh(p)
    assert np.array_equal(s.get("X"), X) and np.array_equal(s.connectivity(), T) and np.array_equal(s.contact_vertices(), C)
    S = oracle_scene(p)
    np.testing.assert_allclose(s.get("mass"), S.mass, rtol=1e-14, atol=0)
    np.testing.assert_array_equal(s.get("v"), S.v0)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 3])
def test_sphere_contact_matches_oracle(gpu, oracle, mode):
    p = small_params()
    s = symx_amd.Sphere(p, symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=mode))
    S = sphere_arrays(s)
    newton = s.newton
    for step in range(8):
        st = s.step()
        r = oracle_step(oracle, S, newton)
        assert st["result"] == r["result"] == 0, (step, st["result_name"], r["result"])
        assert abs(st["newton_iterations"] - r["newton_iterations"]) <= 1, (step, st["newton_iterations"], r["newton_iterations"])
        x = s.get("x")
        disp = np.abs(S.x - S.X).max()
        assert np.abs(x - S.x).max() <= 1e-6 * disp, (step, np.abs(x - S.x).max(), disp)
        assert x[:, 2].min() > 0.0
    assert S.x[:, 2].min() < p.dhat


@pytest.mark.gpu
def test_sphere_contact_mixed_precision_tracks_fp64(gpu):
    p = small_params()
    s64 = symx_amd.Sphere(p, symx_amd.default_newton_params(precision=symx_amd.FP64))
    s32 = symx_amd.Sphere(p, symx_amd.default_newton_params(precision=symx_amd.MIXED))
    for step in range(8):
        a, b = s64.step(), s32.step()
        assert a["result"] == 0 and b["result"] == 0, (step, a["result_name"], b["result_name"])
    x64, x32 = s64.get("x"), s32.get("x")
    assert np.abs(x32 - x64).max() <= 1e-5 * np.abs(x64 - s64.get("X")).max()
    assert x32[:, 2].min() > 0.0


@pytest.mark.gpu
def test_device_plane_step_filter_matches_numpy(gpu):
    import ctypes

    p = small_params()
    s = symx_amd.Sphere(p)
    C = s.contact_vertices()
    x = s.get("x").reshape(-1)
    rng = np.random.default_rng(3)
    du = rng.uniform(-0.02, 0.01, x.size)
    s.set_dofs(x)
    s.upload(symx_amd.VEC_DU, du)
    t = ctypes.c_double()
    symx_amd.check_hip(symx_amd.hip().symx_hip_problem_plane_max_step(s.problem, ctypes.c_double(0.9), ctypes.byref(t)))
    z, dz = x.reshape(-1, 3)[C, 2], du.reshape(-1, 3)[C, 2]
    neg = dz < 0
    ref = min(1.0, np.min(0.9 * z[neg] / -dz[neg])) if neg.any() else 1.0
    assert t.value == ref
