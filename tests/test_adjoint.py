"""Adjoint of the quasistatic cantilever: dL/dlog_mu for L = 0.5 |x_eq - x_target|^2.

CPU tests pin the oracle (closed-form mixed partial, dense solve, central finite
differences through re-solved equilibria, the reference-composed adjoint of
oracle/_ref in tests/golden/adjoint_ref_*.npz) and check the generated adjoint kernel
on the host. GPU tests run the HIP path through the C ABI against the oracle, the
reference fixture, finite differences at full size (64 x 16 x 16 vertices, 85,050
tets) and a torch.autograd.Function harness.
"""
import os

import numpy as np
import pytest

from adjoint_harness import (GpuCantilever, loss_difference, block_params, log_mu_field, loss, oracle_equilibrium, oracle_hessian_csr,
                             oracle_newton_polish, scene_np, torch_equilibrium_function)

HERE = os.path.dirname(os.path.abspath(__file__))
SMALL = (15, 3, 3)     # 16 x 4 x 4 vertices, 270 tets
FULL = (63, 15, 15)    # 64 x 16 x 16 vertices, 85,050 tets


def fixture():
    return np.load(os.path.join(HERE, "golden", "adjoint_ref_cantilever16.npz"))


def random_tet(rng):
    X = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float) * 0.05 + rng.uniform(-0.004, 0.004, 12)
    return X + rng.uniform(-0.004, 0.004, 12), X


# ---- oracle pins (CPU) --------------------------------------------------------------
def test_mixed_partial_closed_form_matches_finite_differences(oracle):
    """d(g_e)/dlog_mu_e = g_e when both Lame parameters scale with mu_e."""
    rng = np.random.default_rng(5)
    ratio = 4.0
    for _ in range(8):
        x, X = random_tet(rng)
        p = np.log(3.5e4) + 0.3 * rng.standard_normal()
        h = 1e-5
        gp = oracle.nh_tet4(x, X, ratio * np.exp(p + h), np.exp(p + h))[1]
        gm = oracle.nh_tet4(x, X, ratio * np.exp(p - h), np.exp(p - h))[1]
        g = oracle.nh_tet4(x, X, ratio * np.exp(p), np.exp(p))[1]
        np.testing.assert_allclose((gp - gm) / (2 * h), g, rtol=1e-8, atol=1e-9 * np.abs(g).max())


@pytest.fixture(scope="module")
def small_problem(oracle):
    S = scene_np(SMALL)
    ne = S.tets.shape[0]
    log_mu_t = log_mu_field(ne, 2)
    x_t = oracle_equilibrium(oracle, S, log_mu_t)
    log_mu = log_mu_field(ne, 1)
    x = oracle_equilibrium(oracle, S, log_mu)
    S.x[...] = x.reshape(S.x.shape)
    S.mu_e = np.exp(log_mu)
    return S, log_mu, x, x_t


def test_oracle_equilibrium_is_stationary(oracle, small_problem):
    S, log_mu, x, _ = small_problem
    _, g = oracle.energy(S, x)
    # interior vertices: round-off level; the penalty rows carry k |x| eps
    free = np.ones(S.X.shape[0], bool)
    free[S.bc_vertex] = False
    assert np.abs(g.reshape(-1, 3)[free]).max() < 1e-12


def test_oracle_adjoint_matches_dense_solve(oracle, small_problem):
    """Dense numpy solve of H lambda = dL/dx and finite-difference mixed partials."""
    S, log_mu, x, x_t = small_problem
    g, lam, info = oracle.adjoint_logmu(S, x - x_t, rel_tol=1e-14)
    assert info["converged"] and not info["found_indefiniteness"]
    H = oracle_hessian_csr(oracle, S, x).toarray()
    lam_dense = np.linalg.solve(H, x - x_t)
    np.testing.assert_allclose(lam, lam_dense, rtol=0, atol=1e-10 * np.abs(lam_dense).max())
    # d^2E_e/dx dp_e by central differences of the element gradient in log(mu)
    xe = x.reshape(-1, 3)[S.tets].reshape(-1, 12)
    Xe = S.X[S.tets].reshape(-1, 12)
    h = 1e-5
    ref = np.zeros(S.tets.shape[0])
    for e in range(S.tets.shape[0]):
        mp, mm = np.exp(log_mu[e] + h), np.exp(log_mu[e] - h)
        gp = oracle.nh_tet4(xe[e], Xe[e], S.lambda_over_mu * mp, mp)[1]
        gm = oracle.nh_tet4(xe[e], Xe[e], S.lambda_over_mu * mm, mm)[1]
        ref[e] = -lam_dense.reshape(-1, 3)[S.tets[e]].reshape(12) @ ((gp - gm) / (2 * h))
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-8 * np.abs(ref).max())


def test_oracle_adjoint_matches_finite_differences(oracle, small_problem):
    """Central differences, h = 1e-6, on 12 entries (the 4 largest and 8 seeded) and one
    random direction; re-solves are plain Newton steps from the equilibrium."""
    S, log_mu, x, x_t = small_problem
    g, _, _ = oracle.adjoint_logmu(S, x - x_t, rel_tol=1e-14)
    rng = np.random.default_rng(32)
    ne = S.tets.shape[0]
    idx = list(np.argsort(-np.abs(g))[:4]) + list(rng.choice(ne, 8, replace=False))
    h = 1e-6
    for j in idx:
        lp, lm = log_mu.copy(), log_mu.copy()
        lp[j] += h
        lm[j] -= h
        fd = (loss(oracle_newton_polish(oracle, S, lp, x), x_t) - loss(oracle_newton_polish(oracle, S, lm, x), x_t)) / (2 * h)
        assert abs(fd - g[j]) <= 1e-5 * np.abs(g).max() + 1e-4 * abs(g[j]), (j, fd, g[j])
    v = rng.standard_normal(ne)
    fd = (loss(oracle_newton_polish(oracle, S, log_mu + h * v, x), x_t) - loss(oracle_newton_polish(oracle, S, log_mu - h * v, x), x_t)) / (2 * h)
    assert abs(fd - g @ v) <= 1e-6 * abs(g @ v)
    S.mu_e = np.exp(log_mu)


def test_oracle_adjoint_matches_reference_composition(oracle):
    """Against the reference's own pieces (oracle/_ref ref_driver adjoint): Hessian from
    SecondOrderCompiledGlobal, Eigen SimplicialLLT, mixed partial from Compiled<double>."""
    f = fixture()
    S = scene_np(tuple(f["cells"]))
    S.mu_e = np.exp(f["log_mu"])
    S.x[...] = f["x"].reshape(S.x.shape)
    g, lam, info = oracle.adjoint_logmu(S, f["x"] - f["x_target"], rel_tol=1e-14)
    np.testing.assert_allclose(g, f["dL_dlog_mu"], rtol=0, atol=1e-10 * np.abs(f["dL_dlog_mu"]).max())
    np.testing.assert_allclose(lam, f["lam"], rtol=0, atol=1e-10 * np.abs(f["lam"]).max())


# ---- generated adjoint kernel on the host (CPU) ----------------------------------------
def test_generated_adjoint_expression_matches_closed_form(oracle):
    """The emitted mixed partial d(lambda . g_e)/dlog_mu_e, evaluated on the host with the
    kernel's gather rules, equals lambda_e . g_e (oracle, closed form)."""
    import symx_amd

    p = block_params((3, 2, 2))
    p.noise = 0.002
    ne = 3 * 2 * 2 * 6
    log_mu = log_mu_field(ne, 3)
    scene = scene_np((3, 2, 2))
    from scene_np import splitmix64_uniform

    nv = scene.X.shape[0]
    fixed = np.abs(scene.X[:, 0]) < 1e-12
    u = splitmix64_uniform(p.seed, 3 * nv).reshape(nv, 3)
    x = scene.X + np.where(fixed[:, None], 0.0, p.noise * (2.0 * u - 1.0))
    lam = np.random.default_rng(9).standard_normal(3 * nv)
    for e in range(ne):
        out = symx_amd.host_scene_adjoint_element(p, log_mu, e, lam)
        t = scene.tets[e]
        mu = np.exp(log_mu[e])
        ge = oracle.nh_tet4(x[t].reshape(12), scene.X[t].reshape(12), scene.lambda_over_mu * mu, mu)[1]
        ref = lam.reshape(-1, 3)[t].reshape(12) @ ge
        assert abs(out[0] - ref) <= 1e-11 * max(1.0, np.abs(lam.reshape(-1, 3)[t].reshape(12) * ge).sum()), (e, out[0], ref)


def test_adjoint_modules_compile_for_gfx950(tmp_path):
    import symx_amd

    n, code, ops = symx_amd.host_scene_jit(block_params((3, 2, 2)), str(tmp_path))
    assert n == 3 and code > 0 and ops > 0


# ---- HIP path (GPU) ---------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_adjoint_matches_reference_fixture(gpu, oracle):
    f = fixture()
    m = GpuCantilever(tuple(f["cells"]))
    m.scene.set_log_mu(f["log_mu"])
    m.scene.set("x", f["x"].reshape(-1, 3))
    g, st = m.adjoint(f["x"] - f["x_target"])
    assert st["converged"] and not st["found_indefiniteness"]
    ref = f["dL_dlog_mu"]
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-8 * np.abs(ref).max())


@pytest.mark.gpu
def test_gpu_equilibrium_and_adjoint_match_oracle(gpu, oracle):
    S = scene_np(SMALL)
    ne = S.tets.shape[0]
    m = GpuCantilever(SMALL)
    log_mu_t = log_mu_field(ne, 2)
    x_t, _ = m.equilibrium(log_mu_t)
    log_mu = log_mu_field(ne, 1)
    x, st = m.equilibrium(log_mu)
    x_o = oracle_equilibrium(oracle, S, log_mu)
    np.testing.assert_allclose(x, x_o, rtol=0, atol=1e-11)
    g, ast = m.adjoint(x - x_t)
    S.mu_e = np.exp(log_mu)
    S.x[...] = x.reshape(S.x.shape)
    g_o, _, _ = oracle.adjoint_logmu(S, x - x_t, rel_tol=1e-14)
    np.testing.assert_allclose(g, g_o, rtol=0, atol=1e-8 * np.abs(g_o).max())


@pytest.mark.gpu
def test_gpu_adjoint_mixed_precision_scene_solves_in_fp64(gpu):
    """A mixed-precision scene's adjoint is solved in FP64 and equals the FP64 scene's;
    the FP32-storage solve through the C ABI is accurate only to the matrix rounding."""
    import symx_amd

    f = fixture()
    out = []
    for prec in (symx_amd.FP64, symx_amd.MIXED):
        m = GpuCantilever(tuple(f["cells"]), precision=prec)
        m.scene.set_log_mu(f["log_mu"])
        m.scene.set("x", f["x"].reshape(-1, 3))
        g, st = m.adjoint(f["x"] - f["x_target"], rel_tol=1e-12)
        assert st["converged"], st
        out.append(g)
    np.testing.assert_allclose(out[1], out[0], rtol=0, atol=1e-9 * np.abs(out[0]).max())
    # FP32 storage: H32 lambda = dL/dx refined in FP64 to 1e-10
    import ctypes

    rhs = np.ascontiguousarray(f["x"] - f["x_target"])
    lam64 = np.zeros_like(rhs)
    lam32 = np.zeros_like(rhs)
    for prec, dst in ((symx_amd.FP64, lam64), (symx_amd.MIXED, lam32)):
        m.scene.set_dofs(f["x"])
        m.scene.eval(symx_amd.EVAL_PGH)
        info = symx_amd.PcgInfo()
        symx_amd.check_hip(symx_amd.hip().symx_hip_problem_adjoint_solve(
            m.scene.problem, ctypes.c_int32(prec), ctypes.c_double(1e-10), ctypes.c_int32(100000),
            rhs.ctypes.data_as(ctypes.c_void_p), dst.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info)))
        assert info.converged
    assert np.abs(lam32 - lam64).max() <= 1e-2 * np.abs(lam64).max()


@pytest.mark.gpu
def test_gpu_adjoint_full_cantilever_matches_finite_differences(gpu):
    """64 x 16 x 16 vertices (85,050 log(mu) parameters): central differences with
    h = 1e-6 on 32 seeded entries plus one random direction."""
    m = GpuCantilever(FULL)
    ne = m.scene.n_elements
    assert ne == 85050
    x_t, st_t = m.equilibrium(log_mu_field(ne, 2))
    log_mu = log_mu_field(ne, 1)
    x, st = m.equilibrium(log_mu)
    print("forward solves:", st_t["newton_iterations"], st["newton_iterations"], "Newton iterations", flush=True)
    g, ast = m.adjoint(x - x_t)
    assert ast["converged"] and not ast["found_indefiniteness"], ast
    rng = np.random.default_rng(20261203)
    idx = rng.choice(ne, 32, replace=False)
    gmax = np.abs(g).max()

    def fd_entry(j, h, steps):
        lp, lm = log_mu.copy(), log_mu.copy()
        lp[j] += h
        lm[j] -= h
        return loss_difference(m.resolve(lp, x, steps), m.resolve(lm, x, steps), x_t) / (2 * h)

    # h = 1e-6 (the configuration's step): the FP64 quantization of the positions (ulp(0.1) =
    # 1.4e-17 per component, summed over 49,152 DoFs against |x - x_t|) leaves a noise of about
    # 1e-12 = 1e-5 max|g| in every finite difference, so the bound is relative to max|g|.
    for j in idx:
        fd = fd_entry(j, 1e-6, 2)
        print(f"h=1e-6 entry {int(j)}: adjoint {g[j]:+.10e} fd {fd:+.10e} |err|/max|g| {abs(fd - g[j]) / gmax:.2e}", flush=True)
        assert abs(fd - g[j]) <= 3e-5 * gmax, (int(j), fd, g[j], gmax)
    # h = 1e-4: the noise drops to ~1e-14 and the central-difference truncation is ~h^2 = 1e-8
    # relative: 1e-5 relative per entry, above a floor of 1e-7 max|g|
    for j in idx:
        fd = fd_entry(j, 1e-4, 3)
        rel = abs(fd - g[j]) / abs(g[j])
        print(f"h=1e-4 entry {int(j)}: adjoint {g[j]:+.10e} fd {fd:+.10e} rel {rel:.2e}", flush=True)
        assert abs(fd - g[j]) <= 1e-5 * abs(g[j]) + 1e-7 * gmax, (int(j), fd, g[j], gmax)
    h = 1e-6
    v = rng.standard_normal(ne)
    fd = loss_difference(m.resolve(log_mu + h * v, x), m.resolve(log_mu - h * v, x), x_t) / (2 * h)
    print(f"random direction: adjoint {g @ v:+.10e} fd {fd:+.10e}", flush=True)
    assert abs(fd - g @ v) <= 1e-6 * abs(g @ v), (fd, g @ v)


@pytest.mark.gpu
def test_torch_autograd_harness(gpu, oracle):
    import torch

    Equilibrium = torch_equilibrium_function()
    m = GpuCantilever(SMALL)
    ne = m.scene.n_elements
    x_t, _ = m.equilibrium(log_mu_field(ne, 2))
    p = torch.tensor(log_mu_field(ne, 1), dtype=torch.float64, requires_grad=True)
    x = Equilibrium.apply(p, m)
    L = 0.5 * ((x - torch.from_numpy(x_t.reshape(-1, 3))) ** 2).sum()
    L.backward()
    S = scene_np(SMALL)
    S.mu_e = np.exp(p.detach().numpy())
    S.x[...] = x.detach().numpy().reshape(S.x.shape)
    g_o, _, _ = oracle.adjoint_logmu(S, x.detach().numpy().reshape(-1) - x_t, rel_tol=1e-14)
    np.testing.assert_allclose(p.grad.numpy(), g_o, rtol=0, atol=1e-8 * np.abs(g_o).max())
