"""Line-search energy test on the device (NewtonsMethod.cpp:565-591).

The Armijo test compares E(x0 + s du) - E(x0) with beta s du.g. Here the left side is
summed element by element against the element energies kept by the last gradient /
Hessian evaluation (symx_hip_problem_energy_change). These tests pin that sum against the
difference of two independent total energies (the oracle's, and the device's own energy
pass) on Tet4, Hex27 and the conditional barrier potential, and check that time steps of
the 6-tet block reach the 1e-6 relative residual in both precisions."""
import numpy as np
import pytest

import symx_amd
from oracle_py import scene_arrays

pytestmark = pytest.mark.gpu


def _perturbed(x, scale, seed):
    return x + scale * np.random.default_rng(seed).uniform(-1.0, 1.0, x.shape)


def test_energy_change_tet4_matches_oracle_difference(gpu, oracle):
    prm = symx_amd.default_block_params(cells=[5, 4, 4], noise=0.2 / 5, seed=3)
    s = symx_amd.Tet4Block(prm, symx_amd.default_newton_params())
    x0 = s.get("x")
    s.set_dofs(x0)
    E0 = s.eval(symx_amd.EVAL_PGH)
    assert abs(s.energy_change()) <= 1e-14 * abs(E0)  # same DoFs: every element difference is ~0
    S0 = scene_arrays(s)
    Eo0, _ = oracle.energy(S0)
    for k, scale in enumerate((1e-3, 1e-6, 1e-9)):
        x1 = _perturbed(x0, scale, 10 + k)
        s.set_dofs(x1)
        dE = s.energy_change()
        E1 = s.eval(symx_amd.EVAL_P)
        s.set("x", x1)
        Eo1, _ = oracle.energy(scene_arrays(s))
        s.set("x", x0)
        # the oracle's two totals carry round-off of about 1e-15 |E|; the device sum of
        # element differences must agree with them to that and with the device totals
        assert abs(dE - (Eo1 - Eo0)) <= 1e-12 * abs(Eo0)
        assert abs(dE - (E1 - E0)) <= 1e-12 * abs(E0)
    # a step whose energy change is ten ulps of the total energy still gives a decrease of
    # the right size: along -grad, dE = -h |g|^2 to first order (the second-order term is
    # below 1e-9 of it here)
    s.set_dofs(x0)
    s.eval(symx_amd.EVAL_PGH)
    g = s.download(symx_amd.VEC_GRAD)
    h = 10 * 2.2e-16 * abs(E0) / float(g @ g)
    x1 = x0.reshape(-1) - h * g
    s.set_dofs(x1)
    dE = s.energy_change()
    lin = float(g @ (x1 - x0.reshape(-1)))  # the step as rounded into x (exact differences)
    assert abs(lin) < 20 * 2.2e-16 * abs(E0)
    assert abs(dE - lin) <= 2e-2 * abs(lin)
    s.close()


def test_energy_change_hex27_structured_kernel(gpu):
    p = symx_amd.default_hex_params(cells=[2, 2, 2], nodes_per_element=27, noise=0.05, seed=5)
    s = symx_amd.HexBlock(p, symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=2))
    x0 = s.get("x")
    s.set_dofs(x0)
    E0 = s.eval(symx_amd.EVAL_PGH)
    assert abs(s.energy_change()) <= 1e-14 * abs(E0)
    s.set_dofs(_perturbed(x0, 1e-4, 4))
    dE = s.energy_change()
    E1 = s.eval(symx_amd.EVAL_P)
    assert abs(dE - (E1 - E0)) <= 1e-12 * abs(E0)
    s.close()


@pytest.mark.parametrize("gap", [0.6e-3, 1.2e-3])
def test_energy_change_conditional_barrier(gpu, gap):
    # The lowest surface vertex of the 8-cell ball starts `gap` above the plane, inside
    # (0.6e-3) or outside (1.2e-3) the barrier range dhat = 1e-3; moving the DoFs down and
    # up switches its barrier element on and off between reference and trial state.
    p0 = symx_amd.default_sphere_params(cells=8, center=[0.0, 0.0, 0.0])
    X, _, C = symx_amd.host_sphere_mesh(p0)
    zmin = X[C, 2].min()
    p = symx_amd.default_sphere_params(cells=8, center=[0.0, 0.0, gap - zmin])
    s = symx_amd.Sphere(p, symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=3))
    x0 = s.get("x")
    assert abs(x0[:, 2].min() - gap) < 1e-12
    s.set_dofs(x0)
    E0 = s.eval(symx_amd.EVAL_PGH)
    for dz in (-0.5e-3, 2e-3, -0.4e-3):
        x1 = x0.copy()
        x1[:, 2] += dz
        s.set_dofs(x1)
        dE = s.energy_change()
        E1 = s.eval(symx_amd.EVAL_P)
        assert np.isfinite(E1)
        assert abs(dE - (E1 - E0)) <= 1e-11 * max(abs(E0), abs(E1))
    s.close()


def test_energy_change_needs_a_reference_evaluation(gpu):
    prm = symx_amd.default_block_params(cells=[2, 2, 2], noise=0.0)
    s = symx_amd.Tet4Block(prm, symx_amd.default_newton_params())
    s.set_dofs(s.get("x"))
    with pytest.raises(RuntimeError, match="reference energies"):
        s.energy_change()
    s.close()


@pytest.mark.parametrize("precision", [symx_amd.FP64, symx_amd.MIXED])
def test_block64_time_steps_reach_1e6(gpu, precision):
    """Five implicit-Euler steps of the 6-tet block at 64^3 cells (1,572,864 tets) from the
    seeded state: every step must end Successful at 1e-6 relative residual."""
    prm = symx_amd.default_block_params(cells=[64] * 3, noise=0.2 / 64, seed=2)
    s = symx_amd.Tet4Block(prm, symx_amd.default_newton_params(precision=precision))
    for i in range(5):
        st = s.step()
        print(f"step {i}: {st['result_name']} newton={st['newton_iterations']} r0={st['first_residual']:.3e} "
              f"r={st['final_residual']:.3e}", flush=True)
        assert st["result_name"] == "Successful"
        assert st["final_residual"] <= 1e-6 * st["first_residual"]
    s.close()


def test_gradient_dot_matches_oracle_and_keeps_state(gpu, oracle):
    prm = symx_amd.default_block_params(cells=[4, 4, 5], noise=0.2 / 4, seed=8)
    s = symx_amd.Tet4Block(prm, symx_amd.default_newton_params())
    x0 = s.get("x")
    s.set_dofs(x0)
    E0 = s.eval(symx_amd.EVAL_PGH)
    g0 = s.download(symx_amd.VEC_GRAD)
    v = np.random.default_rng(3).standard_normal(g0.shape)
    s.upload(symx_amd.VEC_DU, v)
    x1 = _perturbed(x0, 1e-3, 21)
    s.set_dofs(x1)
    gv = s.gradient_dot(symx_amd.VEC_DU)
    s.set("x", x1)
    _, go1 = oracle.energy(scene_arrays(s))
    assert abs(gv - float(go1 @ v)) <= 1e-12 * float(np.abs(go1) @ np.abs(v))
    # VEC_GRAD and the reference energies of energy_change belong to x0 still
    assert np.array_equal(s.download(symx_amd.VEC_GRAD), g0)
    s.set_dofs(x0)
    assert abs(s.energy_change()) <= 1e-14 * abs(E0)
    s.close()
