"""Isoparametric elements (Hex8, Hex27 and the quadratic tetrahedron Tet10): the structured
FEM path (deferred isoparametric Jacobian, generated material derivatives, hand-written
contraction), the PD projection of 24x24 / 30x30 / 81x81 element Hessians, and Newton on
hexahedral and Tet10 blocks.

CPU tests pin the oracle's closed-form Neo-Hookean elements (central differences, rest
state, affine deformations, the reference's own fem_integrator and NewtonsMethod outputs)
and check the structured decomposition on the host against the generic expansion and the
oracle. GPU tests run the HIP path through the C ABI against the oracle and the reference.
"""
import os

import numpy as np
import pytest

import symx_amd
from scene_np import HEX_NODES, TET10_NODES, hex_block


def tol_rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def ref_element(nn, scale=0.05):
    if nn == 10:
        return np.array(TET10_NODES, float) * scale  # a right tetrahedron with legs 2 * scale
    return np.array(HEX_NODES[:nn], float) * scale  # a cube of side 2 * scale


# ---- CPU: oracle ---------------------------------------------------------------------------
@pytest.mark.parametrize("nn", [8, 10, 27])
def test_oracle_hex_matches_finite_differences(oracle, nn):
    rng = np.random.default_rng(nn)
    X = ref_element(nn)
    x = X * 1.05 + rng.uniform(-0.003, 0.003, X.shape)
    lam, mu = 2e5, 3e4
    E, g, H = oracle.nh_hex(x, X, lam, mu)
    h = 1e-7
    xf = x.reshape(-1)
    gfd = np.zeros_like(g)
    Hfd = np.zeros_like(H)
    for i in range(3 * nn):
        xp, xm = xf.copy(), xf.copy()
        xp[i] += h
        xm[i] -= h
        Ep, gp, _ = oracle.nh_hex(xp, X, lam, mu)
        Em, gm, _ = oracle.nh_hex(xm, X, lam, mu)
        gfd[i] = (Ep - Em) / (2 * h)
        Hfd[i] = (gp - gm) / (2 * h)
    assert tol_rel(g, gfd) < 1e-7
    assert tol_rel(H, Hfd) < 1e-7
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


@pytest.mark.parametrize("nn", [8, 10, 27])
def test_oracle_hex_affine_deformation_closed_form(oracle, nn):
    """x = A X: F = A at every quadrature point, so E = V psi(A), g sums to zero
    (translation invariance) and the rest state has E = 0, g = 0."""
    X = ref_element(nn)
    lam, mu = 2e5, 3e4
    E0, g0, _ = oracle.nh_hex(X, X, lam, mu)
    assert abs(E0) < 1e-12 and np.abs(g0).max() < 1e-9
    A = np.array([[1.1, 0.05, 0.0], [0.02, 0.95, 0.03], [0.0, -0.04, 1.02]])
    E, g, _ = oracle.nh_hex(X @ A.T, X, lam, mu)
    V = 0.1 ** 3 / (6.0 if nn == 10 else 1.0)
    J = np.linalg.det(A)
    psi = 0.5 * mu * (np.sum(A * A) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2
    # the truncated 10-digit Hex27 weights sum to 8 only to about 1e-10; the Tet10 rule's
    # four weights 0.041666667 sum to 1/6 (1 + 1.2e-8)
    if nn == 10:
        assert abs(E / (V * psi) - 4 * 0.041666667 * 6) <= 1e-12
    else:
        assert abs(E - V * psi) <= 1e-9 * abs(V * psi)
    assert np.abs(g.reshape(-1, 3).sum(axis=0)).max() <= 1e-9 * np.abs(g).max()


# ---- CPU: structured decomposition on the host ---------------------------------------------
@pytest.mark.parametrize("nn", [8, 10])
def test_structured_hex8_matches_generic_expansion(nn):
    p = symx_amd.default_hex_params(cells=[2, 2, 2], nodes_per_element=nn)
    for e in [0, 5]:
        E, g, H, st = symx_amd.host_hex_element(p, e, structured=True)
        E2, g2, H2, st2 = symx_amd.host_hex_element(p, e, structured=False)
        assert st and not st2
        assert abs(E - E2) <= 1e-13 * abs(E2)
        assert tol_rel(g, g2) < 1e-13 and tol_rel(H, H2) < 1e-13


@pytest.mark.parametrize("nn", [8, 10, 27])
def test_structured_hex_matches_oracle(oracle, nn):
    p = symx_amd.default_hex_params(cells=[2, 2, 2], nodes_per_element=nn)
    S = hex_block(cells=(2, 2, 2), nn=nn, noise=p.noise, seed=p.seed)
    for e in range(S.tets.shape[0]):
        E, g, H, st = symx_amd.host_hex_element(p, e, structured=True)
        mu = S.mu_e[e]
        Eo, go, Ho = oracle.nh_hex(S.x[S.tets[e]], S.X[S.tets[e]], S.lambda_over_mu * mu, mu)
        assert st
        assert abs(E - Eo) <= 1e-12 * abs(Eo)
        assert tol_rel(g, go) < 1e-12 and tol_rel(H, Ho) < 1e-12


@pytest.mark.parametrize("nn", [8, 10, 27])
def test_structured_hex_modules_compile_for_gfx950(tmp_path, nn):
    n, code, n_structured = symx_amd.host_hex_jit(symx_amd.default_hex_params(cells=[2, 2, 2], nodes_per_element=nn), str(tmp_path))
    assert n == 3 and code > 0 and n_structured == 1


# ---- GPU ------------------------------------------------------------------------------------
def upper_blocks_to_full(Hub, nb):
    n = 3 * nb
    H = np.zeros((Hub.shape[0], n, n))
    u = 0
    for a in range(nb):
        for b in range(a, nb):
            blk = Hub[:, 9 * u:9 * u + 9].reshape(-1, 3, 3)
            H[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            H[:, 3 * b:3 * b + 3, 3 * a:3 * a + 3] = np.transpose(blk, (0, 2, 1))
            u += 1
    return H


def hex_scene(nn, cells=(2, 2, 2), precision=symx_amd.FP64, **kw):
    p = symx_amd.default_hex_params(cells=list(cells), nodes_per_element=nn, **kw)
    s = symx_amd.HexBlock(p, symx_amd.default_newton_params(precision=precision))
    s.set_dofs(s.get("x"))
    return s, p


@pytest.mark.gpu
@pytest.mark.parametrize("nn", [8, 10, 27])
def test_hex_scene_inputs_match_numpy_restatement(gpu, nn):
    s, p = hex_scene(nn, cells=(3, 2, 2))
    S = hex_block(cells=(3, 2, 2), nn=nn, noise=p.noise, seed=p.seed)
    assert np.array_equal(s.connectivity(), S.tets)
    assert np.array_equal(s.get("X"), S.X) and np.array_equal(s.get("x"), S.x)
    np.testing.assert_allclose(s.young_moduli(), S.E, rtol=1e-15)
    np.testing.assert_allclose(s.get("mass"), S.mass, rtol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("nn", [8, 10, 27])
def test_hex_element_outputs_match_oracle(gpu, oracle, nn):
    s, p = hex_scene(nn)
    S = hex_block(cells=(2, 2, 2), nn=nn, noise=p.noise, seed=p.seed)
    E = s.eval(symx_amd.EVAL_PGH)
    Eo, go = oracle.energy(S)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    assert tol_rel(s.download(symx_amd.VEC_GRAD), go) < 1e-11
    G, Hub = s.element_outputs(0, S.tets.shape[0], nn)
    H = upper_blocks_to_full(Hub, nn)
    Ho = oracle.element_hessians(S)
    for e in range(S.tets.shape[0]):
        assert tol_rel(H[e], Ho[e]) < 1e-12, e


@pytest.mark.gpu
@pytest.mark.parametrize("nn", [8, 10, 27])
@pytest.mark.parametrize("mirror", [False, True])
def test_hex_projection_matches_oracle(gpu, oracle, nn, mirror):
    # a strongly perturbed state (every element indefinite, none inverted)
    s, p = hex_scene(nn, noise=0.25)
    S = hex_block(cells=(2, 2, 2), nn=nn, noise=0.25, seed=p.seed)
    ne = S.tets.shape[0]
    s.eval(symx_amd.EVAL_PGH)
    _, Hub0 = s.element_outputs(0, ne, nn)
    changed = s.project(1e-10, mirror)
    _, Hub = s.element_outputs(0, ne, nn)
    H = upper_blocks_to_full(Hub, nn)
    H0 = oracle.element_hessians(S)
    Ho, ch = oracle.project(H0, 1e-10, mirror)
    gpu_changed = np.any(Hub != Hub0, axis=1)
    assert changed == int(gpu_changed.sum())
    # The three translations are exact null vectors: their computed eigenvalues are
    # round-off of the size of eps for stiff elements, so whether they alone count as
    # below eps is decided by round-off (here as in the reference's Eigen solver). An
    # element with a clearly negative eigenvalue must be rebuilt by both.
    n_neg = 0
    for e in range(ne):
        lam = np.linalg.eigvalsh(H0[e])
        if lam.min() < -1e-13 * np.abs(lam).max():
            n_neg += 1
            assert gpu_changed[e] and ch[e], e
        assert tol_rel(H[e], Ho[e]) < 1e-9, e
    assert n_neg > 0


@pytest.mark.gpu
@pytest.mark.parametrize("noise,squash", [(0.05, 1.0), (0.08, 1.0), (0.15, 0.8), (0.15, 0.5)])
def test_hex27_projection_of_compressed_states_matches_oracle(gpu, oracle, noise, squash):
    """81x81 projection over a range of negative-eigenvalue counts: node noise 0.05 and 0.08
    give 0 to 10 negative eigenvalues per element (the register-tiled kernel's multisection,
    inverse iteration and rank-k update; above 8 the element goes to the full-size list
    kernel), the compressed states (z scaled by 0.8 and 0.5) 15 to 26 (all to the list
    kernel after the tiled tridiagonalisation and count). Every element matches the oracle's
    Jacobi projection (FP64 storage, clamp)."""
    s, p = hex_scene(27, noise=noise)
    S = hex_block(cells=(2, 2, 2), nn=27, noise=noise, seed=p.seed)
    x = S.x.copy()
    x[:, 2] *= squash
    s.set_dofs(x)
    S.x[...] = x
    ne = S.tets.shape[0]
    s.eval(symx_amd.EVAL_PGH)
    s.project(1e-10, False)
    _, Hub = s.element_outputs(0, ne, 27)
    H = upper_blocks_to_full(Hub, 27)
    H0 = oracle.element_hessians(S)
    Ho, _ = oracle.project(H0, 1e-10, False)
    counts = [int((np.linalg.eigvalsh(H0[e]) < -1e-13 * np.abs(H0[e]).max()).sum()) for e in range(ne)]
    assert max(counts) > 0, counts
    for e in range(ne):
        assert tol_rel(H[e], Ho[e]) < 1e-9, (e, counts[e])


@pytest.mark.gpu
@pytest.mark.parametrize("nn,mode", [(8, 1), (10, 1), (10, 2), (10, 3), (27, 1), (27, 2)])
def test_hex_newton_matches_oracle(gpu, oracle, nn, mode):
    s, p = hex_scene(nn)
    s.set_newton(symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=mode))
    S = hex_block(cells=(2, 2, 2), nn=nn, noise=p.noise, seed=p.seed)
    st = s.solve()
    res = oracle.newton(S, s.newton)
    assert st["result"] == res["result"] == 0, (st["result_name"], res["result"])
    assert st["newton_iterations"] == res["newton_iterations"]
    assert abs(st["cg_iterations"] - res["cg_iterations"]) <= max(2, st["newton_iterations"])
    assert np.abs(s.get("x") - res["x"]).max() <= 1e-7 * np.abs(res["x"] - S.X).max()


@pytest.mark.gpu
@pytest.mark.parametrize("nn", [10, 27])
def test_hex27_mixed_precision_reaches_fp64_solution(gpu, nn):
    """FP32 element Hessians, FP32 PCG and FP64 refinement reach the FP64 path's solution
    (Hex27 and Tet10 blocks, ProjectOnDemand, 1e-8 relative residual)."""
    tight = dict(residual_tolerance_rel=1e-8, cg_rel_tolerance=1e-6, projection_mode=2)
    s64, _ = hex_scene(nn, cells=(3, 3, 3))
    s64.set_newton(symx_amd.default_newton_params(precision=symx_amd.FP64, **tight))
    s32, _ = hex_scene(nn, cells=(3, 3, 3))
    s32.set_newton(symx_amd.default_newton_params(precision=symx_amd.MIXED, **tight))
    a, b = s64.solve(), s32.solve()
    assert a["result"] == 0 and b["result"] == 0, (a["result_name"], b["result_name"])
    disp = np.abs(s64.get("x") - s64.get("X")).max()
    # two solves stopped at a 1e-8 relative residual differ by up to about 2e-8 times the
    # system's condition number, relative to the displacement (bound: condition number 50)
    diff = np.abs(s32.get("x") - s64.get("x")).max()
    assert diff <= 1e-6 * disp, diff / disp


# ---- the reference's own Hex8 / Hex27 (tests/golden/*hex*, written by oracle/_ref) -----------
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("nn", [8, 27])
def test_oracle_hex_matches_reference_fem_integrator(oracle, nn):
    """E / gradient / Hessian of the reference's fem_integrator(FEM_Element::Hex8|Hex27)
    (fem_integrators.cpp:84-116, its truncated 10-digit Hex27 quadrature included) evaluated
    by its SecondOrderCompiledGlobal, on rotated, stretched and jittered cubes of sizes 1e-2..1
    with random materials: the oracle's closed form agrees to FP64 round-off."""
    G = np.load(os.path.join(GOLDEN, "hex_ref_elements.npz"))
    for i in range(G[f"E{nn}"].shape[0]):
        E, g, H = oracle.nh_hex(G[f"x{nn}"][i], G[f"X{nn}"][i], G[f"lam{nn}"][i], G[f"mu{nn}"][i])
        Er, gr, Hr = G[f"E{nn}"][i], G[f"g{nn}"][i], G[f"H{nn}"][i]
        assert abs(E - Er) <= 1e-12 * abs(Er), i
        assert tol_rel(g, gr) < 1e-11, i
        assert tol_rel(H, Hr) < 1e-11, i


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [symx_amd.FP64, symx_amd.MIXED])
def test_hex27_newton_matches_reference_solve(gpu, precision):
    """The 2^3-cell Hex27 block (per-element Young's moduli, ProjectOnDemand, 1e-6 relative
    residual) against the reference's own NewtonsMethod on the same inputs
    (hex27_ref_block2.npz): same outcome and Newton count, positions within 1e-6 of the
    displacement (FP64) and within the tolerance's reach for the mixed path."""
    R = np.load(os.path.join(GOLDEN, "hex27_ref_block2.npz"))
    s, p = hex_scene(27, cells=tuple(R["cells"]))
    s.set_newton(symx_amd.default_newton_params(precision=precision, projection_mode=int(R["projection_mode"])))
    X = s.get("X")
    st = s.solve()
    assert st["result_name"] == "Successful" and int(R["result"]) == 0
    if precision == symx_amd.FP64:
        assert st["newton_iterations"] == int(R["newton_iterations"])
    else:
        assert abs(st["newton_iterations"] - int(R["newton_iterations"])) <= 1
    disp = np.abs(R["x"] - X).max()
    # two solves stopped at a 1e-6 relative residual differ by the error that residual
    # leaves (5.7e-6 of the displacement measured for the mixed path); the FP64 path takes
    # the reference's iterates and agrees far closer
    tol = 1e-6 if precision == symx_amd.FP64 else 1e-4
    assert np.abs(s.get("x") - R["x"]).max() <= tol * disp
    s.close()


# ---- the reference's own Tet10 (tests/golden/tet10_*, written by oracle/_ref) ----------------
def test_oracle_tet10_matches_reference_fem_integrator(oracle):
    """E / gradient / Hessian of the reference's fem_integrator(FEM_Element::Tet10)
    (fem_integrators.h:56-77 shape functions, fem_integrators.cpp:60-68 4-point rule with its
    8-digit constants) evaluated by its SecondOrderCompiledGlobal, on rotated, stretched and
    jittered quadratic tetrahedra of sizes 1e-2..1 with random materials: the oracle's closed
    form agrees to FP64 round-off."""
    G = np.load(os.path.join(GOLDEN, "tet10_ref_elements.npz"))
    for i in range(G["E"].shape[0]):
        E, g, H = oracle.nh_hex(G["x"][i], G["X"][i], G["lam"][i], G["mu"][i])
        assert abs(E - G["E"][i]) <= 1e-12 * abs(G["E"][i]), i
        assert tol_rel(g, G["g"][i]) < 1e-11, i
        assert tol_rel(H, G["H"][i]) < 1e-11, i


def test_tet10_mesh_geometry():
    """The Tet10 cuboid mesh (restated in scene_np; the product's mesh equals it, checked
    through the element outputs above and the scene inputs on the GPU): every mid-edge node
    halves its edge, every element has positive volume, the elements fill the cuboid and
    every lattice node is used."""
    from scene_np import tet10_mesh

    X, T = tet10_mesh((2, 3, 1), (1.0, 1.5, 0.5))
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for m, (a, b) in enumerate(edges):
        assert np.allclose(X[T[:, 4 + m]], 0.5 * (X[T[:, a]] + X[T[:, b]]), atol=1e-15)
    V = np.linalg.det(np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=1)) / 6.0
    assert V.min() > 0 and abs(V.sum() - 0.75) <= 1e-14
    assert len(np.unique(T)) == X.shape[0]


def test_oracle_tet10_newton_matches_reference_solve(oracle):
    """One implicit step of the 2^3-cell Tet10 block (48 elements, per-element Young's moduli,
    ProjectOnDemand, 1e-6 relative residual) by the oracle against the reference's own
    NewtonsMethod on the same inputs (tet10_ref_block2.npz): same outcome and Newton count,
    CG count within one per Newton iteration, positions within 1e-6 of the displacement."""
    from symx_amd import NewtonParams

    R = np.load(os.path.join(GOLDEN, "tet10_ref_block2.npz"))
    S = hex_block(cells=tuple(R["cells"]), nn=10)
    newton = NewtonParams(0, int(R["projection_mode"]), 0, 100, 0.0, float(R["rel_tol"]), 1e-3, 1e-12, 10000, 1e-10, 0.0)
    res = oracle.newton(S, newton)
    assert res["result"] == int(R["result"]) == 0
    assert res["newton_iterations"] == int(R["newton_iterations"])
    assert abs(res["cg_iterations"] - int(R["cg_iterations"])) <= res["newton_iterations"]
    disp = np.abs(R["x"] - S.X).max()
    assert np.abs(res["x"] - R["x"]).max() <= 1e-6 * disp


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [symx_amd.FP64, symx_amd.MIXED])
def test_tet10_newton_matches_reference_solve(gpu, precision):
    """The HIP path on the 2^3-cell Tet10 block against the reference's NewtonsMethod
    (tet10_ref_block2.npz), with the bounds of the Hex27 comparison above."""
    R = np.load(os.path.join(GOLDEN, "tet10_ref_block2.npz"))
    s, p = hex_scene(10, cells=tuple(R["cells"]))
    s.set_newton(symx_amd.default_newton_params(precision=precision, projection_mode=int(R["projection_mode"])))
    X = s.get("X")
    st = s.solve()
    assert st["result_name"] == "Successful" and int(R["result"]) == 0
    if precision == symx_amd.FP64:
        assert st["newton_iterations"] == int(R["newton_iterations"])
    else:
        assert abs(st["newton_iterations"] - int(R["newton_iterations"])) <= 1
    disp = np.abs(R["x"] - X).max()
    tol = 1e-6 if precision == symx_amd.FP64 else 1e-4
    assert np.abs(s.get("x") - R["x"]).max() <= tol * disp
    s.close()
