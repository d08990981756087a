"""GPU parity tests: the HIP path (through include/symx_hip.h and include/symx_scenes.h)
against the C oracle and the reference's own outputs (tests/golden). Integer/index work
is compared bitwise; FP64 results within the tolerances written in each test."""
import os

import numpy as np
import pytest

import symx_amd
from oracle_py import scene_arrays
from scene_np import tet4_block

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def make_scene(cells=(4, 4, 4), noise=None, precision=symx_amd.FP64, **kw):
    h = 1.0 / cells[0]
    prm = symx_amd.default_block_params(cells=list(cells), noise=0.2 * h if noise is None else noise, **kw)
    newton = symx_amd.default_newton_params(precision=precision)
    s = symx_amd.Tet4Block(prm, newton)
    s.set_dofs(s.get("x"))  # direct evaluations read the device DoFs
    return s


def test_scene_inputs_match_numpy_restatement(gpu):
    s = make_scene((5, 4, 3), noise=0.04, seed=9)
    S = tet4_block(cells=(5, 4, 3), noise=0.04, seed=9)
    assert np.array_equal(s.connectivity(), S.tets)
    assert np.array_equal(s.get("X"), S.X)
    assert np.array_equal(s.get("x"), S.x)
    assert np.allclose(s.get("mass"), S.mass, rtol=1e-15, atol=0)


def test_element_outputs_match_oracle(gpu, oracle):
    s = make_scene((4, 4, 4))
    S = scene_arrays(s)
    E = s.eval(symx_amd.EVAL_PGH)
    Eo, go = oracle.energy(S)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    grad = s.download(symx_amd.VEC_GRAD)
    assert np.abs(grad - go).max() <= 1e-11 * np.abs(go).max()
    G, H = s.element_outputs(0, s.n_elements, 4)
    He = oracle.element_hessians(S)
    # unpack upper 3x3 blocks (a <= b) to full 12x12
    Hf = np.zeros((s.n_elements, 12, 12))
    ub = 0
    for a in range(4):
        for b in range(a, 4):
            blk = H[:, 9 * ub:9 * ub + 9].reshape(-1, 3, 3)
            Hf[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            Hf[:, 3 * b:3 * b + 3, 3 * a:3 * a + 3] = np.transpose(blk, (0, 2, 1))
            ub += 1
    scale = np.abs(He).max(axis=(1, 2))
    err = np.abs(Hf - He).max(axis=(1, 2)) / scale
    assert err.max() <= 1e-12


def test_energy_only_and_gradient_modes_agree(gpu):
    s = make_scene((3, 3, 3))
    E_pgh = s.eval(symx_amd.EVAL_PGH)
    g_pgh = s.download(symx_amd.VEC_GRAD)
    E_pg = s.eval(symx_amd.EVAL_PG)
    g_pg = s.download(symx_amd.VEC_GRAD)
    E_p = s.eval(symx_amd.EVAL_P)
    assert E_p == E_pg == E_pgh
    assert np.array_equal(g_pg, g_pgh)


def test_evaluation_and_assembly_are_bitwise_repeatable(gpu):
    s = make_scene((6, 6, 6))
    E1 = s.eval(symx_amd.EVAL_PGH)
    g1 = s.download(symx_amd.VEC_GRAD)
    s.project()
    s.assemble(symx_amd.FP64)
    _, _, v1 = s.export_bsr()
    E2 = s.eval(symx_amd.EVAL_PGH)
    g2 = s.download(symx_amd.VEC_GRAD)
    s.project()
    s.assemble(symx_amd.FP64)
    _, _, v2 = s.export_bsr()
    assert E1 == E2 and np.array_equal(g1, g2) and np.array_equal(v1, v2)


def assemble_np(S, He, nv):
    """Dense sum of element blocks (ElementHessians.cpp:224-256) + inertia + penalty diagonals."""
    n = 3 * nv
    A = np.zeros((n, n))
    idx = (3 * S.tets[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 12)
    for e in range(S.tets.shape[0]):
        A[np.ix_(idx[e], idx[e])] += He[e]
    if S.dt > 0:
        A[np.arange(n), np.arange(n)] += np.repeat(S.mass, 3) / S.dt ** 2
    for v in S.bc_vertex:
        A[3 * v + np.arange(3), 3 * v + np.arange(3)] += S.k
    return A


@pytest.mark.parametrize("noise,seed", [(0.3 / 3, 1), (0.0, 1), (0.12, 2)])
@pytest.mark.parametrize("mirror", [False, True])
def test_projection_and_assembly_match_oracle(gpu, oracle, mirror, noise, seed):
    """noise 0: every element at rest, with the three rotations as an exactly degenerate
    null triple of the reduced 9x9 (inverse iteration must separate them). noise 0.1 and
    0.12 (no element inverted): 2 and 5 of the 162 elements have five eigenvalues below
    eps in the reduced 9x9, which the register path hands to the workgroup-per-matrix
    projection."""
    s = make_scene((3, 3, 3), noise=noise, seed=seed)
    S = scene_arrays(s)
    s.eval(symx_amd.EVAL_PGH)
    n_changed = s.project(1e-10, mirror)
    s.assemble(symx_amd.FP64)
    rp, cols, vals = s.export_bsr()
    He = oracle.element_hessians(S)
    Hp, changed = oracle.project(He, 1e-10, mirror)
    A = assemble_np(S, Hp, s.n_vertices)
    B = np.zeros_like(A)
    for r in range(s.n_vertices):
        for k in range(rp[r], rp[r + 1]):
            B[3 * r:3 * r + 3, 3 * cols[k]:3 * cols[k] + 3] = vals[k]
    assert np.abs(A - B).max() <= 1e-11 * np.abs(A).max()
    # block pattern: exactly the blocks touched by elements (plus diagonals)
    pattern = np.zeros((s.n_vertices, s.n_vertices), bool)
    for t in S.tets:
        pattern[np.ix_(t, t)] = True
    pattern[np.arange(s.n_vertices), np.arange(s.n_vertices)] = True
    got = np.zeros_like(pattern)
    for r in range(s.n_vertices):
        got[r, cols[rp[r]:rp[r + 1]]] = True
    assert np.array_equal(pattern, got)
    assert n_changed >= 1


@pytest.mark.parametrize("n", [3, 6, 9, 12, 24, 30, 81])
def test_dense_projection_matches_oracle(gpu, oracle, n):
    """n <= 12: register Jacobi; n >= 24: the LDS kernel (tridiagonalisation, Sturm
    multisection, inverse iteration, rank-k update). Besides random matrices: repeated
    eigenvalues (an orthogonal similarity of a spectrum with triples), -3 I (every
    eigenvalue below eps and all equal) and the zero matrix."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(n)
    M = rng.normal(size=(200, n, n)) * 10.0 ** rng.uniform(-2, 4, size=(200, 1, 1))
    A = M + np.transpose(M, (0, 2, 1))
    A[:20] = A[:20] @ np.transpose(A[:20], (0, 2, 1)) + np.eye(n)  # some already PD
    for i in range(20, 30):
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        lam = np.repeat(rng.normal(size=n // 3) * 10.0 ** rng.uniform(-2, 4), 3)
        A[i] = (Q * lam) @ Q.T
        A[i] = 0.5 * (A[i] + A[i].T)
    A[30] = -3.0 * np.eye(n)
    A[31] = 0.0
    for mirror in (False, True):
        ref, ch = oracle.project(A, 1e-10, mirror)
        d = torch.tensor(A, dtype=torch.float64, device="cuda")
        changed = torch.zeros(A.shape[0], dtype=torch.uint8, device="cuda")
        ctx = symx_amd.Context()
        symx_amd.project_dense(ctx, d.data_ptr(), A.shape[0], n, 1e-10, mirror, changed.data_ptr())
        out = d.cpu().numpy()
        ctx.close()
        scale = np.maximum(np.abs(A).max(axis=(1, 2)), 1e-300)
        assert (np.abs(out - ref).max(axis=(1, 2)) / scale).max() <= 1e-12
        assert np.array_equal(changed.cpu().numpy().astype(bool), ch)
        assert np.array_equal(out[:20], A[:20])  # PD matrices untouched bitwise


def test_pcg_matches_oracle_iteration_by_iteration(gpu, oracle):
    s = make_scene((5, 5, 5))
    s.eval(symx_amd.EVAL_PGH)
    s.project()
    s.assemble(symx_amd.FP64)
    rp, cols, vals = s.export_bsr()
    g = s.download(symx_amd.VEC_GRAD)
    s.upload(symx_amd.VEC_RHS, -g)
    for rel in (1e-3, 1e-8):
        info = s.linear_solve(symx_amd.FP64, 1e-14, rel)
        du = s.download(symx_amd.VEC_DU)
        x, oinfo = oracle.pcg(rp, cols, vals, -g, 1e-14, rel)
        assert info["converged"] and oinfo["converged"]
        assert abs(info["n_iterations"] - oinfo["n_iterations"]) <= 1
        assert np.abs(du - x).max() <= 1e-8 * np.abs(x).max()
    # zero rhs exits at once with du = 0
    s.upload(symx_amd.VEC_RHS, np.zeros_like(g))
    info = s.linear_solve(symx_amd.FP64, 1e-12, 1e-3)
    assert info["converged"] and info["n_iterations"] == 0 and not s.download(symx_amd.VEC_DU).any()


def test_mixed_precision_solve_meets_the_same_stopping_rule(gpu, oracle):
    s = make_scene((6, 6, 6))
    s.eval(symx_amd.EVAL_PGH)
    s.project()
    s.assemble(symx_amd.FP64)
    rp, cols, vals = s.export_bsr()
    s.assemble(symx_amd.MIXED)
    g = s.download(symx_amd.VEC_GRAD)
    s.upload(symx_amd.VEC_RHS, -g)
    for rel in (1e-3, 1e-6, 1e-9):
        info = s.linear_solve(symx_amd.MIXED, 1e-14, rel)
        du = s.download(symx_amd.VEC_DU)
        assert info["converged"], info
        # true FP64 residual of the returned solution against the FP64 matrix
        Adu = np.zeros_like(du)
        for r in range(s.n_vertices):
            for k in range(rp[r], rp[r + 1]):
                Adu[3 * r:3 * r + 3] += vals[k] @ du[3 * cols[k]:3 * cols[k] + 3]
        err = np.linalg.norm(-g - Adu) / np.linalg.norm(g)
        assert err < rel * 1.05 + 1e-7, (rel, err)


@pytest.mark.parametrize("name", ["block6", "block6_tight", "block20", "block6_pdn", "block6_ppn"])
def test_newton_matches_reference_fixture(gpu, oracle, name):
    f = np.load(os.path.join(GOLDEN, f"newton_ref_{name}.npz"))
    cells = tuple(int(c) for c in f["cells"])
    prm = symx_amd.default_block_params(cells=list(cells), noise=float(f["noise"]), seed=int(f["seed"]))
    mode = int(f["projection_mode"]) if "projection_mode" in f.files else 1
    newton = symx_amd.default_newton_params(residual_tolerance_rel=float(f["rel_tol"]), cg_rel_tolerance=float(f["cg_rel"]),
                                            projection_mode=mode)
    s = symx_amd.Tet4Block(prm, newton)
    S = scene_arrays(s)
    st = s.solve()
    assert st["result_name"] == "Successful"
    assert st["newton_iterations"] == int(f["newton_iterations"])
    assert abs(st["cg_iterations"] - int(f["cg_iterations"])) <= st["newton_iterations"]
    x_ref = f["x"]
    x = s.get("x")
    disp = np.abs(x_ref - S.X).max()
    tol = 1e-8 if name == "block6_tight" else 1e-6
    assert np.abs(x - x_ref).max() <= tol * disp
    E_ref = oracle.energy(S, x_ref, with_grad=False)
    E = oracle.energy(S, x, with_grad=False)
    assert abs(E - E_ref) <= 1e-10 * abs(E_ref)


def test_plain_newton_failure_matches_reference_fixture(gpu):
    """Plain Newton (no projection) on the indefinite block6 problem: the reference stops
    at the first linear solve with LinearSystemSolveFailure; so must the device path."""
    f = np.load(os.path.join(GOLDEN, "newton_ref_block6_newton.npz"))
    assert int(f["projection_mode"]) == 0
    cells = tuple(int(c) for c in f["cells"])
    prm = symx_amd.default_block_params(cells=list(cells), noise=float(f["noise"]), seed=int(f["seed"]))
    newton = symx_amd.default_newton_params(residual_tolerance_rel=float(f["rel_tol"]), cg_rel_tolerance=float(f["cg_rel"]),
                                            projection_mode=0, precision=symx_amd.FP64)
    s = symx_amd.Tet4Block(prm, newton)
    st = s.solve()
    assert st["result"] == int(f["result"]) and st["result_name"] == "LinearSystemSolveFailure"
    assert st["newton_iterations"] == int(f["newton_iterations"])


def test_newton_fp64_matches_oracle_trajectory(gpu, oracle):
    s = make_scene((6, 6, 6))
    S = scene_arrays(s)
    st = s.solve()
    res = oracle.newton(S, s.newton)
    assert st["result_name"] == "Successful" and res["result"] == 0
    assert st["newton_iterations"] == res["newton_iterations"]
    assert abs(st["cg_iterations"] - res["cg_iterations"]) <= st["newton_iterations"]
    assert np.abs(s.get("x") - res["x"]).max() <= 1e-7 * np.abs(res["x"] - S.X).max()


@pytest.mark.parametrize("mode", [0, 2, 3])
@pytest.mark.parametrize("dt", [0.01, 0.1])
def test_projection_modes_match_oracle(gpu, oracle, mode, dt):
    """Newton (no projection; fails on the indefinite first system like the oracle),
    ProjectOnDemand and Progressive projection (NewtonsMethod.cpp:254-386) against the
    oracle's restatement of the same retry logic."""
    s = make_scene((6, 6, 6), seed=11, dt=dt)
    s.set_newton(symx_amd.default_newton_params(projection_mode=mode))
    S = scene_arrays(s)
    st = s.solve()
    res = oracle.newton(S, s.newton)
    assert st["result"] == res["result"], (st["result_name"], res["result"])
    assert st["newton_iterations"] == res["newton_iterations"]
    assert abs(st["cg_iterations"] - res["cg_iterations"]) <= max(2, st["newton_iterations"])
    if res["result"] == 0:
        assert np.abs(s.get("x") - res["x"]).max() <= 1e-7 * np.abs(res["x"] - S.X).max()


def test_mixed_newton_reaches_fp64_solution(gpu, oracle):
    tight = dict(residual_tolerance_rel=1e-8, cg_rel_tolerance=1e-6)
    s64 = make_scene((6, 6, 6))
    s64.set_newton(symx_amd.default_newton_params(precision=symx_amd.FP64, **tight))
    s32 = make_scene((6, 6, 6))
    s32.set_newton(symx_amd.default_newton_params(precision=symx_amd.MIXED, **tight))
    S = scene_arrays(s64)
    a = s64.solve()
    b = s32.solve()
    assert a["result_name"] == "Successful" and b["result_name"] == "Successful"
    x64, x32 = s64.get("x"), s32.get("x")
    disp = np.abs(x64 - S.X).max()
    assert np.abs(x32 - x64).max() <= 1e-7 * disp
    assert abs(oracle.energy(S, x32, False) - oracle.energy(S, x64, False)) <= 1e-10 * abs(oracle.energy(S, x64, False))


def test_time_stepping_updates_state(gpu):
    s = make_scene((3, 3, 3))
    x0 = s.get("x")
    st = s.step()
    assert st["result_name"] == "Successful"
    v = s.get("v")
    assert np.allclose(v, (s.get("x") - x0) / s.params.dt)
    assert np.array_equal(s.get("x0"), s.get("x"))


@pytest.mark.parametrize("cells", [32, 64])
@pytest.mark.parametrize("precision", [symx_amd.FP64, symx_amd.MIXED])
def test_unstructured_numbering_gives_the_same_step(gpu, precision, cells):
    """A block with its vertices renumbered at random: column offsets no longer fit 16 bits
    (int32 column ids) and rows of different length interleave (SELL rows sorted by length
    inside windows). One implicit step must match the structured numbering's vertex for
    vertex, to the Newton tolerance. 32^3 cells (35,937 vertices) run the cache-resident
    PCG; 64^3 cells (274,625 vertices, a 148 MB FP32 / 297 MB FP64 matrix) run the
    streamed one with the separate direction pass."""
    out, mats = [], []
    for renumber in (0, 5):
        prm = symx_amd.default_block_params(cells=[cells] * 3, noise=0.2 / cells, seed=3, renumber_seed=renumber)
        s = symx_amd.Tet4Block(prm, symx_amd.default_newton_params(precision=precision))
        X = s.get("X")
        order = np.lexsort((X[:, 2], X[:, 1], X[:, 0]))  # grid order, independent of numbering
        rank = np.empty_like(order)
        rank[order] = np.arange(order.size)
        if precision == symx_amd.FP64:
            # assembled matrix in grid numbering: same contributions in the same order per
            # slot, so equal bit for bit between the two numberings
            s.set_dofs(s.get("x"))
            s.eval(symx_amd.EVAL_PGH)
            s.project()
            s.assemble(symx_amd.FP64)
            rp, cols, vals = s.export_bsr()
            rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
            key = rank[rows].astype(np.int64) * order.size + rank[cols]
            srt = np.argsort(key)
            mats.append((key[srt], vals[srt]))
        x_start = s.get("x")
        st = s.step()
        x = s.get("x")
        out.append((st, x[order], x_start[order]))
        s.close()
    if mats:
        assert np.array_equal(mats[0][0], mats[1][0])
        assert np.array_equal(mats[0][1], mats[1][1])
    (st0, x0, xs0), (st1, x1, xs1) = out
    assert np.array_equal(xs0, xs1)  # same geometry after the renumbering
    assert st0["result_name"] == st1["result_name"] == "Successful"
    assert abs(st0["newton_iterations"] - st1["newton_iterations"]) <= 1
    moved = np.abs(x0 - xs0).max()
    assert moved > 0
    assert np.abs(x1 - x0).max() <= 1e-5 * moved, (np.abs(x1 - x0).max(), moved)


@pytest.mark.parametrize("precision", [symx_amd.FP64, symx_amd.MIXED])
def test_direct_llt_matches_reference_fixture(gpu, oracle, precision):
    """LinearSolver::DirectLLT (NewtonsMethod.cpp:395-418; the reference factors the
    projected FP64 matrix with Eigen's SimplicialLLT) against the reference's own DirectLLT
    solve of the block6 case to 1e-8 (newton_ref_block6_llt.npz). The device solves the same
    FP64 matrix to round-off (in both precision settings), so the Newton iterates are the
    reference's."""
    f = np.load(os.path.join(GOLDEN, "newton_ref_block6_llt.npz"))
    assert bool(f["llt"])
    cells = tuple(int(c) for c in f["cells"])
    prm = symx_amd.default_block_params(cells=list(cells), noise=float(f["noise"]), seed=int(f["seed"]))
    newton = symx_amd.default_newton_params(residual_tolerance_rel=float(f["rel_tol"]), cg_rel_tolerance=float(f["cg_rel"]),
                                            precision=precision, linear_solver=1)
    s = symx_amd.Tet4Block(prm, newton)
    S = scene_arrays(s)
    st = s.solve()
    assert st["result_name"] == "Successful" and int(f["result"]) == 0
    assert st["newton_iterations"] == int(f["newton_iterations"])
    disp = np.abs(f["x"] - S.X).max()
    assert np.abs(s.get("x") - f["x"]).max() <= 1e-9 * disp
    s.close()
