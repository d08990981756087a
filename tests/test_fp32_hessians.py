"""FP32 element-Hessian storage of the mixed path (SYMX_HIP_EVAL_PGH_F32).

The reference stores the block matrix as SYMX_HESS_STORAGE_FLOAT (ElementHessians.cpp:
236-249 casts each element block when it is added); here the cast moves to the element
Hessians themselves. Energies and gradients stay FP64. Where the element kernel
differentiates through the linear cut of the energy (the deformation gradient of a Tet4
element), the material law's second derivatives and the coefficients dF/dx are computed in
FP64 and rounded once, and the contraction H = C^T A C runs in FP32 (nine products per
entry); otherwise the FP64 values are rounded once. Projection reads the floats, works in
FP64 and rounds its result once; assembly sums them in FP64. Tolerances follow from a few
float roundings (2^-24 = 6e-8 relative) of the largest terms."""
import numpy as np
import pytest

import symx_amd
from oracle_py import scene_arrays
from test_gpu_parity import assemble_np, make_scene

pytestmark = pytest.mark.gpu
F32_ULP = 2.0 ** -23


def unpack_ub(H, nb):
    n = 3 * nb
    Hf = np.zeros((H.shape[0], n, n))
    ub = 0
    for a in range(nb):
        for b in range(a, nb):
            blk = H[:, 9 * ub:9 * ub + 9].reshape(-1, 3, 3)
            Hf[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            Hf[:, 3 * b:3 * b + 3, 3 * a:3 * a + 3] = np.transpose(blk, (0, 2, 1))
            ub += 1
    return Hf


def test_f32_element_hessians_are_the_fp64_ones_rounded(gpu, oracle):
    s = make_scene((4, 4, 4))
    S = scene_arrays(s)
    E64 = s.eval(symx_amd.EVAL_PGH)
    g64 = s.download(symx_amd.VEC_GRAD)
    _, H64 = s.element_outputs(0, s.n_elements, 4)
    E32 = s.eval(symx_amd.EVAL_PGH_F32)
    g32 = s.download(symx_amd.VEC_GRAD)
    _, H32 = s.element_outputs(0, s.n_elements, 4)
    # energies and gradients stay FP64
    assert E32 == E64 and np.array_equal(g32, g64)
    assert np.array_equal(H32, H32.astype(np.float32).astype(np.float64))  # stored as floats
    # a few float roundings of the largest term (measured 1.3e-7 of max|H|)
    scale64 = np.abs(H64).max(axis=1)
    err = (np.abs(H32 - H64).max(axis=1) / scale64).max()
    assert err <= 1e-6, err
    He = oracle.element_hessians(S)
    scale = np.abs(He).max(axis=(1, 2))
    err_o = (np.abs(unpack_ub(H32, 4) - He).max(axis=(1, 2)) / scale).max()
    assert err_o <= 1e-6, err_o
    # translation invariance survives the FP32 contraction at the storage's resolution: the
    # block-row sums stay below the projection's threshold for the translation shortcut
    # (1e-6 max|H|; measured 1.2e-7)
    rows = np.abs(unpack_ub(H32, 4).reshape(-1, 12, 4, 3).sum(axis=2)).max(axis=(1, 2)) / scale
    assert rows.max() <= 1e-6, rows.max()


@pytest.mark.parametrize("noise,seed", [(0.3 / 3, 1), (0.0, 1), (0.12, 2)])
@pytest.mark.parametrize("mirror", [False, True])
def test_f32_projection_and_assembly_match_oracle(gpu, oracle, mirror, noise, seed):
    """Projection of the FP32-stored element Hessians against the oracle's FP64 projection of
    the same (float) values, and the FP32 block matrix against the dense FP64 sum of the
    projected blocks; both within a few float roundings. Cases as in the FP64 test (rest
    state with the degenerate rotation triple; elements handed to the workgroup fallback)."""
    s = make_scene((3, 3, 3), noise=noise, seed=seed)
    S = scene_arrays(s)
    s.eval(symx_amd.EVAL_PGH_F32)
    _, H32 = s.element_outputs(0, s.n_elements, 4)
    s.project(1e-10, mirror)
    _, P32 = s.element_outputs(0, s.n_elements, 4)
    Hp, _ = oracle.project(unpack_ub(H32, 4), 1e-10, mirror)
    scale = np.abs(Hp).max(axis=(1, 2))
    # The stored floats are translation invariant only to their rounding (block-row sums
    # ~1e-7 max|H|): the device maps the translations to f(0) and drops that coupling, the
    # oracle projects the perturbed null space. The device also leaves eigenvalues in
    # [-12 * 2^-24 max|H|, eps) alone (inside the storage rounding). Both differ from the
    # oracle at that level, a few float ulps of max|H|.
    assert (np.abs(unpack_ub(P32, 4) - Hp).max(axis=(1, 2)) / scale).max() <= 16 * F32_ULP
    s.assemble(symx_amd.MIXED)
    rp, cols, vals = s.export_bsr()
    A = assemble_np(S, unpack_ub(P32, 4), s.n_vertices)
    B = np.zeros_like(A)
    for r in range(s.n_vertices):
        for k in range(rp[r], rp[r + 1]):
            B[3 * r:3 * r + 3, 3 * cols[k]:3 * cols[k] + 3] = vals[k]
    # FP64 sums of the float blocks, rounded once to the FP32 matrix (the summation orders of
    # numpy and the device differ by FP64 round-off, visible where a sum cancels)
    assert np.all(np.abs(A - B) <= F32_ULP * np.abs(A) + 1e-14 * np.abs(A).max())
    # the FP64 assembly of the same float element Hessians: the dense sum up to the float
    # rounding of the inertia and penalty blocks (assemble_np adds their FP64 values)
    s.assemble(symx_amd.FP64)
    _, _, v64 = s.export_bsr()
    B64 = np.zeros_like(A)
    for r in range(s.n_vertices):
        for k in range(rp[r], rp[r + 1]):
            B64[3 * r:3 * r + 3, 3 * cols[k]:3 * cols[k] + 3] = v64[k]
    assert np.all(np.abs(A - B64) <= F32_ULP * np.abs(A) + 1e-14 * np.abs(A).max())


def test_f32_structured_hex27_kernel(gpu):
    p = symx_amd.default_hex_params(cells=[2, 2, 2], nodes_per_element=27, noise=0.05, seed=5)
    s = symx_amd.HexBlock(p, symx_amd.default_newton_params(precision=symx_amd.MIXED, projection_mode=2))
    s.set_dofs(s.get("x"))
    E64 = s.eval(symx_amd.EVAL_PGH)
    _, H64 = s.element_outputs(0, s.n_elements, 27)
    E32 = s.eval(symx_amd.EVAL_PGH_F32)
    _, H32 = s.element_outputs(0, s.n_elements, 27)
    assert E32 == E64
    r = H64.astype(np.float32).astype(np.float64)
    assert np.all(np.abs(H32 - r) <= F32_ULP * np.abs(r))
    # projection of the FP32 storage (LDS kernel) against the FP64 one of the same values
    s.project(1e-10, False)
    _, P32 = s.element_outputs(0, s.n_elements, 27)
    ctx = symx_amd.Context()
    torch = pytest.importorskip("torch")
    full = unpack_ub(H32, 27)
    d = torch.tensor(full, dtype=torch.float64, device="cuda")
    symx_amd.project_dense(ctx, d.data_ptr(), full.shape[0], 81, 1e-10, False)
    ref = d.cpu().numpy()
    ctx.close()
    scale = np.abs(ref).max(axis=(1, 2))
    # eigenvalues in [-81 * 2^-24 max|H|, eps) are left alone with FP32 storage (40 float ulps)
    assert (np.abs(unpack_ub(P32, 27) - ref).max(axis=(1, 2)) / scale).max() <= 64 * F32_ULP
    s.close()


@pytest.mark.parametrize("cells", [(6, 6, 6)])
def test_mixed_newton_with_f32_hessians_matches_fp64(gpu, cells):
    """The mixed solve (FP32 element Hessians, FP32 matrix and PCG, FP64 refinement and line
    search) against the FP64 solve of the same step: same outcome, positions within the
    1e-6 residual tolerance's reach."""
    out = {}
    for prec in (symx_amd.FP64, symx_amd.MIXED):
        s = make_scene(cells, precision=prec)
        st = s.solve()
        out[prec] = (st, s.get("x"), s.get("X"))
        s.close()
    (s64, x64, X), (s32, x32, _) = out[symx_amd.FP64], out[symx_amd.MIXED]
    assert s64["result_name"] == s32["result_name"] == "Successful"
    assert abs(s64["newton_iterations"] - s32["newton_iterations"]) <= 1
    assert np.abs(x32 - x64).max() <= 1e-6 * np.abs(x64 - X).max() + 1e-12
