"""CPU tests of the host side: symbolic expressions and derivatives (the inputs of the HIP
emitter), hiprtc compilation of the generated element module for gfx950, and the C ABI
symbol tables. No GPU needed."""
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import symx_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][a-zA-Z_0-9]*\s*\*?\s+\**(symx_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("symx_hip.h", "libsymx_hip.so"), ("symx_scenes.h", "libsymx_amd.so")])
def test_c_abi_exports_every_declared_symbol(header, lib):
    names = declared_functions(header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "symx_amd", "lib", lib)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the library loads
    (symx_amd.hip() if lib == "libsymx_hip.so" else symx_amd.amd())


def test_nh_element_symbolic_matches_oracle(oracle):
    rng = np.random.default_rng(1)
    lam, mu = 142857.14285714287, 35714.28571428572
    for _ in range(10):
        sc = rng.uniform(0.02, 2.0)
        X = sc * (np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float) + rng.uniform(-0.1, 0.1, 12))
        x = X + sc * rng.uniform(-0.1, 0.1, 12)
        E, g, H = symx_amd.host_tet4_element(x, X, lam, mu)
        Eo, go, Ho = oracle.nh_tet4(x, X, lam, mu)
        assert abs(E - Eo) <= 1e-12 * abs(Eo)
        assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
        assert np.abs(H - Ho).max() <= 1e-11 * np.abs(Ho).max()


def test_nh_element_matches_reference_compiled_golden():
    """Golden vectors written by the reference's own Compiled<double> (tests/golden/make_golden.py)."""
    data = json.load(open(os.path.join(GOLDEN, "nh_tet4_reference.json")))
    for case in data["cases"]:
        E, g, H = symx_amd.host_tet4_element(np.array(case["x"]), np.array(case["X"]), case["lambda"], case["mu"])
        out = np.array(case["out"])
        mine = np.concatenate([[E], g, H.ravel()])
        assert np.abs(mine - out).max() <= 1e-12 * np.abs(out).max()


def test_point_triangle_matches_sympy_golden():
    """Reference test test_diff.cpp:74-137 against its sympy-generated function (values in tests/golden)."""
    data = json.load(open(os.path.join(GOLDEN, "point_triangle_sympy.json")))
    for case in data["cases"]:
        out = symx_amd.host_point_triangle(np.array(case["in"]))
        ref = np.array(case["out"])
        assert np.abs(out - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


def test_stable_nh_is_translation_invariant_and_symmetric():
    rng = np.random.default_rng(5)
    X = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float)
    x = X + rng.uniform(-0.1, 0.1, 12)
    E, g, H = symx_amd.host_tet4_element(x, X, 1e5, 0.4, material=1)
    assert np.abs(g.reshape(4, 3).sum(0)).max() <= 1e-9 * np.abs(g).max()
    assert np.abs(H - H.T).max() == 0.0
    h = 1e-6
    gfd = np.array([(symx_amd.host_tet4_element(x + h * e, X, 1e5, 0.4, 1)[0] - symx_amd.host_tet4_element(x - h * e, X, 1e5, 0.4, 1)[0]) / (2 * h)
                    for e in np.eye(12)])
    assert np.abs(g - gfd).max() <= 1e-6 * np.abs(g).max()


@pytest.mark.parametrize("material", [0, 1])
def test_chain_rule_derivatives_match_flat_derivatives_and_oracle(oracle, material):
    """The element kernels of a Neo-Hookean block differentiate through the linear cut of the
    energy (F = Ds Dm^-1: 9 quantities for 12 DoFs): E, gradient and Hessian equal the flat
    symbolic derivatives to round-off, and the oracle's closed form (NH). The inertia and
    penalty potentials have no cut narrower than their DoFs and keep the flat form."""
    p = symx_amd.default_block_params(cells=[2, 2, 2], noise=0.15, seed=3, material=material)
    S = None
    if material == 0:
        from scene_np import tet4_block
        S = tet4_block(cells=(2, 2, 2), noise=0.15, seed=3)  # the same scene inputs (pinned by a GPU test)
    for e in (0, 5, 17, 47):
        E1, g1, H1, fac = symx_amd.host_scene_element(p, 0, e, chain_rule=True)
        E0, g0, H0, fac0 = symx_amd.host_scene_element(p, 0, e, chain_rule=False)
        assert fac and not fac0
        assert E1 == E0
        assert np.abs(g1 - g0).max() <= 1e-13 * np.abs(g0).max()
        assert np.abs(H1 - H0).max() <= 1e-13 * np.abs(H0).max()
        assert np.abs(H1 - H1.T).max() == 0.0
        if S is not None:
            t = S.tets[e]
            Eo, go, Ho = oracle.nh_tet4(S.x[t].ravel(), S.X[t].ravel(), S.lam, S.mu)
            assert abs(E1 - Eo) <= 1e-12 * abs(Eo)
            assert np.abs(g1 - go).max() <= 1e-11 * np.abs(go).max()
            assert np.abs(H1 - Ho).max() <= 1e-11 * np.abs(Ho).max()
    for pot in (1, 2):
        _, _, _, fac = symx_amd.host_scene_element(p, pot, 0, chain_rule=True)
        assert not fac


def test_generated_module_compiles_for_gfx950():
    with tempfile.TemporaryDirectory() as d:
        size, ops = symx_amd.host_jit_tet4(d)
        assert size > 1000
        assert 1000 < ops < 4000  # the reference reports 1,833 FP ops for the NH Tet4 Hessian alone
        files = os.listdir(d)
        assert any(f.endswith(".hsaco") for f in files)
        size2, _ = symx_amd.host_jit_tet4(d)  # cache hit
        assert size2 == size


def test_jit_cache_validates_cached_objects():
    """A cached code object is used only when its stored size and SHA-256 match its
    contents; a truncated or altered file is compiled again (Compilation.cpp:128-226
    validates its embedded SHA the same way)."""
    import ctypes

    lib = symx_amd.hip()
    src = (lib.symx_hip_eval_prelude().decode() +
           'extern "C" __global__ void k_cache_probe(double* x) { x[threadIdx.x] = 2.0 * x[threadIdx.x]; }\n').encode()
    with tempfile.TemporaryDirectory() as d:
        key = ctypes.create_string_buffer(80)
        cached, size = ctypes.c_int(), ctypes.c_size_t()

        def jit():
            assert lib.symx_hip_jit_compile(src, b"gfx950", d.encode(), key, ctypes.byref(cached), ctypes.byref(size)) == 0
            return cached.value, size.value

        c0, n0 = jit()
        assert c0 == 0 and n0 > 0
        assert jit() == (1, n0)
        path = os.path.join(d, key.value.decode() + ".hsaco")
        raw = open(path, "rb").read()
        assert raw[:8] == b"SYMXHSA1" and len(raw) == 8 + 16 + 64 + n0
        open(path, "wb").write(raw[:-100])  # truncated
        assert jit() == (0, n0)
        raw = bytearray(open(path, "rb").read())
        raw[-1] ^= 0xFF  # one byte altered
        open(path, "wb").write(bytes(raw))
        assert jit() == (0, n0)
        assert jit() == (1, n0)


def test_reference_element_code_compiles_and_matches_reference_outputs():
    """The Compiled<double> element code of the reference's NeoHookean example (the same body
    as oracle/ref_driver.cpp run_element: Workspace, fem Jacobians, DiffCache gradients,
    collect_scalars, Compiled<double>(outs, name, folder), set, View<double> r = c.run())
    compiled against this front end (tests/cpp/api_tests.cpp, mode `element`) and run on
    the host: outputs equal the reference's own (nh_tet4_reference.json) to round-off."""
    import json
    import subprocess

    exe = os.path.join(ROOT, "symx_amd", "lib", "api_tests")
    cases = json.load(open(os.path.join(ROOT, "tests", "golden", "nh_tet4_reference.json")))["cases"]
    with tempfile.TemporaryDirectory() as d:
        env = dict(os.environ, SYMX_AMD_CODEGEN_DIR=d)
        for c in cases:
            args = ["%.17g" % v for v in [c["lambda"], c["mu"], *c["x"], *c["X"]]]
            out = subprocess.run([exe, "element", *args], capture_output=True, text=True, check=True, env=env).stdout.split()
            got, ref = np.array([float(v) for v in out]), np.array(c["out"])
            assert got.shape == (157,)
            assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
            assert np.abs(got[1:13] - ref[1:13]).max() <= 1e-11 * np.abs(ref[1:13]).max()
            assert np.abs(got[13:] - ref[13:]).max() <= 1e-11 * np.abs(ref[13:]).max()
