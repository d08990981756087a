"""Writes the golden fixtures of tests/golden/ by running the REAL reference built from
its sources by oracle/Makefile (`make -C oracle ref`, needs /root/reference). Run in the
build container:  python tests/golden/make_golden.py

  nh_tet4_reference.json     E, gradient, Hessian of the Neo-Hookean Tet4 element from the
                             reference's Compiled<double> (examples/NeoHookean.cpp:12-117)
  point_triangle_sympy.json  the reference's sympy-generated point-triangle derivatives
                             (tests/sympy_derivatives.cpp, used by tests/test_diff.cpp:74-137)
  newton_ref_*.npz           final positions and iteration counts of the reference
                             NewtonsMethod on Tet4 block problems (n_threads = 1)
  adjoint_ref_*.npz          dL/dlog_mu of the quasistatic cantilever from the reference's
                             own pieces (ref_driver adjoint) at the oracle's equilibrium
                             (`python tests/golden/make_golden.py adjoint` writes only these)
  newton_ref_block6_{newton,pdn,ppn}.npz  the block6 case under the other projection
                             modes: Newton, ProjectOnDemand, Progressive
                             (`python tests/golden/make_golden.py modes`)
  hex_ref_elements.npz       E, gradient, Hessian of Neo-Hookean Hex8 and Hex27 elements from
                             the reference's fem_integrator (ref_driver hex_element)
  hex27_ref_block2.npz       the reference NewtonsMethod (ProjectOnDemand) on a 2^3-cell Hex27
                             block with per-element Young's moduli (ref_driver solve_hex)
                             (`python tests/golden/make_golden.py hex` writes these two)
  newton_ref_block6_llt.npz  the block6 case with LinearSolver::DirectLLT to 1e-8 relative
                             (`python tests/golden/make_golden.py llt`)
  tet10_ref_elements.npz     E, gradient, Hessian of Neo-Hookean Tet10 elements from the
                             reference's fem_integrator (ref_driver hex_element, 10 nodes)
  tet10_ref_block2.npz       the reference NewtonsMethod (ProjectOnDemand) on a 2^3-cell Tet10
                             block (48 elements) with per-element Young's moduli
                             (`python tests/golden/make_golden.py tet10` writes these two)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
ENV = {k: v for k, v in os.environ.items() if k != "CXX"}


def fmt(vals):
    return ["%.17g" % v for v in vals]


def nh_cases():
    rng = np.random.default_rng(20261203)
    cases = []
    E_, nu = 1e6, 0.3
    lam, mu = E_ * nu / ((1 + nu) * (1 - 2 * nu)), E_ / (2 * (1 + nu))
    inputs = [(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float), np.array([0, 0, 0, .9, 0, 0, 0, .8, 0, 0, 0, .7], float), lam, mu)]
    for _ in range(7):
        X = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], float) * rng.uniform(0.05, 1.0) + rng.uniform(-0.02, 0.02, 12)
        x = X + rng.uniform(-0.1, 0.1, 12) * np.abs(X).max()
        inputs.append((x, X, 142857.14285714287, 35714.28571428572))
    for x, X, l, m in inputs:
        out = subprocess.run([os.path.join(REF, "ref_driver"), "element"] + fmt([l, m, *x, *X]), capture_output=True, text=True,
                             check=True, env=ENV).stdout.split()
        cases.append({"x": x.tolist(), "X": X.tolist(), "lambda": l, "mu": m, "out": [float(v) for v in out]})
    json.dump({"source": "reference Compiled<double> via oracle/ref_driver element", "cases": cases},
              open(os.path.join(HERE, "nh_tet4_reference.json"), "w"))


def point_triangle_cases():
    rng = np.random.default_rng(74)
    cases = []
    for _ in range(8):
        inp = np.concatenate([rng.uniform(-1, 1, 12), [rng.choice([-1.0, 1.0])]])
        out = subprocess.run([os.path.join(REF, "sympy_golden")] + fmt(inp), capture_output=True, text=True, check=True).stdout.split()
        cases.append({"in": inp.tolist(), "out": [float(v) for v in out]})
    json.dump({"source": "reference tests/sympy_derivatives.cpp", "cases": cases}, open(os.path.join(HERE, "point_triangle_sympy.json"), "w"))


def newton_case(name, cells, noise, seed, rel_tol, cg_rel, threads=1, mode=1, llt=False):
    from ref_io import read_solve_output, write_solve_input
    from scene_np import tet4_block
    from symx_amd import NewtonParams

    S = tet4_block(cells=cells, noise=noise, seed=seed)
    newton = NewtonParams(0, mode, 0, 100, 0.0, rel_tol, cg_rel, 1e-12, 10000, 1e-10, 0.0)
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_solve_input(inp, S, newton, threads)
        subprocess.run([os.path.join(REF, "ref_driver"), "solve", inp, outp] + (["llt"] if llt else []), check=True, env=ENV)
        res = read_solve_output(outp, S.X.shape[0])
    np.savez_compressed(os.path.join(HERE, f"newton_ref_{name}.npz"), x=res["x"], cells=np.array(cells), noise=noise, seed=seed, llt=llt,
                        rel_tol=rel_tol, cg_rel=cg_rel, result=res["result"], newton_iterations=res["newton_iterations"],
                        cg_iterations=res["cg_iterations"], ls_bt_iterations=res["ls_bt_iterations"], energy=res["energy"],
                        seconds=res["seconds"], threads=threads, projection_mode=mode)
    print(name, {k: v for k, v in res.items() if k != "x"})


def adjoint_case(name, cells):
    """Reference-composed adjoint of the cantilever at the oracle's equilibrium."""
    from adjoint_harness import log_mu_field, oracle_equilibrium, scene_np
    from oracle_py import Oracle
    from ref_io import read_adjoint_output, write_adjoint_input

    O = Oracle()
    S = scene_np(cells)
    ne = S.tets.shape[0]
    log_mu_t = log_mu_field(ne, 2)
    x_t = oracle_equilibrium(O, S, log_mu_t)
    log_mu = log_mu_field(ne, 1)
    x = oracle_equilibrium(O, S, log_mu)
    S.x[...] = x.reshape(S.x.shape)
    dL_dx = x - x_t
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_adjoint_input(inp, S, log_mu, dL_dx)
        subprocess.run([os.path.join(REF, "ref_driver"), "adjoint", inp, outp], check=True, env=ENV)
        res = read_adjoint_output(outp, S.X.shape[0], ne)
    assert res["ok"] == 1
    np.savez_compressed(os.path.join(HERE, f"adjoint_ref_{name}.npz"), cells=np.array(cells), log_mu=log_mu, x=x, x_target=x_t,
                        dL_dlog_mu=res["dL_dlog_mu"], lam=res["lambda"], energy=res["energy"])
    print(name, "adjoint", "max|dL/dp|", np.abs(res["dL_dlog_mu"]).max())


def hex_cases():
    """Perturbed Hex8 / Hex27 elements: scaled, rotated-and-stretched and noisy reference
    cubes (HEX_NODES order), random materials."""
    from ref_io import read_hex_element_output, write_hex_element_input
    from scene_np import HEX_NODES

    rng = np.random.default_rng(20261204)
    out = {}
    for nn in (8, 27):
        base = np.array(HEX_NODES[:nn], float) / 2.0  # unit cube, nodes at 0, 0.5, 1
        if nn == 8:
            base = np.array([[a // 2, b // 2, c // 2] for (a, b, c) in HEX_NODES[:8]], float)
        cases = []
        for i in range(6):
            h = 10.0 ** rng.uniform(-2, 0)
            X = base * h + rng.uniform(-0.05, 0.05, base.shape) * h
            A = np.eye(3) + rng.uniform(-0.15, 0.15, (3, 3))
            x = X @ A.T + rng.uniform(-0.03, 0.03, X.shape) * h
            E_, nu = 10.0 ** rng.uniform(4, 6), rng.uniform(0.2, 0.45)
            lam, mu = E_ * nu / ((1 + nu) * (1 - 2 * nu)), E_ / (2 * (1 + nu))
            cases.append((lam, mu, x, X))
        with tempfile.TemporaryDirectory() as d:
            inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            write_hex_element_input(inp, nn, cases)
            subprocess.run([os.path.join(REF, "ref_driver"), "hex_element", inp, outp], check=True, env=ENV)
            E, g, H = read_hex_element_output(outp, nn, len(cases))
        out[f"x{nn}"] = np.stack([c[2] for c in cases])
        out[f"X{nn}"] = np.stack([c[3] for c in cases])
        out[f"lam{nn}"] = np.array([c[0] for c in cases])
        out[f"mu{nn}"] = np.array([c[1] for c in cases])
        out[f"E{nn}"], out[f"g{nn}"], out[f"H{nn}"] = E, g, H
    np.savez_compressed(os.path.join(HERE, "hex_ref_elements.npz"), **out)
    print("hex elements", {k: v.shape for k, v in out.items()})


TET10_REF_NODES = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.5, 0, 0), (0.5, 0.5, 0), (0, 0.5, 0), (0, 0, 0.5), (0.5, 0, 0.5),
                   (0, 0.5, 0.5)]


def tet10_cases():
    """Perturbed Tet10 elements: scaled, rotated-and-stretched and noisy unit tetrahedra
    (vertices, then mid-edges 01 12 02 03 13 23), random materials."""
    from ref_io import read_hex_element_output, write_hex_element_input

    rng = np.random.default_rng(20261205)
    base = np.array(TET10_REF_NODES, float)
    cases = []
    for i in range(6):
        h = 10.0 ** rng.uniform(-2, 0)
        X = base * h + rng.uniform(-0.03, 0.03, base.shape) * h
        A = np.eye(3) + rng.uniform(-0.15, 0.15, (3, 3))
        x = X @ A.T + rng.uniform(-0.02, 0.02, X.shape) * h
        E_, nu = 10.0 ** rng.uniform(4, 6), rng.uniform(0.2, 0.45)
        lam, mu = E_ * nu / ((1 + nu) * (1 - 2 * nu)), E_ / (2 * (1 + nu))
        cases.append((lam, mu, x, X))
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_hex_element_input(inp, 10, cases)
        subprocess.run([os.path.join(REF, "ref_driver"), "hex_element", inp, outp], check=True, env=ENV)
        E, g, H = read_hex_element_output(outp, 10, len(cases))
    out = dict(x=np.stack([c[2] for c in cases]), X=np.stack([c[3] for c in cases]), lam=np.array([c[0] for c in cases]),
               mu=np.array([c[1] for c in cases]), E=E, g=g, H=H)
    np.savez_compressed(os.path.join(HERE, "tet10_ref_elements.npz"), **out)
    print("tet10 elements", {k: v.shape for k, v in out.items()})


def hex_newton_case(name, cells, mode=2, rel_tol=1e-6, threads=1, nn=27):
    from ref_io import read_solve_output, write_solve_hex_input
    from scene_np import hex_block
    from symx_amd import NewtonParams

    S = hex_block(cells=cells, nn=nn)
    newton = NewtonParams(0, mode, 0, 100, 0.0, rel_tol, 1e-3, 1e-12, 10000, 1e-10, 0.0)
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_solve_hex_input(inp, S, newton, threads)
        subprocess.run([os.path.join(REF, "ref_driver"), "solve_hex", inp, outp], check=True, env=ENV)
        res = read_solve_output(outp, S.X.shape[0])
    prefix = "tet10" if nn == 10 else f"hex{nn}"
    np.savez_compressed(os.path.join(HERE, f"{prefix}_ref_{name}.npz"), x=res["x"], cells=np.array(cells), rel_tol=rel_tol, result=res["result"],
                        newton_iterations=res["newton_iterations"], cg_iterations=res["cg_iterations"],
                        ls_bt_iterations=res["ls_bt_iterations"], seconds=res["seconds"], threads=threads, projection_mode=mode)
    print(name, {k: v for k, v in res.items() if k != "x"})


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "llt":
        newton_case("block6_llt", (6, 6, 6), 0.2 / 6, 11, 1e-8, 1e-3, llt=True)
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "hex":
        hex_cases()
        hex_newton_case("block2", (2, 2, 2))
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "tet10":
        tet10_cases()
        hex_newton_case("block2", (2, 2, 2), nn=10)
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "adjoint":
        adjoint_case("cantilever16", (15, 3, 3))
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "modes":
        for tag, mode in (("newton", 0), ("pdn", 2), ("ppn", 3)):
            newton_case(f"block6_{tag}", (6, 6, 6), 0.2 / 6, 11, 1e-6, 1e-3, mode=mode)
        sys.exit(0)
    nh_cases()
    point_triangle_cases()
    newton_case("block6", (6, 6, 6), 0.2 / 6, 11, 1e-6, 1e-3)
    newton_case("block6_tight", (6, 6, 6), 0.2 / 6, 11, 1e-8, 1e-10)
    newton_case("block20", (20, 20, 20), 0.01, 1, 1e-6, 1e-3)
