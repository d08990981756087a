"""Benchmark: Newton steps/s and element Hessians/s on the Tet4 Neo-Hookean block,
mixed precision vs FP64, solved to a 1e-6 relative Newton residual.

Workload (default): 101x101x101 vertex grid over [0,1]^3, each hex cell split into 6
tets (6,000,000 tets, 3,090,903 DoF), Neo-Hookean E=1e5 Pa, nu=0.4, density 1000,
gravity -9.81 z, backward-Euler inertia dt=0.01, z=0 face held by a penalty
(k=1e8), free vertices perturbed by seeded uniform noise of 0.2 cell widths.
One step = one pass of the workload: 5 implicit-Euler time steps from the seeded state,
each solved by projected Newton to a 1e-6 relative residual (block-Jacobi PCG to 1e-3
relative, forcing sequence) and followed by the velocity/position update. Every pass
starts from the same seeded state. Inputs are resident in HBM before timing starts; JIT
and topology setup are excluded. `value` = Newton iterations per second of the
mixed-precision solver (one Newton iteration = element E/grad/Hessian evaluation, PD
projection, assembly, PCG solve, line search); `time_to_tolerance_s` = seconds per pass.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the default
# workload (101^3 block, mixed), each counter scaled by its reading for a 1 GiB device copy
# (--calibration-copy-mib): FETCH x 2.0, WRITE x 1.0. The per-kernel table is in profiles/.
MEASURED_TRAFFIC = {("block", "pcg_iteration", 1): 759.3e6}
# what the mixed path computes in: element energies, gradients and material-law derivatives
# in FP64, the element-Hessian contraction in FP32 (FP32 storage), projection in FP64
# arithmetic on the FP32 records, FP32 block matrix and inner PCG, FP64 outer refinement,
# FP64 energies, gradients and line search
DTYPE_MIXED = "f32 element Hessians (f64 material-law derivatives, f32 contraction)/BSR/PCG, f64 projection arithmetic + f64 refinement"


def block_params(cells):
    import symx_amd

    return symx_amd.default_block_params(cells=[cells] * 3, noise=0.2 / cells, seed=2)


def algorithmic_bytes(n_elem, n_vert, nnzb, precision):
    """Compulsory HBM bytes per unit of work (each array's unique footprint once)."""
    sH = 4 if precision == 1 else 8  # element-Hessian and block-matrix storage
    nv_per_e = n_vert / n_elem
    return {
        # conn 16 + x,X 48 per vertex amortised + E partial + E_e 8 (line-search reference) + grad 96
        # + 78 unique Hessian entries
        "element": 16 + 48 * nv_per_e + 8 / 256 + 8 + 96 + 78 * sH,
        # read + write of the 78 unique entries
        "projection": 2 * 78 * sH,
        # per PCG iteration: matrix + column ids (16-bit offsets from the row, which every mesh
        # here fits) + slice offsets + D^-1 (6 distinct values of each symmetric block) + 12
        # vector passes
        "pcg_iteration": nnzb * (9 * sH + 2) + (n_vert + 1) * 4 + n_vert * 6 * sH + 12 * 3 * n_vert * sH,
        # per assembly: element Hessians (78 unique entries + 10 slot ids) + matrix write
        "assembly": n_elem * (78 * sH + 10 * 4) + nnzb * 9 * sH,
    }


def run_precision(scene, precision, steps, warmup, log, keep_newton=False):
    """`warmup` untimed then `steps` timed time steps continuing one trajectory (hex27)."""
    import symx_amd

    if not keep_newton:
        scene.set_newton(symx_amd.default_newton_params(precision=precision))
    for i in range(warmup):
        st = scene.step()
        log(f"warmup step {i}: {st['result_name']} newton={st['newton_iterations']} cg={st['cg_iterations']} {st['seconds']:.3f}s")
    newton = cg = 0
    phases = dict(eval=0.0, project=0.0, assembly=0.0, linear=0.0, energy=0.0)
    results = []
    t0 = time.perf_counter()
    for i in range(steps):
        st = scene.step()
        newton += st["newton_iterations"]
        cg += st["cg_iterations"]
        for k in phases:
            phases[k] += st["seconds_" + k]
        results.append(st["result_name"])
        log(f"step {i}: {st['result_name']} newton={st['newton_iterations']} cg={st['cg_iterations']} refinements={st['n_refinements']} "
            f"r0={st['first_residual']:.3e} r={st['final_residual']:.3e} {st['seconds']:.3f}s")
    t1 = time.perf_counter()
    return {"seconds": t1 - t0, "newton_iterations": newton, "cg_iterations": cg, "phases": phases, "results": results}


class SeededTrajectory:
    """The block workload of one precision: `time_steps` implicit-Euler steps from the seeded
    state. Every pass restores x, x0 and v of the seeded state first (outside the timing), so
    each pass is the same solve from the same inputs."""

    def __init__(self, cells, precision, time_steps):
        import symx_amd

        self.precision = precision
        self.time_steps = time_steps
        self.scene = symx_amd.Tet4Block(block_params(cells), symx_amd.default_newton_params(precision=precision))
        self.seed_state = {k: self.scene.get(k) for k in ("x", "x0", "v")}

    def run_pass(self):
        for k, a in self.seed_state.items():
            self.scene.set(k, a)
        steps = []
        newton = cg = 0
        phases = dict(eval=0.0, project=0.0, assembly=0.0, linear=0.0, energy=0.0)
        t0 = time.perf_counter()
        for _ in range(self.time_steps):
            st = self.scene.step()
            newton += st["newton_iterations"]
            cg += st["cg_iterations"]
            for k in phases:
                phases[k] += st["seconds_" + k]
            steps.append({"result": st["result_name"], "newton": st["newton_iterations"], "cg": st["cg_iterations"],
                          "refinements": st["n_refinements"], "trapezoid": st["ls_trapezoid_accepts"],
                          "r0": st["first_residual"], "r": st["final_residual"],
                          "r_over_r0": st["final_residual"] / st["first_residual"] if st["first_residual"] > 0 else 0.0,
                          "seconds": st["seconds"]})
        t1 = time.perf_counter()
        return {"seconds": t1 - t0, "newton_iterations": newton, "cg_iterations": cg, "phases": phases, "steps": steps}

    def run(self, passes, warmup, log, tag):
        for i in range(warmup):
            r = self.run_pass()
            log(f"[{tag}] warmup pass {i}: {r['seconds']:.3f}s " + " ".join(f"{s['result']}/{s['newton']}" for s in r["steps"]))
        runs = []
        for i in range(passes):
            r = self.run_pass()
            runs.append(r)
            log(f"[{tag}] pass {i}: {r['seconds']:.3f}s newton={r['newton_iterations']} cg={r['cg_iterations']}")
        for j, s in enumerate(runs[-1]["steps"]):
            log(f"[{tag}] time step {j}: {s['result']} newton={s['newton']} cg={s['cg']} refinements={s['refinements']} "
                f"trapezoid={s['trapezoid']} "
                f"r0={s['r0']:.3e} r={s['r']:.3e} r/r0={s['r_over_r0']:.2e} {s['seconds']:.3f}s")
        secs = [r["seconds"] for r in runs]
        # every pass solves the same inputs with fixed summation orders: the iteration
        # counts (and the final positions) repeat exactly
        same = all(r["newton_iterations"] == runs[0]["newton_iterations"] and r["cg_iterations"] == runs[0]["cg_iterations"] for r in runs)
        return {
            "seconds": sum(secs),
            "passes": passes,
            "median_pass_seconds": float(np.median(secs)),
            "newton_iterations": sum(r["newton_iterations"] for r in runs),
            "cg_iterations": sum(r["cg_iterations"] for r in runs),
            "phases": {k: sum(r["phases"][k] for r in runs) for k in runs[0]["phases"]},
            "steps": runs[-1]["steps"],
            "repeatable": same,
        }

    def positions(self):
        return self.scene.get("x")

    def close(self):
        self.scene.close()


def step_summary(r):
    steps = r["steps"]
    return {
        "time_steps_successful": sum(s["result"] == "Successful" for s in steps),
        "time_steps": len(steps),
        "worst_r_over_r0": max(s["r_over_r0"] for s in steps),
        "per_time_step": [{"result": s["result"], "newton": s["newton"], "cg": s["cg"], "trapezoid_ls": s["trapezoid"],
                           "r0": float("%.4e" % s["r0"]),
                           "r_over_r0": float("%.3e" % s["r_over_r0"])} for s in steps],
    }


def profile_kernels(scene, precision, reps=5):
    """Average kernel durations (ms) with HIP events on the runtime's own stream."""
    import ctypes

    import symx_amd

    p = scene.problem
    out = {}
    scene.set_dofs(scene.get("x"))
    scene.eval(symx_amd.EVAL_PGH_F32 if precision == symx_amd.MIXED else symx_amd.EVAL_PGH)
    # element kernel with the Hessian storage of `precision`; projection and assembly then
    # work on that storage
    for name, phase, arg in [("element", 0, precision), ("energy", 4, 0), ("projection", 1, 0), ("assembly", 2, precision),
                             ("pcg_iteration", 3, precision)]:
        ms = ctypes.c_double()
        symx_amd.check_hip(symx_amd.hip().symx_hip_problem_profile(p, ctypes.c_int32(phase), ctypes.c_int32(0), ctypes.c_int32(arg),
                                                                   ctypes.c_int32(reps), ctypes.byref(ms)))
        out[name] = ms.value
    return out


def host_threads():
    """(threads used by the CPU legs, physical cores from lscpu). A GPU box gives one GPU a
    share of 16 host cores; the threads are the CPUs this process may run on, capped there."""
    import subprocess

    n = len(os.sched_getaffinity(0))
    threads = max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n)), 16))
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.CalledProcessError):
        pass
    return threads, phys


def reference_solve(S, threads, log):
    """One Newton solve of scene S by the reference itself (oracle/_ref/ref_driver, SymX
    compiled from its sources), with its own timer around NewtonsMethod::solve. None when
    the reference build is absent."""
    import subprocess
    import tempfile

    from ref_io import read_solve_output, write_solve_input

    import symx_amd

    exe = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.access(exe, os.X_OK):
        return None
    newton = symx_amd.NewtonParams(0, 1, 0, 100, 0.0, 1e-6, 1e-3, 1e-12, 10000, 1e-10, 0.0)
    env = {k: v for k, v in os.environ.items() if k != "CXX"}
    env.update(OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_solve_input(inp, S, newton, threads)
        r = subprocess.run([exe, "solve", inp, outp], capture_output=True, text=True, env=env)
        if r.returncode != 0:
            log(f"reference solve failed ({r.returncode}): {r.stderr[-500:]}")
            return None
        return read_solve_output(outp, S.X.shape[0])


def cpu_baseline(log, cells, threads):
    """CPU legs on a bounded sample, one implicit step of a cells^3 block (same material,
    loads and solver settings): the reference itself (SymX's OpenMP path) and the oracle
    (this repository's C restatement of it), both on `threads` host threads."""
    from oracle_py import Oracle
    from scene_np import tet4_block

    import symx_amd

    S = tet4_block(cells=(cells,) * 3, noise=0.2 / cells, seed=2)
    n_elem = S.tets.shape[0]
    out = {"n_elem": n_elem, "cells": cells}
    ref = reference_solve(S, threads, log)
    if ref is not None:
        log(f"cpu reference ({threads} threads): {cells}^3 cells, {n_elem} tets: result={ref['result']} newton={ref['newton_iterations']} "
            f"cg={ref['cg_iterations']} {ref['seconds']:.2f}s")
        out["reference"] = ref
    orc = Oracle()
    newton = symx_amd.NewtonParams(0, 1, 0, 100, 0.0, 1e-6, 1e-3, 1e-12, 10000, 1e-10, 0.0)
    S2 = tet4_block(cells=(cells,) * 3, noise=0.2 / cells, seed=2)
    t0 = time.perf_counter()
    res = orc.newton(S2, newton)
    t1 = time.perf_counter()
    log(f"cpu oracle ({orc.num_threads()} threads): newton={res['newton_iterations']} cg={res['cg_iterations']} {t1 - t0:.2f}s")
    out["port"] = {"newton_iterations": res["newton_iterations"], "cg_iterations": res["cg_iterations"], "seconds": t1 - t0,
                   "threads": orc.num_threads(), "x": S2.x.copy()}
    out["S"] = S
    return out


def gpu_sample_errors(cells, cpu, log):
    """The same one-step sample solved on the GPU in FP64 and mixed precision, against the
    reference's own positions (and the oracle's): relative to max|x| and to the largest
    displacement of the solve."""
    import symx_amd

    S = cpu["S"]
    ref_x = cpu["reference"]["x"] if "reference" in cpu else None
    out = {}
    for name, prec in (("fp64", symx_amd.FP64), ("mixed", symx_amd.MIXED)):
        s = symx_amd.Tet4Block(block_params(cells), symx_amd.default_newton_params(precision=prec))
        st = s.solve()
        x = s.get("x")
        s.close()
        e = {"result": st["result_name"], "newton": st["newton_iterations"]}
        for tag, xr in (("reference", ref_x), ("oracle", cpu["port"]["x"])):
            if xr is None:
                continue
            d = np.abs(x - xr).max()
            e[f"vs_{tag}_rel_max_x"] = float("%.3e" % (d / np.abs(xr).max()))
            e[f"vs_{tag}_rel_displacement"] = float("%.3e" % (d / np.abs(xr - S.x).max()))
        out[name] = e
        log(f"gpu {name} on the {cells}^3 sample: {e}")
    return out


def bench_adjoint(args, log):
    """Adjoint gradients/s on the quasistatic Neo-Hookean cantilever: 64x16x16 vertex grid
    (63x15x15 cells, 85,050 tets, 49,152 DoF) of length 0.1 m, x=0 face held, gravity,
    per-tet log(mu) ~ N(log 3.5e4, 0.3^2) (seed 1), nu = 0.4, L = 0.5 |x_eq - x_target|^2
    with x_target the equilibrium of a second field (seed 2). One step = one gradient
    dL/dlog_mu at the equilibrium: E/grad/Hessian evaluation, unprojected assembly, PCG
    for H lambda = dL/dx to 1e-10 relative, mixed-partial kernel, scatter."""
    import ctypes

    import symx_amd

    cells = [63, 15, 15]
    L = 0.1
    params = symx_amd.default_block_params(cells=cells, size=[L, L * 15 / 63, L * 15 / 63], youngs_modulus=1e5, poissons_ratio=0.4,
                                           density=1000.0, gravity=[0.0, 0.0, -9.81], dt=0.0, noise=0.0, fixed_axis=0, log_mu_field=1)
    prec = symx_amd.FP64 if args.adjoint_precision == "fp64" else symx_amd.MIXED
    newton = symx_amd.default_newton_params(precision=prec, projection_mode=1, residual_tolerance_rel=1e-4, max_iterations=300,
                                            cg_max_iterations=100000)
    scene = symx_amd.Tet4Block(params, newton)
    ne, nv = scene.n_elements, scene.n_vertices
    X = scene.get("X")

    def field(seed):
        return np.log(3.5e4) + 0.3 * np.random.default_rng(seed).standard_normal(ne)

    def equilibrium(log_mu):
        scene.set_log_mu(log_mu)
        scene.set("x", X)
        t = time.perf_counter()
        st = scene.solve()
        log(f"forward solve: {st['result_name']} newton={st['newton_iterations']} cg={st['cg_iterations']} {time.perf_counter() - t:.3f}s")
        return scene.get("x").reshape(-1), st

    x_t, _ = equilibrium(field(2))
    x, fwd = equilibrium(field(1))
    rhs = x - x_t
    for i in range(args.warmup):
        g, st = scene.adjoint(rhs, rel_tol=1e-10)
        log(f"warmup adjoint {i}: cg={st['cg_iterations']} converged={st['converged']} {st['seconds']:.4f}s")
    cg = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        g, st = scene.adjoint(rhs, rel_tol=1e-10)
        cg += st["cg_iterations"]
    t1 = time.perf_counter()
    log(f"adjoint: {args.steps} gradients in {t1 - t0:.4f}s, cg/gradient={cg / args.steps:.0f}, max|dL/dp|={np.abs(g).max():.3e}")

    # kernel times (HIP events on the runtime stream) at the equilibrium
    p = scene.problem
    prof = {}
    scene.set_dofs(x)
    scene.eval(symx_amd.EVAL_PGH)
    for name, phase, arg in [("element", 0, 0), ("assembly", 2, prec), ("pcg_iteration", 3, prec)]:
        ms = ctypes.c_double()
        symx_amd.check_hip(symx_amd.hip().symx_hip_problem_profile(p, ctypes.c_int32(phase), ctypes.c_int32(0), ctypes.c_int32(arg),
                                                                   ctypes.c_int32(5), ctypes.byref(ms)))
        prof[name] = ms.value
    nnzb_c = ctypes.c_int64()
    symx_amd.check_hip(symx_amd.hip().symx_hip_problem_export_bsr(p, ctypes.byref(nnzb_c), None, None, None))
    ab = algorithmic_bytes(ne, nv, nnzb_c.value, prec)
    per_grad_ms = {"element": prof["element"], "assembly": prof["assembly"], "pcg_iteration": prof["pcg_iteration"] * cg / args.steps}
    dominant = max(per_grad_ms, key=per_grad_ms.get)
    unit_bytes = ab[dominant] * (ne if dominant == "element" else 1)
    achieved = unit_bytes / (prof[dominant] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "kernel": dominant, "avg_ms": round(prof[dominant], 4), "algorithmic_bytes_per_launch": int(unit_bytes)}
    cpu = None
    if not args.no_cpu:
        from oracle_py import Oracle, SceneData

        orc = Oracle()
        tets = scene.connectivity()
        bc = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
        mu, lam = 1e5 / (2 * 1.4), 1e5 * 0.4 / (1.4 * 0.2)
        S = SceneData(tets, X, x.reshape(-1, 3), X, np.zeros_like(X), scene.get("mass"), bc, X[bc], lam, mu, 0.0, 1e8, [0.0, 0.0, -9.81],
                      mu_e=np.exp(field(1)), lambda_over_mu=lam / mu)
        tc = time.perf_counter()
        g_cpu, _, info = orc.adjoint_logmu(S, rhs, rel_tol=1e-10)
        tcpu = time.perf_counter() - tc
        log(f"cpu adjoint ({orc.num_threads()} threads): {tcpu:.2f}s cg={info['n_iterations']} "
            f"max|g_gpu-g_cpu|/max|g|={np.abs(g - g_cpu).max() / np.abs(g_cpu).max():.2e}")
        cpu = {"value": round(1.0 / tcpu, 4), "unit": "adjoint gradients/s", "cores": orc.num_threads(), "kind": "port",
               "sample": f"oracle/symx_oracle.c oracle_adjoint_logmu, one full gradient of the same problem ({tcpu:.2f}s)"}
    line = {
        "metric": "adjoint gradients/s (dL/dlog_mu, 85,050 parameters)",
        "value": round(args.steps / (t1 - t0), 4),
        "unit": "adjoint gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (t1 - t0) / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == symx_amd.FP64 else DTYPE_MIXED,
        "data": "synthetic",
        "config": {"workload": "quasistatic Tet4 Neo-Hookean cantilever 64x16x16 vertices (85050 tets, 49152 DoF), per-tet log(mu), "
                               "adjoint of L=0.5|x-x_target|^2, PCG to 1e-10 rel.", "n_elem": ne, "n_dof": 3 * nv, "nnzb": nnzb_c.value},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forward_solve": {"newton_iterations": fwd["newton_iterations"], "cg_iterations": fwd["cg_iterations"], "seconds": round(fwd["seconds"], 4)},
        "cg_iterations_per_gradient": cg / args.steps,
        "kernel_ms": {k: round(v, 4) for k, v in prof.items()},
    }
    print(json.dumps(line), flush=True)
    scene.close()


def bench_contact(args, log):
    """Ball drop with contact: Tet4 Neo-Hookean ball of radius 0.5 (jittered-grid
    tetrahedralisation, 50 cells across, seed 7: 393,000 tets) dropped at 2 m/s from 1 cm
    above the plane z = 0; IPC log-barrier per surface vertex (dhat 1e-3, kappa 1e6),
    plane step filter (eta 0.9), backward Euler dt = 1/200, Newton with Progressive
    projection (the projection mode of SymX's contact example) to 1e-5 relative residual
    (1e-6 is below the round-off floor of the energy-based line search here) with
    mixed-precision PCG (FP64 barrier energies, gradients and line search). One step = one
    time step; the workload is `steps` time steps (50 for the reference configuration)."""
    import symx_amd

    p = symx_amd.default_sphere_params()
    newton = symx_amd.default_newton_params(precision=symx_amd.MIXED, projection_mode=3, residual_tolerance_rel=1e-5)
    t0 = time.perf_counter()
    s = symx_amd.Sphere(p, newton)
    log(f"setup: {s.n_elements} tets, {s.n_vertices} vertices, {len(s.contact_vertices())} contact vertices in {time.perf_counter() - t0:.1f}s")
    if args.warmup > 0:
        # warm up on a second ball so that the timed trajectory starts at t = 0 (re-solving
        # an already converged step would make its relative residual target unreachable)
        w = symx_amd.Sphere(p, newton)
        for i in range(args.warmup):
            st = w.step()
            log(f"warmup step {i}: {st['result_name']} newton={st['newton_iterations']} {st['seconds']:.3f}s")
        w.close()
    newton_it = cg = 0
    results = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = s.step()
        newton_it += st["newton_iterations"]
        cg += st["cg_iterations"]
        results.append(st["result_name"])
        if i % 10 == 0 or st["result"] != 0:
            log(f"step {i}: {st['result_name']} newton={st['newton_iterations']} cg={st['cg_iterations']} bt={st['ls_bt_iterations']} "
                f"zmin={s.get('x')[:, 2].min():.3e} {st['seconds']:.3f}s")
    t1 = time.perf_counter()
    x = s.get("x")
    log(f"{args.steps} steps in {t1 - t0:.3f}s: newton={newton_it} cg={cg} zmin={x[:, 2].min():.3e} vz={s.get('v')[:, 2].mean():.3f}")
    import ctypes

    prof = profile_kernels(s, symx_amd.MIXED)
    nnzb_c = ctypes.c_int64()
    symx_amd.check_hip(symx_amd.hip().symx_hip_problem_export_bsr(s.problem, ctypes.byref(nnzb_c), None, None, None))
    ab = algorithmic_bytes(s.n_elements, s.n_vertices, nnzb_c.value, symx_amd.MIXED)
    per_iter = {"element": prof["element"], "projection": prof["projection"], "assembly": prof["assembly"],
                "pcg_iteration": prof["pcg_iteration"] * cg / max(newton_it, 1)}
    dominant = max(per_iter, key=per_iter.get)
    unit_bytes = ab[dominant] * (s.n_elements if dominant in ("element", "projection") else 1)
    achieved = unit_bytes / (prof[dominant] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "kernel": dominant, "avg_ms": round(prof[dominant], 4), "algorithmic_bytes_per_launch": int(unit_bytes)}
    cpu = None
    if not args.no_cpu:
        from oracle_py import Oracle, sphere_arrays

        orc = Oracle()
        s2 = symx_amd.Sphere(p, symx_amd.default_newton_params(precision=symx_amd.FP64))
        S = sphere_arrays(s2)
        s2.close()
        tc = time.perf_counter()
        r = orc.newton(S, symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=3, residual_tolerance_rel=1e-5))
        tcpu = time.perf_counter() - tc
        log(f"cpu ({orc.num_threads()} threads): first time step {r['result']} newton={r['newton_iterations']} {tcpu:.2f}s")
        cpu = {"value": round(1.0 / tcpu, 4), "unit": "time steps/s", "cores": orc.num_threads(), "kind": "port",
               "sample": f"oracle/symx_oracle.c, the first time step of the same ball drop ({r['newton_iterations']} Newton iterations, {tcpu:.1f}s)"}
    line = {
        "metric": "time steps/s, ball drop with IPC contact",
        "value": round(args.steps / (t1 - t0), 4),
        "unit": "time steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (t1 - t0) / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_MIXED,
        "data": "synthetic",
        "config": {"workload": f"Tet4 Neo-Hookean ball r=0.5 ({s.n_elements} tets) dropped at 2 m/s on z=0, IPC barrier dhat=1e-3 kappa=1e6, "
                               "step filter 0.9, dt=1/200, Progressive projection, 1e-5 rel. residual", "n_elem": s.n_elements, "n_dof": 3 * s.n_vertices, "nnzb": nnzb_c.value},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "newton_iterations": newton_it,
        "cg_iterations": cg,
        "results": sorted(set(results)),
        "kernel_ms": {k: round(v, 4) for k, v in prof.items()},
    }
    print(json.dumps(line), flush=True)
    s.close()


def bench_hex27(args, log):
    """High-order: Hex27 Neo-Hookean cube of 40^3 cells over [0,1]^3 (64,000 elements,
    531,441 nodes), per-element Young's modulus log-uniform in [1e4, 1e6] (seed 3),
    nu = 0.45, density 1000, gravity, backward Euler dt = 0.01, z = 0 face held by the
    penalty (k = 1e8), free nodes perturbed by 0.1 node spacings; ProjectOnDemand
    projection, Newton to 1e-6 relative residual; mixed precision (FP32 Hessians/BSR/PCG
    with FP64 refinement) against FP64. One step = one time step."""
    import ctypes

    import symx_amd

    cells = args.hex_cells
    p = symx_amd.default_hex_params(cells=[cells] * 3, nodes_per_element=27)
    out = {}
    for name, prec in (("mixed", symx_amd.MIXED), ("fp64", symx_amd.FP64)):
        if name == "fp64" and args.no_fp64:
            continue
        newton = symx_amd.default_newton_params(precision=prec, projection_mode=2)
        t0 = time.perf_counter()
        s = symx_amd.HexBlock(p, newton)
        log(f"[{name}] setup: {s.n_elements} elements, {s.n_vertices} nodes in {time.perf_counter() - t0:.1f}s "
            f"(JIT + topology {symx_amd.amd().symx_scene_compile_seconds(s.ptr):.1f}s)")
        r = run_precision(s, prec, args.steps, args.warmup, log, keep_newton=True)
        prof = profile_kernels(s, prec, reps=3)
        log(f"[{name}] kernel ms: {json.dumps({k: round(v, 4) for k, v in prof.items()})}")
        nnzb_c = ctypes.c_int64()
        symx_amd.check_hip(symx_amd.hip().symx_hip_problem_export_bsr(s.problem, ctypes.byref(nnzb_c), None, None, None))
        ops = s.op_counts(0)
        out[name] = {"run": r, "prof": prof, "nnzb": nnzb_c.value, "ops": ops, "n_elem": s.n_elements, "n_vert": s.n_vertices}
        s.close()
    m = out["mixed"]
    n_elem, nnzb = m["n_elem"], m["nnzb"]
    # algorithmic FLOPs of one structured Hex27 PGH element (FMA = 2): Jacobians of x and X
    # per quadrature point, the generated material law per point, gradient and Hessian
    # contractions (378 blocks x 9 entries x 3 x (3 + 1) FMA per point)
    flops_elem = 27 * (2 * 27 * 9 * 2 + m["ops"][2]) + 27 * 81 * 3 * 2 + 27 * 378 * 9 * 3 * 4 * 2
    el_ms = m["prof"]["element"]
    achieved = flops_elem * n_elem / (el_ms * 1e-3) / 1e12
    per_iter = {"element": el_ms, "projection": m["prof"]["projection"], "assembly": m["prof"]["assembly"],
                "pcg_iteration": m["prof"]["pcg_iteration"] * m["run"]["cg_iterations"] / max(m["run"]["newton_iterations"], 1)}
    roofline = {"bound": "valu", "achieved": round(achieved, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None, "kernel": "element (structured Hex27 PGH)",
                "avg_ms": round(el_ms, 4), "algorithmic_flops_per_launch": int(flops_elem * n_elem),
                "per_newton_iteration_ms": {k: round(v, 3) for k, v in per_iter.items()}}
    cpu = None
    if not args.no_cpu:
        from oracle_py import Oracle
        from scene_np import hex_block

        orc = Oracle()
        c = args.cpu_hex_cells
        S = hex_block(cells=(c, c, c), nn=27, noise=p.noise, seed=p.seed)
        t0 = time.perf_counter()
        res = orc.newton(S, symx_amd.default_newton_params(precision=symx_amd.FP64, projection_mode=2))
        t1 = time.perf_counter()
        ne_c = S.tets.shape[0]
        log(f"cpu ({orc.num_threads()} threads): {c}^3 Hex27 cells, newton={res['newton_iterations']} cg={res['cg_iterations']} {t1 - t0:.2f}s")
        cpu = {"value": round(res["newton_iterations"] / (t1 - t0) * ne_c / n_elem, 4), "unit": "Newton iterations/s", "cores": orc.num_threads(),
               "kind": "port", "sample": f"oracle/symx_oracle.c, one time step of a {c}^3-cell Hex27 cube ({ne_c} elements, {t1 - t0:.1f}s), "
                                         f"scaled linearly by element count to {n_elem} elements"}
    line = {
        "metric": "Newton steps/s, elem Hessians/s; Hex27 mixed vs FP64 at 1e-6 residual",
        "value": round(m["run"]["newton_iterations"] / m["run"]["seconds"], 4),
        "unit": "Newton iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * m["run"]["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_MIXED,
        "data": "synthetic",
        "config": {"workload": f"Hex27 Neo-Hookean cube {cells}^3 cells ({n_elem} elements, {3 * m['n_vert']} DoF), E log-uniform [1e4,1e6], "
                               "nu=0.45, implicit Euler dt=0.01, ProjectOnDemand, 1e-6 rel. residual", "n_elem": n_elem,
                   "n_dof": 3 * m["n_vert"], "nnzb": nnzb},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "elem_hessians_per_s": round(n_elem / (el_ms * 1e-3), 1),
        "newton_iterations": m["run"]["newton_iterations"],
        "cg_iterations": m["run"]["cg_iterations"],
        "results": m["run"]["results"],
        "kernel_ms": {k: round(v, 4) for k, v in m["prof"].items()},
        "generated_ops": {"P": m["ops"][0], "PG": m["ops"][1], "PGH_per_quadrature_point": m["ops"][2]},
    }
    if "fp64" in out:
        f = out["fp64"]
        line["fp64"] = {"newton_per_s": round(f["run"]["newton_iterations"] / f["run"]["seconds"], 4), "seconds": round(f["run"]["seconds"], 4),
                        "newton_iterations": f["run"]["newton_iterations"], "cg_iterations": f["run"]["cg_iterations"],
                        "kernel_ms": {k: round(v, 4) for k, v in f["prof"].items()},
                        "speedup_mixed_over_fp64_wall": round(f["run"]["seconds"] / m["run"]["seconds"], 3)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100)
    ap.add_argument("--time-steps", type=int, default=5, help="implicit-Euler steps per pass of the block workload")
    ap.add_argument("--fp64-steps", type=int, default=0, help="timed passes of the FP64 comparison (0: as --steps)")
    ap.add_argument("--no-north-star", action="store_true", help="skip the 64^3 mixed/FP64 comparison")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cells", type=int, default=40)
    ap.add_argument("--workload", choices=["block", "cantilever_adjoint", "sphere_contact", "hex27"], default="block",
                    help="block: the headline Newton benchmark; cantilever_adjoint: adjoint gradients/s; sphere_contact: ball drop; "
                         "hex27: high-order cube")
    ap.add_argument("--hex-cells", type=int, default=40)
    ap.add_argument("--cpu-hex-cells", type=int, default=6)
    ap.add_argument("--adjoint-precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (the known-size kernel that scales rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE readings into HBM bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import symx_amd

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if symx_amd.device_count() < 1:
        raise RuntimeError("bench.py needs a GPU")
    if args.calibration_copy_mib > 0:
        import torch

        n = args.calibration_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: 3 x {args.calibration_copy_mib} MiB read + {args.calibration_copy_mib} MiB written")
        del src, dst
    if args.workload == "cantilever_adjoint":
        bench_adjoint(args, log)
        return
    if args.workload == "sphere_contact":
        bench_contact(args, log)
        return
    if args.workload == "hex27":
        bench_hex27(args, log)
        return
    bench_block(args, log)


def bench_block(args, log):
    """The headline workload (see the module docstring): one step = one pass of the 5
    implicit-Euler time steps from the seeded state, mixed precision timed; FP64 beside it."""
    import ctypes

    import symx_amd

    cells, T = args.cells, args.time_steps
    t0 = time.perf_counter()
    mixed_tr = SeededTrajectory(cells, symx_amd.MIXED, T)
    scene = mixed_tr.scene
    log(f"setup {cells}^3 cells: {scene.n_elements} tets, {scene.n_vertices} vertices in {time.perf_counter() - t0:.1f}s "
        f"(JIT + topology {symx_amd.amd().symx_scene_compile_seconds(scene.ptr):.1f}s)")
    mixed = mixed_tr.run(args.steps, args.warmup, log, "mixed")
    x_mixed = mixed_tr.positions()
    X = scene.get("X")
    prof_mixed = profile_kernels(scene, symx_amd.MIXED)
    log(f"kernel ms (mixed): {json.dumps({k: round(v, 4) for k, v in prof_mixed.items()})}")
    fp64 = prof64 = None
    if not args.no_fp64:
        tr64 = SeededTrajectory(cells, symx_amd.FP64, T)
        fp64 = tr64.run(args.fp64_steps if args.fp64_steps > 0 else args.steps, min(args.warmup, 1), log, "fp64")
        x64 = tr64.positions()
        prof64 = profile_kernels(tr64.scene, symx_amd.FP64)
        log(f"kernel ms (fp64): {json.dumps({k: round(v, 4) for k, v in prof64.items()})}")
        tr64.close()

    nnzb_c = ctypes.c_int64()
    symx_amd.check_hip(symx_amd.hip().symx_hip_problem_export_bsr(scene.problem, ctypes.byref(nnzb_c), None, None, None))
    nnzb = nnzb_c.value
    n_elem, n_vert = scene.n_elements, scene.n_vertices
    mixed_tr.close()

    newton_per_s = mixed["newton_iterations"] / mixed["seconds"]
    # Roofline of the dominant kernel of a mixed Newton iteration (by measured share).
    ab = algorithmic_bytes(n_elem, n_vert, nnzb, symx_amd.MIXED)
    per_iter_ms = {
        "element": prof_mixed["element"],
        "projection": prof_mixed["projection"],
        "assembly": prof_mixed["assembly"],
        "pcg_iteration": prof_mixed["pcg_iteration"] * (mixed["cg_iterations"] / max(mixed["newton_iterations"], 1)),
    }
    dominant = max(per_iter_ms, key=per_iter_ms.get)
    unit_bytes = ab[dominant] * (n_elem if dominant in ("element", "projection") else 1)
    achieved = unit_bytes / (prof_mixed[dominant] * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get(("block", dominant, symx_amd.MIXED)) if cells == 100 else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else int(traffic), "kernel": dominant, "avg_ms": round(prof_mixed[dominant], 4),
                "algorithmic_bytes_per_launch": int(unit_bytes)}
    elem_hessians_per_s = n_elem / (prof_mixed["element"] * 1e-3)

    # north-star size: 64^3 vertices (1,500,282 tets), the same workload once per precision
    north = None
    if not args.no_north_star and not args.no_fp64:
        north = {}
        for name, prec in (("mixed", symx_amd.MIXED), ("fp64", symx_amd.FP64)):
            tr = SeededTrajectory(63, prec, T)
            r = tr.run(1, 1, log, f"64^3 {name}")
            north[name] = {"seconds": round(r["seconds"], 4), "newton_iterations": r["newton_iterations"], **{
                k: v for k, v in step_summary(r).items() if k != "per_time_step"}}
            tr.close()
        north["fp64_over_mixed_wall"] = round(north["fp64"]["seconds"] / north["mixed"]["seconds"], 3)
        north["workload"] = f"the same {T} time steps on 64^3 vertices (1500282 tets, 786432 DoF)"

    threads, phys = host_threads()
    cpu = None
    sample_err = None
    if not args.no_cpu:
        c = cpu_baseline(log, args.cpu_cells, threads)
        scale = c["n_elem"] / n_elem  # linear extrapolation in element count (optimistic for the CPU)
        port = c["port"]
        port_rate = port["newton_iterations"] / port["seconds"] * scale
        if "reference" in c:
            ref = c["reference"]
            cpu = {"value": round(ref["newton_iterations"] / ref["seconds"] * scale, 4), "unit": "Newton iterations/s", "cores": threads,
                   "kind": "reference",
                   "sample": f"SymX itself (oracle/_ref/ref_driver, built from the reference sources), NewtonsMethod::solve of one implicit "
                             f"step of a {args.cpu_cells}^3-cell block ({c['n_elem']} tets, {ref['newton_iterations']} Newton iterations, "
                             f"{ref['seconds']:.2f}s on {threads} threads), scaled linearly by element count to {n_elem} tets",
                   "lscpu_physical_cores": phys,
                   "port": {"value": round(port_rate, 4), "seconds": round(port["seconds"], 3), "threads": port["threads"],
                            "newton_iterations": port["newton_iterations"],
                            "port_over_reference_time": round(port["seconds"] / ref["seconds"], 3)}}
        else:
            cpu = {"value": round(port_rate, 4), "unit": "Newton iterations/s", "cores": port["threads"], "kind": "port",
                   "sample": f"oracle/symx_oracle.c, one implicit step of a {args.cpu_cells}^3-cell block ({c['n_elem']} tets, "
                             f"{port['seconds']:.1f}s), scaled linearly by element count to {n_elem} tets",
                   "lscpu_physical_cores": phys}
        sample_err = gpu_sample_errors(args.cpu_cells, c, log)

    extra = {
        "elem_hessians_per_s": round(elem_hessians_per_s, 1),
        "time_steps_per_s": round(args.steps * T / mixed["seconds"], 4),
        "time_to_tolerance_s": round(mixed["median_pass_seconds"], 4),
        "newton_iterations": mixed["newton_iterations"],
        "cg_iterations": mixed["cg_iterations"],
        **step_summary(mixed),
        "repeatable": mixed["repeatable"],
        "phase_seconds": {k: round(v, 4) for k, v in mixed["phases"].items()},
        "kernel_ms": {k: round(v, 4) for k, v in prof_mixed.items()},
    }
    if fp64:
        extra["fp64"] = {
            "newton_per_s": round(fp64["newton_iterations"] / fp64["seconds"], 4),
            "passes": fp64["passes"],
            "seconds": round(fp64["seconds"], 4),
            "time_to_tolerance_s": round(fp64["median_pass_seconds"], 4),
            "newton_iterations": fp64["newton_iterations"],
            "cg_iterations": fp64["cg_iterations"],
            **step_summary(fp64),
            "elem_hessians_per_s": round(n_elem / (prof64["element"] * 1e-3), 1),
            "kernel_ms": {k: round(v, 4) for k, v in prof64.items()},
            "fp64_over_mixed_time_to_tolerance": round(fp64["median_pass_seconds"] / mixed["median_pass_seconds"], 3),
        }
        d = np.abs(x_mixed - x64).max()
        extra["error"] = {
            "mixed_vs_fp64_after_workload_rel_max_x": float("%.3e" % (d / np.abs(x64).max())),
            "mixed_vs_fp64_after_workload_rel_displacement": float("%.3e" % (d / np.abs(x64 - X).max())),
        }
    if sample_err is not None:
        extra.setdefault("error", {})[f"sample_{args.cpu_cells}cube_one_step"] = sample_err
    if north is not None:
        extra["north_star_64cube"] = north
    line = {
        "metric": "Newton steps/s, elem Hessians/s; mixed vs FP64 at 1e-6 residual",
        "value": round(newton_per_s, 4),
        "unit": "Newton iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_MIXED,
        "data": "synthetic",
        "config": {"workload": f"Tet4 Neo-Hookean block {cells + 1}^3 vertices ({n_elem} tets, {3 * n_vert} DoF): {T} implicit-Euler steps "
                               f"(dt=0.01) from the seeded state, projected Newton to 1e-6 rel. residual; one step = that whole pass",
                   "n_elem": n_elem, "n_dof": 3 * n_vert, "nnzb": nnzb, "time_steps_per_step": T},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
