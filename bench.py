"""Benchmark: Newton steps/s and element Hessians/s on the Tet4 Neo-Hookean block,
mixed precision vs FP64, solved to a 1e-6 relative Newton residual.

Workload (default): 101x101x101 vertex grid over [0,1]^3, each hex cell split into 6
tets (6,000,000 tets, 3,090,903 DoF), Neo-Hookean E=1e5 Pa, nu=0.4, density 1000,
gravity -9.81 z, backward-Euler inertia dt=0.01, z=0 face held by a penalty
(k=1e8), free vertices perturbed by seeded uniform noise of 0.2 cell widths.
One step = one implicit-Euler time step solved by projected Newton to a 1e-6
relative residual (block-Jacobi PCG to 1e-3 relative, forcing sequence), followed
by the velocity/position update. Inputs are resident in HBM before timing starts;
JIT and topology setup are excluded. `value` = Newton iterations per second of the
mixed-precision solver (one Newton iteration = element E/grad/Hessian evaluation,
PD projection, assembly, PCG solve, line search).
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


def block_params(cells):
    import symx_amd

    return symx_amd.default_block_params(cells=[cells] * 3, noise=0.2 / cells, seed=2)


def algorithmic_bytes(n_elem, n_vert, nnzb, precision):
    """Compulsory HBM bytes per unit of work (each array's unique footprint once)."""
    sH = 4 if precision == 1 else 8
    nv_per_e = n_vert / n_elem
    return {
        # conn 16 + x,X 48 per vertex amortised + E partial + grad 96 + 78 unique Hessian entries
        "element": 16 + 48 * nv_per_e + 8 / 256 + 96 + 78 * sH,
        # read + write of the 78 unique entries (FP64 projection)
        "projection": 2 * 78 * 8,
        # per PCG iteration: matrix + column ids + row offsets + D^-1 + 12 vector passes
        "pcg_iteration": nnzb * (9 * sH + 4) + (n_vert + 1) * 4 + n_vert * 9 * sH + 12 * 3 * n_vert * sH,
        # per assembly: element Hessians (10 upper blocks incl. slot ids) + matrix write
        "assembly": n_elem * (78 * 8 + 10 * 4) + nnzb * 9 * sH,
    }


def run_precision(scene, precision, steps, warmup, log):
    import symx_amd

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


def profile_kernels(scene, precision, reps=5):
    """Average kernel durations (ms) with HIP events on the runtime's own stream."""
    import ctypes

    import symx_amd

    p = scene.problem
    out = {}
    scene.set_dofs(scene.get("x"))
    scene.eval(symx_amd.EVAL_PGH)
    for name, phase, arg in [("element", 0, 0), ("energy", 4, 0), ("projection", 1, 0), ("assembly", 2, precision),
                             ("pcg_iteration", 3, precision)]:
        ms = ctypes.c_double()
        symx_amd.check_hip(symx_amd.hip().symx_hip_problem_profile(p, ctypes.c_int32(phase), ctypes.c_int32(0), ctypes.c_int32(arg),
                                                                   ctypes.c_int32(reps), ctypes.byref(ms)))
        out[name] = ms.value
    return out


def cpu_baseline(log, cells=40):
    """The C oracle (OpenMP restatement of the reference path) on a bounded sample:
    one implicit step of a cells^3 block (same material, loads and solver settings)."""
    from oracle_py import Oracle
    from scene_np import tet4_block

    import symx_amd

    orc = Oracle()
    S = tet4_block(cells=(cells,) * 3, noise=0.2 / cells, seed=2)
    newton = symx_amd.NewtonParams(0, 1, 0, 100, 0.0, 1e-6, 1e-3, 1e-12, 10000, 1e-10, 0.0)
    t0 = time.perf_counter()
    res = orc.newton(S, newton)
    t1 = time.perf_counter()
    # element Hessians/s of the evaluate + project + assemble unit
    S2 = tet4_block(cells=(cells,) * 3, noise=0.2 / cells, seed=2)
    t2 = time.perf_counter()
    orc.eval_project_assemble(S2, project=True)
    t3 = time.perf_counter()
    n_elem = S.tets.shape[0]
    log(f"cpu baseline ({orc.num_threads()} threads): {cells}^3 cells, {n_elem} tets: newton={res['newton_iterations']} "
        f"cg={res['cg_iterations']} {t1 - t0:.2f}s; eval+project+assemble {t3 - t2:.2f}s")
    return {
        "newton_per_s": res["newton_iterations"] / (t1 - t0),
        "elem_hessians_per_s": n_elem / (t3 - t2),
        "cores": orc.num_threads(),
        "n_elem": n_elem,
        "seconds": t1 - t0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cells", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import symx_amd

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if symx_amd.device_count() < 1:
        raise RuntimeError("bench.py needs a GPU")
    cells = args.cells
    t0 = time.perf_counter()
    scene = symx_amd.Tet4Block(block_params(cells), symx_amd.default_newton_params(precision=symx_amd.MIXED))
    log(f"setup {cells}^3 cells: {scene.n_elements} tets, {scene.n_vertices} vertices in {time.perf_counter() - t0:.1f}s "
        f"(JIT + topology {symx_amd.amd().symx_scene_compile_seconds(scene.ptr):.1f}s)")
    mixed = run_precision(scene, symx_amd.MIXED, args.steps, args.warmup, log)
    prof_mixed = profile_kernels(scene, symx_amd.MIXED)
    log(f"kernel ms (mixed): {json.dumps({k: round(v, 4) for k, v in prof_mixed.items()})}")
    fp64 = None
    prof64 = None
    if not args.no_fp64:
        scene64 = symx_amd.Tet4Block(block_params(cells), symx_amd.default_newton_params(precision=symx_amd.FP64))
        fp64 = run_precision(scene64, symx_amd.FP64, args.steps, args.warmup, log)
        prof64 = profile_kernels(scene64, symx_amd.FP64)
        log(f"kernel ms (fp64): {json.dumps({k: round(v, 4) for k, v in prof64.items()})}")
        scene64.close()

    import ctypes

    nnzb_c = ctypes.c_int64()
    symx_amd.check_hip(symx_amd.hip().symx_hip_problem_export_bsr(scene.problem, ctypes.byref(nnzb_c), None, None, None))
    nnzb = nnzb_c.value
    n_elem, n_vert = scene.n_elements, scene.n_vertices

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
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "kernel": dominant, "avg_ms": round(prof_mixed[dominant], 4),
                "algorithmic_bytes_per_launch": int(unit_bytes)}
    elem_hessians_per_s = n_elem / (prof_mixed["element"] * 1e-3)

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(log, args.cpu_cells)
        scale = c["n_elem"] / n_elem  # linear extrapolation in element count (optimistic for the CPU)
        cpu = {"value": round(c["newton_per_s"] * scale, 4), "unit": "Newton iterations/s", "cores": c["cores"], "kind": "port",
               "sample": f"oracle/symx_oracle.c, one implicit step of a {args.cpu_cells}^3-cell block ({c['n_elem']} tets, "
                         f"{c['seconds']:.1f}s), scaled linearly by element count to {n_elem} tets",
               "elem_hessians_per_s": round(c["elem_hessians_per_s"], 1)}

    extra = {
        "elem_hessians_per_s": round(elem_hessians_per_s, 1),
        "time_steps_per_s": round(args.steps / mixed["seconds"], 4),
        "newton_iterations": mixed["newton_iterations"],
        "cg_iterations": mixed["cg_iterations"],
        "phase_seconds": {k: round(v, 4) for k, v in mixed["phases"].items()},
        "kernel_ms": {k: round(v, 4) for k, v in prof_mixed.items()},
        "results": mixed["results"],
    }
    if fp64:
        extra["fp64"] = {
            "newton_per_s": round(fp64["newton_iterations"] / fp64["seconds"], 4),
            "seconds": round(fp64["seconds"], 4),
            "newton_iterations": fp64["newton_iterations"],
            "cg_iterations": fp64["cg_iterations"],
            "elem_hessians_per_s": round(n_elem / (prof64["element"] * 1e-3), 1),
            "kernel_ms": {k: round(v, 4) for k, v in prof64.items()},
            "speedup_mixed_over_fp64_wall": round(fp64["seconds"] / mixed["seconds"], 3),
        }
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
        "dtype": "f32 Hessian/BSR/PCG + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"Tet4 Neo-Hookean block {cells + 1}^3 vertices ({n_elem} tets, {3 * n_vert} DoF), implicit Euler dt=0.01, "
                               "projected Newton to 1e-6 rel. residual", "n_elem": n_elem, "n_dof": 3 * n_vert, "nnzb": nnzb},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)
    scene.close()


if __name__ == "__main__":
    main()
