"""Binary input/output layouts of oracle/ref_driver's `solve` and `adjoint` modes (test infrastructure)."""
import numpy as np


def write_solve_input(path, S, newton, threads=1):
    hdr = np.array([S.X.shape[0], S.tets.shape[0], S.bc_vertex.shape[0], 1 if S.dt > 0 else 0, threads,
                    newton.projection_mode, newton.use_mirroring, newton.max_iterations], np.int32)
    prm = np.array([S.lam, S.mu, S.dt, S.k, *S.gravity, newton.residual_tolerance_abs, newton.residual_tolerance_rel,
                    newton.cg_rel_tolerance, newton.cg_abs_tolerance, newton.projection_eps], np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(prm.tobytes())
        f.write(S.tets.astype(np.int32).tobytes())
        for a in (S.X, S.x, S.x0, S.v0):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(S.mass.astype(np.float64).tobytes())
        f.write(S.bc_vertex.astype(np.int32).tobytes())
        f.write(np.ascontiguousarray(S.bc_target, np.float64).tobytes())


def read_solve_output(path, nv):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:16], np.int32)
    E, secs = np.frombuffer(raw[16:32], np.float64)
    x = np.frombuffer(raw[32:32 + 24 * nv], np.float64).reshape(nv, 3).copy()
    return {"result": int(hdr[0]), "newton_iterations": int(hdr[1]), "cg_iterations": int(hdr[2]), "ls_bt_iterations": int(hdr[3]),
            "energy": float(E), "seconds": float(secs), "x": x}


def write_adjoint_input(path, S, log_mu, dL_dx, threads=1):
    """S: SceneData of the quasistatic cantilever with S.x the equilibrium."""
    hdr = np.array([S.X.shape[0], S.tets.shape[0], S.bc_vertex.shape[0], threads], np.int32)
    prm = np.array([S.lambda_over_mu, S.k, *S.gravity], np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(prm.tobytes())
        f.write(S.tets.astype(np.int32).tobytes())
        for a in (S.X, S.x):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(S.mass.astype(np.float64).tobytes())
        f.write(S.bc_vertex.astype(np.int32).tobytes())
        f.write(np.ascontiguousarray(S.bc_target, np.float64).tobytes())
        f.write(np.ascontiguousarray(log_mu, np.float64).tobytes())
        f.write(np.ascontiguousarray(dL_dx, np.float64).reshape(-1).tobytes())


def read_adjoint_output(path, nv, ne):
    raw = open(path, "rb").read()
    ok = int(np.frombuffer(raw[:4], np.int32)[0])
    E = float(np.frombuffer(raw[4:12], np.float64)[0])
    g = np.frombuffer(raw[12:12 + 8 * ne], np.float64).copy()
    lam = np.frombuffer(raw[12 + 8 * ne:12 + 8 * ne + 24 * nv], np.float64).copy()
    return {"ok": ok, "energy": E, "dL_dlog_mu": g, "lambda": lam}


def write_hex_element_input(path, nn, cases):
    """cases: list of (lam, mu, x [nn,3], X [nn,3])."""
    with open(path, "wb") as f:
        f.write(np.array([nn, len(cases)], np.int32).tobytes())
        for lam, mu, x, X in cases:
            f.write(np.array([lam, mu], np.float64).tobytes())
            f.write(np.ascontiguousarray(x, np.float64).reshape(-1).tobytes())
            f.write(np.ascontiguousarray(X, np.float64).reshape(-1).tobytes())


def read_hex_element_output(path, nn, n_cases):
    n = 3 * nn
    raw = np.frombuffer(open(path, "rb").read(), np.float64).reshape(n_cases, 1 + n + n * n)
    return raw[:, 0].copy(), raw[:, 1:1 + n].copy(), raw[:, 1 + n:].reshape(n_cases, n, n).copy()


def write_solve_hex_input(path, S, newton, threads=1):
    """S: SceneData of a hexahedral block (scene_np.hex_block), per-element mu in S.mu_e."""
    nn = S.tets.shape[1]
    hdr = np.array([S.X.shape[0], S.tets.shape[0], S.bc_vertex.shape[0], 1 if S.dt > 0 else 0, threads, newton.projection_mode,
                    newton.use_mirroring, newton.max_iterations, nn], np.int32)
    prm = np.array([S.lambda_over_mu, S.dt, S.k, *S.gravity, newton.residual_tolerance_abs, newton.residual_tolerance_rel,
                    newton.cg_rel_tolerance, newton.cg_abs_tolerance, newton.projection_eps], np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(prm.tobytes())
        f.write(S.tets.astype(np.int32).tobytes())
        for a in (S.X, S.x, S.x0, S.v0):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        f.write(S.mass.astype(np.float64).tobytes())
        f.write(np.ascontiguousarray(S.mu_e, np.float64).tobytes())
        f.write(S.bc_vertex.astype(np.int32).tobytes())
        f.write(np.ascontiguousarray(S.bc_target, np.float64).tobytes())
