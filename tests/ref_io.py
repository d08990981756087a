"""Binary input/output layout of oracle/ref_driver's `solve` mode (test infrastructure)."""
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
