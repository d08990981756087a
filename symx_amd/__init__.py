"""ctypes bindings of the two C ABIs of this package.

* ``include/symx_hip.h``    - device runtime (JIT, element evaluation, projection,
  assembly, PCG, Newton vector ops), library ``lib/libsymx_hip.so``.
* ``include/symx_scenes.h`` - example problems built on the C++ front end
  (GlobalPotential / NewtonsMethod), library ``lib/libsymx_amd.so``.

Loading fails loudly when the libraries are missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

FP64 = 0
MIXED = 1
EVAL_P, EVAL_PG, EVAL_PGH, EVAL_PGH_F32 = 0, 1, 2, 3
VEC_DOFS, VEC_GRAD, VEC_DU, VEC_DOFS_LS, VEC_RHS = 0, 1, 2, 3, 4

SOLVER_RETURN = [
    "Successful", "Running", "InvalidInitialState", "TooManyIterations", "TooManyArmijoIterations",
    "LinearSystemSolveFailure", "TooManyInvalidIntermediateIterations", "StepDoesNotDescend", "InvalidConvergedState",
]


class PcgInfo(Structure):
    _fields_ = [("n_iterations", c_int32), ("converged", c_int32), ("found_indefiniteness", c_int32),
                ("n_refinements", c_int32), ("error", c_double)]


class Tet4BlockParams(Structure):
    _fields_ = [("cells", c_int32 * 3), ("size", c_double * 3), ("youngs_modulus", c_double), ("poissons_ratio", c_double),
                ("density", c_double), ("gravity", c_double * 3), ("dt", c_double), ("noise", c_double), ("seed", c_uint64),
                ("bc_stiffness", c_double), ("fixed_axis", c_int32), ("material", c_int32), ("log_mu_field", c_int32),
                ("renumber_seed", c_uint64)]


class SphereParams(Structure):
    _fields_ = [("radius", c_double), ("cells", c_int32), ("jitter", c_double), ("seed", c_uint64), ("center", c_double * 3),
                ("velocity", c_double * 3), ("youngs_modulus", c_double), ("poissons_ratio", c_double), ("density", c_double),
                ("gravity", c_double * 3), ("dt", c_double), ("material", c_int32), ("contact_stiffness", c_double), ("dhat", c_double),
                ("step_filter", c_double)]


class HexBlockParams(Structure):
    _fields_ = [("cells", c_int32 * 3), ("size", c_double * 3), ("nodes_per_element", c_int32), ("youngs_min", c_double),
                ("youngs_max", c_double), ("poissons_ratio", c_double), ("density", c_double), ("gravity", c_double * 3), ("dt", c_double),
                ("noise", c_double), ("seed", c_uint64), ("bc_stiffness", c_double), ("fixed_axis", c_int32)]


class NewtonParams(Structure):
    _fields_ = [("precision", c_int32), ("projection_mode", c_int32), ("use_mirroring", c_int32), ("max_iterations", c_int32),
                ("residual_tolerance_abs", c_double), ("residual_tolerance_rel", c_double), ("cg_rel_tolerance", c_double),
                ("cg_abs_tolerance", c_double), ("cg_max_iterations", c_int32), ("projection_eps", c_double),
                ("step_tolerance", c_double), ("verbosity", c_int32), ("linear_solver", c_int32)]


class SolveStats(Structure):
    _fields_ = [("result", c_int32), ("newton_iterations", c_int32), ("cg_iterations", c_int32), ("ls_bt_iterations", c_int32),
                ("n_refinements", c_int32), ("final_energy", c_double), ("final_residual", c_double), ("first_residual", c_double),
                ("seconds", c_double), ("seconds_eval", c_double), ("seconds_project", c_double), ("seconds_assembly", c_double),
                ("seconds_linear", c_double), ("seconds_energy", c_double), ("ls_trapezoid_accepts", c_int32), ("pad_", c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["result_name"] = SOLVER_RETURN[self.result] if 0 <= self.result < len(SOLVER_RETURN) else str(self.result)
        return d


class AdjointStats(Structure):
    _fields_ = [("cg_iterations", c_int32), ("n_refinements", c_int32), ("converged", c_int32), ("found_indefiniteness", c_int32),
                ("cg_error", c_double), ("energy", c_double), ("seconds", c_double)]


_hip = None
_amd = None
_DP = POINTER(c_double)


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"symx_amd: {path} is missing; build it with __graft_entry__.build() (no CPU fallback exists)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def hip():
    """The device runtime library (include/symx_hip.h)."""
    global _hip
    if _hip is None:
        lib = _load("libsymx_hip.so")
        lib.symx_hip_last_error.restype = c_char_p
        lib.symx_hip_eval_prelude.restype = c_char_p
        lib.symx_hip_ctx_destroy.restype = None
        lib.symx_hip_problem_destroy.restype = None
        lib.symx_hip_module_unload.restype = None
        _hip = lib
    return _hip


def amd():
    """The C++ front end + scenes library (include/symx_scenes.h)."""
    global _amd
    if _amd is None:
        hip()
        lib = _load("libsymx_amd.so")
        lib.symx_scene_last_error.restype = c_char_p
        lib.symx_scene_problem.restype = c_void_p
        lib.symx_scene_compile_seconds.restype = c_double
        lib.symx_scene_destroy.restype = None
        lib.symx_newton_params_default.restype = None
        lib.symx_tet4_block_params_default.restype = None
        lib.symx_sphere_params_default.restype = None
        lib.symx_hex_block_params_default.restype = None
        _amd = lib
    return _amd


def device_count() -> int:
    n = c_int(0)
    hip().symx_hip_device_count(ctypes.byref(n))
    return n.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def check_hip(rc, ctx=None):
    if rc != 0:
        raise RuntimeError(f"symx_hip error {rc}: {hip().symx_hip_last_error(ctx).decode()}")


def check_scene(rc):
    if rc != 0:
        raise RuntimeError(f"symx scene error: {amd().symx_scene_last_error().decode()}")


def default_newton_params(**kw) -> NewtonParams:
    p = NewtonParams()
    amd().symx_newton_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def default_sphere_params(**kw) -> SphereParams:
    p = SphereParams()
    amd().symx_sphere_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("center", "velocity", "gravity"):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


def default_hex_params(**kw) -> HexBlockParams:
    p = HexBlockParams()
    amd().symx_hex_block_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("cells", "size", "gravity"):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


def host_hex_element(params: HexBlockParams, e: int = 0, structured: bool = True):
    """(E, gradient, Hessian, structured?) of element e of a hexahedral block, evaluated on the
    host from the generated expressions (no GPU)."""
    n = 3 * params.nodes_per_element
    E = c_double()
    g = np.zeros(n)
    H = np.zeros(n * n)
    ws = c_int32()
    check_scene(amd().symx_host_hex_element(ctypes.byref(params), c_int32(int(structured)), c_int32(e), ctypes.byref(E), _ptr(g), _ptr(H),
                                            ctypes.byref(ws)))
    return E.value, g, H.reshape(n, n), bool(ws.value)


def host_hex_jit(params: HexBlockParams, cache_dir: str):
    n, cs, ns = c_int32(), c_int64(), c_int32()
    check_scene(amd().symx_host_hex_jit(ctypes.byref(params), cache_dir.encode(), ctypes.byref(n), ctypes.byref(cs), ctypes.byref(ns)))
    return n.value, cs.value, ns.value


def host_sphere_mesh(params: SphereParams):
    """(X [nv,3], tets [ne,4], contact vertices) of the sphere scene's mesh (no GPU)."""
    nv, ne, nc = c_int32(), c_int32(), c_int32()
    check_scene(amd().symx_host_sphere_mesh(ctypes.byref(params), ctypes.byref(nv), ctypes.byref(ne), ctypes.byref(nc), None, None, None))
    X = np.zeros((nv.value, 3))
    T = np.zeros((ne.value, 4), np.int32)
    C = np.zeros(nc.value, np.int32)
    check_scene(amd().symx_host_sphere_mesh(ctypes.byref(params), ctypes.byref(nv), ctypes.byref(ne), ctypes.byref(nc), _ptr(X), _ptr(T), _ptr(C)))
    return X, T, C


def default_block_params(**kw) -> Tet4BlockParams:
    p = Tet4BlockParams()
    amd().symx_tet4_block_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("cells", "size", "gravity"):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


class Context:
    """A HIP context of the device runtime."""

    def __init__(self, device: int = 0):
        self.ptr = c_void_p()
        check_hip(hip().symx_hip_ctx_create(c_int(device), ctypes.byref(self.ptr)))

    def close(self):
        if self.ptr:
            hip().symx_hip_ctx_destroy(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        self.close()


class Tet4Block:
    """Tet4 elastic block scene (include/symx_scenes.h) driven through the C++ front end."""

    def __init__(self, params: Tet4BlockParams | None = None, newton: NewtonParams | None = None, _sphere: SphereParams | None = None,
                 _hex: HexBlockParams | None = None):
        self.newton = newton or default_newton_params()
        self.ptr = c_void_p()
        self.sphere = _sphere
        self.hex = _hex
        if _sphere is not None:
            self.params = None
            check_scene(amd().symx_scene_sphere_create(ctypes.byref(_sphere), ctypes.byref(self.newton), ctypes.byref(self.ptr)))
        elif _hex is not None:
            self.params = None
            check_scene(amd().symx_scene_hex_block_create(ctypes.byref(_hex), ctypes.byref(self.newton), ctypes.byref(self.ptr)))
        else:
            self.params = params or default_block_params()
            check_scene(amd().symx_scene_tet4_block_create(ctypes.byref(self.params), ctypes.byref(self.newton), ctypes.byref(self.ptr)))
        n = c_int32()
        amd().symx_scene_n_vertices(self.ptr, ctypes.byref(n))
        self.n_vertices = n.value
        amd().symx_scene_n_elements(self.ptr, ctypes.byref(n))
        self.n_elements = n.value
        amd().symx_scene_nodes_per_element(self.ptr, ctypes.byref(n))
        self.nodes_per_element = n.value

    def close(self):
        if self.ptr:
            amd().symx_scene_destroy(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        self.close()

    @property
    def problem(self):
        return c_void_p(amd().symx_scene_problem(self.ptr))

    def connectivity(self) -> np.ndarray:
        c = np.zeros((self.n_elements, self.nodes_per_element), np.int32)
        amd().symx_scene_connectivity(self.ptr, _ptr(c))
        return c

    def get(self, name: str) -> np.ndarray:
        shape = (self.n_vertices,) if name == "mass" else (self.n_vertices, 3)
        a = np.zeros(shape)
        check_scene(amd().symx_scene_get_array(self.ptr, name.encode(), _ptr(a)))
        return a

    def set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check_scene(amd().symx_scene_set_array(self.ptr, name.encode(), _ptr(a)))

    def contact_vertices(self) -> np.ndarray:
        n = c_int32()
        check_scene(amd().symx_scene_contact_vertices(self.ptr, ctypes.byref(n), None))
        out = np.zeros(n.value, np.int32)
        check_scene(amd().symx_scene_contact_vertices(self.ptr, ctypes.byref(n), _ptr(out)))
        return out

    def set_log_mu(self, log_mu: np.ndarray):
        a = np.ascontiguousarray(log_mu, dtype=np.float64)
        check_scene(amd().symx_scene_set_log_mu(self.ptr, _ptr(a)))

    def set_newton(self, newton: NewtonParams):
        self.newton = newton
        check_scene(amd().symx_scene_set_newton_params(self.ptr, ctypes.byref(newton)))

    def solve(self) -> dict:
        st = SolveStats()
        check_scene(amd().symx_scene_solve(self.ptr, ctypes.byref(st)))
        return st.as_dict()

    def step(self) -> dict:
        st = SolveStats()
        check_scene(amd().symx_scene_step(self.ptr, ctypes.byref(st)))
        return st.as_dict()

    def adjoint(self, dL_dx: np.ndarray, rel_tol: float = 1e-10, max_iter: int = 100000):
        """dL/dlog_mu (one value per tet) at the current positions, an equilibrium of the
        quasistatic energy: H lambda = dL_dx on the unprojected Hessian, then
        dL/dlog_mu_e = -lambda^T d^2E/dx dlog_mu_e. Returns (gradient, stats)."""
        rhs = np.ascontiguousarray(dL_dx, dtype=np.float64).reshape(-1)
        if rhs.size != 3 * self.n_vertices:
            raise ValueError("dL_dx must have one entry per DoF")
        out = np.zeros(self.n_elements)
        st = AdjointStats()
        check_scene(amd().symx_scene_adjoint(self.ptr, _ptr(rhs), _ptr(out), c_double(rel_tol), c_int32(max_iter), ctypes.byref(st)))
        return out, {k: getattr(st, k) for k, _ in st._fields_}

    def source(self, i: int = 0) -> str:
        n = c_int64()
        amd().symx_scene_source(self.ptr, c_int32(i), None, c_int64(0), ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        amd().symx_scene_source(self.ptr, c_int32(i), buf, c_int64(n.value + 1), ctypes.byref(n))
        return buf.value.decode()

    def op_counts(self, i: int = 0):
        a = (c_int64 * 3)()
        amd().symx_scene_op_counts(self.ptr, c_int32(i), a)
        return list(a)

    # ---- direct access to the device problem (include/symx_hip.h) -------------------
    def _p(self):
        return self.problem

    def eval(self, mode: int = EVAL_PGH) -> float:
        E = c_double()
        check_hip(hip().symx_hip_problem_eval(self._p(), c_int32(mode), ctypes.byref(E)))
        return E.value

    def gradient_dot(self, vec: int) -> float:
        """g(x) . vec with the gradient at the current device DoFs (VEC_GRAD untouched)."""
        out = c_double()
        check_hip(hip().symx_hip_problem_gradient_dot(self._p(), c_int32(vec), ctypes.byref(out)))
        return out.value

    def energy_change(self) -> float:
        """Sum over elements of E_e(current DoFs) - E_e(DoFs of the last EVAL_PG/EVAL_PGH)."""
        dE = c_double()
        check_hip(hip().symx_hip_problem_energy_change(self._p(), ctypes.byref(dE)))
        return dE.value

    def set_dofs(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        check_hip(hip().symx_hip_problem_set_dofs(self._p(), _ptr(x)))

    def download(self, vec: int) -> np.ndarray:
        a = np.zeros(3 * self.n_vertices)
        check_hip(hip().symx_hip_problem_download(self._p(), c_int32(vec), _ptr(a)))
        return a

    def upload(self, vec: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        check_hip(hip().symx_hip_problem_upload(self._p(), c_int32(vec), _ptr(a)))

    def get_dofs(self) -> np.ndarray:
        a = np.zeros(3 * self.n_vertices)
        check_hip(hip().symx_hip_problem_get_dofs(self._p(), _ptr(a)))
        return a

    def axpy(self, dst: int, a: int, s: float, b: int):
        """Device vectors: dst = a + s * b."""
        check_hip(hip().symx_hip_problem_axpy(self._p(), c_int32(dst), c_int32(a), c_double(s), c_int32(b)))

    def scale(self, dst: int, a: int, s: float):
        """Device vectors: dst = s * a."""
        check_hip(hip().symx_hip_problem_scale(self._p(), c_int32(dst), c_int32(a), c_double(s)))

    def element_outputs(self, pot: int, n_elem: int, nb: int):
        G = np.zeros((n_elem, 3 * nb))
        H = np.zeros((n_elem, nb * (nb + 1) // 2 * 9))
        check_hip(hip().symx_hip_problem_element_outputs(self._p(), c_int32(pot), _ptr(G), _ptr(H)))
        return G, H

    def project(self, eps: float = 1e-10, mirror: bool = False) -> int:
        n = c_int64()
        check_hip(hip().symx_hip_problem_project(self._p(), c_double(eps), c_int32(int(mirror)), ctypes.byref(n)))
        return n.value

    def assemble(self, precision: int = FP64):
        check_hip(hip().symx_hip_problem_assemble(self._p(), c_int32(precision)))

    def export_bsr(self):
        nnzb = c_int64()
        check_hip(hip().symx_hip_problem_export_bsr(self._p(), ctypes.byref(nnzb), None, None, None))
        rp = np.zeros(self.n_vertices + 1, np.int64)
        cols = np.zeros(nnzb.value, np.int32)
        vals = np.zeros((nnzb.value, 3, 3))
        check_hip(hip().symx_hip_problem_export_bsr(self._p(), ctypes.byref(nnzb), _ptr(rp), _ptr(cols), _ptr(vals)))
        return rp, cols, vals

    def linear_solve(self, precision: int, abs_tol: float, rel_tol: float, max_iter: int = 10000, stop_on_indef: bool = True) -> dict:
        info = PcgInfo()
        check_hip(hip().symx_hip_problem_solve(self._p(), c_int32(precision), c_double(abs_tol), c_double(rel_tol),
                                               c_int32(max_iter), c_int32(int(stop_on_indef)), ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}


class HexBlock(Tet4Block):
    """Hex8 / Hex27 block scene with per-element Young's moduli (include/symx_scenes.h)."""

    def __init__(self, params: HexBlockParams | None = None, newton: NewtonParams | None = None):
        super().__init__(None, newton, _hex=params or default_hex_params())

    def young_moduli(self) -> np.ndarray:
        out = np.zeros(self.n_elements)
        check_scene(amd().symx_scene_young_moduli(self.ptr, _ptr(out)))
        return out


class Sphere(Tet4Block):
    """Ball dropped on the ground plane (contact scene of include/symx_scenes.h)."""

    def __init__(self, params: SphereParams | None = None, newton: NewtonParams | None = None):
        super().__init__(None, newton, _sphere=params or default_sphere_params())


def host_tet4_element(x12, X12, a, b, material=0):
    """E, gradient, Hessian of the Tet4 elastic potential from the host expression evaluator (no GPU)."""
    x12 = np.ascontiguousarray(x12, np.float64).reshape(12)
    X12 = np.ascontiguousarray(X12, np.float64).reshape(12)
    E = c_double()
    g = np.zeros(12)
    H = np.zeros(144)
    check_scene(amd().symx_host_tet4_element(c_int32(material), _ptr(x12), _ptr(X12), c_double(a), c_double(b), ctypes.byref(E),
                                             _ptr(g), _ptr(H)))
    return E.value, g, H.reshape(12, 12)


def host_point_triangle(in13):
    a = np.ascontiguousarray(in13, np.float64).reshape(13)
    out = np.zeros(157)
    check_scene(amd().symx_host_point_triangle(_ptr(a), _ptr(out)))
    return out


def host_jit_tet4(cache_dir: str, material: int = 0):
    cs = c_int64()
    ops = c_int64()
    check_scene(amd().symx_host_jit_tet4(cache_dir.encode(), c_int32(material), ctypes.byref(cs), ctypes.byref(ops)))
    return cs.value, ops.value


def host_scene_jit(params: Tet4BlockParams, cache_dir: str):
    """Generate and hiprtc-compile every potential module of a scene (no GPU)."""
    n = c_int32()
    cs = c_int64()
    ops = c_int64()
    check_scene(amd().symx_host_scene_jit(ctypes.byref(params), cache_dir.encode(), ctypes.byref(n), ctypes.byref(cs), ctypes.byref(ops)))
    return n.value, cs.value, ops.value


def host_scene_element(params: Tet4BlockParams, potential: int, e: int, chain_rule: bool = True):
    """(E, gradient, Hessian, factorised?) of element e of a potential of a Tet4 block scene,
    evaluated on the host from the generated expressions (no GPU)."""
    E = ctypes.c_double()
    g = np.zeros(36)
    H = np.zeros(36 * 36)
    n = c_int32()
    fac = c_int32()
    check_scene(amd().symx_host_scene_element(ctypes.byref(params), c_int32(int(chain_rule)), c_int32(potential), c_int32(e),
                                              ctypes.byref(E), _ptr(g), _ptr(H), ctypes.byref(n), ctypes.byref(fac)))
    k = n.value
    return E.value, g[:k].copy(), H[: k * k].reshape(k, k).copy(), bool(fac.value)


def host_scene_adjoint_element(params: Tet4BlockParams, log_mu: np.ndarray, e: int, lam: np.ndarray) -> np.ndarray:
    """Generated adjoint outputs of element e evaluated on the host expression graph."""
    log_mu = np.ascontiguousarray(log_mu, np.float64)
    lam = np.ascontiguousarray(lam, np.float64).reshape(-1)
    out = np.zeros(1)
    check_scene(amd().symx_host_scene_adjoint_element(ctypes.byref(params), _ptr(log_mu), c_int32(e), _ptr(lam), _ptr(out)))
    return out


def project_dense(ctx: Context, mats_dev_ptr: int, n_mats: int, n: int, eps: float, mirror: bool, changed_dev_ptr: int | None = None):
    check_hip(hip().symx_hip_project_dense(ctx.ptr, c_void_p(mats_dev_ptr), c_int64(n_mats), c_int32(n), c_double(eps),
                                           c_int32(int(mirror)), c_void_p(changed_dev_ptr) if changed_dev_ptr else None), ctx.ptr)
