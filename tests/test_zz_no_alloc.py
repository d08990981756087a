"""Allocation-free loop evaluation (symx_hip_loop_eval_noalloc) against the allocating
entry point symx_hip_loop_eval, through the C ABI: the same loop kernel on the same inputs
must give bitwise the same outputs, on two calls in a row (the second smaller, with a
different activity mask), and the scratch limit must be enforced on the host."""
import ctypes
import tempfile
from ctypes import c_double, c_int32, c_void_p

import numpy as np
import pytest

import symx_amd
from scene_np import tet4_block

pytestmark = pytest.mark.gpu

LOOP_KERNEL = r"""
extern "C" __global__ void __launch_bounds__(256) tet_volume_loop(const symx_hip_eval_params p, const symx_hip_adjoint_params a)
{
    const int e = symx_block() * 256 + threadIdx.x;
    if (e >= p.n_elem) return;
    double* o = a.out + (long long)e * a.n_out;
    if (p.active != nullptr && p.active[e] == 0) { o[0] = 0.0; o[1] = 0.0; return; }
    double x[4][3];
    for (int c = 0; c < 4; c++) {
        const long long v = p.conn[(long long)c * p.n_elem + e];
        for (int k = 0; k < 3; k++) x[c][k] = p.maps[0][v * 3 + k];
    }
    double u[3], w[3], d[3];
    for (int k = 0; k < 3; k++) { u[k] = x[1][k] - x[0][k]; w[k] = x[2][k] - x[0][k]; d[k] = x[3][k] - x[0][k]; }
    const double vol = (u[0] * (w[1] * d[2] - w[2] * d[1]) - u[1] * (w[0] * d[2] - w[2] * d[0]) + u[2] * (w[0] * d[1] - w[1] * d[0])) / 6.0;
    o[0] = vol * p.maps[1][0];
    o[1] = (double)p.conn[(long long)3 * p.n_elem + e];
}
"""


class MapDesc(ctypes.Structure):
    _fields_ = [("is_dof", c_int32), ("dof_offset", c_int32), ("host", c_void_p), ("n_values", ctypes.c_int64)]


def _module(ctx, cache):
    lib = symx_amd.hip()
    src = lib.symx_hip_eval_prelude().decode() + LOOP_KERNEL
    mod = c_void_p()
    symx_amd.check_hip(lib.symx_hip_module_load(ctx.ptr, src.encode(), cache.encode(), ctypes.byref(mod)), ctx.ptr)
    return mod


def _host_loop(ctx, mod, conn, x, scale, active):
    lib = symx_amd.hip()
    n = conn.shape[0]
    maps = (MapDesc * 2)()
    maps[0] = MapDesc(0, 0, x.ctypes.data, x.size)
    maps[1] = MapDesc(0, 0, scale.ctypes.data, scale.size)
    out = np.zeros((n, 2))
    symx_amd.check_hip(lib.symx_hip_loop_eval(ctx.ptr, mod, b"tet_volume_loop", conn.ctypes.data_as(c_void_p), c_int32(n), c_int32(4),
                                              maps, c_int32(2), active.ctypes.data_as(c_void_p), c_int32(2),
                                              out.ctypes.data_as(c_void_p)), ctx.ptr)
    return out


def _device_loop(ctx, mod, conn, x, scale, active):
    import torch

    lib = symx_amd.hip()
    n = conn.shape[0]
    d_conn = torch.from_numpy(np.ascontiguousarray(conn)).cuda()
    d_x = torch.from_numpy(x).cuda()
    d_s = torch.from_numpy(scale).cuda()
    d_act = torch.from_numpy(active).cuda()
    d_out = torch.full((n, 2), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ptrs = (c_void_p * 2)(d_x.data_ptr(), d_s.data_ptr())
    print(f"loop_eval_noalloc: n_elem={n}", flush=True)
    rc = lib.symx_hip_loop_eval_noalloc(ctx.ptr, mod, b"tet_volume_loop", c_void_p(d_conn.data_ptr()), c_int32(n), c_int32(4), ptrs,
                                        c_int32(2), c_void_p(d_act.data_ptr()), c_int32(2), c_void_p(d_out.data_ptr()))
    symx_amd.check_hip(rc, ctx.ptr)
    return d_out.cpu().numpy()


def test_loop_eval_noalloc_matches_allocating_entry(gpu):
    ctx = symx_amd.Context()
    with tempfile.TemporaryDirectory() as cache:
        mod = _module(ctx, cache)
        S = tet4_block(cells=(9, 8, 7), noise=0.02, seed=4)
        conn = S.tets.astype(np.int32)
        x = np.ascontiguousarray(S.x.reshape(-1))
        rng = np.random.default_rng(12)
        for n, scale in ((conn.shape[0], 2.5), (conn.shape[0] // 3 + 5, -1.25)):
            c = np.ascontiguousarray(conn[:n])
            act = (rng.uniform(size=n) < 0.8).astype(np.uint8)
            s = np.array([scale])
            ref = _host_loop(ctx, mod, c, x, s, act)
            got = _device_loop(ctx, mod, c, x, s, act)
            assert np.array_equal(got, ref)
            # closed form: signed tet volume times the fixed map, last node id; zeros when inactive
            P = S.x[c]
            vol = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6.0
            assert np.allclose(got[:, 0], np.where(act != 0, vol * scale, 0.0), rtol=1e-13, atol=1e-18)
            assert np.array_equal(got[:, 1], np.where(act != 0, c[:, 3], 0).astype(np.float64))
        symx_amd.hip().symx_hip_module_unload(mod)
    ctx.close()


def test_loop_eval_noalloc_rejects_oversized_loops(gpu):
    import torch

    ctx = symx_amd.Context()
    with tempfile.TemporaryDirectory() as cache:
        mod = _module(ctx, cache)
        buf = torch.zeros(16, dtype=torch.int32, device="cuda")
        out = torch.zeros(16, dtype=torch.float64, device="cuda")
        lib = symx_amd.hip()
        limit = 1 << 22
        # rejected on the host before any launch (the buffers are never read)
        rc = lib.symx_hip_loop_eval_noalloc(ctx.ptr, mod, b"tet_volume_loop", c_void_p(buf.data_ptr()), c_int32(limit // 4 + 1), c_int32(4),
                                            None, c_int32(0), None, c_int32(2), c_void_p(out.data_ptr()))
        assert rc == 1
        assert "SYMX_HIP_LOOP_SCRATCH_ENTRIES" in lib.symx_hip_last_error(ctx.ptr).decode()
        symx_amd.hip().symx_hip_module_unload(mod)
    ctx.close()
