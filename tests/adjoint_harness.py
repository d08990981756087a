"""Test harness of the adjoint path: the quasistatic Neo-Hookean cantilever with per-tet
log(mu) parameters, driven on the GPU through the C ABI, and a torch.autograd.Function
that wraps "log_mu -> equilibrium positions" so that torch computes dL/dlog_mu of any
loss of the positions with the GPU adjoint. Test infrastructure only.

Cantilever: a Tet4 block of cells[0] x cells[1] x cells[2] hex cells (6 tets each) and
length L along x, with the x = 0 face held by the penalty boundary potential,
quasistatic gravity -9.81 z, density 1000, lambda_e = (lambda/mu) mu_e with nu = 0.4,
and mu_e = exp(log_mu_e), log_mu ~ N(log(3.5e4), 0.3^2) from a seed.

Finite differences of an equilibrium need re-solves far below the energy resolution
of the line search (a 1e-6 change of one tet's log(mu) moves the energy by about
1e-20 of itself), so re-solves start at the unperturbed equilibrium and take plain
Newton steps on the unprojected Hessian with a tight PCG.
"""
from __future__ import annotations

import numpy as np

import symx_amd

MU_MEAN = np.log(3.5e4)
MU_STD = 0.3
NU = 0.4
LENGTH = 0.1


def log_mu_field(n_elem: int, seed: int) -> np.ndarray:
    return MU_MEAN + MU_STD * np.random.default_rng(seed).standard_normal(n_elem)


def block_params(cells, length=LENGTH):
    cells = list(cells)
    return symx_amd.default_block_params(
        cells=cells, size=[length, length * cells[1] / cells[0], length * cells[2] / cells[0]], youngs_modulus=1e5,
        poissons_ratio=NU, density=1000.0, gravity=[0.0, 0.0, -9.81], dt=0.0, noise=0.0, fixed_axis=0, log_mu_field=1)


def scene_np(cells, length=LENGTH):
    """The same cantilever as numpy arrays for the oracle (mu_e set by the caller)."""
    from scene_np import tet4_block

    cells = tuple(cells)
    S = tet4_block(cells=cells, size=(length, length * cells[1] / cells[0], length * cells[2] / cells[0]), E=1e5, nu=NU,
                   dt=0.0, noise=0.0, fixed_axis=0)
    S.lambda_over_mu = S.lam / S.mu
    return S


def loss(x, x_target):
    d = np.asarray(x).reshape(-1) - np.asarray(x_target).reshape(-1)
    return 0.5 * float(d @ d)


def loss_difference(x_plus, x_minus, x_target):
    """L(x_plus) - L(x_minus) without forming either loss: 0.5 (x+ - x-) . (x+ + x- - 2 x_t).
    The two losses agree to about 1e-10 of themselves in a finite difference, so their
    separate FP64 values would cancel most of the digits of the difference."""
    xp = np.asarray(x_plus).reshape(-1)
    xm = np.asarray(x_minus).reshape(-1)
    return 0.5 * float((xp - xm) @ (xp + xm - 2.0 * np.asarray(x_target).reshape(-1)))


class GpuCantilever:
    """Forward equilibrium, Newton polish and adjoint of the cantilever on the GPU."""

    def __init__(self, cells, precision=symx_amd.FP64, length=LENGTH):
        self.params = block_params(cells, length)
        self.newton = symx_amd.default_newton_params(precision=precision, projection_mode=1, residual_tolerance_rel=1e-4,
                                                     max_iterations=300, cg_max_iterations=100000)
        self.scene = symx_amd.Tet4Block(self.params, self.newton)
        self.precision = precision
        self.X = self.scene.get("X").reshape(-1)

    def equilibrium(self, log_mu, x_start=None, polish_steps=2):
        """Projected Newton from x_start (rest shape by default) to 1e-4 of the first residual
        (the penalty rows' round-off floor is near 1e-5 of it), then plain Newton steps."""
        s = self.scene
        s.set_log_mu(log_mu)
        s.set("x", (self.X if x_start is None else np.asarray(x_start)).reshape(-1, 3))
        st = s.solve()
        x = self.polish(s.get("x").reshape(-1), polish_steps)
        return x, st

    def polish(self, x, steps, cg_rel_tol=1e-13):
        """Plain Newton steps x -= H^-1 g with the unprojected Hessian, all on the device.
        Leaves x in the scene's positions and returns it."""
        s = self.scene
        s.set_dofs(x)
        for _ in range(steps):
            s.eval(symx_amd.EVAL_PGH)
            s.assemble(symx_amd.FP64)
            s.scale(symx_amd.VEC_RHS, symx_amd.VEC_GRAD, -1.0)
            info = s.linear_solve(symx_amd.FP64, 0.0, cg_rel_tol, 200000, stop_on_indef=True)
            if info["found_indefiniteness"]:
                raise RuntimeError("unprojected Hessian is indefinite at the equilibrium")
            s.axpy(symx_amd.VEC_DOFS, symx_amd.VEC_DOFS, 1.0, symx_amd.VEC_DU)
        x = s.get_dofs()
        s.set("x", x.reshape(-1, 3))
        return x

    def resolve(self, log_mu, x_eq, steps=2):
        """Equilibrium for log_mu close to the one of x_eq (finite differences)."""
        self.scene.set_log_mu(log_mu)
        return self.polish(x_eq, steps)

    def adjoint(self, dL_dx, rel_tol=1e-12):
        return self.scene.adjoint(dL_dx, rel_tol=rel_tol)


def torch_equilibrium_function():
    """torch.autograd.Function: log_mu (per tet) -> equilibrium positions [n_vertices, 3]."""
    import torch

    class Equilibrium(torch.autograd.Function):
        @staticmethod
        def forward(ctx, log_mu, model: GpuCantilever):
            x, _ = model.equilibrium(log_mu.detach().cpu().numpy().astype(np.float64))
            ctx.model = model
            return torch.from_numpy(x.reshape(-1, 3).copy()).to(device=log_mu.device, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_x):
            g, st = ctx.model.adjoint(grad_x.detach().cpu().numpy().astype(np.float64).reshape(-1))
            if not st["converged"]:
                raise RuntimeError(f"adjoint PCG did not converge: {st}")
            return torch.from_numpy(g).to(device=grad_x.device, dtype=torch.float64), None

    return Equilibrium


# ---- CPU (oracle) side -----------------------------------------------------------
def oracle_hessian_csr(O, S, x):
    """Unprojected global Hessian of the cantilever (elastic + penalty) as scipy CSR."""
    import scipy.sparse as sp

    He = O.element_hessians(S, x)
    n = S.X.size
    t = S.tets
    dof = (3 * t[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 12)
    rows = np.repeat(dof, 12, axis=1).ravel()
    cols = np.tile(dof, (1, 12)).ravel()
    H = sp.coo_matrix((He.reshape(-1), (rows, cols)), shape=(n, n)).tocsr()
    bc = (3 * S.bc_vertex[:, None] + np.arange(3)).ravel()
    return H + sp.coo_matrix((np.full(bc.size, S.k), (bc, bc)), shape=(n, n)).tocsr()


def oracle_newton_polish(O, S, log_mu, x, steps=3):
    """Plain Newton steps with a sparse direct solve (scipy)."""
    import scipy.sparse.linalg as spla

    S.mu_e = np.exp(np.asarray(log_mu, np.float64))
    x = np.asarray(x, np.float64).reshape(-1).copy()
    for _ in range(steps):
        _, g = O.energy(S, x)
        x -= spla.spsolve(oracle_hessian_csr(O, S, x).tocsc(), g)
    return x


def oracle_equilibrium(O, S, log_mu, x_start=None, polish_steps=3):
    """Projected Newton of the oracle from x_start (rest shape by default), then polish.
    The projected Newton stops at 1e-4 of the first residual: the penalty rows cannot go
    much lower (k |x| eps is about 1e-9 N against gravity loads of 4e-5 N per vertex);
    the plain Newton polish then converges the interior to round-off."""
    S.mu_e = np.exp(np.asarray(log_mu, np.float64))
    S.x[...] = (S.X if x_start is None else np.asarray(x_start)).reshape(S.x.shape)
    O.newton(S, _OracleNewton())
    return oracle_newton_polish(O, S, log_mu, S.x.reshape(-1), polish_steps)


class _OracleNewton:
    precision = 0
    projection_mode = 1
    use_mirroring = 0
    max_iterations = 300
    residual_tolerance_abs = 0.0
    residual_tolerance_rel = 1e-4
    cg_rel_tolerance = 1e-3
    cg_abs_tolerance = 1e-12
    cg_max_iterations = 100000
    projection_eps = 1e-10
    step_tolerance = 0.0
