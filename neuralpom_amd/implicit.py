"""Batched semi-implicit (P0-fixed) POM step on the HIP extension, with its adjoint.

``ImplicitEnsemble`` extends the explicit ``Ensemble`` (same member packing
and field tables) with the spectral-preconditioner cache and drives
``pom_implicit_step`` / ``pom_implicit_step_adj`` (include/neuralpom_amd.h).
``ImplicitStep`` is the ``torch.autograd.Function`` of one internal step:
backward recomputes the step (per-step checkpointing) and runs the adjoint
kernels, including the eta-solve adjoint (a second PCG on the area-weighted
transpose).  No reference to a live core is kept.

Reference: src/neuralpom/cores/implicit_p0_fixed/time_loop.py:38-688 (the
step), core.py:24-59 (eta-solve autograd), src/neuralpom/v1/core_spec.py:22-33
(production settings: alpha .333, smoth .15, Shapiro gamma 0, eta relative
tolerance 1e-5, spectral preconditioner).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib as L
from .explicit import Ensemble, StepScalars


@dataclass
class ImplicitScalars(StepScalars):
    """Scalars of the semi-implicit core (explicit defaults + core_spec.py:22-33 + core.py:62-119)."""
    alpha: float = 0.333
    smoth: float = 0.15
    eta_tol: float = 1.0e-12
    eta_rel_tol: float = 1.0e-5
    eta_max_iter: int = 200
    eta_preconditioner: str = "spectral"
    barotropic_shapiro: float = 0.0
    tracer_shapiro: float = 0.0     # numerics/double_gyre.py:78-130 (closed basin)
    velocity_shapiro: float = 0.0   # numerics/double_gyre.py:132-184 (closed basin)
    aam_min: float = 0.0            # Smagorinsky viscosity floor (time_loop.py:186-188)
    tide_wt: float = 0.0                    # (speed pi / 180) t / 3600 of the tide this step
    # capture-safe eta solves (core.py:504-555): exactly eta_max_iter iterations
    # are launched, no host synchronisation; converged members do nothing
    eta_capture: bool = False

    def struct(self) -> L.Params:
        p = super().struct()
        p.core = L.POM_CORE_IMPLICIT
        p.eta_tol = float(self.eta_tol)
        p.eta_rel_tol = float(self.eta_rel_tol)
        p.eta_max_iter = int(self.eta_max_iter)
        pre = str(self.eta_preconditioner).strip().lower()
        if pre not in ("spectral", "jacobi"):
            raise ValueError(f"eta preconditioner must be 'spectral' or 'jacobi', got {self.eta_preconditioner!r}")
        p.eta_precond = L.POM_PRECOND_SPECTRAL if pre == "spectral" else L.POM_PRECOND_JACOBI
        p.gamma_baro = float(self.barotropic_shapiro)
        p.tracer_shapiro = float(self.tracer_shapiro)
        p.velocity_shapiro = float(self.velocity_shapiro)
        p.aam_min = float(self.aam_min)
        p.tide_wt = float(self.tide_wt)
        p.eta_capture = int(bool(self.eta_capture))
        return p


class ImplicitEnsemble(Ensemble):
    """Members of the semi-implicit core; ``prepare(geom, sc)`` builds the
    preconditioner cache once after initialisation (core.py:254-259)."""

    def __init__(self, shapes, dtype=torch.float32, device="cuda"):
        super().__init__(shapes, dtype=dtype, device=device)
        n = C.c_size_t()
        L.check(self.lib.pom_implicit_cache_bytes(self._h, C.byref(n)), "pom_implicit_cache_bytes")
        self.cache = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        self._prepared = False

    def prepare(self, geom, sc: ImplicitScalars):
        prm = sc.struct()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_implicit_prepare(self._h, C.byref(prm), self._geom_ptrs(geom),
                                              self.cache.data_ptr(), self.cache.numel(), stream),
                "pom_implicit_prepare")
        self._prepared = True

    def _workspace(self, prm, adjoint, pool=None):
        """Scratch of one semi-implicit call from ``pool`` (default: the ensemble's own)."""
        n = C.c_size_t()
        L.check(self.lib.pom_implicit_workspace(self._h, C.byref(prm), int(adjoint), C.byref(n)),
                "pom_implicit_workspace")
        pool = self._ws if pool is None else pool
        key = ("imp", adjoint)
        ws = pool.get(key)
        if ws is None or ws.numel() < n.value:
            ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
            pool[key] = ws
        return ws, n.value

    def step(self, state, geom, sc: ImplicitScalars, ramp, vgrid, internal=True, iters=None, ws_pool=None,
             out=None):
        """One internal step; returns (new state dict, vamax).  ``iters``: optional
        int32 device tensor [members] receiving the eta-solve iteration counts;
        ``out``: optional buffers for the new state."""
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the first step")
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        self._check(ins, L.FIELDS, self.numel)
        gp = self._geom_ptrs(geom)
        out = self._outputs(state, out)
        vamax = torch.zeros(1, dtype=self.dtype, device=self.device)
        ws, nb = self._workspace(prm, False, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        rc = self.lib.pom_implicit_step(
            self._h, C.byref(prm), self.cache.data_ptr(), ramp.data_ptr(), vgrid.data_ptr(),
            L.ptr_array([t.data_ptr() for t in ins]), L.ptr_array([out[n].data_ptr() for n in L.FIELDS]),
            gp, int(bool(internal)), vamax.data_ptr(),
            iters.data_ptr() if iters is not None else None, ws.data_ptr(), nb, stream)
        L.check(rc, "pom_implicit_step")
        return out, vamax

    def step_adj(self, state, geom, sc: ImplicitScalars, ramp, vgrid, internal, cot, ws_pool=None):
        """Gradients of the step inputs for output cotangents ``cot`` (dict, may omit fields)."""
        return self._reverse_call("pom_implicit_step_adj", state, None, geom, sc, ramp, vgrid, internal, cot,
                                  ws_pool)

    def step_taped(self, state, geom, sc: ImplicitScalars, ramp, vgrid, internal=True, iters=None, ws_pool=None,
                   out=None):
        """``step`` that keeps its intermediates in the adjoint workspace and its
        outputs for ``step_reverse``."""
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the first step")
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        self._check(ins, L.FIELDS, self.numel)
        gp = self._geom_ptrs(geom)
        out = self._outputs(state, out)
        vamax = torch.zeros(1, dtype=self.dtype, device=self.device)
        ws, nb = self._workspace(prm, True, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        rc = self.lib.pom_implicit_step_taped(
            self._h, C.byref(prm), self.cache.data_ptr(), ramp.data_ptr(), vgrid.data_ptr(),
            L.ptr_array([t.data_ptr() for t in ins]), L.ptr_array([out[n].data_ptr() for n in L.FIELDS]),
            gp, int(bool(internal)), vamax.data_ptr(),
            iters.data_ptr() if iters is not None else None, ws.data_ptr(), nb, stream)
        L.check(rc, "pom_implicit_step_taped")
        return out, vamax

    def step_reverse(self, state, out, geom, sc: ImplicitScalars, ramp, vgrid, internal, cot, ws_pool=None):
        """``step_adj`` of the last ``step_taped`` of the same arguments, from its
        kept intermediates; ``state``, ``out`` and ``geom`` unchanged since."""
        return self._reverse_call("pom_implicit_step_reverse", state, out, geom, sc, ramp, vgrid, internal, cot,
                                  ws_pool)

    def _reverse_call(self, fn, state, out, geom, sc, ramp, vgrid, internal, cot, ws_pool):
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the first step")
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        self._check(ins, L.FIELDS, self.numel)
        gp = self._geom_ptrs(geom)
        grads, cot_ptrs, grad_ptrs = self._adj_buffers(cot)
        ws, nb = self._workspace(prm, True, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        args = [self._h, C.byref(prm), self.cache.data_ptr(), ramp.data_ptr(), vgrid.data_ptr(),
                L.ptr_array([t.data_ptr() for t in ins])]
        if out is not None:
            self._check([out[n] for n in L.FIELDS], L.FIELDS, self.numel)
            args.append(L.ptr_array([out[n].data_ptr() for n in L.FIELDS]))
        args += [gp, int(bool(internal)), cot_ptrs, grad_ptrs, ws.data_ptr(), nb, stream]
        L.check(getattr(self.lib, fn)(*args), fn)
        return grads

    # ---------------------------------------------------------------- eta solve on its own
    def _ws_eta(self):
        n = C.c_size_t()
        L.check(self.lib.pom_eta_workspace(self._h, C.byref(n)), "pom_eta_workspace")
        ws = self._ws.get("eta")
        if ws is None or ws.numel() < n.value:
            ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
            self._ws["eta"] = ws
        return ws, n.value

    def _geom_ptrs(self, geom):
        names = [n for n in L.GEOM_ALL if n in L.GEOM or geom.get(n) is not None]
        self._check([geom[n] for n in names], names, self.numel)
        return L.geom_ptrs(geom)

    def _flat2(self, x, name):
        x = x.detach().to(self.device, self.dtype).contiguous().reshape(-1)
        if x.numel() != self.numel("el"):
            raise ValueError(f"{name}: expected {self.numel('el')} elements, got {x.numel()}")
        return x

    def elliptic_apply(self, geom, sc: ImplicitScalars, eta, d):
        """A(d) eta (implicit_p0_fixed/core.py:591-643) for every member."""
        eta, d = self._flat2(eta, "eta"), self._flat2(d, "d")
        out = torch.empty_like(eta)
        prm = sc.struct()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_elliptic_apply(self._h, C.byref(prm), self._geom_ptrs(geom), eta.data_ptr(),
                                            d.data_ptr(), out.data_ptr(), stream), "pom_elliptic_apply")
        return out

    def elliptic_apply_adj(self, geom, sc: ImplicitScalars, eta, d, g):
        """(A^T g, (d(A eta)/dd)^T g)."""
        eta, d, g = self._flat2(eta, "eta"), self._flat2(d, "d"), self._flat2(g, "cotangent")
        b_eta, b_d = torch.zeros_like(eta), torch.zeros_like(d)
        prm = sc.struct()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_elliptic_apply_adj(self._h, C.byref(prm), self._geom_ptrs(geom), eta.data_ptr(),
                                                d.data_ptr(), g.data_ptr(), b_eta.data_ptr(), b_d.data_ptr(), stream),
                "pom_elliptic_apply_adj")
        return b_eta, b_d

    def eta_solve(self, geom, sc: ImplicitScalars, rhs, d, x0=None, iters=None):
        """x = A(d)^-1 rhs by the batched PCG (core.py:460-589); ``x0`` is the initial
        guess (interior values; None: zero); ``iters`` an optional int32 device
        tensor [members] receiving the iteration counts."""
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the eta solve")
        rhs, d = self._flat2(rhs, "rhs"), self._flat2(d, "d")
        x0 = self._flat2(x0, "initial_guess") if x0 is not None else None
        x = torch.empty_like(rhs)
        prm = sc.struct()
        ws, nb = self._ws_eta()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_eta_solve(self._h, C.byref(prm), self.cache.data_ptr(), self._geom_ptrs(geom),
                                       rhs.data_ptr(), d.data_ptr(), x0.data_ptr() if x0 is not None else None,
                                       x.data_ptr(), iters.data_ptr() if iters is not None else None,
                                       ws.data_ptr(), nb, stream), "pom_eta_solve")
        return x

    def eta_solve_noalloc(self, geom, sc: ImplicitScalars, rhs, d, x, x0=None, iters=None):
        """``eta_solve`` into the caller's ``x`` through pom_eta_solve_noalloc:
        no allocation and no host synchronisation inside the library (its work
        arrays are the library's static per-GPU scratch; exactly
        ``sc.eta_max_iter`` iterations are launched).  Calls that overlap on
        one GPU share that scratch and must not run at the same time."""
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the eta solve")
        rhs, d = self._flat2(rhs, "rhs"), self._flat2(d, "d")
        x0 = self._flat2(x0, "initial_guess") if x0 is not None else None
        if x.dtype != self.dtype or x.device != self.device or not x.is_contiguous() or x.numel() != rhs.numel():
            raise ValueError("x: expected a contiguous tensor of the ensemble's dtype, device and 2-D size")
        prm = sc.struct()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_eta_solve_noalloc(self._h, C.byref(prm), self.cache.data_ptr(), self._geom_ptrs(geom),
                                               rhs.data_ptr(), d.data_ptr(),
                                               x0.data_ptr() if x0 is not None else None, x.data_ptr(),
                                               iters.data_ptr() if iters is not None else None, stream),
                "pom_eta_solve_noalloc")
        return x

    def eta_solve_adj(self, geom, sc: ImplicitScalars, x, d, g, iters=None):
        """Backward of the solve (core.py:24-59): (lambda, -(d(A x)/dd)^T lambda)."""
        if not self._prepared:
            raise RuntimeError("ImplicitEnsemble.prepare(geom, scalars) must run before the eta solve")
        x, d, g = self._flat2(x, "x"), self._flat2(d, "d"), self._flat2(g, "cotangent")
        lam, b_d = torch.empty_like(x), torch.zeros_like(d)
        prm = sc.struct()
        ws, nb = self._ws_eta()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pom_eta_solve_adj(self._h, C.byref(prm), self.cache.data_ptr(), self._geom_ptrs(geom),
                                           x.data_ptr(), d.data_ptr(), g.data_ptr(), lam.data_ptr(), b_d.data_ptr(),
                                           iters.data_ptr() if iters is not None else None, ws.data_ptr(), nb,
                                           stream), "pom_eta_solve_adj")
        return lam, b_d


class EtaSolve(torch.autograd.Function):
    """``_ImplicitEtaSolveFunction`` (core.py:24-59): x = A(d)^-1 rhs; backward
    returns (lambda, -grad_D) from a second PCG.  ctx keeps the ensemble, the
    detached geometry and the scalars, never a core."""

    @staticmethod
    def forward(ctx, ens, geom, sc, rhs, d, x0, iters=None):
        x = ens.eta_solve(geom, sc, rhs, d, x0, iters).view(rhs.shape)
        ctx.ens, ctx.geom, ctx.sc = ens, geom, sc
        ctx.save_for_backward(x, d)
        return x

    @staticmethod
    def backward(ctx, g):
        x, d = ctx.saved_tensors
        lam, b_d = ctx.ens.eta_solve_adj(ctx.geom, ctx.sc, x, d, g)
        return None, None, None, lam.view(x.shape), b_d.view(d.shape), None, None


class Elliptic(torch.autograd.Function):
    """A(d) eta as an autograd node (linear in eta; bilinear with d)."""

    @staticmethod
    def forward(ctx, ens, geom, sc, eta, d):
        out = ens.elliptic_apply(geom, sc, eta, d).view(eta.shape)
        ctx.ens, ctx.geom, ctx.sc = ens, geom, sc
        ctx.save_for_backward(eta, d)
        return out

    @staticmethod
    def backward(ctx, g):
        eta, d = ctx.saved_tensors
        b_eta, b_d = ctx.ens.elliptic_apply_adj(ctx.geom, ctx.sc, eta, d, g)
        return None, None, None, b_eta.view(eta.shape), b_d.view(d.shape)


class ImplicitStep(torch.autograd.Function):
    """One semi-implicit internal step as an autograd node (inputs/outputs in FIELDS order)."""

    @staticmethod
    def forward(ctx, ens, geom, sc, ramp, vgrid, internal, iters, *inputs):
        state = dict(zip(L.FIELDS, inputs))
        out, vamax = ens.step(state, geom, sc, ramp, vgrid, internal, iters)
        ctx.ens, ctx.geom, ctx.sc, ctx.internal = ens, geom, sc, internal
        ctx.save_for_backward(ramp, vgrid, *inputs)
        ctx.mark_non_differentiable(vamax)
        return tuple(out[n] for n in L.FIELDS) + (vamax,)

    @staticmethod
    def backward(ctx, *gouts):
        ramp, vgrid, *inputs = ctx.saved_tensors
        state = dict(zip(L.FIELDS, inputs))
        cot = {n: g for n, g in zip(L.FIELDS, gouts[:len(L.FIELDS)]) if g is not None}
        grads = ctx.ens.step_adj(state, ctx.geom, ctx.sc, ramp, vgrid, ctx.internal, cot)
        return (None,) * 7 + tuple(grads.get(n) for n in L.FIELDS)


def step_autograd(ens, state, geom, sc, ramp, vgrid, internal=True, iters=None):
    outs = ImplicitStep.apply(ens, geom, sc, ramp, vgrid, internal, iters, *[state[n] for n in L.FIELDS])
    return dict(zip(L.FIELDS, outs[:-1])), outs[-1]
