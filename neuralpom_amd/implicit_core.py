"""Drop-in semi-implicit (P0-fixed) POM core: the reference's implicit API on the HIP step.

Mirrors ``neuralpom.cores.implicit_p0_fixed.POM2K_Core``:
  src/neuralpom/cores/implicit_p0_fixed/core.py:62-119   class constants and environment settings
  src/neuralpom/cores/implicit_p0_fixed/base.py:93-101   alpha .333, smoth .15, aam_min 0
  src/neuralpom/cores/implicit_p0_fixed/time_loop.py:38-688
                                                         run_time_loop; FloatingPointError above vmaxl
  src/neuralpom/cores/implicit_p0_fixed/forcing/surface.py:73-84
                                                         double-gyre wind rewritten before every step
  src/neuralpom/cores/implicit_p0_fixed/core.py:775-793  solve_implicit_eta_cg (+ its autograd, :24-59)
  src/neuralpom/cores/implicit_p0_fixed/core.py:641-643  apply_elliptic_operator
and the construction protocol of src/neuralpom/hybrid/core_pool.py:30-48.

The closed-basin boundary, double-gyre forcing and double-gyre numerics
policies (case_policy.py:14-60) are what the HIP step implements, and, in
place of the closed basin, the open-boundary tide policy of the case adapter
(case_adapters/coastal_boundary.py:9-67) installed with
``neuralpom_amd.boundary.install``.  A core whose policy objects, eta solver or
elliptic operator were replaced otherwise cannot run the fused step, and
``run_time_loop`` raises instead of silently ignoring the replacement.  There
is no CPU path.
"""
from __future__ import annotations

import os
from dataclasses import replace

import torch

from . import _lib as L
from .boundary import TideBoundary
from .core import POM2K_Core as _ExplicitCore, _STEP_SCALARS
from .implicit import Elliptic, EtaSolve, ImplicitEnsemble, ImplicitScalars, step_autograd

_FLAG_VALUES = {"0", "1", "false", "true", "off", "on", "no", "yes"}
_POLICY_ATTRS = ("_implicit_boundary_policy", "_implicit_forcing_provider", "_implicit_numerics_policy")


class POM2K_Core(_ExplicitCore):
    """Semi-implicit mode-split POM core (double-gyre problem 4) on the MI355X kernels."""

    NUMERICS_VERSION = "implicit-dg-v2-aligned-1e5-p0-barotropic-coupling-fixed"
    BAROTROPIC_ENDPOINT_TRANSPORT_COEFFICIENT = 0.25
    AUTOGRAD_VERSION = "area-transpose-scaled-adjoint-zero-strain-v2"
    DEFAULT_ETA_SOLVER_REL_TOL = 1e-5
    DEFAULT_ETA_SOLVER_PRECONDITIONER = "spectral"
    DEFAULT_SPATIAL_SHAPIRO_GAMMA = 0.05

    def __init__(self, im, jm, problem_id=4, settings_overrides=None, case_name=None, device=None,
                 dtype=torch.float32):
        super().__init__(im, jm, problem_id=problem_id, settings_overrides=settings_overrides,
                         case_name=case_name, device=device, dtype=dtype)
        flag = os.environ.get("NEURALPOM_IMPLICIT_SPATIAL_SHAPIRO", "0").strip().lower()
        if flag not in _FLAG_VALUES:
            raise ValueError("NEURALPOM_IMPLICIT_SPATIAL_SHAPIRO must be one of 0/1, false/true, off/on, or no/yes")
        self.spatial_shapiro_enabled = flag in {"1", "true", "on", "yes"}
        gamma = self.DEFAULT_SPATIAL_SHAPIRO_GAMMA if self.spatial_shapiro_enabled else 0.0
        self.barotropic_shapiro = gamma
        self.tracer_shapiro = gamma
        self.velocity_shapiro = gamma
        self.eta_solver_rel_tol = float(os.environ.get("NEURALPOM_ETA_SOLVER_REL_TOL",
                                                       str(self.DEFAULT_ETA_SOLVER_REL_TOL)))
        self.eta_solver_preconditioner = os.environ.get("NEURALPOM_ETA_SOLVER_PRECONDITIONER",
                                                        self.DEFAULT_ETA_SOLVER_PRECONDITIONER).strip().lower()
        self._eta_solver_iteration_history: list[int] = []

    def _core_defaults(self) -> None:
        self.smoth = 0.150
        self.alpha = 0.333
        self.aam_min = 0.0

    # ------------------------------------------------------------ kernels
    def _make_ensemble(self):
        return ImplicitEnsemble([(self.im, self.jm)], dtype=self.dtype, device=self.device)

    def _ensemble(self) -> ImplicitEnsemble:
        """The member's ensemble handle with its preconditioner cache built for
        the current geometry and operator coefficient (core.py:121-199, 730-760)."""
        e = super()._ensemble()
        acc = self._accel()
        geo = tuple((t.data_ptr(), t._version) for t in (self.dx, self.dy, self.h))
        ck = (float(self.dti), float(self.grav), float(self.alpha), geo)
        if acc.get("eta_key") != ck:
            e.prepare(self._geom(), self.step_scalars())
            acc["eta_key"] = ck
        return e

    def step_scalars(self, max_iter: int | None = None, tol: float | None = None) -> ImplicitScalars:
        if self.nadv != 1:
            raise NotImplementedError("only the centred advection scheme (nadv = 1) is implemented")
        kw = {n: getattr(self, n) for n in _STEP_SCALARS}
        kw.update(eta_tol=float(getattr(self, "eta_solver_tol", 1.0e-12) if tol is None else tol),
                  eta_rel_tol=float(self.eta_solver_rel_tol),
                  eta_max_iter=int(getattr(self, "eta_solver_max_iter", 200) if max_iter is None else max_iter),
                  eta_preconditioner=self.eta_solver_preconditioner,
                  barotropic_shapiro=float(self.barotropic_shapiro),
                  tracer_shapiro=float(self.tracer_shapiro), velocity_shapiro=float(self.velocity_shapiro),
                  aam_min=float(getattr(self, "aam_min", 0.0) or 0.0))
        tide = self._tide_policy()
        if tide is not None:
            if float(self.tracer_shapiro) != 0.0 or float(self.velocity_shapiro) != 0.0:
                raise NotImplementedError("the tide policy's numerics have no spatial Shapiro filters "
                                          "(tracer_shapiro / velocity_shapiro must be 0)")
            kw.update(boundary=L.POM_BOUNDARY_TIDE, tide_wt=tide.wt(tide.seconds(self)))
        sc = ImplicitScalars(**kw)
        if abs(sc.dti - self.dti) > 1e-9 * abs(self.dti):
            raise ValueError(f"dti {self.dti} != dte * isplit {sc.dti}")
        return sc

    def _tide_policy(self):
        pol = getattr(self, "_implicit_boundary_policy", None)
        return pol if isinstance(pol, TideBoundary) else None

    def _scalars_at_step(self, sc):
        """The tide's time argument at this step's self.time (coastal_boundary.py:36, 45)."""
        tide = self._tide_policy()
        return sc if tide is None else replace(sc, tide_wt=tide.wt(tide.seconds(self)))

    # ------------------------------------------------------------ time loop
    def _before_time_loop(self) -> None:
        for name in _POLICY_ATTRS:
            pol = getattr(self, name, None)
            if pol is not None and not (name == "_implicit_boundary_policy" and isinstance(pol, TideBoundary)):
                raise NotImplementedError(f"{name} replaced: the HIP step implements the closed-basin (or tide) "
                                          "boundary and the double-gyre forcing and numerics policies only")
        for name in ("solve_implicit_eta_cg", "apply_elliptic_operator"):
            if name in vars(self) or getattr(type(self), name) is not getattr(POM2K_Core, name):
                raise NotImplementedError(f"{name} is overridden: the fused step cannot call a replaced operator")
        self._apply_surface_wind()
        self._iters().clear()

    def _iters(self) -> list:
        """Per-step iteration counters (device), kept in the PCG workspace dict."""
        work = getattr(self, "_cg_work", None)
        if not isinstance(work, dict):
            work = {}
            self._cg_work = work
        return work.setdefault("iters", [])

    def _apply_surface_wind(self) -> None:
        """DoubleGyreWindForcing.update_before_step (forcing/surface.py:73-84)."""
        if not hasattr(self, "north_e"):
            return
        tau0 = float(getattr(self, "double_gyre_wind_stress_pa", 0.1))
        arg = 2.0 * float(self.pi) * (self.north_e - 30.0) / 20.0
        self.wusurf = (-(tau0 / self.rhoref) * (1.0 - torch.cos(arg)) * self.dum).to(self.dtype)
        self.wvsurf = torch.zeros_like(self.wvsurf)

    def _after_time_loop(self) -> None:
        its = self._iters()
        if its:
            self._eta_solver_iteration_history.extend(int(x) for x in torch.cat(its).cpu())
        its.clear()

    def _vamax_checked(self) -> bool:
        fast = getattr(self, "generation_fast_path", False) or getattr(self, "runtime_no_grad_fast_path", False)
        return not (fast and getattr(self, "runtime_vamax_safe_fast_path", False))

    def _vamax_exceeded(self, vamax: float) -> None:
        """time_loop.py:685-688."""
        raise FloatingPointError(f"Barotropic velocity {vamax} exceeds limit {self.vmaxl} at time {self.time}")

    def _advance(self, ens, state, geom, sc, ramp, vg, internal):
        iters = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._iters().append(iters)
        if torch.is_grad_enabled() and any(v.requires_grad for v in state.values()):
            return step_autograd(ens, state, geom, sc, ramp, vg, internal, iters)
        with torch.no_grad():
            return ens.step({n: v.detach() for n, v in state.items()}, geom, sc, ramp, vg, internal, iters)

    # ------------------------------------------------------------ operator overrides
    def solve_implicit_eta_cg(self, rhs, d_2d, max_iter=200, tol=1e-7, print_cg_residual=True, initial_guess=None):
        """PCG solve of A(d_2d) eta = rhs (core.py:775-793); differentiable in rhs and d_2d.

        The initial guess defaults to 2 el - elb (core.py:474-477); its edge
        ring is ignored and the solution is zero there.  ``print_cg_residual``
        is accepted for signature compatibility; iteration counts are kept in
        ``_eta_solver_iteration_history``."""
        ens = self._ensemble()
        sc = self.step_scalars(max_iter=max_iter, tol=tol)
        geom = self._geom()
        x0 = (2.0 * self.el - self.elb) if initial_guess is None else initial_guess
        x0 = x0.detach()
        iters = torch.zeros(1, dtype=torch.int32, device=self.device)
        if torch.is_grad_enabled() and (rhs.requires_grad or d_2d.requires_grad):
            x = EtaSolve.apply(ens, geom, sc, rhs, d_2d, x0, iters)
        else:
            with torch.no_grad():
                x = ens.eta_solve(geom, sc, rhs, d_2d, x0, iters).view(rhs.shape)
        self._eta_solver_iteration_history.append(int(iters[0]))
        return x

    def apply_elliptic_operator(self, eta, d_2d):
        """A(d_2d) eta (core.py:591-643); differentiable in eta and d_2d."""
        ens = self._ensemble()
        sc = self.step_scalars()
        geom = self._geom()
        if torch.is_grad_enabled() and (eta.requires_grad or d_2d.requires_grad):
            return Elliptic.apply(ens, geom, sc, eta, d_2d)
        with torch.no_grad():
            return ens.elliptic_apply(geom, sc, eta, d_2d).view(eta.shape)

