"""Drop-in explicit POM core: the reference's dynamical-core API on the HIP step.

Mirrors the reference's explicit ``POM2K_Core``:
  src/neuralpom/cores/explicit/core.py:17-61     class, settings overrides, derived parameters
  src/neuralpom/cores/explicit/base.py:8-189     attribute names and class defaults
  src/neuralpom/cores/explicit/time_loop.py:20-709  run_time_loop (internal-step loop, ramp,
                                                    early return when vamax > vmaxl)
  src/neuralpom/cores/explicit/momentum.py:467-541, 543-700  dens(), baropg()
and the construction protocol of src/neuralpom/hybrid/core_pool.py:30-48
(instantiate, set ``device``, ``configure_core_runtime``, ``run_initialization``).

Every state field is a plain torch tensor attribute, (im, jm, 16) or (im, jm),
on ``self.device``.  ``run_time_loop`` rebinds the attributes to the outputs of
the HIP step (one ``pom_explicit_step`` launch sequence per internal step), so
``vars(core)``-based capture / restore and ``torch.utils.checkpoint`` over
cloned attributes keep working.  When autograd is recording and any state
tensor requires grad, each internal step is one ``ExplicitStep`` node whose
backward recomputes the step and runs the hand-written adjoint kernels; the
node keeps no reference to the core.  There is no CPU path: without the HIP
library or a GPU every entry point raises.
"""
from __future__ import annotations

import math
import os
from typing import Any

import torch

from . import _lib as L
from . import cases
from .explicit import Ensemble, StepScalars, step_autograd

KB = 16
_STEP_SCALARS = ("dte", "isplit", "alpha", "smoth", "horcon", "grav", "rhoref", "kappa", "umol", "tprni",
                 "small", "tbias", "sbias", "hmax", "rfe", "rfw", "rfn", "rfs", "ispadv", "ntp", "nbct", "nbcs",
                 "vmaxl")


class POM2K_Core:
    """Explicit mode-split POM core (double-gyre problem 4) running on the MI355X kernels."""

    NUMERICS_VERSION = "explicit-dg-v2-split-smag-v2align-closure"

    def __init__(self, im: int, jm: int, problem_id: int = 4, settings_overrides: dict[str, Any] | None = None,
                 case_name: str | None = None, device=None, dtype=torch.float32):
        self._requested_problem_id = problem_id
        self._settings_overrides = dict(settings_overrides or {})
        self.case_name = case_name
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else "cpu"
        self.device = torch.device(device)
        self.iproblem = problem_id
        if self.iproblem != 4:
            raise ValueError(f"Only the Double-gyre problem_id=4 is supported; got {self.iproblem}.")
        self.im, self.jm, self.kb = int(im), int(jm), KB
        self._refresh_derived_indices()
        # scalars and defaults (base.py:38-117)
        self.small = 1.e-9
        self.pi = math.atan(1.0) * 4.0
        self.mode = 3
        self.nadv = 1
        self.nitera = 2
        self.sw = 0.50
        self.nread = 0
        self.lramp = True
        self.rhoref = 1035.0
        self.hmax = 2000.0
        self.dte = 30.0
        self.isplit = 30
        self.days = 3650.0
        self.prtd1 = 1.0
        self.prtd2 = 1.0
        self.grav = 9.806
        self.kappa = 0.40
        self.z0b = 0.010
        self.cbcmin = 0.00250
        self.cbcmax = 1.0
        self.horcon = 0.20
        self.tprni = 0.20
        self.umol = 2.e-5
        self.vmaxl = 100.0
        self.slmax = 2.0
        self.kl1 = 2
        self.kl2 = 2
        self.swtch = 99999.0
        self.iskp, self.jskp = 4, 3
        self.ntp = 2
        self.nbct = 1
        self.nbcs = 1
        self.ispadv = 5
        self.smoth = 0.100
        self.alpha = 0.2250
        self.aam_init = 500.0
        self.ramp = 0.0
        self.tatm = 0.0
        self.satm = 0.0
        self.tbias = 0.0
        self.sbias = 0.0
        self.rfe = self.rfw = self.rfn = self.rfs = 0.0
        self.time = 0.0
        self.time0 = 0.0
        self.iint = 0
        self.period = 1.0
        self._refresh_derived_time_params()
        # fields, zero until run_initialization (base.py:119-189)
        st, g = cases._empty(self.im, self.jm, dtype, self.device)
        for n, v in {**st, **g}.items():
            setattr(self, n, v)
        for n in ("z", "zz", "dz", "dzz"):
            setattr(self, n, torch.zeros(KB, dtype=dtype, device=self.device))
        self._core_defaults()
        self._apply_settings_overrides()

    def _core_defaults(self) -> None:
        """Class defaults that differ between the cores (set before settings overrides)."""

    # ------------------------------------------------------------ construction (core.py:28-61)
    @classmethod
    def from_settings(cls, settings) -> "POM2K_Core":
        return cls(im=settings.im, jm=settings.jm, problem_id=settings.problem_id,
                   settings_overrides=settings.overrides, case_name=settings.name)

    def _apply_settings_overrides(self) -> None:
        if not self._settings_overrides:
            return
        for key, value in self._settings_overrides.items():
            setattr(self, key, value)
        self._refresh_derived_indices()
        self._refresh_derived_time_params()

    def _refresh_derived_indices(self) -> None:
        self.imm1, self.imm2 = self.im - 1, self.im - 2
        self.jmm1, self.jmm2 = self.jm - 1, self.jm - 2
        self.kbm1, self.kbm2 = self.kb - 1, self.kb - 2

    def _refresh_derived_time_params(self) -> None:
        self.dti = self.dte * self.isplit
        self.dte2 = self.dte * 2.0
        self.dti2 = self.dti * 2.0
        self.iend = max(int((self.days * 24.0 * 3600.0 / self.dti) + 0.5), 2)
        self.iprint = int((self.prtd1 * 24.0 * 3600.0 / self.dti) + 0.5)
        self.iswtch = int((self.swtch * 24.0 * 3600.0 / self.dti) + 0.5)
        self.ispi = 1.0 / self.isplit
        self.isp2i = 1.0 / (2.0 * self.isplit)

    def configure_runtime(self, dte: float = 20.0, isplit: int = 30) -> None:
        """The production clock (src/neuralpom/hybrid/core_pool.py:19-26)."""
        self.dte = dte
        self.isplit = isplit
        self.dti = self.dte * self.isplit
        self.dte2 = self.dte * 2.0
        self.dti2 = self.dti * 2.0
        self.ispi = 1.0 / self.isplit
        self.isp2i = 1.0 / (2.0 * self.isplit)

    def run_initialization(self) -> None:
        """Double-gyre basin and initial state (cases.py:30-164, initialization.py:10-99)."""
        st, g, vg, period = cases.double_gyre(self.im, self.jm, dtype=self.dtype, device=self.device)
        for n, v in {**st, **g}.items():
            setattr(self, n, v)
        self.z, self.zz, self.dz, self.dzz = (vg[i].clone() for i in range(4))
        self.period = period
        self.horcon = float(os.environ.get("NEURALPOM_HORCON", "0.05"))   # cases.py:152

    def load_case(self, case) -> None:
        """Fields, vertical grid and inertial period of a case built by
        ``neuralpom_amd.cases`` (state, geom, vgrid, period_days), in place of
        the double gyre of ``run_initialization``."""
        st, g, vg, period = case
        if tuple(st["u"].shape[:2]) != (self.im, self.jm):
            raise ValueError(f"case grid {tuple(st['u'].shape[:2])} != core grid {(self.im, self.jm)}")
        dtype = self.dtype
        for n, v in {**st, **g}.items():
            setattr(self, n, v.to(self.device, dtype))
        self.z, self.zz, self.dz, self.dzz = (vg[i].to(self.device, dtype).clone() for i in range(4))
        self.period = float(period)

    # ------------------------------------------------------------ kernels
    @property
    def dtype(self) -> torch.dtype:
        """Arithmetic type of the step: that of the state tensors."""
        return self.u.dtype

    def _accel(self) -> dict:
        """The kernel handles live in ``_accel_cache``, a name the reference's
        capture / restore treats as rebuildable workspace (v1/checkpoint.py:17-18,
        93-95), so ``vars(core)`` holds physical state only."""
        acc = getattr(self, "_accel_cache", None)
        if not isinstance(acc, dict):
            acc = {}
            self._accel_cache = acc
        return acc

    def _ensemble(self) -> Ensemble:
        acc = self._accel()
        key = (self.im, self.jm, self.dtype, str(self.device))
        if acc.get("key") != key:
            acc.clear()
            acc["key"] = key
            acc["ensemble"] = self._make_ensemble()
        return acc["ensemble"]

    def _make_ensemble(self):
        return Ensemble([(self.im, self.jm)], dtype=self.dtype, device=self.device)

    def step_scalars(self) -> StepScalars:
        if self.nadv != 1:
            raise NotImplementedError("only the centred advection scheme (nadv = 1) is implemented")
        sc = StepScalars(**{n: getattr(self, n) for n in _STEP_SCALARS})
        if abs(sc.dti - self.dti) > 1e-9 * abs(self.dti):
            raise ValueError(f"dti {self.dti} != dte * isplit {sc.dti}")
        return sc

    def _vgrid(self):
        return torch.stack([self.z, self.zz, self.dz, self.dzz]).to(self.device, self.dtype).reshape(1, 4, KB) \
            .detach().contiguous()

    def _flat(self, n):
        return self._as_flat(getattr(self, n))

    def run_time_loop(self):
        """Advance ``iend`` internal steps from ``time0`` (time_loop.py:20-709).

        Sets ``iint``, ``time`` and ``ramp`` as the reference does.  When the
        depth-mean velocity bound ``vmaxl`` is exceeded it prints
        "Abnormal end" and returns before ``iend`` (time_loop.py:702-709).
        As there, the check belongs to the diagnostics branch that
        ``generation_fast_path`` / ``runtime_no_grad_fast_path`` skip
        (time_loop.py:30-34, 702): with either flag the loop runs to ``iend``.
        """
        self._before_time_loop()
        ens = self._ensemble()
        sc = self.step_scalars()
        vg = self._vgrid()
        geom = self._geom()
        check = self._vamax_checked()
        shapes = {n: tuple(getattr(self, n).shape) for n in L.FIELDS}
        for iint in range(1, int(self.iend) + 1):
            self.iint = iint
            self.time = self.dti * iint * 1.0 / 86400.0 + self.time0
            if self.lramp:
                self.ramp = min(self.time / float(self.period), 1.0)
            else:
                self.ramp = 1.0
            ramp = torch.tensor([self.ramp], dtype=self.dtype, device=self.device)
            internal = not (iint == 1 and self.time0 == 0.0)
            state = {n: self._flat(n) for n in L.FIELDS}
            out, vamax = self._advance(ens, state, geom, self._scalars_at_step(sc), ramp, vg, internal)
            for n in L.FIELDS:
                setattr(self, n, out[n].view(shapes[n]))
            if check and not float(vamax) <= self.vmaxl:
                self._after_time_loop()
                self._vamax_exceeded(float(vamax))
                return
        self._after_time_loop()

    def _geom(self):
        return {n: self._flat(n).detach() for n in L.GEOM}

    # hooks of the time loop (the semi-implicit core overrides them)
    def _scalars_at_step(self, sc):
        return sc

    def _before_time_loop(self) -> None:
        pass

    def _after_time_loop(self) -> None:
        pass

    def _vamax_checked(self) -> bool:
        """time_loop.py:30-34, 702: only the path with diagnostics checks vamax."""
        fast = getattr(self, "generation_fast_path", False) or getattr(self, "runtime_no_grad_fast_path", False)
        return not fast

    def _vamax_exceeded(self, vamax: float) -> None:
        """time_loop.py:702-709: print and return before iend."""
        print("Abnormal end")

    def _advance(self, ens, state, geom, sc, ramp, vg, internal):
        if torch.is_grad_enabled() and any(v.requires_grad for v in state.values()):
            return step_autograd(ens, state, geom, sc, ramp, vg, internal)
        with torch.no_grad():
            return ens.step({n: v.detach() for n, v in state.items()}, geom, sc, ramp, vg, internal)

    # ------------------------------------------------------------ operator overrides
    def dens(self):
        """Equation of state into ``self.rho`` (momentum.py:467-541); returns (s, t, rho)."""
        self.rho = self.eos(self.t, self.s)
        return self.s, self.t, self.rho

    def eos(self, t, s):
        """The core's equation of state for given t, s without touching the core:
        levels 0-14 from (t, s), level 15 kept from ``self.rho`` (the ghost rule
        of dens, as src/neuralpom/hybrid/sparse_coupling.py:10-15 evaluates it)."""
        rho = _Dens.apply(self._ensemble(), self._params(), self._vgrid(), self._as_flat(t), self._as_flat(s),
                          self._flat("rho"), self._flat("h").detach(), self._flat("fsm").detach())
        return rho.view(self.im, self.jm, KB)

    def baropg(self):
        """Baroclinic pressure gradient (momentum.py:543-700); returns (rho, drhox, drhoy)."""
        rho, drhox, drhoy = self.baropg_of(self.rho)
        self.rho = rho
        return self.rho, drhox, drhoy

    def baropg_of(self, rho, rmean=None):
        """baropg for a given density (and mean density, default ``self.rmean``)
        without touching the core; returns (rho, drhox, drhoy)."""
        ramp = torch.tensor([float(self.ramp)], dtype=self.dtype, device=self.device)
        rm = self._flat("rmean") if rmean is None else self._as_flat(rmean)
        out = _Baropg.apply(self._ensemble(), self._params(), ramp, self._vgrid(), self._as_flat(rho), rm.detach(),
                            self._flat("dt"), *(self._flat(n).detach() for n in ("dx", "dy", "dum", "dvm")))
        shape = (self.im, self.jm, KB)
        return tuple(x.view(shape) for x in out)

    def bcond(self, idx, elf, uaf, vaf, uf, vf, w):
        """One lateral boundary condition (baroclinic.py:496-928) on the HIP
        boundary operators; returns (elf, uaf, vaf, uf, vf, w) with the fields
        of ``idx`` replaced, reading the core's el and ramp (idx 2), u and v
        (3, 4, 6), t, s, w and dt (4), q2 and q2l (6) as the reference does.  The
        exterior values of open edges (ele, uabe, tbe, ...) are zero: the
        double gyre's closed basin.  Differentiable."""
        F = self._as_flat
        if idx == 1:
            fields = [F(elf)]
        elif idx == 2:
            fields = [F(uaf), F(vaf), F(self.el)]
        elif idx == 3:
            fields = [F(uf), F(vf), F(self.u), F(self.v)]
        elif idx == 4:
            fields = [F(uf), F(vf), F(self.t), F(self.s), F(self.u), F(self.v), F(self.w), F(self.dt)]
        elif idx == 5:
            fields = [F(w)]
        elif idx == 6:
            fields = [F(uf), F(vf), F(self.q2), F(self.q2l), F(self.u), F(self.v)]
        else:
            raise ValueError(f"bcond index {idx} is not one of 1-6")
        ramp = torch.tensor([float(self.ramp)], dtype=self.dtype, device=self.device)
        outs = _Bcond.apply(self._ensemble(), self._params(), ramp, self._vgrid(), int(idx), self._geom(), *fields)
        if torch.is_tensor(outs):
            outs = (outs,)
        shape3, shape2 = (self.im, self.jm, self.kb), (self.im, self.jm)
        if idx == 1:
            elf = outs[0].view(shape2)
        elif idx == 2:
            uaf, vaf = outs[0].view(shape2), outs[1].view(shape2)
        elif idx == 5:
            w = outs[0].view(shape3)
        else:
            uf, vf = outs[0].view(shape3), outs[1].view(shape3)
        return elf, uaf, vaf, uf, vf, w

    def _as_flat(self, v):
        if v.dtype != self.dtype or v.device != self.device:
            v = v.to(self.device, self.dtype)
        return v.contiguous().reshape(-1)

    def _params(self):
        return self.step_scalars().struct()


def _ptrs(ts):
    return L.ptr_array([t.data_ptr() if t is not None else None for t in ts])


class _Dens(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ens, prm, vg, t, s, rho, h, fsm):
        out = torch.empty_like(rho)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_dens(ens._h, prm, vg.data_ptr(), t.data_ptr(), s.data_ptr(), rho.data_ptr(),
                                 h.data_ptr(), fsm.data_ptr(), out.data_ptr(), stream), "pom_dens")
        ctx.ens, ctx.prm = ens, prm
        ctx.save_for_backward(vg, t, s, rho, h, fsm)
        return out

    @staticmethod
    def backward(ctx, g):
        vg, t, s, rho, h, fsm = ctx.saved_tensors
        ens = ctx.ens
        g = g.contiguous()
        bt, bs, br = torch.zeros_like(t), torch.zeros_like(s), torch.zeros_like(rho)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_dens_adj(ens._h, ctx.prm, vg.data_ptr(), t.data_ptr(), s.data_ptr(), rho.data_ptr(),
                                     h.data_ptr(), fsm.data_ptr(), g.data_ptr(), bt.data_ptr(), bs.data_ptr(),
                                     br.data_ptr(), stream), "pom_dens_adj")
        return None, None, None, bt, bs, br, None, None


class _Bcond(torch.autograd.Function):
    """pom_bcond / pom_bcond_adj: inputs the idx-specific field list, outputs
    the replaced fields."""

    @staticmethod
    def forward(ctx, ens, prm, ramp, vg, idx, geom, *fields):
        nout = 1 if idx in (1, 5) else 2
        outs = [torch.empty_like(fields[q]) for q in range(nout)]
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_bcond(ens._h, prm, ramp.data_ptr(), vg.data_ptr(), idx, _ptrs(fields),
                                  L.geom_ptrs(geom), _ptrs(outs), stream), "pom_bcond")
        ctx.ens, ctx.prm, ctx.idx, ctx.geom = ens, prm, idx, geom
        ctx.save_for_backward(ramp, vg, *fields)
        return tuple(outs) if nout > 1 else outs[0]

    @staticmethod
    def backward(ctx, *g):
        ramp, vg, *fields = ctx.saved_tensors
        ens = ctx.ens
        cots = [x.contiguous() if x is not None else None for x in g]
        grads = [torch.zeros_like(f) if ctx.needs_input_grad[6 + q] else None for q, f in enumerate(fields)]
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_bcond_adj(ens._h, ctx.prm, ramp.data_ptr(), vg.data_ptr(), ctx.idx, _ptrs(fields),
                                      L.geom_ptrs(ctx.geom), _ptrs(cots), _ptrs(grads), stream), "pom_bcond_adj")
        return (None,) * 6 + tuple(grads)


class _Baropg(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ens, prm, ramp, vg, rho, rmean, dt, dx, dy, dum, dvm):
        outs = [torch.empty_like(rho) for _ in range(3)]
        ins = (rho, rmean, dt, dx, dy, dum, dvm)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_baropg(ens._h, prm, ramp.data_ptr(), vg.data_ptr(), _ptrs(ins), _ptrs(outs), stream),
                "pom_baropg")
        ctx.ens, ctx.prm = ens, prm
        ctx.save_for_backward(ramp, vg, *ins)
        return tuple(outs)

    @staticmethod
    def backward(ctx, g_rho, g_x, g_y):
        ramp, vg, *ins = ctx.saved_tensors
        ens = ctx.ens
        cot = [None if g is None else g.contiguous() for g in (g_rho, g_x, g_y)]
        b_rho, b_dt = torch.zeros_like(ins[0]), torch.zeros_like(ins[2])
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_baropg_adj(ens._h, ctx.prm, ramp.data_ptr(), vg.data_ptr(), _ptrs(ins), _ptrs(cot),
                                       _ptrs((b_rho, b_dt)), stream), "pom_baropg_adj")
        return None, None, None, None, b_rho, None, b_dt, None, None, None, None
