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

import ctypes as C
import math
import os
from dataclasses import replace
from typing import Any

import torch

from . import _lib as L
from . import cases
from .boundary import TideBoundary
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
        tide = self._tide_policy()
        if tide is not None:
            sc = replace(sc, boundary=L.POM_BOUNDARY_TIDE, tide_coef=tide.coef)
        return sc

    def _tide_policy(self):
        """The tide policy's explicit family installed by
        ``neuralpom_amd.boundary.install`` (the reference replaces ``bcond``
        instead, case_adapters/coastal_boundary.py:62), or None."""
        pol = getattr(self, "_explicit_boundary_policy", None)
        return pol if isinstance(pol, TideBoundary) else None

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
        geom = {n: self._flat(n).detach() for n in L.GEOM}
        tide = self._tide_policy()
        if tide is not None:   # node, amplitude, phase, fixed set and exterior T / S, rebuilt when they change
            acc = self._accel()
            src = tuple((t.data_ptr(), t._version) for t in
                        (self._tide_i, self._tide_j, self._tide_amp, self._tide_phase, self._fixed, self.tbe,
                         self.tbw, self.tbn, self.tbs, self.sbe, self.sbw, self.sbn, self.sbs))
            key = (src, tide.speed, tide.pi, self.dtype)
            if acc.get("tide_key") != key:
                acc["tide_geom"] = {n: self._as_flat(v).detach() for n, v in tide.geometry(self).items()}
                acc["tide_key"] = key
            geom.update(acc["tide_geom"])
        return geom

    # hooks of the time loop (the semi-implicit core overrides them)
    def _scalars_at_step(self, sc):
        """The explicit tide family's clock: the adapter advances its substep
        counter once per substep and never resets it itself
        (coastal_boundary.py:13-14), so a step starts from the counter's
        current value; callers that follow the case registry's one-step
        protocol set ``_tide_substep = 0`` before each step
        (case_adapters/registry.py:57-58)."""
        if self._tide_policy() is None:
            return sc
        n0 = int(self._tide_substep)
        self._tide_substep = n0 + int(sc.isplit)
        return replace(sc, tide_sec=float(self.time) * 86400 - float(self.dti), tide_substep0=n0)

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
            if sc.boundary == L.POM_BOUNDARY_TIDE:
                raise NotImplementedError("the tide policy's explicit family is forward-only (the reference's case "
                                          "registry advances it under torch.no_grad() only, registry.py:39-40)")
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
        double gyre's closed basin.  Differentiable.  With the tide policy
        installed it raises: the adapter's bcond 1 advances its substep
        counter and runs inside ``run_time_loop`` only."""
        if self._tide_policy() is not None:
            raise NotImplementedError("the tide policy's bcond runs inside run_time_loop only")
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

    def profq(self, sm, sh, dh, cc, ee, gg, l, kq, km, kh, uf, vf, q2, q2b, q2lb, a, c,
              h, etf, dti2, umol, dzz, grav, rho, kappa, u, v, dt, small, fsm,
              im, jm, kb, imm1, jmm1, kbm1, tbias, sbias, dz,
              wusurf, wubot, wvsurf, wvbot, t, s, rhoref, zz, z):
        """Mellor-Yamada 2.5 closure, the reference's operator method with its
        positional signature (tracers.py:158-163) and returned tuple
        (tracers.py:431): (sm, sh, dh, cc, ee, gg, l, kq, km, kh, uf, vf, q2b,
        q2lb, a, c), uf / vf being the new q2 / q2l, cc the sound speed on
        levels 0-14, dh, ee and gg the arguments passed in (the reference
        computes h + etf and its sweep coefficients in locals) and q2b, q2lb
        copies of the inputs.  The work arrays passed in are not read.
        Runs the step's HIP kernels (pom_profq); l, kq, km, kh, uf, vf are
        differentiable with respect to uf, vf, q2, q2b, q2lb, km, kh, kq, u, v,
        t, s, rho, wubot, wvbot, etf and dt (pom_profq_adj); the work arrays
        are returned detached."""
        if (im, jm, kb) != (self.im, self.jm, self.kb) or (imm1, jmm1, kbm1) != (im - 1, jm - 1, kb - 1):
            raise ValueError(f"profq: grid {(im, jm, kb)} is not the core's {(self.im, self.jm, self.kb)}")
        F = self._as_flat
        prm = self._params()
        prm.dti2, prm.umol, prm.grav, prm.kappa, prm.small = float(dti2), float(umol), float(grav), float(kappa), \
            float(small)
        prm.tbias, prm.sbias, prm.rhoref = float(tbias), float(sbias), float(rhoref)
        vg = torch.stack([F(z), F(zz), F(dz), F(dzz)]).unsqueeze(0).contiguous()
        ins = [F(x) for x in (uf, vf, q2, q2b, q2lb, km, kh, kq, u, v, t, s, rho)] + \
              [F(x).detach() for x in (wusurf, wvsurf)] + [F(x) for x in (wubot, wvbot, etf, dt)] + \
              [F(x).detach() for x in (h, fsm)]
        outs = _Profq.apply(self._ensemble(), prm, vg, *ins)
        s3, s2 = (self.im, self.jm, self.kb), (self.im, self.jm)
        q2n, q2ln, l_o, kq_o, km_o, kh_o, sm_o, sh_o, cc_o, a_o, c_o, _dh = outs
        kbm1_ = self.kb - 1
        return (sm_o.view(s3), sh_o.view(s3), dh, cc_o.view(s3)[:, :, :kbm1_], ee, gg, l_o.view(s3),
                kq_o.view(s3), km_o.view(s3), kh_o.view(s3), q2n.view(s3), q2ln.view(s3), q2b.clone(), q2lb.clone(),
                a_o.view(s3), c_o.view(s3))

    def proft(self, f, wfsurf, fsurf, nbc, dh, a, c, ee, gg, h, etf, dti2, dz, dzz, swrad, ntp,
              im, jm, kb, kbm1, kbm2, kh, umol):
        """Vertical diffusion of one tracer, the reference's operator method with
        its positional signature (tracers.py:890-892) and returned tuple
        (tracers.py:1010): (f, wfsurf, fsurf, nbc, dh, a, c, ee, gg) with f the
        diffused field, dh = h + etf and a, c, ee, gg the sweep's coefficients.
        Surface conditions nbc 1 (flux) and 3 (prescribed value); 2 and 4
        (shortwave absorption, ntp) raise NotImplementedError.  Runs the step's
        HIP kernel (pom_proft); f is differentiable with respect to f, kh and
        etf (pom_proft_adj)."""
        if (im, jm, kb) != (self.im, self.jm, self.kb) or (kbm1, kbm2) != (kb - 1, kb - 2):
            raise ValueError(f"proft: grid {(im, jm, kb)} is not the core's {(self.im, self.jm, self.kb)}")
        nbc = int(nbc)
        if nbc in (2, 4):
            raise NotImplementedError("proft: surface conditions with shortwave absorption (nbc 2, 4) are not built")
        if nbc not in (1, 3):
            raise ValueError(f"proft: nbc must be 1-4, got {nbc}")
        F = self._as_flat
        prm = self._params()
        prm.dti2, prm.umol = float(dti2), float(umol)
        vg = torch.stack([F(self.z), F(self.zz), F(dz), F(dzz)]).unsqueeze(0).contiguous()
        outs = _Proft.apply(self._ensemble(), prm, vg, nbc, F(f), F(kh), F(wfsurf).detach(), F(fsurf).detach(), F(etf),
                            F(h).detach())
        s3, s2 = (self.im, self.jm, self.kb), (self.im, self.jm)
        fo, ao, co, eeo, ggo, dho = outs
        return fo.view(s3), wfsurf, fsurf, nbc, dho.view(s2), ao.view(s3), co.view(s3), eeo.view(s3), ggo.view(s3)

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


class _Profq(torch.autograd.Function):
    """pom_profq / pom_profq_adj.  Inputs (flat): uf vf q2 q2b q2lb km kh kq u v t s
    rho wusurf wvsurf wubot wvbot etf dt h fsm; outputs: q2 q2l l kq km kh (3-D,
    differentiable) then sm sh cc a c dh (work arrays, not differentiable)."""

    N3_OUT = 11

    @staticmethod
    def forward(ctx, ens, prm, vg, *ins):
        n3, n2 = ins[0].numel(), ins[-1].numel()
        outs = [torch.empty(n3, dtype=ins[0].dtype, device=ins[0].device) for _ in range(_Profq.N3_OUT)]
        outs.append(torch.empty(n2, dtype=ins[0].dtype, device=ins[0].device))
        nb = C.c_size_t()
        L.check(ens.lib.pom_profq_workspace(ens._h, C.byref(nb)), "pom_profq_workspace")
        ws = torch.empty(nb.value, dtype=torch.uint8, device=ens.device)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_profq(ens._h, C.byref(prm), vg.data_ptr(), _ptrs(ins), _ptrs(outs), ws.data_ptr(),
                                  nb.value, stream), "pom_profq")
        ctx.ens, ctx.prm = ens, prm
        ctx.save_for_backward(vg, *ins)
        ctx.mark_non_differentiable(*outs[6:])
        return tuple(outs)

    @staticmethod
    def backward(ctx, *g):
        vg, *ins = ctx.saved_tensors
        ens = ctx.ens
        cots = [x.contiguous() if x is not None else None for x in g[:6]]
        diff = set(range(13)) | {15, 16, 17, 18}   # uf .. rho, wubot, wvbot, etf, dt
        grads = [torch.zeros_like(x) if (q in diff and ctx.needs_input_grad[3 + q]) else None
                 for q, x in enumerate(ins)]
        nb = C.c_size_t()
        L.check(ens.lib.pom_profq_workspace(ens._h, C.byref(nb)), "pom_profq_workspace")
        ws = torch.empty(nb.value, dtype=torch.uint8, device=ens.device)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_profq_adj(ens._h, C.byref(ctx.prm), vg.data_ptr(), _ptrs(ins), _ptrs(cots), _ptrs(grads),
                                      ws.data_ptr(), nb.value, stream), "pom_profq_adj")
        return (None, None, None) + tuple(grads)


class _Proft(torch.autograd.Function):
    """pom_proft / pom_proft_adj.  Inputs (flat): f kh wfsurf fsurf etf h; outputs
    f (differentiable) then a c ee gg dh (work arrays)."""

    @staticmethod
    def forward(ctx, ens, prm, vg, nbc, f, kh, wfsurf, fsurf, etf, h):
        outs = [torch.empty_like(f) for _ in range(5)] + [torch.empty_like(etf)]
        ins = (f, kh, wfsurf, fsurf, etf, h)
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_proft(ens._h, C.byref(prm), vg.data_ptr(), _ptrs(ins), int(nbc), _ptrs(outs), stream),
                "pom_proft")
        ctx.ens, ctx.prm, ctx.nbc = ens, prm, int(nbc)
        ctx.save_for_backward(vg, *ins)
        ctx.mark_non_differentiable(*outs[1:])
        return tuple(outs)

    @staticmethod
    def backward(ctx, g, *_):
        vg, *ins = ctx.saved_tensors
        ens = ctx.ens
        if g is None:
            return (None,) * 10
        g = g.contiguous()
        need = ctx.needs_input_grad
        grads = [torch.zeros_like(ins[0]) if need[4] else None, torch.zeros_like(ins[1]) if need[5] else None,
                 torch.zeros_like(ins[4]) if need[8] else None]
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(ens.lib.pom_proft_adj(ens._h, C.byref(ctx.prm), vg.data_ptr(), _ptrs(ins), ctx.nbc, g.data_ptr(),
                                      _ptrs(grads), stream), "pom_proft_adj")
        return (None, None, None, None, grads[0], grads[1], None, None, grads[2], None)
