"""Write golden vectors for the explicit POM step from the reference implementation.

Runs only in a container that has the reference checkout at /root/reference
(read-only).  The reference's core modules import ``netCDF4`` without using it,
so a stub module is installed before import.  The reference hard-codes float32
scratch allocations (explicit/momentum.py:27,30,146,149,557,558;
explicit/tracers.py:722,723,924; explicit/cases.py:55-56,107); to obtain an
FP64 trajectory the factory functions ``torch.zeros/ones/full/tensor/arange/
linspace`` are wrapped so that an explicit ``dtype=torch.float32`` becomes
float64 while this script runs.  Nothing else in the reference is changed.

Outputs (``tests/golden/*.npz``), all float64:

* ``explicit_init_<im>x<jm>.npz``   -- every tensor attribute after
  ``build_initialized_core`` + ``core_spec('explicit').configure`` (the
  double-gyre case, explicit/cases.py:30-164, explicit/initialization.py:10-99).
* ``explicit_ops_<im>x<jm>.npz``    -- a seeded perturbation of that state and
  the inputs/outputs of the first call of every operator inside one
  ``run_time_loop`` step (advct, baropg, advave, vertvl, advq, profq, advt1,
  proft, dens, advu, advv, profu, profv, bcond 1..6).
* ``explicit_steps_<im>x<jm>.npz``  -- full state after 1 and 3 internal steps
  from the perturbed state (time0 = 0.1 d: ramp < 1), and after 1 step from
  time0 = 0 (the internal-mode skip at explicit/time_loop.py:438).
* ``explicit_vjp_<im>x<jm>.npz``    -- a seeded-cotangent vector-Jacobian
  product of one step (torch.autograd through the reference step).
"""
from __future__ import annotations

import os
import sys
import types
from pathlib import Path

import numpy as np
import torch

REF_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent


def _install_stubs() -> None:
    # the reference tree is read-only: importing it must leave no bytecode there
    sys.dont_write_bytecode = True
    sys.modules.setdefault("netCDF4", types.ModuleType("netCDF4"))
    if str(REF_SRC) not in sys.path:
        sys.path.insert(0, str(REF_SRC))


_F32_FACTORIES = ("zeros", "ones", "full", "tensor", "arange", "linspace", "empty")


def _force_fp64():
    originals = {}
    for name in _F32_FACTORIES:
        fn = getattr(torch, name)
        originals[name] = fn

        def wrapped(*args, __fn=fn, **kwargs):
            if kwargs.get("dtype", None) is torch.float32:
                kwargs["dtype"] = torch.float64
            return __fn(*args, **kwargs)

        setattr(torch, name, wrapped)
    return originals


def _tensors(core):
    return {k: v for k, v in vars(core).items() if torch.is_tensor(v) and v.is_floating_point()}


def _np(t):
    return t.detach().to(torch.float64).cpu().numpy().copy()


def build_core(im, jm):
    from neuralpom.hybrid.core_pool import build_initialized_core
    from neuralpom.v1.core_spec import core_spec

    spec = core_spec("explicit")
    core = build_initialized_core(spec.core_class, torch.device("cpu"), im, jm)
    spec.configure(core)
    return core


# Fields perturbed for the operator and step goldens, with (mean shift, scale,
# positive?) -- chosen so every branch of the step (inflow/outflow boundary
# selection, clipping in the turbulence closure) is exercised.
PERTURB = {
    "u": (0.0, 0.05, False), "ub": (0.0, 0.05, False),
    "v": (0.0, 0.05, False), "vb": (0.0, 0.05, False),
    "t": (0.0, 0.5, False), "tb": (0.0, 0.5, False),
    "s": (0.0, 0.05, False), "sb": (0.0, 0.05, False),
    "el": (0.0, 0.02, False), "elb": (0.0, 0.02, False),
    "et": (0.0, 0.02, False), "etb": (0.0, 0.02, False), "etf": (0.0, 0.02, False),
    "ua": (0.0, 0.03, False), "uab": (0.0, 0.03, False),
    "va": (0.0, 0.03, False), "vab": (0.0, 0.03, False),
    "q2": (0.0, 1e-4, True), "q2b": (0.0, 1e-4, True),
    "q2l": (0.0, 1e-3, True), "q2lb": (0.0, 1e-3, True),
    "km": (0.0, 1e-3, True), "kh": (0.0, 1e-3, True), "kq": (0.0, 1e-3, True),
    "aam": (200.0, 100.0, True), "w": (0.0, 1e-5, False),
    "wubot": (0.0, 1e-5, False), "wvbot": (0.0, 1e-5, False),
    "wvsurf": (0.0, 2e-5, False), "utb": (0.0, 10.0, False), "vtb": (0.0, 10.0, False),
    "egb": (0.0, 0.02, False), "vfluxf": (0.0, 1e-6, False), "vfluxb": (0.0, 1e-6, False),
    "e_atmos": (0.0, 1e-3, False), "swrad": (0.0, 1e-6, False),
    "wtsurf": (0.0, 1e-6, False), "wssurf": (0.0, 1e-6, False),
}


def perturb(core, seed):
    g = torch.Generator().manual_seed(seed)
    for name, (shift, scale, positive) in PERTURB.items():
        x = getattr(core, name)
        noise = torch.randn(x.shape, generator=g, dtype=torch.float64) * scale
        if positive:
            noise = noise.abs()
        setattr(core, name, (x + shift + noise).to(x.dtype))
    core.q2 = core.q2 + core.small
    core.q2b = core.q2b + core.small
    core.d = core.h + core.el
    core.dt = core.h + core.et


# Attributes an operator reads from ``self`` rather than from its arguments.
_BCOND_READS = ("u", "v", "t", "s", "w", "el", "h", "dt", "zz", "dx", "dy", "fsm", "dum",
                "dvm", "q2", "q2l", "ele", "elw", "eln", "els", "uabe", "uabw", "vabn",
                "vabs", "tbe", "tbw", "tbn", "tbs", "sbe", "sbw", "sbn", "sbs")
SELF_READS = {
    "dens": ("t", "s", "h", "zz", "fsm", "rho"),
    "baropg": ("rho", "rmean", "dt", "zz", "dx", "dy", "dum", "dvm"),
    "bcond": _BCOND_READS,
    "advave": ("d",),
    "proft": ("z",),
}

OPS = ("advct", "baropg", "advave", "vertvl", "advq", "profq", "advt1", "proft",
       "dens", "advu", "advv", "profu", "profv", "bcond")


def record_ops(core):
    """Wrap every operator so its first call (first per bcond index) is saved."""
    records = {}

    def wrap(name, fn):
        def inner(*args, **kwargs):
            key = name if name != "bcond" else f"bcond{args[0]}"
            first = key not in records
            if first:
                snap_args = [(_np(a) if torch.is_tensor(a) else a) for a in args]
                # self.* attributes read implicitly (dens, baropg, bcond)
                names = SELF_READS.get(name, ())
                snap_self = {k: _np(getattr(core, k)) for k in names}
            out = fn(*args, **kwargs)
            if first:
                outs = out if isinstance(out, tuple) else (out,)
                records[key] = {
                    "args": snap_args,
                    "self": snap_self,
                    "out": [(_np(o) if torch.is_tensor(o) else o) for o in outs],
                }
            return out
        return inner

    for name in OPS:
        setattr(core, name, wrap(name, getattr(core, name)))
    return records


def flatten_records(records):
    flat = {}
    for key, rec in records.items():
        for i, a in enumerate(rec["args"]):
            if isinstance(a, np.ndarray):
                flat[f"{key}/arg{i}"] = a
            elif isinstance(a, (int, float)):
                flat[f"{key}/arg{i}"] = np.array(a, dtype=np.float64)
        for i, o in enumerate(rec["out"]):
            if isinstance(o, np.ndarray):
                flat[f"{key}/out{i}"] = o
            elif isinstance(o, (int, float)):
                flat[f"{key}/out{i}"] = np.array(o, dtype=np.float64)
        for name, v in rec["self"].items():
            flat[f"{key}/self/{name}"] = v
    return flat


SCALARS = ("dte", "dti", "dte2", "dti2", "isplit", "ispi", "isp2i", "alpha", "smoth",
           "horcon", "grav", "rhoref", "kappa", "umol", "tprni", "small", "vmaxl",
           "tbias", "sbias", "hmax", "period", "ispadv", "ntp", "nbct", "nbcs",
           "rfe", "rfw", "rfn", "rfs", "z0b", "cbcmin", "cbcmax", "aam_init", "slmax")


def scalars(core):
    out = {}
    for k in SCALARS:
        v = getattr(core, k)
        out[f"scalar/{k}"] = np.array(float(v), dtype=np.float64)
    return out


def main(im=12, jm=10, seed=1234):
    _install_stubs()
    import neuralpom.cores.explicit  # noqa: F401  (imports every mixin)
    originals = _force_fp64()
    torch.set_default_dtype(torch.float64)
    try:
        tag = f"{im}x{jm}"
        core = build_core(im, jm)
        init = {f"t/{k}": _np(v) for k, v in _tensors(core).items()}
        init.update(scalars(core))
        np.savez_compressed(OUT / f"explicit_init_{tag}.npz", **init)

        # ---- operator records inside one step from a perturbed state ----
        core = build_core(im, jm)
        perturb(core, seed)
        core.time0 = core.time = 0.1
        state0 = {f"t/{k}": _np(v) for k, v in _tensors(core).items()}
        state0.update(scalars(core))
        records = record_ops(core)
        core.iint = 0
        core.iend = 1
        with torch.no_grad():
            core.run_time_loop()
        ops = flatten_records(records)
        ops.update(state0)
        np.savez_compressed(OUT / f"explicit_ops_{tag}.npz", **ops)

        # ---- full steps ----
        steps = {}
        for label, time0, nstep in (("a", 0.1, 1), ("b", 0.1, 3), ("c", 0.0, 1)):
            core = build_core(im, jm)
            perturb(core, seed)
            core.time0 = core.time = time0
            if label == "a":
                for k, v in _tensors(core).items():
                    steps[f"in/{k}"] = _np(v)
                steps.update(scalars(core))
            core.iint = 0
            core.iend = nstep
            with torch.no_grad():
                core.run_time_loop()
            for k, v in _tensors(core).items():
                steps[f"{label}/{k}"] = _np(v)
            steps[f"{label}/time"] = np.array(core.time)
            steps[f"{label}/time0"] = np.array(time0)
            steps[f"{label}/nstep"] = np.array(nstep)
        np.savez_compressed(OUT / f"explicit_steps_{tag}.npz", **steps)

        # ---- VJP of one step ----
        core = build_core(im, jm)
        perturb(core, seed)
        core.time0 = core.time = 0.1
        leaves = {}
        for name in VJP_INPUTS:
            leaf = getattr(core, name).detach().clone().requires_grad_(True)
            leaves[name] = leaf
            setattr(core, name, leaf.clone())
        core.iint = 0
        core.iend = 1
        core.run_time_loop()
        g = torch.Generator().manual_seed(seed + 1)
        loss = 0.0
        vjp = {}
        for name in VJP_OUTPUTS:
            out = getattr(core, name)
            cot = torch.randn(out.shape, generator=g, dtype=torch.float64)
            vjp[f"cot/{name}"] = _np(cot)
            loss = loss + (out * cot).sum()
        grads = torch.autograd.grad(loss, [leaves[n] for n in VJP_INPUTS], allow_unused=True)
        for name, gr in zip(VJP_INPUTS, grads):
            vjp[f"grad/{name}"] = _np(gr) if gr is not None else np.zeros(leaves[name].shape)
        np.savez_compressed(OUT / f"explicit_vjp_{tag}.npz", **vjp)
    finally:
        for name, fn in originals.items():
            setattr(torch, name, fn)


# Dynamic state read by a step (inputs) and written by it (outputs).
VJP_INPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
              "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
              "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")
VJP_OUTPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
               "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
               "et", "etb", "etf", "rho", "egb", "utb", "vtb", "wubot", "wvbot", "w")


if __name__ == "__main__":
    im = int(os.environ.get("GOLDEN_IM", "12"))
    jm = int(os.environ.get("GOLDEN_JM", "10"))
    main(im, jm)
