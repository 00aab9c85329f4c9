"""Endpoint corrector coupling on the drop-in cores (the corrector hook).

Follows src/neuralpom/hybrid/sparse_coupling.py:1-82: the network's raw
increment (60 channels: u, v, t, s on the 15 active levels) is scaled, made
budget-neutral, capped by one shared factor per pair, added to both time
levels, and the density is refreshed through the core's equation of state.

On a drop-in core (one that owns a HIP ensemble) ``apply_joint_correction``
runs the fused kernels (pom_joint_correction, five launches, with a
hand-written adjoint) and then the HIP EOS.  For any other core object with
the reference's attributes it composes the same constraints from torch ops
(``budget_project``, ``paired_limit``); that composition restates the
reference's functions one for one and is what the fused kernels are tested
against.  The equation of state and the pressure gradient run the HIP
kernels through the core's functional ``eos`` / ``baropg_of`` (no proxy
core, no state mutation).  Every function is differentiable end to end.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L

OUTPUT_FIELDS = ("u", "v", "t", "s")
LEVELS = 15
OUTPUT_CHANNELS = 4 * LEVELS


def eos(core, t, s):
    """rho of (t, s) by the core's own EOS and ghost-level rule (sparse_coupling.py:10-15)."""
    return core.eos(t, s)


def volume_weights(core, field="t"):
    """Cell volumes of the active levels, wet mask of the field's grid (:18-20)."""
    if field in ("u", "ub"):
        mask = core.dum
    elif field in ("v", "vb"):
        mask = core.dvm
    else:
        mask = core.fsm
    column = core.art * mask * (core.h + core.et)
    return column.unsqueeze(2) * core.dz[:LEVELS].view(1, 1, LEVELS)


def weighted_mse(x, y, w):
    """Volume-weighted mean squared difference (:23-24)."""
    return ((x - y) ** 2 * w).sum() / w.expand_as(x).sum().clamp_min(1e-20)


def budget_project(increment, weights):
    """Remove each level's volume-weighted mean and zero dry cells (:27-32):
    the increment redistributes a tracer within a layer without changing its
    inventory."""
    total = weights.sum(dim=(0, 1), keepdim=True).clamp_min(1e-20)
    layer_mean = (increment * weights).sum(dim=(0, 1), keepdim=True) / total
    return (increment - layer_mean) * (weights > 0)


def paired_limit(increments, scales, cap):
    """One multiplier for a pair of increments so that max |x / scale| <= cap (:35-39)."""
    peak = torch.stack([(x / sc).abs().amax() for x, sc in zip(increments, scales)]).amax()
    factor = torch.clamp(cap / torch.clamp(peak, min=cap), max=1.0)
    return [factor * x for x in increments], factor


def _split(raw, scales):
    return {f: raw[n * LEVELS:(n + 1) * LEVELS].permute(1, 2, 0) * scales[f] for n, f in enumerate(OUTPUT_FIELDS)}


STATE_FIELDS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb")
CORR_GEOM = ("dum", "dvm", "fsm", "art", "h")


class _JointCorrection(torch.autograd.Function):
    """pom_joint_correction on one ensemble: inputs raw (60 planes per member),
    et and the 8 state fields; outputs the 8 corrected fields and diag [M, 3].
    Backward runs pom_joint_correction_adj; the state inputs' cotangents are
    those of the outputs (out = in + increment)."""

    @staticmethod
    def forward(ctx, ens, vgrid, scales, cap, geom, raw, et, *state):
        lib = ens.lib
        n = C.c_size_t()
        L.check(lib.pom_correction_workspace(ens._h, C.byref(n)), "pom_correction_workspace")
        work = torch.empty(n.value, dtype=torch.uint8, device=ens.device)
        out = [torch.empty_like(x) for x in state]
        diag = torch.empty(3 * len(ens.shapes), dtype=ens.dtype, device=ens.device)
        sc = (C.c_double * 4)(*scales)
        gp = L.ptr_array([g.data_ptr() for g in geom] + [et.data_ptr()])
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(lib.pom_joint_correction(ens._h, raw.data_ptr(), gp, vgrid.data_ptr(), sc, float(cap),
                                         work.data_ptr(), L.ptr_array([x.data_ptr() for x in state]),
                                         L.ptr_array([x.data_ptr() for x in out]), diag.data_ptr(), stream),
                "pom_joint_correction")
        ctx.ens, ctx.scales, ctx.cap = ens, tuple(scales), float(cap)
        ctx.save_for_backward(vgrid, raw, et, *geom)
        return tuple(out) + (diag,)

    @staticmethod
    def backward(ctx, *g):
        vgrid, raw, et, *geom = ctx.saved_tensors
        ens = ctx.ens
        lib = ens.lib
        n = C.c_size_t()
        L.check(lib.pom_correction_workspace(ens._h, C.byref(n)), "pom_correction_workspace")
        work = torch.empty(n.value, dtype=torch.uint8, device=ens.device)
        cots = [x.contiguous() if x is not None else None for x in g[:8]]
        cdiag = g[8].contiguous() if g[8] is not None else None
        graw = torch.empty_like(raw)
        get = torch.zeros_like(et) if ctx.needs_input_grad[6] else None
        sc = (C.c_double * 4)(*ctx.scales)
        gp = L.ptr_array([x.data_ptr() for x in geom] + [et.data_ptr()])
        stream = torch.cuda.current_stream(ens.device).cuda_stream
        L.check(lib.pom_joint_correction_adj(ens._h, raw.data_ptr(), gp, vgrid.data_ptr(), sc, ctx.cap,
                                             work.data_ptr(),
                                             L.ptr_array([x.data_ptr() if x is not None else None for x in cots]),
                                             cdiag.data_ptr() if cdiag is not None else None, graw.data_ptr(),
                                             get.data_ptr() if get is not None else None, stream),
                "pom_joint_correction_adj")
        return (None, None, None, None, None, graw, get) + tuple(cots)


def _fused(core):
    return hasattr(core, "_ensemble") and getattr(core, "device", torch.device("cpu")).type == "cuda"


def apply_joint_correction(core, raw, scales, cap=5.0):
    """Apply the corrector's raw output (60, im, jm) to the core (:42-67).

    Velocity increments lose their dz-weighted column mean (the barotropic
    part belongs to the core) and are masked; tracer increments are
    budget-projected per layer; T/S and u/v are each capped by a shared
    factor; levels 15.. (the ghost level) are left as the core made them."""
    if core.kbm1 != LEVELS or tuple(raw.shape) != (OUTPUT_CHANNELS, core.im, core.jm):
        raise ValueError(f"Unexpected grid/output contract: {core.kbm1}, {tuple(raw.shape)}")
    if _fused(core):
        return _apply_joint_correction_fused(core, raw, scales, cap)
    return _apply_joint_correction_composed(core, raw, scales, cap)


def _apply_joint_correction_fused(core, raw, scales, cap):
    ens = core._ensemble()
    sc = [float(scales[f]) for f in OUTPUT_FIELDS]
    geom = [core._flat(n).detach() for n in CORR_GEOM]
    state = [core._flat(n) for n in STATE_FIELDS]
    outs = _JointCorrection.apply(ens, core._vgrid(), sc, cap, geom, core._as_flat(raw), core._flat("et"), *state)
    shape = (core.im, core.jm, core.kb)
    for n, x in zip(STATE_FIELDS, outs[:8]):
        setattr(core, n, x.view(shape))
    core.rho = eos(core, core.t, core.s)
    diag = outs[8]
    return {"ts_limit_factor": diag[0], "t_budget": diag[1], "s_budget": diag[2]}


def _apply_joint_correction_composed(core, raw, scales, cap):
    """The constraints as torch ops on any core object with the reference's attributes."""
    inc = _split(raw, scales)
    dz = core.dz[:LEVELS].view(1, 1, LEVELS)
    for f, mask in (("u", core.dum), ("v", core.dvm)):
        column_mean = (inc[f] * dz).sum(dim=2, keepdim=True) / dz.sum()
        inc[f] = (inc[f] - column_mean) * mask.unsqueeze(2)
    w = volume_weights(core)
    inc["t"], inc["s"] = budget_project(inc["t"], w), budget_project(inc["s"], w)
    (inc["t"], inc["s"]), factor = paired_limit([inc["t"], inc["s"]], [scales["t"], scales["s"]], cap)
    (inc["u"], inc["v"]), _ = paired_limit([inc["u"], inc["v"]], [scales["u"], scales["v"]], cap)
    for f, older in (("u", "ub"), ("v", "vb"), ("t", "tb"), ("s", "sb")):
        for name in (f, older):
            cur = getattr(core, name)
            setattr(core, name, torch.cat((cur[:, :, :LEVELS] + inc[f], cur[:, :, LEVELS:]), dim=2))
    core.rho = eos(core, core.t, core.s)
    wsum = w.sum()
    return {"ts_limit_factor": factor,
            "t_budget": (w * inc["t"]).sum() / wsum,
            "s_budget": (w * inc["s"]).sum() / wsum}


def pressure_velocity_error(core, reference_t, reference_s):
    """Velocity kick of the baroclinic pressure-gradient difference between the
    core's (t, s) and a reference (t, s) (:70-82).  baropg is linear in rho at
    fixed geometry, so it runs on the density difference with zero mean density."""
    drho = eos(core, core.t, core.s) - eos(core, reference_t, reference_s)
    _, px, py = core.baropg_of(drho, torch.zeros_like(core.rmean))
    depth = core.h + core.etf
    du = (depth + torch.roll(depth, 1, 0)) * core.aru
    dv = (depth + torch.roll(depth, 1, 1)) * core.arv
    kick_u = -2.0 * core.dti2 * px[:, :, :LEVELS] / du.unsqueeze(2).clamp_min(1e-20)
    kick_v = -2.0 * core.dti2 * py[:, :, :LEVELS] / dv.unsqueeze(2).clamp_min(1e-20)
    return kick_u * core.dum.unsqueeze(2), kick_v * core.dvm.unsqueeze(2)
