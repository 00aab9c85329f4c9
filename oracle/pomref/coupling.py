"""Eager PyTorch restatement of the corrector hook and the density-75
projection (test infrastructure only: the parity checker of the fused HIP
kernels, never called by the product).

* ``joint_correction``: src/neuralpom/hybrid/sparse_coupling.py:18-67
  (volume_weights, budget_project, paired_limit, apply_joint_correction).
* ``density75``: src/neuralpom/evaluation/density.py:14-64 (eos_gradient,
  constrained_amplitude, apply_density75).

Both work on an oracle ``State`` (explicit.State) and return new tensors
instead of mutating it; the EOS is ``explicit.dens``.  Pinned against the
reference's own outputs in tests/golden/corrector_12x10.npz
(tests/test_oracle_coupling.py).
"""
from __future__ import annotations

import torch

from . import explicit as X

LEVELS = 15
FIELDS = ("u", "v", "t", "s")


def volume_weights(S, et):
    """art fsm (h + et) dz on the 15 active levels (sparse_coupling.py:18-20)."""
    return (S.art * S.fsm * (S.h + et))[:, :, None] * S.dz[:LEVELS][None, None, :]


def paired_limit(xs, scales, cap):
    """One factor min(cap / max(peak, cap), 1) for a pair (:35-39)."""
    peak = torch.stack([(x / s).abs().amax() for x, s in zip(xs, scales)]).amax()
    factor = (cap / peak.clamp_min(cap)).clamp(max=1.0)
    return [factor * x for x in xs], factor


def joint_correction(S, state, raw, scales, cap=5.0):
    """state: dict of u, ub, v, vb, t, tb, s, sb, et, rho.  Returns the
    corrected fields (rho refreshed by the EOS, ghost level kept) and the
    diagnostics (tracer factor, t budget, s budget) (:42-67)."""
    inc = {f: raw[n * LEVELS:(n + 1) * LEVELS].permute(1, 2, 0) * scales[f] for n, f in enumerate(FIELDS)}
    dz = S.dz[:LEVELS].view(1, 1, -1)
    for f, mask in (("u", S.dum), ("v", S.dvm)):
        inc[f] = (inc[f] - (inc[f] * dz).sum(2, keepdim=True) / dz.sum()) * mask[:, :, None]
    w = volume_weights(S, state["et"])
    for f in ("t", "s"):
        mean = (inc[f] * w).sum((0, 1), keepdim=True) / w.sum((0, 1), keepdim=True).clamp_min(1e-20)
        inc[f] = (inc[f] - mean) * (w > 0)
    (inc["t"], inc["s"]), factor = paired_limit([inc["t"], inc["s"]], [scales["t"], scales["s"]], cap)
    (inc["u"], inc["v"]), _ = paired_limit([inc["u"], inc["v"]], [scales["u"], scales["v"]], cap)
    out = {}
    for f, older in (("u", "ub"), ("v", "vb"), ("t", "tb"), ("s", "sb")):
        for name in (f, older):
            x = state[name]
            out[name] = torch.cat((x[:, :, :LEVELS] + inc[f], x[:, :, LEVELS:]), dim=2)
    out["rho"] = X.dens(S, out["t"], out["s"], state["rho"])
    diag = (factor, (w * inc["t"]).sum() / w.sum(), (w * inc["s"]).sum() / w.sum())
    return out, diag


def eos_gradient(S, t, s, rho):
    """d(sum of rho over the active levels)/d(t, s) in float64 (density.py:14-22)."""
    with torch.enable_grad():
        tt = t.detach().double().requires_grad_(True)
        ss = s.detach().double().requires_grad_(True)
        dens = X.dens(S, tt, ss, rho.double())[..., :LEVELS].sum()
        dt, ds = torch.autograd.grad(dens, (tt, ss))
    return dt[..., :LEVELS], ds[..., :LEVELS]


def constrained_amplitude(amplitude, direction, weights):
    """Remove from the amplitude the part whose weighted moments against the
    direction's two components are nonzero, per level, through the
    pseudo-inverse of the 2 x 2 weighted Gram matrix (eigenvalues below 1e-10
    of the largest dropped) (density.py:25-36)."""
    w = weights / weights.sum((0, 1), keepdim=True).clamp_min(1e-30)
    gram = torch.einsum("ijk,ijka,ijkb->kab", w, direction, direction)
    rhs = torch.einsum("ijk,ijka,ijk->ka", w, direction, amplitude)
    lam, basis = torch.linalg.eigh(gram)
    cutoff = lam.amax(-1, keepdim=True) * 1e-10
    inv = torch.where(lam > cutoff, lam.clamp_min(1e-30).reciprocal(), 0.0)
    coef = torch.einsum("kai,ki,kbi,kb->ka", basis, inv, basis, rhs)
    adjusted = amplitude - torch.einsum("ijka,ka->ijk", direction, coef)
    residual = torch.einsum("ijk,ijka,ijk->ka", w, direction, adjusted)
    return adjusted, residual


def density75(S, state, raw, scales, cap=5.0):
    """The joint correction (cap 5), then the density component of the T/S
    increment shrunk to 75 % under the per-level constraints, one shared cap
    on the result, increments on both time levels from the state before the
    correction, and the EOS refresh (density.py:39-64).  Returns the new
    fields and (ts limit factor, density factor, residual max) with NaN for
    the two keys the early return (no T/S change) leaves out."""
    before = {f: state[f] for f in ("t", "tb", "s", "sb")}
    out, diag = joint_correction(S, state, raw, scales, 5.0)
    nan = torch.tensor(float("nan"), dtype=torch.float64)
    sc = torch.tensor([scales["t"], scales["s"]], dtype=torch.float64)
    x = torch.stack([(out[f] - before[f])[..., :LEVELS].double() for f in ("t", "s")], -1) / sc
    if not bool(x.any()):
        return out, (diag[0], nan, nan)
    dt, ds = eos_gradient(S, before["t"], before["s"], state["rho"])
    g = torch.stack((dt, ds), -1) * sc
    direction = g / torch.linalg.vector_norm(g, dim=-1, keepdim=True).clamp_min(1e-30)
    weights = volume_weights(S, state["et"]).double()
    amplitude = (x * direction).sum(-1)
    adjusted, residual = constrained_amplitude(amplitude, direction, weights)
    proposed = (x + (0.75 - 1.0) * adjusted[..., None] * direction) * sc
    (inc_t, inc_s), factor = paired_limit([proposed[..., 0], proposed[..., 1]], list(sc), cap)
    for f, older, delta in (("t", "tb", inc_t), ("s", "sb", inc_s)):
        for name in (f, older):
            x0 = before[name]
            out[name] = torch.cat((x0[..., :LEVELS] + delta.to(x0.dtype), x0[..., LEVELS:]), 2)
    out["rho"] = X.dens(S, out["t"], out["s"], state["rho"])
    return out, (diag[0] * factor, factor, residual.abs().max())
