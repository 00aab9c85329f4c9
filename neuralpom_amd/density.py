"""The DND density-75 projection on the drop-in cores
(src/neuralpom/evaluation/density.py:14-64).

After the corrector hook (``coupling.apply_joint_correction``, cap 5), the
T/S increment is split along the scaled EOS gradient of the state before the
correction; its density component, less the per-level part fixed by two
weighted moment constraints, is shrunk to 75 %; one cap factor covers the
result, T and S and their history levels become the state before the
correction plus it, and the density is refreshed.  The reference runs this
without gradients (``@torch.no_grad``), and so does this module.

* ``apply_density75``: the whole policy step on a drop-in core, on the fused
  kernels (pom_joint_correction, pom_density75, pom_dens).
* ``density75``: pom_density75 on a batch of members (packed fields).
* ``eos_gradient``: d(sum of rho over the active levels)/d(t, s) through the
  core's EOS adjoint (pom_dens_adj).
* ``constrained_amplitude``: the per-level constraint of density.py:25-36 as
  tensor operations, for callers that hold amplitudes and directions
  themselves (the fused kernel does not call it).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from . import coupling

LEVELS = coupling.LEVELS
D75_GEOM = ("fsm", "art", "h", "et")
D75_STATE = ("t", "tb", "s", "sb")


def density75(ens, prm, vgrid, scales, cap, geom, before, after):
    """pom_density75 on the members of ``ens``.  ``scales`` = (t, s);
    ``geom``: packed fsm, art, h, et; ``before`` / ``after``: packed t, tb,
    s, sb.  Returns the four new fields and diag [M, 3] (float64: factor,
    max constraint residual, 1 where T/S changed)."""
    lib = ens.lib
    n = C.c_size_t()
    L.check(lib.pom_density75_workspace(ens._h, C.byref(n)), "pom_density75_workspace")
    work = torch.empty(n.value, dtype=torch.uint8, device=ens.device)
    out = [torch.empty_like(x) for x in before]
    diag = torch.empty(3 * len(ens.shapes), dtype=torch.float64, device=ens.device)
    sc = (C.c_double * 2)(*[float(x) for x in scales])
    stream = torch.cuda.current_stream(ens.device).cuda_stream
    L.check(lib.pom_density75(ens._h, C.byref(prm), vgrid.data_ptr(), L.ptr_array([g.data_ptr() for g in geom]),
                              sc, float(cap), work.data_ptr(), L.ptr_array([x.data_ptr() for x in before]),
                              L.ptr_array([x.data_ptr() for x in after]), L.ptr_array([x.data_ptr() for x in out]),
                              diag.data_ptr(), stream), "pom_density75")
    return out, diag.view(-1, 3)


def eos_gradient(core, t, s):
    """(drho/dt, drho/ds) on the 15 active levels at (t, s), in float64
    (density.py:14-22: the gradient of the summed density)."""
    with torch.enable_grad():
        tt = t.detach().to(core.dtype).requires_grad_(True)
        ss = s.detach().to(core.dtype).requires_grad_(True)
        rho = core.eos(tt, ss)[..., :LEVELS].sum()
        dt, ds = torch.autograd.grad(rho, (tt, ss))
    return dt[..., :LEVELS].double(), ds[..., :LEVELS].double()


def constrained_amplitude(amplitude, direction, weights):
    """Per level k, remove from ``amplitude`` (i, j, k) its projection on the
    two components of ``direction`` (i, j, k, 2) under the weights normalised
    over (i, j): coefficients from the pseudo-inverse of the weighted Gram
    matrix (eigenvalues at or below 1e-10 of the largest dropped).  Returns
    the adjusted amplitude and the residual moments (k, 2) (density.py:25-36)."""
    w = weights / weights.sum((0, 1), keepdim=True).clamp_min(1e-30)
    gram = torch.einsum("ijk,ijka,ijkb->kab", w, direction, direction)
    rhs = torch.einsum("ijk,ijka,ijk->ka", w, direction, amplitude)
    lam, basis = torch.linalg.eigh(gram)
    keep = lam > lam.amax(-1, keepdim=True) * 1e-10
    inv = torch.where(keep, lam.clamp_min(1e-30).reciprocal(), torch.zeros_like(lam))
    coef = torch.einsum("kai,ki,kbi,kb->ka", basis, inv, basis, rhs)
    adjusted = amplitude - torch.einsum("ijka,ka->ijk", direction, coef)
    return adjusted, torch.einsum("ijk,ijka,ijk->ka", w, direction, adjusted)


@torch.no_grad()
def apply_density75(core, raw, correction_scales):
    """The DND policy step on a drop-in core (density.py:38-64): returns the
    corrector hook's diagnostics with its T/S factor multiplied by the
    projection's, plus ``density_component_limit_factor`` and
    ``density_constraint_residual_max_abs`` (both absent, as in the
    reference, when the hook left T and S unchanged)."""
    if not coupling._fused(core):
        raise RuntimeError("apply_density75 runs on a drop-in core on a HIP device")
    before = [core._flat(n).detach().clone() for n in D75_STATE]
    diag = coupling.apply_joint_correction(core, raw, correction_scales, cap=5.0)
    after = [core._flat(n).detach() for n in D75_STATE]
    geom = [core._flat(n).detach() for n in D75_GEOM]
    sc = (float(correction_scales["t"]), float(correction_scales["s"]))
    out, d = density75(core._ensemble(), core._params(), core._vgrid(), sc, 5.0, geom, before, after)
    shape = (core.im, core.jm, core.kb)
    for n, x in zip(D75_STATE, out):
        setattr(core, n, x.view(shape))
    core.rho = coupling.eos(core, core.t, core.s)
    if float(d[0, 2]) == 0.0:
        return diag
    diag["ts_limit_factor"] = diag["ts_limit_factor"] * d[0, 0].to(diag["ts_limit_factor"].dtype)
    diag["density_component_limit_factor"] = d[0, 0]
    diag["density_constraint_residual_max_abs"] = d[0, 1]
    return diag
