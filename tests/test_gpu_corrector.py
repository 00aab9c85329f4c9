"""The fused corrector hook (pom_joint_correction and its adjoint) against the
same constraints composed from torch ops (coupling._apply_joint_correction_composed,
a one-for-one restatement of src/neuralpom/hybrid/sparse_coupling.py:18-67,
itself pinned by the invariant tests of test_coupling.py).

Float64; raw increments below and above the cap (the factor then depends on
the peak element); outputs, the reference's diagnostics and gradients of the
raw output, of et and of the state; a ragged batch of two members in one
launch against each member's composition.
"""
import pytest
import torch

from neuralpom_amd import coupling as CP
from neuralpom_amd.core import POM2K_Core
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu

SCALES = {"u": 0.05, "v": 0.04, "t": 0.2, "s": 0.05}
OUT = CP.STATE_FIELDS + ("rho",)


def _core(im, jm, seed):
    core = POM2K_Core(im, jm, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    gen = torch.Generator().manual_seed(seed)
    for n, a in (("t", 0.5), ("s", 0.05), ("u", 0.02), ("v", 0.02), ("et", 0.01)):
        x = getattr(core, n)
        setattr(core, n, x + a * torch.randn(x.shape, generator=gen, dtype=torch.float64).cuda())
    core.tb, core.sb = core.t.clone(), core.s.clone()
    core.rho = core.eos(core.t, core.s)
    return core


def _raw(im, jm, amp, seed):
    gen = torch.Generator().manual_seed(seed)
    return (amp * torch.randn(CP.OUTPUT_CHANNELS, im, jm, generator=gen, dtype=torch.float64)).cuda()


def _run(core, raw, cap, fused, leaves=()):
    vals = {n: getattr(core, n).clone().requires_grad_(n in leaves) for n in CP.STATE_FIELDS + ("et",)}
    for n, v in vals.items():
        setattr(core, n, v)
    r = raw.clone().requires_grad_("raw" in leaves)
    f = CP._apply_joint_correction_fused if fused else CP._apply_joint_correction_composed
    diag = f(core, r, SCALES, cap)
    return {n: getattr(core, n) for n in OUT}, diag, dict(vals, raw=r)


@pytest.mark.parametrize("amp", [0.5, 40.0])
def test_fused_matches_composed(amp):
    print(f"[gpu] fused corrector hook vs composition, amplitude {amp}", flush=True)
    im, jm, cap = 26, 22, 5.0
    raw = _raw(im, jm, amp, 3)
    leaves = ("raw", "et", "u", "t", "sb")
    outs, grads = [], []
    for fused in (True, False):
        core = _core(im, jm, 7)
        out, diag, lv = _run(core, raw, cap, fused, leaves)
        gen = torch.Generator().manual_seed(11)
        loss = sum((out[n] * torch.randn(out[n].shape, generator=gen, dtype=torch.float64).cuda()).sum()
                   for n in OUT)
        loss = loss + 3.0 * diag["ts_limit_factor"] + 1e3 * diag["t_budget"] - 2e3 * diag["s_budget"]
        grads.append(torch.autograd.grad(loss, [lv[n] for n in leaves]))
        outs.append((out, diag))
    (fo, fd), (co, cd) = outs
    if amp > 10.0:
        assert float(cd["ts_limit_factor"]) < 1.0, "the cap must be active in this case"
    for n in OUT:
        assert rel_err(fo[n].detach().cpu(), co[n].detach().cpu()) < 1e-12, n
    # the budgets are zero up to rounding (the projection is budget-neutral):
    # compared on the scale of the increments, not relative to themselves
    for k in ("ts_limit_factor", "t_budget", "s_budget"):
        a, b = float(fd[k]), float(cd[k])
        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-3), (k, a, b)
    bad = [f"{n}: {rel_err(a.cpu(), b.cpu()):.2e}" for n, a, b in zip(leaves, grads[0], grads[1])
           if not rel_err(a.cpu(), b.cpu()) <= 1e-10]
    assert not bad, bad


def test_ragged_batch_matches_per_member():
    """Two members of different shapes in one launch of the fused kernels."""
    from neuralpom_amd.explicit import Ensemble
    print("[gpu] fused corrector hook, ragged batch (26x22, 19x31)", flush=True)
    shapes, cap = ((26, 22), (19, 31)), 5.0
    cores = [_core(im, jm, 20 + m) for m, (im, jm) in enumerate(shapes)]
    raws = [_raw(im, jm, 30.0 * (m + 1), 40 + m) for m, (im, jm) in enumerate(shapes)]
    ens = Ensemble(list(shapes), dtype=torch.float64, device="cuda")
    pk = {n: ens.pack([getattr(c, n) for c in cores], n) for n in CP.STATE_FIELDS + ("et",) + CP.CORR_GEOM}
    vg = torch.stack([c._vgrid()[0] for c in cores]).contiguous()
    raw = torch.cat([r.reshape(-1) for r in raws]).requires_grad_(True)
    et = pk["et"].clone().requires_grad_(True)
    state = [pk[n].clone().requires_grad_(True) for n in CP.STATE_FIELDS]
    sc = [SCALES[f] for f in CP.OUTPUT_FIELDS]
    outs = CP._JointCorrection.apply(ens, vg, sc, cap, [pk[n] for n in CP.CORR_GEOM], raw, et, *state)
    gen = torch.Generator().manual_seed(5)
    cot = [torch.randn(o.shape, generator=gen, dtype=torch.float64).cuda() for o in outs]
    g = torch.autograd.grad(sum((o * c).sum() for o, c in zip(outs, cot)), [raw, et] + state)
    off = 0
    for m, (c, r) in enumerate(zip(cores, raws)):
        im, jm = shapes[m]
        out, diag, lv = _run(c, r, cap, False, ("raw", "et") + CP.STATE_FIELDS)
        ref_cot = {n: ens.view(cot[q], m, n) for q, n in enumerate(CP.STATE_FIELDS)}
        dg = cot[8][3 * m:3 * m + 3]
        loss = sum((out[n] * ref_cot[n]).sum() for n in CP.STATE_FIELDS) + \
            dg[0] * diag["ts_limit_factor"] + dg[1] * diag["t_budget"] + dg[2] * diag["s_budget"]
        rg = torch.autograd.grad(loss, [lv["raw"], lv["et"]] + [lv[n] for n in CP.STATE_FIELDS])
        for q, n in enumerate(CP.STATE_FIELDS):
            assert rel_err(ens.view(outs[q], m, n).detach().cpu(), out[n].detach().cpu()) < 1e-12, (m, n)
        for q, k in enumerate(("ts_limit_factor", "t_budget", "s_budget")):
            assert abs(float(outs[8][3 * m + q]) - float(diag[k])) <= 1e-12 * max(abs(float(diag[k])), 1e-3)
        n2 = im * jm
        assert rel_err(g[0][60 * off:60 * (off + n2)].view(60, im, jm).cpu(), rg[0].cpu()) < 1e-10, m
        assert rel_err(ens.view(g[1], m, "et").cpu(), rg[1].cpu()) < 1e-10, m
        for q, n in enumerate(CP.STATE_FIELDS):
            assert rel_err(ens.view(g[2 + q], m, n).cpu(), rg[2 + q].cpu()) < 1e-12, (m, n)
        off += n2
