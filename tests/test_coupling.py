"""The corrector hook (neuralpom_amd.coupling) on the drop-in core.

Pinned by what the constraints guarantee, not by retyping them:
  * budget_project: zero volume-weighted layer mean, dry cells zero, idempotent;
  * paired_limit: factor 1 below the cap, peak exactly the cap above it, one
    factor for both members of a pair;
  * apply_joint_correction: ghost level untouched, the same increment on both
    time levels, zero dz-weighted column mean and masking of the velocity
    increments, tracer inventories unchanged per layer, the cap, and rho the
    oracle's EOS of the new (t, s) with the old ghost level
    (src/neuralpom/hybrid/sparse_coupling.py:42-67);
  * pressure_velocity_error: zero for the core's own (t, s), and the oracle's
    baropg of the oracle EOS difference with zero mean density (:70-82).
The EOS and baropg themselves are pinned to the reference in test_core.py.
"""
from types import SimpleNamespace

import pytest
import torch

from neuralpom_amd import coupling as CP
from oracle.pomref import explicit as X
from oracle.pomref.cases import double_gyre as oracle_dg
from tests._helpers import rel_err

IM, JM = 12, 10


def _weights(gen, wet_frac=0.8):
    w = torch.rand(IM, JM, CP.LEVELS, generator=gen, dtype=torch.float64) + 0.1
    dry = torch.rand(IM, JM, 1, generator=gen, dtype=torch.float64) > wet_frac
    return w.masked_fill(dry, 0.0)


def test_budget_project_invariants():
    gen = torch.Generator().manual_seed(1)
    w = _weights(gen)
    x = torch.randn(IM, JM, CP.LEVELS, generator=gen, dtype=torch.float64)
    y = CP.budget_project(x, w)
    layer = (y * w).sum((0, 1)) / w.sum((0, 1))
    assert float(layer.abs().max()) < 1e-14
    assert float(y[w == 0].abs().max()) == 0.0
    # idempotent: a wet-only increment with zero layer means passes unchanged
    assert rel_err(CP.budget_project(y, w), y) < 1e-14


def test_paired_limit_invariants():
    gen = torch.Generator().manual_seed(2)
    a = torch.randn(5, 6, generator=gen, dtype=torch.float64)
    b = torch.randn(5, 6, generator=gen, dtype=torch.float64)
    sa, sb = 0.1, 0.02
    peak = max(float((a / sa).abs().max()), float((b / sb).abs().max()))
    (la, lb), f = CP.paired_limit([a, b], [sa, sb], peak * 2.0)
    assert float(f) == 1.0 and torch.equal(la, a) and torch.equal(lb, b)
    cap = peak / 3.0
    (la, lb), f = CP.paired_limit([a, b], [sa, sb], cap)
    new_peak = max(float((la / sa).abs().max()), float((lb / sb).abs().max()))
    assert abs(new_peak - cap) < 1e-12 * cap
    assert rel_err(la / a, torch.full_like(a, float(f))) < 1e-14
    assert rel_err(lb / b, torch.full_like(b, float(f))) < 1e-14


def test_volume_weights_and_weighted_mse():
    gen = torch.Generator().manual_seed(3)
    core = SimpleNamespace(art=torch.rand(IM, JM, generator=gen, dtype=torch.float64) + 1.0,
                           h=torch.rand(IM, JM, generator=gen, dtype=torch.float64) * 100.0,
                           et=torch.rand(IM, JM, generator=gen, dtype=torch.float64),
                           dz=torch.rand(16, generator=gen, dtype=torch.float64),
                           fsm=(torch.rand(IM, JM, generator=gen) > 0.2).double(),
                           dum=(torch.rand(IM, JM, generator=gen) > 0.3).double(),
                           dvm=(torch.rand(IM, JM, generator=gen) > 0.4).double())
    for field, mask in (("t", core.fsm), ("u", core.dum), ("vb", core.dvm)):
        w = CP.volume_weights(core, field)
        i, j, k = 3, 4, 7
        assert float(w[i, j, k]) == float(core.art[i, j] * mask[i, j] * (core.h[i, j] + core.et[i, j]) * core.dz[k])
        assert w.shape == (IM, JM, CP.LEVELS)
    x = torch.randn(IM, JM, CP.LEVELS, generator=gen, dtype=torch.float64)
    y = torch.randn(IM, JM, CP.LEVELS, generator=gen, dtype=torch.float64)
    w = CP.volume_weights(core)
    brute = sum(float((x[i, j, k] - y[i, j, k]) ** 2 * w[i, j, k]) for i in range(IM) for j in range(JM)
                for k in range(CP.LEVELS)) / float(w.sum())
    assert abs(float(CP.weighted_mse(x, y, w)) - brute) < 1e-12 * brute


def test_output_contract_is_checked():
    core = SimpleNamespace(kbm1=15, im=IM, jm=JM)
    with pytest.raises(ValueError):
        CP.apply_joint_correction(core, torch.zeros(59, IM, JM), {})


# ---------------------------------------------------------------- GPU
SCALES = {"u": 0.05, "v": 0.05, "t": 0.2, "s": 0.05}


def _core():
    from neuralpom_amd.core import POM2K_Core
    core = POM2K_Core(IM, JM, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    gen = torch.Generator().manual_seed(4)
    core.t = core.t + 0.5 * torch.randn(core.t.shape, generator=gen, dtype=torch.float64).cuda()
    core.et = core.et + 0.01 * torch.randn(core.et.shape, generator=gen, dtype=torch.float64).cuda()
    core.rho = core.eos(core.t, core.s)
    return core


def _snapshot(core):
    return {n: getattr(core, n).clone() for n in ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "rho")}


def _oracle_rho(t, s, rho_old):
    S = oracle_dg(IM, JM)
    return X.dens(S, t.cpu(), s.cpu(), rho_old.cpu())


@pytest.mark.gpu
def test_zero_network_changes_nothing_but_refreshes_rho():
    print("[gpu] corrector hook, zero increment", flush=True)
    core = _core()
    before = _snapshot(core)
    out = CP.apply_joint_correction(core, torch.zeros(CP.OUTPUT_CHANNELS, IM, JM, dtype=torch.float64,
                                                      device="cuda"), SCALES)
    for n in ("u", "ub", "v", "vb", "t", "tb", "s", "sb"):
        assert torch.equal(getattr(core, n), before[n]), n
    assert float(out["ts_limit_factor"]) == 1.0
    assert rel_err(core.rho.cpu(), _oracle_rho(before["t"], before["s"], before["rho"])) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("amp", [0.5, 50.0])
def test_correction_invariants(amp):
    print(f"[gpu] corrector hook, amplitude {amp}", flush=True)
    core = _core()
    before = _snapshot(core)
    gen = torch.Generator().manual_seed(6)
    raw = amp * torch.randn(CP.OUTPUT_CHANNELS, IM, JM, generator=gen, dtype=torch.float64).cuda()
    cap = 5.0
    out = CP.apply_joint_correction(core, raw, SCALES, cap)
    L = CP.LEVELS
    inc = {}
    for f, older in (("u", "ub"), ("v", "vb"), ("t", "tb"), ("s", "sb")):
        for n in (f, older):
            assert torch.equal(getattr(core, n)[:, :, L:], before[n][:, :, L:]), f"ghost level of {n}"
        inc[f] = getattr(core, f)[:, :, :L] - before[f][:, :, :L]
        d_old = getattr(core, older)[:, :, :L] - before[older][:, :, :L]
        assert rel_err(d_old, inc[f]) < 1e-12, f"{older} increment"
    dz = core.dz[:L].view(1, 1, L)
    for f, mask in (("u", core.dum), ("v", core.dvm)):
        col = (inc[f] * dz).sum(2) / dz.sum()
        assert float(col.abs().max()) < 1e-12 * max(float(inc[f].abs().max()), 1e-30)
        assert float(inc[f][mask == 0].abs().max() if (mask == 0).any() else 0.0) == 0.0
    w = CP.volume_weights(core)
    for f in ("t", "s"):
        layer = (inc[f] * w).sum((0, 1))
        assert float(layer.abs().max()) < 1e-10 * float((inc[f].abs() * w).sum((0, 1)).max())
    peak_ts = max(float((inc["t"] / SCALES["t"]).abs().max()), float((inc["s"] / SCALES["s"]).abs().max()))
    peak_uv = max(float((inc["u"] / SCALES["u"]).abs().max()), float((inc["v"] / SCALES["v"]).abs().max()))
    assert peak_ts <= cap * (1 + 1e-12) and peak_uv <= cap * (1 + 1e-12)
    if amp > 10.0:
        assert abs(peak_ts - cap) < 1e-9 * cap and float(out["ts_limit_factor"]) < 1.0
    else:
        assert float(out["ts_limit_factor"]) == 1.0
    assert abs(float(out["t_budget"])) < 1e-12 and abs(float(out["s_budget"])) < 1e-12
    ref = _oracle_rho(core.t, core.s, before["rho"])
    assert rel_err(core.rho.cpu(), ref) < 1e-13


@pytest.mark.gpu
def test_correction_is_differentiable():
    print("[gpu] corrector hook, directional derivative", flush=True)
    gen = torch.Generator().manual_seed(8)
    raw = 0.5 * torch.randn(CP.OUTPUT_CHANNELS, IM, JM, generator=gen, dtype=torch.float64).cuda()
    d = torch.randn(raw.shape, generator=gen, dtype=torch.float64).cuda()

    def loss(r):
        # increments only, so the loss (and the rounding of the difference quotient) stays O(1)
        core = _core()
        b = _snapshot(core)
        CP.apply_joint_correction(core, r, SCALES)
        return (((core.rho - b["rho"]) ** 2).sum() + ((core.u - b["u"]) ** 2).sum() +
                ((core.t - b["t"]) * (core.s - b["s"])).sum() * 10.0)

    r = raw.clone().requires_grad_(True)
    (g,) = torch.autograd.grad(loss(r), r)
    eps = 1e-5
    fd = (loss(raw + eps * d) - loss(raw - eps * d)) / (2 * eps)
    assert abs(float((g * d).sum()) - float(fd)) < 1e-6 * max(abs(float(fd)), 1.0)


@pytest.mark.gpu
def test_pressure_velocity_error():
    print("[gpu] pressure_velocity_error", flush=True)
    core = _core()
    ku, kv = CP.pressure_velocity_error(core, core.t, core.s)
    assert float(ku.abs().max()) == 0.0 and float(kv.abs().max()) == 0.0
    gen = torch.Generator().manual_seed(9)
    ref_t = core.t + 0.3 * torch.randn(core.t.shape, generator=gen, dtype=torch.float64).cuda()
    core.ramp = 1.0
    ku, kv = CP.pressure_velocity_error(core, ref_t, core.s)
    S = oracle_dg(IM, JM)
    S.rmean = torch.zeros_like(S.rmean)
    S.et, S.dt = core.et.cpu(), core.dt.cpu()
    drho = _oracle_rho(core.t, core.s, core.rho) - _oracle_rho(ref_t, core.s, core.rho)
    _, px, _ = X.baropg(S, drho, S.dt, 1.0)
    i, j, k = 5, 4, 3
    depth = core.h + core.etf
    du = float((depth[i, j] + depth[i - 1, j]) * core.aru[i, j])
    want = -2.0 * core.dti2 * float(px[i, j, k]) / du * float(core.dum[i, j])
    assert abs(float(ku[i, j, k]) - want) < 1e-12 * max(abs(want), 1e-30)
