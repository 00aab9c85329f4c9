"""The differentiable coupled rollout (physics intervals on the drop-in core,
the residual U-Net corrector and the corrector hook after each interval)
against the same rollout on the oracle under torch autograd, and against
central finite differences.

Float64 on a perturbed 26 x 22 double gyre: 3 intervals of 2 internal steps,
a width-8 U-Net with a seeded head (the reference's zero head would make
every correction zero), loss = volume-weighted MSE of u, v, t, s against a
seeded perturbation of the start state.  The oracle runs the reference
algorithm on the CPU; the network (a deep copy) and the hook are the same
torch code on both sides, the hook's equation of state being the oracle's
dens there and the HIP EOS here.
"""
import copy

import pytest
import torch

from neuralpom_amd import coupling as CP
from neuralpom_amd.core import POM2K_Core
from neuralpom_amd.hybrid import CorrectorNet, coupled_rollout
from oracle.pomref import explicit as X
from oracle.pomref.cases import double_gyre as oracle_dg
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu

IM, JM = 26, 22
INTERVALS, STEPS = 3, 2
SCALES = {"u": 0.02, "v": 0.02, "t": 0.1, "s": 0.02}
LEAVES = ("t", "s", "u", "v", "el")
FIELDS = ("u", "v", "t", "s")


def _start():
    S = oracle_dg(IM, JM)
    gen = torch.Generator().manual_seed(11)
    for n, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02)):
        x = getattr(S, n)
        setattr(S, n, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
    S.d = S.h + S.el
    S.time0 = S.time = 0.25
    return S


def _net():
    gen = torch.Generator().manual_seed(21)
    torch.manual_seed(21)
    return CorrectorNet(hidden_dim=8, head_std=0.05, generator=gen).double()


def _targets(S):
    gen = torch.Generator().manual_seed(31)
    tg = {f: getattr(S, f)[:, :, :CP.LEVELS] + 0.01 * torch.randn((IM, JM, CP.LEVELS), generator=gen,
                                                                   dtype=torch.float64) for f in FIELDS}
    w = {f: CP.volume_weights(S, f) for f in FIELDS}
    return tg, w


def _loss(st, tg, w):
    return sum(((getattr(st, f)[:, :, :CP.LEVELS] - tg[f]) ** 2 * w[f]).sum() / w[f].sum() for f in FIELDS)


def _core_from(S):
    core = POM2K_Core(IM, JM, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    for n in ("t", "s", "el", "u", "v", "d"):
        setattr(core, n, getattr(S, n).to("cuda").clone())
    core.time0 = core.time = S.time
    return core


def _oracle_rollout(S, net):
    S.eos = lambda t, s: X.dens(S, t, s, S.rho)
    for _ in range(INTERVALS):
        S.time0 = S.time
        S.iend = STEPS
        assert X.run_time_loop(S)
        planes = [getattr(S, n).permute(2, 0, 1) for n in ("u", "ub", "v", "vb", "t", "s")]
        planes += [getattr(S, n).unsqueeze(0) for n in ("el", "ua", "uab", "va", "vab")]
        raw = net(torch.cat(planes).unsqueeze(0))[0]
        CP.apply_joint_correction(S, raw, SCALES)


def _hip_loss(S0, net, tg, w, over=None):
    core = _core_from(S0)
    for n, v in (over or {}).items():
        setattr(core, n, v)
    coupled_rollout(core, net, SCALES, INTERVALS, STEPS)
    return _loss(core, tg, w)


def test_rollout_loss_and_gradients_match_oracle():
    print(f"[gpu] coupled rollout {IM}x{JM} fp64, {INTERVALS}x{STEPS} steps, U-Net width 8", flush=True)
    S0 = _start()
    tg, w = _targets(S0)
    net = _net().cuda()
    core = _core_from(S0)
    leaves = {n: getattr(core, n).clone().requires_grad_(True) for n in LEAVES}
    for n, v in leaves.items():
        setattr(core, n, v)
    coupled_rollout(core, net, SCALES, INTERVALS, STEPS)
    loss = _loss(core, {f: x.cuda() for f, x in tg.items()}, {f: x.cuda() for f, x in w.items()})
    params = list(net.parameters())
    grads = torch.autograd.grad(loss, list(leaves.values()) + params)

    S = S0.copy()
    onet = copy.deepcopy(net).cpu()
    oleaves = {n: getattr(S, n).clone().requires_grad_(True) for n in LEAVES}
    for n, v in oleaves.items():
        setattr(S, n, v)
    _oracle_rollout(S, onet)
    oloss = _loss(S, tg, w)
    ograds = torch.autograd.grad(oloss, list(oleaves.values()) + list(onet.parameters()))

    assert rel_err(loss.detach().cpu(), oloss.detach()) < 1e-10, (float(loss), float(oloss))
    names = list(LEAVES) + [n for n, _ in net.named_parameters()]
    bad = [f"{n}: {rel_err(g.detach().cpu(), r):.2e}" for n, g, r in zip(names, grads, ograds)
           if not rel_err(g.detach().cpu(), r) <= 1e-7]
    assert not bad, bad


def test_rollout_gradient_matches_central_differences():
    print(f"[gpu] coupled rollout central FD, {IM}x{JM} fp64", flush=True)
    S0 = _start()
    tg, w = _targets(S0)
    w = {f: x.cuda() for f, x in w.items()}
    tg = {f: x.cuda() for f, x in tg.items()}
    net = _net().cuda()
    base = {n: getattr(S0, n).to("cuda") for n in LEAVES}
    leaves = {n: v.clone().requires_grad_(True) for n, v in base.items()}
    loss = _hip_loss(S0, net, tg, w, leaves)
    params = list(net.parameters())
    g = torch.autograd.grad(loss, list(leaves.values()) + params)
    gs, gp = g[:len(LEAVES)], g[len(LEAVES):]
    rel = []

    def quotient(ds, dp, eps):
        vals = []
        for sgn in (1.0, -1.0):
            netd = copy.deepcopy(net)
            for p, q in zip(netd.parameters(), dp):
                p.add_(sgn * eps * q)
            vals.append(float(_hip_loss(S0, netd, tg, w, {n: base[n] + sgn * eps * ds[n] for n in LEAVES})))
        return (vals[0] - vals[1]) / (2.0 * eps)

    with torch.no_grad():
        for d in range(8):
            gen = torch.Generator(device="cuda").manual_seed(500 + d)
            ds = {n: (v.abs() + 1e-3 * v.abs().max()) * torch.randn(v.shape, generator=gen, device="cuda",
                                                                    dtype=v.dtype) for n, v in base.items()}
            dp = [p.abs().mean() * torch.randn(p.shape, generator=gen, device="cuda", dtype=p.dtype) for p in params]
            ad = sum(float((a * b).sum()) for a, b in zip(gs, ds.values())) + \
                sum(float((a * b).sum()) for a, b in zip(gp, dp))
            # The network's ReLUs and max pools (and the hook's cap) are kinks:
            # a direction whose +-eps points straddle one gives a quotient off
            # by O(eps / distance to the kink).  Such a direction is taken again
            # at a 10x smaller step; an adjoint error would not shrink with eps.
            errs = []
            for eps in (1e-8, 1e-9):
                fd = quotient(ds, dp, eps)
                errs.append(abs(fd - ad) / max(abs(ad), abs(fd), 1e-300))
                if errs[-1] <= 1e-5:
                    break
            rel.append(min(errs))
    print(f"[gpu] coupled rollout FD rel err per direction: {['%.1e' % r for r in rel]}", flush=True)
    assert max(rel) <= 1e-4, rel
