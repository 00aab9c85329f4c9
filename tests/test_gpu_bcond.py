"""The drop-in core's bcond(idx, elf, uaf, vaf, uf, vf, w) (HIP boundary
operators) against the reference's own outputs for each idx, recorded inside
a perturbed float64 step (tests/golden/explicit_ops_12x10.npz), and its
gradients against torch autograd through the oracle's bcond1-6 (pinned to the
same records by tests/test_oracle_golden.py)."""
import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.core import POM2K_Core
from oracle.pomref import explicit as X
from oracle.pomref.state import SCALAR_NAMES, load_state
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

ARGS = ("elf", "uaf", "vaf", "uf", "vf", "w")
READS = {2: ("el",), 3: ("u", "v"), 4: ("t", "s", "u", "v", "w", "dt"), 5: (), 6: ("q2", "q2l", "u", "v"), 1: ()}
CHANGED = {1: (0,), 2: (1, 2), 3: (3, 4), 4: (3, 4), 5: (5,), 6: (3, 4)}


def _setup():
    ops = golden("explicit_ops_12x10.npz")
    S0 = load_state(ops, "t")
    S0.time = S0.dti / 86400.0 + 0.1
    S0.ramp = min(S0.time / S0.period, 1.0)
    core = POM2K_Core(S0.im, S0.jm, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=S0.dte, isplit=S0.isplit)
    core.run_initialization()
    for n in L.FIELDS + L.GEOM + ("z", "zz", "dz", "dzz"):
        setattr(core, n, getattr(S0, n).to("cuda").clone())
    for n in SCALAR_NAMES:
        if hasattr(S0, n):
            setattr(core, n, getattr(S0, n))
    core.ramp = S0.ramp
    return ops, S0, core


def _rec(ops, idx, key):
    return torch.from_numpy(np.array(ops[f"bcond{idx}/{key}"]))


def _oracle(S, idx, args, rd):
    elf, uaf, vaf, uf, vf, w = args
    if idx == 1:
        return (X.bcond1(S, elf),)
    if idx == 2:
        return X.bcond2(S, uaf, vaf, rd["el"], S.ramp)
    if idx == 3:
        return X.bcond3(S, uf, vf, rd["u"], rd["v"])
    if idx == 4:
        return X.bcond4(S, uf, vf, rd["t"], rd["s"], rd["u"], rd["v"], rd["w"], rd["dt"])
    if idx == 5:
        return (X.bcond5(S, w),)
    return X.bcond6(S, uf, vf, rd["q2"], rd["q2l"], rd["u"], rd["v"])


@pytest.mark.parametrize("idx", [1, 2, 3, 4, 5, 6])
def test_bcond_matches_reference_and_its_gradient(idx):
    print(f"[gpu] core.bcond({idx}) 12x10 fp64", flush=True)
    ops, S0, core = _setup()
    args = [_rec(ops, idx, f"arg{q}") for q in range(1, 7)]
    rd = {n: _rec(ops, idx, f"self/{n}") for n in READS[idx]}
    for n, v in rd.items():
        setattr(core, n, v.to("cuda").clone().requires_grad_(True))
    gargs = [a.to("cuda").clone().requires_grad_(True) for a in args]
    out = core.bcond(idx, *gargs)
    for q in range(6):
        want = _rec(ops, idx, f"out{q}")
        e = rel_err(out[q].detach().cpu(), want)
        assert e <= 1e-13, f"bcond{idx} output {ARGS[q]}: {e:.2e}"
    # gradients through the changed outputs against the oracle's autograd
    gen = torch.Generator().manual_seed(idx)
    cot = {q: torch.randn(args[q].shape, generator=gen, dtype=torch.float64) for q in CHANGED[idx]}
    leaves_hip = gargs + [getattr(core, n) for n in READS[idx]]
    loss = sum((out[q] * cot[q].cuda()).sum() for q in CHANGED[idx])
    g_hip = torch.autograd.grad(loss, leaves_hip, allow_unused=True)
    oargs = [a.clone().requires_grad_(True) for a in args]
    ord_ = {n: v.clone().requires_grad_(True) for n, v in rd.items()}
    ores = _oracle(S0, idx, oargs, ord_)
    oout = {CHANGED[idx][q]: r for q, r in enumerate(ores)}
    oloss = sum((oout[q] * cot[q]).sum() for q in CHANGED[idx])
    g_ref = torch.autograd.grad(oloss, oargs + list(ord_.values()), allow_unused=True)
    names = list(ARGS) + list(READS[idx])
    bad = []
    for n, a, b in zip(names, g_hip, g_ref):
        a = torch.zeros(1) if a is None else a.detach().cpu()
        b = torch.zeros(1) if b is None else b
        if float(b.abs().max()) == 0.0 and float(a.abs().max()) == 0.0:
            continue
        e = rel_err(a.expand_as(b) if a.numel() == 1 else a, b)
        if not e <= 1e-12:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, f"bcond{idx} gradients: {bad}"
