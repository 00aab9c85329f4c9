"""The oracle's corrector hook, density-75 projection (oracle/pomref/coupling.py)
and the product's CorrectorNet module against the reference's own outputs
(tests/golden/make_golden_corrector.py; src/neuralpom/hybrid/sparse_coupling.py:18-67,
src/neuralpom/evaluation/density.py:14-64, src/neuralpom/models/corrector_net.py:7-100)."""
import math

import numpy as np
import pytest
import torch

from oracle.pomref import coupling as OC
from oracle.pomref.state import load_state
from tests._helpers import golden, rel_err

STATE = ("u", "ub", "v", "vb", "t", "tb", "s", "sb")


@pytest.fixture(scope="module")
def g():
    return golden("corrector_12x10.npz")


def _scales(g):
    return dict(zip(("u", "v", "t", "s"), (float(x) for x in g["scales"])))


def _setup(g, leaves=()):
    S = load_state(g, "in")
    st = {n: getattr(S, n).clone().requires_grad_(n in leaves) for n in STATE + ("et", "rho")}
    return S, st


def T(g, k):
    return torch.from_numpy(np.array(g[k]))


def test_joint_correction_matches_reference(g):
    S, st = _setup(g, leaves=STATE + ("et",))
    raw = T(g, "jc/raw").requires_grad_(True)
    out, diag = OC.joint_correction(S, st, raw, _scales(g), 5.0)
    for n in STATE + ("rho",):
        assert rel_err(out[n].detach(), T(g, f"jc/{n}")) < 1e-13, n
    ref = g["jc/diag"]
    dv = [float(x.detach()) for x in diag]
    assert dv[0] < 1.0, "the tracer cap binds in this fixture"
    assert abs(dv[0] - ref[0]) < 1e-14
    for q in (1, 2):     # budgets are zero up to rounding: compared on the increments' scale
        assert abs(dv[q] - ref[q]) < 1e-15, (q, dv[q], ref[q])
    loss = sum((out[n] * T(g, f"jccot/{n}")).sum() for n in STATE + ("rho",))
    cd = T(g, "jccot/diag")
    loss = loss + sum(cd[q] * diag[q] for q in range(3))
    names = ["raw", "et"] + list(STATE)
    grads = torch.autograd.grad(loss, [raw] + [st[n] for n in names[1:]])
    for n, gr in zip(names, grads):
        assert rel_err(gr, T(g, f"jcgrad/{n}")) < 1e-11, n


@pytest.mark.parametrize("case", ["a", "b", "z"])
def test_density75_matches_reference(g, case):
    S, st = _setup(g)
    out, diag = OC.density75(S, st, T(g, f"d75{case}/raw"), _scales(g))
    for n in ("t", "tb", "s", "sb", "rho", "u", "ub", "v", "vb"):
        assert rel_err(out[n], T(g, f"d75{case}/{n}")) < 1e-13, (case, n)
    ref = g[f"d75{case}/diag"]
    assert abs(float(diag[0]) - ref[0]) < 1e-14
    if case == "z":
        assert math.isnan(float(diag[1])) and math.isnan(ref[1])
    else:
        assert abs(float(diag[1]) - ref[1]) < 1e-14
        # the constraint residual is zero up to rounding
        assert float(diag[2]) < 1e-13 and ref[2] < 1e-13
    if case == "b":
        assert ref[0] < 1.0   # the joint-correction cap binds


def test_density75_removes_a_quarter_of_the_density_component(g):
    """What the projection fixes, checked without the reference: with no cap
    binding, the T/S increment's density component (along the scaled EOS
    gradient) is 75 % of the corrected one's after the per-level constraint
    part is taken out, the orthogonal component is unchanged, and the
    weighted moments of the removed part against the direction are zero."""
    S, st = _setup(g)
    raw = T(g, "d75a/raw")
    sc = _scales(g)
    jc, _ = OC.joint_correction(S, st, raw, sc, 5.0)
    out, diag = OC.density75(S, st, raw, sc)
    assert float(diag[1]) == 1.0
    s2 = torch.tensor([sc["t"], sc["s"]], dtype=torch.float64)
    x = torch.stack([(jc[f] - st[f])[..., :15] for f in ("t", "s")], -1) / s2
    y = torch.stack([(out[f] - st[f])[..., :15] for f in ("t", "s")], -1) / s2
    dt, ds = OC.eos_gradient(S, st["t"], st["s"], st["rho"])
    d = torch.stack((dt, ds), -1) * s2
    d = d / torch.linalg.vector_norm(d, dim=-1, keepdim=True).clamp_min(1e-30)
    perp = torch.stack((-d[..., 1], d[..., 0]), -1)
    assert rel_err((y * perp).sum(-1), (x * perp).sum(-1)) < 1e-12
    removed = ((x - y) * d).sum(-1)                   # 0.25 * adjusted amplitude
    w = OC.volume_weights(S, st["et"])
    mom = torch.einsum("ijk,ijka,ijk->ka", w, d, removed)
    assert float(mom.abs().max()) < 1e-12 * float((w[..., None] * d.abs() * removed.abs()[..., None]).sum())


def test_corrector_net_matches_reference(g):
    from neuralpom_amd.hybrid import CorrectorNet
    net = CorrectorNet(hidden_dim=4).double()
    sd = {k[6:]: torch.from_numpy(np.array(g[k])) for k in g.files if k.startswith("net/p/")}
    net.load_state_dict(sd, strict=True)
    net.eval()
    with torch.no_grad():
        y = net(T(g, "net/x"))
    assert rel_err(y, T(g, "net/y")) < 1e-12
