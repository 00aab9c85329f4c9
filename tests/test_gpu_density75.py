"""The fused corrector hook and the DND density-75 projection on the MI355X
against the reference's own outputs (tests/golden/corrector_12x10.npz from
tests/golden/make_golden_corrector.py: src/neuralpom/hybrid/sparse_coupling.py:42-67,
src/neuralpom/evaluation/density.py:38-64, src/neuralpom/models/corrector_net.py:72-100)
and, on a ragged batch, against the oracle (oracle/pomref/coupling.py, pinned
to the same fixture by test_oracle_coupling.py).

The density-component cap never binds in the reference's draws (the hook's
cap leaves |x| <= 5 and the 25 % shrink rarely moves a component past it), so
its binding branch is checked against the oracle with a smaller cap passed to
pom_density75 directly.
"""
import math

import numpy as np
import pytest
import torch

from neuralpom_amd import coupling as CP
from neuralpom_amd import density as DN
from neuralpom_amd.core import POM2K_Core
from oracle.pomref import coupling as OC
from oracle.pomref.state import load_state
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

STATE = ("u", "ub", "v", "vb", "t", "tb", "s", "sb")
LOADED = STATE + ("et", "rho", "h", "art", "fsm", "dum", "dvm", "z", "zz", "dz", "dzz")


@pytest.fixture(scope="module")
def g():
    return golden("corrector_12x10.npz")


def _scales(g):
    return dict(zip(("u", "v", "t", "s"), (float(x) for x in g["scales"])))


def _core(g):
    core = POM2K_Core(12, 10, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=30.0, isplit=30)
    core.run_initialization()
    for n in LOADED:
        setattr(core, n, torch.from_numpy(np.array(g[f"in/{n}"])).to("cuda", torch.float64))
    return core


def T(g, k):
    return torch.from_numpy(np.array(g[k]))


def test_fused_hook_matches_reference_forward_and_vjp(g):
    print("[gpu] fused corrector hook vs the reference (12x10, fwd + VJP)", flush=True)
    core = _core(g)
    leaves = {n: getattr(core, n).clone().requires_grad_(True) for n in STATE + ("et",)}
    for n, v in leaves.items():
        setattr(core, n, v)
    raw = T(g, "jc/raw").to("cuda").requires_grad_(True)
    diag = CP.apply_joint_correction(core, raw, _scales(g), cap=5.0)
    bad = []
    for n in STATE + ("rho",):
        e = rel_err(getattr(core, n).detach().cpu(), T(g, f"jc/{n}"))
        if not e <= 1e-12:
            bad.append(f"{n}: {e:.1e}")
    dv = [float(diag[k].detach()) for k in ("ts_limit_factor", "t_budget", "s_budget")]
    ref = g["jc/diag"]
    assert abs(dv[0] - ref[0]) < 1e-14 and abs(dv[1] - ref[1]) < 1e-15 and abs(dv[2] - ref[2]) < 1e-15, (dv, ref)
    loss = sum((getattr(core, n) * T(g, f"jccot/{n}").cuda()).sum() for n in STATE + ("rho",))
    cd = T(g, "jccot/diag")
    loss = loss + cd[0] * diag["ts_limit_factor"] + cd[1] * diag["t_budget"] + cd[2] * diag["s_budget"]
    names = ["raw", "et"] + list(STATE)
    grads = torch.autograd.grad(loss, [raw] + [leaves[n] for n in names[1:]])
    for n, gr in zip(names, grads):
        e = rel_err(gr.cpu(), T(g, f"jcgrad/{n}"))
        if not e <= 1e-10:
            bad.append(f"grad {n}: {e:.1e}")
    assert not bad, bad


@pytest.mark.parametrize("case", ["a", "b", "z"])
def test_density75_matches_reference(g, case):
    print(f"[gpu] apply_density75 vs the reference (12x10), case {case}", flush=True)
    core = _core(g)
    diag = DN.apply_density75(core, T(g, f"d75{case}/raw").cuda(), _scales(g))
    bad = []
    for n in ("t", "tb", "s", "sb", "rho", "u", "ub", "v", "vb"):
        e = rel_err(getattr(core, n).cpu(), T(g, f"d75{case}/{n}"))
        if not e <= 1e-12:
            bad.append(f"{n}: {e:.1e}")
    assert not bad, bad
    ref = g[f"d75{case}/diag"]
    assert abs(float(diag["ts_limit_factor"]) - ref[0]) < 1e-13
    if case == "z":
        assert math.isnan(ref[1]) and "density_component_limit_factor" not in diag
    else:
        assert abs(float(diag["density_component_limit_factor"]) - ref[1]) < 1e-13
        assert float(diag["density_constraint_residual_max_abs"]) < 1e-13


def _oracle_member(im, jm, seed):
    from oracle.pomref.cases import double_gyre
    S = double_gyre(im, jm)
    gen = torch.Generator().manual_seed(seed)
    for n, a in (("t", 0.5), ("s", 0.05), ("et", 0.01)):
        x = getattr(S, n)
        setattr(S, n, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
    S.tb = S.t + 0.01 * torch.randn(S.t.shape, generator=gen, dtype=S.t.dtype)
    S.sb = S.s + 0.001 * torch.randn(S.s.shape, generator=gen, dtype=S.s.dtype)
    S.rho = __import__("oracle.pomref.explicit", fromlist=["dens"]).dens(S, S.t, S.s, S.rho)
    raw = 3.0 * torch.randn((60, im, jm), generator=gen, dtype=torch.float64)
    return S, raw


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("cap", [5.0, 1.0])
def test_density75_ragged_batch_matches_oracle(dtype, cap):
    """Three members of different shapes in one launch; each member's (before,
    after) from the oracle's hook, the projection against the oracle's.  cap 1
    makes the density-component cap bind."""
    from neuralpom_amd.explicit import Ensemble
    print(f"[gpu] pom_density75 ragged batch vs oracle, {dtype}, cap {cap}", flush=True)
    shapes = ((26, 22), (19, 31), (12, 10))
    sc = {"u": 0.02, "v": 0.02, "t": 0.05, "s": 0.01}
    members = [_oracle_member(im, jm, 70 + m) for m, (im, jm) in enumerate(shapes)]
    refs, befores, afters = [], [], []
    for S, raw in members:
        st = {n: getattr(S, n) for n in STATE + ("et", "rho")}
        after, _ = OC.joint_correction(S, st, raw, sc, 5.0)
        # the oracle's projection of the same (before, after) with the cap under test
        x = torch.stack([(after[f] - st[f])[..., :15] for f in ("t", "s")], -1) / torch.tensor([sc["t"], sc["s"]],
                                                                                              dtype=torch.float64)
        dt, ds = OC.eos_gradient(S, st["t"], st["s"], st["rho"])
        gg = torch.stack((dt, ds), -1) * torch.tensor([sc["t"], sc["s"]], dtype=torch.float64)
        d = gg / torch.linalg.vector_norm(gg, dim=-1, keepdim=True).clamp_min(1e-30)
        amp = (x * d).sum(-1)
        adj, res = OC.constrained_amplitude(amp, d, OC.volume_weights(S, st["et"]))
        prop = (x - 0.25 * adj[..., None] * d) * torch.tensor([sc["t"], sc["s"]], dtype=torch.float64)
        (it, is_), f = OC.paired_limit([prop[..., 0], prop[..., 1]], [sc["t"], sc["s"]], cap)
        ref = {}
        for n, inc in (("t", it), ("tb", it), ("s", is_), ("sb", is_)):
            ref[n] = torch.cat((st[n][..., :15] + inc, st[n][..., 15:]), 2)
        refs.append((ref, float(f), float(res.abs().max())))
        befores.append({n: st[n] for n in DN.D75_STATE})
        afters.append({n: after[n] for n in DN.D75_STATE})
    if cap == 1.0:
        assert all(r[1] < 1.0 for r in refs), "the density-component cap must bind"
    ens = Ensemble(list(shapes), dtype=dtype, device="cuda")
    pk = lambda xs, n: ens.pack([x.to(dtype) for x in xs], n)  # noqa: E731
    geom = [pk([getattr(S, n) if n != "et" else S.et for S, _ in members], n) for n in DN.D75_GEOM]
    before = [pk([b[n] for b in befores], n) for n in DN.D75_STATE]
    after = [pk([a[n] for a in afters], n) for n in DN.D75_STATE]
    vg = torch.stack([torch.stack([S.z, S.zz, S.dz, S.dzz]) for S, _ in members]).to("cuda", dtype).contiguous()
    core = POM2K_Core(12, 10, device="cuda", dtype=dtype)
    core.configure_runtime(dte=30.0, isplit=30)
    out, diag = DN.density75(ens, core._params(), vg, (sc["t"], sc["s"]), cap, geom, before, after)
    torch.cuda.synchronize()
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    for m, (ref, f, res) in enumerate(refs):
        # the factor follows the peak element, with the increments' fp32 error (below)
        assert abs(float(diag[m, 0]) - f) <= (1e-13 if dtype == torch.float64 else 1e-3 * f), (m, float(diag[m, 0]), f)
        assert float(diag[m, 2]) == 1.0
        assert float(diag[m, 1]) < (1e-13 if dtype == torch.float64 else 1e-6)
        for q, n in enumerate(DN.D75_STATE):
            got = ens.view(out[q], m, n).double().cpu()
            inc_ref = (ref[n] - befores[m][n])[..., :15]
            inc_got = got[..., :15] - befores[m][n][..., :15].to(dtype).double()
            e = rel_err(inc_got, inc_ref)
            # fp32: x = after - before is formed in the fields' precision, as
            # the reference forms it (density.py:43); ulp(35) = 3.8e-6 against
            # salinity increments of 1e-2 .. 5e-2 gives about 1e-4
            assert e <= (1e-11 if dtype == torch.float64 else 1e-3), (m, n, e)
            assert rel_err(got, ref[n]) <= tol, (m, n)
            assert torch.equal(got[..., 15], befores[m][n][..., 15].to(dtype).double())


def test_density75_no_change_member_keeps_the_corrected_state():
    """A member whose T/S increments are zero gets the corrected state back
    (the reference's early return) while its batch neighbour is projected."""
    from neuralpom_amd.explicit import Ensemble
    print("[gpu] pom_density75: a member without T/S change next to one with", flush=True)
    shapes = ((19, 31), (12, 10))
    members = [_oracle_member(im, jm, 90 + m) for m, (im, jm) in enumerate(shapes)]
    ens = Ensemble(list(shapes), dtype=torch.float64, device="cuda")
    before, after = [], []
    for n in DN.D75_STATE:
        b = [getattr(S, n) for S, _ in members]
        a = [b[0] + 0.01 * torch.randn(b[0].shape, dtype=torch.float64, generator=torch.Generator().manual_seed(3)),
             b[1].clone()]
        a[1][..., 15] += 1.0     # a ghost-level difference only: x is zero
        before.append(ens.pack(b, n))
        after.append(ens.pack(a, n))
    geom = [ens.pack([getattr(S, n) for S, _ in members], n) for n in DN.D75_GEOM]
    vg = torch.stack([torch.stack([S.z, S.zz, S.dz, S.dzz]) for S, _ in members]).to("cuda").contiguous()
    core = POM2K_Core(12, 10, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=30.0, isplit=30)
    out, diag = DN.density75(ens, core._params(), vg, (0.05, 0.01), 5.0, geom, before, after)
    torch.cuda.synchronize()
    assert float(diag[0, 2]) == 1.0 and float(diag[1, 2]) == 0.0
    assert float(diag[1, 0]) == 1.0 and float(diag[1, 1]) == 0.0
    for q, n in enumerate(DN.D75_STATE):
        assert torch.equal(ens.view(out[q], 1, n), ens.view(after[q], 1, n)), n
        assert not torch.equal(ens.view(out[q], 0, n), ens.view(after[q], 0, n)), n


def test_corrector_net_matches_reference_on_gpu(g):
    from neuralpom_amd.hybrid import CorrectorNet
    print("[gpu] CorrectorNet (width 4, seeded) vs the reference forward, fp64", flush=True)
    net = CorrectorNet(hidden_dim=4).double()
    net.load_state_dict({k[6:]: torch.from_numpy(np.array(g[k])) for k in g.files if k.startswith("net/p/")})
    net = net.cuda().eval()
    with torch.no_grad():
        y = net(T(g, "net/x").cuda())
    assert rel_err(y.cpu(), T(g, "net/y")) < 1e-11
