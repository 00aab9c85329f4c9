"""Pin the oracle's open-boundary tide policy (oracle/pomref/implicit.py, ``install_tide``).

Against the reference: vectors written from the reference's semi-implicit core
with its coastal tide adapter installed, in float64
(tests/golden/make_golden_tide.py): 1- and 3-step trajectories of an open
channel with tide nodes on the first interior column, the adapter's CG
iteration counts, and a seeded VJP through the whole step (autograd through
the adapter's CG).

Against mathematics: the lifted CG equals a dense solve of the free-node
system with the nodes held at the tide; the radiation conditions copy the
neighbouring transport when the surface is level; the tide is amp cos(-phase)
at t = 0.
"""
import numpy as np
import pytest
import torch

from oracle.pomref import implicit as IM
from oracle.pomref.state import load_state
from tests._helpers import assert_rel, golden, rel_err

TAG = "14x12"
DYNAMIC = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab", "va", "vab",
           "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam", "et", "etb", "etf", "rho",
           "w", "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "dt", "d", "l",
           "uaf", "vaf", "elf", "advua", "advva", "adx2d", "ady2d", "drx2d", "dry2d", "aam2d",
           "advx", "advy", "drhox", "drhoy", "fluxua", "fluxva", "wusurf", "wvsurf")
VJP_IN = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
          "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
          "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")


@pytest.fixture(scope="module")
def tide():
    return golden(f"implicit_tide_{TAG}.npz")


def T(g, key):
    return torch.from_numpy(np.array(g[key]))


def tide_state(g):
    S = load_state(g, "in")
    IM.configure(S)
    IM.install_tide(S, T(g, "tide/i").long(), T(g, "tide/j").long(), T(g, "tide/amp"), T(g, "tide/phase"))
    S.time0 = S.time = 0.1
    return S


def test_install_matches_reference_settings(tide):
    S = tide_state(tide)
    assert (S.rfe, S.rfw, S.rfn, S.rfs) == (1.0, 1.0, 1.0, 1.0)
    assert float(tide["scalar/rfe"]) == 1.0
    # nodes on the first interior column; the fixed set is the ring, dry cells and the nodes
    assert set(T(tide, "tide/i").long().tolist()) == {1}
    fixed = S.tide.fixed
    assert bool(fixed[0].all()) and bool(fixed[:, 0].all()) and bool(fixed[1, 1:-1].all())
    assert not bool(fixed[2:-1, 1:-1].any())


@pytest.mark.parametrize("label", ["a", "b"])
def test_tide_steps_match_reference(tide, label):
    S = tide_state(tide)
    S.iend = 1 if label == "a" else 3
    S.eta_iters = []
    IM.run_time_loop(S)
    assert S.eta_iters == list(tide[f"{label}/iters"])
    bad = []
    for n in DYNAMIC:
        e = rel_err(getattr(S, n), T(tide, f"{label}/{n}"))
        if not e <= 1e-10:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def test_tide_vjp_matches_reference(tide):
    S = tide_state(tide)
    S.iend = 1
    leaves = {n: getattr(S, n).clone().requires_grad_(True) for n in VJP_IN}
    for n, x in leaves.items():
        setattr(S, n, x.clone())
    IM.run_time_loop(S)
    loss = sum((getattr(S, k[4:]) * T(tide, k)).sum() for k in tide.files if k.startswith("cot/"))
    grads = torch.autograd.grad(loss, [leaves[n] for n in VJP_IN], allow_unused=True)
    bad = []
    for n, g in zip(VJP_IN, grads):
        g = torch.zeros_like(leaves[n]) if g is None else g
        e = rel_err(g, T(tide, f"grad/{n}"))
        if not e <= 1e-9:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


# ---------------------------------------------------------------- closed forms
def test_lifted_cg_equals_dense_free_node_solve(tide):
    S = tide_state(tide)
    g = torch.Generator().manual_seed(21)
    rhs = torch.randn((S.im, S.jm), generator=g, dtype=torch.float64)
    d = S.h + S.el
    x = IM.tide_eta_solve(S, rhs, d, max_iter=500)
    free = ~S.tide.fixed
    idx = free.reshape(-1).nonzero().view(-1)
    lift = torch.zeros_like(rhs)
    lift[S.tide.i, S.tide.j] = IM.tide_value(S, IM.tide_seconds(S))
    n = S.im * S.jm
    cols = []
    for q in range(n):
        e = torch.zeros(n, dtype=torch.float64)
        e[q] = 1.0
        cols.append(IM.elliptic_apply(S, e.view(S.im, S.jm), d).reshape(-1))
    A = torch.stack(cols, dim=1)
    Aff = A[idx][:, idx]
    b = rhs.reshape(-1)[idx] - (A @ lift.reshape(-1))[idx]
    want = lift.reshape(-1).clone()
    want[idx] = torch.linalg.solve(Aff, b)
    assert_rel(x.reshape(-1), want, 1e-10, "lifted solve")
    assert torch.equal(x[S.tide.fixed], lift[S.tide.fixed])


def test_radiation_copies_neighbour_transport_on_a_level_surface(tide):
    S = tide_state(tide)
    S.el = torch.full_like(S.el, 0.25)
    g = torch.Generator().manual_seed(22)
    ua = torch.randn((S.im, S.jm), generator=g, dtype=torch.float64)
    va = torch.randn((S.im, S.jm), generator=g, dtype=torch.float64)
    u, v = IM.tide_bc2(S, ua, va)
    J, I = slice(1, -1), slice(1, -1)
    assert_rel(u[-1, J], ua[-2, J] * S.dum[-1, J], 1e-15, "east")
    assert_rel(u[0, J], ua[1, J] * S.dum[0, J], 1e-15, "west")
    assert_rel(v[I, -1], va[I, -2] * S.dvm[I, -1], 1e-15, "north")
    assert_rel(v[I, 0], va[I, 1] * S.dvm[I, 0], 1e-15, "south")
    assert float(v[-1, J].abs().max()) == 0.0 and float(u[I, 0].abs().max()) == 0.0


def test_tide_value_at_zero_time_is_amp_cos_phase(tide):
    S = tide_state(tide)
    got = IM.tide_value(S, 0.0)
    want = S.tide.amp * torch.cos(-S.tide.phase * (IM.TIDE_PI / 180.0))
    assert_rel(got, want, 1e-15, "tide(0)")
    # one period later (360 / speed hours) the value repeats up to the adapter's pi
    per = 360.0 / IM.TIDE_SPEED * 3600.0
    assert_rel(IM.tide_value(S, per), got, 1e-4, "periodic")
