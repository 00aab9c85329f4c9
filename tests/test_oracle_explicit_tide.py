"""Pin the oracle's explicit family of the open-boundary tide policy
(oracle/pomref/explicit.py ``tide_bcond1``, oracle/pomref/implicit.py
``install_tide(..., family="explicit")`` and ``tide_bc2``).

Against the reference: vectors written from the reference's explicit core
with its coastal tide adapter installed (explicit family), in float64
(tests/golden/make_golden_explicit_tide.py): an open channel with tide nodes
on the first interior column, a 1-step trajectory, 3 steps in one
``run_time_loop`` call (the substep counter carried on), and 3 steps of the
case registry's one-step protocol (counter reset per step, time0 the current
time; case_adapters/registry.py:54-59).

Against mathematics: with a level surface at rest, bcond 2 copies the
neighbouring transport; the nodes carry amp cos(w t - phase) at the last
substep's time; edges that are not nodes keep their previous elevation.
"""
import numpy as np
import pytest
import torch

from oracle.pomref import explicit as X
from oracle.pomref import implicit as IM
from oracle.pomref.state import load_state
from tests._helpers import golden, rel_err

TAG = "14x12"
DYNAMIC = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab", "va", "vab",
           "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam", "et", "etb", "etf", "rho",
           "w", "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "dt", "d", "l",
           "uaf", "vaf", "elf", "advua", "advva", "adx2d", "ady2d", "drx2d", "dry2d", "aam2d",
           "advx", "advy", "drhox", "drhoy", "fluxua", "fluxva")


@pytest.fixture(scope="module")
def gt():
    return golden(f"explicit_tide_{TAG}.npz")


def T(g, key):
    return torch.from_numpy(np.array(g[key]))


def tide_state(g):
    S = load_state(g, "in")
    IM.install_tide(S, T(g, "tide/i").long(), T(g, "tide/j").long(), T(g, "tide/amp"), T(g, "tide/phase"),
                    family="explicit")
    S.time0 = S.time = 0.1
    return S


def _compare(S, g, label):
    bad = []
    for n in DYNAMIC:
        e = rel_err(getattr(S, n), T(g, f"{label}/{n}"))
        if not e <= 1e-10:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def test_install_matches_reference_settings(gt):
    S = tide_state(gt)
    assert (S.rfe, S.rfw, S.rfn, S.rfs) == (1.0, 1.0, 1.0, 1.0)
    assert float(gt["scalar/rfe"]) == 1.0
    assert S.tide_family == "explicit" and S.tide_substep == 0


@pytest.mark.parametrize("label,nstep", [("a", 1), ("b", 3)])
def test_tide_steps_match_reference(gt, label, nstep):
    S = tide_state(gt)
    S.iend = nstep
    with torch.no_grad():
        assert X.run_time_loop(S)
    assert S.tide_substep == int(S.isplit) * nstep
    if label == "b":
        assert S.tide_substep == int(gt["b/substep"])
    _compare(S, gt, label)


def test_registry_protocol_matches_reference(gt):
    S = tide_state(gt)
    with torch.no_grad():
        for _ in range(3):
            S.time0 = S.time
            S.iend = 1
            S.tide_substep = 0
            assert X.run_time_loop(S)
    assert S.tide_substep == int(gt["c/substep"])
    assert abs(S.time - float(gt["c/time"])) <= 1e-15
    _compare(S, gt, "c")


# ---------------------------------------------------------------- closed forms
def test_nodes_carry_the_last_substeps_tide(gt):
    S = tide_state(gt)
    S.iend = 1
    with torch.no_grad():
        X.run_time_loop(S)
    sec = S.time * 86400 - S.dti + S.dte * (int(S.isplit) - 1)
    w = (IM.TIDE_SPEED * IM.TIDE_PI / 180) / 3600
    want = S.tide.amp * torch.cos(w * sec - S.tide.phase * IM.TIDE_PI / 180)
    got = S.elf[S.tide.i, S.tide.j]
    assert rel_err(got, want * S.fsm[S.tide.i, S.tide.j]) <= 1e-13


def test_non_node_edges_keep_their_elevation(gt):
    S = tide_state(gt)
    elf0 = S.elf.clone()
    S.iend = 1
    with torch.no_grad():
        X.run_time_loop(S)
    # the east edge column and the west edge column are not nodes: bcond 1
    # neither copies the interior onto them nor changes them (other than fsm)
    for i in (0, S.im - 1):
        assert torch.equal(S.elf[i], elf0[i] * S.fsm[i])


def test_bcond2_copies_the_neighbour_at_rest():
    g = torch.Generator().manual_seed(5)
    S = X.State()
    im, jm = 7, 6
    S.grav = 9.806
    S.rfe = S.rfw = S.rfn = S.rfs = 1.0
    S.dum = torch.ones(im, jm, dtype=torch.float64)
    S.dvm = torch.ones(im, jm, dtype=torch.float64)
    el = torch.full((im, jm), 0.25, dtype=torch.float64)
    d = 10.0 + torch.rand((im, jm), generator=g, dtype=torch.float64)
    uaf = torch.randn((im, jm), generator=g, dtype=torch.float64)
    vaf = torch.randn((im, jm), generator=g, dtype=torch.float64)
    u2, v2 = IM.tide_bc2(S, uaf, vaf, el=el, d=d)
    J = slice(1, -1)
    assert torch.equal(u2[-1, J], uaf[-2, J])
    assert torch.equal(u2[1, J], uaf[1, J]) and torch.equal(u2[0, J], uaf[1, J])
    assert torch.equal(v2[J, -1], vaf[J, -2])
    assert torch.equal(v2[J, 1], vaf[J, 1]) and torch.equal(v2[J, 0], vaf[J, 1])
    assert not u2[J, 0].any() and not u2[J, -1].any() and not v2[0, J].any() and not v2[-1, J].any()
