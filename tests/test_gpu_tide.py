"""The open-boundary tide policy on the HIP step (neuralpom_amd.boundary).

Golden vectors: tests/golden/implicit_tide_14x12.npz, written from the
reference's semi-implicit core with its coastal tide adapter installed
(src/neuralpom/cores/case_adapters/coastal_boundary.py:60-67), float64: an
open channel with tide nodes on the first interior column, seeded exterior
tracer values, 1- and 3-step trajectories and a seeded VJP of one step.  The
drop-in core gets the same state, the policy through
``neuralpom_amd.boundary.install`` and a tight eta tolerance (the adapter's
CG stops at a relative residual of 1e-12); the standalone eta solve is
compared with the oracle's lifted CG (pinned to the reference in
tests/test_oracle_tide.py) and with autograd through it.
"""
import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import boundary as B
from neuralpom_amd.implicit_core import POM2K_Core
from oracle.pomref import implicit as IM
from oracle.pomref.state import load_state
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

TAG = "14x12"
EDGE = ("ele", "elw", "uabe", "uabw", "eln", "els", "vabn", "vabs", "sbe", "sbw", "tbe", "tbw", "ube", "ubw",
        "sbn", "sbs", "tbn", "tbs", "vbn", "vbs")
VJP_IN = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
          "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
          "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")


@pytest.fixture(scope="module")
def gt():
    return golden(f"implicit_tide_{TAG}.npz")


def _t(g, key, dtype=torch.float64):
    return torch.from_numpy(np.array(g[key])).to("cuda", dtype)


def _core(g, dtype=torch.float64):
    im, jm = g["in/u"].shape[:2]
    core = POM2K_Core(im, jm, device="cuda", dtype=dtype)
    core.configure_runtime(dte=float(g["scalar/dte"]), isplit=int(g["scalar/isplit"]))
    for n in ("alpha", "smoth", "horcon", "hmax", "vmaxl", "umol", "tprni", "small", "period"):
        setattr(core, n, float(g[f"scalar/{n}"]))
    for n in L.FIELDS + L.GEOM + EDGE + ("z", "zz", "dz", "dzz", "north_e"):
        setattr(core, n, _t(g, f"in/{n}", dtype))
    nodes = np.stack([g["tide/i"] + 1, g["tide/j"] + 1, g["tide/amp"], g["tide/phase"]], axis=1)
    B.install(core, "implicit", nodes=nodes)
    core.eta_solver_rel_tol, core.eta_solver_tol = 1e-13, 0.0
    core.time0 = core.time = 0.1
    return core


def test_install_mirrors_the_adapter(gt):
    core = _core(gt)
    assert isinstance(core._implicit_boundary_policy, B.TideBoundary)
    assert (core.rfe, core.rfw, core.rfn, core.rfs) == (1.0, 1.0, 1.0, 1.0)
    assert core._tide_i.tolist() == [1] * (core.jm - 2)
    assert torch.equal(core._tide_amp.cpu(), torch.from_numpy(np.array(gt["in/_tide_amp"])))
    sc = core.step_scalars()
    assert sc.boundary == L.POM_BOUNDARY_TIDE
    sec = 0.1 * 86400 - core.dte
    assert sc.tide_wt == (B.TIDE_SPEED * B.TIDE_PI / 180) * sec / 3600


@pytest.mark.parametrize("label,nstep", [("a", 1), ("b", 3)])
def test_tide_steps_match_reference(gt, label, nstep):
    core = _core(gt)
    core.iint, core.iend = 0, nstep
    print(f"[gpu] tide policy run_time_loop {TAG} steps={nstep}", flush=True)
    core.run_time_loop()
    torch.cuda.synchronize()
    bad = []
    for n in L.FIELDS:
        e = rel_err(getattr(core, n).cpu(), torch.from_numpy(np.array(gt[f"{label}/{n}"])))
        if not e <= 1e-9:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad
    # the nodes carry the tide of the last step
    last = core.time * 86400 - core.dte
    want = core._implicit_boundary_policy.tide(core, last)
    assert rel_err(core.el[core._tide_i, core._tide_j].cpu(), want.cpu()) <= 1e-12


def test_tide_adjoint_matches_reference_vjp(gt):
    core = _core(gt)
    print(f"[gpu] tide policy adjoint {TAG}", flush=True)
    ens = core._ensemble()
    core.time = core.time0 + core.dti / 86400.0
    sc = core.step_scalars()
    geom = core._geom()
    state = {n: core._flat(n).detach() for n in L.FIELDS}
    ramp = torch.tensor([min(core.time / core.period, 1.0)], dtype=torch.float64, device="cuda")
    core._apply_surface_wind()
    geom.update({n: core._flat(n).detach() for n in ("wusurf", "wvsurf")})
    cot = {k[4:]: _t(gt, k).reshape(-1) for k in gt.files if k.startswith("cot/")}
    grads = ens.step_adj(state, geom, sc, ramp, core._vgrid(), True, cot)
    torch.cuda.synchronize()
    bad = []
    for n in VJP_IN:
        ref = torch.from_numpy(np.array(gt[f"grad/{n}"]))
        e = rel_err(grads[n].cpu().reshape(ref.shape), ref)
        if not e <= 1e-7:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def _oracle(g):
    S = load_state(g, "in")
    IM.configure(S)
    IM.install_tide(S, torch.from_numpy(np.array(g["tide/i"])).long(), torch.from_numpy(np.array(g["tide/j"])).long(),
                    torch.from_numpy(np.array(g["tide/amp"])), torch.from_numpy(np.array(g["tide/phase"])))
    S.time0 = S.time = 0.1
    return S


def test_tide_eta_solve_and_gradient(gt):
    core = _core(gt)
    S = _oracle(gt)
    gen = torch.Generator().manual_seed(31)
    rhs = torch.randn((core.im, core.jm), generator=gen, dtype=torch.float64)
    d = (S.h + S.el).clone()
    w = torch.randn((core.im, core.jm), generator=gen, dtype=torch.float64)
    print(f"[gpu] tide policy eta solve {TAG}", flush=True)
    r1 = rhs.clone().requires_grad_(True)
    d1 = d.clone().requires_grad_(True)
    x_ref = IM.tide_eta_solve(S, r1, d1, max_iter=500)
    g_rhs, g_d = torch.autograd.grad((x_ref * w).sum(), [r1, d1])
    r2 = rhs.cuda().requires_grad_(True)
    d2 = d.cuda().requires_grad_(True)
    x = core.solve_implicit_eta_cg(r2, d2, max_iter=500, tol=0.0)
    assert rel_err(x.detach().cpu(), x_ref.detach()) <= 1e-10
    fixed = S.tide.fixed   # held at the tide (device cos against host cos: last-bit differences)
    assert rel_err(x.detach().cpu()[fixed], x_ref.detach()[fixed]) <= 1e-15
    h_rhs, h_d = torch.autograd.grad((x * w.cuda()).sum(), [r2, d2])
    assert rel_err(h_rhs.cpu(), g_rhs) <= 1e-8
    assert rel_err(h_d.cpu(), g_d) <= 1e-8
