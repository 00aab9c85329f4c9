"""Windows of many steps trained on their own (the daily truncated update of the
coupled rollout): fine semi-implicit members of the open-boundary tidal
channel, their end state restricted 4 x 4 to the forecast grid, the corrector
network and the fused corrector hook there, a loss on the corrected state,
backward through all of it to the window's start, the state detached between
windows.  Float64, two members of 42 x 34 (forecast grid 12 x 10), windows of
six 30 s steps.

* Recomputing runs of three steps in backward (``segment``), or runs of four
  recomputed as runs of two, gives bitwise the losses and gradients of
  per-step recompute.
* Detaching between windows: the second window's loss and start gradient are
  bitwise those of a fresh run started from the first window's end state.
* The start-state gradient of the whole chain (physics, restriction, pack,
  network, hook, loss) against central finite differences of the same chain.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import boundary as B
from neuralpom_amd import cases
from neuralpom_amd import coupling as CP
from neuralpom_amd.explicit import Ensemble
from neuralpom_amd.hybrid import CorrectorNet, integrate, restrict_state, truncated_windows
from neuralpom_amd.implicit import ImplicitEnsemble
from neuralpom_amd.implicit_core import POM2K_Core
from neuralpom_amd.state import INPUT_CHANNELS, pack_input

pytestmark = pytest.mark.gpu

FINE = [(42, 34), (42, 34)]
COARSE = [(12, 10), (12, 10)]
STEPS, SEG = 6, 3
NAMES = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "et", "ua", "uab", "va", "vab")
SCALES = [0.05, 0.04, 0.2, 0.05]
SPEED = 360.0 / 12.42


def _member(im, jm, seed):
    st, g, vg, period, nodes = cases.tidal_channel(im, jm, seed=seed, dtdz=0.05, dtype=torch.float64,
                                                   device="cuda")
    core = POM2K_Core(im, jm, device="cuda", dtype=torch.float64)
    core.configure_runtime(dte=1.0, isplit=30)
    for n, v in {**st, **g}.items():
        setattr(core, n, v)
    core.z, core.zz, core.dz, core.dzz = vg
    core.period = period
    core.horcon = 0.05
    core.double_gyre_wind_stress_pa = 0.0
    core.eta_solver_rel_tol = 1e-13
    core.eta_solver_tol = 0.0
    core.eta_solver_max_iter = 2000
    B.install(core, "implicit", nodes=nodes, speed=SPEED, radiation=0.0)
    core.time0 = core.time = 0.0
    return core


def _schedule(core, members):
    sc = core.step_scalars()

    def at(it):
        core.time = core.dti * it / 86400.0 + core.time0
        r = min(core.time / float(core.period), 1.0) if core.lramp else 1.0
        ramp = torch.full((members,), r, dtype=core.dtype, device=core.device)
        return core._scalars_at_step(sc), ramp, not (it == 1 and core.time0 == 0.0)
    return at


class Setup:
    def __init__(self):
        cores = [_member(im, jm, 3 + m) for m, (im, jm) in enumerate(FINE)]
        self.ens = ImplicitEnsemble(FINE, dtype=torch.float64, device="cuda")
        g0 = [c._geom() for c in cores]
        self.geom = {n: self.ens.pack([g[n] for g in g0], n) for n in g0[0]}
        self.state = {n: self.ens.pack([c._flat(n).detach() for c in cores], n) for n in L.FIELDS}
        self.vg = torch.stack([c._vgrid()[0] for c in cores]).contiguous()
        self.schedule = _schedule(cores[0], len(FINE))
        self.ens.prepare(self.geom, self.schedule(1)[0])
        self.coarse = Ensemble(COARSE, dtype=torch.float64, device="cuda")
        self.cgeom = cases.forecast_geometry(self.ens, self.coarse, self.geom, 2000.0)
        gen = torch.Generator().manual_seed(0)
        self.net = CorrectorNet(hidden_dim=8, head_std=1e-2, generator=gen).to("cuda", torch.float64)
        with torch.no_grad():
            c0 = restrict_state(self.ens, self.coarse, self.state, self.geom, self.vg, NAMES)
        self.target = {n: c0[n] + 0.01 * torch.randn(c0[n].shape, generator=gen, dtype=torch.float64).cuda()
                       for n in ("u", "v", "t", "s")}

    def loss(self, end, window=0):
        cs = restrict_state(self.ens, self.coarse, end, self.geom, self.vg, NAMES)
        x = pack_input(self.coarse, cs)
        (im, jm), m = COARSE[0], len(COARSE)
        raw = self.net(x.view(m, INPUT_CHANNELS, im, jm)).reshape(-1)
        outs = CP._JointCorrection.apply(self.coarse, self.vg, SCALES, 5.0, [self.cgeom[n] for n in CP.CORR_GEOM],
                                         raw, cs["et"], *[cs[n] for n in CP.STATE_FIELDS])
        got = dict(zip(CP.STATE_FIELDS, outs[:8]))
        return sum(((got[n] - self.target[n]) ** 2).mean() / s ** 2
                   for n, s in zip(("u", "v", "t", "s"), SCALES))

    def run(self, windows, segment, state=None, first=1):
        self.net.zero_grad(set_to_none=True)
        losses, grads, end = truncated_windows(self.ens, self.state if state is None else state, self.geom,
                                               self.vg, self.schedule, windows, STEPS, self.loss, first=first,
                                               segment=segment)
        pgrads = [p.grad.clone() for p in self.net.parameters()]
        return losses, grads, end, pgrads


@pytest.fixture(scope="module")
def setup():
    return Setup()


def test_segment_recompute_is_bitwise(setup):
    print(f"[gpu] windows: 2 members {FINE[0]} -> {COARSE[0]}, 2 windows of {STEPS} steps, segment 1 vs {SEG} and (4, 2)",
          flush=True)
    lb, gb, eb, pb = setup.run(2, 1)
    print(f"  window losses {lb}", flush=True)
    for seg in (SEG, (4, 2)):
        la, ga, ea, pa = setup.run(2, seg)
        assert la == lb, seg
        for w in range(2):
            assert set(ga[w]) == set(gb[w]) and ga[w]
            for n in ga[w]:
                assert torch.equal(ga[w][n], gb[w][n]), (seg, w, n)
            assert float(ga[w]["t"].abs().max()) > 0.0
        for a, b in zip(pa, pb):
            assert torch.equal(a, b), seg
        assert any(float(a.abs().max()) > 0.0 for a in pa)
        for n in L.FIELDS:
            assert torch.equal(ea[n], eb[n]), (seg, n)


def test_detach_between_windows(setup):
    print("[gpu] windows: second window equals a fresh run from the first window's end", flush=True)
    la, ga, _, _ = setup.run(2, SEG)
    _, _, e1, _ = setup.run(1, SEG)
    ld, gd, _, _ = setup.run(1, SEG, state=e1, first=STEPS + 1)
    assert ld[0] == la[1]
    for n in gd[0]:
        assert torch.equal(gd[0][n], ga[1][n]), n


def test_window_gradient_matches_finite_differences(setup):
    print("[gpu] windows: start-state gradient of the chain vs central differences", flush=True)
    _, grads, _, _ = setup.run(1, SEG)
    g = grads[0]
    gen = torch.Generator().manual_seed(9)
    scale = {"t": 0.1, "s": 0.01, "u": 0.01, "v": 0.01, "el": 0.01, "ub": 0.01, "tb": 0.1}
    worst = 0.0
    for d in range(3):
        dirs = {n: (torch.randn(setup.state[n].shape, generator=gen, dtype=torch.float64) * s).cuda()
                for n, s in scale.items()}
        dot = sum(float((g[n] * dirs[n]).sum()) for n in dirs)

        def f(eps):
            st = dict(setup.state)
            for n in dirs:
                st[n] = setup.state[n] + eps * dirs[n]
            with torch.no_grad():
                end, _ = integrate(setup.ens, st, setup.geom, setup.vg, setup.schedule, 1, STEPS)
                return float(setup.loss(end))
        eps = 1e-6
        fd = (f(eps) - f(-eps)) / (2 * eps)
        err = abs(fd - dot) / max(abs(fd), 1e-30)
        print(f"  direction {d}: adjoint {dot:.10e} fd {fd:.10e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst <= 1e-4
