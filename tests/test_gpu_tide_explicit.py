"""The explicit family of the open-boundary tide policy on the HIP step.

Golden vectors: tests/golden/explicit_tide_14x12.npz, written from the
reference's explicit core with its coastal tide adapter installed
(explicit family, src/neuralpom/cores/case_adapters/coastal_boundary.py:11-30,
55-62), float64: an open channel with tide nodes on the first interior
column and seeded exterior tracer values; one step, three steps in one
``run_time_loop`` call (the adapter's substep counter carried on) and three
steps of the case registry's one-step protocol (counter reset, time0 the
current time; case_adapters/registry.py:54-59).  The drop-in explicit core
gets the same state and the policy through ``neuralpom_amd.boundary.install``.
The family is forward-only, as in the reference: its adjoint entry points
refuse it and autograd through it raises.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import boundary as B
from neuralpom_amd.core import POM2K_Core
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

TAG = "14x12"
EDGE = ("ele", "elw", "uabe", "uabw", "eln", "els", "vabn", "vabs", "sbe", "sbw", "tbe", "tbw", "ube", "ubw",
        "sbn", "sbs", "tbn", "tbs", "vbn", "vbs")


@pytest.fixture(scope="module")
def gt():
    return golden(f"explicit_tide_{TAG}.npz")


def _t(g, key, dtype=torch.float64):
    return torch.from_numpy(np.array(g[key])).to("cuda", dtype)


def _core(g, dtype=torch.float64):
    im, jm = g["in/u"].shape[:2]
    core = POM2K_Core(im, jm, device="cuda", dtype=dtype)
    core.configure_runtime(dte=float(g["scalar/dte"]), isplit=int(g["scalar/isplit"]))
    for n in ("alpha", "smoth", "horcon", "hmax", "vmaxl", "umol", "tprni", "small", "period"):
        setattr(core, n, float(g[f"scalar/{n}"]))
    for n in L.FIELDS + L.GEOM + EDGE + ("z", "zz", "dz", "dzz"):
        setattr(core, n, _t(g, f"in/{n}", dtype))
    nodes = np.stack([g["tide/i"] + 1, g["tide/j"] + 1, g["tide/amp"], g["tide/phase"]], axis=1)
    B.install(core, "explicit", nodes=nodes)
    core.time0 = core.time = 0.1
    return core


def _compare(core, g, label, tol):
    bad = []
    for n in L.FIELDS:
        e = rel_err(getattr(core, n).cpu().double(), torch.from_numpy(np.array(g[f"{label}/{n}"])))
        if not e <= tol:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def test_install_mirrors_the_adapter(gt):
    core = _core(gt)
    assert isinstance(core._explicit_boundary_policy, B.TideBoundary)
    assert core._validation_family == "explicit" and core._tide_substep == 0
    assert (core.rfe, core.rfw, core.rfn, core.rfs) == (1.0, 1.0, 1.0, 1.0)
    sc = core.step_scalars()
    assert sc.boundary == L.POM_BOUNDARY_TIDE
    assert sc.tide_coef == B.TIDE_SPEED * B.TIDE_PI / 180
    with pytest.raises(NotImplementedError):
        B.install(POM2K_Core(6, 6, device="cpu"), "implicit", nodes=[[2, 2, 0.1, 0.0]])


@pytest.mark.parametrize("label,nstep", [("a", 1), ("b", 3)])
def test_tide_steps_match_reference(gt, label, nstep):
    core = _core(gt)
    core.iint, core.iend = 0, nstep
    print(f"[gpu] explicit tide run_time_loop {TAG} steps={nstep}", flush=True)
    with torch.no_grad():
        core.run_time_loop()
    torch.cuda.synchronize()
    assert core._tide_substep == nstep * core.isplit
    if label == "b":
        assert core._tide_substep == int(gt["b/substep"])
    _compare(core, gt, label, 1e-9)
    # the nodes carry the tide of the last substep
    tide = core._explicit_boundary_policy
    sec = tide.substep_seconds(core, core._tide_substep)
    want = tide.tide(core, sec)
    assert rel_err(core.elf[core._tide_i, core._tide_j].cpu(), want.cpu()) <= 1e-12


def test_registry_protocol_matches_reference(gt):
    core = _core(gt)
    print(f"[gpu] explicit tide registry protocol {TAG} steps=3", flush=True)
    with torch.no_grad():
        for _ in range(3):   # case_adapters/registry.py:54-59
            core.time0 = float(core.time)
            core.iint, core.iend = 0, 1
            core._tide_substep = 0
            core.run_time_loop()
    torch.cuda.synchronize()
    assert core._tide_substep == int(gt["c/substep"])
    assert abs(core.time - float(gt["c/time"])) <= 1e-15
    _compare(core, gt, "c", 1e-9)


def test_fp32_tracks_reference(gt):
    core = _core(gt, torch.float32)
    core.iint, core.iend = 0, 3
    print(f"[gpu] explicit tide run_time_loop {TAG} fp32 steps=3", flush=True)
    with torch.no_grad():
        core.run_time_loop()
    torch.cuda.synchronize()
    for n in ("el", "elb", "ua", "va", "u", "v", "t", "s"):
        e = rel_err(getattr(core, n).cpu().double(), torch.from_numpy(np.array(gt[f"b/{n}"])))
        assert e <= 2e-3, f"{n}: {e:.2e}"


def test_forward_only(gt):
    core = _core(gt)
    core.iint, core.iend = 0, 1
    core.u = core.u.clone().requires_grad_(True)
    with pytest.raises(NotImplementedError):
        core.run_time_loop()
    # the adjoint entry points refuse the family
    ens = core._ensemble()
    sc = core.step_scalars()
    n = C.c_size_t()
    assert ens.lib.pom_explicit_workspace(ens._h, C.byref(sc.struct()), 1, C.byref(n)) == L.POM_E_UNSUPPORTED
    assert ens.lib.pom_explicit_workspace(ens._h, C.byref(sc.struct()), 0, C.byref(n)) == L.POM_OK
    with pytest.raises(NotImplementedError):
        core.bcond(1, core.elf, core.uaf, core.vaf, core.u, core.v, core.w)
