"""24 model hours of the shelf basin (90 x 82 x 16, dx = dy = 10 km, the
tanh shelf, the cos wind, T = 5 + 15 exp(z / 500 m), S = 35; dte = 30 s,
isplit = 30, so 96 internal steps of 900 s from rest) on the HIP explicit
step against the oracle (the reference algorithm, float64, run with torch on
the same GPU).

* float64 HIP against float64 oracle: the same algorithm in the same
  precision, so the surface elevation, velocities and tracers agree to
  rounding after 2 880 barotropic substeps.
* float32 HIP against the float64 oracle, beside the oracle itself run in
  float32: the float32 error of the reference algorithm over a day is what a
  float32 implementation can reach (the surface elevation enters the step
  through d = h + el, whose float32 resolution at h = 2000 m is 1.2e-4 m
  against |el| of a few mm), so the HIP float32 error is bounded by a
  multiple of it, field by field.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import cases
from neuralpom_amd.explicit import Ensemble, StepScalars
from oracle.pomref import explicit as X
from oracle.pomref.cases import double_gyre as oracle_dg
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu

IM, JM, NSTEP = 90, 82, 96
FIELDS = ("el", "u", "v", "t", "s")


def _oracle(dtype=torch.float64):
    case = cases.shelf_basin(IM, JM, dtype=torch.float64)
    S = oracle_dg(IM, JM, dte=30.0, isplit=30)
    st, g, vg, period = case
    for n, v in {**st, **g}.items():
        setattr(S, n, v.clone())
    S.z, S.zz, S.dz, S.dzz = (vg[i].clone() for i in range(4))
    S.period = period
    for k, v in list(vars(S).items()):
        if torch.is_tensor(v):
            setattr(S, k, v.to("cuda", dtype))
    S.time0 = S.time = 0.0
    return S, case


def _hip(case, S, dtype):
    st, g, vg, period = case
    ens = Ensemble([(IM, JM)], dtype=dtype, device="cuda")
    state = {n: ens.pack([st[n]], n) for n in L.FIELDS}
    geom = {n: ens.pack([g[n]], n) for n in L.GEOM}
    vgr = vg.unsqueeze(0).to("cuda", dtype).contiguous()
    sc = StepScalars(dte=30.0, isplit=30, alpha=S.alpha, smoth=S.smoth, horcon=S.horcon, hmax=S.hmax)
    for it in range(1, NSTEP + 1):
        t = sc.dti * it / 86400.0
        ramp = torch.tensor([min(t / period, 1.0)], dtype=dtype, device="cuda")
        state, vamax = ens.step(state, geom, sc, ramp, vgr, it > 1)
    assert float(vamax) <= sc.vmaxl
    torch.cuda.synchronize()
    return {n: ens.view(state[n], 0, n).double().cpu() for n in FIELDS}


def test_shelf_basin_24h_fp64_and_fp32_against_oracle():
    print(f"[gpu] shelf basin {IM}x{JM}, 24 h ({NSTEP} steps): oracle fp64 on cuda, HIP fp64 and fp32", flush=True)
    S, case = _oracle()
    S.iend = NSTEP
    with torch.no_grad():
        assert X.run_time_loop(S)
    ref = {n: getattr(S, n).cpu() for n in FIELDS}
    S32, _ = _oracle(torch.float32)
    S32.iend = NSTEP
    with torch.no_grad():
        assert X.run_time_loop(S32)
    o32 = {n: rel_err(getattr(S32, n).double().cpu(), ref[n]) for n in FIELDS}
    h64 = _hip(case, S, torch.float64)
    h32 = _hip(case, S, torch.float32)
    e64 = {n: rel_err(h64[n], ref[n]) for n in FIELDS}
    e32 = {n: rel_err(h32[n], ref[n]) for n in FIELDS}
    print("[gpu] 24 h max rel err, HIP fp64:    " + ", ".join(f"{n} {e:.1e}" for n, e in e64.items()), flush=True)
    print("[gpu] 24 h max rel err, HIP fp32:    " + ", ".join(f"{n} {e:.1e}" for n, e in e32.items()), flush=True)
    print("[gpu] 24 h max rel err, oracle fp32: " + ", ".join(f"{n} {e:.1e}" for n, e in o32.items()), flush=True)
    assert all(e <= 1e-9 for e in e64.values()), e64
    bad = [n for n in FIELDS if not e32[n] <= max(3.0 * o32[n], 1e-6)]
    assert not bad, f"fp32 error beyond 3x the algorithm's own: {[(n, e32[n], o32[n]) for n in bad]}"
