"""The headline workload, the stratified seamount ensemble (32 x 90x82,
24 x 120x110, 8 x 177x161 members, 16 levels, dx = 8 km, dte = 6 s,
isplit = 30, T from N^2 = 1e-5 s^-2 through the model EOS with seeded
noise), on the HIP explicit step against the oracle (the reference algorithm,
pinned bitwise to the reference by test_oracle_golden.py), run with torch on
the same GPU.

* One member of each shape, float64, three chained steps from time0 = 1 day:
  every output field, and the vector-Jacobian product of the three steps
  (through the autograd node, per-step recompute), against the oracle and
  its autograd.  The seamount is ill-conditioned at the rounding level: the
  sigma-coordinate pressure gradient over 2 700 m of relief subtracts nearly
  equal terms of |drhox| ~ 4e3 to drive velocities of ~1e-5 m/s from rest,
  so a one-ulp change of the input temperature moves u by ~1e-9 of its
  magnitude in the reference algorithm itself, and two correct evaluations
  of it differ at that level.  The bound per field is therefore
  max(1e-10, 20 x spread) for the forward and max(1e-8, 20 x spread) for the
  VJP, the spread being the difference between the oracle evaluated by torch
  on the GPU and on the host CPU (other reduction orders and libm).
* The full 64-member float32 batch: every member's state after two steps
  and its input cotangents equal, bitwise, those of the same member run
  alone (batching changes neither the arithmetic nor the order of any
  reduction), all finite, and max |vaf| below vmaxl.
* Six model hours (120 steps of 180 s) of a 90x82 member against the
  float64 oracle: the float64 step within 1e-6 (the rounding spread above,
  grown over 120 steps), the float32 step within 3x the oracle's own float32
  error, field by field.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import cases
from neuralpom_amd.explicit import Ensemble, StepScalars, step_autograd
from oracle.pomref import explicit as X
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu

DTE, ISPLIT, TIME0 = 6.0, 30, 1.0
SHAPES = [(90, 82)] * 32 + [(120, 110)] * 24 + [(177, 161)] * 8
DIFF = ("u", "v", "t", "s", "el", "ua", "va", "q2", "km")


def _scalars():
    return StepScalars(dte=DTE, isplit=ISPLIT, alpha=0.225, smoth=0.10, horcon=0.05, hmax=4500.0)


def _oracle_state(member, dtype, device="cuda"):
    from oracle.pomref.cases import base_state
    st, g, vg, period = member
    im, jm = st["u"].shape[:2]
    S = base_state(im, jm, dtype)
    for k, v in {**st, **g}.items():
        setattr(S, k, v.detach().to(device, dtype).clone())
    for k, v in list(vars(S).items()):
        if torch.is_tensor(v):
            setattr(S, k, v.to(device))
    S.z, S.zz, S.dz, S.dzz = (vg[i].to(device, dtype) for i in range(4))
    sc = _scalars()
    S.hmax, S.dte, S.isplit = sc.hmax, sc.dte, sc.isplit
    S.dti, S.dte2, S.dti2 = sc.dti, 2.0 * sc.dte, 2.0 * sc.dti
    S.ispi, S.isp2i = 1.0 / sc.isplit, 0.5 / sc.isplit
    S.alpha, S.smoth, S.horcon = sc.alpha, sc.smoth, sc.horcon
    S.period = period
    S.time0 = S.time = TIME0
    return S


def _ramps(periods, it, dtype):
    t = TIME0 + DTE * ISPLIT * it / 86400.0
    return torch.tensor([min(t / p, 1.0) for p in periods], dtype=dtype, device="cuda")


def _ensemble(members, dtype):
    ens = Ensemble([tuple(m[0]["u"].shape[:2]) for m in members], dtype=dtype, device="cuda")
    state = {n: ens.pack([m[0][n] for m in members], n) for n in L.FIELDS}
    geom = {n: ens.pack([m[1][n] for m in members], n) for n in L.GEOM}
    vg = torch.stack([m[2] for m in members]).to("cuda", dtype).contiguous()
    return ens, state, geom, vg, [m[3] for m in members]


def _oracle_fwd_vjp(member, w, device="cuda"):
    """Oracle fp64 on ``device``: 3 steps from the member, outputs and the VJP of sum(out * w)."""
    S = _oracle_state(member, torch.float64, device)
    w = {n: v.to(device) for n, v in w.items()}
    ol = {n: getattr(S, n).clone().requires_grad_(True) for n in DIFF}
    for n, v in ol.items():
        setattr(S, n, v)
    S.iend = 3
    assert X.run_time_loop(S)
    out = {n: getattr(S, n).detach().cpu() for n in L.FIELDS}
    loss = sum((getattr(S, n) * w[n].view(getattr(S, n).shape)).sum() for n in w)
    gr = torch.autograd.grad(loss, list(ol.values()))
    return out, {n: g.cpu() for n, g in zip(DIFF, gr)}


@pytest.mark.parametrize("shape", [(90, 82), (120, 110), (177, 161)])
def test_seamount_member_fp64_three_steps_and_vjp_match_oracle(shape):
    print(f"[gpu] seamount member {shape}, fp64, 3 steps fwd + VJP vs oracle", flush=True)
    member = cases.seamount(*shape, seed=1000 + shape[0], dtype=torch.float64)
    ens, state, geom, vg, periods = _ensemble([member], torch.float64)
    sc = _scalars()
    leaves = {n: state[n].clone().requires_grad_(True) for n in DIFF}
    st = {**state, **leaves}
    for it in (1, 2, 3):
        st, vamax = step_autograd(ens, st, geom, sc, _ramps(periods, it, torch.float64), vg, True)
    assert float(vamax) <= sc.vmaxl
    gen = torch.Generator().manual_seed(3)
    w = {n: torch.randn(st[n].shape, generator=gen, dtype=torch.float64).cuda() for n in ("u", "t", "el", "q2")}
    loss = sum((st[n] * w[n]).sum() for n in w)
    grads = dict(zip(DIFF, torch.autograd.grad(loss, list(leaves.values()))))

    ref, rgr = _oracle_fwd_vjp(member, w)
    alt, agr = _oracle_fwd_vjp(member, w, device="cpu")
    bad, worst = [], 0.0
    for n in L.FIELDS:
        spread = rel_err(alt[n], ref[n])
        e = rel_err(ens.view(st[n].detach(), 0, n).cpu(), ref[n])
        worst = max(worst, e)
        if not e <= max(1e-10, 20.0 * spread):
            bad.append(f"{n} {e:.1e} (oracle spread {spread:.1e})")
    print(f"  forward: max rel err {worst:.1e}; oracle cpu/gpu spread up to "
          f"{max(rel_err(alt[n], ref[n]) for n in L.FIELDS):.1e}", flush=True)
    assert not bad, "forward off: " + ", ".join(bad)
    bad, worst = [], 0.0
    for n in DIFF:
        spread = rel_err(agr[n], rgr[n])
        e = rel_err(ens.view(grads[n], 0, n).cpu(), rgr[n])
        worst = max(worst, e)
        if not e <= max(1e-8, 20.0 * spread):
            bad.append(f"{n} {e:.1e} (oracle spread {spread:.1e})")
    print(f"  VJP: max rel err {worst:.1e}; oracle cpu/gpu spread up to "
          f"{max(rel_err(agr[n], rgr[n]) for n in DIFF):.1e}", flush=True)
    assert not bad, "VJP off: " + ", ".join(bad)


def test_seamount_batch_fp32_members_equal_their_single_runs():
    print("[gpu] seamount batch, 64 members fp32: each member vs its own run, 2 steps + adjoint", flush=True)
    members = [cases.seamount(im, jm, seed=1000 + m, dtype=torch.float32, device="cuda")
               for m, (im, jm) in enumerate(SHAPES)]
    sc = _scalars()
    ens, state, geom, vg, periods = _ensemble(members, torch.float32)
    gen = torch.Generator(device="cuda").manual_seed(7)
    cot = {n: 1e-3 * torch.randn(ens.numel(n), generator=gen, device="cuda") for n in ("u", "v", "t", "el", "q2")}
    st = state
    for it in (1, 2):
        st_in = st
        st, vamax = ens.step(st_in, geom, sc, _ramps(periods, it, torch.float32), vg, True)
    grads = ens.step_adj(st_in, geom, sc, _ramps(periods, 2, torch.float32), vg, True, dict(cot))
    torch.cuda.synchronize()
    assert float(vamax) <= sc.vmaxl
    for n in ("u", "t", "el", "q2"):
        assert torch.isfinite(st[n]).all() and torch.isfinite(grads[n]).all(), n
    seen = set()
    for m, (im, jm) in enumerate(SHAPES):
        one, ost, og, ovg, op = _ensemble([members[m]], torch.float32)
        s1 = ost
        for it in (1, 2):
            s1_in = s1
            s1, _ = one.step(s1_in, og, sc, _ramps(op, it, torch.float32), ovg, True)
        c1 = {n: ens.view(c, m, n).reshape(-1).contiguous() for n, c in cot.items()}
        g1 = one.step_adj(s1_in, og, sc, _ramps(op, 2, torch.float32), ovg, True, c1)
        bad = [n for n in L.FIELDS if not torch.equal(one.view(s1[n], 0, n), ens.view(st[n], m, n))]
        bad += [f"grad {n}" for n in L.DIFF_INPUTS if not torch.equal(one.view(g1[n], 0, n), ens.view(grads[n], m, n))]
        assert not bad, f"member {m} {im}x{jm} differs from its own run in {bad}"
        seen.add((im, jm))
    assert seen == {(90, 82), (120, 110), (177, 161)}


def test_seamount_six_hours_fp32_within_the_algorithms_own_error():
    nstep = 120   # 6 h at dti = 180 s
    print(f"[gpu] seamount 90x82, {nstep} steps: HIP fp32 / fp64 vs oracle fp64, oracle fp32", flush=True)
    member = cases.seamount(90, 82, seed=1000, dtype=torch.float64)
    fields = ("el", "u", "v", "t", "s", "ua", "va")
    ref, o32 = {}, {}
    for dtype, dst in ((torch.float64, ref), (torch.float32, o32)):
        S = _oracle_state(member, dtype)
        S.iend = nstep
        with torch.no_grad():
            assert X.run_time_loop(S)
        for n in fields:
            dst[n] = getattr(S, n).double().cpu()
    err_o32 = {n: rel_err(o32[n], ref[n]) for n in fields}
    res = {}
    for dtype in (torch.float64, torch.float32):
        ens, state, geom, vg, periods = _ensemble([member], dtype)
        sc = _scalars()
        for it in range(1, nstep + 1):
            state, vamax = ens.step(state, geom, sc, _ramps(periods, it, dtype), vg, True)
        assert float(vamax) <= sc.vmaxl
        res[dtype] = {n: rel_err(ens.view(state[n], 0, n).double().cpu(), ref[n]) for n in fields}
    print("  6 h max rel err HIP fp64:    " + ", ".join(f"{n} {e:.1e}" for n, e in res[torch.float64].items()))
    print("  6 h max rel err HIP fp32:    " + ", ".join(f"{n} {e:.1e}" for n, e in res[torch.float32].items()))
    print("  6 h max rel err oracle fp32: " + ", ".join(f"{n} {e:.1e}" for n, e in err_o32.items()), flush=True)
    assert all(e <= 1e-6 for e in res[torch.float64].values()), res[torch.float64]
    bad = [n for n in fields if not res[torch.float32][n] <= max(3.0 * err_o32[n], 1e-6)]
    assert not bad, [(n, res[torch.float32][n], err_o32[n]) for n in bad]
