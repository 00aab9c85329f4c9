"""Parity of the HIP explicit step (forward and adjoint) with the reference.

Golden vectors: tests/golden/explicit_{steps,vjp}_12x10.npz written from the
reference's float64 explicit core (tests/golden/make_golden.py).  Batched
ragged members are checked against the oracle (oracle/pomref), itself pinned
bitwise to the same golden vectors by tests/test_oracle_golden.py.
"""
import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from oracle.pomref import explicit as X
from oracle.pomref.state import load_state
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of, unpack
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

SKIP_CMP = set()   # every output field is compared

# max relative error (max |a-b| / max |b|) per precision
TOL = {torch.float64: 1e-10, torch.float32: 2e-4}


def _state(label="a"):
    g = golden("explicit_steps_12x10.npz")
    S = load_state(g, "in")
    S.time0 = float(g[f"{label}/time0"])
    S.time = S.time0
    return S, g


def _run(states, dtype, nstep, internal_first):
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    S0 = states[0]
    for it in range(1, nstep + 1):
        ramp = torch.tensor([ramp_of(S, it) for S in states], dtype=dtype, device="cuda")
        internal = internal_first or it > 1
        state, vamax = ens.step(state, geom, sc, ramp, vg, internal)
    torch.cuda.synchronize()
    return ens, state, vamax


def _compare(ens, state, refs, tol, floor=None):
    """max rel error per field <= tol, or <= 5x the oracle's own error when
    ``floor`` maps field -> error of the oracle evaluated in the same precision."""
    bad = []
    for n in L.FIELDS:
        if n in SKIP_CMP:
            continue
        lim = tol if floor is None else max(tol, 5.0 * floor.get(n, 0.0))
        for m, got in enumerate(unpack(ens, state[n], n)):
            ref = refs[m](n)
            e = rel_err(got.double().cpu(), ref)
            if not (e <= lim):
                bad.append(f"{n}[m{m}]: {e:.2e} (limit {lim:.1e})")
    assert not bad, "fields off: " + ", ".join(bad)


def _oracle_fp32_error(label, nstep):
    """Error of the reference algorithm itself evaluated in float32 (oracle on
    the CPU) against the float64 golden: FP32 cancellation in rho - rmean makes
    drhox/drhoy inherently ~1e-3 relative, whatever the implementation."""
    S, g = _state(label)
    for k, v in list(vars(S).items()):
        if torch.is_tensor(v):
            setattr(S, k, v.float())
    S.iend = nstep
    X.run_time_loop(S)
    return {n: rel_err(getattr(S, n).double(), torch.from_numpy(g[f"{label}/{n}"])) for n in L.FIELDS}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("label", ["a", "b", "c"])
def test_step_matches_reference(dtype, label):
    S, g = _state(label)
    nstep = int(g[f"{label}/nstep"])
    print(f"[gpu] explicit step 12x10 label={label} nstep={nstep} dtype={dtype}", flush=True)
    ens, state, vamax = _run([S], dtype, nstep, internal_first=(S.time0 != 0.0))
    ref = lambda n: torch.from_numpy(g[f"{label}/{n}"])
    floor = _oracle_fp32_error(label, nstep) if dtype == torch.float32 else None
    _compare(ens, state, [ref], TOL[dtype] * (10 if label == "b" else 1), floor)
    assert float(vamax) <= S.vmaxl


def _perturbed(seed, im=None, jm=None):
    S, g = _state("a")
    gen = torch.Generator().manual_seed(seed)
    for n in ("u", "v", "t", "el", "ua", "va"):
        x = getattr(S, n)
        setattr(S, n, x + 0.01 * torch.randn(x.shape, generator=gen, dtype=x.dtype) * x.abs().max())
    S.d = S.h + S.el
    return S


def test_batched_members_match_oracle():
    """Three members in one launch, each checked against its own oracle run."""
    states = [_perturbed(s) for s in (1, 2, 3)]
    print("[gpu] batched 3x(12x10) float64", flush=True)
    ens, state, _ = _run([s.copy() for s in states], torch.float64, 1, True)
    refs = []
    for S in states:
        R = S.copy()
        R.iend = 1
        assert X.run_time_loop(R)
        refs.append(lambda n, R=R: getattr(R, n))
    _compare(ens, state, refs, 1e-10)


VJP_IN = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
          "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
          "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")
VJP_OUT = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
           "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
           "et", "etb", "etf", "rho", "egb", "utb", "vtb", "wubot", "wvbot", "w")


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-8), (torch.float32, 2e-3)])
def test_adjoint_matches_reference_vjp(dtype, tol):
    S, g = _state("a")
    v = golden("explicit_vjp_12x10.npz")
    print(f"[gpu] explicit adjoint 12x10 dtype={dtype}", flush=True)
    ens, state, geom, vg = ensemble_inputs([S], dtype)
    sc = scalars_of(S)
    ramp = torch.tensor([ramp_of(S, 1)], dtype=dtype, device="cuda")
    cot = {n: torch.from_numpy(v[f"cot/{n}"]).reshape(-1).to("cuda", dtype) for n in VJP_OUT}
    grads = ens.step_adj(state, geom, sc, ramp, vg, True, cot)
    torch.cuda.synchronize()
    bad = []
    for n in VJP_IN:
        e = rel_err(grads[n].double().cpu().reshape(v[f"grad/{n}"].shape), torch.from_numpy(v[f"grad/{n}"]))
        if not (e <= tol):
            bad.append(f"{n}: {e:.2e}")
    assert not bad, "gradients off: " + ", ".join(bad)


def test_adjoint_matches_oracle_on_a_larger_ragged_batch():
    """VJP of one step for members large enough to have columns far from every
    edge (the gathers of boundary operators skip off-column offsets there) and
    sizes that are not multiples of the launch blocks, against torch autograd
    through the oracle (itself pinned to the reference's VJP on 12 x 10)."""
    from oracle.pomref.cases import double_gyre
    shapes = ((26, 22), (19, 31))
    states = []
    for n, (im, jm) in enumerate(shapes):
        S = double_gyre(im, jm)
        gen = torch.Generator().manual_seed(40 + n)
        for f, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02), ("ua", 0.01), ("va", 0.01)):
            x = getattr(S, f)
            setattr(S, f, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
        S.d = S.h + S.el
        S.time0 = S.time = 0.25
        states.append(S)
    print(f"[gpu] explicit adjoint, members {shapes}, float64", flush=True)
    ens, state, geom, vg = ensemble_inputs(states, torch.float64)
    sc = scalars_of(states[0])
    ramp = torch.tensor([ramp_of(S, 1) for S in states], dtype=torch.float64, device="cuda")
    gen = torch.Generator().manual_seed(77)
    cots = [{n: torch.randn(getattr(S, n).shape, generator=gen, dtype=torch.float64) for n in VJP_OUT}
            for S in states]
    cot = {n: ens.pack([c[n] for c in cots], n) for n in VJP_OUT}
    grads = ens.step_adj(state, geom, sc, ramp, vg, True, cot)
    torch.cuda.synchronize()
    bad = []
    for m, S in enumerate(states):
        R = S.copy()
        leaves = {n: getattr(R, n).clone().requires_grad_(True) for n in VJP_IN}
        for n, v in leaves.items():
            setattr(R, n, v)
        R.iend = 1
        assert X.run_time_loop(R)
        loss = sum((getattr(R, n) * cots[m][n]).sum() for n in VJP_OUT)
        ref = torch.autograd.grad(loss, list(leaves.values()), allow_unused=True)
        for n, r in zip(VJP_IN, ref):
            if r is None:
                r = torch.zeros_like(leaves[n])
            e = rel_err(unpack(ens, grads[n], n)[m].cpu(), r)
            if not e <= 1e-8:
                bad.append(f"{n}[m{m}]: {e:.2e}")
    assert not bad, "gradients off: " + ", ".join(bad)
