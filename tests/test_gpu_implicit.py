"""Parity of the HIP semi-implicit (P0-fixed) step, forward and adjoint, with the reference.

Golden vectors: tests/golden/implicit_{steps,vjp}_12x10.npz written from the
reference's float64 implicit core (tests/golden/make_golden_implicit.py).
Ragged batches are checked against the oracle (oracle/pomref/implicit.py),
itself pinned to those vectors and to dense solves (test_oracle_implicit.py).
The double-gyre wind of the implicit forcing policy (forcing/surface.py:73-84)
is time-invariant, so it enters as the geometry fields wusurf / wvsurf.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.implicit import ImplicitEnsemble, ImplicitScalars
from oracle.pomref import implicit as IM
from oracle.pomref.state import load_state
from tests._gpu_inputs import SCAL, ramp_of, unpack
from tests._helpers import golden, rel_err

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-10, torch.float32: 2e-4}


def _state(label="a", tight=False):
    g = golden("implicit_steps_12x10.npz")
    S = load_state(g, "in")
    IM.configure(S)
    if tight:
        S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    S.time0 = S.time = float(g[f"{label}/time0"])
    _wind(S)
    return S, g


def _wind(S):
    """The implicit core's forcing rewrite (time-invariant): wusurf from the wind, wvsurf = 0."""
    arg = 2.0 * S.pi * (S.north_e - 30.0) / 20.0
    S.wusurf = -(0.1 / S.rhoref) * (1.0 - torch.cos(arg)) * S.dum
    S.wvsurf = torch.zeros_like(S.wvsurf)


def _scalars(S):
    kw = {n: getattr(S, n) for n in SCAL}
    kw.update(eta_tol=S.eta_solver_tol, eta_rel_tol=S.eta_solver_rel_tol, eta_max_iter=S.eta_solver_max_iter,
              eta_preconditioner=S.eta_solver_preconditioner, barotropic_shapiro=S.barotropic_shapiro)
    return ImplicitScalars(**kw)


def _inputs(states, dtype):
    ens = ImplicitEnsemble([(S.im, S.jm) for S in states], dtype=dtype, device="cuda")
    state = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.FIELDS}
    geom = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.GEOM}
    vg = torch.stack([torch.stack([S.z, S.zz, S.dz, S.dzz]) for S in states]).to("cuda", dtype).contiguous()
    sc = _scalars(states[0])
    ens.prepare(geom, sc)
    return ens, state, geom, vg, sc


def _run(states, dtype, nstep, internal_first):
    ens, state, geom, vg, sc = _inputs(states, dtype)
    iters = torch.zeros(len(states), dtype=torch.int32, device="cuda")
    its = []
    for it in range(1, nstep + 1):
        ramp = torch.tensor([ramp_of(S, it) for S in states], dtype=dtype, device="cuda")
        state, vamax = ens.step(state, geom, sc, ramp, vg, internal_first or it > 1, iters)
        its.append(iters.cpu().tolist())
    torch.cuda.synchronize()
    return ens, state, vamax, its


def _oracle_fp32_error(label, nstep, tight):
    S, g = _state(label, tight)
    for k, v in list(vars(S).items()):
        if torch.is_tensor(v) and v.is_floating_point():
            setattr(S, k, v.float())
    S.iend = nstep
    IM.run_time_loop(S)
    return {n: rel_err(getattr(S, n).double(), torch.from_numpy(g[f"{label}/{n}"])) for n in L.FIELDS}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("label", ["a", "b", "c", "t"])
def test_step_matches_reference(dtype, label):
    S, g = _state(label, tight=(label == "t"))
    nstep = int(g[f"{label}/nstep"])
    print(f"[gpu] implicit step 12x10 label={label} nstep={nstep} dtype={dtype}", flush=True)
    ens, state, vamax, its = _run([S], dtype, nstep, internal_first=(S.time0 != 0.0))
    if dtype == torch.float64:
        assert [i[0] for i in its] == list(g[f"{label}/iters"]), "eta-solve iteration counts differ"
    floor = _oracle_fp32_error(label, nstep, label == "t") if dtype == torch.float32 else None
    tol = TOL[dtype] * (10 if label == "b" else 1)
    bad = []
    for n in L.FIELDS:
        ref = torch.from_numpy(g[f"{label}/{n}"])
        got = unpack(ens, state[n], n)[0].double().cpu()
        lim = tol if floor is None else max(tol, 5.0 * floor.get(n, 0.0))
        e = rel_err(got, ref)
        if not e <= lim:
            bad.append(f"{n}: {e:.2e} (limit {lim:.1e})")
    assert not bad, "fields off: " + ", ".join(bad)
    assert float(vamax) <= S.vmaxl


def _perturbed(seed):
    S, g = _state("a")
    gen = torch.Generator().manual_seed(seed)
    for n in ("u", "v", "t", "el", "ua", "va"):
        x = getattr(S, n)
        setattr(S, n, x + 0.01 * torch.randn(x.shape, generator=gen, dtype=x.dtype) * x.abs().max())
    S.d = S.h + S.el
    return S


@pytest.mark.parametrize("members", [3, 12])
def test_batched_members_match_oracle(members):
    """Members each stopping its eta solve at its own convergence.  Twelve
    members make twelve 2-D blocks, more than the eight XCDs, so the logical
    block order differs from the dispatch order and each member's dot-product
    partials must still be summed over that member's blocks alone."""
    states = [_perturbed(s) for s in range(1, members + 1)]
    print(f"[gpu] implicit batched {members}x(12x10) float64", flush=True)
    ens, state, _, its = _run([s.copy() for s in states], torch.float64, 2, True)
    bad = []
    for m, S in enumerate(states):
        R = S.copy()
        R.iend = 2
        R.eta_iters = []
        IM.run_time_loop(R)
        assert [i[m] for i in its] == R.eta_iters, f"member {m} iterations"
        for n in L.FIELDS:
            e = rel_err(unpack(ens, state[n], n)[m].cpu(), getattr(R, n))
            if not e <= 1e-10:
                bad.append(f"{n}[m{m}]: {e:.2e}")
    assert not bad, bad


VJP_IN = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
          "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
          "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-8), (torch.float32, 5e-3)])
def test_adjoint_matches_reference_vjp(dtype, tol):
    S, g = _state("a", tight=True)
    v = golden("implicit_vjp_12x10.npz")
    print(f"[gpu] implicit adjoint 12x10 dtype={dtype}", flush=True)
    ens, state, geom, vg, sc = _inputs([S], dtype)
    ramp = torch.tensor([ramp_of(S, 1)], dtype=dtype, device="cuda")
    cot = {k[4:]: torch.from_numpy(v[k]).reshape(-1).to("cuda", dtype) for k in v.files if k.startswith("cot/")}
    grads = ens.step_adj(state, geom, sc, ramp, vg, True, cot)
    torch.cuda.synchronize()
    bad = []
    for n in VJP_IN:
        e = rel_err(grads[n].double().cpu().reshape(v[f"grad/{n}"].shape), torch.from_numpy(v[f"grad/{n}"]))
        if not e <= tol:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, "gradients off: " + ", ".join(bad)


def _shapiro_state():
    g = golden("implicit_shapiro_12x10.npz")
    S = load_state(g, "in")
    IM.configure(S)
    S.barotropic_shapiro = S.tracer_shapiro = S.velocity_shapiro = float(g["scalar/shapiro_gamma"])
    S.time0 = S.time = 0.1
    _wind(S)
    return S, g


def _shapiro_scalars(S):
    import dataclasses
    return dataclasses.replace(_scalars(S), tracer_shapiro=S.tracer_shapiro, velocity_shapiro=S.velocity_shapiro)


@pytest.mark.parametrize("label", ["a", "b"])
def test_shapiro_step_matches_reference(label):
    """Spatial Shapiro filters on (weights 0.05 on eta, T, S and the filtered
    u, v): the fp64 step against the reference (1 and 3 steps)."""
    S, g = _shapiro_state()
    nstep = 1 if label == "a" else 3
    print(f"[gpu] implicit step with Shapiro filters, {nstep} step(s)", flush=True)
    ens, state, geom, vg, _ = _inputs([S], torch.float64)
    sc = _shapiro_scalars(S)
    iters = torch.zeros(1, dtype=torch.int32, device="cuda")
    its = []
    for it in range(1, nstep + 1):
        ramp = torch.tensor([ramp_of(S, it)], dtype=torch.float64, device="cuda")
        state, vamax = ens.step(state, geom, sc, ramp, vg, True, iters)
        its.append(int(iters[0]))
    assert its == list(g[f"{label}/iters"])
    bad = []
    for n in L.FIELDS:
        e = rel_err(unpack(ens, state[n], n)[0].cpu(), torch.from_numpy(g[f"{label}/{n}"]))
        if not e <= (1e-9 if label == "b" else 1e-10):
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def test_shapiro_adjoint_matches_reference_vjp():
    S, g = _shapiro_state()
    S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    print("[gpu] implicit adjoint with Shapiro filters", flush=True)
    ens, state, geom, vg, _ = _inputs([S], torch.float64)
    sc = _shapiro_scalars(S)
    ramp = torch.tensor([ramp_of(S, 1)], dtype=torch.float64, device="cuda")
    cot = {k[4:]: torch.from_numpy(g[k]).reshape(-1).to("cuda", torch.float64) for k in g.files
           if k.startswith("cot/")}
    grads = ens.step_adj(state, geom, sc, ramp, vg, True, cot)
    torch.cuda.synchronize()
    bad = []
    for n in VJP_IN:
        e = rel_err(grads[n].cpu().reshape(g[f"grad/{n}"].shape), torch.from_numpy(g[f"grad/{n}"]))
        if not e <= 1e-8:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, "gradients off: " + ", ".join(bad)


def _floor_state():
    g = golden("implicit_shapiro_12x10.npz")
    S = load_state(g, "in")
    IM.configure(S)
    S.aam_min = float(g["scalar/aam_min"])
    S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    S.time0 = S.time = 0.1
    _wind(S)
    return S, g


def test_smagorinsky_floor_step_and_vjp_match_reference():
    """aam_min = 800 m^2/s (implicit_p0_fixed/time_loop.py:186-188): one fp64
    step and its VJP against the reference, Shapiro filters off."""
    import dataclasses
    S, g = _floor_state()
    print("[gpu] implicit step + adjoint with the Smagorinsky floor", flush=True)
    ens, state, geom, vg, sc = _inputs([S], torch.float64)
    sc = dataclasses.replace(sc, aam_min=S.aam_min)
    iters = torch.zeros(1, dtype=torch.int32, device="cuda")
    ramp = torch.tensor([ramp_of(S, 1)], dtype=torch.float64, device="cuda")
    out, _ = ens.step(state, geom, sc, ramp, vg, True, iters)
    assert [int(iters[0])] == list(g["m/iters"])
    bad = []
    for n in L.FIELDS:
        e = rel_err(unpack(ens, out[n], n)[0].cpu(), torch.from_numpy(g[f"m/{n}"]))
        if not e <= 1e-10:
            bad.append(f"{n}: {e:.2e}")
    cot = {k[5:]: torch.from_numpy(g[k]).reshape(-1).to("cuda", torch.float64) for k in g.files
           if k.startswith("mcot/")}
    grads = ens.step_adj(state, geom, sc, ramp, vg, True, cot)
    torch.cuda.synchronize()
    for n in VJP_IN:
        e = rel_err(grads[n].cpu().reshape(g[f"mgrad/{n}"].shape), torch.from_numpy(g[f"mgrad/{n}"]))
        if not e <= 1e-8:
            bad.append(f"grad {n}: {e:.2e}")
    assert not bad, bad
