"""The explicit step recorded into a HIP graph (``StepGraph``) against the
same steps run eagerly.

Each replay is checked, with eager work and host synchronisation between the
replays (graph nodes that only the first replay gets right would show here),
for the advanced state and the adjoint's input cotangents.  The graph runs
the same kernels in the same order, so results are bitwise equal.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.explicit import StepGraph
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of

pytestmark = pytest.mark.gpu


def _members():
    from oracle.pomref.cases import double_gyre
    states = []
    for n, (im, jm) in enumerate(((26, 22), (19, 31), (12, 10))):
        S = double_gyre(im, jm)
        gen = torch.Generator().manual_seed(60 + n)
        for f, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02), ("ua", 0.01), ("va", 0.01)):
            x = getattr(S, f)
            setattr(S, f, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
        S.d = S.h + S.el
        S.time0 = S.time = 0.01   # ramp still rising: each step has its own ramp
        states.append(S)
    return states


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_graph_replays_match_eager_steps(dtype):
    print(f"[gpu] StepGraph vs eager, 3 members, dtype={dtype}", flush=True)
    states = _members()
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    ramps = [torch.tensor([ramp_of(S, it) for S in states], dtype=dtype, device="cuda") for it in (1, 2, 3, 4)]
    gen = torch.Generator(device="cuda").manual_seed(3)
    cot = {n: torch.randn(ens.numel(n), generator=gen, device="cuda", dtype=dtype)
           for n in ("u", "v", "t", "s", "el", "ua", "va", "q2", "km")}

    G = StepGraph(ens, state, geom, sc, ramps[0], vg, True, cot)
    eager = {n: state[n].clone() for n in L.FIELDS}
    for r in range(3):
        G.ramp.copy_(ramps[r])
        vamax = G.replay()
        # eager reference of the same step, with unrelated eager work in between
        out, vamax_e = ens.step(eager, geom, sc, ramps[r], vg, True)
        grads = ens.step_adj(eager, geom, sc, ramps[r], vg, True, dict(cot))
        torch.cuda.synchronize()
        bad = [n for n in L.FIELDS if not torch.equal(G.state[n], out[n])]
        bad += [f"grad {n}" for n in L.DIFF_INPUTS if not torch.equal(G.grads[n], grads[n])]
        assert not bad, f"replay {r}: differs from the eager step in {bad}"
        assert torch.equal(vamax, vamax_e)
        eager = out


def test_graph_keeps_its_workspace_after_larger_eager_step():
    """An eager step with a larger isplit between replays needs more scratch
    than the recorded step; the graph's kernels must still use memory the
    graph owns (the eager call may not replace it), so replays stay bitwise
    equal to eager steps."""
    print("[gpu] StepGraph workspace ownership, isplit 30 graph, isplit 60 eager", flush=True)
    import dataclasses
    dtype = torch.float64
    states = _members()
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    sc_big = dataclasses.replace(sc, isplit=2 * sc.isplit, dte=0.5 * sc.dte)
    r0 = torch.tensor([ramp_of(S, 1) for S in states], dtype=dtype, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(4)
    cot = {n: torch.randn(ens.numel(n), generator=gen, device="cuda", dtype=dtype) for n in ("u", "t", "el")}
    G = StepGraph(ens, state, geom, sc, r0, vg, True, cot)
    eager = {n: state[n].clone() for n in L.FIELDS}
    for r in range(2):
        # unrelated eager work with a larger workspace, then a fresh allocation
        # that would land on any memory the ensemble freed
        ens.step(eager, geom, sc_big, r0, vg, True)
        ens.step_adj(eager, geom, sc_big, r0, vg, True, dict(cot))
        junk = torch.full((1 << 22,), float("nan"), dtype=dtype, device="cuda")
        vamax = G.replay()
        out, vamax_e = ens.step(eager, geom, sc, r0, vg, True)
        grads = ens.step_adj(eager, geom, sc, r0, vg, True, dict(cot))
        torch.cuda.synchronize()
        del junk
        bad = [n for n in L.FIELDS if not torch.equal(G.state[n], out[n])]
        bad += [f"grad {n}" for n in L.DIFF_INPUTS if not torch.equal(G.grads[n], grads[n])]
        assert not bad, f"replay {r}: differs from the eager step in {bad}"
        assert torch.equal(vamax, vamax_e)
        eager = out


@pytest.mark.parametrize("precond", ["spectral", "jacobi"])
def test_implicit_graph_replays_match_eager_steps(precond):
    """The semi-implicit fwd+adjoint step recorded with capture-safe eta solves
    (a fixed launch count, no host polls) against eager steps whose solves
    poll the host and stop early: converged members do nothing in the extra
    iterations, so every replay is bitwise equal, iteration counts included."""
    import dataclasses
    from tests.test_gpu_implicit import _inputs, _state
    print(f"[gpu] implicit StepGraph vs eager, 2 members, {precond}", flush=True)
    S1, _ = _state("a")
    S2, _ = _state("a")
    gen = torch.Generator().manual_seed(11)
    S2.t = S2.t + 0.3 * torch.randn(S2.t.shape, generator=gen, dtype=S2.t.dtype)
    states = [S1, S2]
    ens, state, geom, vg, sc = _inputs(states, torch.float64)
    sc = dataclasses.replace(sc, eta_preconditioner=precond, eta_max_iter=120)
    ens.prepare(geom, sc)
    sc_cap = dataclasses.replace(sc, eta_capture=True)
    with pytest.raises(ValueError):
        StepGraph(ens, state, geom, sc, torch.ones(2, dtype=torch.float64, device="cuda"), vg, True)
    ramps = [torch.tensor([ramp_of(S, it) for S in states], dtype=torch.float64, device="cuda") for it in (1, 2, 3)]
    gcot = torch.Generator(device="cuda").manual_seed(5)
    cot = {n: torch.randn(ens.numel(n), generator=gcot, device="cuda", dtype=torch.float64)
           for n in ("u", "t", "el", "ua", "va")}
    G = StepGraph(ens, state, geom, sc_cap, ramps[0], vg, True, cot)
    eager = {n: state[n].clone() for n in L.FIELDS}
    for r in range(3):
        G.ramp.copy_(ramps[r])
        vamax = G.replay()
        out, vamax_e = ens.step(eager, geom, sc, ramps[r], vg, True)
        grads = ens.step_adj(eager, geom, sc, ramps[r], vg, True, dict(cot))
        torch.cuda.synchronize()
        bad = [n for n in L.FIELDS if not torch.equal(G.state[n], out[n])]
        bad += [f"grad {n}" for n in L.DIFF_INPUTS if not torch.equal(G.grads[n], grads[n])]
        assert not bad, f"replay {r}: differs from the eager step in {bad}"
        assert torch.equal(vamax, vamax_e)
        eager = out
