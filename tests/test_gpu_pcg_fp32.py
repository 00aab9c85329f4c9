"""The float32 spectral preconditioner (the DST-I products on the matrix cores,
k_spec_gemm_mfma, 32 x 32 output tiles) against the float64 one (vector ALUs,
k_spec_gemm) on the same eta solves: the 354 x 322 tidal channel (352 x 320
interior: whole 32 x 32 tiles, and a K extent of 352 that is not a multiple of
the 64-wide K stage), the 90 x 82 double-gyre basin (88 x 80: partial tiles
both ways), and a ragged batch of three members (12 x 10, 26 x 22, 90 x 82)
whose smaller members leave most tiles of the batch's grid empty; seeded
right-hand sides, relative tolerance 1e-5.

A preconditioner that is applied correctly in float32 takes the float64
iteration count to within one iteration and converges to the same solution;
a wrong operand or result layout would still converge, slower, as CG does
with any SPD preconditioner, so the iteration count is what pins it.
"""
import pytest
import torch

from neuralpom_amd import boundary as B
from neuralpom_amd import cases
from neuralpom_amd.implicit_core import POM2K_Core
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu


def _channel(dtype):
    st, g, vg, period, nodes = cases.tidal_channel(354, 322, dtype=dtype, device="cuda")
    core = POM2K_Core(354, 322, device="cuda", dtype=dtype)
    core.configure_runtime(dte=1.0, isplit=30)
    for n, v in {**st, **g}.items():
        setattr(core, n, v)
    core.z, core.zz, core.dz, core.dzz = vg
    core.period = period
    B.install(core, "implicit", nodes=nodes, speed=360.0 / 12.42, radiation=0.0)
    return core


def _basin(dtype):
    core = POM2K_Core(90, 82, device="cuda", dtype=dtype)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    return core


@pytest.mark.parametrize("case", ["channel", "basin"])
def test_fp32_spectral_preconditioner_matches_fp64(case):
    print(f"[gpu] eta solve, spectral preconditioner, fp32 (matrix cores) vs fp64: {case}", flush=True)
    make = _channel if case == "channel" else _basin
    res = {}
    for dtype in (torch.float64, torch.float32):
        core = make(dtype)
        core.eta_solver_rel_tol = 1e-5
        gen = torch.Generator().manual_seed(17)
        rhs = (1e-3 * torch.randn(core.im, core.jm, generator=gen, dtype=torch.float64)).to("cuda", dtype)
        rhs = rhs * core.fsm
        d = core.h + core.el
        its = []
        for k in range(3):   # three solves from different guesses: zero, the previous solution, its half
            x0 = torch.zeros_like(rhs) if k == 0 else (x if k == 1 else 0.5 * x)
            x = core.solve_implicit_eta_cg(rhs * (1.0 + 0.1 * k), d, max_iter=500, tol=0.0, initial_guess=x0)
            its.append(core._eta_solver_iteration_history[-1])
        torch.cuda.synchronize()
        res[dtype] = (x.double().cpu(), its)
    (x64, it64), (x32, it32) = res[torch.float64], res[torch.float32]
    err = rel_err(x32, x64)
    print(f"  iterations fp64 {it64} fp32 {it32}; solution rel err {err:.2e}", flush=True)
    assert all(abs(a - b) <= 1 for a, b in zip(it32, it64)), (it32, it64)
    assert max(it64) > 3
    assert err <= 1e-4


def test_fp32_spectral_batched_ragged_members_match_fp64():
    """One batched solve of three members of different sizes: each member's
    iteration count within one of its float64 count and its solution within
    1e-4, so the tiles past a small member's extent (skipped) and its partial
    tiles are both right."""
    import dataclasses
    from neuralpom_amd import _lib as L
    from neuralpom_amd.implicit import ImplicitEnsemble
    from tests.test_gpu_implicit import _scalars
    from tests.test_zz_no_alloc import _members
    shapes = ((12, 10), (26, 22), (90, 82))
    print(f"[gpu] eta solve, spectral, fp32 vs fp64, ragged batch {shapes}", flush=True)
    states = _members(shapes, 70)
    res = {}
    for dtype in (torch.float64, torch.float32):
        ens = ImplicitEnsemble(list(shapes), dtype=dtype, device="cuda")
        geom = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.GEOM}
        sc = dataclasses.replace(_scalars(states[0]), eta_preconditioner="spectral", eta_max_iter=500,
                                 eta_rel_tol=1e-5, eta_tol=0.0)
        ens.prepare(geom, sc)
        gen = torch.Generator().manual_seed(71)
        rhs = ens.pack([1e-3 * torch.randn((S.im, S.jm), generator=gen, dtype=torch.float64) * S.fsm
                        for S in states], "el")
        d = ens.pack([S.d for S in states], "d")
        its = torch.zeros(len(shapes), dtype=torch.int32, device="cuda")
        x = ens.eta_solve(geom, sc, rhs, d, None, its)
        torch.cuda.synchronize()
        res[dtype] = ([ens.view(x, m, "el").double().cpu() for m in range(len(shapes))], its.cpu().tolist())
    (x64, it64), (x32, it32) = res[torch.float64], res[torch.float32]
    print(f"  iterations fp64 {it64} fp32 {it32}", flush=True)
    assert all(abs(a - b) <= 1 for a, b in zip(it32, it64)), (it32, it64)
    assert min(it64) > 2
    for m in range(len(shapes)):
        assert rel_err(x32[m], x64[m]) <= 1e-4, (m, rel_err(x32[m], x64[m]))
