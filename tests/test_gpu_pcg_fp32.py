"""The float32 spectral preconditioner (the DST-I products on the matrix cores,
k_spec_gemm_mfma) against the float64 one (vector ALUs, k_spec_gemm) on the
same eta solves: the 354 x 322 tidal channel (352 x 320 interior: a partial
64-row tile) and the 90 x 82 double-gyre basin (88 x 80: partial tiles both
ways), seeded right-hand sides, relative tolerance 1e-5.

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
