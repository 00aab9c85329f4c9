"""The semi-implicit drop-in core (neuralpom_amd.implicit_core.POM2K_Core) against the oracle.

Construction, class constants and the capture contract of the reference's
checkpointing (src/neuralpom/v1/checkpoint.py:51-72: every attribute of
vars(core) is a tensor, a plain scalar, a known policy object or scratch)
run on the CPU; the time loop, solve_implicit_eta_cg, apply_elliptic_operator
and their gradients run the HIP kernels (marked gpu).  The oracle is pinned
to the reference's float64 golden vectors and to dense solves
(tests/test_oracle_implicit.py).
"""
import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.implicit_core import POM2K_Core
from oracle.pomref import implicit as IM
from oracle.pomref.cases import double_gyre as oracle_dg
from tests._helpers import rel_err

IM_, JM_ = 12, 10

# the reference's scratch names and the objects it deep-copies (v1/checkpoint.py:17-18, 47-48, 64-66)
SCRATCH = {"device", "_accel_cache", "_cg_work", "_eta_solver_iteration_history",
           "_spectral_eta_preconditioner_cache"}
OBJECTS = {"_settings_overrides", "_implicit_forcing_provider", "_implicit_numerics_policy",
           "_implicit_boundary_policy", "_gradient_kernel_names"}


def _core(device, dtype=torch.float64):
    core = POM2K_Core(IM_, JM_, device=device, dtype=dtype)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    core.horcon, core.alpha, core.smoth = 0.05, 0.333, 0.15     # v1/core_spec.py:32-33
    return core


def _unclassified(core):
    bad = []
    for name, value in vars(core).items():
        if name in SCRATCH or name.startswith(("_compiled_", "_cuda_graph_")):
            continue
        if torch.is_tensor(value) or value is None or isinstance(value, (bool, int, float, str, np.generic)):
            continue
        if name in OBJECTS:
            continue
        bad.append(f"{name}: {type(value).__name__}")
    return bad


def test_class_contract():
    core = _core("cpu")
    assert core.NUMERICS_VERSION == "implicit-dg-v2-aligned-1e5-p0-barotropic-coupling-fixed"
    assert core.BAROTROPIC_ENDPOINT_TRANSPORT_COEFFICIENT == 0.25
    assert core.alpha == 0.333 and core.smoth == 0.15 and core.aam_min == 0.0
    assert core.eta_solver_rel_tol == 1e-5 and core.eta_solver_preconditioner == "spectral"
    assert core.barotropic_shapiro == 0.0 and core.tracer_shapiro == 0.0
    assert core.dti == 600.0 and core.kbm1 == 15
    assert not _unclassified(core), _unclassified(core)


def test_explicit_core_capture_contract():
    from neuralpom_amd.core import POM2K_Core as Explicit
    core = Explicit(IM_, JM_, device="cpu", dtype=torch.float64)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    assert not _unclassified(core), _unclassified(core)


def test_overrides_are_applied_after_core_defaults():
    core = POM2K_Core(IM_, JM_, device="cpu", settings_overrides={"alpha": 0.3, "smoth": 0.2})
    assert core.alpha == 0.3 and core.smoth == 0.2


def test_invalid_shapiro_flag(monkeypatch):
    monkeypatch.setenv("NEURALPOM_IMPLICIT_SPATIAL_SHAPIRO", "maybe")
    with pytest.raises(ValueError):
        POM2K_Core(IM_, JM_, device="cpu")


def test_replaced_eta_solver_is_refused():
    core = _core("cpu")
    core.solve_implicit_eta_cg = lambda *a, **k: None
    core.iend = 1
    with pytest.raises(NotImplementedError):
        core.run_time_loop()


# ---------------------------------------------------------------- GPU
def _perturbed(seed=7):
    S = oracle_dg(IM_, JM_)
    gen = torch.Generator().manual_seed(seed)
    for n, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02), ("ua", 0.01), ("va", 0.01)):
        x = getattr(S, n)
        setattr(S, n, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
    S.d = S.h + S.el
    return S


def _core_from(S, device="cuda"):
    core = _core(device)
    for n in ("t", "el", "u", "v", "ua", "va", "d"):
        setattr(core, n, getattr(S, n).to(device).clone())
    return core


@pytest.mark.gpu
@pytest.mark.parametrize("time0", [0.0, 0.25])
def test_time_loop_matches_oracle(time0):
    S = _perturbed()
    core = _core_from(S)
    IM.configure(S)
    S.time0 = S.time = core.time0 = time0
    S.iend = core.iend = 2
    S.eta_iters = []
    print(f"[gpu] implicit core run_time_loop 12x10 time0={time0}", flush=True)
    IM.run_time_loop(S)
    core.run_time_loop()
    torch.cuda.synchronize()
    assert core.iint == 2 and abs(core.time - (time0 + 2 * 600.0 / 86400.0)) < 1e-12
    assert core._eta_solver_iteration_history == S.eta_iters
    bad = [f"{n}: {rel_err(getattr(core, n).cpu(), getattr(S, n)):.2e}" for n in L.FIELDS
           if not rel_err(getattr(core, n).cpu(), getattr(S, n)) <= 1e-10]
    assert not bad, bad


@pytest.mark.gpu
def test_time_loop_gradient_matches_oracle():
    S = _perturbed(3)
    core = _core_from(S)
    IM.configure(S)
    S.time0 = core.time0 = 0.25
    S.iend = core.iend = 1
    print("[gpu] implicit core autograd 12x10", flush=True)
    t0 = S.t.clone().requires_grad_(True)
    S.t = t0
    IM.run_time_loop(S)
    gen = torch.Generator().manual_seed(11)
    w = {n: torch.randn(getattr(S, n).shape, generator=gen, dtype=torch.float64) for n in ("u", "t", "el")}
    (gt,) = torch.autograd.grad(sum((getattr(S, n) * w[n]).sum() for n in w), t0)
    t1 = core.t.clone().requires_grad_(True)
    core.t = t1
    core.run_time_loop()
    (ht,) = torch.autograd.grad(sum((getattr(core, n) * w[n].cuda()).sum() for n in w), t1)
    assert rel_err(ht.cpu(), gt) <= 1e-8


def _solve_inputs(seed=2):
    S = _perturbed(seed)
    IM.configure(S)
    gen = torch.Generator().manual_seed(seed + 100)
    rhs = 0.1 * torch.randn(S.el.shape, generator=gen, dtype=torch.float64)
    x0 = 0.05 * torch.randn(S.el.shape, generator=gen, dtype=torch.float64)
    return S, rhs, x0


@pytest.mark.gpu
@pytest.mark.parametrize("precond", ["spectral", "jacobi"])
def test_eta_solve_matches_oracle(precond):
    S, rhs, x0 = _solve_inputs()
    core = _core_from(S)
    core.eta_solver_preconditioner = S.eta_solver_preconditioner = precond
    print(f"[gpu] solve_implicit_eta_cg 12x10 {precond}", flush=True)
    info = []
    ref = IM.pcg(S, rhs, S.d, x0, 200, 1e-12, 1e-5, precond, info)
    got = core.solve_implicit_eta_cg(rhs.cuda(), core.d, max_iter=200, tol=1e-12, print_cg_residual=False,
                                     initial_guess=x0.cuda())
    assert core._eta_solver_iteration_history[-1] == info[-1]
    assert rel_err(got.cpu(), ref) <= 1e-10
    # tight tolerance: the dense solution of the interior system
    core.eta_solver_rel_tol = 1e-13
    tight = core.solve_implicit_eta_cg(rhs.cuda(), core.d, max_iter=500, tol=0.0, initial_guess=x0.cuda()).cpu()
    A = _dense(S)
    I = _interior(S)
    xd = torch.linalg.solve(A, rhs[1:-1, 1:-1].reshape(-1))
    assert rel_err(tight[1:-1, 1:-1].reshape(-1), xd) <= 1e-9
    assert float(tight[~I].abs().max()) == 0.0


def _interior(S):
    m = torch.zeros(S.el.shape, dtype=torch.bool)
    m[1:-1, 1:-1] = True
    return m


def _dense(S):
    """A(d) restricted to the interior unknowns (edges held at zero)."""
    n = (S.im - 2) * (S.jm - 2)
    A = torch.zeros(n, n, dtype=torch.float64)
    for c in range(n):
        e = torch.zeros(S.el.shape, dtype=torch.float64)
        e[1 + c // (S.jm - 2), 1 + c % (S.jm - 2)] = 1.0
        A[:, c] = IM.elliptic_apply(S, e, S.d)[1:-1, 1:-1].reshape(-1)
    return A


@pytest.mark.gpu
def test_eta_solve_gradient_matches_oracle():
    S, rhs, x0 = _solve_inputs(4)
    core = _core_from(S)
    S.eta_solver_tol, S.eta_solver_rel_tol = 1e-12, 1e-5
    print("[gpu] solve_implicit_eta_cg backward 12x10", flush=True)
    gen = torch.Generator().manual_seed(9)
    w = torch.randn(S.el.shape, generator=gen, dtype=torch.float64)
    r0, d0 = rhs.clone().requires_grad_(True), S.d.clone().requires_grad_(True)
    x = IM.eta_solve(S, r0, d0, x0)
    gr, gd = torch.autograd.grad((x * w).sum(), (r0, d0))
    r1, d1 = rhs.cuda().requires_grad_(True), core.d.clone().requires_grad_(True)
    y = core.solve_implicit_eta_cg(r1, d1, max_iter=200, tol=1e-12, initial_guess=x0.cuda())
    hr, hd = torch.autograd.grad((y * w.cuda()).sum(), (r1, d1))
    assert rel_err(hr.cpu(), gr) <= 1e-8
    assert rel_err(hd.cpu(), gd) <= 1e-8


@pytest.mark.gpu
def test_elliptic_operator_and_gradient():
    S, eta, _ = _solve_inputs(6)
    core = _core_from(S)
    print("[gpu] apply_elliptic_operator 12x10", flush=True)
    e0, d0 = eta.clone().requires_grad_(True), S.d.clone().requires_grad_(True)
    ref = IM.elliptic_apply(S, e0, d0)
    gen = torch.Generator().manual_seed(12)
    w = torch.randn(ref.shape, generator=gen, dtype=torch.float64)
    ge, gd = torch.autograd.grad((ref * w).sum(), (e0, d0))
    e1, d1 = eta.cuda().requires_grad_(True), core.d.clone().requires_grad_(True)
    out = core.apply_elliptic_operator(e1, d1)
    assert rel_err(out.detach().cpu(), ref.detach()) <= 1e-13
    he, hd = torch.autograd.grad((out * w.cuda()).sum(), (e1, d1))
    assert rel_err(he.cpu(), ge) <= 1e-12
    assert rel_err(hd.cpu(), gd) <= 1e-12
