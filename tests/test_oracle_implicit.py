"""Pin the implicit-core oracle (oracle/pomref/implicit.py).

Against the reference: vectors written from the reference's semi-implicit core
in float64 (tests/golden/make_golden_implicit.py): barotropic stages of one
step, eta-solve solutions and iteration counts, 1- and 3-step trajectories,
the time0 = 0 skip, a tight-tolerance step and a seeded VJP through the
reference's custom eta-solve adjoint.

Against mathematics (no reference needed): the PCG solution equals a dense
solve of the assembled operator; the eta-solve adjoint equals a dense solve of
the transpose and its depth term equals autograd through the dense solve; the
DST-I spectral preconditioner is the exact inverse of the constant-coefficient
Dirichlet operator.
"""
import numpy as np
import pytest
import torch

from oracle.pomref import implicit as IM
from oracle.pomref.state import load_state
from tests._helpers import assert_rel, golden, rel_err

TAG = "12x10"


@pytest.fixture(scope="module")
def steps():
    return golden(f"implicit_steps_{TAG}.npz")


@pytest.fixture(scope="module")
def ops():
    return golden(f"implicit_ops_{TAG}.npz")


def _state(g, prefix="in"):
    S = load_state(g, prefix)
    IM.configure(S)
    return S


def T(g, key):
    return torch.from_numpy(np.array(g[key]))


DYNAMIC = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab", "va", "vab",
           "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam", "et", "etb", "etf", "rho",
           "w", "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "dt", "d", "l",
           "uaf", "vaf", "elf", "advua", "advva", "adx2d", "ady2d", "drx2d", "dry2d", "aam2d",
           "advx", "advy", "drhox", "drhoy", "fluxua", "fluxva", "wusurf", "wvsurf", "wtsurf", "wssurf")


@pytest.mark.parametrize("label", ["a", "b", "c", "t"])
def test_steps_match_reference(steps, label):
    S = _state(steps)
    if label == "t":
        S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    S.time0 = S.time = float(steps[f"{label}/time0"])
    S.iend = int(steps[f"{label}/nstep"])
    S.eta_iters = []
    IM.run_time_loop(S)
    assert S.eta_iters == list(steps[f"{label}/iters"])
    bad = []
    for n in DYNAMIC:
        e = rel_err(getattr(S, n), T(steps, f"{label}/{n}"))
        if not e <= (1e-10 if label == "b" else 1e-11):
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def _ops_state(steps, ops, stage):
    S = _state(steps)
    pre = f"{stage}/self/"
    for key in ops.files:
        if key.startswith(pre):
            setattr(S, key[len(pre):], T(ops, key))
    return S


def test_barotropic_stages_match_reference(steps, ops):
    S = _ops_state(steps, ops, "_barotropic_predictor_with_bc")
    us, vs, du, dv = IM.predictor(S)
    for got, i in ((us, 0), (vs, 1), (du, 2), (dv, 3)):
        assert_rel(got, T(ops, f"_barotropic_predictor_with_bc/out{i}"), 1e-13, f"predictor out{i}")
    S = _ops_state(steps, ops, "_barotropic_transport_rhs_full")
    fu, fv, rhs = IM.transport_rhs(S, T(ops, "_barotropic_transport_rhs_full/arg0"),
                                   T(ops, "_barotropic_transport_rhs_full/arg1"))
    for got, i in ((fu, 0), (fv, 1), (rhs, 2)):
        assert_rel(got, T(ops, f"_barotropic_transport_rhs_full/out{i}"), 1e-13, f"rhs out{i}")
    S = _ops_state(steps, ops, "solve_implicit_eta_cg")
    info = []
    x = IM.pcg(S, T(ops, "solve_implicit_eta_cg/arg0"), T(ops, "solve_implicit_eta_cg/arg1"),
               T(ops, "solve_implicit_eta_cg/kw/initial_guess"), **{**IM.solver_options(S), "info": info})
    assert info == [int(ops["solve_implicit_eta_cg/iters"])]
    assert_rel(x, T(ops, "solve_implicit_eta_cg/out0"), 1e-12, "eta solve")
    S = _ops_state(steps, ops, "_surface_bc_after_solver")
    assert_rel(IM.surface_bc_after_solver(S, T(ops, "_surface_bc_after_solver/arg0")),
               T(ops, "_surface_bc_after_solver/out0"), 1e-15, "post-solve bc")
    S = _ops_state(steps, ops, "_barotropic_corrector_with_bc")
    a = [T(ops, f"_barotropic_corrector_with_bc/arg{i}") for i in range(4)]
    uaf, vaf = IM.corrector(S, *a)
    assert_rel(uaf, T(ops, "_barotropic_corrector_with_bc/out0"), 1e-13, "uaf")
    assert_rel(vaf, T(ops, "_barotropic_corrector_with_bc/out1"), 1e-13, "vaf")
    S = _state(steps)
    assert_rel(IM.elliptic_apply(S, T(ops, "elliptic/eta"), T(ops, "elliptic/d")), T(ops, "elliptic/out"),
               1e-15, "elliptic")


VJP_IN = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
          "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
          "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")


def test_vjp_matches_reference(steps):
    v = golden(f"implicit_vjp_{TAG}.npz")
    S = _state(steps)
    S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    S.time0 = S.time = 0.1
    S.iend = 1
    leaves = {n: getattr(S, n).clone().requires_grad_(True) for n in VJP_IN}
    for n, x in leaves.items():
        setattr(S, n, x.clone())
    IM.run_time_loop(S)
    loss = sum((getattr(S, k[4:]) * T(v, k)).sum() for k in v.files if k.startswith("cot/"))
    grads = torch.autograd.grad(loss, [leaves[n] for n in VJP_IN], allow_unused=True)
    bad = []
    for n, g in zip(VJP_IN, grads):
        g = torch.zeros_like(leaves[n]) if g is None else g
        e = rel_err(g, T(v, f"grad/{n}"))
        if not e <= 1e-9:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


# ---------------------------------------------------------------- closed forms
def _dense(S, d):
    """Interior matrix of A(D) (Dirichlet-0 edges), assembled by unit vectors."""
    im, jm = S.im, S.jm
    n = (im - 2) * (jm - 2)
    cols = []
    for q in range(n):
        e = torch.zeros((im, jm), dtype=d.dtype)
        e[1 + q // (jm - 2), 1 + q % (jm - 2)] = 1.0
        cols.append(IM.elliptic_apply(S, e, d)[1:-1, 1:-1].reshape(-1))
    return torch.stack(cols, dim=1)


def test_pcg_equals_dense_solve(steps):
    S = _state(steps)
    d = S.h + S.el
    g = torch.Generator().manual_seed(11)
    rhs = torch.randn((S.im, S.jm), generator=g, dtype=torch.float64)
    for pre in ("spectral", "jacobi"):
        x = IM.pcg(S, rhs, d, torch.zeros_like(rhs), max_iter=400, tol=0.0, rel_tol=1e-14, precond=pre)
        ref = torch.linalg.solve(_dense(S, d), rhs[1:-1, 1:-1].reshape(-1)).view(S.im - 2, S.jm - 2)
        assert_rel(x[1:-1, 1:-1], ref, 1e-11, f"pcg/{pre}")
        assert float(x[0].abs().max()) == 0.0 and float(x[:, -1].abs().max()) == 0.0


def test_eta_adjoint_equals_transpose_solve(steps):
    S = _state(steps)
    S.eta_solver_rel_tol, S.eta_solver_tol = 1e-14, 0.0
    gen = torch.Generator().manual_seed(12)
    rhs = torch.randn((S.im, S.jm), generator=gen, dtype=torch.float64).requires_grad_(True)
    d = (S.h + S.el).clone().requires_grad_(True)
    w = torch.randn((S.im, S.jm), generator=gen, dtype=torch.float64)
    x = IM.eta_solve(S, rhs, d, torch.zeros_like(w))
    g_rhs, g_d = torch.autograd.grad((x * w).sum(), [rhs, d])
    # dense differentiable reference: x_int = A(d)^-1 rhs_int
    d2 = d.detach().clone().requires_grad_(True)
    rhs2 = rhs.detach().clone().requires_grad_(True)
    Ad = _dense_autograd(S, d2)
    xr = torch.linalg.solve(Ad, rhs2[1:-1, 1:-1].reshape(-1))
    r_rhs, r_d = torch.autograd.grad((xr * w[1:-1, 1:-1].reshape(-1)).sum(), [rhs2, d2])
    assert_rel(g_rhs, r_rhs, 1e-10, "adjoint rhs")
    assert_rel(g_d, r_d, 1e-9, "adjoint depth")


def _dense_autograd(S, d):
    im, jm = S.im, S.jm
    n = (im - 2) * (jm - 2)
    eye = torch.eye(n, dtype=d.dtype).view(n, im - 2, jm - 2)
    cols = []
    for q in range(n):
        e = torch.zeros((im, jm), dtype=d.dtype)
        e[1:-1, 1:-1] = eye[q]
        cols.append(IM.elliptic_apply(S, e, d)[1:-1, 1:-1].reshape(-1))
    return torch.stack(cols, dim=1)


def test_spectral_preconditioner_inverts_constant_operator(steps):
    S = _state(steps)
    S.dx = torch.full_like(S.dx, 7000.0)
    S.dy = torch.full_like(S.dy, 5000.0)
    S.h = torch.full_like(S.h, 800.0)
    S.art = S.dx * S.dy
    S.dum = torch.ones_like(S.dum)
    S.dvm = torch.ones_like(S.dvm)
    P = IM.spectral_cache(S)
    g = torch.Generator().manual_seed(13)
    x = torch.zeros((S.im, S.jm), dtype=torch.float64)
    x[1:-1, 1:-1] = torch.randn((S.im - 2, S.jm - 2), generator=g, dtype=torch.float64)
    y = IM.spectral_apply(S, IM.elliptic_apply(S, x, S.h), P)
    assert_rel(y, x, 1e-12, "M^-1 A x")


@pytest.mark.parametrize("label", ["a", "b"])
def test_shapiro_steps_match_reference(label):
    """The spatial Shapiro filters (barotropic, tracer and velocity weights
    0.05, numerics/double_gyre.py:78-184): 1 and 3 steps against the
    reference (tests/golden/make_golden_shapiro.py)."""
    g = golden(f"implicit_shapiro_{TAG}.npz")
    S = _state(g)
    gamma = float(g["scalar/shapiro_gamma"])
    S.barotropic_shapiro = S.tracer_shapiro = S.velocity_shapiro = gamma
    S.time0 = S.time = 0.1
    S.iend = 1 if label == "a" else 3
    S.eta_iters = []
    IM.run_time_loop(S)
    assert S.eta_iters == list(g[f"{label}/iters"])
    bad = []
    for n in DYNAMIC:
        e = rel_err(getattr(S, n), T(g, f"{label}/{n}"))
        if not e <= (1e-10 if label == "b" else 1e-11):
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad


def test_smagorinsky_floor_step_matches_reference():
    """aam_min = 800 m^2/s floors the Smagorinsky viscosity
    (implicit_p0_fixed/time_loop.py:186-188; the perturbed state has cells
    below it): one step against the reference with the Shapiro filters off."""
    g = golden(f"implicit_shapiro_{TAG}.npz")
    S = _state(g)
    S.aam_min = float(g["scalar/aam_min"])
    S.time0 = S.time = 0.1
    S.eta_solver_rel_tol, S.eta_solver_tol = 1e-12, 0.0
    S.iend = 1
    S.eta_iters = []
    IM.run_time_loop(S)
    assert S.eta_iters == list(g["m/iters"])
    bad = []
    for n in DYNAMIC:
        e = rel_err(getattr(S, n), T(g, f"m/{n}"))
        if not e <= 1e-11:
            bad.append(f"{n}: {e:.2e}")
    assert not bad, bad
