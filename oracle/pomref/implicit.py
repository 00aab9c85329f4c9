"""Oracle: the reference's semi-implicit (P0-fixed) POM step, restated in eager torch.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's baseline legs as the checker; never by the product path.

Restates the double-gyre (closed basin) configuration of
src/neuralpom/cores/implicit_p0_fixed/:
  time_loop.py:38-688           run_time_loop
  core.py:121-199               acceleration cache (face metric ratios, implicit coefficient)
  core.py:460-589               eager PCG (eta solve), stopping rule and update order
  core.py:591-675               elliptic operator A(D) and its depth adjoint
  core.py:677-773               Jacobi and spectral (DST-I) preconditioners
  core.py:24-59, 205-252        custom autograd of the eta solve (adjoint PCG on the
                                area-weighted transpose, plus the depth term)
  core.py:795-943               barotropic predictor, transport/RHS, post-solve BC, corrector
  boundary/closed_basin.py:27-101  closed-basin bcond 1-6 and eta boundary
  forcing/surface.py:20-84      double-gyre wind and surface heat/salt fluxes
  numerics/double_gyre.py:78-191  tracer / velocity filters (Shapiro gamma 0 in production)
  autograd.py:6-29              strain magnitude with zero subgradient at zero strain
and the open-boundary tide policy of the case adapter
src/neuralpom/cores/case_adapters/coastal_boundary.py:9-67 (implicit family),
installed with ``install_tide``.
The 3-D operators (advct, baropg, advave, vertvl, advq, profq, advt1, proft,
dens, advu/advv, profu/profv) are those of the explicit oracle; the implicit
copies differ only in the profq wall guard (IM/tracers.py:360, 1e-8 instead of
1e-20), selected here by ``S.wall_eps``.

The restatement is pinned against vectors written from the reference itself
(tests/golden/make_golden_implicit.py, tests/test_oracle_implicit.py; the tide
policy by tests/golden/make_golden_tide.py, tests/test_oracle_tide.py).
"""
from __future__ import annotations

import math

import torch

from . import explicit as X
from .explicit import EPS_DIV, _b, vsum

WALL_EPS = 1e-8           # IM/tracers.py:360
ENDPOINT_COEF = 0.25      # IM/time_loop.py:21


# --------------------------------------------------------------------------
# closed-basin lateral boundary policy: IM/boundary/closed_basin.py:27-101
# --------------------------------------------------------------------------
def _zero_edges(x):
    x[0] = 0.0
    x[-1] = 0.0
    x[:, 0] = 0.0
    x[:, -1] = 0.0
    return x


def bc1(S, elf):
    if getattr(S, "tide", None) is not None:
        return tide_bc1(S, elf)
    return _zero_edges(elf * S.fsm)


def bc2(S, uaf, vaf):
    if getattr(S, "tide", None) is not None:
        return tide_bc2(S, uaf, vaf)
    return _zero_edges(uaf * S.dum), _zero_edges(vaf * S.dvm)


# --------------------------------------------------------------------------
# open boundaries with a prescribed tidal elevation at fixed nodes: the
# implicit family of src/neuralpom/cores/case_adapters/coastal_boundary.py:9-64
# (bcond 1 and 2 :11-30, the lifted area-weighted CG solve :41-58, node and
# fixed-set construction :60-67).  bcond 3-6 are the explicit core's open
# forms (case_adapters/kernels/coastal_explicit_baroclinic.py:496-928, the
# edge weight of bcond 3 from d instead of h).
# --------------------------------------------------------------------------
TIDE_SPEED = 29.03225806451612   # degrees per hour (coastal_boundary.py:9)
TIDE_PI = 3.1416                 # the adapter's pi (coastal_boundary.py:9)


class Tide:
    """Tide nodes (0-based i, j), amplitudes, phases (degrees) and the fixed
    set of the eta solve: the outer ring, dry cells and the nodes
    (coastal_boundary.py:60-67)."""

    def __init__(self, S, i, j, amp, phase, speed=TIDE_SPEED, pi=TIDE_PI):
        self.i = torch.as_tensor(i, dtype=torch.long)
        self.j = torch.as_tensor(j, dtype=torch.long)
        self.amp = torch.as_tensor(amp, dtype=S.el.dtype)
        self.phase = torch.as_tensor(phase, dtype=S.el.dtype)
        self.speed, self.pi = float(speed), float(pi)
        fixed = torch.ones(S.fsm.shape, dtype=torch.bool)
        fixed[1:-1, 1:-1] = False
        fixed |= S.fsm == 0
        fixed[self.i, self.j] = True
        self.fixed = fixed


def install_tide(S, i, j, amp, phase, family="implicit", **kw):
    """coastal_boundary.py:55-62: nodes, rfe = rfw = rfn = rfs = 1, and the
    explicit family's substep counter, set to 0."""
    if family not in ("implicit", "explicit"):
        raise ValueError(family)
    S.tide = Tide(S, i, j, amp, phase, **kw)
    S.tide_family = family
    S.tide_substep = 0
    S.rfe = S.rfw = S.rfn = S.rfs = 1.0
    return S


def tide_value(S, seconds):
    """coastal_boundary.py:9-10"""
    T = S.tide
    return T.amp * torch.cos((T.speed * T.pi / 180) * seconds / 3600 - T.phase * T.pi / 180)


def tide_seconds(S):
    """the implicit family's evaluation time (coastal_boundary.py:15, 36, 45)"""
    return S.time * 86400 - S.dte


def tide_bc1(S, elf):
    """coastal_boundary.py:12-23 (implicit family)"""
    elf = elf.clone()
    elf[S.tide.i, S.tide.j] = tide_value(S, tide_seconds(S))
    return elf * S.fsm


def tide_bc2(S, uaf, vaf, el=None, d=None):
    """radiation conditions on c.el and c.d (coastal_boundary.py:24-30); the
    explicit family's substeps pass their current el and d"""
    uaf = uaf.clone()
    vaf = vaf.clone()
    J = I = slice(1, -1)
    g = S.grav
    d = S.d if d is None else d
    el = S.el if el is None else el
    uaf[-1, J] = uaf[-2, J] + S.rfe * torch.sqrt(g / d[-2, J]) * (el[-2, J] - el[-1, J])
    vaf[-1, J] = 0
    uaf[1, J] = uaf[1, J] - S.rfw * torch.sqrt(g / d[1, J]) * (el[1, J] - el[0, J])
    uaf[0, J] = uaf[1, J]
    vaf[0, J] = 0
    vaf[I, -1] = vaf[I, -2] + S.rfn * torch.sqrt(g / d[I, -2]) * (el[I, -2] - el[I, -1])
    uaf[I, -1] = 0
    vaf[I, 1] = vaf[I, 1] - S.rfs * torch.sqrt(g / d[I, 1]) * (el[I, 1] - el[I, 0])
    vaf[I, 0] = vaf[I, 1]
    uaf[I, 0] = 0
    return uaf * S.dum, vaf * S.dvm


def tide_eta_solve(S, rhs, d, max_iter=200, cap=500):
    """coastal_boundary.py:41-58: subtract the tidal lift, solve the homogeneous
    problem on the free nodes by CG in the area-weighted inner product from a
    zero start (stop when r.r <= max(1e-30, 1e-24 r0.r0), at most
    min(max_iter, 500) iterations), add the lift back.  Differentiated by
    autograd through the iterations, as the reference is.  ``cap`` replaces the
    reference's 500 so that a test can compare converged solutions on grids
    where 500 plain CG iterations do not reach the tolerance."""
    T = S.tide
    lift = torch.zeros_like(rhs)
    lift[T.i, T.j] = tide_value(S, tide_seconds(S))
    fixed = T.fixed
    weight = S.art * (~fixed)

    def op(v):
        return elliptic_apply(S, v.masked_fill(fixed, 0), d).masked_fill(fixed, 0)

    b = (rhs - elliptic_apply(S, lift, d)).masked_fill(fixed, 0)
    x = torch.zeros_like(rhs)
    r = b.clone()
    p = r.clone()
    rr = (weight * r * r).sum()
    rr0 = rr
    it = 0
    for it in range(min(max_iter, cap)):
        if float(rr.detach()) <= max(1e-30, float(rr0.detach()) * 1e-24):
            break
        ap = op(p)
        den = (weight * p * ap).sum()
        a = rr / (den + 1e-40)
        x = x + a * p
        r = r - a * ap
        new = (weight * r * r).sum()
        p = r + (new / (rr + 1e-40)) * p
        rr = new
    info = getattr(S, "eta_iters", None)
    if info is not None:
        info.append(it + 1)
    return x + lift


def bc3(S, uf, vf):
    return _zero_edges(uf * _b(S.dum)), _zero_edges(vf * _b(S.dvm))


def bc4(S, uf, vf):
    uf = uf.clone()
    vf = vf.clone()
    for f in (uf, vf):
        f[0] = f[1]
        f[-1] = f[-2]
        f[:, 0] = f[:, 1]
        f[:, -1] = f[:, -2]
    return uf, vf


def bc5(S, w):
    return w * _b(S.fsm)


def bc6(S, uf, vf):
    return uf * _b(S.fsm), vf * _b(S.fsm)


# --------------------------------------------------------------------------
# Smagorinsky with the zero-subgradient strain norm: IM/time_loop.py:152-201,
# IM/autograd.py:6-29
# --------------------------------------------------------------------------
class _Strain(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inv):
        m = torch.sqrt(inv)
        ctx.save_for_backward(m)
        return m

    @staticmethod
    def backward(ctx, g):
        m, = ctx.saved_tensors
        pos = m > 0
        return g * torch.where(pos, 0.5 / torch.where(pos, m, torch.ones_like(m)), torch.zeros_like(m))


def smagorinsky(S, u, v, aam):
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    K = slice(0, kbm1)
    dx = _b(S.dx[I, J])
    dy = _b(S.dy[I, J])
    t1 = ((u[Ip, J, K] - u[I, J, K]) / dx) ** 2
    t2 = ((v[I, Jp, K] - v[I, J, K]) / dy) ** 2
    ut = 0.25 * (u[I, Jp, K] + u[Ip, Jp, K] - u[I, Jm, K] - u[Ip, Jm, K]) / dy
    vt = 0.25 * (v[Ip, J, K] + v[Ip, Jp, K] - v[Im, J, K] - v[Im, Jp, K]) / dx
    t3 = 0.5 * (ut + vt) ** 2
    visc = S.horcon * dx * dy * _Strain.apply(t1 + t2 + t3)
    aam_min = float(getattr(S, "aam_min", 0.0) or 0.0)
    if aam_min > 0.0:
        visc = torch.clamp(visc, min=aam_min)
    out = aam.clone()
    out[I, J, K] = visc
    return out


# --------------------------------------------------------------------------
# surface forcing: IM/forcing/surface.py:73-84 (wind) + :20-70 (fluxes)
# --------------------------------------------------------------------------
def forcing(S):
    tau0 = float(getattr(S, "double_gyre_wind_stress_pa", 0.1))
    arg = 2.0 * S.pi * (S.north_e - 30.0) / 20.0
    S.wusurf = -(tau0 / S.rhoref) * (1.0 - torch.cos(arg)) * S.dum
    S.wvsurf = torch.zeros_like(S.wvsurf)
    im, jm = S.im, S.jm
    I, J = slice(1, im - 1), slice(1, jm - 1)
    inner = torch.zeros_like(S.e_atmos, dtype=torch.bool)
    inner[I, J] = True
    S.e_atmos = S.e_atmos.masked_fill(inner, 0.0)
    S.vfluxf = S.vfluxf.masked_fill(inner, 0.0)
    S.swrad = S.swrad.masked_fill(inner, 0.0)
    w = S.w.clone()
    w[I, J, 0] = S.vfluxf[I, J]
    S.w = w
    ts = S.t[I, J, 0]
    wt = S.wtsurf.clone()
    wt[I, J] = wt[I, J] + S.vfluxf[I, J] * ((ts + S.tbias) - ts - S.tbias)
    S.wtsurf = wt
    ws = S.wssurf.clone()
    ws[I, J] = S.vfluxf[I, J] * (0.0 - S.s[I, J, 0] - S.sbias)
    S.wssurf = ws


# --------------------------------------------------------------------------
# elliptic operator and preconditioners: IM/core.py:121-199, 591-773
# --------------------------------------------------------------------------
def accel_cache(S):
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    dx, dy = S.dx, S.dy
    dx_e = 0.5 * (dx[Ip, J] + dx[I, J])
    dx_w = 0.5 * (dx[I, J] + dx[Im, J])
    dy_e = 0.5 * (dy[Ip, J] + dy[I, J])
    dy_w = 0.5 * (dy[I, J] + dy[Im, J])
    dy_n = 0.5 * (dy[I, Jp] + dy[I, J])
    dy_s = 0.5 * (dy[I, J] + dy[I, Jm])
    dx_n = 0.5 * (dx[I, Jp] + dx[I, J])
    dx_s = 0.5 * (dx[I, J] + dx[I, Jm])
    return {
        "ex": dy_e / (dx_e + 1e-20), "wx": dy_w / (dx_w + 1e-20),
        "ny": dx_n / (dy_n + 1e-20), "sy": dx_s / (dy_s + 1e-20),
        "me": S.dum[Ip, J], "mw": S.dum[I, J], "mn": S.dvm[I, Jp], "ms": S.dvm[I, J],
        "inv_art": 1.0 / (S.art[I, J] + 1e-20),
        "coeff": 4.0 * (S.dti ** 2) * S.grav * S.alpha,
    }


def elliptic_apply(S, eta, d, C=None):
    """A(D) eta = eta - 4 dti^2 g alpha div(D grad eta) / art on the interior (core.py:591-639)."""
    C = C or accel_cache(S)
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    he = 0.5 * (d[Ip, J] + d[I, J])
    hw = 0.5 * (d[I, J] + d[Im, J])
    hn = 0.5 * (d[I, Jp] + d[I, J])
    hs = 0.5 * (d[I, J] + d[I, Jm])
    fe = he * (eta[Ip, J] - eta[I, J]) * C["ex"] * C["me"]
    fw = hw * (eta[I, J] - eta[Im, J]) * C["wx"] * C["mw"]
    fn = hn * (eta[I, Jp] - eta[I, J]) * C["ny"] * C["mn"]
    fs = hs * (eta[I, J] - eta[I, Jm]) * C["sy"] * C["ms"]
    div = (fe - fw + fn - fs) * C["inv_art"]
    out = eta.clone()
    out[I, J] = eta[I, J] - C["coeff"] * div
    return out


def elliptic_depth_adjoint(S, eta, lam, d, C=None):
    """(d(A eta)/dD)^T lam (core.py:645-675)."""
    C = C or accel_cache(S)
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    g = torch.zeros_like(d)
    lm = lam[I, J] * C["coeff"] * C["inv_art"]
    e = 0.5 * lm * (eta[Ip, J] - eta[I, J]) * C["ex"] * C["me"]
    w = 0.5 * lm * (eta[I, J] - eta[Im, J]) * C["wx"] * C["mw"]
    n = 0.5 * lm * (eta[I, Jp] - eta[I, J]) * C["ny"] * C["mn"]
    s = 0.5 * lm * (eta[I, J] - eta[I, Jm]) * C["sy"] * C["ms"]
    g[Ip, J] += -e
    g[I, J] += -e
    g[I, J] += w
    g[Im, J] += w
    g[I, Jp] += -n
    g[I, J] += -n
    g[I, J] += s
    g[I, Jm] += s
    return g


def jacobi_diag(S, d, C=None):
    C = C or accel_cache(S)
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    he = 0.5 * (d[Ip, J] + d[I, J])
    hw = 0.5 * (d[I, J] + d[Im, J])
    hn = 0.5 * (d[I, Jp] + d[I, J])
    hs = 0.5 * (d[I, J] + d[I, Jm])
    diag = torch.ones_like(d)
    diag[I, J] = 1.0 + C["coeff"] * C["inv_art"] * (
        he * C["ex"] * C["me"] + hw * C["wx"] * C["mw"] + hn * C["ny"] * C["mn"] + hs * C["sy"] * C["ms"])
    return diag


def dst1(x, dim):
    """2 x DST-I along ``dim`` through an odd extension and an FFT (core.py:710-728)."""
    n = x.shape[dim]
    shape = list(x.shape)
    shape[dim] = 2 * (n + 1)
    y = x.new_zeros(shape)
    sl = [slice(None)] * x.ndim
    sl[dim] = slice(1, n + 1)
    y[tuple(sl)] = x
    sl[dim] = slice(n + 2, 2 * (n + 1))
    y[tuple(sl)] = -torch.flip(x, dims=[dim])
    return -torch.fft.fft(y, dim=dim).imag.narrow(dim, 1, n)


def spectral_cache(S, C=None):
    """Exact inverse eigenvalues of the constant-coefficient operator at the median
    dx, dy, h of the interior (core.py:730-760)."""
    C = C or accel_cache(S)
    im, jm = S.im, S.jm
    I, J = slice(1, im - 1), slice(1, jm - 1)
    ni, nj = im - 2, jm - 2
    dt_, dev = S.el.dtype, S.el.device
    kx = torch.arange(1, ni + 1, device=dev, dtype=dt_).view(-1, 1)
    ky = torch.arange(1, nj + 1, device=dev, dtype=dt_).view(1, -1)
    dx0 = S.dx[I, J].median()
    dy0 = S.dy[I, J].median()
    h0 = S.h[I, J].median()
    lx = 4.0 * torch.sin(torch.pi * kx / (2.0 * (ni + 1))).square() / (dx0 * dx0 + 1e-30)
    ly = 4.0 * torch.sin(torch.pi * ky / (2.0 * (nj + 1))).square() / (dy0 * dy0 + 1e-30)
    return {"inv_eigs": 1.0 / (1.0 + C["coeff"] * h0 * (lx + ly)), "scale": 4.0 * (ni + 1) * (nj + 1)}


def spectral_apply(S, r, P):
    I, J = slice(1, S.im - 1), slice(1, S.jm - 1)
    c = dst1(dst1(r[I, J], 0), 1) * P["inv_eigs"]
    out = torch.zeros_like(r)
    out[I, J] = dst1(dst1(c, 0), 1) / P["scale"]
    return out


def pcg(S, rhs, d, x0, max_iter, tol, rel_tol, precond="spectral", info=None):
    """Preconditioned CG of the eager path (core.py:460-589, non-capture branch)."""
    C = accel_cache(S)
    I, J = slice(1, S.im - 1), slice(1, S.jm - 1)
    x = _zero_edges(x0.detach().clone())
    b = _zero_edges(rhs.clone())
    r = b - elliptic_apply(S, x, d, C)
    if precond == "spectral":
        P = spectral_cache(S, C)
        apply_m = lambda v: spectral_apply(S, v, P)
    else:
        diag = jacobi_diag(S, d, C)

        def apply_m(v):
            o = v.clone()
            o[I, J] = v[I, J] / (diag[I, J] + 1e-30)
            return o
    z = apply_m(r)
    p = z.clone()
    rho_old = (r[I, J] * z[I, J]).sum()
    rs0 = float((r[I, J] * r[I, J]).sum())
    tol_sq = tol * tol
    rel_sq = rel_tol ** 2 * max(rs0, 1e-30)
    it = 0
    for it in range(1, max_iter + 1):
        p = _zero_edges(p)
        ap = elliptic_apply(S, p, d, C)
        den = (p[I, J] * ap[I, J]).sum()
        alpha = rho_old / (den + 1e-30)
        x = _zero_edges(x + alpha * p)
        r = r - alpha * ap
        rs = float((r[I, J] * r[I, J]).sum())
        if rs <= tol_sq or rs <= rel_sq or rs <= 1e-30:
            break
        z = apply_m(r)
        rho_new = (r[I, J] * z[I, J]).sum()
        beta = rho_new / (rho_old + 1e-30)
        p = _zero_edges(p * beta + z)
        rho_old = rho_new
    if info is not None:
        info.append(it)
    return _zero_edges(x)


class _EtaSolve(torch.autograd.Function):
    """x = A(D)^-1 rhs with the reference's adjoint (core.py:24-59, 231-252)."""

    @staticmethod
    def forward(ctx, S, rhs, d, x0, opts):
        x = pcg(S, rhs.detach(), d.detach(), x0, **opts)
        ctx.S, ctx.opts = S, opts
        ctx.save_for_backward(x, d)
        return x

    @staticmethod
    def backward(ctx, g):
        x, d = ctx.saved_tensors
        S, opts = ctx.S, ctx.opts
        with torch.no_grad():
            area = torch.ones_like(g)
            area[1:-1, 1:-1] = S.art[1:-1, 1:-1] / S.art[1, 1]
            rhs = g / area
            scale = rhs.abs().amax().clamp_min(1.0)
            o = dict(opts)
            o["tol"] = float(opts["tol"]) / float(scale)
            o["info"] = None
            lam = pcg(S, rhs / scale, d, torch.zeros_like(g), **o) * area * scale
            gd = elliptic_depth_adjoint(S, x, lam, d)
        return None, lam, -gd, None, None


def solver_options(S):
    """IM/numerics/double_gyre.py:186-191 and the production settings of
    src/neuralpom/v1/core_spec.py:22-29."""
    return {"max_iter": int(getattr(S, "eta_solver_max_iter", 200)),
            "tol": float(getattr(S, "eta_solver_tol", 1.0e-12)),
            "rel_tol": float(getattr(S, "eta_solver_rel_tol", 1.0e-5)),
            "precond": str(getattr(S, "eta_solver_preconditioner", "spectral"))}


def eta_solve(S, rhs, d, x0):
    opts = solver_options(S)
    if getattr(S, "tide", None) is not None:
        return tide_eta_solve(S, rhs, d, opts["max_iter"], getattr(S, "tide_iteration_cap", 500))
    opts["info"] = getattr(S, "eta_iters", None)   # list collecting iteration counts
    if torch.is_grad_enabled() and (rhs.requires_grad or d.requires_grad):
        return _EtaSolve.apply(S, rhs, d, x0, opts)
    return pcg(S, rhs, d, x0, **opts)


# --------------------------------------------------------------------------
# barotropic predictor / transport / corrector: IM/core.py:795-943
# --------------------------------------------------------------------------
def predictor(S):
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    cor, d, ua, va, el, elb, h = S.cor, S.d, S.ua, S.va, S.el, S.elb, S.h
    grav, alpha, dti = S.grav, S.alpha, S.dti
    ctu = 0.25 * (cor[I, J] * d[I, J] * (va[I, Jp] + va[I, J]) +
                  cor[Im, J] * d[Im, J] * (va[Im, Jp] + va[Im, J]))
    ctv = 0.25 * (cor[I, J] * d[I, J] * (ua[Ip, J] + ua[I, J]) +
                  cor[I, Jm] * d[I, Jm] * (ua[Ip, Jm] + ua[I, Jm]))
    gu = 0.25 * grav * (S.dy[I, J] + S.dy[Im, J]) * (d[I, J] + d[Im, J]) * (
        (1.0 - 2.0 * alpha) * (el[I, J] - el[Im, J]) + alpha * (elb[I, J] - elb[Im, J]) +
        S.e_atmos[I, J] - S.e_atmos[Im, J])
    gv = 0.25 * grav * (S.dx[I, J] + S.dx[I, Jm]) * (d[I, J] + d[I, Jm]) * (
        (1.0 - 2.0 * alpha) * (el[I, J] - el[I, Jm]) + alpha * (elb[I, J] - elb[I, Jm]) +
        S.e_atmos[I, J] - S.e_atmos[I, Jm])
    pu = S.adx2d[I, J] + S.advua[I, J] - S.aru[I, J] * ctu + gu + S.drx2d[I, J] - \
        S.aru[I, J] * (-S.wusurf[I, J] + S.wubot[I, J])
    pv = S.ady2d[I, J] + S.advva[I, J] + S.arv[I, J] * ctv + gv + S.dry2d[I, J] - \
        S.arv[I, J] * (-S.wvsurf[I, J] + S.wvbot[I, J])
    us = ua.clone()
    vs = va.clone()
    du = (h[I, J] + el[I, J] + h[Im, J] + el[Im, J]) * S.aru[I, J]
    nu = (h[I, J] + elb[I, J] + h[Im, J] + elb[Im, J]) * S.aru[I, J] * S.uab[I, J] - 4.0 * dti * pu
    dv = (h[I, J] + el[I, J] + h[I, Jm] + el[I, Jm]) * S.arv[I, J]
    nv = (h[I, J] + elb[I, J] + h[I, Jm] + elb[I, Jm]) * S.arv[I, J] * S.vab[I, J] - 4.0 * dti * pv
    us[I, J] = (nu / (du + 1e-20)) * S.dum[I, J]
    vs[I, J] = (nv / (dv + 1e-20)) * S.dvm[I, J]
    us, vs = bc2(S, us, vs)
    return us, vs, du, dv


def transport_rhs(S, us, vs):
    im, jm = S.im, S.jm
    I, J = slice(1, im - 1), slice(1, jm - 1)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    fu = 0.25 * (S.d[A, B] + S.d[Am, B]) * (S.dy[A, B] + S.dy[Am, B]) * us[A, B]
    fv = 0.25 * (S.d[A, B] + S.d[A, Bm]) * (S.dx[A, B] + S.dx[A, Bm]) * vs[A, B]
    div = fu[1:, :-1] - fu[:-1, :-1] + fv[:-1, 1:] - fv[:-1, :-1]
    rhs = S.elb.clone()
    rhs[I, J] = S.elb[I, J] - 2.0 * S.dti * (div / (S.art[I, J] + 1e-20) + S.vfluxf[I, J])
    fluxua = torch.zeros_like(S.fluxua)
    fluxva = torch.zeros_like(S.fluxva)
    fluxua[A, B] = fu
    fluxva[A, B] = fv
    return fluxua, fluxva, rhs


def surface_bc_after_solver(S, elf):
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    gamma = float(getattr(S, "barotropic_shapiro", 0.0))
    sm = elf.clone()
    sm[I, J] = elf[I, J] + (gamma / 4.0) * (elf[Ip, J] + elf[Im, J] + elf[I, Jp] + elf[I, Jm] - 4.0 * elf[I, J])
    out = elf.clone()
    out[I, J] = sm[I, J] * S.fsm[I, J]
    return bc1(S, out)


def corrector(S, us, vs, du, dv):
    im, jm = S.im, S.jm
    I, Im = slice(1, im - 1), slice(0, im - 2)
    J, Jm = slice(1, jm - 1), slice(0, jm - 2)
    d, elf, h, grav, alpha, dti = S.d, S.elf, S.h, S.grav, S.alpha, S.dti
    gu = 0.25 * grav * (S.dy[I, J] + S.dy[Im, J]) * (d[I, J] + d[Im, J]) * alpha * (elf[I, J] - elf[Im, J])
    gv = 0.25 * grav * (S.dx[I, J] + S.dx[I, Jm]) * (d[I, J] + d[I, Jm]) * alpha * (elf[I, J] - elf[I, Jm])
    duf = (h[I, J] + elf[I, J] + h[Im, J] + elf[Im, J]) * S.aru[I, J]
    dvf = (h[I, J] + elf[I, J] + h[I, Jm] + elf[I, Jm]) * S.arv[I, J]
    uaf = S.uaf.clone()
    vaf = S.vaf.clone()
    uaf[I, J] = (us[I, J] * du - 4.0 * dti * gu) / (duf + 1e-20) * S.dum[I, J]
    vaf[I, J] = (vs[I, J] * dv - 4.0 * dti * gv) / (dvf + 1e-20) * S.dvm[I, J]
    return bc2(S, uaf, vaf)


def external_mode(S):
    """Semi-implicit barotropic step (IM/time_loop.py:247-398).  Returns vamax."""
    im, jm = S.im, S.jm
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    S.egf = S.el * S.ispi
    d_baro = S.h + S.elb
    utf = S.utf.clone()
    vtf = S.vtf.clone()
    utf[A, :] = S.ua[A, :] * (S.d[A, :] + S.d[Am, :]) * S.isp2i
    vtf[:, B] = S.va[:, B] * (S.d[:, B] + S.d[:, Bm]) * S.isp2i
    S.utf, S.vtf = utf, vtf
    us, vs, du, dv = predictor(S)
    S.fluxua, S.fluxva, rhs = transport_rhs(S, us, vs)
    x0 = S.elf
    S.elf = eta_solve(S, rhs, S.d, x0)
    S.elf = surface_bc_after_solver(S, S.elf)
    S.d = S.h + S.elf
    S.uaf, S.vaf = corrector(S, us, vs, du, dv)
    S.etf = S.elf * S.fsm
    vamax = torch.max(torch.abs(S.vaf))
    if float(vamax) <= S.vmaxl:
        sm = S.smoth
        ua = S.ua + 0.5 * sm * (S.uab - 2.0 * S.ua + S.uaf)
        va = S.va + 0.5 * sm * (S.vab - 2.0 * S.va + S.vaf)
        el = S.el + 0.5 * sm * (S.elb - 2.0 * S.el + S.elf)
        S.elb, S.el = el, S.elf
        S.d = S.h + S.el
        S.uab, S.ua, S.vab, S.va = ua, S.uaf, va, S.vaf
        S.egf = (S.elb + S.el) * 0.5
        utf = S.utf.clone()
        vtf = S.vtf.clone()
        su_new = S.d[A, :] + S.d[Am, :]
        su_old = d_baro[A, :] + d_baro[Am, :]
        utf[A, :] = ENDPOINT_COEF * (S.uab[A, :] * su_old + S.ua[A, :] * su_new)
        sv_new = S.d[:, B] + S.d[:, Bm]
        sv_old = d_baro[:, B] + d_baro[:, Bm]
        vtf[:, B] = ENDPOINT_COEF * (S.vab[:, B] * sv_old + S.va[:, B] * sv_new)
        S.utf, S.vtf = utf, vtf
    return vamax


def internal_mode(S):
    """IM/time_loop.py:400-650 with the closed-basin boundary policy, or the
    open forms of bcond 3-6 when a tide policy is installed."""
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    K = slice(0, kbm1)
    tps_u = vsum(S, S.u)
    corr = (S.utb[A, :] + S.utf[A, :]) / (S.dt[A, :] + S.dt[Am, :] + EPS_DIV)
    u = S.u.clone()
    u[A, :, K] = (S.u[A, :, K] - _b(tps_u[A, :])) + _b(corr)
    S.u = u
    tps_v = vsum(S, S.v)
    corr = (S.vtb[:, B] + S.vtf[:, B]) / (S.dt[:, B] + S.dt[:, Bm] + EPS_DIV)
    v = S.v.clone()
    v[:, B, K] = (S.v[:, B, K] - _b(tps_v[:, B])) + _b(corr)
    S.v = v

    tide = getattr(S, "tide", None) is not None
    w = X.vertvl(S, S.u, S.v, S.w, S.dt, S.etf, S.etb, S.vfluxb, S.vfluxf)
    S.w = X.bcond5(S, w) if tide else bc5(S, w)
    q2f = X.advq(S, S.q2b, S.q2, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    q2lf = X.advq(S, S.q2lb, S.q2l, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    q2f, q2lf, S.l, S.kq, S.km, S.kh = X.profq(
        S, q2f, q2lf, S.q2, S.q2b, S.q2lb, S.km, S.kh, S.kq, S.u, S.v, S.t, S.s, S.rho,
        S.wusurf, S.wvsurf, S.wubot, S.wvbot, S.etf, S.dt)
    q2f, q2lf = X.bcond6(S, q2f, q2lf, S.q2, S.q2l, S.u, S.v) if tide else bc6(S, q2f, q2lf)
    q2 = S.q2 + 0.5 * S.smoth * (q2f + S.q2b - 2.0 * S.q2)
    q2l = S.q2l + 0.5 * S.smoth * (q2lf + S.q2lb - 2.0 * S.q2l)
    S.q2b, S.q2, S.q2lb, S.q2l = q2, q2f, q2l, q2lf
    vmix = getattr(S, "vertical_mixing_max", None)
    if vmix is not None:
        S.kh = torch.clamp(S.kh, min=0.0, max=float(vmix))
        S.km = torch.clamp(S.km, min=0.0, max=float(vmix))
        S.kq = torch.clamp(S.kq, min=0.0, max=float(vmix))

    tb, t, tf = X.advt1(S, S.tb, S.t, S.tclim, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    sb, s, sf = X.advt1(S, S.sb, S.s, S.sclim, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    S.tb, S.t, S.sb, S.s = tb, t, sb, s
    tf = X.proft(S, tf, S.wtsurf, S.tsurf, S.nbct, S.kh, S.etf, S.swrad)
    sf = X.proft(S, sf, S.wssurf, S.ssurf, S.nbcs, S.kh, S.etf, S.swrad)
    tf, sf = X.bcond4(S, tf, sf, S.t, S.s, S.u, S.v, S.w, S.dt) if tide else bc4(S, tf, sf)
    t = S.t + 0.5 * S.smoth * (tf + S.tb - 2.0 * S.t)
    s = S.s + 0.5 * S.smoth * (sf + S.sb - 2.0 * S.s)
    tf, sf = _shapiro3(S, tf, "tracer_shapiro", S.fsm), _shapiro3(S, sf, "tracer_shapiro", S.fsm)
    S.tb, S.t, S.sb, S.s = t, tf, s, sf
    S.rho = X.dens(S, S.t, S.s, S.rho)

    uf = X.advu(S, S.advx, S.drhox, S.u, S.v, S.w, S.ub, S.dt, S.egf, S.egb, S.etf, S.etb, S.e_atmos)
    vf = X.advv(S, S.advy, S.drhoy, S.u, S.v, S.w, S.vb, S.dt, S.egf, S.egb, S.etf, S.etb, S.e_atmos)
    uf, S.wubot, _ = X.profu(S, uf, S.wusurf, S.wubot, S.km, S.etf, S.ub, S.vb)
    vf, S.wvbot, S.tps = X.profv(S, vf, S.wvsurf, S.wvbot, S.km, S.etf, S.ub, S.vb)
    uf, vf = X.bcond3(S, uf, vf, S.u, S.v, depth=S.d) if tide else bc3(S, uf, vf)
    dzw = S.dz[:kbm1].view(1, 1, -1)
    term = uf[:, :, K] + S.ub[:, :, K] - 2.0 * S.u[:, :, K]
    tps = torch.sum(term * dzw, dim=2)
    ubn = S.u.clone()
    ubn[:, :, K] = S.u[:, :, K] + 0.5 * S.smoth * (term - _b(tps))
    term = vf[:, :, K] + S.vb[:, :, K] - 2.0 * S.v[:, :, K]
    tps = torch.sum(term * dzw, dim=2)
    vbn = S.v.clone()
    vbn[:, :, K] = S.v[:, :, K] + 0.5 * S.smoth * (term - _b(tps))
    ubn = _shapiro3(S, ubn, "velocity_shapiro", S.dum)
    vbn = _shapiro3(S, vbn, "velocity_shapiro", S.dvm)
    S.ub, S.u, S.vb, S.v = ubn, uf, vbn, vf


def _shapiro3(S, f, attr, mask):
    """numerics/double_gyre.py:78-184: Shapiro filter (gamma) on the interior, then the mask."""
    gamma = float(getattr(S, attr, 0.0))
    out = f.clone()
    if gamma > 0.0:
        im, jm, K = S.im, S.jm, slice(0, S.kbm1)
        I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
        J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
        out[I, J, K] = f[I, J, K] + (gamma / 4.0) * (f[Ip, J, K] + f[Im, J, K] + f[I, Jp, K] + f[I, Jm, K]
                                                     - 4.0 * f[I, J, K])
    return out * _b(mask)


def step(S, iint):
    """One internal step iint (IM/time_loop.py:95-688).  Raises FloatingPointError
    when vamax > vmaxl (time_loop.py:685-688)."""
    S.iint = iint
    S.time = S.dti * iint * 1.0 / 86400.0 + S.time0
    S.ramp = min(S.time / float(S.period), 1.0) if S.lramp else 1.0
    S.wall_eps = WALL_EPS
    forcing(S)
    S.advx, S.advy, S.curv = X.advct(S, S.u, S.v, S.ub, S.vb, S.aam, S.dt)
    S.rho, S.drhox, S.drhoy = X.baropg(S, S.rho, S.dt, S.ramp)
    S.aam = smagorinsky(S, S.u, S.v, S.aam)
    S.adx2d = vsum(S, S.advx)
    S.ady2d = vsum(S, S.advy)
    S.drx2d = vsum(S, S.drhox)
    S.dry2d = vsum(S, S.drhoy)
    S.aam2d = vsum(S, S.aam)
    S.advua, S.advva, S.fluxua, S.fluxva, S.tps = X.advave(
        S, S.ua, S.va, S.uab, S.vab, S.d, S.aam2d, S.fluxua, S.fluxva, S.tps)
    S.adx2d = S.adx2d - S.advua
    S.ady2d = S.ady2d - S.advva
    vamax = external_mode(S)
    ok = float(vamax) <= S.vmaxl
    if ok and (iint != 1 or S.time0 != 0.0):
        internal_mode(S)
    S.egb = S.egf
    S.etb = S.et
    S.et = S.etf
    S.dt = S.h + S.et
    S.utb = S.utf
    S.vtb = S.vtf
    S.vfluxb = S.vfluxf
    if not ok:
        raise FloatingPointError(f"Barotropic velocity {float(vamax)} exceeds limit {S.vmaxl}")
    return True


def run_time_loop(S):
    for iint in range(1, int(S.iend) + 1):
        step(S, iint)
    return True


def configure(S):
    """Production settings of the implicit core (src/neuralpom/v1/core_spec.py:22-33,
    src/neuralpom/hybrid/core_pool.py:19-26): alpha .333, smoth .15, horcon .05,
    Shapiro gamma 0, eta relative tolerance 1e-5, spectral preconditioner."""
    S.alpha, S.smoth, S.horcon = 0.333, 0.15, 0.05
    S.barotropic_shapiro = S.tracer_shapiro = S.velocity_shapiro = 0.0
    S.eta_solver_rel_tol = 1e-5
    S.eta_solver_tol = 1e-12
    S.eta_solver_max_iter = 200
    S.eta_solver_preconditioner = "spectral"
    S.pi = math.atan(1.0) * 4.0
    return S
