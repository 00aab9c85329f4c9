"""CPU restatement of NeuralPOM's explicit mode-split POM step (the oracle).

TEST INFRASTRUCTURE ONLY.  This module is the parity checker for the HIP
kernels in ``neuralpom_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import it.  The product path never
calls it.

It restates, in dtype-generic eager PyTorch (run in float64 on the CPU for
parity), the explicit core of the reference:

* ``run_time_loop``        -- src/neuralpom/cores/explicit/time_loop.py:20-709
* ``advct``                -- cores/explicit/baroclinic.py:14-260
* ``advave`` (mode 3)      -- cores/explicit/baroclinic.py:262-432
* ``bcond`` 1..6           -- cores/explicit/baroclinic.py:496-928
* ``vertvl``               -- cores/explicit/baroclinic.py:930-1040
* ``advq``                 -- cores/explicit/tracers.py:14-156
* ``profq``                -- cores/explicit/tracers.py:158-431
* ``advt1``                -- cores/explicit/tracers.py:433-569
* ``proft``                -- cores/explicit/tracers.py:890-1010
* ``advu`` / ``advv``      -- cores/explicit/momentum.py:14-248
* ``profu`` / ``profv``    -- cores/explicit/momentum.py:250-465
* ``dens``                 -- cores/explicit/momentum.py:467-541
* ``baropg``               -- cores/explicit/momentum.py:543-700
* Smagorinsky ``aam``      -- cores/explicit/time_loop.py:10-16, 192-233
* double-gyre set-up       -- cores/explicit/cases.py:30-164,
                              cores/explicit/initialization.py:10-99,
                              cores/explicit/grid.py:113-166
* the tide adapter's explicit family -- case_adapters/coastal_boundary.py:11-30,
                              55-62 (installed with ``implicit.install_tide(...,
                              family="explicit")``): ``tide_bcond1`` at the
                              substep counter's time, the radiation bcond 2 of
                              ``implicit.tide_bc2`` on the substep's el and d,
                              and bcond 3 weighted by d
                              (kernels/coastal_explicit_baroclinic.py:586-641)

Every operator keeps the reference's association order of floating-point
operations so that the float64 results agree with the reference's own float64
trajectory to round-off.  Operators are written out-of-place (fresh tensors)
so the whole step is differentiable with ``torch.autograd``; this gives the
gradient oracle for the hand-written adjoint kernels.

Parity pins (``tests/test_oracle_golden.py``): every operator against the
golden vectors written from the reference by ``tests/golden/make_golden.py``,
the 1- and 3-step trajectories, the time0 = 0 internal-mode skip, and the
vector-Jacobian product of a step; the tide family by
``tests/test_oracle_explicit_tide.py`` against
``tests/golden/make_golden_explicit_tide.py``.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import torch

EPS_DIV = 1e-20      # src/neuralpom/utils/safety.py:6
EPS_SQRT = 1e-8      # src/neuralpom/utils/safety.py:7
EPS_LOG = 1e-12      # src/neuralpom/utils/safety.py:8


def _b(x):
    return x.unsqueeze(2)


def safe_sqrt(x):
    # utils/safety.py:25-26
    return torch.sqrt(torch.clamp(x, min=EPS_SQRT))


def safe_abs_pow(x, e):
    # utils/safety.py:29-30
    return torch.pow(torch.clamp(torch.abs(x), min=EPS_SQRT), e)


# --------------------------------------------------------------------------
# State container
# --------------------------------------------------------------------------
class State(SimpleNamespace):
    """Attribute bag with the reference's field names (cores/explicit/base.py:120-189)."""

    def tensors(self):
        return {k: v for k, v in vars(self).items() if torch.is_tensor(v)}

    def copy(self):
        out = State()
        for k, v in vars(self).items():
            setattr(out, k, v.clone() if torch.is_tensor(v) else v)
        return out


# --------------------------------------------------------------------------
# 3-D momentum advection (advct): baroclinic.py:14-260
# --------------------------------------------------------------------------
def advct(S, u, v, ub, vb, aam, dt):
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    dx, dy, aru, arv = S.dx, S.dy, S.aru, S.arv
    K = slice(0, kbm1)
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)

    curv = torch.zeros_like(u)
    t1 = (v[I, Jp, K] + v[I, J, K]) * _b(dy[Ip, J] - dy[Im, J])
    t2 = (u[Ip, J, K] + u[I, J, K]) * _b(dx[I, Jp] - dx[I, Jm])
    curv[I, J, K] = 0.25 * (t1 - t2) / (_b(dx[I, J] * dy[I, J]) + EPS_DIV)

    # ---- x-momentum fluxes (baroclinic.py:75-141)
    xflux = torch.zeros_like(u)
    yflux = torch.zeros_like(u)
    tu1 = (_b(dt[Ip, :]) + _b(dt[I, :])) * u[Ip, :, K]
    tu2 = (_b(dt[I, :]) + _b(dt[Im, :])) * u[I, :, K]
    xflux[I, :, K] = 0.125 * (tu1 + tu2) * (u[Ip, :, K] + u[I, :, K])

    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    tv1 = (_b(dt[A, B]) + _b(dt[A, Bm])) * v[A, B, K]
    tv2 = (_b(dt[Am, B]) + _b(dt[Am, Bm])) * v[Am, B, K]
    yflux[A, B, K] = 0.125 * (tv1 + tv2) * (u[A, B, K] + u[A, Bm, K])

    grad_u = (ub[Ip, B, K] - ub[I, B, K]) / (_b(dx[I, B]) + EPS_DIV)
    xflux[I, B, K] = xflux[I, B, K] - _b(dt[I, B]) * aam[I, B, K] * 2.0 * grad_u
    dt_avg = 0.25 * (dt[I, B] + dt[Im, B] + dt[I, Bm] + dt[Im, Bm])
    aam_sum = aam[I, B, K] + aam[Im, B, K] + aam[I, Bm, K] + aam[Im, Bm, K]
    dtaam = _b(dt_avg) * aam_sum
    den_dy = _b(dy[I, B] + dy[Im, B] + dy[I, Bm] + dy[Im, Bm])
    den_dx = _b(dx[I, B] + dx[Im, B] + dx[I, Bm] + dx[Im, Bm])
    du_dy = (ub[I, B, K] - ub[I, Bm, K]) / (den_dy + EPS_DIV)
    dv_dx = (vb[I, B, K] - vb[Im, B, K]) / (den_dx + EPS_DIV)
    yflux[I, B, K] = yflux[I, B, K] - dtaam * (du_dy + dv_dx)
    xflux[I, B, K] = xflux[I, B, K] * _b(dy[I, B])
    dx_sum = dx[I, B] + dx[Im, B] + dx[I, Bm] + dx[Im, Bm]
    yflux[I, B, K] = yflux[I, B, K] * _b(0.25 * dx_sum)

    advx = torch.zeros_like(u)
    advx[I, J, K] = xflux[I, J, K] - xflux[Im, J, K] + yflux[I, Jp, K] - yflux[I, J, K]
    Ic, Icm = slice(2, im - 1), slice(1, im - 2)
    c1 = curv[Ic, J, K] * _b(dt[Ic, J]) * (v[Ic, Jp, K] + v[Ic, J, K])
    c2 = curv[Icm, J, K] * _b(dt[Icm, J]) * (v[Icm, Jp, K] + v[Icm, J, K])
    advx[Ic, J, K] = advx[Ic, J, K] - aru[Ic, J].unsqueeze(2) * 0.25 * (c1 + c2)

    # ---- y-momentum fluxes (baroclinic.py:165-256)
    xflux = torch.zeros_like(u)
    yflux = torch.zeros_like(u)
    tu1 = (_b(dt[A, B]) + _b(dt[Am, B])) * u[A, B, K]
    tu2 = (_b(dt[A, Bm]) + _b(dt[Am, Bm])) * u[A, Bm, K]
    xflux[A, B, K] = 0.125 * (tu1 + tu2) * (v[A, B, K] + v[Am, B, K])
    tv1 = (_b(dt[:, Jp]) + _b(dt[:, J])) * v[:, Jp, K]
    tv2 = (_b(dt[:, J]) + _b(dt[:, Jm])) * v[:, J, K]
    yflux[:, J, K] = 0.125 * (tv1 + tv2) * (v[:, Jp, K] + v[:, J, K])

    dt_avg = 0.25 * (dt[A, J] + dt[Am, J] + dt[A, Jm] + dt[Am, Jm])
    aam_sum = aam[A, J, K] + aam[Am, J, K] + aam[A, Jm, K] + aam[Am, Jm, K]
    dtaam = _b(dt_avg) * aam_sum
    den_dy = _b(dy[A, J] + dy[Am, J] + dy[A, Jm] + dy[Am, Jm])
    den_dx = _b(dx[A, J] + dx[Am, J] + dx[A, Jm] + dx[Am, Jm])
    du_dy = (ub[A, J, K] - ub[A, Jm, K]) / (den_dy + EPS_DIV)
    dv_dx = (vb[A, J, K] - vb[Am, J, K]) / (den_dx + EPS_DIV)
    xflux[A, J, K] = xflux[A, J, K] - dtaam * (du_dy + dv_dx)
    grad_v = (vb[A, Jp, K] - vb[A, J, K]) / (_b(dy[A, J]) + EPS_DIV)
    yflux[A, J, K] = yflux[A, J, K] - _b(dt[A, J]) * aam[A, J, K] * 2.0 * grad_v
    dy_sum = dy[A, J] + dy[Am, J] + dy[A, Jm] + dy[Am, Jm]
    xflux[A, J, K] = xflux[A, J, K] * _b(0.25 * dy_sum)
    yflux[A, J, K] = yflux[A, J, K] * _b(dx[A, J])

    advy = torch.zeros_like(v)
    advy[I, J, K] = xflux[Ip, J, K] - xflux[I, J, K] + yflux[I, J, K] - yflux[I, Jm, K]
    Jc, Jcm = slice(2, jm - 1), slice(1, jm - 2)
    c1 = curv[I, Jc, K] * _b(dt[I, Jc]) * (u[Ip, Jc, K] + u[I, Jc, K])
    c2 = curv[I, Jcm, K] * _b(dt[I, Jcm]) * (u[Ip, Jcm, K] + u[I, Jcm, K])
    advy[I, Jc, K] = advy[I, Jc, K] + arv[I, Jc].unsqueeze(2) * 0.25 * (c1 + c2)
    return advx, advy, curv


# --------------------------------------------------------------------------
# Baroclinic pressure gradient: momentum.py:543-700
# --------------------------------------------------------------------------
def baropg(S, rho, dt, ramp):
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    zz, grav = S.zz, S.grav
    rho_a = rho - S.rmean
    I, Im = slice(1, im - 1), slice(0, im - 2)
    J, Jm = slice(1, jm - 1), slice(0, jm - 2)
    Kk, Km = slice(1, kbm1), slice(0, kbm1 - 1)
    zk = zz[Kk].view(1, 1, -1)
    zkm = zz[Km].view(1, 1, -1)

    def one(r0, r1, dt0, dt1, dd, mask):
        term_dt = dt0 + dt1
        surf = 0.5 * grav * (-zz[0]) * term_dt * (r0[..., 0] - r1[..., 0])
        diff = (r0[..., Kk] - r1[..., Kk]) + (r0[..., Km] - r1[..., Km])
        t1 = grav * 0.25 * (zkm - zk) * _b(term_dt) * diff
        sm = (r0[..., Kk] + r1[..., Kk]) - (r0[..., Km] + r1[..., Km])
        t2 = grav * 0.25 * (zkm + zk) * _b(dt0 - dt1) * sm
        integ = torch.cumsum(t1 + t2, dim=2)
        col = torch.cat([_b(surf), _b(surf) + integ], dim=2)
        scale = 0.25 * term_dt * mask * dd
        return col * _b(scale)

    drhox = torch.zeros((im, jm, kb), dtype=rho.dtype, device=rho.device)
    drhoy = torch.zeros((im, jm, kb), dtype=rho.dtype, device=rho.device)
    drhox[I, J, :kbm1] = one(rho_a[I, J], rho_a[Im, J], dt[I, J], dt[Im, J],
                             S.dy[I, J] + S.dy[Im, J], S.dum[I, J])
    drhoy[I, J, :kbm1] = one(rho_a[I, J], rho_a[I, Jm], dt[I, J], dt[I, Jm],
                             S.dx[I, J] + S.dx[I, Jm], S.dvm[I, J])
    drhox = drhox * ramp
    drhoy = drhoy * ramp
    # The reference subtracts and re-adds rmean in place (momentum.py:554, 698);
    # the round trip can change rho in its last bit, so it is kept.
    return rho_a + S.rmean, drhox, drhoy


# --------------------------------------------------------------------------
# Smagorinsky viscosity: time_loop.py:10-16, 192-233
# --------------------------------------------------------------------------
def smagorinsky(S, u, v, aam):
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    K = slice(0, kbm1)
    dx = _b(S.dx[I, J])
    dy = _b(S.dy[I, J])
    du_dx = (u[Ip, J, K] - u[I, J, K]) / dx
    dv_dy = (v[I, Jp, K] - v[I, J, K]) / dy
    ut = 0.25 * (u[I, Jp, K] + u[Ip, Jp, K] - u[I, Jm, K] - u[Ip, Jm, K]) / dy
    vt = 0.25 * (v[Ip, J, K] + v[Ip, Jp, K] - v[Im, J, K] - v[Im, Jp, K]) / dx
    shear = ut + vt
    comps = torch.stack((du_dx, dv_dy, shear * math.sqrt(0.5)), dim=0)
    strain = torch.linalg.vector_norm(comps, dim=0)
    out = aam.clone()
    out[I, J, K] = S.horcon * dx * dy * strain
    return out


def vsum(S, f):
    """Depth integral sum_k f*dz over k < kbm1 (time_loop.py:240-246)."""
    return torch.sum(f[:, :, :S.kbm1] * S.dz[:S.kbm1].view(1, 1, -1), dim=2)


# --------------------------------------------------------------------------
# 2-D advection/diffusion of the depth-mean flow: baroclinic.py:262-432 (mode 3)
# --------------------------------------------------------------------------
def advave(S, ua, va, uab, vab, d, aam2d, fluxua, fluxva, tps):
    """Returns (advua, advva, fluxua, fluxva, tps).  fluxua/fluxva/tps are the
    reference's in-place work arrays: entries the operator does not write keep
    their input values, and advua reads fluxua[0, :] (baroclinic.py:364)."""
    im, jm = S.im, S.jm
    dx, dy = S.dx, S.dy
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)

    fluxua = fluxua.clone()
    fluxva = fluxva.clone()
    tps = tps.clone()
    t1 = (d[Ip, B] + d[I, B]) * ua[Ip, B]
    t2 = (d[I, B] + d[Im, B]) * ua[I, B]
    fluxua[I, B] = 0.125 * (t1 + t2) * (ua[Ip, B] + ua[I, B])
    t1 = (d[A, B] + d[A, Bm]) * va[A, B]
    t2 = (d[Am, B] + d[Am, Bm]) * va[Am, B]
    fluxva[A, B] = 0.125 * (t1 + t2) * (ua[A, B] + ua[A, Bm])
    diff_u = (uab[Ip, B] - uab[I, B]) / dx[I, B]
    fluxua[I, B] = fluxua[I, B] - d[I, B] * 2.0 * aam2d[I, B] * diff_u

    d_sum = d[A, B] + d[Am, B] + d[A, Bm] + d[Am, Bm]
    aam_sum = aam2d[A, B] + aam2d[A, Bm] + aam2d[Am, B] + aam2d[Am, Bm]
    dy_sum = dy[A, B] + dy[Am, B] + dy[A, Bm] + dy[Am, Bm]
    dx_sum = dx[A, B] + dx[Am, B] + dx[A, Bm] + dx[Am, Bm]
    du_dy = (uab[A, B] - uab[A, Bm]) / dy_sum
    dv_dx = (vab[A, B] - vab[Am, B]) / dx_sum
    tps[A, B] = 0.25 * d_sum * aam_sum * (du_dy + dv_dx)
    fluxua[A, B] = fluxua[A, B] * dy[A, B]
    fluxva[A, B] = (fluxva[A, B] - tps[A, B]) * 0.25 * dx_sum

    advua = torch.zeros_like(d)
    advua[I, J] = fluxua[I, J] - fluxua[Im, J] + fluxva[I, Jp] - fluxva[I, J]

    t1 = (d[A, B] + d[Am, B]) * ua[A, B]
    t2 = (d[A, Bm] + d[Am, Bm]) * ua[A, Bm]
    fluxua[A, B] = 0.125 * (t1 + t2) * (va[Am, B] + va[A, B])
    t1 = (d[A, Jp] + d[A, J]) * va[A, Jp]
    t2 = (d[A, J] + d[A, Jm]) * va[A, J]
    fluxva[A, J] = 0.125 * (t1 + t2) * (va[A, Jp] + va[A, J])
    diff_v = (vab[A, Jp] - vab[A, J]) / dy[A, J]
    fluxva[A, J] = fluxva[A, J] - d[A, J] * 2.0 * aam2d[A, J] * diff_v
    fluxva[A, B] = fluxva[A, B] * dx[A, B]
    fluxua[A, B] = (fluxua[A, B] - tps[A, B]) * 0.25 * dy_sum

    advva = torch.zeros_like(d)
    advva[I, J] = fluxua[Ip, J] - fluxua[I, J] + fluxva[I, J] - fluxva[I, Jm]
    return advua, advva, fluxua, fluxva, tps


# --------------------------------------------------------------------------
# Lateral boundary conditions: baroclinic.py:496-928
# --------------------------------------------------------------------------
def bcond1(S, elf):
    elf = elf.clone()
    elf[-1, :] = elf[-2, :]
    elf[0, :] = elf[1, :]
    elf[:, -1] = elf[:, -2]
    elf[:, 0] = elf[:, 1]
    return elf * S.fsm


def tide_bcond1(S, elf):
    """bcond 1 of the tide adapter's explicit family (coastal_boundary.py:12-22):
    each call advances the substep counter and sets the nodes to the tide at
    time * 86400 - dti + dte (counter - 1); no edge copy."""
    from . import implicit as IM
    S.tide_substep += 1
    sec = S.time * 86400 - S.dti + S.dte * (S.tide_substep - 1)
    elf = elf.clone()
    elf[S.tide.i, S.tide.j] = IM.tide_value(S, sec)
    return elf * S.fsm


def bcond2(S, uaf, vaf, el, ramp):
    uaf = uaf.clone()
    vaf = vaf.clone()
    J = slice(1, S.jm - 1)
    I = slice(1, S.im - 1)
    g, h = S.grav, S.h

    def c(hh):
        return safe_sqrt(g / torch.clamp(hh, min=EPS_SQRT))

    uaf[-1, J] = (S.uabe[J] + S.rfe * c(h[-2, J]) * (el[-2, J] - S.ele[J])) * ramp
    vaf[-1, J] = 0.0
    uaf[1, J] = (S.uabw[J] - S.rfw * c(h[1, J]) * (el[1, J] - S.elw[J])) * ramp
    uaf[0, J] = uaf[1, J]
    vaf[0, J] = 0.0
    vaf[I, -1] = (S.vabn[I] + S.rfn * c(h[I, -2]) * (el[I, -2] - S.eln[I])) * ramp
    uaf[I, -1] = 0.0
    vaf[I, 1] = (S.vabs[I] - S.rfs * c(h[I, 1]) * (el[I, 1] - S.els[I])) * ramp
    vaf[I, 0] = vaf[I, 1]
    uaf[I, 0] = 0.0
    return uaf * S.dum, vaf * S.dvm


def bcond3(S, uf, vf, u, v, depth=None):
    """``depth`` replaces h in the edge weight ga = sqrt(depth / hmax): the
    case-adapter kernel copy uses the total depth d
    (case_adapters/kernels/coastal_explicit_baroclinic.py:586, 602, 627, 641)."""
    uf = uf.clone()
    vf = vf.clone()
    K = slice(0, S.kbm1)
    J, Jm, Jp = slice(1, S.jm - 1), slice(0, S.jm - 2), slice(2, S.jm)
    I, Im, Ip = slice(1, S.im - 1), slice(0, S.im - 2), slice(2, S.im)
    h, hmax = (S.h if depth is None else depth), S.hmax

    def sm_j(f, i):
        return 0.25 * f[i, Jm, K] + 0.50 * f[i, J, K] + 0.25 * f[i, Jp, K]

    def sm_i(f, j):
        return 0.25 * f[Im, j, K] + 0.50 * f[I, j, K] + 0.25 * f[Ip, j, K]

    ga = safe_sqrt(h[-1, J] / hmax).unsqueeze(1)
    uf[-1, J, K] = ga * sm_j(u, -2) + (1.0 - ga) * sm_j(u, -1)
    vf[-1, J, K] = 0.0
    ga = safe_sqrt(h[0, J] / hmax).unsqueeze(1)
    uf[1, J, K] = ga * sm_j(u, 2) + (1.0 - ga) * sm_j(u, 1)
    uf[0, J, K] = uf[1, J, K]
    vf[0, J, K] = 0.0
    ga = safe_sqrt(h[I, -1] / hmax).unsqueeze(1)
    vf[I, -1, K] = ga * sm_i(v, -2) + (1.0 - ga) * sm_i(v, -1)
    uf[I, -1, K] = 0.0
    ga = safe_sqrt(h[I, 0] / hmax).unsqueeze(1)
    vf[I, 1, K] = ga * sm_i(v, 2) + (1.0 - ga) * sm_i(v, 1)
    vf[I, 0, K] = vf[I, 1, K]
    uf[I, 0, K] = 0.0
    return uf * _b(S.dum), vf * _b(S.dvm)


def bcond4(S, uf, vf, t, s, u, v, w, dt):
    uf = uf.clone()
    vf = vf.clone()
    kbm1, dti, zz = S.kbm1, S.dti, S.zz
    K = slice(0, kbm1)
    Kmid, Kmi, Kpl = slice(1, kbm1 - 1), slice(0, kbm1 - 2), slice(2, kbm1)
    dzz2 = (zz[Kmi] - zz[Kpl]).view(1, -1)

    def side(cur, inner, ext_t, ext_s, vel, dsum, inflow_le, wi, dti_, dtv, src_sign):
        # cur/inner: (n, kb) slices of t and s; vel: (n, kb) velocity for u1
        u1 = 2.0 * vel[:, K] * dti_ / dsum.unsqueeze(1)
        out = []
        for f, ext, fin in ((cur[0], ext_t, inner[0]), (cur[1], ext_s, inner[1])):
            fc = f[:, K]
            if src_sign > 0:   # east / north form
                f_in = fc - u1 * (ext[:, K] - fc)
                f_out = fc - u1 * (fc - fin[:, K])
            else:              # west / south form
                f_in = fc - u1 * (fc - ext[:, K])
                f_out = fc - u1 * (fin[:, K] - fc)
            inflow = (u1 <= 0.0) if inflow_le else (u1 >= 0.0)
            res = torch.where(inflow, f_in, f_out)
            if kbm1 > 2:
                wm = 0.5 * (wi[:, Kmid] + wi[:, Kpl]) * dti_ / (dzz2 * dtv.unsqueeze(1))
                corr = wm * (fin[:, Kmi] - fin[:, Kpl])
                u1m = u1[:, Kmid]
                outflow = (u1m > 0.0) if inflow_le else (u1m < 0.0)
                mid = torch.where(outflow, res[:, Kmid] - corr, res[:, Kmid])
                res = torch.cat([res[:, :1], mid, res[:, kbm1 - 1:kbm1]], dim=1)
            out.append(res)
        return out

    e = side((t[-1], s[-1]), (t[-2], s[-2]), S.tbe, S.sbe, u[-1], S.dx[-1, :] + S.dx[-2, :],
             True, w[-2], dti, dt[-2, :], +1)
    uf[-1, :, K], vf[-1, :, K] = e
    wst = side((t[0], s[0]), (t[1], s[1]), S.tbw, S.sbw, u[1], S.dx[0, :] + S.dx[1, :],
               False, w[1], dti, dt[1, :], -1)
    uf[0, :, K], vf[0, :, K] = wst
    n = side((t[:, -1], s[:, -1]), (t[:, -2], s[:, -2]), S.tbn, S.sbn, v[:, -1],
             S.dy[:, -1] + S.dy[:, -2], True, w[:, -2], dti, dt[:, -2], +1)
    uf[:, -1, K], vf[:, -1, K] = n
    so = side((t[:, 0], s[:, 0]), (t[:, 1], s[:, 1]), S.tbs, S.sbs, v[:, 1],
              S.dy[:, 0] + S.dy[:, 1], False, w[:, 1], dti, dt[:, 1], -1)
    uf[:, 0, K], vf[:, 0, K] = so
    return uf * _b(S.fsm), vf * _b(S.fsm)


def bcond5(S, w):
    w = w.clone()
    w[:, :, :S.kbm1] = w[:, :, :S.kbm1] * _b(S.fsm)
    return w


def bcond6(S, uf, vf, q2, q2l, u, v):
    uf = uf.clone()
    vf = vf.clone()
    small, dti = S.small, S.dti

    def side(qc, qi, lc, li, vel, dsum, east_form):
        u1 = 2.0 * vel * dti / dsum.unsqueeze(1)
        if east_form:
            inflow = u1 <= 0.0
            a_in = qc - u1 * (small - qc)
            b_in = lc - u1 * (small - lc)
            a_out = qc - u1 * (qc - qi)
            b_out = lc - u1 * (lc - li)
        else:
            inflow = u1 >= 0.0
            a_in = qc - u1 * (qc - small)
            b_in = lc - u1 * (lc - small)
            a_out = qc - u1 * (qi - qc)
            b_out = lc - u1 * (li - lc)
        return torch.where(inflow, a_in, a_out), torch.where(inflow, b_in, b_out)

    uf[-1], vf[-1] = side(q2[-1], q2[-2], q2l[-1], q2l[-2], u[-1], S.dx[-1, :] + S.dx[-2, :], True)
    uf[0], vf[0] = side(q2[0], q2[1], q2l[0], q2l[1], u[1], S.dx[0, :] + S.dx[1, :], False)
    uf[:, -1], vf[:, -1] = side(q2[:, -1], q2[:, -2], q2l[:, -1], q2l[:, -2], v[:, -1],
                                S.dy[:, -1] + S.dy[:, -2], True)
    uf[:, 0], vf[:, 0] = side(q2[:, 0], q2[:, 1], q2l[:, 0], q2l[:, 1], v[:, 1],
                              S.dy[:, 0] + S.dy[:, 1], False)
    return uf * _b(S.fsm) + 1.e-10, vf * _b(S.fsm) + 1.e-10


# --------------------------------------------------------------------------
# Vertical velocity: baroclinic.py:930-1040
# --------------------------------------------------------------------------
def vertvl(S, u, v, w, dt, etf, etb, vfluxb, vfluxf):
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    dx, dy = S.dx, S.dy
    K = slice(0, kbm1)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    I, Ip = slice(1, im - 1), slice(2, im)
    J, Jp = slice(1, jm - 1), slice(2, jm)
    xflux = torch.zeros_like(u)
    yflux = torch.zeros_like(u)
    xflux[A, B, K] = 0.25 * _b(dy[A, B] + dy[Am, B]) * _b(dt[A, B] + dt[Am, B]) * u[A, B, K]
    yflux[A, B, K] = 0.25 * _b(dx[A, B] + dx[A, Bm]) * _b(dt[A, B] + dt[A, Bm]) * v[A, B, K]
    w = w.clone()
    w[I, J, 0] = 0.5 * (vfluxb[I, J] + vfluxf[I, J])
    div = (xflux[Ip, J, K] - xflux[I, J, K] + yflux[I, Jp, K] - yflux[I, J, K])
    inc = (div / _b(dx[I, J] * dy[I, J]) + _b((etf[I, J] - etb[I, J]) / S.dti2)) \
        * S.dz[K].view(1, 1, -1)
    w[I, J, 1:kbm1 + 1] = w[I, J, 0].unsqueeze(2) + torch.cumsum(inc, dim=2)
    return w


# --------------------------------------------------------------------------
# Turbulence advection: tracers.py:14-156
# --------------------------------------------------------------------------
def advq(S, qb, q, u, v, w, aam, dt, etb, etf):
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    dx, dy, h, art = S.dx, S.dy, S.h, S.art
    Kk, Km, Kp = slice(1, kbm1), slice(0, kbm1 - 1), slice(2, kbm1 + 1)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    I, Ip = slice(1, im - 1), slice(2, im)
    J, Jp = slice(1, jm - 1), slice(2, jm)
    xflux = torch.zeros_like(q)
    yflux = torch.zeros_like(q)
    xflux[A, B, Kk] = 0.125 * (q[A, B, Kk] + q[Am, B, Kk]) * _b(dt[A, B] + dt[Am, B]) \
        * (u[A, B, Kk] + u[A, B, Km])
    yflux[A, B, Kk] = 0.125 * (q[A, B, Kk] + q[A, Bm, Kk]) * _b(dt[A, B] + dt[A, Bm]) \
        * (v[A, B, Kk] + v[A, B, Km])
    aam_x = aam[A, B, Kk] + aam[Am, B, Kk] + aam[A, B, Km] + aam[Am, B, Km]
    xflux[A, B, Kk] = xflux[A, B, Kk] - 0.25 * aam_x * _b(h[A, B] + h[Am, B]) \
        * (qb[A, B, Kk] - qb[Am, B, Kk]) * _b(S.dum[A, B]) / _b(dx[A, B] + dx[Am, B])
    aam_y = aam[A, B, Kk] + aam[A, Bm, Kk] + aam[A, B, Km] + aam[A, Bm, Km]
    yflux[A, B, Kk] = yflux[A, B, Kk] - 0.25 * aam_y * _b(h[A, B] + h[A, Bm]) \
        * (qb[A, B, Kk] - qb[A, Bm, Kk]) * _b(S.dvm[A, B]) / _b(dy[A, B] + dy[A, Bm])
    xflux[A, B, Kk] = xflux[A, B, Kk] * (0.5 * _b(dy[A, B] + dy[Am, B]))
    yflux[A, B, Kk] = yflux[A, B, Kk] * (0.5 * _b(dx[A, B] + dx[A, Bm]))

    dzs = (S.dz[Kk] + S.dz[Km]).view(1, 1, -1)
    vert = (w[I, J, Km] * q[I, J, Km] - w[I, J, Kp] * q[I, J, Kp]) * _b(art[I, J]) / dzs
    hdiv = xflux[Ip, J, Kk] - xflux[I, J, Kk] + yflux[I, Jp, Kk] - yflux[I, J, Kk]
    qf = torch.zeros_like(q)
    tb = _b((h[I, J] + etb[I, J]) * art[I, J])
    tf = _b((h[I, J] + etf[I, J]) * art[I, J])
    qf[I, J, Kk] = (tb * qb[I, J, Kk] - S.dti2 * (vert + hdiv)) / (tf + EPS_DIV)
    return qf


# --------------------------------------------------------------------------
# Mellor-Yamada 2.5 closure: tracers.py:158-431
# --------------------------------------------------------------------------
A1, B1, A2, B2, C1 = 0.92, 16.6, 0.74, 10.1, 0.08
E1, E2, SEF = 1.8, 1.33, 1.0
CBCNST, SURFL, SHIW = 100.0, 2.0e5, 0.0


def profq(S, q2f, q2lf, q2, q2b, q2lb, km, kh, kq, u, v, t, s, rho,
          wusurf, wvsurf, wubot, wvbot, etf, dt):
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    dti2, umol, grav, kappa, small = S.dti2, S.umol, S.grav, S.kappa, S.small
    dz, dzz, z, zz = S.dz, S.dzz, S.z, S.zz
    dh = S.h + etf
    dt_ = q2.dtype
    dev_ = q2.device

    a = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    c = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    for k in range(1, kbm1):
        a[:, :, k] = -dti2 * (kq[:, :, k + 1] + kq[:, :, k] + 2.0 * umol) * 0.5 \
            / (dzz[k - 1] * dz[k] * dh * dh)
        c[:, :, k] = -dti2 * (kq[:, :, k - 1] + kq[:, :, k] + 2.0 * umol) * 0.5 \
            / (dzz[k - 1] * dz[k - 1] * dh * dh)

    const1 = (16.6 ** (2.0 / 3.0)) * SEF
    I0, I1 = slice(0, im - 1), slice(1, im)
    J0, J1 = slice(0, jm - 1), slice(1, jm)
    ust = 0.5 * (wusurf[I0, J0] + wusurf[I1, J0])
    vst = 0.5 * (wvsurf[I0, J0] + wvsurf[I0, J1])
    utau2 = torch.sqrt(ust ** 2 + vst ** 2 + 1e-8)
    l0 = torch.zeros((im, jm), dtype=dt_, device=dev_)
    l0[I0, J0] = SURFL * utau2 / grav
    ubt = 0.5 * (wubot[I0, J0] + wubot[I1, J0])
    vbt = 0.5 * (wvbot[I0, J0] + wvbot[I0, J1])
    q2bot = torch.zeros((im, jm), dtype=dt_, device=dev_)
    q2bot[I0, J0] = torch.sqrt(ubt ** 2 + vbt ** 2 + 1e-8) * const1

    # buoyancy with sound-speed compressibility (tracers.py:245-261)
    tp = t[:, :, :kbm1] + S.tbias
    sp = s[:, :, :kbm1] + S.sbias
    p = grav * S.rhoref * (-zz[:kbm1].view(1, 1, -1) * _b(dh)) * 1.e-4
    cc = 1449.1 + 0.00821 * p + 4.55 * tp - 0.045 * (tp ** 2) + 1.34 * (sp - 35.0)
    cc = torch.where(torch.abs(cc) < 1e-6, torch.full_like(cc, 1e-6), cc)
    rad = (1.0 - 0.01642 * p / cc) * (1.0 - 0.40 * p / (cc ** 2))
    cc = cc / torch.sqrt(torch.clamp(rad, min=1e-8))
    boygr = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    for k in range(1, kbm1):
        boygr[:, :, k] = grav * (rho[:, :, k - 1] - rho[:, :, k]) / (dzz[k - 1] * dh) \
            + (grav ** 2) * 2.0 / (cc[:, :, k - 1] ** 2 + cc[:, :, k] ** 2)

    # length scale (tracers.py:263-285)
    q2b_abs = torch.abs(q2b)
    l = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    l[:, :, 1:kbm1] = torch.abs(q2lb[:, :, 1:kbm1]) / (q2b_abs[:, :, 1:kbm1] + 1e-20)
    surf = (z[:kbm1] > -0.5).view(1, 1, -1)
    lk = l[:, :, :kbm1].clone()
    l[:, :, :kbm1] = torch.where(surf, torch.max(lk, kappa * _b(l0)), lk)
    gh = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    gh[:, :, 1:kbm1] = (l[:, :, 1:kbm1].clone() ** 2) * boygr[:, :, 1:kbm1] \
        / (q2b_abs[:, :, 1:kbm1] + 1e-20)
    gh = torch.clamp(gh, max=0.028)
    l[:, :, 0] = kappa * l0
    l[:, :, kb - 1] = 0.0
    gh[:, :, 0] = 0.0
    gh[:, :, kb - 1] = 0.0

    # shear production (tracers.py:287-300)
    kn = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    I, Ip = slice(1, im - 1), slice(2, im)
    J, Jp = slice(1, jm - 1), slice(2, jm)
    for k in range(1, kbm1):
        du = (u[I, J, k] - u[I, J, k - 1] + u[Ip, J, k] - u[Ip, J, k - 1]) ** 2
        dv = (v[I, J, k] - v[I, J, k - 1] + v[I, Jp, k] - v[I, Jp, k - 1]) ** 2
        val = km[I, J, k] * 0.25 * SEF * (du + dv) / ((dzz[k - 1] * dh[I, J]) ** 2)
        kn[I, J, k] = val - SHIW * km[I, J, k] * boygr[I, J, k] + kh[I, J, k] * boygr[I, J, k]

    # q2 Thomas solve (tracers.py:302-346)
    dtef = torch.sqrt(q2b_abs + 1e-8) * 1.0 / (B1 * l + small)
    ee = [torch.zeros((im, jm), dtype=dt_, device=dev_)]
    g0 = torch.zeros((im, jm), dtype=dt_, device=dev_)
    g0[I0, J0] = (15.8 * CBCNST) ** (2.0 / 3.0) * utau2
    gg = [g0]
    for k in range(1, kbm1):
        den = a[:, :, k] + c[:, :, k] * (1.0 - ee[k - 1]) - (2.0 * dti2 * dtef[:, :, k] + 1.0)
        ee.append(a[:, :, k] / (den + 1e-20))
        gg.append((-2.0 * dti2 * kn[:, :, k] + c[:, :, k] * gg[k - 1] - q2f[:, :, k]) / (den + 1e-20))
    col = [None] * kb
    col[kb - 1] = q2bot
    for k in range(kbm1 - 1, -1, -1):
        col[k] = ee[k] * col[k + 1] + gg[k]
    q2n = torch.stack(col, dim=2)

    # q2l Thomas solve with wall proximity (tracers.py:348-384)
    ee = [torch.zeros((im, jm), dtype=dt_, device=dev_)]
    gg = [torch.zeros((im, jm), dtype=dt_, device=dev_)]
    ones = torch.ones_like(dh)
    # wall-proximity guard: 1e-20 (explicit, tracers.py:362), 1e-8 (implicit core, IM/tracers.py:360)
    weps = getattr(S, "wall_eps", 1e-20)
    for k in range(1, kbm1):
        wall = (ones / (torch.abs(z[k] - z[0]) + weps) + ones / (torch.abs(z[k] - z[kb - 1]) + weps)) \
            * l[:, :, k] / (dh * kappa)
        dtk = dtef[:, :, k] * (1.0 + E2 * (wall ** 2))
        den = a[:, :, k] + c[:, :, k] * (1.0 - ee[k - 1]) - (dti2 * dtk + 1.0)
        ee.append(a[:, :, k] / (den + 1e-20))
        gg.append((dti2 * (-kn[:, :, k] * l[:, :, k] * E1) + c[:, :, k] * gg[k - 1] - q2lf[:, :, k])
                  / (den + 1e-20))
    col = [None] * kb
    col[kb - 1] = torch.zeros((im, jm), dtype=dt_, device=dev_)
    for k in range(kbm1 - 1, -1, -1):
        col[k] = ee[k] * col[k + 1] + gg[k]
    q2ln = torch.stack(col, dim=2)

    # clipping (tracers.py:386-403)
    mask = (q2n[:, :, 1:kbm1] <= small) | (q2ln[:, :, 1:kbm1] <= small)
    q2c = q2n.clone()
    q2lc = q2ln.clone()
    q2c[:, :, 1:kbm1] = q2n[:, :, 1:kbm1].masked_fill(mask, small)
    q2lc[:, :, 1:kbm1] = torch.where(mask, 0.1 * _b(dt) * small, q2ln[:, :, 1:kbm1])

    # stability functions and mixing coefficients (tracers.py:405-429)
    coef4 = 18.0 * A1 * A1 + 9.0 * A1 * A2
    coef5 = 9.0 * A1 * A2
    coef1 = A2 * (1.0 - 6.0 * A1 / B1)
    coef2 = 3.0 * A2 * B2 + 18.0 * A1 * A2
    coef3 = A1 * (1.0 - 3.0 * C1 - 6.0 * A1 / B1)
    sh = coef1 / (1.0 - coef2 * gh + 1e-20)
    sm = (coef3 + sh * coef4 * gh) / (1.0 - coef5 * gh + 1e-20)
    knf = l * torch.sqrt(torch.abs(q2) + 1e-8)
    kq_o = (knf * 0.41 * sh + kq) * 0.5
    km_o = (knf * sm + km) * 0.5
    kh_o = (knf * sh + kh) * 0.5

    def edges(arr):
        arr = arr.clone()
        fsm = S.fsm
        arr[im - 1, :, :] = arr[im - 2, :, :] * fsm[im - 1, :].unsqueeze(1)
        arr[0, :, :] = arr[1, :, :] * fsm[0, :].unsqueeze(1)
        arr[:, jm - 1, :] = arr[:, jm - 2, :] * fsm[:, jm - 1].unsqueeze(1)
        arr[:, 0, :] = arr[:, 1, :] * fsm[:, 0].unsqueeze(1)
        return arr

    return q2c, q2lc, l, kq_o, edges(km_o), edges(kh_o)


# --------------------------------------------------------------------------
# Tracer advection (centred): tracers.py:433-569
# --------------------------------------------------------------------------
def advt1(S, fb, f, fclim, u, v, w, aam, dt, etb, etf):
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    dx, dy, h, art = S.dx, S.dy, S.h, S.art
    f = f.clone()
    fb = fb.clone()
    f[:, :, kb - 1] = f[:, :, kbm1 - 1]
    fb[:, :, kb - 1] = fb[:, :, kbm1 - 1]
    K = slice(0, kbm1)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    I, Ip = slice(1, im - 1), slice(2, im)
    J, Jp = slice(1, jm - 1), slice(2, jm)
    xflux = torch.zeros_like(f)
    yflux = torch.zeros_like(f)
    xflux[A, B, K] = 0.25 * _b(dt[A, B] + dt[Am, B]) * (f[A, B, K] + f[Am, B, K]) * u[A, B, K]
    yflux[A, B, K] = 0.25 * _b(dt[A, B] + dt[A, Bm]) * (f[A, B, K] + f[A, Bm, K]) * v[A, B, K]
    fa = fb - fclim
    tp = S.tprni
    xflux[A, B, K] = xflux[A, B, K] - 0.5 * (aam[A, B, K] + aam[Am, B, K]) * _b(h[A, B] + h[Am, B]) \
        * tp * (fa[A, B, K] - fa[Am, B, K]) * _b(S.dum[A, B]) / _b(dx[A, B] + dx[Am, B])
    xflux[A, B, K] = xflux[A, B, K] * _b(0.5 * (dy[A, B] + dy[Am, B]))
    yflux[A, B, K] = yflux[A, B, K] - 0.5 * (aam[A, B, K] + aam[A, Bm, K]) * _b(h[A, B] + h[A, Bm]) \
        * tp * (fa[A, B, K] - fa[A, Bm, K]) * _b(S.dvm[A, B]) / _b(dy[A, B] + dy[A, Bm])
    yflux[A, B, K] = yflux[A, B, K] * _b(0.5 * (dx[A, B] + dx[A, Bm]))

    zflux = torch.zeros_like(f)
    zflux[I, J, 0] = f[I, J, 0] * w[I, J, 0] * art[I, J]
    zflux[I, J, 1:kbm1] = 0.5 * (f[I, J, 0:kbm1 - 1] + f[I, J, 1:kbm1]) * w[I, J, 1:kbm1] * _b(art[I, J])
    hdiv = xflux[Ip, J, K] - xflux[I, J, K] + yflux[I, Jp, K] - yflux[I, J, K]
    vdiv = (zflux[I, J, K] - zflux[I, J, 1:kbm1 + 1]) / S.dz[K].view(1, 1, -1)
    ff = torch.zeros_like(f)
    tb_ = _b((h[I, J] + etb[I, J]) * art[I, J])
    tf_ = _b((h[I, J] + etf[I, J]) * art[I, J])
    ff[I, J, K] = (tb_ * fb[I, J, K] - S.dti2 * (hdiv + vdiv)) / (tf_ + EPS_DIV)
    return fb, f, ff


# --------------------------------------------------------------------------
# Vertical diffusion of a tracer: tracers.py:890-1010
# --------------------------------------------------------------------------
R_ = [0.58, 0.62, 0.67, 0.77, 0.78]
AD1_ = [0.35, 0.60, 1.00, 1.50, 1.40]
AD2_ = [0.23, 20.0, 17.0, 14.0, 7.90]


def proft(S, f_in, wfsurf, fsurf, nbc, kh, etf, swrad):
    im, jm, kb, kbm1, kbm2 = S.im, S.jm, S.kb, S.kbm1, S.kbm2
    dti2, dz, dzz, umol = S.dti2, S.dz, S.dzz, S.umol
    dh = S.h + etf
    dt_ = f_in.dtype
    dev_ = f_in.device
    a = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    c = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    for k in range(1, kbm1):
        dh2 = dh ** 2
        num = -dti2 * (kh[:, :, k] + umol)
        a[:, :, k - 1] = num / (dz[k - 1] * dzz[k - 1] * dh2)
        c[:, :, k] = num / (dz[k] * dzz[k - 1] * dh2)
    rad = torch.zeros((im, jm, kb), dtype=dt_, device=dev_)
    i = int(S.ntp) - 1
    if nbc in (2, 4):
        for k in range(kbm1):
            zt = S.z[k] * dh
            rad[:, :, k] = swrad * (R_[i] * torch.exp(zt / AD1_[i]) + (1.0 - R_[i]) * torch.exp(zt / AD2_[i]))
    ee = [None] * kb
    gg = [None] * kb
    if nbc == 1:
        den = a[:, :, 0] - 1.0
        ee[0] = a[:, :, 0] / den
        gg[0] = (-dti2 * wfsurf / (-dz[0] * dh) - f_in[:, :, 0]) / den
    elif nbc == 2:
        den = a[:, :, 0] - 1.0
        ee[0] = a[:, :, 0] / den
        gg[0] = (dti2 * (wfsurf + rad[:, :, 0] - rad[:, :, 1]) / (dz[0] * dh) - f_in[:, :, 0]) / den
    else:
        ee[0] = torch.zeros((im, jm), dtype=dt_, device=dev_)
        gg[0] = fsurf
    for k in range(1, kbm2):
        den = a[:, :, k] + c[:, :, k] * (1.0 - ee[k - 1]) - 1.0
        inv = 1.0 / (den + 1e-20)
        ee[k] = a[:, :, k] * inv
        rt = dti2 * (rad[:, :, k] - rad[:, :, k + 1]) / (dh * dz[k])
        gg[k] = (c[:, :, k] * gg[k - 1] - f_in[:, :, k] + rt) * inv
    kb_ = kbm1 - 1
    rt = dti2 * (rad[:, :, kb_] - rad[:, :, kb_ + 1]) / (dh * dz[kb_])
    num = c[:, :, kb_] * gg[kb_ - 1] - f_in[:, :, kb_] + rt
    den = c[:, :, kb_] * (1.0 - ee[kb_ - 1]) - 1.0
    col = [None] * kb
    col[kb_] = num / (den + 1e-20)
    for k in range(kbm1 - 2, -1, -1):
        col[k] = ee[k] * col[k + 1] + gg[k]
    for k in range(kbm1, kb):
        col[k] = f_in[:, :, k]
    return torch.stack(col, dim=2)


# --------------------------------------------------------------------------
# Equation of state: momentum.py:467-541
# --------------------------------------------------------------------------
def dens(S, t, s, rho):
    kbm1 = S.kbm1
    tr = t[:, :, :kbm1] + S.tbias
    sr = s[:, :, :kbm1] + S.sbias
    tr2 = tr * tr
    tr3 = tr2 * tr
    tr4 = tr3 * tr
    p = S.grav * S.rhoref * (-S.zz[:kbm1].view(1, 1, -1) * _b(S.h)) * 1.0e-5
    rhor = -0.157406 + 6.793952e-2 * tr - 9.095290e-3 * tr2 + \
        1.001685e-4 * tr3 - 1.120083e-6 * tr4 + 6.536332e-9 * tr4 * tr
    rhor = rhor + (0.824493 - 4.0899e-3 * tr + 7.6438e-5 * tr2 - 8.2467e-7 * tr3 +
                   5.3875e-9 * tr4) * sr
    rhor = rhor + (-5.72466e-3 + 1.0227e-4 * tr - 1.6546e-6 * tr2) * safe_abs_pow(sr, 1.5)
    rhor = rhor + 4.8314e-4 * sr * sr
    cr = 1449.1 + 0.0821 * p + 4.55 * tr - 0.045 * tr2 + 1.34 * (sr - 35.0)
    cr2 = cr * cr
    tp = 1.0e5 * p / cr2
    rhor = rhor + tp * (1.0 - 2.0 * p / cr2)
    out = rho.clone()
    out[:, :, :kbm1] = (rhor / S.rhoref) * _b(S.fsm)
    return out


# --------------------------------------------------------------------------
# Internal momentum: momentum.py:14-248
# --------------------------------------------------------------------------
def advu(S, advx, drhox, u, v, w, ub, dt, egf, egb, etf, etb, e_atmos):
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    aru, dz, cor, dy, h, grav, dti2 = S.aru, S.dz, S.cor, S.dy, S.h, S.grav, S.dti2
    vfl = torch.zeros((im, jm, kb), dtype=u.dtype, device=u.device)
    A, Am = slice(1, im), slice(0, im - 1)
    Kk, Km = slice(1, kbm1), slice(0, kbm1 - 1)
    vfl[A, :, Kk] = 0.25 * (w[A, :, Kk] + w[Am, :, Kk]) * (u[A, :, Kk] + u[A, :, Km])
    K, Kp = slice(0, kbm1), slice(1, kbm1 + 1)
    I, Im = slice(1, im - 1), slice(0, im - 2)
    J, Jp = slice(1, jm - 1), slice(2, jm)
    vdiv = (vfl[I, J, K] - vfl[I, J, Kp]) * _b(aru[I, J]) / dz[K].view(1, 1, -1)
    c1 = _b(cor[I, J] * dt[I, J]) * (v[I, Jp, K] + v[I, J, K])
    c2 = _b(cor[Im, J] * dt[Im, J]) * (v[Im, Jp, K] + v[Im, J, K])
    corio = -_b(aru[I, J]) * 0.25 * (c1 + c2)
    geg = (egf[I, J] - egf[Im, J]) + (egb[I, J] - egb[Im, J])
    gpa = 2.0 * (e_atmos[I, J] - e_atmos[Im, J])
    pg = grav * 0.125 * _b((dt[I, J] + dt[Im, J]) * (geg + gpa) * (dy[I, J] + dy[Im, J]))
    rhs = advx[I, J, K] + vdiv + corio + pg + drhox[I, J, K]
    hs = h[I, J] + h[Im, J]
    dnew = _b((hs + etf[I, J] + etf[Im, J]) * aru[I, J])
    dold = _b((hs + etb[I, J] + etb[Im, J]) * aru[I, J])
    uf = torch.zeros((im, jm, kb), dtype=u.dtype, device=u.device)
    uf[I, J, K] = (dold * ub[I, J, K] - 2.0 * dti2 * rhs) / (dnew + EPS_DIV)
    return uf


def advv(S, advy, drhoy, u, v, w, vb, dt, egf, egb, etf, etb, e_atmos):
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    arv, dz, cor, dx, h, grav, dti2 = S.arv, S.dz, S.cor, S.dx, S.h, S.grav, S.dti2
    vfl = torch.zeros((im, jm, kb), dtype=u.dtype, device=u.device)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    Kk, Km = slice(1, kbm1), slice(0, kbm1 - 1)
    vfl[:, B, Kk] = 0.25 * (w[:, B, Kk] + w[:, Bm, Kk]) * (v[:, B, Kk] + v[:, B, Km])
    K, Kp = slice(0, kbm1), slice(1, kbm1 + 1)
    I, Ip = slice(1, im - 1), slice(2, im)
    J, Jm = slice(1, jm - 1), slice(0, jm - 2)
    vdiv = (vfl[I, J, K] - vfl[I, J, Kp]) * _b(arv[I, J]) / dz[K].view(1, 1, -1)
    c1 = _b(cor[I, J] * dt[I, J]) * (u[Ip, J, K] + u[I, J, K])
    c2 = _b(cor[I, Jm] * dt[I, Jm]) * (u[Ip, Jm, K] + u[I, Jm, K])
    corio = _b(arv[I, J]) * 0.25 * (c1 + c2)
    geg = (egf[I, J] - egf[I, Jm]) + (egb[I, J] - egb[I, Jm])
    gpa = 2.0 * (e_atmos[I, J] - e_atmos[I, Jm])
    pg = grav * 0.125 * _b((dt[I, J] + dt[I, Jm]) * (geg + gpa) * (dx[I, J] + dx[I, Jm]))
    rhs = advy[I, J, K] + vdiv + corio + pg + drhoy[I, J, K]
    hs = h[I, J] + h[I, Jm]
    dnew = _b((hs + etf[I, J] + etf[I, Jm]) * arv[I, J])
    dold = _b((hs + etb[I, J] + etb[I, Jm]) * arv[I, J])
    vf = torch.zeros((im, jm, kb), dtype=u.dtype, device=u.device)
    vf[I, J, K] = (dold * vb[I, J, K] - 2.0 * dti2 * rhs) / (dnew + EPS_DIV)
    return vf


# --------------------------------------------------------------------------
# Vertical viscosity with bottom drag: momentum.py:250-465
# --------------------------------------------------------------------------
def _profuv(S, f_in, wsurf, wbot, km, etf, cbc, ub, vb, mask, along_i):
    im, jm, kb, kbm1, kbm2 = S.im, S.jm, S.kb, S.kbm1, S.kbm2
    dti2, umol, dz, dzz, h = S.dti2, S.umol, S.dz, S.dzz, S.h
    A, B = slice(1, im), slice(1, jm)
    if along_i:
        Am, Bm = slice(0, im - 1), B
    else:
        Am, Bm = A, slice(0, jm - 1)
    dhf = 0.5 * (h[A, B] + etf[A, B] + h[Am, Bm] + etf[Am, Bm])   # (im-1, jm-1)
    kmf = 0.5 * (km[A, B, :] + km[Am, Bm, :])
    dh2 = dhf ** 2
    a = torch.zeros((im - 1, jm - 1, kb), dtype=f_in.dtype, device=f_in.device)
    c = torch.zeros((im - 1, jm - 1, kb), dtype=f_in.dtype, device=f_in.device)
    for k in range(1, kbm1):
        num = -dti2 * (kmf[:, :, k] + umol)
        a[:, :, k - 1] = num / (dz[k - 1] * dzz[k - 1] * dh2)
        c[:, :, k] = num / (dz[k] * dzz[k - 1] * dh2)
    fi = f_in[A, B]
    ee = [None] * kb
    gg = [None] * kb
    den = a[:, :, 0] - 1.0
    ee[0] = a[:, :, 0] / (den + 1e-20)
    gg[0] = (-dti2 * wsurf[A, B] / (-dz[0] * dhf) - fi[:, :, 0]) / (den + 1e-20)
    for k in range(1, kbm2):
        den = a[:, :, k] + c[:, :, k] * (1.0 - ee[k - 1]) - 1.0
        inv = 1.0 / (den + 1e-20)
        ee[k] = a[:, :, k] * inv
        gg[k] = (c[:, :, k] * gg[k - 1] - fi[:, :, k]) * inv
    # bottom boundary on the interior (i, j) in [1, im-2] x [1, jm-2]
    n0, n1 = im - 2, jm - 2
    I, J = slice(1, im - 1), slice(1, jm - 1)
    kb_ = kbm1 - 1
    if along_i:
        cbca = 0.5 * (cbc[I, J] + cbc[0:im - 2, J])
        vs = vb[I, J, kb_] + vb[I, 2:jm, kb_] + vb[0:im - 2, J, kb_] + vb[0:im - 2, 2:jm, kb_]
        mag = torch.sqrt(ub[I, J, kb_] ** 2 + (0.25 * vs) ** 2 + 1e-8)
    else:
        cbca = 0.5 * (cbc[I, J] + cbc[I, 0:jm - 2])
        us = ub[I, J, kb_] + ub[2:im, J, kb_] + ub[I, 0:jm - 2, kb_] + ub[2:im, 0:jm - 2, kb_]
        mag = torch.sqrt((0.25 * us) ** 2 + vb[I, J, kb_] ** 2 + 1e-8)
    tps = torch.zeros((im, jm), dtype=f_in.dtype, device=f_in.device)
    tps[I, J] = cbca * mag
    sub = (slice(0, n0), slice(0, n1))
    fric = tps[I, J] * dti2 / (-dz[kb_] * dhf[sub])
    den = fric - 1.0 - (ee[kb_ - 1][sub] - 1.0) * c[:, :, kb_][sub]
    num = c[:, :, kb_][sub] * gg[kb_ - 1][sub] - fi[:, :, kb_][sub]
    m = mask[I, J]
    col = [None] * kb
    col[kb_] = num / (den + 1e-20) * m
    for k in range(kb_ - 1, -1, -1):
        col[k] = (ee[k][sub] * col[k + 1] + gg[k][sub]) * m
    out = f_in.clone()
    out[I, J, :kbm1] = torch.stack(col[:kbm1], dim=2)
    wbot_o = wbot.clone()
    wbot_o[I, J] = -tps[I, J] * out[I, J, kb_]
    return out, wbot_o, tps


def profu(S, uf, wusurf, wubot, km, etf, ub, vb):
    return _profuv(S, uf, wusurf, wubot, km, etf, S.cbc, ub, vb, S.dum, True)


def profv(S, vf, wvsurf, wvbot, km, etf, ub, vb):
    return _profuv(S, vf, wvsurf, wvbot, km, etf, S.cbc, ub, vb, S.dvm, False)


# --------------------------------------------------------------------------
# One internal step and the loop: time_loop.py:75-709
# --------------------------------------------------------------------------
def external_mode(S, X):
    """isplit barotropic substeps (time_loop.py:275-431); X holds the 2-D state."""
    im, jm = S.im, S.jm
    I, Ip, Im = slice(1, im - 1), slice(2, im), slice(0, im - 2)
    J, Jp, Jm = slice(1, jm - 1), slice(2, jm), slice(0, jm - 2)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    dx, dy, art, aru, arv, cor, h, grav = S.dx, S.dy, S.art, S.aru, S.arv, S.cor, S.h, S.grav
    alpha, dte, smoth = S.alpha, S.dte, S.smoth
    tide = getattr(S, "tide", None) is not None
    vamax = None
    for iext in range(1, int(S.isplit) + 1):
        d = X.d
        fluxua = X.fluxua.clone()
        fluxva = X.fluxva.clone()
        fluxua[A, B] = 0.25 * (d[A, B] + d[Am, B]) * (dy[A, B] + dy[Am, B]) * X.ua[A, B]
        fluxva[A, B] = 0.25 * (d[A, B] + d[A, Bm]) * (dx[A, B] + dx[A, Bm]) * X.va[A, B]
        div = fluxua[Ip, J] - fluxua[I, J] + fluxva[I, Jp] - fluxva[I, J]
        elf = X.elf.clone()
        elf[I, J] = X.elb[I, J] + S.dte2 * (-(div / art[I, J]) - X.vfluxf[I, J])
        elf = tide_bcond1(S, elf) if tide else bcond1(S, elf)
        X.fluxua, X.fluxva = fluxua, fluxva
        if iext % int(S.ispadv) == 0:
            X.advua, X.advva, X.fluxua, X.fluxva, X.tps = advave(
                S, X.ua, X.va, X.uab, X.vab, d, X.aam2d, X.fluxua, X.fluxva, X.tps)
        el, elb, ua, va, uab, vab = X.el, X.elb, X.ua, X.va, X.uab, X.vab
        ct = 0.25 * (cor[I, J] * d[I, J] * (va[I, Jp] + va[I, J]) +
                     cor[Im, J] * d[Im, J] * (va[Im, Jp] + va[Im, J]))
        gt = 0.25 * grav * (dy[I, J] + dy[Im, J]) * (d[I, J] + d[Im, J]) * (
            (1.0 - 2.0 * alpha) * (el[I, J] - el[Im, J]) +
            alpha * (elb[I, J] - elb[Im, J] + elf[I, J] - elf[Im, J]) +
            X.e_atmos[I, J] - X.e_atmos[Im, J])
        prov = X.adx2d[I, J] + X.advua[I, J] - aru[I, J] * ct + gt + \
            X.drx2d[I, J] - aru[I, J] * (-X.wusurf[I, J] + X.wubot[I, J])
        den = (h[I, J] + elf[I, J] + h[Im, J] + elf[Im, J]) * aru[I, J]
        num = (h[I, J] + elb[I, J] + h[Im, J] + elb[Im, J]) * aru[I, J] * uab[I, J] - 4.0 * dte * prov
        uaf = X.uaf.clone()
        uaf[I, J] = num / (den + EPS_DIV)
        ct = 0.25 * (cor[I, J] * d[I, J] * (ua[Ip, J] + ua[I, J]) +
                     cor[I, Jm] * d[I, Jm] * (ua[Ip, Jm] + ua[I, Jm]))
        gt = 0.25 * grav * (dx[I, J] + dx[I, Jm]) * (d[I, J] + d[I, Jm]) * (
            (1.0 - 2.0 * alpha) * (el[I, J] - el[I, Jm]) +
            alpha * (elb[I, J] - elb[I, Jm] + elf[I, J] - elf[I, Jm]) +
            X.e_atmos[I, J] - X.e_atmos[I, Jm])
        prov = X.ady2d[I, J] + X.advva[I, J] + arv[I, J] * ct + gt + \
            X.dry2d[I, J] - arv[I, J] * (-X.wvsurf[I, J] + X.wvbot[I, J])
        den = (h[I, J] + elf[I, J] + h[I, Jm] + elf[I, Jm]) * arv[I, J]
        num = (h[I, J] + elb[I, J] + h[I, Jm] + elb[I, Jm]) * arv[I, J] * vab[I, J] - 4.0 * dte * prov
        vaf = X.vaf.clone()
        vaf[I, J] = num / (den + EPS_DIV)
        if tide:
            from . import implicit as IM
            uaf, vaf = IM.tide_bc2(S, uaf, vaf, el=el, d=d)
        else:
            uaf, vaf = bcond2(S, uaf, vaf, el, X.ramp)
        if iext == int(S.isplit) - 2:
            X.etf = 0.25 * smoth * elf
        elif iext == int(S.isplit) - 1:
            X.etf = X.etf + 0.5 * (1.0 - 0.5 * smoth) * elf
        elif iext == int(S.isplit):
            X.etf = (X.etf + 0.5 * elf) * S.fsm
        X.elf, X.uaf, X.vaf = elf, uaf, vaf
        vamax = torch.max(torch.abs(vaf))
        if float(vamax) <= S.vmaxl:
            ua = ua + 0.5 * smoth * (uab - 2.0 * ua + uaf)
            va = va + 0.5 * smoth * (vab - 2.0 * va + vaf)
            el = el + 0.5 * smoth * (elb - 2.0 * el + elf)
            X.elb, X.el, X.d = el, elf, h + elf
            X.uab, X.ua, X.vab, X.va = ua, uaf, va, vaf
            if iext != int(S.isplit):
                X.egf = X.egf + X.el * S.ispi
                utf = X.utf.clone()
                vtf = X.vtf.clone()
                utf[A, :] = utf[A, :] + X.ua[A, :] * (X.d[A, :] + X.d[Am, :]) * S.isp2i
                vtf[:, B] = vtf[:, B] + X.va[:, B] * (X.d[:, B] + X.d[:, Bm]) * S.isp2i
                X.utf, X.vtf = utf, vtf
    return vamax


def step(S, iint):
    """One internal step iint (time_loop.py:75-696).  Mutates S by rebinding."""
    im, jm, kbm1 = S.im, S.jm, S.kbm1
    I, J = slice(1, im - 1), slice(1, jm - 1)
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    S.iint = iint
    S.time = S.dti * iint * 1.0 / 86400.0 + S.time0
    if S.lramp:
        S.ramp = min(S.time / float(S.period), 1.0)
    else:
        S.ramp = 1.0

    # surface forcing (time_loop.py:118-164)
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

    # 3-D preparation (time_loop.py:169-270)
    S.advx, S.advy, S.curv = advct(S, S.u, S.v, S.ub, S.vb, S.aam, S.dt)
    S.rho, S.drhox, S.drhoy = baropg(S, S.rho, S.dt, S.ramp)
    S.aam = smagorinsky(S, S.u, S.v, S.aam)
    S.adx2d = vsum(S, S.advx)
    S.ady2d = vsum(S, S.advy)
    S.drx2d = vsum(S, S.drhox)
    S.dry2d = vsum(S, S.drhoy)
    S.aam2d = vsum(S, S.aam)
    S.advua, S.advva, S.fluxua, S.fluxva, S.tps = advave(
        S, S.ua, S.va, S.uab, S.vab, S.d, S.aam2d, S.fluxua, S.fluxva, S.tps)
    S.adx2d = S.adx2d - S.advua
    S.ady2d = S.ady2d - S.advva

    S.egf = S.el * S.ispi
    utf = S.utf.clone()
    vtf = S.vtf.clone()
    utf[A, :] = S.ua[A, :] * (S.d[A, :] + S.d[Am, :]) * S.isp2i
    vtf[:, B] = S.va[:, B] * (S.d[:, B] + S.d[:, Bm]) * S.isp2i
    S.utf, S.vtf = utf, vtf

    vamax = external_mode(S, S)
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
    return ok


def internal_mode(S):
    """3-D part of the internal step (time_loop.py:438-678)."""
    im, jm, kb, kbm1 = S.im, S.jm, S.kb, S.kbm1
    A, Am = slice(1, im), slice(0, im - 1)
    B, Bm = slice(1, jm), slice(0, jm - 1)
    K = slice(0, kbm1)
    # transport adjustment (time_loop.py:440-477)
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

    S.w = vertvl(S, S.u, S.v, S.w, S.dt, S.etf, S.etb, S.vfluxb, S.vfluxf)
    S.w = bcond5(S, S.w)

    q2f = advq(S, S.q2b, S.q2, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    q2lf = advq(S, S.q2lb, S.q2l, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    q2f, q2lf, S.l, S.kq, S.km, S.kh = profq(
        S, q2f, q2lf, S.q2, S.q2b, S.q2lb, S.km, S.kh, S.kq, S.u, S.v, S.t, S.s, S.rho,
        S.wusurf, S.wvsurf, S.wubot, S.wvbot, S.etf, S.dt)
    q2f, q2lf = bcond6(S, q2f, q2lf, S.q2, S.q2l, S.u, S.v)
    q2 = S.q2 + 0.5 * S.smoth * (q2f + S.q2b - 2.0 * S.q2)
    q2l = S.q2l + 0.5 * S.smoth * (q2lf + S.q2lb - 2.0 * S.q2l)
    S.q2b, S.q2, S.q2lb, S.q2l = q2, q2f, q2l, q2lf

    tb, t, tf = advt1(S, S.tb, S.t, S.tclim, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    sb, s, sf = advt1(S, S.sb, S.s, S.sclim, S.u, S.v, S.w, S.aam, S.dt, S.etb, S.etf)
    S.tb, S.t, S.sb, S.s = tb, t, sb, s
    tf = proft(S, tf, S.wtsurf, S.tsurf, S.nbct, S.kh, S.etf, S.swrad)
    sf = proft(S, sf, S.wssurf, S.ssurf, S.nbcs, S.kh, S.etf, S.swrad)
    tf, sf = bcond4(S, tf, sf, S.t, S.s, S.u, S.v, S.w, S.dt)
    t = S.t + 0.5 * S.smoth * (tf + S.tb - 2.0 * S.t)
    s = S.s + 0.5 * S.smoth * (sf + S.sb - 2.0 * S.s)
    S.tb, S.t, S.sb, S.s = t, tf, s, sf
    S.rho = dens(S, S.t, S.s, S.rho)

    uf = advu(S, S.advx, S.drhox, S.u, S.v, S.w, S.ub, S.dt, S.egf, S.egb, S.etf, S.etb, S.e_atmos)
    vf = advv(S, S.advy, S.drhoy, S.u, S.v, S.w, S.vb, S.dt, S.egf, S.egb, S.etf, S.etb, S.e_atmos)
    uf, S.wubot, _ = profu(S, uf, S.wusurf, S.wubot, S.km, S.etf, S.ub, S.vb)
    vf, S.wvbot, S.tps = profv(S, vf, S.wvsurf, S.wvbot, S.km, S.etf, S.ub, S.vb)
    uf, vf = bcond3(S, uf, vf, S.u, S.v, depth=S.d if getattr(S, "tide", None) is not None else None)

    dzw = S.dz[:kbm1].view(1, 1, -1)
    term = uf[:, :, K] + S.ub[:, :, K] - 2.0 * S.u[:, :, K]
    tps = torch.sum(term * dzw, dim=2)
    ubn = S.u.clone()
    ubn[:, :, K] = S.u[:, :, K] + 0.5 * S.smoth * (term - _b(tps))
    term = vf[:, :, K] + S.vb[:, :, K] - 2.0 * S.v[:, :, K]
    tps = torch.sum(term * dzw, dim=2)
    vbn = S.v.clone()
    vbn[:, :, K] = S.v[:, :, K] + 0.5 * S.smoth * (term - _b(tps))
    S.ub, S.u, S.vb, S.v = ubn, uf, vbn, vf


def run_time_loop(S):
    """Advance S.iend internal steps from S.time0 (time_loop.py:75-709).

    Returns False if the barotropic velocity bound vmaxl was exceeded (the
    reference prints "Abnormal end" and returns early, time_loop.py:702-709).
    """
    for iint in range(1, int(S.iend) + 1):
        if not step(S, iint):
            return False
    return True
