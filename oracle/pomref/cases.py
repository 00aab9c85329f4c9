"""Case set-up for the oracle (test infrastructure only).

``double_gyre`` restates the reference's explicit-core construction:
BaseCore defaults (src/neuralpom/cores/explicit/base.py:8-189), the
production clock and numerics (src/neuralpom/hybrid/core_pool.py:20-48,
src/neuralpom/v1/core_spec.py:8-47), the double-gyre case
(cores/explicit/cases.py:30-164), areas and masks (cores/explicit/grid.py:
113-166) and run_initialization (cores/explicit/initialization.py:10-99).
Pinned by tests/test_oracle_golden.py against the reference's own state after
construction (tests/golden/explicit_init_12x10.npz).
"""
from __future__ import annotations

import math

import torch

from .explicit import State, baropg, dens, vsum

SCALARS = dict(
    small=1e-9, rhoref=1035.0, hmax=2000.0, grav=9.806, kappa=0.40, z0b=0.010, cbcmin=0.00250,
    cbcmax=1.0, tprni=0.20, umol=2.e-5, vmaxl=100.0, slmax=2.0, ntp=2, nbct=1, nbcs=1, ispadv=5,
    aam_init=500.0, tbias=0.0, sbias=0.0, rfe=0.0, rfw=0.0, rfn=0.0, rfs=0.0)

F2 = ("aam2d", "advua", "advva", "adx2d", "ady2d", "art", "aru", "arv", "cbc", "cor", "d", "drx2d",
      "dry2d", "dt", "dum", "dvm", "dx", "dy", "h", "east_c", "east_e", "east_u", "east_v", "e_atmos",
      "egb", "egf", "el", "elb", "elf", "et", "etb", "etf", "fluxua", "fluxva", "fsm", "north_c",
      "north_e", "north_u", "north_v", "psi", "rot", "ssurf", "swrad", "vfluxb", "tps", "tsurf", "ua",
      "vfluxf", "uab", "uaf", "utb", "utf", "va", "vab", "vaf", "vtb", "vtf", "wssurf", "wtsurf",
      "wubot", "wusurf", "wvbot", "wvsurf")
F3 = ("aam", "advx", "advy", "a", "c", "drhox", "drhoy", "dtef", "ee", "gg", "kh", "km", "kq", "l",
      "q2b", "q2", "q2lb", "q2l", "rho", "rmean", "sb", "sclim", "s", "tb", "tclim", "t", "ub", "uf",
      "u", "vb", "vf", "v", "w", "zflux", "curv")


def _vertical_interp(z_src, val_src, z_dst):
    """cases.py:10-27 (torch.bucketize + linear weights)."""
    zf = torch.flip(z_src, dims=[0])
    vf = torch.flip(val_src, dims=[0])
    idx = torch.bucketize(z_dst, zf)
    idx = torch.clamp(idx, min=1, max=z_src.shape[0] - 1)
    z0, z1 = zf[idx - 1], zf[idx]
    v0, v1 = vf[idx - 1], vf[idx]
    den = z1 - z0
    den = torch.where(torch.abs(den) < 1e-6, torch.ones_like(den), den)
    w = (z_dst - z0) / den
    return v0 * (1.0 - w) + v1 * w


def _areas_masks(S):
    """grid.py:113-166"""
    S.art = S.dx * S.dy
    S.aru = torch.zeros_like(S.dx)
    S.arv = torch.zeros_like(S.dx)
    S.aru[1:, 1:] = 0.25 * (S.dx[1:, 1:] + S.dx[:-1, 1:]) * (S.dy[1:, 1:] + S.dy[:-1, 1:])
    S.arv[1:, 1:] = 0.25 * (S.dx[1:, 1:] + S.dx[1:, :-1]) * (S.dy[1:, 1:] + S.dy[1:, :-1])
    S.aru[0, :] = S.aru[1, :]
    S.arv[0, :] = S.arv[1, :]
    S.aru[:, 0] = S.aru[:, 1]
    S.arv[:, 0] = S.arv[:, 1]
    S.fsm = (S.h > 1.0).to(S.h.dtype)
    S.dum = torch.zeros_like(S.dx)
    S.dvm = torch.zeros_like(S.dx)
    S.dum[1:, 1:] = S.fsm[1:, 1:] * S.fsm[:-1, 1:]
    S.dvm[1:, 1:] = S.fsm[1:, 1:] * S.fsm[1:, :-1]
    S.dum[0, :] = S.dum[1, :]
    S.dvm[0, :] = S.dvm[1, :]
    S.dum[:, 0] = S.dum[:, 1]
    S.dvm[:, 0] = S.dvm[:, 1]


def base_state(im, jm, dtype=torch.float64, dte=20.0, isplit=30, alpha=0.225, smoth=0.10, horcon=0.05):
    S = State()
    S.im, S.jm, S.kb = im, jm, 16
    S.kbm1, S.kbm2 = S.kb - 1, S.kb - 2
    for k, v in SCALARS.items():
        setattr(S, k, v)
    S.dte, S.isplit = dte, isplit
    S.dti = dte * isplit
    S.dte2, S.dti2 = dte * 2.0, S.dti * 2.0
    S.ispi, S.isp2i = 1.0 / isplit, 1.0 / (2.0 * isplit)
    S.alpha, S.smoth, S.horcon = alpha, smoth, horcon
    S.lramp = True
    S.ramp = 0.0
    for n in F2:
        setattr(S, n, torch.zeros((im, jm), dtype=dtype))
    for n in F3:
        setattr(S, n, torch.zeros((im, jm, S.kb), dtype=dtype))
    for n in ("z", "zz", "dz", "dzz"):
        setattr(S, n, torch.zeros(S.kb, dtype=dtype))
    S.dx = torch.ones((im, jm), dtype=dtype)
    S.dy = torch.ones((im, jm), dtype=dtype)
    S.h = torch.ones((im, jm), dtype=dtype)
    # open-boundary values (base.py:181-189); zero for closed basins
    for n in ("ele", "elw", "uabe", "uabw"):
        setattr(S, n, torch.zeros(jm, dtype=dtype))
    for n in ("eln", "els", "vabn", "vabs"):
        setattr(S, n, torch.zeros(im, dtype=dtype))
    for n in ("sbe", "sbw", "tbe", "tbw", "ube", "ubw"):
        setattr(S, n, torch.zeros((jm, S.kb), dtype=dtype))
    for n in ("sbn", "sbs", "tbn", "tbs", "vbn", "vbs"):
        setattr(S, n, torch.zeros((im, S.kb), dtype=dtype))
    return S


def uniform_sigma(S):
    """cases.py:84-99"""
    kb = S.kb
    z = torch.tensor([-float(k) / float(kb - 1) for k in range(kb)], dtype=S.z.dtype)
    S.z = z
    S.dz = torch.zeros_like(z)
    S.zz = torch.zeros_like(z)
    S.dzz = torch.zeros_like(z)
    S.dz[:kb - 1] = z[:kb - 1] - z[1:]
    S.zz[:kb - 1] = 0.5 * (z[:kb - 1] + z[1:])
    S.dzz[:kb - 1] = S.dz[:kb - 1]
    S.zz[kb - 1] = 2.0 * S.zz[kb - 2] - S.zz[kb - 3]


def double_gyre(im, jm, dtype=torch.float64, **clock):
    """Construct and initialise the double-gyre basin (cases.py:30-164 + initialization.py:10-99)."""
    S = base_state(im, jm, dtype, **clock)
    deg2rad = math.pi / 180.0
    re, omega = 6371.0e3, 7.2921e-5
    dlat = 20.0 / float(jm - 2)
    dlon = 22.0 / float(im - 2)
    j_idx = torch.arange(jm, dtype=dtype).view(1, -1).repeat(im, 1)
    i_idx = torch.arange(im, dtype=dtype).view(-1, 1).repeat(1, jm)
    S.north_e = 30.0 + (j_idx - 1.0) * dlat
    S.east_e = 0.0 + (i_idx - 1.0) * dlon
    S.dy = torch.full((im, jm), re * dlat * deg2rad, dtype=dtype)
    S.dx = re * torch.cos(S.north_e * deg2rad) * dlon * deg2rad
    S.cor = 2.0 * omega * torch.sin(S.north_e * deg2rad)
    S.h = torch.full((im, jm), 2500.0, dtype=dtype)
    S.h[0, :] = 1.0
    S.h[im - 1, :] = 1.0
    S.h[:, 0] = 1.0
    S.h[:, jm - 1] = 1.0
    hmax_real = float(S.h.max())
    uniform_sigma(S)
    _areas_masks(S)
    S.sb = torch.full((im, jm, S.kb), 35.0, dtype=dtype)
    zref = torch.linspace(0, -(hmax_real + 100.0), steps=1000, dtype=dtype)
    tref = 2.0 + (30.0 - 2.0) * torch.exp(zref / 800.0)
    S.tb = _vertical_interp(zref, tref, S.z.view(1, 1, -1) * S.h.unsqueeze(2))
    S.tclim = S.tb.clone()
    S.sclim = S.sb.clone()
    # dens() here still sees t = s = 0 (cases.py:132 runs before t, s are set)
    S.rho = dens(S, S.t, S.s, S.rho)
    S.rmean = S.rho.clone()
    S.wusurf = -(0.1 / S.rhoref) * (1.0 - torch.cos(2.0 * math.pi * (S.north_e - 30.0) / 20.0)) * S.dum
    S.wvsurf = torch.zeros_like(S.h)
    _initialize(S)
    return S


def _initialize(S):
    """initialization.py:16-99"""
    kb, kbm1 = S.kb, S.kbm1
    mi, mj = S.im // 2, S.jm // 2
    S.period = float((2.0 * math.pi) / (torch.abs(S.cor[mi, mj]) + 1e-20) / 86400.0)
    S.time0 = S.time = 0.0
    S.ua, S.va, S.el, S.et = S.uab.clone(), S.vab.clone(), S.elb.clone(), S.etb.clone()
    S.etf = S.et.clone()
    S.d = S.h + S.el
    S.dt = S.h + S.et
    S.w = S.w.clone()
    S.w[:, :, 0] = S.vfluxf
    S.l = 0.1 * S.dt.unsqueeze(2).expand(-1, -1, kb).clone()
    S.q2b = torch.full_like(S.q2b, S.small)
    S.q2lb = S.l * S.q2b
    S.kh = S.l * torch.sqrt(S.q2b + 1e-8)
    S.km = S.kh.clone()
    S.kq = S.kh.clone()
    S.aam = torch.full_like(S.aam, S.aam_init)
    for dst, src in (("q2", "q2b"), ("q2l", "q2lb"), ("t", "tb"), ("s", "sb"), ("u", "ub"), ("v", "vb")):
        x = getattr(S, dst).clone()
        x[:, :, :kbm1] = getattr(S, src)[:, :, :kbm1]
        setattr(S, dst, x)
    S.rho = dens(S, S.t, S.s, S.rho)
    S.rho, S.drhox, S.drhoy = baropg(S, S.rho, S.dt, S.ramp)
    S.drx2d = vsum(S, S.drhox)
    S.dry2d = vsum(S, S.drhoy)
    term = torch.clamp((1.0 + S.zz[kbm1 - 1]) * S.h / S.z0b, min=1e-12)
    S.cbc = torch.clamp((S.kappa / torch.log(torch.clamp(term, min=1e-12))) ** 2, min=S.cbcmin, max=S.cbcmax)
    dxs = torch.clamp(S.dx, min=1e-6)
    dys = torch.clamp(S.dy, min=1e-6)
    S.tps = (0.5 / torch.sqrt(1.0 / dxs ** 2 + 1.0 / dys ** 2)) / torch.sqrt(S.grav * (S.h + S.small)) * S.fsm
    S.d = S.h + S.el
    S.dt = S.h + S.et
    return S
