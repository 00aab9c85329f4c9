"""Case construction for the explicit core (one-time set-up, torch on device).

``double_gyre`` follows the reference's construction of the production basin:
src/neuralpom/cores/explicit/base.py:8-189 (defaults),
src/neuralpom/cores/explicit/cases.py:30-164 (double gyre),
src/neuralpom/cores/explicit/grid.py:113-166 (areas, masks),
src/neuralpom/cores/explicit/initialization.py:10-99 (initial state).

``seamount`` is the stratified seamount ensemble member of the headline
benchmark: closed basin, uniform dx = dy, h = H - 0.6 H exp(-r^2/L^2) with
H = 4500 m and L = 25 km, temperature with constant N^2 = 1e-5 s^-2 through
the model's equation of state (the potential density of the EOS polynomial,
momentum.py:507-532 without its pressure term, rises by rhoref N^2 / g per
metre of depth from T = 25 C at the surface, S = 35), an f-plane at 40 N, and
seeded Gaussian temperature noise; the mean density rmean is the density of
the noise-free stratification, so the sigma-coordinate pressure gradient of
the background is zero and the noise drives the flow.  It reuses the
double-gyre initialisation sequence.

Set-up is not on the per-step path, so it is written with torch tensor ops;
the equation of state used here mirrors the kernel's (momentum.py:467-541).
"""
from __future__ import annotations

import math

import torch

KB = 16
SCALARS = dict(small=1e-9, rhoref=1035.0, hmax=2000.0, grav=9.806, kappa=0.40, z0b=0.010,
               cbcmin=0.00250, cbcmax=1.0, aam_init=500.0, tbias=0.0, sbias=0.0)

STATE_3D = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq",
            "aam", "rho", "w", "l", "advx", "advy", "drhox", "drhoy")
STATE_2D = ("el", "elb", "elf", "ua", "uab", "uaf", "va", "vab", "vaf", "d", "dt", "et", "etb", "etf",
            "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "advua", "advva", "adx2d", "ady2d",
            "drx2d", "dry2d", "aam2d", "fluxua", "fluxva", "tps", "e_atmos", "vfluxf", "vfluxb", "swrad",
            "wtsurf", "wssurf")
GEOM_2D = ("h", "dx", "dy", "art", "aru", "arv", "fsm", "dum", "dvm", "cor", "cbc", "wusurf", "wvsurf",
           "tsurf", "ssurf")
GEOM_3D = ("rmean", "tclim", "sclim")
# open-boundary values along the edges (base.py:181-194): per-j arrays of the
# east / west edges, per-i arrays of the north / south edges
EDGE_J = ("ele", "elw", "uabe", "uabw")
EDGE_I = ("eln", "els", "vabn", "vabs")
EDGE_J3 = ("sbe", "sbw", "tbe", "tbw", "ube", "ubw")
EDGE_I3 = ("sbn", "sbs", "tbn", "tbs", "vbn", "vbs")


def eos(t, s, h, zz, fsm, grav=9.806, rhoref=1035.0):
    """(density - 1000)/rhoref on levels k < kb-1 (momentum.py:467-541)."""
    kbm1 = KB - 1
    tr = t[:, :, :kbm1]
    sr = s[:, :, :kbm1]
    tr2 = tr * tr
    tr3 = tr2 * tr
    tr4 = tr3 * tr
    p = grav * rhoref * (-zz[:kbm1].view(1, 1, -1) * h.unsqueeze(2)) * 1.0e-5
    r = -0.157406 + 6.793952e-2 * tr - 9.095290e-3 * tr2 + 1.001685e-4 * tr3 - 1.120083e-6 * tr4 + \
        6.536332e-9 * tr4 * tr
    r = r + (0.824493 - 4.0899e-3 * tr + 7.6438e-5 * tr2 - 8.2467e-7 * tr3 + 5.3875e-9 * tr4) * sr
    r = r + (-5.72466e-3 + 1.0227e-4 * tr - 1.6546e-6 * tr2) * torch.pow(torch.clamp(sr.abs(), min=1e-8), 1.5)
    r = r + 4.8314e-4 * sr * sr
    cr = 1449.1 + 0.0821 * p + 4.55 * tr - 0.045 * tr2 + 1.34 * (sr - 35.0)
    cr2 = cr * cr
    r = r + 1.0e5 * p / cr2 * (1.0 - 2.0 * p / cr2)
    out = torch.zeros_like(t)
    out[:, :, :kbm1] = (r / rhoref) * fsm.unsqueeze(2)
    return out


def eos_potential(t, s):
    """Density - 1000 (kg/m^3) of the EOS polynomial at zero pressure
    (momentum.py:507-532 without the pressure term)."""
    t2 = t * t
    t3 = t2 * t
    t4 = t3 * t
    r = -0.157406 + 6.793952e-2 * t - 9.095290e-3 * t2 + 1.001685e-4 * t3 - 1.120083e-6 * t4 + \
        6.536332e-9 * t4 * t
    r = r + (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t2 - 8.2467e-7 * t3 + 5.3875e-9 * t4) * s
    r = r + (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t2) * torch.pow(torch.clamp(s.abs(), min=1e-8), 1.5)
    return r + 4.8314e-4 * s * s


def constant_n2_temperature(depth, n2, t0=25.0, s0=35.0, grav=9.806, rhoref=1035.0):
    """T(depth) (depth >= 0 in metres, any shape) at salinity s0 whose
    potential density under the model EOS rises linearly, rhoref n2 / grav
    per metre, from t0 at the surface: the buoyancy frequency is n2 at every
    depth.  Newton iterations on the EOS polynomial in float64."""
    d = torch.as_tensor(depth, dtype=torch.float64)
    s = torch.full_like(d, s0)
    target = eos_potential(torch.full_like(d, t0), s) + rhoref * n2 / grav * d
    t = torch.full_like(d, t0) - n2 / (grav * 2.0e-4) * d   # linear-expansion first guess
    for _ in range(50):
        tv = t.detach().requires_grad_(True)
        f = eos_potential(tv, s) - target
        (g,) = torch.autograd.grad(f.sum(), tv)
        step = f.detach() / g
        t = tv.detach() - step
        if float(step.abs().max()) < 1e-12:
            break
    return t


def _uniform_sigma(dtype, device):
    z = torch.tensor([-float(k) / float(KB - 1) for k in range(KB)], dtype=dtype, device=device)
    dz = torch.zeros_like(z)
    zz = torch.zeros_like(z)
    dz[:KB - 1] = z[:KB - 1] - z[1:]
    zz[:KB - 1] = 0.5 * (z[:KB - 1] + z[1:])
    dzz = dz.clone()
    zz[KB - 1] = 2.0 * zz[KB - 2] - zz[KB - 3]
    return z, zz, dz, dzz


def _areas_masks(g):
    dx, dy, h = g["dx"], g["dy"], g["h"]
    g["art"] = dx * dy
    aru = torch.zeros_like(dx)
    arv = torch.zeros_like(dx)
    aru[1:, 1:] = 0.25 * (dx[1:, 1:] + dx[:-1, 1:]) * (dy[1:, 1:] + dy[:-1, 1:])
    arv[1:, 1:] = 0.25 * (dx[1:, 1:] + dx[1:, :-1]) * (dy[1:, 1:] + dy[1:, :-1])
    aru[0, :] = aru[1, :]
    arv[0, :] = arv[1, :]
    aru[:, 0] = aru[:, 1]
    arv[:, 0] = arv[:, 1]
    fsm = (h > 1.0).to(h.dtype)
    dum = torch.zeros_like(dx)
    dvm = torch.zeros_like(dx)
    dum[1:, 1:] = fsm[1:, 1:] * fsm[:-1, 1:]
    dvm[1:, 1:] = fsm[1:, 1:] * fsm[1:, :-1]
    dum[0, :] = dum[1, :]
    dvm[0, :] = dvm[1, :]
    dum[:, 0] = dum[:, 1]
    dvm[:, 0] = dvm[:, 1]
    g.update(aru=aru, arv=arv, fsm=fsm, dum=dum, dvm=dvm)


def _initialize(st, g, vg, im, jm, period_cor):
    """initialization.py:16-99 on dict state ``st`` and geometry ``g``."""
    z, zz, dz, dzz = vg
    small = SCALARS["small"]
    kbm1 = KB - 1
    st["d"] = g["h"] + st["el"]
    st["dt"] = g["h"] + st["et"]
    st["w"][:, :, 0] = st["vfluxf"]
    l = 0.1 * st["dt"].unsqueeze(2).expand(-1, -1, KB).clone()
    st["l"] = l
    st["q2b"] = torch.full_like(l, small)
    st["q2lb"] = l * st["q2b"]
    st["kh"] = l * torch.sqrt(st["q2b"] + 1e-8)
    st["km"] = st["kh"].clone()
    st["kq"] = st["kh"].clone()
    st["aam"] = torch.full_like(l, SCALARS["aam_init"])
    for dst, src in (("q2", "q2b"), ("q2l", "q2lb"), ("t", "tb"), ("s", "sb"), ("u", "ub"), ("v", "vb")):
        st[dst][:, :, :kbm1] = st[src][:, :, :kbm1]
    st["rho"] = eos(st["t"], st["s"], g["h"], zz, g["fsm"])
    # baropg at ramp = 0 leaves drhox = drhoy = 0; rho passes through (rho - rmean) + rmean
    st["rho"] = (st["rho"] - g["rmean"]) + g["rmean"]
    term = torch.clamp((1.0 + zz[kbm1 - 1]) * g["h"] / SCALARS["z0b"], min=1e-12)
    g["cbc"] = torch.clamp((SCALARS["kappa"] / torch.log(term)) ** 2, min=SCALARS["cbcmin"], max=SCALARS["cbcmax"])
    dxs = torch.clamp(g["dx"], min=1e-6)
    dys = torch.clamp(g["dy"], min=1e-6)
    st["tps"] = (0.5 / torch.sqrt(1.0 / dxs ** 2 + 1.0 / dys ** 2)) / torch.sqrt(
        SCALARS["grav"] * (g["h"] + small)) * g["fsm"]
    period = (2.0 * math.pi) / (abs(period_cor) + 1e-20) / 86400.0
    return period


def _empty(im, jm, dtype, device):
    st = {n: torch.zeros((im, jm, KB), dtype=dtype, device=device) for n in STATE_3D}
    st.update({n: torch.zeros((im, jm), dtype=dtype, device=device) for n in STATE_2D})
    g = {n: torch.zeros((im, jm), dtype=dtype, device=device) for n in GEOM_2D}
    g.update({n: torch.zeros((im, jm, KB), dtype=dtype, device=device) for n in GEOM_3D})
    g.update({n: torch.zeros(jm, dtype=dtype, device=device) for n in EDGE_J})
    g.update({n: torch.zeros(im, dtype=dtype, device=device) for n in EDGE_I})
    g.update({n: torch.zeros((jm, KB), dtype=dtype, device=device) for n in EDGE_J3})
    g.update({n: torch.zeros((im, KB), dtype=dtype, device=device) for n in EDGE_I3})
    return st, g


def _interp_profile(zref, tref, zdst):
    zf = torch.flip(zref, dims=[0])
    vf = torch.flip(tref, dims=[0])
    idx = torch.clamp(torch.bucketize(zdst, zf), min=1, max=zref.shape[0] - 1)
    z0, z1 = zf[idx - 1], zf[idx]
    v0, v1 = vf[idx - 1], vf[idx]
    den = z1 - z0
    den = torch.where(den.abs() < 1e-6, torch.ones_like(den), den)
    w = (zdst - z0) / den
    return v0 * (1.0 - w) + v1 * w


def double_gyre(im, jm, dtype=torch.float32, device="cpu"):
    """Reference double-gyre basin; returns (state, geom, vgrid, period_days)."""
    st, g = _empty(im, jm, dtype, device)
    deg2rad = math.pi / 180.0
    re, omega = 6371.0e3, 7.2921e-5
    dlat, dlon = 20.0 / float(jm - 2), 22.0 / float(im - 2)
    j_idx = torch.arange(jm, dtype=dtype, device=device).view(1, -1).repeat(im, 1)
    north = 30.0 + (j_idx - 1.0) * dlat
    i_idx = torch.arange(im, dtype=dtype, device=device).view(-1, 1).repeat(1, jm)
    g["north_e"] = north.clone()                   # cases.py:54-55 (grid-point latitude / longitude)
    g["east_e"] = (i_idx - 1.0) * dlon
    g["dy"] = torch.full((im, jm), re * dlat * deg2rad, dtype=dtype, device=device)
    g["dx"] = re * torch.cos(north * deg2rad) * dlon * deg2rad
    g["cor"] = 2.0 * omega * torch.sin(north * deg2rad)
    h = torch.full((im, jm), 2500.0, dtype=dtype, device=device)
    h[0, :] = h[-1, :] = 1.0
    h[:, 0] = h[:, -1] = 1.0
    g["h"] = h
    vg = _uniform_sigma(dtype, device)
    _areas_masks(g)
    st["sb"].fill_(35.0)
    zref = torch.linspace(0, -(float(h.max()) + 100.0), steps=1000, dtype=dtype, device=device)
    st["tb"] = _interp_profile(zref, 2.0 + 28.0 * torch.exp(zref / 800.0), vg[0].view(1, 1, -1) * h.unsqueeze(2))
    g["tclim"] = st["tb"].clone()
    g["sclim"] = st["sb"].clone()
    g["rmean"] = eos(st["t"], st["s"], h, vg[1], g["fsm"])   # t = s = 0 at this point (cases.py:132)
    g["wusurf"] = -(0.1 / SCALARS["rhoref"]) * (1.0 - torch.cos(2.0 * math.pi * (north - 30.0) / 20.0)) * g["dum"]
    period = _initialize(st, g, vg, im, jm, float(g["cor"][im // 2, jm // 2]))
    return st, g, torch.stack(vg), period


def shelf_basin(im=90, jm=82, dx=10.0e3, dtype=torch.float32, device="cpu"):
    """Closed rectangular basin with a shelf: uniform dx = dy, depth
    200 + 1800 tanh(d / 5) m with d the distance in cells from the wet cells next to the land ring
    (h = 1 there), zonal wind stress 0.1 cos(2 pi y / Ly) N/m^2 with y from the
    southern wet row and Ly = (jm - 2) dy, T = 5 + 15 exp(z / 500 m) on the
    interface depths z h, S = 35, Coriolis 2 Omega sin(lat) over 30-50 N as in
    the double gyre.  The mean density rmean is the initial density (the double
    gyre's rmean = rho(T = 0, S = 0) puts a large sigma-coordinate pressure
    gradient error on the shelf break).  Returns (state, geom, vgrid, period_days)."""
    st, g = _empty(im, jm, dtype, device)
    deg2rad = math.pi / 180.0
    omega = 7.2921e-5
    dlat = 20.0 / float(jm - 2)
    j_idx = torch.arange(jm, dtype=dtype, device=device).view(1, -1).repeat(im, 1)
    i_idx = torch.arange(im, dtype=dtype, device=device).view(-1, 1).repeat(1, jm)
    north = 30.0 + (j_idx - 1.0) * dlat
    g["north_e"] = north.clone()
    g["east_e"] = (i_idx - 1.0) * dx / 1.0e3
    g["dx"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["dy"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["cor"] = 2.0 * omega * torch.sin(north * deg2rad)
    dist = torch.minimum(torch.minimum(i_idx, (im - 1) - i_idx), torch.minimum(j_idx, (jm - 1) - j_idx))
    h = 200.0 + 1800.0 * torch.tanh((dist - 1.0) / 5.0)   # 200 m on the cells next to the land ring
    h[0, :] = h[-1, :] = 1.0
    h[:, 0] = h[:, -1] = 1.0
    g["h"] = h
    vg = _uniform_sigma(dtype, device)
    _areas_masks(g)
    st["sb"].fill_(35.0)
    st["tb"] = 5.0 + 15.0 * torch.exp(vg[0].view(1, 1, -1) * h.unsqueeze(2) / 500.0)
    g["tclim"] = st["tb"].clone()
    g["sclim"] = st["sb"].clone()
    ly = (jm - 2) * dx
    y = (j_idx - 1.0) * dx
    g["wusurf"] = -(0.1 * torch.cos(2.0 * math.pi * y / ly) / SCALARS["rhoref"]) * g["dum"]
    g["rmean"] = eos(st["tb"], st["sb"], h, vg[1], g["fsm"])
    period = _initialize(st, g, vg, im, jm, float(g["cor"][im // 2, jm // 2]))
    return st, g, torch.stack(vg), period


def seamount(im, jm, seed=0, dx=8000.0, H=4500.0, L=25.0e3, n2=1.0e-5, noise=0.1,
             dtype=torch.float32, device="cpu"):
    """Stratified seamount member; returns (state, geom, vgrid, period_days)."""
    st, g = _empty(im, jm, dtype, device)
    omega = 7.2921e-5
    g["dx"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["dy"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["cor"] = torch.full((im, jm), 2.0 * omega * math.sin(math.radians(40.0)), dtype=dtype, device=device)
    x = (torch.arange(im, dtype=dtype, device=device) - 0.5 * (im - 1)) * dx
    y = (torch.arange(jm, dtype=dtype, device=device) - 0.5 * (jm - 1)) * dx
    r2 = x.view(-1, 1) ** 2 + y.view(1, -1) ** 2
    h = H - 0.6 * H * torch.exp(-r2 / L ** 2)
    h[0, :] = h[-1, :] = 1.0
    h[:, 0] = h[:, -1] = 1.0
    g["h"] = h
    vg = _uniform_sigma(dtype, device)
    _areas_masks(g)
    st["sb"].fill_(35.0)
    # constant N^2 through the model EOS, at the depths of the levels (zz h)
    depth = -(vg[1].view(1, 1, -1).double().cpu() * h.unsqueeze(2).double().cpu())
    tbg = constant_n2_temperature(depth, n2, t0=25.0, s0=35.0, grav=SCALARS["grav"],
                                  rhoref=SCALARS["rhoref"]).to(device, dtype)
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    st["tb"] = tbg + noise * torch.randn(tbg.shape, generator=gen, dtype=torch.float64).to(device, dtype)
    g["tclim"] = st["tb"].clone()
    g["sclim"] = st["sb"].clone()
    # mean density = density of the noise-free stratification: the background's
    # sigma-coordinate pressure gradient over the seamount is then exactly zero,
    # and the seeded noise (a density anomaly of ~2e-5 relative) drives the flow
    g["rmean"] = eos(tbg, st["sb"], h, vg[1], g["fsm"])
    g["hmax"] = H
    period = _initialize(st, g, vg, im, jm, float(g["cor"][im // 2, jm // 2]))
    g.pop("hmax")
    return st, g, torch.stack(vg), period


def gaussian_random_field(im, jm, corr, seed, dtype=torch.float32, device="cpu"):
    """Seeded stationary Gaussian random field with correlation exp(-r^2 / (2 corr^2))
    (r in cells): white noise filtered by a Gaussian of width corr / sqrt(2), by FFT
    on a periodic domain padded by 4 corr so the ends do not wrap."""
    pad = int(math.ceil(4 * corr))
    ni, nj = im + 2 * pad, jm + 2 * pad
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    w = torch.randn((ni, nj), generator=gen, dtype=torch.float64)
    sig = corr / math.sqrt(2.0)
    ki = torch.fft.fftfreq(ni, dtype=torch.float64).view(-1, 1)
    kj = torch.fft.fftfreq(nj, dtype=torch.float64).view(1, -1)
    filt = torch.exp(-2.0 * (math.pi * sig) ** 2 * (ki ** 2 + kj ** 2))
    f = torch.fft.ifft2(torch.fft.fft2(w) * filt).real[pad:pad + im, pad:pad + jm]
    return f.to(device, dtype)


def tidal_channel(im=354, jm=322, seed=0, dx=500.0, hmin=10.0, hmax=60.0, corr=20.0, amp=1.0, phase=90.0,
                  t0=15.0, s0=32.0, dtdz=0.0, lat=36.0, open_east=False, dtype=torch.float32, device="cpu"):
    """Open-boundary tidal channel of the semi-implicit core; returns (state, geom,
    vgrid, period_days, nodes).

    Land on the north, south and east edges; the west end is open.  Depth:
    a seeded Gaussian random field (correlation length ``corr`` cells) mapped
    to [hmin, hmax]; dx = dy; an f-plane at ``lat``; T = t0 + dtdz z (z the
    depth in metres, negative downwards; dtdz = 0: uniform) and uniform S, the
    exterior values of the open edges from the same profiles, the initial
    density as the mean density; no wind.  ``nodes`` is the tide table of the open-boundary policy
    (1-based i j, amplitude, phase in degrees): the wet edge cells of the west
    end, amplitude ``amp``, phase ``phase`` (90: the tide rises from zero at
    t = 0, amp sin(omega t), so the start carries no elevation jump); the
    policy's speed for an M2-like period P hours is 360 / P degrees per hour.

    Install the policy with ``radiation=0``: the elevation is clamped at the
    west edge and the face velocity next to it is the corrector's, so the
    barotropic transport stays consistent with the eta solve.  With the
    adapter's radiation coefficients of 1 the corrected face velocity is
    overwritten after the solve; the column next to an open edge then gains
    or loses volume the continuity equation does not see, and its bottom
    layer drifts in T (a property of the reference adapter, reproduced by the
    HIP step)."""
    st, g = _empty(im, jm, dtype, device)
    omega = 7.2921e-5
    g["dx"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["dy"] = torch.full((im, jm), dx, dtype=dtype, device=device)
    g["cor"] = torch.full((im, jm), 2.0 * omega * math.sin(math.radians(lat)), dtype=dtype, device=device)
    g["north_e"] = torch.full((im, jm), lat, dtype=dtype, device=device)
    f = gaussian_random_field(im, jm, corr, seed, torch.float64, device)
    h = (hmin + (f - f.min()) / (f.max() - f.min()) * (hmax - hmin)).to(dtype)
    h[:, 0] = h[:, -1] = 1.0
    if not open_east:
        h[-1, :] = 1.0
    g["h"] = h
    vg = _uniform_sigma(dtype, device)
    _areas_masks(g)
    st["tb"] = t0 + dtdz * vg[0].view(1, 1, -1) * h.unsqueeze(2)
    st["sb"].fill_(s0)
    g["tclim"] = st["tb"].clone()
    g["sclim"] = st["sb"].clone()
    g["rmean"] = eos(st["tb"], st["sb"], h, vg[1], g["fsm"])
    g["tbe"], g["tbw"] = st["tb"][-1].clone(), st["tb"][0].clone()
    g["tbn"], g["tbs"] = st["tb"][:, -1].clone(), st["tb"][:, 0].clone()
    for n in ("sbe", "sbw", "sbn", "sbs"):
        g[n].fill_(s0)
    period = _initialize(st, g, vg, im, jm, float(g["cor"][im // 2, jm // 2]))
    js = torch.arange(1, jm - 1, dtype=torch.float64)
    nodes = torch.stack([torch.full_like(js, 1.0), js + 1.0, torch.full_like(js, float(amp)),
                         torch.full_like(js, float(phase))], dim=1).numpy()
    return st, g, torch.stack(vg), period, nodes



def forecast_geometry(fine, coarse, fine_geom, dx):
    """Geometry of the 4x coarser forecast grid of the fine members (the
    354 x 322 -> 90 x 82 pair of the coupled rollout): the depth is the
    area-weighted restriction of the fine depth over wet cells, land (h = 1)
    where a 4 x 4 block holds no wet cell and on the edge ring; masks and
    areas follow from it as for every case here (dx = dy = ``dx``).  Returns
    flat fields in the coarse batch's layout: h, dx, dy, art, aru, arv, fsm,
    dum, dvm."""
    from .state import restrict4
    hc = restrict4(fine, coarse, fine_geom["h"].detach(), "h", fine_geom, volume=False)
    out = {n: [] for n in ("h", "dx", "dy", "art", "aru", "arv", "fsm", "dum", "dvm")}
    for m, (im, jm) in enumerate(coarse.shapes):
        h = coarse.view(hc, m, "h").clone()
        h[h <= 1.0] = 1.0
        g = {"h": h, "dx": torch.full_like(h, float(dx)), "dy": torch.full_like(h, float(dx))}
        _areas_masks(g)
        for n in out:
            out[n].append(g[n])
    return {n: coarse.pack(v, n) for n, v in out.items()}
