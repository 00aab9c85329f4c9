"""Build an oracle ``State`` from a golden ``.npz`` (test infrastructure only)."""
from __future__ import annotations

import numpy as np
import torch

from .explicit import State

SCALAR_NAMES = ("dte", "dti", "dte2", "dti2", "isplit", "ispi", "isp2i", "alpha", "smoth",
                "horcon", "grav", "rhoref", "kappa", "umol", "tprni", "small", "vmaxl",
                "tbias", "sbias", "hmax", "period", "ispadv", "ntp", "nbct", "nbcs",
                "rfe", "rfw", "rfn", "rfs", "z0b", "cbcmin", "cbcmax", "aam_init", "slmax")
_INT = {"isplit", "ispadv", "ntp", "nbct", "nbcs"}


def load_state(npz, prefix: str, dtype=torch.float64) -> State:
    """Tensors under ``prefix/`` and scalars under ``scalar/`` become attributes."""
    S = State()
    pre = prefix.rstrip("/") + "/"
    for key in npz.files:
        if key.startswith(pre):
            setattr(S, key[len(pre):], torch.from_numpy(np.array(npz[key])).to(dtype))
    for name in SCALAR_NAMES:
        key = f"scalar/{name}"
        if key in npz.files:
            val = float(npz[key])
            setattr(S, name, int(round(val)) if name in _INT else val)
    S.im, S.jm, S.kb = (int(x) for x in S.u.shape)
    S.kbm1, S.kbm2 = S.kb - 1, S.kb - 2
    S.lramp = True
    S.tsurf = torch.zeros((S.im, S.jm), dtype=dtype) if not hasattr(S, "tsurf") else S.tsurf
    S.ssurf = torch.zeros((S.im, S.jm), dtype=dtype) if not hasattr(S, "ssurf") else S.ssurf
    return S
