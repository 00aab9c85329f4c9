"""Build HIP-step inputs (packed dicts) from oracle states (test helper)."""
import math

import numpy as np
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.explicit import Ensemble, StepScalars

SCAL = ("dte", "isplit", "alpha", "smoth", "horcon", "grav", "rhoref", "kappa", "umol", "tprni", "small",
        "tbias", "sbias", "hmax", "rfe", "rfw", "rfn", "rfs", "ispadv", "ntp", "nbct", "nbcs", "vmaxl")


def scalars_of(S):
    return StepScalars(**{n: getattr(S, n) for n in SCAL})


def ramp_of(S, iint=1):
    t = S.dti * iint / 86400.0 + S.time0
    return min(t / float(S.period), 1.0)


def ensemble_inputs(states, dtype, device="cuda"):
    """Pack oracle States (one per member) into (ens, state, geom, vgrid)."""
    ens = Ensemble([(S.im, S.jm) for S in states], dtype=dtype, device=device)
    state = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.FIELDS}
    geom = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.GEOM}
    vg = torch.stack([torch.stack([S.z, S.zz, S.dz, S.dzz]) for S in states]).to(device, dtype).contiguous()
    return ens, state, geom, vg


def unpack(ens, flat, name):
    return [ens.view(flat, m, name) for m in range(len(ens.shapes))]
