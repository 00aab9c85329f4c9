"""Lateral boundary policies of the drop-in cores.

Each core runs one of two policies inside its HIP step:

* its closed basin (explicit: cores/explicit/baroclinic.py:496-928 with zero
  exterior values; semi-implicit: src/neuralpom/cores/implicit_p0_fixed/
  boundary/closed_basin.py:27-101), the default;
* the open-boundary tide policy of the reference's case adapter
  (src/neuralpom/cores/case_adapters/coastal_boundary.py:9-67), installed with
  :func:`install` exactly as the adapter's ``install`` is: tide nodes read
  from an ``i j amplitude phase`` table (1-based indices, phase in degrees, at
  most 99 rows), all four radiation coefficients set to 1 (or to
  ``radiation``), and the fixed set of the eta solve (edges, dry cells,
  nodes).  The implicit family runs on the semi-implicit core, the explicit
  family (forward-only, as the reference's case registry runs it) on the
  explicit core.

Where the adapter replaces ``bcond`` and ``solve_implicit_eta_cg`` with
Python closures, here the policy object is a description the core hands to
the step: elevation at the nodes (bcond 1; the explicit family on every
substep at the time of its substep counter, the other edge points keeping
their elevation), radiation conditions on the depth-mean transports
(bcond 2), the lifted eta solve (implicit family), and the open forms of
bcond 3-6 with the exterior tracer values tbe tbw tbn tbs / sbe sbw sbn sbs.
The attributes ``install`` sets on the core carry the adapter's names
(``_tide_i``, ``_tide_j``, ``_tide_amp``, ``_tide_phase``, ``_fixed``,
``_validation_family``, ``_tide_substep``), so state capture over
``vars(core)`` sees what it sees with the reference; the explicit core also
advances ``_tide_substep`` by ``isplit`` per step, as the adapter's bcond 1
does once per substep.
"""
from __future__ import annotations

import numpy as np
import torch

TIDE_SPEED = 29.03225806451612   # degrees per hour (coastal_boundary.py:9)
TIDE_PI = 3.1416                 # the adapter's value of pi (coastal_boundary.py:9)


class TideBoundary:
    """Open edges with a prescribed tidal elevation at fixed nodes.

    ``speed`` (degrees per hour) and ``pi`` default to the adapter's
    constants; an M2 tide of period P hours is ``speed = 360 / P``."""

    name = "independent_original_pom08_tide_adapter"

    def __init__(self, speed: float = TIDE_SPEED, pi: float = TIDE_PI):
        self.speed = float(speed)
        self.pi = float(pi)

    @property
    def coef(self) -> float:
        """speed pi / 180, the factor of t / 3600 in the argument (coastal_boundary.py:10)."""
        return self.speed * self.pi / 180

    def wt(self, seconds: float) -> float:
        """(speed pi / 180) t / 3600, the argument's time part (coastal_boundary.py:9-10)."""
        return self.coef * seconds / 3600

    @staticmethod
    def substep_seconds(core, counter: int) -> float:
        """Evaluation time of the explicit family at substep counter value
        ``counter`` (1-based, after the increment; coastal_boundary.py:14)."""
        return float(core.time) * 86400 - float(core.dti) + float(core.dte) * (counter - 1)

    @staticmethod
    def seconds(core) -> float:
        """Evaluation time of the implicit family (coastal_boundary.py:15, 36, 45)."""
        return float(core.time) * 86400 - float(core.dte)

    def tide(self, core, seconds: float | None = None):
        """The tide at the nodes, amp cos(wt - phase) (coastal_boundary.py:9-10)."""
        sec = self.seconds(core) if seconds is None else seconds
        return core._tide_amp * torch.cos(self.wt(sec) - core._tide_phase * self.pi / 180)

    def geometry(self, core):
        """Per-cell arrays of the step: node mask, amplitude, phase (radians),
        fixed mask of the eta solve, exterior T / S on the edge cells (j edges
        take the corners, as the north / south statements of bcond 4 run last)."""
        dt, dev = core.el.dtype, core.el.device
        im, jm, kb = core.im, core.jm, core.kb
        i, j = core._tide_i.to(dev), core._tide_j.to(dev)
        node = torch.zeros((im, jm), dtype=dt, device=dev)
        amp = torch.zeros_like(node)
        phase = torch.zeros_like(node)
        node[i, j] = 1.0
        amp[i, j] = core._tide_amp.to(dev, dt)
        ph = core._tide_phase.to(dev, torch.float64) * self.pi / 180
        phase[i, j] = ph.to(dt)
        fixed = core._fixed.to(dev).to(dt)
        out = {"tide_node": node, "tide_amp": amp, "tide_phase": phase, "eta_fixed": fixed}
        for name, (e, w, n, s) in (("tbnd", ("tbe", "tbw", "tbn", "tbs")), ("sbnd", ("sbe", "sbw", "sbn", "sbs"))):
            x = torch.zeros((im, jm, kb), dtype=dt, device=dev)
            x[im - 1] = getattr(core, e).to(dev, dt)
            x[0] = getattr(core, w).to(dev, dt)
            x[:, jm - 1] = getattr(core, n).to(dev, dt)
            x[:, 0] = getattr(core, s).to(dev, dt)
            out[name] = x
        return out


def fixed_set(fsm, i, j):
    """Points the eta solve holds: the edges, dry cells and the nodes (coastal_boundary.py:64-66)."""
    fixed = torch.ones(fsm.shape, dtype=torch.bool, device=fsm.device)
    fixed[1:-1, 1:-1] = False
    fixed |= fsm == 0
    fixed[i.to(fsm.device), j.to(fsm.device)] = True
    return fixed


def install(core, family: str, path=None, nodes=None, speed: float = TIDE_SPEED, pi: float = TIDE_PI,
            radiation: float = 1.0):
    """Install the tide policy on a core (coastal_boundary.py:55-64).

    ``family``: ``"implicit"`` on the semi-implicit core, ``"explicit"`` on
    the explicit core (forward-only).  ``path``: text table ``i j amplitude
    phase`` (1-based, at most 99 rows are read, as the adapter's
    ``np.loadtxt(path, max_rows=99)``); or ``nodes``, an (n, 4) array of the
    same columns.  ``radiation`` is the value given to rfe rfw rfn rfs (the
    adapter's: 1)."""
    from .implicit_core import POM2K_Core as ImplicitCore
    if family not in ("implicit", "explicit"):
        raise ValueError(f"tide family must be 'implicit' or 'explicit', got {family!r}")
    if (family == "implicit") != isinstance(core, ImplicitCore):
        raise NotImplementedError(f"the tide policy's {family} family runs on the "
                                  f"{'semi-implicit' if family == 'implicit' else 'explicit'} core")
    if (path is None) == (nodes is None):
        raise ValueError("give exactly one of path and nodes")
    data = np.loadtxt(path, max_rows=99) if path is not None else np.asarray(nodes, dtype=np.float64)
    data = np.atleast_2d(data)
    if data.ndim != 2 or data.shape[1] < 4:
        raise ValueError("tide table rows must be i j amplitude phase")
    i = torch.as_tensor(data[:, 0] - 1, dtype=torch.long, device=core.device)
    j = torch.as_tensor(data[:, 1] - 1, dtype=torch.long, device=core.device)
    if bool((i < 0).any() or (i >= core.im).any() or (j < 0).any() or (j >= core.jm).any()):
        raise ValueError("tide node outside the grid")
    core._validation_family = family
    core._tide_substep = 0
    core._tide_i, core._tide_j = i, j
    core._tide_amp = torch.as_tensor(data[:, 2], dtype=core.el.dtype, device=core.device)
    core._tide_phase = torch.as_tensor(data[:, 3], dtype=core.el.dtype, device=core.device)
    core.rfe = core.rfw = core.rfn = core.rfs = float(radiation)
    core._fixed = fixed_set(core.fsm, i, j)
    if family == "implicit":
        core._implicit_boundary_policy = TideBoundary(speed, pi)
    else:
        core._explicit_boundary_policy = TideBoundary(speed, pi)
    return core
