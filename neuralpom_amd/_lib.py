"""ctypes binding of libneuralpom_amd.so (declarations: include/neuralpom_amd.h).

The library is the only compute path of this package: if it is missing or a
HIP device is absent, every entry point raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libneuralpom_amd.so"

POM_OK, POM_E_ARG, POM_E_HIP, POM_E_UNSUPPORTED, POM_E_VMAX, POM_E_SCRATCH = 0, 1, 2, 3, 4, 5
NOALLOC_SCRATCH_BYTES = 64 << 20   # POM_NOALLOC_SCRATCH_BYTES
POM_F32, POM_F64 = 0, 1
POM_CORE_EXPLICIT, POM_CORE_IMPLICIT = 0, 1
POM_PRECOND_JACOBI, POM_PRECOND_SPECTRAL = 0, 1
POM_BOUNDARY_CLOSED, POM_BOUNDARY_TIDE = 0, 1

# pom_field order (include/neuralpom_amd.h)
FIELDS_3D = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "q2", "q2b", "q2l", "q2lb",
             "km", "kh", "kq", "aam", "rho", "w", "l", "advx", "advy", "drhox", "drhoy")
FIELDS_2D = ("el", "elb", "elf", "ua", "uab", "uaf", "va", "vab", "vaf", "d", "dt", "et", "etb",
             "etf", "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "advua", "advva",
             "adx2d", "ady2d", "drx2d", "dry2d", "aam2d", "fluxua", "fluxva", "tps",
             "e_atmos", "vfluxf", "vfluxb", "swrad", "wtsurf", "wssurf")
FIELDS = FIELDS_3D + FIELDS_2D
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
GEOM_2D = ("h", "dx", "dy", "art", "aru", "arv", "fsm", "dum", "dvm", "cor", "cbc",
           "wusurf", "wvsurf", "tsurf", "ssurf")
GEOM_3D = ("rmean", "tclim", "sclim")
GEOM = GEOM_2D + GEOM_3D
# geometry of the open-boundary tide policy (read by the step only in that mode)
GEOM_OPEN_2D = ("tide_node", "tide_amp", "tide_phase", "eta_fixed")
GEOM_OPEN_3D = ("tbnd", "sbnd")
GEOM_ALL = GEOM + GEOM_OPEN_2D + GEOM_OPEN_3D   # pom_geom order
THREE_D = frozenset(FIELDS_3D + GEOM_3D + GEOM_OPEN_3D)

# inputs with a cotangent (pom_explicit_step_adj)
DIFF_INPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "q2", "q2b", "q2l", "q2lb", "km",
               "kh", "kq", "aam", "rho", "el", "elb", "ua", "uab", "va", "vab", "d", "dt", "et",
               "etb", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot")


class Member(C.Structure):
    _fields_ = [("im", C.c_int32), ("jm", C.c_int32), ("off2", C.c_int64), ("off3", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("map2_blocks_per_group", "map2_min_groups", "flat2_blocks_per_group",
                                         "flat2_min_groups", "flat3_blocks_per_group", "flat3_min_groups")]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dti", "dte", "dti2", "dte2", "ispi", "isp2i", "alpha", "smoth", "horcon", "grav",
        "rhoref", "kappa", "umol", "tprni", "small", "tbias", "sbias", "hmax", "rfe", "rfw",
        "rfn", "rfs")] + [(n, C.c_int32) for n in ("isplit", "ispadv", "ntp", "nbct", "nbcs", "iext")] + [
        ("vmaxl", C.c_double)] + [(n, C.c_int32) for n in ("core", "eta_max_iter", "eta_precond", "boundary")] + [
        (n, C.c_double) for n in ("eta_tol", "eta_rel_tol", "gamma_baro", "tide_wt")] + [
        (n, C.c_int32) for n in ("eta_capture", "reserved0")] + [
        (n, C.c_double) for n in ("tracer_shapiro", "velocity_shapiro", "aam_min", "tide_coef", "tide_sec")] + [
        (n, C.c_int32) for n in ("tide_substep0", "reserved1")]


_lib = None


class HipExtensionError(RuntimeError):
    pass


def load():
    """Load the library; raise HipExtensionError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise HipExtensionError(
            f"{_LIB_PATH} is not built; run `python -m neuralpom_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(_LIB_PATH))
    vp, vpp = C.c_void_p, C.POINTER(C.c_void_p)
    lib.pom_batch_create.argtypes = [C.POINTER(Member), C.c_int32, C.c_int32, C.POINTER(vp)]
    lib.pom_batch_destroy.argtypes = [vp]
    lib.pom_batch_tune.argtypes = [vp, C.POINTER(Tuning)]
    lib.pom_batch_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.pom_explicit_workspace.argtypes = [vp, C.POINTER(Params), C.c_int32, C.POINTER(C.c_size_t)]
    lib.pom_explicit_step.argtypes = [vp, C.POINTER(Params), vp, vp, vpp, vpp, vpp, C.c_int32, vp, vp,
                                      C.c_size_t, vp]
    lib.pom_explicit_step_adj.argtypes = [vp, C.POINTER(Params), vp, vp, vpp, vpp, C.c_int32, vpp, vpp,
                                          vp, C.c_size_t, vp]
    lib.pom_explicit_step_taped.argtypes = lib.pom_explicit_step.argtypes
    lib.pom_explicit_step_reverse.argtypes = [vp, C.POINTER(Params), vp, vp, vpp, vpp, vpp, C.c_int32, vpp, vpp,
                                              vp, C.c_size_t, vp]
    lib.pom_dens.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pom_dens_adj.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pom_baropg.argtypes = [vp, C.POINTER(Params), vp, vp, vpp, vpp, vp]
    lib.pom_baropg_adj.argtypes = [vp, C.POINTER(Params), vp, vp, vpp, vpp, vpp, vp]
    lib.pom_implicit_cache_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.pom_implicit_prepare.argtypes = [vp, C.POINTER(Params), vpp, vp, C.c_size_t, vp]
    lib.pom_implicit_workspace.argtypes = [vp, C.POINTER(Params), C.c_int32, C.POINTER(C.c_size_t)]
    lib.pom_implicit_step.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vpp, vpp, vpp, C.c_int32, vp, vp, vp,
                                      C.c_size_t, vp]
    lib.pom_implicit_step_adj.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vpp, vpp, C.c_int32, vpp, vpp, vp,
                                          C.c_size_t, vp]
    lib.pom_implicit_step_taped.argtypes = lib.pom_implicit_step.argtypes
    lib.pom_implicit_step_reverse.argtypes = [vp, C.POINTER(Params), vp, vp, vp, vpp, vpp, vpp, C.c_int32, vpp,
                                              vpp, vp, C.c_size_t, vp]
    lib.pom_elliptic_apply.argtypes = [vp, C.POINTER(Params), vpp, vp, vp, vp, vp]
    lib.pom_elliptic_apply_adj.argtypes = [vp, C.POINTER(Params), vpp, vp, vp, vp, vp, vp, vp]
    lib.pom_eta_workspace.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.pom_eta_solve.argtypes = [vp, C.POINTER(Params), vp, vpp, vp, vp, vp, vp, vp, vp, C.c_size_t, vp]
    lib.pom_eta_solve_adj.argtypes = [vp, C.POINTER(Params), vp, vpp, vp, vp, vp, vp, vp, vp, vp, C.c_size_t, vp]
    lib.pom_eta_solve_noalloc.argtypes = [vp, C.POINTER(Params), vp, vpp, vp, vp, vp, vp, vp, vp]
    lib.pom_pack_input.argtypes = [vp, vpp, vpp, vpp, vp, vp]
    lib.pom_pack_input_adj.argtypes = [vp, vpp, vp, vpp, vp]
    lib.pom_restrict4.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pom_restrict4_adj.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pom_copy_arrays.argtypes = [C.c_int32, vpp, vpp, C.POINTER(C.c_int64), vp]
    lib.pom_bcond.argtypes = [vp, C.POINTER(Params), vp, vp, C.c_int32, vpp, vpp, vpp, vp]
    lib.pom_profq_workspace.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.pom_profq.argtypes = [vp, C.POINTER(Params), vp, vpp, vpp, vp, C.c_size_t, vp]
    lib.pom_profq_adj.argtypes = [vp, C.POINTER(Params), vp, vpp, vpp, vpp, vp, C.c_size_t, vp]
    lib.pom_proft.argtypes = [vp, C.POINTER(Params), vp, vpp, C.c_int32, vpp, vp]
    lib.pom_proft_adj.argtypes = [vp, C.POINTER(Params), vp, vpp, C.c_int32, vp, vpp, vp]
    lib.pom_bcond_adj.argtypes = [vp, C.POINTER(Params), vp, vp, C.c_int32, vpp, vpp, vpp, vpp, vp]
    dp = C.POINTER(C.c_double)
    lib.pom_correction_workspace.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.pom_joint_correction.argtypes = [vp, vp, vpp, vp, dp, C.c_double, vp, vpp, vpp, vp, vp]
    lib.pom_joint_correction_adj.argtypes = [vp, vp, vpp, vp, dp, C.c_double, vp, vpp, vp, vp, vp, vp]
    lib.pom_density75_workspace.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.pom_density75.argtypes = [vp, C.POINTER(Params), vp, vpp, dp, C.c_double, vp, vpp, vpp, vpp, vp, vp]
    lib.pom_profile_stage.argtypes = [C.c_int32]
    lib.pom_profile_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    for fn in EXPORTED:
        getattr(lib, fn).restype = C.c_int
    _lib = lib
    return lib


EXPORTED = ("pom_batch_create", "pom_batch_destroy", "pom_batch_tune", "pom_batch_get_tuning",
            "pom_explicit_workspace", "pom_explicit_step",
            "pom_explicit_step_adj", "pom_explicit_step_taped", "pom_explicit_step_reverse",
            "pom_implicit_cache_bytes", "pom_implicit_prepare", "pom_implicit_workspace",
            "pom_implicit_step", "pom_implicit_step_adj", "pom_implicit_step_taped", "pom_implicit_step_reverse", "pom_dens", "pom_dens_adj", "pom_baropg",
            "pom_baropg_adj", "pom_elliptic_apply", "pom_elliptic_apply_adj", "pom_eta_workspace", "pom_eta_solve",
            "pom_eta_solve_adj", "pom_eta_solve_noalloc", "pom_pack_input", "pom_pack_input_adj", "pom_restrict4", "pom_restrict4_adj", "pom_copy_arrays",
            "pom_bcond", "pom_bcond_adj", "pom_profq_workspace", "pom_profq", "pom_profq_adj", "pom_proft",
            "pom_proft_adj",
            "pom_correction_workspace", "pom_joint_correction", "pom_joint_correction_adj",
            "pom_density75_workspace", "pom_density75",
            "pom_profile_stage", "pom_profile_read")
STAGES = {"profq": 0, "profq_adj": 1, "extuv": 2, "advct": 3}


def profile_stage(name):
    """Start timing every launch of one stage (None stops)."""
    lib = load()
    check(lib.pom_profile_stage(-1 if name is None else STAGES[name]), "pom_profile_stage")


def profile_read():
    """(total kernel ms, launch count) of the selected stage since the last read."""
    lib = load()
    ms, n = C.c_double(), C.c_int32()
    check(lib.pom_profile_read(C.byref(ms), C.byref(n)), "pom_profile_read")
    return ms.value, n.value


def check(rc: int, what: str):
    if rc == POM_OK:
        return
    msg = {POM_E_ARG: "invalid argument", POM_E_HIP: "HIP runtime error",
           POM_E_UNSUPPORTED: "unsupported configuration",
           POM_E_SCRATCH: f"the batch needs more than the {NOALLOC_SCRATCH_BYTES >> 20} MiB static scratch "
                          "of this entry point"}.get(rc, f"error {rc}")
    if rc in (POM_E_ARG, POM_E_SCRATCH):
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def ptr_array(ptrs):
    arr = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def geom_ptrs(geom):
    """The pom_geom pointer table: every base entry from ``geom`` (required),
    the tide-policy entries when present (null otherwise)."""
    return ptr_array([geom[n].data_ptr() for n in GEOM] +
                     [geom[n].data_ptr() if geom.get(n) is not None else None for n in GEOM_ALL[len(GEOM):]])
