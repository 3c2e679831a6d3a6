"""ctypes mirror of include/sphnj.h (structs, enums, function prototypes).

The same struct definitions describe the CPU oracle's API, which shares the POD parameter
and particle structs of the C ABI.
"""
from __future__ import annotations

import ctypes as C

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)

# ---- enums (include/sphnj.h) --------------------------------------------------------------
OK, E_ARG, E_PARAM, E_IO, E_COUNT, E_HIP, E_NODEVICE, E_STATE, E_NOMEM, E_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6, -7, -8, -9
PREC_P64, PREC_M32 = 0, 1
LIQUID, GAS = 0, 1
SLIP, NO_SLIP, MOVING = 0, 1, 2
SPLASH, FREE_SURFACE, VICINITY, INNER = 0, 1, 2, 3
NO_PARTICLE, HAS_PARTICLE, HAS_NEAR_PARTICLE = 0, 1, 2
LINEAR_ELASTIC, AUXETIC, NEO_HOOKEAN = 0, 1, 2
UNCONSTRAINED, FIXED = 0, 1
ST_OUT_OF_GRID, ST_CELL_OVER_64, ST_NAN, ST_REFLIST_LEFT, ST_SYNC_TIMEOUT = 1, 2, 4, 8, 16
PROBES = ("force_x", "force_y", "front", "height", "vm_elastic", "vm_auxetic", "vm_max", "mean_pressure")
STAGE_CELLS, STAGE_COUPLE, STAGE_WALL, STAGE_FLUID, STAGE_SOLID, STAGE_SHIFT = 1, 2, 3, 4, 5, 6
FLUID, SOLID, WALL = 0, 1, 2

STAGES = {
    "cells": STAGE_CELLS, "couple": STAGE_COUPLE, "wall": STAGE_WALL,
    "fluid": STAGE_FLUID, "solid": STAGE_SOLID, "shift": STAGE_SHIFT,
}


class Params(C.Structure):
    _fields_ = [
        ("gravity", C.c_double), ("is_steady", C.c_int32),
        ("particle_shift_coef", C.c_double), ("is_viscid", C.c_int32), ("min_pressure", C.c_double),
        ("liquid_sound_speed", C.c_double), ("liquid_density", C.c_double), ("liquid_gamma", C.c_double),
        ("liquid_viscosity", C.c_double),
        ("gas_sound_speed", C.c_double), ("gas_density", C.c_double), ("gas_gamma", C.c_double),
        ("gas_viscosity", C.c_double),
        ("solid_density", C.c_double), ("youngs_modulus", C.c_double), ("poisson_ratio", C.c_double),
        ("arti_vis_alpha", C.c_double), ("arti_vis_beta", C.c_double),
        ("solid_fixed_speed_x", C.c_double), ("solid_fixed_speed_y", C.c_double),
        ("is_2D_strain", C.c_int32), ("is_plastic", C.c_int32),
        ("auxetic_density", C.c_double), ("auxetic_modulus", C.c_double), ("auxetic_poisson_ratio", C.c_double),
        ("yield_strain", C.c_double), ("yield_stress", C.c_double), ("harden_strain", C.c_double),
        ("harden_coef", C.c_double),
        ("cell_xmin", C.c_double), ("cell_ymin", C.c_double), ("cell_xmax", C.c_double), ("cell_ymax", C.c_double),
        ("domain_xmin", C.c_double), ("domain_ymin", C.c_double), ("domain_xmax", C.c_double),
        ("domain_ymax", C.c_double),
        ("dt", C.c_double), ("istart", C.c_int32), ("iend", C.c_int32),
        ("solid_sub_step", C.c_int32), ("result_interval", C.c_int32), ("sort_every_step", C.c_int32),
        ("shift_per_step", C.c_int32),
        ("smoothing_length", C.c_double), ("particle_diameter", C.c_double),
        ("impact_length_hsml_ratio", C.c_double), ("thread_num", C.c_int32),
        ("background_pressure", C.c_double), ("kernel_2d_coef", C.c_double), ("adaptive_cfl", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p


class FluidIn(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", c_double_p), ("y", c_double_p), ("u", c_double_p), ("v", c_double_p),
                ("rho", c_double_p), ("p", c_double_p), ("phase", c_int32_p)]


class SolidIn(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", c_double_p), ("y", c_double_p), ("u", c_double_p), ("v", c_double_p),
                ("X", c_double_p), ("Y", c_double_p), ("material", c_int32_p), ("bc", c_int32_p)]


class WallIn(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", c_double_p), ("y", c_double_p), ("bc", c_int32_p)]


class FluidOut(C.Structure):
    _fields_ = [("x", c_double_p), ("y", c_double_p), ("u", c_double_p), ("v", c_double_p), ("rho", c_double_p),
                ("p", c_double_p), ("phase", c_int32_p), ("boundary_type", c_int32_p), ("near", c_int32_p),
                ("color", c_double_p), ("du_dx", c_double_p)]


class SolidOut(C.Structure):
    _fields_ = [("x", c_double_p), ("y", c_double_p), ("u", c_double_p), ("v", c_double_p), ("X", c_double_p),
                ("Y", c_double_p), ("pressure", c_double_p), ("normal", c_double_p),
                ("deformation", c_double_p), ("correction", c_double_p), ("stress", c_double_p),
                ("max_strain", c_double_p), ("du_dt", c_double_p), ("couple_dudt", c_double_p)]


class WallOut(C.Structure):
    _fields_ = [("x", c_double_p), ("y", c_double_p), ("u", c_double_p), ("v", c_double_p), ("p", c_double_p),
                ("rho", c_double_p), ("normal", c_double_p), ("phase", c_int32_p), ("bc", c_int32_p)]


# Fields of each output struct with their per-particle width and numpy dtype.
FLUID_FIELDS = [("x", 1, "f8"), ("y", 1, "f8"), ("u", 1, "f8"), ("v", 1, "f8"), ("rho", 1, "f8"), ("p", 1, "f8"),
                ("phase", 1, "i4"), ("boundary_type", 1, "i4"), ("near", 1, "i4"), ("color", 1, "f8"),
                ("du_dx", 4, "f8")]
SOLID_FIELDS = [("x", 1, "f8"), ("y", 1, "f8"), ("u", 1, "f8"), ("v", 1, "f8"), ("X", 1, "f8"), ("Y", 1, "f8"),
                ("pressure", 1, "f8"), ("normal", 2, "f8"), ("deformation", 4, "f8"), ("correction", 4, "f8"),
                ("stress", 4, "f8"), ("max_strain", 4, "f8"), ("du_dt", 2, "f8"), ("couple_dudt", 2, "f8")]
WALL_FIELDS = [("x", 1, "f8"), ("y", 1, "f8"), ("u", 1, "f8"), ("v", 1, "f8"), ("p", 1, "f8"), ("rho", 1, "f8"),
               ("normal", 2, "f8"), ("phase", 1, "i4"), ("bc", 1, "i4")]

VP = C.c_void_p
P = C.POINTER

# Context-level functions shared by libsphnj (prefix "sphnj_") and the oracle ("oracle_").
CTX_PROTOS = {
    "upload": (C.c_int, [VP, C.c_int32, P(FluidIn), P(SolidIn), P(WallIn)]),
    "init": (C.c_int, [VP]),
    "step": (C.c_int, [VP, C.c_int32]),
    "stage": (C.c_int, [VP, C.c_int32]),
    "download": (C.c_int, [VP, C.c_int32, P(FluidOut), P(SolidOut), P(WallOut)]),
    "status": (C.c_int, [VP, P(C.c_uint32)]),
    "stage_sorted": (C.c_int, [VP, C.c_int32, C.c_int32, c_int32_p, c_int32_p, c_int32_p]),
    "stage_neighbors": (C.c_int, [VP, C.c_int32, c_int32_p, c_int32_p, C.c_int64, P(C.c_int64)]),
    "solid_reflist": (C.c_int, [VP, C.c_int32, c_int32_p, c_int32_p, C.c_int64]),
    "energy": (C.c_int, [VP, C.c_int32, c_double_p]),
    "probe_global": (C.c_int, [VP, C.c_int32, C.c_double, c_double_p]),
    "time": (C.c_int, [VP, C.c_int32, c_double_p, c_double_p]),
    "probe_solid": (C.c_int, [VP, C.c_int32, C.c_double, C.c_double, c_double_p, c_int32_p]),
    "destroy": (None, [VP]),
}

LIB_PROTOS = {
    "sphnj_abi_version": (C.c_int, []),
    "sphnj_global_error": (C.c_char_p, []),
    "sphnj_params_default": (C.c_int, [P(Params)]),
    "sphnj_params_from_ini": (C.c_int, [C.c_char_p, P(Params)]),
    "sphnj_params_validate": (C.c_int, [P(Params)]),
    "sphnj_params_grid": (C.c_int, [P(Params), c_int32_p, c_int32_p, c_double_p, c_double_p]),
    "sphnj_particles_read": (C.c_int, [C.c_char_p, P(VP)]),
    "sphnj_particles_write": (C.c_int, [VP, C.c_char_p]),
    "sphnj_particles_views": (C.c_int, [VP, P(FluidIn), P(SolidIn), P(WallIn), c_int32_p]),
    "sphnj_particles_destroy": (None, [VP]),
    "sphnj_case_create": (C.c_int, [C.c_char_p, C.c_char_p, P(Params), P(VP)]),
    "sphnj_case_write": (C.c_int, [P(Params), VP, C.c_char_p]),
    "sphnj_create": (C.c_int, [P(Params), C.c_int, P(VP)]),
    "sphnj_batch_create": (C.c_int, [P(Params), C.c_int32, C.c_int, P(VP)]),
    "sphnj_sync": (C.c_int, [VP]),
    "sphnj_counts": (C.c_int, [VP, C.c_int32, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
    "sphnj_last_error": (C.c_char_p, [VP]),
    "sphnj_set_graph": (C.c_int, [VP, C.c_int]),
    "sphnj_set_persistent_solid": (C.c_int, [VP, C.c_int]),
    "sphnj_set_shift_runs": (C.c_int, [VP, C.c_int]),
    "sphnj_energy_noalloc": (C.c_int, [VP, C.c_int32, c_double_p]),
    "sphnj_snapshot_write": (C.c_int, [VP, C.c_int32, C.c_int64, C.c_char_p]),
    "sphnj_snapshot_info": (C.c_int, [C.c_char_p, P(Params), P(C.c_int64), c_int32_p]),
    "sphnj_snapshot_load": (C.c_int, [VP, C.c_int32, C.c_char_p]),
    "sphnj_autotune": (C.c_int, [VP]),
    "sphnj_autotune_report": (C.c_int, [VP, C.c_int32, c_double_p]),
    "sphnj_get_block": (C.c_int, [VP, C.c_int32, c_int32_p]),
    "sphnj_set_block": (C.c_int, [VP, C.c_int32, C.c_int32]),
    "sphnj_last_step_ms": (C.c_int, [VP, c_double_p]),
    "sphnj_set_stage_timing": (C.c_int, [VP, C.c_int]),
    "sphnj_stage_time": (C.c_int, [VP, C.c_int32, c_int32_p, c_double_p]),
}
for _k, _v in CTX_PROTOS.items():
    LIB_PROTOS["sphnj_" + _k] = _v


def bind(lib: C.CDLL, protos: dict) -> None:
    for name, (res, args) in protos.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
