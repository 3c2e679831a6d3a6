"""TEST INFRASTRUCTURE ONLY -- Python driver of the FP64 CPU oracle (liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker / the
timed CPU baseline.  The oracle API mirrors libsphnj's, so `Oracle` offers the same methods as
`sphnj_amd.Simulation` (step, stage, download, sorted, neighbors, reflist, energy).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from sphnj_amd import _abi as A
from sphnj_amd.sim import Case, ContextBase, SphnjError

ORACLE_LIB = Path(__file__).resolve().parent / "liboracle.so"
_OLIB = None


def oracle_library() -> C.CDLL:
    global _OLIB
    if _OLIB is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} is missing: build it with `python -m sphnj_amd.build --oracle`")
        lib = C.CDLL(str(ORACLE_LIB))
        protos = {"oracle_" + k: v for k, v in A.CTX_PROTOS.items()}
        protos.update({
            "oracle_create": (C.c_int, [C.POINTER(A.Params), C.c_int32, C.POINTER(C.c_void_p)]),
            "oracle_last_error": (C.c_char_p, []),
            "oracle_set_threads": (C.c_int, [C.c_int32]),
            "oracle_derived": (C.c_int, [C.c_void_p, A.c_double_p]),
        })
        A.bind(lib, protos)
        _OLIB = lib
    return _OLIB


def set_threads(n: int) -> int:
    return oracle_library().oracle_set_threads(int(n))


class Oracle(ContextBase):
    prefix = "oracle_"

    def __init__(self, cases: Case | list[Case]):
        lib = oracle_library()
        cases = [cases] if isinstance(cases, Case) else list(cases)
        arr = (A.Params * len(cases))(*[c.params for c in cases])
        h = C.c_void_p()
        rc = lib.oracle_create(arr, len(cases), C.byref(h))
        if rc:
            raise SphnjError(rc, "oracle_create: " + lib.oracle_last_error().decode())
        super().__init__(lib, h, cases)
        self._upload_all()

    def _err(self) -> str:
        return self._lib.oracle_last_error().decode()

    def derived(self):
        import numpy as np

        out = np.zeros(15)
        self._call("derived", out.ctypes.data_as(A.c_double_p))
        keys = ["rc2", "hinv2", "kdc", "k2c", "eR", "nx", "ny", "cdx", "cdy", "lambda", "mu_s", "c_s",
                "yield_strain", "aux_modulus", "umax_cap"]
        return dict(zip(keys, out))

    # the reference's phase names, as on Simulation
    def particle_cell_update(self):
        self.stage("cells")

    def couple_calculation(self):
        self.stage("couple")

    def virt_update(self):
        self.stage("wall")

    def fluid_calculation(self):
        self.stage("fluid")

    def solid_calculation(self):
        self.stage("solid")

    def fluid_correction(self):
        self.stage("shift")
