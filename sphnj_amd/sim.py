"""Host-side Python mirror of SPHNJ's operator API over libsphnj.

The reference drives one simulation through the static class DeviceCalculation
(include/calculation.cuh:9-23): particleCellUpdate / sortParticle / fillCell,
coupleCalculation, virtUpdate, fluidCalculation, solidCalculation, fluidCorrection, called in
that order by DeviceCalculation::calculation (src/calculation.cu:20-48).  `Simulation` exposes
the same phases (snake_case) on an explicit context, plus whole steps, downloads in input
order, and the inspection calls used by the parity tests.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Iterable

import numpy as np

from . import _abi as A

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libsphnj.so"


class SphnjError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sphnj error {code}: {msg}")
        self.code = code


def library() -> C.CDLL:
    """Load libsphnj.so (built in-tree).  There is no fallback: without the HIP library the
    package cannot run."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sphnj_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        A.bind(lib, A.LIB_PROTOS)
        _LIB = lib
    return _LIB


def _check(rc: int, where: str, msg: str | None = None) -> None:
    if rc != 0:
        if msg is None:
            msg = library().sphnj_global_error().decode()
        raise SphnjError(rc, f"{where}: {msg}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(A.c_double_p)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(A.c_int32_p)


def default_params() -> A.Params:
    p = A.Params()
    _check(library().sphnj_params_default(C.byref(p)), "params_default")
    return p


def params_from_ini(path: str | Path) -> A.Params:
    p = A.Params()
    _check(library().sphnj_params_from_ini(str(path).encode(), C.byref(p)), "params_from_ini")
    return p


def validate_params(p: A.Params) -> int:
    return library().sphnj_params_validate(C.byref(p))


def grid(p: A.Params) -> tuple[int, int, float, float]:
    nx, ny = C.c_int32(), C.c_int32()
    dx, dy = C.c_double(), C.c_double()
    _check(library().sphnj_params_grid(C.byref(p), C.byref(nx), C.byref(ny), C.byref(dx), C.byref(dy)), "grid")
    return nx.value, ny.value, dx.value, dy.value


def snapshot_info(path: str | Path):
    """(params, step, (n_fluid, n_solid, n_wall)) of a snapshot file."""
    p = A.Params()
    step = C.c_int64()
    n = (C.c_int32 * 3)()
    _check(library().sphnj_snapshot_info(str(path).encode(), C.byref(p), C.byref(step), n), "snapshot_info")
    return p, step.value, tuple(n)


class Case:
    """One case: a parameter set plus its particles (the contents of config/input.ini and
    data/*.dat), as numpy arrays in input (original index) order."""

    def __init__(self, params: A.Params, fluid: dict, solid: dict, wall: dict, n_liquid: int | None = None):
        self.params = params
        self.fluid = {k: np.ascontiguousarray(v) for k, v in fluid.items()}
        self.solid = {k: np.ascontiguousarray(v) for k, v in solid.items()}
        self.wall = {k: np.ascontiguousarray(v) for k, v in wall.items()}
        nf = len(self.fluid.get("x", []))
        self.n_liquid = int((self.fluid["phase"] == A.LIQUID).sum()) if n_liquid is None and nf else (n_liquid or 0)

    @property
    def counts(self) -> tuple[int, int, int]:
        return len(self.fluid["x"]), len(self.solid["x"]), len(self.wall["x"])

    @staticmethod
    def _from_handle(params: A.Params, h: C.c_void_p) -> "Case":
        lib = library()
        f, s, w = A.FluidIn(), A.SolidIn(), A.WallIn()
        nl = C.c_int32()
        _check(lib.sphnj_particles_views(h, C.byref(f), C.byref(s), C.byref(w), C.byref(nl)), "particles_views")

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        fluid = {k: arr(getattr(f, k), f.n, "f8") for k in ("x", "y", "u", "v", "rho", "p")}
        fluid["phase"] = arr(f.phase, f.n, "i4")
        solid = {k: arr(getattr(s, k), s.n, "f8") for k in ("x", "y", "u", "v", "X", "Y")}
        solid["material"] = arr(s.material, s.n, "i4")
        solid["bc"] = arr(s.bc, s.n, "i4")
        wall = {k: arr(getattr(w, k), w.n, "f8") for k in ("x", "y")}
        wall["bc"] = arr(w.bc, w.n, "i4")
        lib.sphnj_particles_destroy(h)
        return Case(params, fluid, solid, wall, nl.value)

    @staticmethod
    def generate(name: str, **options) -> "Case":
        """A case from the built-in generators ("dam_break", "gate_fsi", "wedge_entry",
        "ensemble_member"); keyword options become "key=value" pairs."""
        opts = ";".join(f"{k}={v}" for k, v in options.items())
        p = A.Params()
        h = C.c_void_p()
        _check(library().sphnj_case_create(name.encode(), opts.encode(), C.byref(p), C.byref(h)), "case_create")
        return Case._from_handle(p, h)

    @staticmethod
    def read(case_dir: str | Path) -> "Case":
        """config/input.ini + data/*.dat of a case directory (the reference's surfaces)."""
        case_dir = Path(case_dir)
        p = params_from_ini(case_dir / "config" / "input.ini")
        h = C.c_void_p()
        _check(library().sphnj_particles_read(str(case_dir / "data").encode(), C.byref(h)), "particles_read")
        return Case._from_handle(p, h)

    def structs(self):
        f, s, w = self.fluid, self.solid, self.wall
        fi = A.FluidIn(len(f["x"]), *(_dptr(f[k]) for k in ("x", "y", "u", "v", "rho", "p")), _iptr(f["phase"]))
        si = A.SolidIn(len(s["x"]), *(_dptr(s[k]) for k in ("x", "y", "u", "v", "X", "Y")), _iptr(s["material"]),
                       _iptr(s["bc"]))
        wi = A.WallIn(len(w["x"]), _dptr(w["x"]), _dptr(w["y"]), _iptr(w["bc"]))
        return fi, si, wi


class ContextBase:
    """Shared driver of a context-style C API (libsphnj's sphnj_* or the oracle's oracle_*)."""

    prefix = "sphnj_"

    def __init__(self, lib: C.CDLL, handle: C.c_void_p, cases: list[Case]):
        self._lib = lib
        self._h = handle
        self.cases = cases

    def _f(self, name):
        return getattr(self._lib, self.prefix + name)

    def _err(self) -> str:
        return ""

    def _call(self, name: str, *args) -> int:
        rc = self._f(name)(self._h, *args)
        if rc != 0:
            raise SphnjError(rc, f"{self.prefix}{name}: {self._err()}")
        return rc

    def close(self) -> None:
        if self._h:
            self._f("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _upload_all(self) -> None:
        for i, cs in enumerate(self.cases):
            fi, si, wi = cs.structs()
            self._call("upload", i, C.byref(fi), C.byref(si), C.byref(wi))
        self._call("init")

    # ---- stepping -----------------------------------------------------------------------
    def step(self, n: int = 1) -> None:
        self._call("step", int(n))

    def stage(self, name: str) -> None:
        self._call("stage", A.STAGES[name])

    # ---- results ------------------------------------------------------------------------
    def download(self, case: int = 0, species: Iterable[str] = ("fluid", "solid", "wall")) -> dict:
        nf, ns, nw = self.cases[case].counts
        out = {}
        structs = {}
        for sp, fields, n, st in (("fluid", A.FLUID_FIELDS, nf, A.FluidOut), ("solid", A.SOLID_FIELDS, ns, A.SolidOut),
                                  ("wall", A.WALL_FIELDS, nw, A.WallOut)):
            if sp not in species:
                structs[sp] = None
                continue
            arrs = {k: np.zeros((n, w) if w > 1 else n, dtype=dt) for k, w, dt in fields}
            s = st(*[(_dptr(arrs[k]) if dt == "f8" else _iptr(arrs[k])) for k, _, dt in fields])
            out[sp] = arrs
            structs[sp] = s
        ptr = lambda s: C.byref(s) if s is not None else None  # noqa: E731
        self._call("download", case, ptr(structs["fluid"]), ptr(structs["solid"]), ptr(structs["wall"]))
        return out

    def status(self) -> int:
        v = C.c_uint32()
        self._call("status", C.byref(v))
        return v.value

    def sorted(self, case: int, species: int):
        """(keys, original index per sorted slot, start table) of the last CELLS stage.

        Solids and walls are binned by cell (n_cells + 1 starts); fluids by half-cell
        sub-cell, x-major (4 n_cells + 1 starts)."""
        n = self.cases[case].counts[species]
        nx, ny, _, _ = grid(self.cases[case].params)
        keys = np.zeros(n, np.int32)
        orig = np.zeros(n, np.int32)
        start = np.zeros((4 if species == 0 else 1) * nx * ny + 1, np.int32)
        self._call("stage_sorted", case, species, _iptr(keys), _iptr(orig), _iptr(start))
        return keys, orig, start

    def neighbors(self, case: int = 0):
        """CSR of in-range neighbours per fluid particle (input order): (counts, entries) with
        entries = species << 28 | original index, in iteration order."""
        n = self.cases[case].counts[0]
        counts = np.zeros(n, np.int32)
        total = C.c_int64()
        self._call("stage_neighbors", case, _iptr(counts), None, 0, C.byref(total))
        lst = np.zeros(max(total.value, 1), np.int32)
        self._call("stage_neighbors", case, _iptr(counts), _iptr(lst), total.value, C.byref(total))
        return counts, lst[: total.value]

    def reflist(self, case: int = 0):
        ns = self.cases[case].counts[1]
        off = np.zeros(ns + 1, np.int32)
        self._call("solid_reflist", case, _iptr(off), None, 0)
        lst = np.zeros(max(int(off[-1]), 1), np.int32)
        self._call("solid_reflist", case, _iptr(off), _iptr(lst), int(off[-1]))
        return off, lst[: int(off[-1])]

    def energy(self, case: int = 0) -> np.ndarray:
        out = np.zeros(6)
        self._call("energy", case, _dptr(out))
        return out

    def time(self, case: int = 0) -> tuple[float, float]:
        """(simulated time, size of the last step) of a case."""
        t, dt = C.c_double(), C.c_double()
        self._call("time", case, C.byref(t), C.byref(dt))
        return t.value, dt.value

    def probe_global(self, case: int = 0, height_x: float = 0.0) -> dict:
        """The reference's global outputs (src/output.cu:283-560) of the current state:
        force_x/y, front, height (at x = height_x), vm_elastic, vm_auxetic, vm_max, mean_pressure."""
        out = np.zeros(len(A.PROBES))
        self._call("probe_global", case, float(height_x), _dptr(out))
        return dict(zip(A.PROBES, out.tolist()))

    def probe_solid(self, X: float, Y: float, case: int = 0) -> tuple[np.ndarray, int]:
        """Current position of the solid particle whose reference position lies within d/2 of
        (X, Y), and its input index."""
        pos = np.zeros(2)
        idx = C.c_int32()
        self._call("probe_solid", case, float(X), float(Y), _dptr(pos), C.byref(idx))
        return pos, idx.value


class Simulation(ContextBase):
    """One libsphnj context over one case (the drop-in) or a batch of cases."""

    prefix = "sphnj_"

    def __init__(self, cases: Case | list[Case], precision: str = "p64", graph: bool = True,
                 persistent_solid: bool = False, shift_runs: int = 1):
        lib = library()
        cases = [cases] if isinstance(cases, Case) else list(cases)
        prec = {"p64": A.PREC_P64, "m32": A.PREC_M32}[precision]
        arr = (A.Params * len(cases))(*[c.params for c in cases])
        h = C.c_void_p()
        _check(lib.sphnj_batch_create(arr, len(cases), prec, C.byref(h)), "batch_create")
        super().__init__(lib, h, cases)
        self.precision = precision
        _check(lib.sphnj_set_graph(h, int(graph)), "set_graph")
        _check(lib.sphnj_set_persistent_solid(h, int(persistent_solid)), "set_persistent_solid")
        _check(lib.sphnj_set_shift_runs(h, int(shift_runs)), "set_shift_runs")
        self._upload_all()

    def _err(self) -> str:
        return self._lib.sphnj_last_error(self._h).decode()

    def sync(self) -> None:
        self._call("sync")

    def last_step_ms(self) -> float:
        v = C.c_double()
        self._call("last_step_ms", C.byref(v))
        return v.value

    def set_block(self, family: int, block: int) -> None:
        self._call("set_block", family, block)

    def autotune(self) -> dict:
        """Pick the block sizes of the pair and shift kernels; returns ms per launch measured
        for block sizes 64, 128 and 256, and the choice."""
        self._call("autotune")
        out = {}
        for fam, name in ((0, "pair"), (1, "shift")):
            ms = np.zeros(3)
            self._call("autotune_report", fam, _dptr(ms))
            blk = C.c_int32()
            self._call("get_block", fam, C.byref(blk))
            out[name] = {"ms": dict(zip((64, 128, 256), ms.tolist())), "block": blk.value}
        return out

    def snapshot(self, path: str | Path, step: int = 0, case: int = 0) -> None:
        """Binary restart file of one case (sphnj_snapshot_write)."""
        self._call("snapshot_write", case, int(step), str(path).encode())

    @classmethod
    def restore(cls, paths, precision: str = "p64", graph: bool = True, persistent_solid: bool = False):
        """A context whose cases come from snapshots (sphnj_snapshot_load in place of uploads)."""
        paths = [paths] if isinstance(paths, (str, Path)) else list(paths)
        infos = [snapshot_info(p) for p in paths]
        lib = library()
        prec = {"p64": A.PREC_P64, "m32": A.PREC_M32}[precision]
        arr = (A.Params * len(paths))(*[i[0] for i in infos])
        h = C.c_void_p()
        _check(lib.sphnj_batch_create(arr, len(paths), prec, C.byref(h)), "batch_create")
        self = cls.__new__(cls)
        z = lambda n: np.zeros(n)  # noqa: E731
        cases = [Case(p, {"x": z(nf)}, {"x": z(ns)}, {"x": z(nw)}, 0) for p, _, (nf, ns, nw) in infos]
        ContextBase.__init__(self, lib, h, cases)
        self.precision = precision
        _check(lib.sphnj_set_graph(h, int(graph)), "set_graph")
        _check(lib.sphnj_set_persistent_solid(h, int(persistent_solid)), "set_persistent_solid")
        for i, p in enumerate(paths):
            self._call("snapshot_load", i, str(p).encode())
        self._call("init")
        return self

    def set_stage_timing(self, on: bool) -> None:
        self._call("set_stage_timing", int(on))

    def set_shift_runs(self, mode: int) -> None:
        """Shift passes: 1 walk the candidate runs recorded by the pair kernel (same block sizes)
        or by the first pass (default), 0 search the full windows, 2 record in the first pass and
        never walk them (sphnj_set_shift_runs)."""
        self._call("set_shift_runs", int(mode))

    def stage_time(self, name: str) -> tuple[int, float]:
        """(kernel launches, total ms) of a stage since timing was enabled."""
        n = C.c_int32()
        ms = C.c_double()
        self._call("stage_time", A.STAGES[name], C.byref(n), C.byref(ms))
        return n.value, ms.value

    # ---- the reference's phase names (DeviceCalculation, include/calculation.cuh:9-23) ----
    def particle_cell_update(self) -> None:
        """particleCellUpdate + sortParticle + fillCell: hash, sort, cell ranges."""
        self.stage("cells")

    def couple_calculation(self) -> None:
        self.stage("couple")

    def virt_update(self) -> None:
        self.stage("wall")

    def fluid_calculation(self) -> None:
        self.stage("fluid")

    def solid_calculation(self) -> None:
        self.stage("solid")

    def fluid_correction(self) -> None:
        self.stage("shift")

    def calculation(self, nsteps: int) -> None:
        """nsteps iterations of the step loop (output excluded)."""
        self.step(nsteps)
