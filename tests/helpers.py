"""Shared builders for the tests: parameter sets and hand-made particle configurations."""
import numpy as np

from sphnj_amd import Case, default_params
from sphnj_amd.sim import grid


def params(**kw):
    """A valid parameter set over a 1 x 1 m box with h = 1.3 dx (dx = 0.02)."""
    p = default_params()
    p.particle_diameter = 0.02
    p.smoothing_length = 0.026
    p.dt = 1e-4
    p.iend = 10
    p.auxetic_modulus = 2e8
    p.cell_xmin, p.cell_ymin, p.cell_xmax, p.cell_ymax = -0.2, -0.2, 1.2, 1.2
    p.domain_xmin, p.domain_ymin, p.domain_xmax, p.domain_ymax = 0.0, 0.0, 1.0, 1.0
    p.liquid_sound_speed = 20.0
    p.gas_sound_speed = 20.0
    p.shift_per_step = 1
    p.particle_shift_coef = 1.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def fluid(x, y, u=None, v=None, rho=None, p=None, phase=None, rho0=1000.0):
    n = len(x)
    z = np.zeros(n)
    return {
        "x": np.asarray(x, float), "y": np.asarray(y, float),
        "u": z.copy() if u is None else np.broadcast_to(np.asarray(u, float), (n,)).copy(),
        "v": z.copy() if v is None else np.broadcast_to(np.asarray(v, float), (n,)).copy(),
        "rho": np.full(n, rho0) if rho is None else np.broadcast_to(np.asarray(rho, float), (n,)).copy(),
        "p": z.copy() if p is None else np.broadcast_to(np.asarray(p, float), (n,)).copy(),
        "phase": np.zeros(n, np.int32) if phase is None else np.broadcast_to(np.asarray(phase, np.int32), (n,)).copy(),
    }


def solid(x, y, X=None, Y=None, u=None, v=None, material=0, bc=0):
    n = len(x)
    z = np.zeros(n)
    return {
        "x": np.asarray(x, float), "y": np.asarray(y, float),
        "u": z.copy() if u is None else np.broadcast_to(np.asarray(u, float), (n,)).copy(),
        "v": z.copy() if v is None else np.broadcast_to(np.asarray(v, float), (n,)).copy(),
        "X": np.asarray(x if X is None else X, float).copy(), "Y": np.asarray(y if Y is None else Y, float).copy(),
        "material": np.broadcast_to(np.asarray(material, np.int32), (n,)).copy(),
        "bc": np.broadcast_to(np.asarray(bc, np.int32), (n,)).copy(),
    }


def wall(x, y, bc=0):
    n = len(x)
    return {"x": np.asarray(x, float), "y": np.asarray(y, float),
            "bc": np.broadcast_to(np.asarray(bc, np.int32), (n,)).copy()}


def empty_fluid():
    return fluid([], [])


def empty_solid():
    return solid([], [])


def empty_wall():
    return wall([], [])


def lattice(nx, ny, dx, x0=0.0, y0=0.0):
    """Cell-centre lattice ((i + 1/2) dx, (j + 1/2) dx), i outer."""
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    return (x0 + (i.ravel() + 0.5) * dx, y0 + (j.ravel() + 0.5) * dx)


def case(p, f=None, s=None, w=None):
    return Case(p, f if f is not None else empty_fluid(), s if s is not None else empty_solid(),
                w if w is not None else empty_wall())


def half_cell_keys(case, x, y):
    """Sub-cell keys (sy + sx 2 ny) of positions, by the hash's own arithmetic (axis_sub)."""
    nx, ny, cdx, cdy = grid(case.params)
    p = case.params

    def axis(v, v0, d, n):
        t = (np.asarray(v, np.float64) - v0) * (1.0 / d)
        c = np.trunc(t)
        h = np.clip(np.trunc(2.0 * t) - 2.0 * c, 0, 1)
        out = (2 * c + h).astype(np.int64)
        out[~(t > -1.0)] = 0
        out[~(t < n)] = 2 * n - 1
        return out

    return axis(y, p.cell_ymin, cdy, ny) + axis(x, p.cell_xmin, cdx, nx) * 2 * ny


GOLDEN = ("dam_liquid", "dam_two_phase", "gate_fsi")


def golden(name):
    """A fixture of tests/golden (written by tests/golden/make_golden.py): the Case rebuilt from
    its stored parameters and inputs, and the fixture's arrays."""
    import json
    from pathlib import Path

    import sphnj_amd._abi as A

    z = dict(np.load(Path(__file__).resolve().parent / "golden" / f"{name}.npz"))
    p = A.Params()
    for k, v in json.loads(str(z["params"])).items():
        setattr(p, k, v)
    sp = {s: {k.split(".", 2)[2]: a for k, a in z.items() if k.startswith(f"in.{s}.")}
          for s in ("fluid", "solid", "wall")}
    return Case(p, sp["fluid"], sp["solid"], sp["wall"]), z


def read_vtk(path):
    """A legacy-VTK POLYDATA file as written by the sphnj driver: (points (n, 3), {name: values})."""
    lines = open(path).read().split("\n")
    assert lines[0].startswith("# vtk DataFile") and lines[2] == "ASCII" and lines[3] == "DATASET POLYDATA"
    head = lines[4].split()
    assert head[0] == "POINTS"
    n = int(head[1])
    pts = np.array([list(map(float, ln.split())) for ln in lines[5:5 + n]]).reshape(n, 3)
    k = 5 + n
    assert lines[k] == f"POINT_DATA {n}"
    k += 1
    fields = {}
    while k < len(lines) and lines[k]:
        tag, name, typ, _ = lines[k].split()
        assert tag == "SCALARS" and lines[k + 1] == "LOOKUP_TABLE default"
        conv = float if typ == "double" else int
        fields[name] = np.array([conv(v) for v in lines[k + 2:k + 2 + n]])
        k += 2 + n
    return pts, fields
