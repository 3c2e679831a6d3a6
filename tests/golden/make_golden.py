"""Writes the golden fixtures in this directory (run from the repository root after
`python -c "import __graft_entry__ as g; g.build()"`):

    python tests/golden/make_golden.py

The reference ships no tests, golden data or fixtures and cannot be built or run here, so these
fixtures come from the FP64 oracle (oracle/), whose arithmetic is pinned separately by the
closed-form and brute-force checks in tests/test_oracle.py.  They are regression anchors: each
holds a small deterministic case (parameters and every input array), the sorted keys, original
indices and cell starts of each species after the first cell hash, and the state after a fixed
number of full steps.  tests/test_oracle.py checks that the oracle still reproduces them;
tests/test_gpu_parity.py checks the HIP library against them.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

# name: (generator, options, steps)
FIXTURES = {
    "dam_liquid": ("dam_break", dict(tank_w=1.0, tank_h=0.8, col_w=0.4, col_h=0.6, dx=0.02), 50),
    "dam_two_phase": ("dam_break", dict(tank_w=1.0, tank_h=0.8, col_w=0.4, col_h=0.6, dx=0.04, air=1), 50),
    "gate_fsi": ("gate_fsi", dict(dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=4), 40),
}
FLUID_IN = ("x", "y", "u", "v", "rho", "p", "phase")
SOLID_IN = ("x", "y", "u", "v", "X", "Y", "material", "bc")
WALL_IN = ("x", "y", "bc")
FLUID_OUT = ("x", "y", "u", "v", "rho", "p")
SOLID_OUT = ("x", "y", "u", "v", "pressure", "stress")


def params_json(p) -> str:
    return json.dumps({name: getattr(p, name) for name, _ in p._fields_}, sort_keys=True)


def write(name: str) -> Path:
    from oracle import oracle_py
    from sphnj_amd import Case

    gen, opts, steps = FIXTURES[name]
    case = Case.generate(gen, **opts)
    out = {"params": np.array(params_json(case.params)), "steps": np.array(steps),
           "generator": np.array(json.dumps([gen, opts]))}
    for sp, keys in (("fluid", FLUID_IN), ("solid", SOLID_IN), ("wall", WALL_IN)):
        for k in keys:
            out[f"in.{sp}.{k}"] = np.asarray(getattr(case, sp)[k])
    ora = oracle_py.Oracle(case)
    ora.particle_cell_update()
    for s in (0, 1, 2):
        keys, order, starts = ora.sorted(0, s)
        out[f"sorted.{s}.keys"], out[f"sorted.{s}.order"], out[f"sorted.{s}.starts"] = keys, order, starts
    ora.close()
    ora = oracle_py.Oracle(case)
    ora.step(steps)
    if ora.status() != 0:
        raise RuntimeError(f"{name}: oracle status {ora.status()}")
    st = ora.download()
    for k in FLUID_OUT:
        out[f"out.fluid.{k}"] = np.asarray(st["fluid"][k])
    if len(case.solid["x"]):
        for k in SOLID_OUT:
            out[f"out.solid.{k}"] = np.asarray(st["solid"][k])
    ora.close()
    path = HERE / f"{name}.npz"
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    for n in FIXTURES:
        p = write(n)
        print(f"{p.name}: {p.stat().st_size} bytes")
