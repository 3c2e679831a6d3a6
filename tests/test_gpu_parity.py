"""GPU parity: libsphnj (HIP, through the C ABI) against the FP64 CPU oracle on the same inputs.

Bars: bit-exact for sorted keys, original-index order, cell ranges, neighbour lists and solid
reference lists; relative L2 per field for the floating-point state (P64: 1e-9 after the step
counts below, where the two sides differ only by rounding; M32: 1e-4, the precision target).
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import half_cell_keys
from sphnj_amd import Case, Simulation

pytestmark = pytest.mark.gpu

FLUID_FIELDS = ("x", "y", "u", "v", "rho", "p")
SOLID_FIELDS = ("x", "y", "u", "v", "pressure", "deformation", "stress")


def small_dam(**kw):
    opts = dict(tank_w=1.0, tank_h=0.8, col_w=0.4, col_h=0.6, dx=0.02, steps=100)
    opts.update(kw)
    return Case.generate("dam_break", **opts)


def small_gate(**kw):
    opts = dict(dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=4)
    opts.update(kw)
    return Case.generate("gate_fsi", **opts)


def compare_state(a, b, tol, species=("fluid",), solid_fields=SOLID_FIELDS, fluid_fields=FLUID_FIELDS):
    errs = {}
    for f in fluid_fields:
        if "fluid" in species:
            errs["fluid." + f] = rel_l2(a["fluid"][f], b["fluid"][f])
    if "solid" in species:
        for f in solid_fields:
            errs["solid." + f] = rel_l2(a["solid"][f], b["solid"][f])
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"fields over tolerance {tol}: {bad}"
    return errs


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_cells_bit_exact(gpu, oracle_mod, precision):
    case = small_gate()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.particle_cell_update()
    ora.particle_cell_update()
    for species in (0, 1, 2):
        ks, os_, ss = sim.sorted(0, species)
        ko, oo, so = ora.sorted(0, species)
        assert np.array_equal(ks, ko), f"species {species}: keys differ"
        assert np.array_equal(os_, oo), f"species {species}: order differs"
        assert np.array_equal(ss, so), f"species {species}: cell starts differ"
        assert np.all(np.diff(ks) >= 0)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_sort_after_motion(gpu, precision):
    """After particles have crossed sub-cells, the CELLS stage's order (a counting sort that
    starts from last step's order) is still the (sub-column, sub-row, original index) order of
    the current positions, with consistent starts."""
    case = small_dam()
    sim = Simulation(case, precision)
    k0 = half_cell_keys(case, case.fluid["x"], case.fluid["y"])
    sim.step(300)
    sim.particle_cell_update()
    ks, os_, ss = sim.sorted(0, 0)
    st = sim.download()["fluid"]
    key = half_cell_keys(case, st["x"], st["y"])
    assert np.count_nonzero(key != k0) > 20, "the case must move particles across sub-cells"
    assert np.array_equal(np.sort(os_), np.arange(len(os_)))
    assert np.array_equal(ks, key[os_])
    assert np.all(np.diff(ks) >= 0)
    same = np.diff(ks) == 0
    assert np.all(np.diff(os_)[same] > 0), "original-index order inside a sub-cell"
    slots = np.arange(len(ks))
    assert np.all(ss[ks] <= slots) and np.all(slots < ss[ks + 1])


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_solid_reflist_bit_exact(gpu, oracle_mod, precision):
    case = small_gate()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    o1, l1 = sim.reflist()
    o2, l2 = ora.reflist()
    assert np.array_equal(o1, o2)
    assert np.array_equal(l1, l2)


def test_neighbors_bit_exact_p64(gpu, oracle_mod):
    case = small_gate()
    sim = Simulation(case, "p64")
    ora = oracle_mod.Oracle(case)
    sim.particle_cell_update()
    ora.particle_cell_update()
    c1, l1 = sim.neighbors()
    c2, l2 = ora.neighbors()
    assert np.array_equal(c1, c2)
    assert np.array_equal(l1, l2)


@pytest.mark.parametrize("stage", ["cells", "couple", "wall", "fluid", "solid", "shift"])
def test_stage_by_stage_p64(gpu, oracle_mod, stage):
    """Run the step stage by stage on both sides; after each stage the state agrees.

    State fields to 1e-9: the two sides evaluate the same FP64 sums in the same order but with
    different (algebraically equal) forms -- e.g. the library forms F - I from displacements
    where the oracle forms F~ L and cancels -- so they differ at the rounding level (seen:
    <= 1e-10 on solid stress).  Colour, wall fields to 1e-12; flags bit-exact."""
    case = small_gate()
    sim = Simulation(case, "p64")
    ora = oracle_mod.Oracle(case)
    order = ["cells", "couple", "wall", "fluid", "solid", "shift"]
    for st in order[: order.index(stage) + 1]:
        sim.stage(st)
        ora.stage(st)
    a, b = sim.download(), ora.download()
    compare_state(a, b, 1e-9, species=("fluid", "solid"))
    for f in ("u", "v", "p", "rho"):
        assert rel_l2(a["wall"][f], b["wall"][f]) <= 1e-12, f
    assert np.array_equal(a["wall"]["phase"], b["wall"]["phase"])
    if order.index(stage) >= order.index("fluid"):
        assert rel_l2(a["fluid"]["color"], b["fluid"]["color"]) <= 1e-12
        assert np.array_equal(a["fluid"]["near"], b["fluid"]["near"])
        assert rel_l2(a["fluid"]["du_dx"], b["fluid"]["du_dx"]) <= 1e-10
    if stage == "shift":
        assert np.array_equal(a["fluid"]["boundary_type"], b["fluid"]["boundary_type"])


@pytest.mark.parametrize("precision,tol", [("p64", 1e-9), ("m32", 1e-4)])
def test_dam_break_steps(gpu, oracle_mod, precision, tol):
    case = small_dam()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.step(100)
    ora.step(100)
    assert sim.status() == ora.status() == 0
    compare_state(sim.download(), ora.download(), tol)


@pytest.mark.parametrize("precision,tol", [("p64", 1e-9), ("m32", 1e-4)])
def test_gate_fsi_steps(gpu, oracle_mod, precision, tol):
    case = small_gate()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.step(40)
    ora.step(40)
    compare_state(sim.download(), ora.download(), tol, species=("fluid", "solid"),
                  solid_fields=("x", "y", "u", "v", "pressure"))


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_graph_equals_eager(gpu, precision):
    """Replays of the captured step, with eager work and downloads between them, match the
    eager step bit for bit."""
    case = small_gate()
    a = Simulation(case, precision, graph=True)
    b = Simulation(case, precision, graph=False)
    for n in (1, 2, 3, 1):
        a.step(n)
        b.step(n)
        da, db = a.download(), b.download()
        for sp in ("fluid", "solid", "wall"):
            for k in da[sp]:
                assert np.array_equal(da[sp][k], db[sp][k]), (sp, k, n)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_batch_equals_single(gpu, precision):
    """A batch of uneven cases (two dam breaks and an FSI gate) steps each case bit for bit as
    it steps alone."""
    cases = [small_dam(), small_dam(col_w=0.3, dx=0.025), small_gate()]
    batch = Simulation(cases, precision)
    batch.step(5)
    for i, c in enumerate(cases):
        one = Simulation(c, precision)
        one.step(5)
        a, b = batch.download(i), one.download()
        for sp in ("fluid", "solid", "wall"):
            for k in a[sp]:
                assert np.array_equal(a[sp][k], b[sp][k]), (i, sp, k)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_batch_equals_single_past_one_block_scan(gpu, precision):
    """The first eight seeded ensemble members (5 k to 184 k fluid each) batched: together they
    have more than 4096 sub-columns, so the batch plans its tiles through the separate tile
    scan across case boundaries, while each member alone stays under that size; every member
    still steps bit for bit as it steps alone."""
    from sphnj_amd.sim import grid

    cases = [Case.generate("ensemble_member", index=i) for i in range(8)]
    cols = [2 * grid(c.params)[0] for c in cases]
    assert sum(cols) > 4096 and max(cols) <= 4096, cols
    batch = Simulation(cases, precision)
    batch.step(3)
    for i, c in enumerate(cases):
        one = Simulation(c, precision)
        one.step(3)
        a, b = batch.download(i), one.download()
        one.close()
        for k in a["fluid"]:
            assert np.array_equal(a["fluid"][k], b["fluid"][k]), (i, k)


@pytest.mark.parametrize("precision,tol", [("p64", 1e-9), ("m32", 1e-4)])
@pytest.mark.parametrize("name", ["dam_liquid", "dam_two_phase", "gate_fsi"])
def test_matches_golden(gpu, name, precision, tol):
    """The library against the committed fixtures (tests/golden): the first hash's keys,
    original-index order and cell starts bit for bit, then the state after the fixture's steps
    to the precision's bar (M32: the solid stress is left to the P64 check)."""
    from helpers import golden

    c, z = golden(name)
    sim = Simulation(c, precision)
    sim.particle_cell_update()
    for s in (0, 1, 2):
        for part, a in zip(("keys", "order", "starts"), sim.sorted(0, s)):
            assert np.array_equal(a, z[f"sorted.{s}.{part}"]), (s, part)
    sim.close()
    sim = Simulation(c, precision)
    sim.step(int(z["steps"]))
    assert sim.status() == 0
    st = sim.download()
    errs = {}
    for k, ref in z.items():
        if k.startswith("out.") and not (precision == "m32" and k == "out.solid.stress"):
            _, sp, f = k.split(".")
            errs[k] = rel_l2(st[sp][f], ref)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"fields over tolerance {tol}: {bad}"


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_cells_bit_exact_multi_block_sort(gpu, oracle_mod, precision):
    """28k wall particles: more than one block's worth (16,384) of the small sort, so the walls
    go through the multi-block radix passes; keys, order and cell starts of every species match
    the oracle bit for bit, and the fields after 100 steps meet the precision's bar."""
    case = Case.generate("dam_break", tank_w=4.0, tank_h=3.0, col_w=0.2, col_h=0.2, dx=0.002)
    assert case.counts[2] > 16384
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.particle_cell_update()
    ora.particle_cell_update()
    for species in (0, 2):
        for part, a, b in zip(("keys", "order", "starts"), sim.sorted(0, species), ora.sorted(0, species)):
            assert np.array_equal(a, b), (species, part)
    sim.close()
    ora.close()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.step(100)
    ora.step(100)
    compare_state(sim.download(), ora.download(), 1e-9 if precision == "p64" else 1e-4)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_large_grid_scan_paths(gpu, oracle_mod, precision):
    """A 17 m x 0.05 m water layer (212k fluid, 104k walls) in a 2205 x 1179-cell grid: more
    than 4096 sub-columns, so the tile counts go through the separate tile scan and the row
    tables through the one-wave blocks, and more than 4096 blocks of sub-cell counts, so the
    counting sort takes its separate block scan.  Keys, order and cell starts of fluid and walls
    match the oracle bit for bit; the fields after 10 steps meet the precision's bar.  The layer
    is near rest, so its horizontal velocity is settling noise (|u| is 1/600 of |v| in the
    oracle): at M32 the velocity is compared as one vector field, as the error metric of the
    precision target defines it, and u, v each to 1e-9 at P64."""
    from sphnj_amd.sim import grid

    case = Case.generate("dam_break", tank_w=17.0, tank_h=9.0, col_w=17.0, col_h=0.05, dx=0.002)
    nx, ny, _, _ = grid(case.params)
    sort_block_keys = 256 * 8
    assert 2 * nx > 4096 and -(-4 * nx * ny // sort_block_keys) > 4096, (nx, ny)
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.particle_cell_update()
    ora.particle_cell_update()
    for species in (0, 2):
        for part, a, b in zip(("keys", "order", "starts"), sim.sorted(0, species), ora.sorted(0, species)):
            assert np.array_equal(a, b), (species, part)
    sim.close()
    ora.close()
    sim = Simulation(case, precision)
    ora = oracle_mod.Oracle(case)
    sim.step(10)
    ora.step(10)
    assert sim.status() == ora.status() == 0
    tol = 1e-9 if precision == "p64" else 1e-4
    a, b = sim.download(), ora.download()
    compare_state(a, b, tol, fluid_fields=FLUID_FIELDS if precision == "p64" else ("x", "y", "rho", "p"))
    uv = [np.stack([s["fluid"]["u"], s["fluid"]["v"]]) for s in (a, b)]
    assert rel_l2(*uv) <= tol


LONG_CASES = {
    # the 5 k-particle liquid dam break (dx = 0.02 m, K = 10) the reference's numbers start from
    "dam5k": lambda: Case.generate("dam_break", dx=0.02),
    # a two-phase dam break: 3.6 k water under 4.4 k air, K = 3
    "two_phase": lambda: Case.generate("dam_break", tank_w=1.0, tank_h=0.8, col_w=0.4, col_h=0.6, dx=0.01,
                                       air=1, K=3),
    # an elastic gate: 3.5 k water, 120 solids, S = 4 solid sub-steps
    "gate": small_gate,
}
_LONG_ORACLE = {}


@pytest.mark.parametrize("precision", ["p64", "m32"])
@pytest.mark.parametrize("name", ["dam5k", "two_phase", "gate"])
def test_precision_target_after_1000_steps(gpu, oracle_mod, name, precision):
    """The precision target where it is stated: relative L2 against the FP64 oracle after 1000
    steps, per field (position, density, pressure, velocity as one vector and per component;
    the gate's solid position, velocity, pressure and stress too): M32 within 1e-4 (seen: at most
    4.6e-6, on the solid v), P64 within 1e-9 (seen: at most 8.9e-12)."""
    if name not in _LONG_ORACLE:  # one oracle run serves both precisions
        ora = oracle_mod.Oracle(LONG_CASES[name]())
        ora.step(1000)
        _LONG_ORACLE[name] = ora.download()
        ora.close()
    ref = _LONG_ORACLE[name]
    sim = Simulation(LONG_CASES[name](), precision)
    sim.step(1000)
    assert sim.status() == 0
    out = sim.download()
    a, b = out["fluid"], ref["fluid"]
    errs = {f: rel_l2(a[f], b[f]) for f in ("x", "y", "u", "v", "rho", "p")}
    errs["vel"] = rel_l2(np.stack([a["u"], a["v"]]), np.stack([b["u"], b["v"]]))
    if len(ref["solid"]["x"]):
        errs.update({"solid." + f: rel_l2(out["solid"][f], ref["solid"][f])
                     for f in ("x", "y", "u", "v", "pressure", "stress")})
    print(name, precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    tol = 1e-9 if precision == "p64" else 1e-4
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, (tol, bad)
