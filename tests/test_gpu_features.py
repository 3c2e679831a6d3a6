"""GPU tests of execution options and the remaining outputs, through the C ABI:
block sizes / autotuner (results independent of the tiling), status flags forced from the
host, FP64 reductions, solid materials against the oracle, and M32 neighbour lists."""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import case, fluid, lattice, params, solid, wall
from sphnj_amd.sim import grid
from sphnj_amd import Case, Simulation, _abi as A

pytestmark = pytest.mark.gpu


def gate():
    return Case.generate("gate_fsi", dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=4)


def two_phase_small():
    return Case.generate("dam_break", tank_w=1.0, tank_h=0.8, col_w=0.4, col_h=0.6, dx=0.01, air=1, K=3)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_block_size_does_not_change_results(gpu, precision):
    """Every block size stages different tiles but performs the same per-particle sums in the
    same order: P64 is bit-identical.  M32 re-bases the staged fluid and wall positions on the
    tile's frame, so tiles of another size round separations differently, by float ulps of a
    tile height.  In this nearly static start, where large pressure terms cancel to tiny
    accelerations, that shows most in u and v (bounded here at 1e-5 relative after 4 steps,
    well inside the 1e-4 M32 bar); the classes stay identical."""
    c = two_phase_small()
    ref = Simulation(c, precision)
    ref.step(4)
    a = ref.download()
    for blk in (64, 128):
        s = Simulation(c, precision)
        s.set_block(0, blk)
        s.set_block(1, blk)
        s.step(4)
        b = s.download()
        for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
            if precision == "p64" or k == "boundary_type":
                assert np.array_equal(a["fluid"][k], b["fluid"][k]), (blk, k)
            else:
                assert rel_l2(a["fluid"][k], b["fluid"][k]) <= 1e-5, (blk, k, rel_l2(a["fluid"][k], b["fluid"][k]))


def test_autotune_keeps_state(gpu):
    c = two_phase_small()
    s = Simulation(c, "m32")
    t = Simulation(c, "m32")
    s.step(3)
    t.step(3)
    before = s.download()["fluid"]
    rep = s.autotune()
    assert rep["pair"]["block"] in (64, 128, 256) and rep["shift"]["block"] in (64, 128, 256)
    assert all(v > 0 for v in rep["pair"]["ms"].values())
    after = s.download()["fluid"]
    for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
        assert np.array_equal(before[k], after[k]), k
    # same tiles as the tuned context, so M32 sums round identically
    t.set_block(0, rep["pair"]["block"])
    t.set_block(1, rep["shift"]["block"])
    s.step(3)
    t.step(3)
    a, b = s.download()["fluid"], t.download()["fluid"]
    for k in ("x", "y", "u", "v", "rho", "p"):
        assert np.array_equal(a[k], b[k]), k


def test_status_flags_from_host(gpu, oracle_mod):
    p = params()
    x, y = lattice(9, 9, 0.008, 0.127, 0.127)          # 81 particles in one cell
    f = fluid(np.append(x, 5.0), np.append(y, 0.5))    # and one outside the cell box
    c = case(p, f)
    s = Simulation(c, "p64")
    s.step(1)
    o = oracle_mod.Oracle(c)
    o.step(1)
    st = s.status()
    assert st & A.ST_OUT_OF_GRID and st & A.ST_CELL_OVER_64
    assert st == o.status()


def test_energy_matches_oracle(gpu, oracle_mod):
    c = gate()
    s = Simulation(c, "p64")
    o = oracle_mod.Oracle(c)
    s.step(5)
    o.step(5)
    e1, e2 = s.energy(), o.energy()
    assert e1[5] == e2[5]
    for k in range(5):
        assert e1[k] == pytest.approx(e2[k], rel=1e-9, abs=1e-300), k


@pytest.mark.parametrize("material,kw", [
    (A.AUXETIC, dict(yield_strain=0.02, yield_stress=2.0e5, harden_strain=0.3, harden_coef=1.0e8, is_plastic=1,
                     auxetic_density=600.0)),
    (A.NEO_HOOKEAN, dict()),
    (A.LINEAR_ELASTIC, dict(is_2D_strain=1)),
])
@pytest.mark.parametrize("precision,tol", [("p64", 1e-9), ("m32", 2e-4)])
def test_solid_materials_match_oracle(gpu, oracle_mod, material, kw, precision, tol):
    """A compressed, sheared block with a fixed top row, under gravity, several steps."""
    X, Y = lattice(8, 12, 0.02, 0.4, 0.1)
    Am = np.array([[0.985, 0.02], [0.0, 0.99]])
    x = Am[0, 0] * (X - 0.48) + Am[0, 1] * (Y - 0.2) + 0.48
    y = Am[1, 0] * (X - 0.48) + Am[1, 1] * (Y - 0.2) + 0.2
    bc = (Y > Y.max() - 0.01).astype(np.int32)
    p = params(youngs_modulus=1.0e6, poisson_ratio=0.3, solid_density=1000.0, solid_sub_step=4, dt=2e-5,
               **kw)
    c = case(p, s=solid(x, y, X, Y, material=material, bc=bc))
    s = Simulation(c, precision)
    o = oracle_mod.Oracle(c)
    s.step(10)
    o.step(10)
    a, b = s.download()["solid"], o.download()["solid"]
    for k in ("x", "y", "u", "v", "stress", "deformation"):
        assert rel_l2(a[k], b[k]) <= tol, (k, rel_l2(a[k], b[k]))
    if material == A.AUXETIC:
        assert rel_l2(a["max_strain"], b["max_strain"]) <= tol


def test_m32_neighbor_lists_match_up_to_the_cutoff_band(gpu, oracle_mod):
    """M32 tests r^2 < rc^2 in float on cell-relative coordinates: lists equal the oracle's
    except for pairs within float rounding of the cut-off."""
    c = gate()
    s = Simulation(c, "m32")
    o = oracle_mod.Oracle(c)
    s.particle_cell_update()
    o.particle_cell_update()
    c1, l1 = s.neighbors()
    c2, l2 = o.neighbors()
    assert np.abs(c1 - c2).max() <= 1
    assert (c1 != c2).mean() < 1e-3
    same = np.repeat(c1 == c2, c1)
    same2 = np.repeat(c1 == c2, c2)
    assert np.array_equal(l1[same], l2[same2])


def test_two_phase_with_walls_and_gas_matches_oracle(gpu, oracle_mod):
    c = two_phase_small()
    s = Simulation(c, "p64")
    o = oracle_mod.Oracle(c)
    s.step(30)
    o.step(30)
    a, b = s.download(), o.download()
    for k in ("x", "y", "u", "v", "rho", "p"):
        assert rel_l2(a["fluid"][k], b["fluid"][k]) <= 1e-9, k
    assert np.array_equal(a["wall"]["phase"], b["wall"]["phase"])
    assert np.array_equal(a["fluid"]["boundary_type"], b["fluid"]["boundary_type"])


def test_dense_tiles_read_global_candidates(gpu, oracle_mod):
    """A block packed at a quarter of the particle spacing overflows every tile's LDS budget,
    so its tiles read their candidates from global memory, in the pair kernel and in every
    shift pass; beside it a block at the normal spacing is staged.  Both match the oracle, and
    P64 is bit-identical at every block size (the two paths sum the same terms in the same
    order)."""
    p = params(shift_per_step=2)
    xd, yd = lattice(40, 40, 0.005, 0.3, 0.3)
    xn, yn = lattice(20, 20, 0.02, 0.6, 0.2)
    x, y = np.r_[xd, xn], np.r_[yd, yn]
    c = case(p, f=fluid(x, y, u=0.5 * (y - 0.4), v=0.3 * (x - 0.5), p=100.0 * x, rho=1000.0 + 0.5 * y))
    o = oracle_mod.Oracle(c)
    o.step(3)
    ref = o.download()["fluid"]
    out = {}
    for prec, tol in (("p64", 1e-9), ("m32", 1e-4)):
        for blk in (64, 128, 256):
            s = Simulation(c, prec)
            s.set_block(0, blk)
            s.set_block(1, blk)
            s.step(3)
            out[prec, blk] = s.download()["fluid"]
            for k in ("x", "y", "u", "v", "rho", "p"):
                assert rel_l2(out[prec, blk][k], ref[k]) <= tol, (prec, blk, k)
    for blk in (128, 256):
        for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
            assert np.array_equal(out["p64", 64][k], out["p64", blk][k]), (blk, k)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_shift_runs_equal_full_windows(gpu, precision):
    """The shift passes walk recorded candidate runs (mode 1: recorded by the pair kernel when
    it shares the shift's tiles, else by the step's first shift pass), search the full pruned
    windows (mode 0), or record the runs in the first pass and always fall back to the windows
    rebuilt from its path bound (mode 2): the same non-zero terms in the same order, so every
    field is bit-identical, through graph replays, on a batch whose cases take 10 and 3 passes
    per step (the one with fewer carries its state over).  Pair tiles of 128 slots beside shift
    tiles of 256 leave the recording to the first shift pass; they are compared with mode 0 on
    the same tiles."""
    cases = [Case.generate("dam_break", dx=0.02, air=1, K=10),
             Case.generate("dam_break", dx=0.025, K=3)]
    sims = {}
    for m, pair_block in ((0, 256), (1, 256), (2, 256), (0, 128), (1, 128)):
        s = Simulation(cases, precision, shift_runs=m)
        s.set_block(0, pair_block)
        s.set_block(1, 256)
        sims[m, pair_block] = s
    for n in (1, 9, 30):
        out = {}
        for key, s in sims.items():
            s.step(n)
            out[key] = [s.download(c)["fluid"] for c in range(2)]
        for key, ref in (((1, 256), (0, 256)), ((2, 256), (0, 256)), ((1, 128), (0, 128))):
            for c in range(2):
                for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
                    assert np.array_equal(out[key][c][k], out[ref][c][k]), (n, key, c, k)
    for s in sims.values():
        assert s.status() == 0
    with pytest.raises(Exception):
        sims[1, 256].set_shift_runs(3)


def test_shift_runs_on_global_tiles(gpu):
    """Tiles too dense for LDS read global memory and record no runs; beside them staged
    tiles walk theirs: P64 with and without runs is bit-identical over several steps."""
    p = params(shift_per_step=4)
    xd, yd = lattice(40, 40, 0.005, 0.3, 0.3)
    xn, yn = lattice(20, 20, 0.02, 0.6, 0.2)
    x, y = np.r_[xd, xn], np.r_[yd, yn]
    c = case(p, f=fluid(x, y, u=0.5 * (y - 0.4), v=0.3 * (x - 0.5), p=100.0 * x, rho=1000.0 + 0.5 * y))
    a = Simulation(c, "p64", shift_runs=1)
    b = Simulation(c, "p64", shift_runs=0)
    a.step(3)
    b.step(3)
    da, db = a.download()["fluid"], b.download()["fluid"]
    for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
        assert np.array_equal(da[k], db[k]), k


def test_band_tiles_and_per_subcolumn_cut(gpu, oracle_mod):
    """A cell column holding a dense 10 x 7 cluster inside one sub-row (70 lanes in one sub-row
    pair, past 64 and below 128) below a sparse line of single particles, beside a normal block:
    at 64 slots its sub-row bands (the cluster's in two pieces) would need more tiles than its
    budget, so it is cut per sub-column, while the other columns take band tiles (at 128 and 256
    the cluster's column takes bands of 9 sub-rows).  Every block size matches the oracle, and
    P64 is bit-identical across them."""
    p = params(shift_per_step=2)
    _, _, cdx, cdy = grid(p)
    x0, y0 = -0.2 + 7 * cdx, -0.2 + 6 * cdy  # cell column 7, sub-row 12
    xb, yb = lattice(10, 10, 0.02, 0.1, 0.1)
    xc, yc = lattice(10, 7, 0.005, x0 + 0.01, y0 + 0.003)
    ys = np.arange(y0 + 0.1, 0.95, 0.045)
    xs = np.full_like(ys, x0 + 0.03)
    x, y = np.r_[xb, xc, xs], np.r_[yb, yc, ys]
    c = case(p, f=fluid(x, y, u=0.3 * (y - 0.4), v=0.2 * (x - 0.5), p=50.0 * x, rho=1000.0 + 0.3 * y))
    o = oracle_mod.Oracle(c)
    o.step(3)
    ref = o.download()["fluid"]
    out = {}
    for prec, tol in (("p64", 1e-9), ("m32", 1e-4)):
        for blk in (64, 128, 256):
            s = Simulation(c, prec)
            s.set_block(0, blk)
            s.set_block(1, blk)
            s.step(3)
            out[prec, blk] = s.download()["fluid"]
            for k in ("x", "y", "u", "v", "rho", "p"):
                assert rel_l2(out[prec, blk][k], ref[k]) <= tol, (prec, blk, k)
    for blk in (128, 256):
        for k in ("x", "y", "u", "v", "rho", "p", "boundary_type"):
            assert np.array_equal(out["p64", 64][k], out["p64", blk][k]), (blk, k)


def test_solid_sort_follows_cell_changes(gpu, oracle_mod):
    """A FIXED solid block driven at 4 m/s through a fluid block: its particles cross cell
    boundaries during the run, so the solids' sort, which is skipped on the steps where no solid
    changes cell, must run on exactly the others.  After every step the solids' keys, order and
    cell starts equal the oracle's bit for bit; the fluid matches at the P64 bar."""
    p = params(solid_fixed_speed_x=4.0)
    xf, yf = lattice(10, 10, 0.02, 0.3, 0.3)
    xs, ys = lattice(4, 10, 0.02, 0.52, 0.3)
    c = case(p, f=fluid(xf, yf), s=solid(xs, ys, bc=A.FIXED))
    sim = Simulation(c, "p64")
    ora = oracle_mod.Oracle(c)
    key0 = None
    changed = 0
    for step in range(40):
        sim.step(1)
        ora.step(1)
        ks, os_, ss = sim.sorted(0, 1)
        ko, oo, so = ora.sorted(0, 1)
        assert np.array_equal(ks, ko) and np.array_equal(os_, oo) and np.array_equal(ss, so), step
        changed += key0 is not None and not np.array_equal(ks, key0)
        key0 = ks
    assert changed > 0  # the block did cross cells
    a, b = sim.download(), ora.download()
    for k in ("x", "y", "u", "v", "rho", "p"):
        assert rel_l2(a["fluid"][k], b["fluid"][k]) <= 1e-9, k
    assert np.array_equal(a["solid"]["x"], b["solid"]["x"]) or rel_l2(a["solid"]["x"], b["solid"]["x"]) <= 1e-12


@pytest.mark.parametrize("what", ["empty", "walls_only", "one_fluid", "solids_only"])
def test_degenerate_inputs(gpu, oracle_mod, what):
    """Degenerate inputs step like the oracle: no particle at all, walls alone, a single fluid
    particle (ballistic under gravity), a solid block alone (the fluid stages are skipped)."""
    p = params()
    c = {
        "empty": lambda: case(p),
        "walls_only": lambda: case(p, w=wall(*lattice(5, 3, 0.02, 0.1, 0.0))),
        "one_fluid": lambda: case(p, f=fluid([0.5], [0.5])),
        "solids_only": lambda: case(p, s=solid(*lattice(4, 4, 0.02, 0.4, 0.4))),
    }[what]()
    sim = Simulation(c, "p64")
    ora = oracle_mod.Oracle(c)
    sim.step(3)
    ora.step(3)
    assert sim.status() == ora.status() == 0
    a, b = sim.download(), ora.download()
    for sp in ("fluid", "solid", "wall"):
        assert len(a[sp]["x"]) == len(b[sp]["x"]) == len(getattr(c, sp)["x"]), sp
        for k in ("x", "y"):
            if len(b[sp]["x"]):
                assert rel_l2(a[sp][k], b[sp][k]) <= 1e-12, (sp, k)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_degenerate_members_in_batch(gpu, precision):
    """A batch that mixes full cases with the degenerate ones (no particle, walls alone, one
    fluid particle, solids alone): the empty slot ranges and cell grids of those members leave
    the others untouched, and every member steps bit for bit as it steps alone."""
    p = params()
    cases = [gate(), case(p), case(p, w=wall(*lattice(5, 3, 0.02, 0.1, 0.0))),
             case(p, f=fluid([0.5], [0.5])), case(p, s=solid(*lattice(4, 4, 0.02, 0.4, 0.4))),
             two_phase_small()]
    batch = Simulation(cases, precision)
    batch.step(3)
    assert batch.status() == 0
    for i, c in enumerate(cases):
        one = Simulation(c, precision)
        one.step(3)
        a, b = batch.download(i), one.download()
        one.close()
        for sp in ("fluid", "solid", "wall"):
            assert len(a[sp]["x"]) == len(getattr(c, sp)["x"]), (i, sp)
            for k in a[sp]:
                assert np.array_equal(a[sp][k], b[sp][k]), (i, sp, k)


def test_destroy_returns_device_memory(gpu):
    """Contexts over a 2.6 M-cell grid (a few hundred MB of device memory each) are created,
    stepped through their graphs, autotuned, downloaded and destroyed eight times: the device's
    free memory comes back to within 64 MB of where it started."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so.7")

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    c = Case.generate("dam_break", tank_w=17.0, tank_h=9.0, col_w=17.0, col_h=0.05, dx=0.002)
    s = Simulation(c, "m32")  # the runtime's own first-use allocations happen here
    s.step(1)
    s.close()
    f0 = free_bytes()
    used = 0
    for k in range(8):
        s = Simulation(c, "m32" if k % 2 else "p64")
        s.step(2)
        if k == 0:
            s.autotune()
        used = max(used, f0 - free_bytes())
        s.download()
        s.close()
    f1 = free_bytes()
    assert used > 64 << 20, used  # the contexts were large enough to show a leak
    assert f0 - f1 < 64 << 20, (f0 - f1, used)


@pytest.mark.parametrize("precision", ["p64", "m32"])
@pytest.mark.parametrize("make", [gate, two_phase_small])
def test_stages_equal_step(gpu, precision, make):
    """The six phases called one by one equal the step, whose cell stage also runs the wall
    update and the coupling in its row-table launch (the FSI gate: fluid, solids and walls; a
    two-phase dam break with 3 shift passes, whose pair kernel records the shift's runs in
    both), over three steps."""
    c = make()
    a = Simulation(c, precision, graph=False)
    b = Simulation(c, precision, graph=False)
    for _ in range(3):
        for st in ("cells", "couple", "wall", "fluid", "solid", "shift"):
            a.stage(st)
        b.step(1)
    da, db = a.download(), b.download()
    for sp in ("fluid", "solid", "wall"):
        for k in da[sp]:
            assert np.array_equal(da[sp][k], db[sp][k]), (sp, k)


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_one_launch_solid_substeps_equal_launch_pairs(gpu, precision):
    """The S solid sub-steps of a fluid step as one launch with grid barriers equal the 2 S
    launch form (the default) bit for bit, through graph replays, on a batch whose cases have
    different sub-step counts (S = 4 and S = 10: the case with fewer carries its state over)."""
    from test_gpu_parity import small_gate

    cases = [small_gate(), small_gate(S=10, gate_t=0.008)]
    a = Simulation(cases, precision, persistent_solid=True)
    b = Simulation(cases, precision)
    for n in (1, 3, 7):
        a.step(n)
        b.step(n)
        for c in range(2):
            da, db = a.download(c), b.download(c)
            for sp in ("fluid", "solid"):
                for k in da[sp]:
                    assert np.array_equal(da[sp][k], db[sp][k]), (n, c, sp, k)
    assert a.status() == b.status() == 0
    # per-stage timing counts the sub-steps as one launch
    a.set_stage_timing(True)
    a.step(1)
    assert a.stage_time("solid")[0] == 1
    b.set_stage_timing(True)
    b.step(1)
    assert b.stage_time("solid")[0] == 2 * 10


def test_driver_run_writes_what_download_returns(gpu, tmp_path):
    """`sphnj run` on a generated case (the reference's main loop: input.ini + data/*.dat in,
    out/*.vtk every result_interval steps): the last VTK files hold what the library returns
    for the same case stepped as many times."""
    import subprocess

    from helpers import read_vtk
    from test_gpu_parity import small_gate

    driver = __import__("pathlib").Path(__file__).resolve().parent.parent / "sphnj_amd" / "bin" / "sphnj"
    case_dir = tmp_path / "case"
    r = subprocess.run([str(driver), "gen", "gate_fsi", str(case_dir), "dx=2.0e-3;tank_w=0.3;tank_h=0.2;S=4;steps=12"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ini = case_dir / "config" / "input.ini"
    ini.write_text(ini.read_text().replace("result_interval = 1000000", "result_interval = 5"))
    r = subprocess.run([str(driver), "run", str(case_dir)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "total time=" in r.stdout
    out = sorted(p.name for p in (case_dir / "out").iterdir())
    assert out == sorted(f"{s}-{k}.vtk" for s in ("fluid", "solid", "virt") for k in (0, 5, 10, 12)), out
    c = Case.read(case_dir)
    sim = Simulation(c, "p64")
    sim.step(12)
    d = sim.download()
    pts, f = read_vtk(case_dir / "out" / "fluid-12.vtk")
    assert np.array_equal(pts[:, 0], d["fluid"]["x"]) and np.array_equal(pts[:, 1], d["fluid"]["y"])
    for k, name in (("rho", "density"), ("p", "pressure"), ("u", "u"), ("v", "v")):
        assert np.array_equal(f[name], d["fluid"][k]), k
    assert np.array_equal(f["boundary_type"], d["fluid"]["boundary_type"])
    pts, f = read_vtk(case_dir / "out" / "solid-12.vtk")
    assert np.array_equal(pts[:, 1], d["solid"]["y"]) and np.array_equal(f["pressure"], d["solid"]["pressure"])
    assert np.array_equal(f["P_yy"], d["solid"]["stress"][:, 3])
    _, f = read_vtk(case_dir / "out" / "virt-12.vtk")
    assert np.array_equal(f["pressure"], d["wall"]["p"])


@pytest.mark.parametrize("precision,tol", [("p64", 1e-9), ("m32", 1e-4)])
def test_probes_match_oracle(gpu, oracle_mod, precision, tol):
    """The device probes after 40 steps of the FSI gate against the oracle's: total force,
    front, water height at the gate, von Mises stresses, mean pressure, and the gate's free end
    (outputTankBeamEnd's point, here the plate's bottom-left particle)."""
    from test_gpu_parity import small_gate

    c = small_gate()
    sim = Simulation(c, precision)
    ora = oracle_mod.Oracle(c)
    sim.step(40)
    ora.step(40)
    # at a lattice column: a point midway between two columns sits exactly d/2 from both,
    # where rounding decides which particles the strict window takes
    xh = float(np.unique(c.fluid["x"])[25])
    a, b = sim.probe_global(0, xh), ora.probe_global(0, xh)
    for k in b:
        if b[k] == 0.0:
            assert a[k] == 0.0, k
        else:
            assert abs(a[k] - b[k]) <= tol * abs(b[k]), (k, a[k], b[k])
    assert b["force_x"] != 0.0 and b["vm_elastic"] > 0.0 and b["height"] > 0.0
    X0, Y0 = c.solid["X"].min(), c.solid["Y"].min()
    pa, ia = sim.probe_solid(X0, Y0)
    pb, ib = ora.probe_solid(X0, Y0)
    assert ia == ib
    assert np.abs(pa - pb).max() <= tol * 1e-2


@pytest.mark.parametrize("precision", ["p64", "m32"])
def test_snapshot_restart_is_bit_exact(gpu, tmp_path, precision):
    """Write a snapshot after 15 steps, restore it into a new context, and step both on: the
    restored run equals the original bit for bit (an auxetic gate, so the max_strain history
    and the fluid boundary class must both come back).  A damaged file or a context with other
    parameters is refused."""
    from sphnj_amd.sim import snapshot_info

    c = Case.generate("gate_fsi", dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=4, material=1, K=2)
    c.params.is_plastic = 1          # a soft, plastic auxetic gate, so it yields within the run
    c.params.yield_stress, c.params.yield_strain = 1.0e2, 1.0e-5
    a = Simulation(c, precision)
    a.step(15)
    f = tmp_path / "gate.snap"
    a.snapshot(f, step=15)
    p, step, counts = snapshot_info(f)
    assert step == 15 and counts == c.counts and p.yield_strain == 1.0e-5
    b = Simulation.restore(f, precision)
    for n in (1, 9):
        a.step(n)
        b.step(n)
        da, db = a.download(), b.download()
        for sp in ("fluid", "solid", "wall"):
            for k in da[sp]:
                assert np.array_equal(da[sp][k], db[sp][k]), (n, sp, k)
    assert np.any(da["solid"]["max_strain"][:, 0] < -1.0e-5), "some particle must have yielded"
    raw = bytearray(f.read_bytes())
    raw[200] ^= 1
    g = tmp_path / "damaged.snap"
    g.write_bytes(bytes(raw))
    with pytest.raises(Exception, match="damaged"):
        Simulation.restore(g, precision)
    other = Case.generate("gate_fsi", dx=2.0e-3, tank_w=0.3, tank_h=0.2, S=5)
    import ctypes as C
    from sphnj_amd import _abi as AB, library
    h = C.c_void_p()
    assert library().sphnj_create(C.byref(other.params), 0, C.byref(h)) == 0
    assert library().sphnj_snapshot_load(h, 0, str(f).encode()) == AB.E_PARAM
    library().sphnj_destroy(h)
