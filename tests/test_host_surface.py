"""Host driver surface (CPU only): the INI reader with the reference's Config semantics
(include/config.h:114-159, src/config.cpp:59-129), parameter defaults and validation
(src/{physics,properties,domain,sim_time,kernel}.cpp), the .dat particle files
(include/host_data.cuh:24-115, src/properties.cpp readNumber) and the case generators."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from sphnj_amd import Case, _abi as A, default_params, grid, library, params_from_ini, validate_params
from sphnj_amd.sim import SphnjError

ROOT = Path(__file__).resolve().parent.parent


def test_defaults_match_reference_table():
    p = default_params()
    expect = dict(gravity=9.81, is_steady=0, particle_shift_coef=0.4, is_viscid=0, min_pressure=-1.013e5,
                  liquid_sound_speed=1500.0, liquid_density=1000.0, liquid_gamma=7.0, liquid_viscosity=1.01e-3,
                  gas_sound_speed=340.0, gas_density=1.29, gas_gamma=1.4, gas_viscosity=17.9e-6,
                  solid_density=2700.0, youngs_modulus=67.5e9, poisson_ratio=0.34, arti_vis_alpha=2.5,
                  arti_vis_beta=2.5, solid_fixed_speed_x=0.0, solid_fixed_speed_y=0.0, is_2D_strain=0, is_plastic=0,
                  auxetic_density=636.0, auxetic_modulus=0.0, auxetic_poisson_ratio=-0.2, yield_strain=0.0,
                  yield_stress=2.0e7, harden_strain=0.5, harden_coef=100.0, dt=0.0, istart=0, iend=0,
                  solid_sub_step=1, result_interval=1, sort_every_step=1, shift_per_step=1, smoothing_length=0.0,
                  particle_diameter=0.0, impact_length_hsml_ratio=3.0, thread_num=32, background_pressure=0.0,
                  kernel_2d_coef=0.0)
    for k, v in expect.items():
        assert getattr(p, k) == pytest.approx(v, rel=0, abs=0), k


def _ini(tmp_path, text):
    f = tmp_path / "input.ini"
    f.write_text(text)
    return params_from_ini(f)


def test_ini_comments_sections_and_repeats(tmp_path):
    p = _ini(tmp_path, """[physics]
gravity = 1.5 ## trailing comment
# a comment line with = sign
is_steady = false

[fluid]
liquid_density = 999.5
liquid_density = 1001.25   # a repeated key overwrites the earlier one
min_pressure=-5e4
""")
    assert p.gravity == 1.5
    assert p.is_steady == 0
    assert p.liquid_density == 1001.25
    assert p.min_pressure == -5e4
    assert p.gas_density == 1.29  # untouched default


@pytest.mark.parametrize("text,val", [("false", 0), ("F", 0), ("no", 0), ("n", 0), ("0", 0), ("NONE", 0),
                                      ("None", 0), ("true", 1), ("1", 1), ("-1", 1), ("yes", 1), ("off", 1)])
def test_ini_bool_parse(tmp_path, text, val):
    # string_as_T<bool>: only false/F/no/n/0/none (any case) read as false
    assert _ini(tmp_path, f"is_viscid = {text}\n").is_viscid == val


def test_ini_numeric_extraction_like_istringstream(tmp_path):
    p = _ini(tmp_path, "iend = 1e5\nistart = 12abc\ndt = 2.5e-4xyz\nsolid_sub_step = junk\n")
    assert p.iend == 1          # an int reads the leading "1" of "1e5" (the reference INI warns about this)
    assert p.istart == 12
    assert p.dt == 2.5e-4
    assert p.solid_sub_step == 0  # failed extraction -> 0


def test_ini_continuation_lines(tmp_path):
    # a line without '=' that follows a key is appended to its value (joined by '\n'); the
    # numeric read still takes the leading number, a bool read sees the whole value
    p = _ini(tmp_path, "gravity = 3.0\n4.0\nis_plastic = false\n[solid]\n\nyield_strain = 0.05\n")
    assert p.gravity == 3.0
    assert p.is_plastic == 1    # value "false\n[solid]" is not one of the false spellings
    assert p.yield_strain == 0.05


@pytest.mark.parametrize("text,viscid", [
    ("is_viscid = no\n   # note\nis_steady = 1\n", 0),    # a comment-only line does not extend the value
    ("is_viscid = no\n  # note\ny\n", 1),                 # ... nor close it: "no\ny" is not a false spelling
    ("is_viscid = no\n  # note\n\ny\n", 0),               # a blank line closes it; "y" then opens nothing
    ("is_viscid = no\n \t \ny\n", 0),                     # white space only is blank
    ("is_viscid =\nfalse\n", 0),                          # an empty first value continues on the next line
    ("stray words\nis_viscid = 0 # off\n", 0),            # lines before the first key are ignored
    ("is_viscid = n\n=\n", 0),                            # '=' alone opens an entry with an empty key
])
def test_ini_entry_grouping(tmp_path, text, viscid):
    """Which lines extend a value, which close it (src/config.cpp:84-118)."""
    p = _ini(tmp_path, text)
    assert p.is_viscid == viscid


def test_ini_last_line_without_newline(tmp_path):
    assert _ini(tmp_path, "gravity = 4.5").gravity == 4.5
    assert _ini(tmp_path, "is_viscid = \nno").is_viscid == 0


def test_ini_missing_file(tmp_path):
    with pytest.raises(SphnjError) as e:
        params_from_ini(tmp_path / "nope.ini")
    assert e.value.code == A.E_IO


def _valid():
    p = default_params()
    p.smoothing_length, p.particle_diameter, p.dt, p.iend = 0.026, 0.02, 1e-4, 10
    p.auxetic_modulus = 2e8
    p.cell_xmin, p.cell_ymin, p.cell_xmax, p.cell_ymax = -0.1, -0.1, 1.1, 1.1
    p.domain_xmin, p.domain_ymin, p.domain_xmax, p.domain_ymax = 0, 0, 1, 1
    return p


@pytest.mark.parametrize("field,value", [("smoothing_length", 0.0), ("particle_diameter", 0.0), ("dt", 0.0),
                                         ("iend", 0), ("cell_xmax", -0.1), ("domain_ymax", 0.0)])
def test_validation_rejects(field, value):
    p = _valid()
    assert validate_params(p) == 0
    setattr(p, field, value)
    assert validate_params(p) == A.E_PARAM
    assert "FAILED TO READ INPUT.INI" in library().sphnj_global_error().decode()


def test_validation_needs_auxetic_or_yield():
    # src/properties.cpp:150-154 rejects both zero, even for pure-fluid inputs
    p = _valid()
    p.auxetic_modulus = 0.0
    assert validate_params(p) == A.E_PARAM
    p.yield_strain = 0.05
    assert validate_params(p) == 0


def test_grid_floor_of_extent_over_cutoff():
    p = _valid()
    nx, ny, dx, dy = grid(p)
    cell = 3.0 * 0.026
    assert nx == math.floor(1.2 / cell) and ny == math.floor(1.2 / cell)
    assert dx == pytest.approx(1.2 / nx) and dx >= cell


def test_dat_round_trip_and_errors(tmp_path):
    c = Case.generate("gate_fsi", dx=4e-3, tank_w=0.3, tank_h=0.2, S=2)
    lib = library()
    h = C.c_void_p()
    assert lib.sphnj_case_create(b"gate_fsi", b"dx=4e-3;tank_w=0.3;tank_h=0.2;S=2", C.byref(A.Params()),
                                 C.byref(h)) == 0
    assert lib.sphnj_case_write(C.byref(c.params), h, str(tmp_path).encode()) == 0
    lib.sphnj_particles_destroy(h)
    r = Case.read(tmp_path)
    for sp in ("fluid", "solid", "wall"):
        a, b = getattr(c, sp), getattr(r, sp)
        for k in a:
            assert np.array_equal(a[k], b[k]), (sp, k)
    for k, _ in A.Params._fields_:
        assert getattr(r.params, k) == getattr(c.params, k), k
    # header count mismatch -> SPHNJ_E_COUNT
    f = tmp_path / "data" / "virt_particle.dat"
    lines = f.read_text().splitlines()
    f.write_text("\n".join([str(int(lines[0]) + 1)] + lines[1:]) + "\n")
    with pytest.raises(SphnjError) as e:
        Case.read(tmp_path)
    assert e.value.code == A.E_COUNT


def test_solid_rows_without_bc_column(tmp_path):
    d = tmp_path / "data"
    d.mkdir()
    (d / "fluid_particle.dat").write_text("0 0\n")
    (d / "virt_particle.dat").write_text("0\n")
    (d / "solid_particle.dat").write_text("2\n0.1 0.2 0 0 0.1 0.2 0\n0.3 0.4 1 2 0.3 0.4 1 1\n")
    h = C.c_void_p()
    assert library().sphnj_particles_read(str(d).encode(), C.byref(h)) == 0
    c = Case._from_handle(default_params(), h)
    assert list(c.solid["bc"]) == [0, 1]          # a 7-column row reads bc = 0 (UNCONSTRAINED)
    assert list(c.solid["material"]) == [0, 1]
    assert c.solid["u"][1] == 1.0 and c.solid["v"][1] == 2.0


def test_missing_dat_file(tmp_path):
    with pytest.raises(SphnjError) as e:
        Case.read(tmp_path)
    assert e.value.code in (A.E_IO,)


# ---- case generators: the configurations of the benchmark suite -----------------------

def test_case_dam_break_5k():
    c = Case.generate("dam_break")
    nf, ns, nw = c.counts
    assert (nf, ns) == (5000, 0)
    assert nw == (200 + 8) * (150 + 8) - 200 * 150     # 4 layers on 4 sides
    p = c.params
    c0 = 10 * math.sqrt(2 * 9.81 * 2.0)
    assert p.liquid_sound_speed == pytest.approx(c0)
    assert p.dt == pytest.approx(0.25 * 1.3 * 0.02 / c0)
    assert p.smoothing_length == pytest.approx(1.3 * 0.02)
    assert p.shift_per_step == 10
    # hydrostatic: p = rho0 g (H - y), rho = rho0 (p / B + 1)^(1/7)
    f = c.fluid
    assert np.allclose(f["p"], 1000 * 9.81 * (2.0 - f["y"]))
    B = 1000 * c0 ** 2 / 7
    assert np.allclose(f["rho"], 1000 * (f["p"] / B + 1) ** (1 / 7), rtol=1e-14)
    assert validate_params(p) == 0


def test_case_two_phase_counts():
    c = Case.generate("dam_break", dx=0.02, air=1)
    f = c.fluid
    assert (f["phase"] == 0).sum() == 5000
    assert (f["phase"] == 1).sum() == 200 * 150 - 5000
    # continuous hydrostatic pressure across the air/water surface
    gas = f["phase"] == 1
    assert np.allclose(f["p"][gas], 1.0 * 9.81 * (3.0 - f["y"][gas]))


def test_case_full_size_counts():
    c = Case.generate("gate_fsi")
    assert c.counts[:2] == (224000, 6320)
    assert (c.solid["bc"] == 1).sum() == 4 * 20            # clamped top rows
    c = Case.generate("wedge_entry", dx=4e-3)
    nf, ns, nw = c.counts
    assert ns > 0 and nf > 0 and np.all(c.solid["v"] == -6.0)


def test_case_ensemble_members():
    a = Case.generate("ensemble_member", index=3)
    b = Case.generate("ensemble_member", index=3)
    c = Case.generate("ensemble_member", index=4)
    assert a.counts == b.counts and np.array_equal(a.fluid["x"], b.fluid["x"])
    assert a.counts != c.counts
    ns = [Case.generate("ensemble_member", index=i).counts[0] for i in range(16)]
    assert min(ns) >= 4000 and max(ns) <= 210000


def test_unknown_case():
    with pytest.raises(SphnjError):
        Case.generate("no_such_case")


def test_library_exports_every_declared_symbol():
    """Every function declared in include/sphnj.h is exported by libsphnj.so (no GPU call)."""
    header = (ROOT / "include" / "sphnj.h").read_text()
    names = set(re.findall(r"^\s*(?:int|void|const char \*)\s*(sphnj_\w+)\s*\(", header, re.M))
    assert len(names) > 30
    lib = library()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.sphnj_abi_version() == 2


# ---- the reference's six scenarios (cases/<name>/particle_generate.cpp + config/input.ini) ----

REF_CASES = {
    # name: (fluid, liquid, solid, wall) at the reference's own spacing, from the generators'
    # loops: e.g. dam_break 1073 x 500 fluid columns/rows with a 2.0 x 1.0 liquid block
    # (400 x 200), walls = (1073 + 16)(500 + 16) - 536500
    "ref_dam_break": (536500, 80000, 0, 25424),
    "ref_water_tank": (320000, 320000, 416 * 20, 416 * 808 - 320000),
    "ref_beam_vibrate": (0, 0, 244 * 24, 0),
    "ref_beam_dam": (320 * 160 - 4 * 120, 80 * 112, 4 * 120, 336 * 176 - 320 * 160),
    "ref_wedge_entry": (1664640, 1200 * 800, 480 * 32, 1216 * 1416 - 1200 * 1400),
}


@pytest.mark.parametrize("name", sorted(REF_CASES))
def test_reference_scenario_counts(name):
    c = Case.generate(name)
    nf, nl, ns, nw = REF_CASES[name]
    assert c.counts == (nf, ns, nw)
    assert c.n_liquid == nl
    p = c.params
    assert p.smoothing_length == pytest.approx(1.2 * p.particle_diameter)
    assert validate_params(p) == 0


def test_reference_scenario_details():
    # water tank: hydrostatic column at c = 200, viscous, S = 20, K = 10; floor plate below y = 0
    c = Case.generate("ref_water_tank")
    f, p = c.fluid, c.params
    assert p.is_viscid == 1 and p.liquid_viscosity == 1.0 and p.solid_sub_step == 20 and p.shift_per_step == 10
    assert np.allclose(f["p"], 1000.0 * 9.81 * (2.0 - f["y"]))
    assert np.allclose(f["rho"], 1000.0 * (f["p"] * 7.0 / 200.0 ** 2 / 1000.0 + 1.0) ** (1 / 7))
    assert c.solid["y"].max() < 0 and np.all(c.solid["bc"] == A.UNCONSTRAINED)
    # wedge: 10 deg, auxetic lower half (j < 16), aluminium upper half, rim columns and the
    # top centre pair FIXED, everything at -30 m/s
    c = Case.generate("ref_wedge_entry")
    s = c.solid
    assert np.all(s["v"] == -30.0) and c.params.solid_fixed_speed_y == -30.0
    assert (s["material"] == A.AUXETIC).sum() == 480 * 16
    fixed = s["bc"] == A.FIXED
    assert fixed.sum() == 2 * 16 + 2 and np.all(s["material"][fixed] == A.LINEAR_ELASTIC)
    apex = np.abs(s["x"] - 1.5) < 2.5e-3
    assert np.allclose(s["y"][apex].min(), math.tan(10 * 3.1415926 / 180) * 1.25e-3 + 2.5 + 1.25e-3)
    # circle: 4 + 10 + 4 layers (elastic, auxetic, elastic) at 50 m/s; no fluid within r + dx
    c = Case.generate("ref_circle_entry")
    r = np.hypot(c.solid["x"] - 0.15, c.solid["y"] - 0.4)
    aux = c.solid["material"] == A.AUXETIC
    assert r[aux].min() > 0.05 - 0.014 and r[aux].max() < 0.05 - 0.004
    assert r.min() > 0.05 - 0.018 and r.max() < 0.05
    assert np.all(c.solid["v"] == -50.0) and c.params.particle_shift_coef == 2.0
    assert np.hypot(c.fluid["x"] - 0.15, c.fluid["y"] - 0.4).min() > 0.051
    # beam: first-mode velocity profile, zero at the root, largest at the tip
    c = Case.generate("ref_beam_vibrate")
    v = c.solid["v"]
    tip = v[np.argmax(c.solid["x"])] / (0.01 * math.sqrt(2e6 / 0.6 / 1000))   # v(L) = xi c_s, the tip row is dx/2 short
    assert c.params.gravity == 0.0 and 0.99 < tip < 1.0
    root = np.abs(c.solid["x"] - 5e-4) < 1e-9
    assert np.abs(v[root]).max() < 1e-4 * np.abs(v).max()
    # dam break: the reference's nondimensional one (g = 1, rho = 1 / 1e-3, c = 20), p = 0
    c = Case.generate("ref_dam_break")
    assert c.params.gravity == 1.0 and np.all(c.fluid["p"] == 0.0)
    assert set(np.unique(c.fluid["rho"])) == {1.0, 1.0e-3}


@pytest.mark.parametrize("name", sorted(REF_CASES) + ["ref_circle_entry"])
def test_reference_scenario_round_trips_through_files(tmp_path, name):
    """sphnj_case_write then the reference's surfaces back (input.ini + data/*.dat)."""
    import ctypes as C

    c = Case.generate(name, dx={"ref_wedge_entry": 0.02, "ref_dam_break": 0.02, "ref_circle_entry": 0.004,
                                "ref_water_tank": 0.01}.get(name, 2.5e-3 if name == "ref_beam_dam" else 1e-3))
    h = C.c_void_p()
    opts = ""
    lib = library()
    p = A.Params()
    assert lib.sphnj_case_create(name.encode(), opts.encode(), C.byref(p), C.byref(h)) == 0
    assert lib.sphnj_case_write(C.byref(c.params), h, str(tmp_path).encode()) == 0
    lib.sphnj_particles_destroy(h)
    back = Case.read(tmp_path)
    for k, _ in A.Params._fields_:
        assert getattr(back.params, k) == pytest.approx(getattr(c.params, k), rel=1e-15), k


DRIVER = ROOT / "sphnj_amd" / "bin" / "sphnj"


def test_driver_writes_input_particles_as_vtk(tmp_path):
    """`sphnj gen` then `sphnj vtk` (no GPU): the files parse back to the generated arrays, and
    an unwritable output directory is an error, not a silent success."""
    import subprocess

    from helpers import read_vtk

    case_dir = tmp_path / "case"
    r = subprocess.run([str(DRIVER), "gen", "gate_fsi", str(case_dir), "dx=4e-3;tank_w=0.3;tank_h=0.2;S=2"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    c = Case.read(case_dir)
    r = subprocess.run([str(DRIVER), "vtk", str(case_dir), str(tmp_path / "vtk")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    pts, f = read_vtk(tmp_path / "vtk" / "fluid-input.vtk")
    assert np.array_equal(pts[:, 0], c.fluid["x"]) and np.array_equal(pts[:, 1], c.fluid["y"])
    assert np.array_equal(f["density"], c.fluid["rho"]) and np.array_equal(f["pressure"], c.fluid["p"])
    assert np.array_equal(f["phase"], c.fluid["phase"])
    pts, f = read_vtk(tmp_path / "vtk" / "solid-input.vtk")
    assert np.array_equal(pts[:, 1], c.solid["y"]) and np.array_equal(f["X"], c.solid["X"])
    assert np.array_equal(f["material"], c.solid["material"])
    pts, f = read_vtk(tmp_path / "vtk" / "virt-input.vtk")
    assert len(pts) == c.counts[2] and np.array_equal(f["boundary_type"], c.wall["bc"])
    # a regular file where the output directory should be
    (tmp_path / "blocker").write_text("x")
    r = subprocess.run([str(DRIVER), "vtk", str(case_dir), str(tmp_path / "blocker" / "out")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "cannot" in r.stderr


def test_snapshot_info_rejects_other_files(tmp_path):
    from sphnj_amd.sim import snapshot_info

    with pytest.raises(SphnjError) as e:
        snapshot_info(tmp_path / "none.snap")
    assert e.value.code == A.E_IO
    f = tmp_path / "bad.snap"
    f.write_bytes(b"SPHNJSNQ" + bytes(200))
    with pytest.raises(SphnjError) as e:
        snapshot_info(f)
    assert e.value.code == A.E_IO and "not a sphnj snapshot" in str(e.value)
    f.write_bytes(b"SPHNJSNP" + bytes(4))
    with pytest.raises(SphnjError) as e:
        snapshot_info(f)
    assert e.value.code == A.E_IO
