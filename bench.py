"""Benchmark of the SPHNJ per-step particle pipeline on one MI355X.

Workload (default): the two-phase dam break at dx = 0.002 m -- 500,000 water + 2,500,000 air +
28,064 dummy wall particles (tank 4.0 x 3.0 m, water column 1.0 x 2.0 m, h = 1.3 dx, K = 10
shift passes per step).  A step is one full fluid time step of the pipeline (cell hash + key
sort + cell ranges, wall update, fluid pair + update, K shift passes) over all particles,
inputs resident in HBM.  The FP32-mixed form (M32) is timed, then the plain FP64 form (P64) on
the same input for the same number of steps; the relative L2 error of M32 against P64 is
reported per field.  Per-stage kernel durations come from HIP events recorded on the
library's stream.  The CPU baseline is the FP64 oracle on a bounded sample of the same workload.

The default run also times the batched form on the ensemble of 256 seeded liquid dam breaks
(15.7 M particles, one context, M32) against every member stepped alone in P64, as the
"ensemble" object of the same JSON line.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--config two_phase|dam5k|gate|wedge|ensemble]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "M particle-updates/s, FP32-mixed+batched vs FP64 plain on 1 GPU at ≤1e-4 rel L2"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per particle per launch: each kernel reads its per-particle
# inputs once and writes its outputs once, neighbour reads counted as ideal reuse.
BYTES = {
    # stage: {precision: bytes per fluid particle per launch}
    # pair + update (fused), with the shift's candidate runs it records (16 B written)
    "fluid": {"m32": 74 + 106, "p64": 136 + 138},
    # gradient pass + shift (fused), per pass, averaged over K = 10 passes, the runs recorded by
    # the pair kernel.  M32: every pass reads the views, path length and class and writes the
    # views and path length (25 B); the first also reads the runs (12 B) and writes the
    # accumulated shift (45 B); a middle pass reads the runs and reads and writes the accumulated
    # shift (53 B); the last also reads the sub-cell, the path since the hash kept with the runs,
    # the double masters, L, u, v and writes the masters, u, v (133 B).  P64: the positions are
    # the views, the accumulated shift and L are double (69, 85 and 141 B).
    "shift": {"m32": 60, "p64": 89},
    "cells": {"m32": 16 + 142, "p64": 20 + 150},   # hash + sort/permute/cell ranges
}
WALL_BYTES = {"m32": 38, "p64": 66}  # wall Shepard, per wall particle


def shift_step_bytes_m32(K):
    """M32 shift bytes per fluid particle per step at K passes, the runs recorded by the pair
    kernel: 45 B for the first pass, 53 B for each middle one and 133 B for the last (60.2 B a
    pass on average at K = 10); a single pass reads the sub-cell, views, path, class, masters,
    L, u, v and writes views, path, masters, u, v."""
    return 109 if K == 1 else 53 * K + 72


def configs():
    from sphnj_amd import Case

    return {
        "two_phase": ("two-phase dam break dx=0.002 m: 500k water + 2.5M air + 28k wall, K=10",
                      lambda: Case.generate("dam_break", dx=0.002, air=1)),
        "dam5k": ("liquid dam break dx=0.02 m: 5k water + 2.9k wall, K=10",
                  lambda: Case.generate("dam_break", dx=0.02)),
        "gate": ("elastic-gate FSI dx=2.5e-4 m: 224k water + 6.3k solid + 58k wall, K=1, S=10",
                 lambda: Case.generate("gate_fsi")),
        "wedge": ("wedge entry dx=1e-3 m: 3.0M water + 1.5M air + 8.3k solid + 36k wall, K=1, S=10",
                  lambda: Case.generate("wedge_entry")),
    }


def measured_traffic(config, stage, block):
    """HBM bytes per launch of the stage's kernel at the block size in use, as measured by the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/pmc_traffic_<workload>.json,
    copy-calibrated); None where no measurement of this workload and block size is committed."""
    kern = {"shift": f"k_fluid_shift<sphnj::M32, {block}>",
            "fluid": f"k_fluid_pair_update<sphnj::M32, false, {block}>", "cells": None}[stage]
    f = ROOT / "profiles" / {"two_phase": "pmc_traffic_two_phase_3M.json"}.get(config, "-")
    if kern is None or not f.exists():
        return None
    for k, v in json.loads(f.read_text())["kernels"].items():
        if kern in k:
            return v["hbm_bytes_per_dispatch"]
    return None


def measured_issue(config, stage, block):
    """Vector-ALU instructions per CU-cycle of the stage's kernel at the block size in use, from
    the committed rocprofv3 SQ_INSTS_VALU / GRBM_GUI_ACTIVE pass (profiles/pmc_valu_<workload>.json;
    a CU issues at most about one wave64 VALU instruction per cycle); None where not measured."""
    kern = {"shift": f"k_fluid_shift<sphnj::M32, {block}>",
            "fluid": f"k_fluid_pair_update<sphnj::M32, false, {block}>", "cells": None}[stage]
    f = ROOT / "profiles" / {"two_phase": "pmc_valu_two_phase_3M.json"}.get(config, "-")
    if kern is None or not f.exists():
        return None
    for k, v in json.loads(f.read_text())["kernels"].items():
        if kern in k:
            return v["valu_per_cu_cycle"]
    return None


def case_probes(config, case, a, b):
    """The configurations' own error figures, M32 (a) against P64 (b) at the compared step: the
    elastic gate's tip displacement (mean horizontal displacement of the free bottom row of the
    plate, which is clamped at the top), the wedge's largest surface pressure, and its mean solid
    pressure (the figure src/output.cu:546 logs)."""
    def rel(m, r):
        return {"m32": float(f"{m:.6e}"), "p64": float(f"{r:.6e}"),
                "rel_err": float(f"{abs(m - r) / abs(r):.3e}") if r else None}
    if config == "gate":
        Y = np.asarray(case.solid["Y"], np.float64)
        row = np.isclose(Y, Y.min())
        X = np.asarray(case.solid["X"], np.float64)[row]
        return {"gate_tip_displacement_m": rel(float(np.mean(a["solid"]["x"][row] - X)),
                                               float(np.mean(b["solid"]["x"][row] - X)))}
    if config == "wedge":
        pa, pb = a["solid"]["pressure"], b["solid"]["pressure"]
        return {"wedge_max_surface_pressure_pa": rel(float(np.max(pa)), float(np.max(pb))),
                "wedge_mean_solid_pressure_pa": rel(float(np.mean(pa)), float(np.mean(pb)))}
    return {}


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a - b))


def timed_steps(sim, steps):
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    return time.perf_counter() - t0


def copy_calibration(mib):
    """A plain device copy of `mib` MiB (torch), so that rocprofv3 FETCH_SIZE / WRITE_SIZE
    readings can be scaled by what the counters report for a copy of known size."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"# copy calibration: 3 x {mib} MiB device copies", file=sys.stderr, flush=True)


def stage_profile(sim, steps):
    if steps <= 0:
        return {st: {"launches": 0, "ms_per_step": 0.0, "ms_per_launch": 0.0}
                for st in ("cells", "couple", "wall", "fluid", "solid", "shift")}
    sim.set_stage_timing(True)
    sim.step(steps)
    out = {}
    for st in ("cells", "couple", "wall", "fluid", "solid", "shift"):
        n, ms = sim.stage_time(st)
        out[st] = {"launches": n, "ms_per_step": ms / steps, "ms_per_launch": (ms / n) if n else 0.0}
    sim.set_stage_timing(False)
    return out


def ensemble_step_bytes(cases):
    """Algorithmic M32 bytes of one step of a batch: fluid and wall stages per case (each case
    at its own shift count)."""
    tot = 0
    for c in cases:
        nf, _, nw = c.counts
        tot += nf * (BYTES["cells"]["m32"] + BYTES["fluid"]["m32"] + shift_step_bytes_m32(c.params.shift_per_step))
        tot += nw * WALL_BYTES["m32"]
    return tot


def run_ensemble(members, steps, warmup, serial_members=None):
    """The batched form: `members` seeded liquid dam breaks in one context (M32, graph-captured
    step, one launch per phase over every member's tiles), against the same members stepped
    one by one in plain FP64 -- every member timed, none extrapolated unless serial_members is
    given.  Returns a dict of the figures."""
    from sphnj_amd import Case, Simulation

    t0 = time.perf_counter()
    cases = [Case.generate("ensemble_member", index=i) for i in range(members)]
    n_total = sum(sum(m.counts) for m in cases)
    batch = Simulation(cases, "m32")
    batch.step(warmup)
    batch.autotune()
    t = timed_steps(batch, steps)
    value = n_total * steps / t / 1e6
    print(f"# ensemble: {members} members, {n_total} particles, batched m32 {1e3 * t / steps:.3f} ms/step "
          f"({value:.1f} M updates/s; set-up {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    nser = members if serial_members is None else min(serial_members, members)
    serial = 0.0
    n_ser = 0
    errs = []
    for i, m in enumerate(cases[:nser]):
        one = Simulation(m, "p64")
        one.step(warmup)
        serial += timed_steps(one, steps)
        n_ser += sum(m.counts)
        a = batch.download(i)["fluid"]
        b = one.download()["fluid"]
        # per field, velocity as one vector (as in the single-case line)
        errs.append(max([rel_l2(a[k], b[k]) for k in ("x", "y", "rho", "p")] +
                        [rel_l2(np.stack([a["u"], a["v"]]), np.stack([b["u"], b["v"]]))]))
        one.close()
    batch.close()
    p64_value = n_ser * steps / serial / 1e6
    gbs = ensemble_step_bytes(cases) * steps / t / 1e9
    out = {"members": members, "particles": n_total, "steps": steps, "warmup": warmup,
           "m32_batched_value": round(value, 3), "m32_batched_ms_per_step": round(1e3 * t / steps, 4),
           "p64_serial_value": round(p64_value, 3),
           "p64_serial_ms_per_step": round(1e3 * serial / steps * n_total / n_ser, 4),
           "speedup_vs_p64": round(value / p64_value, 3),
           "max_rel_l2_vs_p64": float(f"{max(errs):.3e}") if errs else None,
           "serial_members_timed": nser,
           "step_hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "species": "fluid + wall"}}
    print(f"# ensemble: serial p64 {p64_value:.1f} M updates/s over {nser} members, "
          f"speedup {value / p64_value:.2f}x, max rel L2 {max(errs):.2e}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="two_phase", choices=["two_phase", "dam5k", "gate", "wedge", "ensemble"])
    ap.add_argument("--profile-steps", type=int, default=5, help="steps timed per stage with HIP events")
    ap.add_argument("--cpu-steps", type=int, default=15,
                    help="oracle steps in the CPU baseline sample, about 10 s on 16 cores (0 = skip)")
    ap.add_argument("--members", type=int, default=256, help="members of the batched ensemble")
    ap.add_argument("--serial-members", type=int, default=None,
                    help="members timed one by one in P64 (default: all; fewer are scaled to all)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble figures")
    ap.add_argument("--no-p64", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps after which M32 is compared with P64 (both continued to this count)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default block sizes")
    ap.add_argument("--solid-one-launch", action="store_true",
                    help="solid sub-steps as one launch with grid barriers instead of 2 S launches")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run device copies of MIB MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    from sphnj_amd import Simulation

    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)

    if args.config == "ensemble":
        e = run_ensemble(args.members, args.steps, args.warmup, args.serial_members)
        line = {"metric": METRIC, "value": e["m32_batched_value"], "unit": "M particle-updates/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": e["m32_batched_ms_per_step"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
                "data": "synthetic",
                "config": {"workload": f"ensemble of {args.members} seeded liquid dam breaks, batched M32",
                           "particles": e["particles"]},
                "ensemble": e}
        print(json.dumps(line))
        return 0

    desc, make = configs()[args.config]
    case = make()
    nf, ns, nw = case.counts
    n_total = nf + ns + nw
    print(f"# {desc}: {nf} fluid, {ns} solid, {nw} wall", file=sys.stderr, flush=True)

    sim = Simulation(case, "m32", persistent_solid=args.solid_one_launch)
    sim.step(args.warmup)
    tune = sim.autotune() if not args.no_tune else None
    t = timed_steps(sim, args.steps)
    ev_ms = sim.last_step_ms()
    value = n_total * args.steps / t / 1e6
    prof = stage_profile(sim, args.profile_steps)
    print(f"# m32: {1e3 * t / args.steps:.3f} ms/step (events {ev_ms / args.steps:.3f}), "
          f"{value:.1f} M updates/s", file=sys.stderr, flush=True)

    extra = {}
    if not args.no_p64:
        ref = Simulation(case, "p64", persistent_solid=args.solid_one_launch)
        ref.step(args.warmup)
        tune64 = ref.autotune() if not args.no_tune else None
        tp = timed_steps(ref, args.steps)
        p64_value = n_total * args.steps / tp / 1e6
        prof64 = stage_profile(ref, args.profile_steps)
        done = args.warmup + args.steps + args.profile_steps
        if args.accuracy_steps > done:  # the metric's accuracy bar is quoted after 1000 steps
            sim.step(args.accuracy_steps - done)
            ref.step(args.accuracy_steps - done)
            done = args.accuracy_steps
        a, b = sim.download(), ref.download()
        errs = {f: rel_l2(a["fluid"][f], b["fluid"][f]) for f in ("x", "y", "u", "v", "rho", "p")}
        # velocity as one vector field: |u_a - u_b| over |u_b|, summed over particles
        errs["vel"] = rel_l2(np.stack([a["fluid"]["u"], a["fluid"]["v"]]), np.stack([b["fluid"]["u"], b["fluid"]["v"]]))
        if ns:
            errs.update({"solid_" + f: rel_l2(a["solid"][f], b["solid"][f]) for f in ("x", "y", "stress")})
        probes = case_probes(args.config, case, a, b)
        extra = {"p64_value": round(p64_value, 3), "p64_fluid_value": round(nf * args.steps / tp / 1e6, 3),
                 "p64_ms_per_step": round(1e3 * tp / args.steps, 4),
                 "speedup_vs_p64": round(value / p64_value, 3),
                 "rel_l2_vs_p64": {k: float(f"{v:.3e}") for k, v in errs.items()},
                 "steps_compared": done,
                 **probes,
                 "p64_stage_ms": {k: round(v["ms_per_step"], 4) for k, v in prof64.items()},
                 "p64_blocks": tune64 and {k: v["block"] for k, v in tune64.items()}}
        print(f"# p64: {1e3 * tp / args.steps:.3f} ms/step, {p64_value:.1f} M updates/s; "
              f"max rel L2 {max(errs.values()):.2e}", file=sys.stderr, flush=True)
        ref.close()

    # roofline of the dominant kernel (largest share of the step)
    dom = max(("fluid", "shift", "cells"), key=lambda s: prof[s]["ms_per_step"]) if args.profile_steps else "shift"
    per_launch_ms = prof[dom]["ms_per_launch"]
    K = case.params.shift_per_step
    bytes_pp = shift_step_bytes_m32(K) / K if dom == "shift" else BYTES[dom]["m32"]
    bytes_launch = bytes_pp * nf
    achieved = bytes_launch / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms else 0.0
    blocks = {"fluid": (tune or {}).get("pair", {}).get("block", 256),
              "shift": (tune or {}).get("shift", {}).get("block", 256), "cells": 256}
    kernel_name = {"fluid": f"k_fluid_pair_update<M32, {blocks['fluid']}>",
                   "shift": f"k_fluid_shift<M32, {blocks['shift']}> (average of the step's {K} passes)",
                   "cells": "cells stage (hash + counting sort + permute + ranges)"}[dom]
    # The neighbour kernels issue vector-ALU instructions at the CU's rate (about one wave64
    # instruction per cycle), so VALU issue, not HBM, is what binds them; achieved / peak / frac
    # are still the algorithmic HBM bytes against the HBM peak.  traffic and valu_per_cu_cycle
    # are read from the committed rocprofv3 counter passes of this workload (profiles/), not
    # measured in this run.
    roofline = {"bound": "valu", "kernel": kernel_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(args.config, dom, blocks[dom]),
                "valu_per_cu_cycle": measured_issue(args.config, dom, blocks[dom]),
                "counters_from": "committed rocprofv3 passes: profiles/pmc_traffic_two_phase_3M.json "
                                 "(FETCH_SIZE / WRITE_SIZE, copy-calibrated), profiles/pmc_valu_two_phase_3M.json "
                                 "(SQ_INSTS_VALU / GRBM_GUI_ACTIVE)",
                "bytes_per_particle": round(bytes_pp, 2), "ms_per_launch": round(per_launch_ms, 5)}

    # the whole step against HBM: algorithmic bytes of every fluid and wall stage (solids, which
    # are at most 0.3% of the particles of these workloads, are not counted) over the step time
    step_bytes = nf * (BYTES["cells"]["m32"] + BYTES["fluid"]["m32"] + shift_step_bytes_m32(K)) \
        + nw * WALL_BYTES["m32"]
    step_gbs = step_bytes * args.steps / t / 1e9
    step_hbm = {"achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(step_gbs / HBM_PEAK_GBS, 4), "bytes_per_step": int(step_bytes),
                "species": "fluid + wall"}

    cpu = None
    if args.cpu_steps > 0:
        from oracle.oracle_py import Oracle, set_threads

        cores = set_threads(0)
        ora = Oracle(case)
        t0 = time.perf_counter()
        ora.step(args.cpu_steps)
        tc = time.perf_counter() - t0
        cpu = {"value": round(n_total * args.cpu_steps / tc / 1e6, 4), "unit": "M particle-updates/s",
               "cores": cores, "kind": "port",
               "sample": f"FP64 oracle (OpenMP), {args.cpu_steps} steps of the same workload, {tc:.1f} s"}
        ora.close()

    line = {"metric": METRIC, "value": round(value, 3), "unit": "M particle-updates/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t / args.steps, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": desc, "fluid": nf, "solid": ns, "wall": nw, "form": "M32 (FP32-mixed)",
                       "dt": case.params.dt, "shift_per_step": case.params.shift_per_step,
                       "solid_sub_step": case.params.solid_sub_step},
            "roofline": roofline, "step_hbm": step_hbm, "cpu_baseline": cpu,
            "fluid_value": round(nf * args.steps / t / 1e6, 3),
            "stage_ms": {k: round(v["ms_per_step"], 4) for k, v in prof.items()},
            "autotune": tune}
    line.update(extra)
    if args.config == "two_phase" and not args.no_ensemble:
        sim.close()
        line["ensemble"] = run_ensemble(args.members, args.steps, args.warmup, args.serial_members)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
