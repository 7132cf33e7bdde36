#!/usr/bin/env python3
"""Benchmark of the Spark RK stage on one MI355X.

Workloads (--workload):
  rt_amr       (default) 3D Rayleigh-Taylor with 8 mass scalars on a 3-level
               block hierarchy (4384 leaves of 16^3: 32 / 256 / 4096 per
               level, 256^3-equivalent resolution over x, y in [0, 1], z in
               [0, 2]), gravity g = -1 as a source, device guard fill over
               the levels and fine/coarse flux correction every step,
               10 guard cells (telescoping RK3).  The configuration the
               FP32-flux target (>= 3x plain FP64 at L1(rho) <= 1e-5) is
               stated for; FP32-flux accuracy after --accuracy-steps
               (default 200) against the plain FP64 form.
  sedov128     3D Sedov blast, uniform 8^3 blocks of 16^3 interior cells
               (128^3 zones), 6 guard cells refilled on the device between
               stages, WENO5 + HLLC, RK3, CFL 0.45 (the reference's own 3D
               Sedov test value: at 0.8 the time step, the largest
               (|u_d| + c)/dx_d over directions, lets the unsplit 3D update
               go unstable within ~15 steps), gamma 1.4.
  species2d    2D two-gas shock tube (256^2 zones, 16^2 blocks, 10 guard
               cells): A 1 / gamma 1.4 against A 4 / gamma 5/3 across
               x + y = 1 with the Sod states, Multigamma EOS, species clamped
               and renormalised at faces; the plain FP64 form only.
  batched4096  4096 independent 3D blocks of 8^3 / 16^3 / 32^3 cells
               (p = 0.25 / 0.625 / 0.125), levels 0-4 (dx = 2^-l / 16),
               smooth random fields plus a Mach 2-5 planar shock in every
               16th block, 10 guard cells (telescoping RK3, frozen guards).
               The FP32-flux accuracy after --accuracy-steps steps is
               reported against the plain FP64 form.
A step is one full RK3 step of all blocks (CFL dt on the device, guard
fill where the workload has one, 3 stages).  Inputs are resident in HBM
before the timed region; the timed region is K device-resident steps
bracketed by device synchronisations.  The headline value is the plain
FP64 form; the batched FP64 and FP32-flux forms are timed beside it.

Prints one JSON line (see the metric/roofline/cpu_baseline fields).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "flash-x_amd"))

METRIC = "zone-updates/s per RK3 step: plain FP64 vs batched FP32-flux; % HBM roofline"
HBM_PEAK_GBS = 8000.0
# the reference's 3D Sedov test runs at 0.45 (Sedov/tests/test_amr_unsplit_3d.par:47)
SEDOV_CFL = 0.45
FP64_PEAK_TFS = 78.6

WORKLOADS = {
    "sedov128": "3D Sedov blast, 8^3 blocks of 16^3 (128^3 zones), 6 guard cells refilled on the device between "
                "stages, WENO5-Z + HLLC, RK3, CFL 0.45, gamma 1.4",
    "rt_amr": "3D Rayleigh-Taylor, 8 mass scalars, 3-level block hierarchy of 4384 16^3 leaves (32/256/4096 per "
              "level, 256^3-equivalent), rho 2 over 1, g = -1 source, hydrostatic p (2.5 at the interface), "
              "seeded interface modes of amplitude 0.01, periodic x/y and reflecting z, device guard fill over levels "
              "(limited linear prolongation) and fine/coarse flux correction every step, 10 guard cells, telescoping RK3, WENO5-Z + HLLC, CFL 0.4, gamma 1.4",
    "species2d": "2D two-gas shock tube, 16x16 blocks of 16^2 (256^2 zones), 10 guard cells, telescoping RK3, "
                 "gas 1 (A 1, gamma 1.4) | gas 2 (A 4, gamma 5/3) across x + y = 1 with the Sod states, Multigamma "
                 "EOS, species clamped and renormalised at faces, WENO5-Z + HLLC, CFL 0.8 (plain form only)",
    "batched4096": "4096 independent 3D blocks of 8^3/16^3/32^3 (p .25/.625/.125, seeds 20261202/20261203), "
                   "levels 0-4, Fourier fields + Mach 2-5 shocks every 16th block, 10 guard cells, telescoping RK3, "
                   "WENO5-Z + HLLC, CFL 0.8, gamma 1.4",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 50 sedov128, 10 otherwise)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 5 sedov128, 2 otherwise)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="rt_amr")
    ap.add_argument("--profile-steps", type=int, default=10, help="steps of the per-kernel timing pass (>= 1)")
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="steps of the FP32-flux vs FP64 accuracy run (default 0 sedov128, 100 batched4096, 200 rt_amr)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--forms", default="batched_fp64,fp32_flux",
                    help="comma list of the other forms timed beside plain FP64 (empty: none)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--profile-form", default=None, metavar="FORMS",
                    help="rocprofv3 pass over a comma list of forms, each alone: --warmup steps (the batched forms "
                         "tune their tiles here), a marker dispatch, --steps steps, a marker dispatch; prints one "
                         "JSON line and exits")
    ap.add_argument("--summarize", default=None, metavar="DIR",
                    help="reduce the rocprofv3 CSV files under DIR (kernel trace and --pmc passes of --profile-form "
                         "runs) to a per-kernel table of the dispatches after the marker; needs --calib DIR and "
                         "--summary-steps; prints the table and exits")
    ap.add_argument("--calib", default=None, metavar="DIR",
                    help="rocprofv3 FETCH_SIZE / WRITE_SIZE CSVs of a --calibrate-copy run (with --summarize)")
    ap.add_argument("--summary-steps", type=int, default=None, help="profiled steps per form (with --summarize)")
    ap.add_argument("--summary-form", type=int, default=0,
                    help="which form of a --profile-form list to summarize (0-based, with --summarize)")
    a = ap.parse_args()
    big = a.workload in ("batched4096", "rt_amr")
    if a.steps is None:
        a.steps = 10 if big else 50
    if a.warmup is None:
        a.warmup = 2 if big else 5
    if a.accuracy_steps is None:
        a.accuracy_steps = {"sedov128": 0, "species2d": 0, "batched4096": 100, "rt_amr": 200}[a.workload]
    a.profile_steps = max(1, a.profile_steps)
    return a


def calibrate_copy(mib: int):
    """Device passes of known size for the counter scale: a 1 GiB read and
    write by an ordinary elementwise kernel (torch.neg, what the scale is
    taken from) and by the runtime's blit copy (reported beside it: its
    reads are counted at half their bytes)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 8, "launches": 3}), flush=True)


def make_problem(workload):
    from sparkhip import problems
    if workload == "batched4096":
        return problems.batched(nblocks=4096, nguard=10)
    if workload == "rt_amr":
        return problems.rt_amr()
    if workload == "species2d":
        return problems.sod2d_species(nbx=16, nby=16, n=16)
    return problems.sedov3d(nb=8, n=16, nguard=6, guard_mode="refill", cfl=SEDOV_CFL)


def cpu_baseline(workload, target_s: float):
    """The oracle (CPU restatement, OpenMP over blocks) on a bounded sample of
    the same workload, about target_s seconds of wall time."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    from sparkhip import problems

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    if workload == "species2d":
        prob = problems.sod2d_species(nbx=16, nby=16, n=16)
        run = lambda n: pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, prob.arrays, n, nthreads=threads)
        what = "the same two-gas shock tube (256^2 zones)"
    elif workload == "batched4096":
        prob = problems.batched(nblocks=256, nguard=10)
        run = lambda n: pyoracle.batch_advance(prob.cfg, prob.blocks, prob.arrays, n, nthreads=threads)
        what = f"the first 256 of the 4096 mixed blocks ({prob.zones} zones)"
    elif workload == "rt_amr":
        prob = problems.rt_amr(nb0=(1, 1, 2))
        run = lambda n: pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"],
                                             prob.arrays, n, nthreads=threads, bc=prob.info["bc"])
        what = (f"the same Rayleigh-Taylor setup on a 1 x 1 x 2 level-1 grid ({len(prob.blocks)} leaves, "
                f"{prob.zones} zones, guard fill and flux correction included)")
    else:
        prob = problems.sedov3d(nb=8, n=16, nguard=6, guard_mode="refill", cfl=SEDOV_CFL)
        run = lambda n: pyoracle.ug_advance(prob.cfg, prob.blocks[0], prob.ug, prob.arrays, n, refill=True,
                                            nthreads=threads)
        what = "3D Sedov 128^3 (8^3 blocks of 16^3, 6 guards, refill)"
    t0 = time.perf_counter()
    run(1)
    one = time.perf_counter() - t0
    nsteps = int(max(1, min(100, (target_s - one) / max(one, 1e-3))))
    t0 = time.perf_counter()
    run(nsteps)
    el = time.perf_counter() - t0
    return {
        "value": prob.zones * nsteps / el,
        "unit": "zone-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what}, {nsteps} RK3 steps after 1 untimed, {el:.1f} s wall, OpenMP over blocks",
    }



def norms(prob, got, ref):
    """Volume-weighted relative L1 and relative L-infinity of rho, rho*u, E."""
    import numpy as np
    from sparkhip import capi
    g = prob.cfg["nguard"]
    V = capi.VI
    fields = {
        "rho": lambda a: a[..., V["dens"]],
        "rhou": lambda a: a[..., V["dens"]] * a[..., V["velx"]],
        "E": lambda a: a[..., V["dens"]] * a[..., V["ener"]],
    }
    out = {}
    for name, f in fields.items():
        num = den = linf = scale = 0.0
        for b, x, r in zip(prob.blocks, got, ref):
            vol = b["dx"] * b["dy"] * b["dz"]
            fx, fr = f(x[g:-g, g:-g, g:-g]), f(r[g:-g, g:-g, g:-g])
            num += float(np.sum(np.abs(fx - fr))) * vol
            den += float(np.sum(np.abs(fr))) * vol
            linf = max(linf, float(np.max(np.abs(fx - fr))))
            scale = max(scale, float(np.max(np.abs(fr))))
        # relative norms; absolute ones where the reference field is zero
        out[name] = {"L1": num / den if den > 0 else num, "Linf": linf / scale if scale > 0 else linf,
                     "relative": bool(den > 0 and scale > 0)}
    return out



# name fragment of the marker dispatch (torch's fill kernel) that separates
# the untimed warmup / tuning dispatches from the profiled steps
MARKER = "FillFunctor"
NUM_CU = 256
CLOCK_HZ = 2.4e9  # MI355X peak engine clock


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            yield {k.lower(): v for k, v in row.items()}


def _segment(rows, i):
    """Dispatches of the i-th profiled form: between its two marker
    dispatches (one before and one after its timed steps)."""
    rows = sorted(rows, key=lambda r: int(r["dispatch_id"]))
    marks = sorted({int(r["dispatch_id"]) for r in rows if MARKER in r["kernel_name"]})
    if len(marks) < 2 * i + 2:
        raise SystemExit(f"profile holds {len(marks)} marker dispatches, form {i} needs {2 * i + 2}")
    lo, hi = marks[2 * i], marks[2 * i + 1]
    return [r for r in rows if lo < int(r["dispatch_id"]) < hi]


def summarize(args):
    """Per-kernel table of the profiled steps: dispatches per step, mean
    duration (kernel trace) and per-dispatch means of the counters (pmc
    passes), HBM bytes scaled by the calibration pass."""
    from collections import defaultdict
    root = Path(args.summarize)
    nsteps = args.summary_steps
    # calibration: the largest dispatch of a --calibrate-copy kernel moves
    # 1 GiB in and 1 GiB out
    def calib_of(match):
        out = {}
        for f in Path(args.calib).rglob("*counter_collection.csv"):
            for r in _csv_rows(f):
                if match(r["kernel_name"].lower()):
                    c = r["counter_name"]
                    out[c] = max(out.get(c, 0.0), float(r["counter_value"]) * 1024.0)  # KiB -> bytes
        return out
    true_bytes = 1024 << 20
    calib = calib_of(lambda k: "neg" in k and "elementwise" in k)
    blit = calib_of(lambda k: "copybuffer" in k)
    scale = {c: true_bytes / v for c, v in calib.items() if v > 0}
    dur = defaultdict(list)
    for f in root.rglob("*kernel_trace.csv"):
        for r in _segment(list(_csv_rows(f)), args.summary_form):
            dur[r["kernel_name"]].append((int(r["end_timestamp"]) - int(r["start_timestamp"])) * 1e-3)
    ctr = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))  # kernel -> counter -> dispatch -> value
    for f in root.rglob("*counter_collection.csv"):
        for r in _segment(list(_csv_rows(f)), args.summary_form):
            ctr[r["kernel_name"]][r["counter_name"]][r["dispatch_id"]] += float(r["counter_value"])

    def mean(k, c):
        d = ctr[k].get(c)
        if not d:
            d = ctr[k].get(c + "_sum")
        return sum(d.values()) / len(d) if d else None

    def short(name):
        name = name.replace("(anonymous namespace)::", "")
        return name.split("(")[0].replace("void ", "").replace("spark::", "").strip()

    lines = []
    cols = ["kernel", "per_step", "avg_us", "fetch_MB", "write_MB", "hbm_MB", "hbm_frac", "valu_insts", "valu_busy",
            "valu_issue", "fp64_gflop", "fp64_frac", "fp32_gflop", "lds_insts", "salu_insts", "waves"]
    lines.append(" | ".join(cols))
    for k in sorted(dur, key=lambda k: -sum(dur[k])):
        n = len(dur[k])
        us = sum(dur[k]) / n
        fe, wr = mean(k, "FETCH_SIZE"), mean(k, "WRITE_SIZE")
        fe = fe * 1024.0 * scale.get("FETCH_SIZE", 1.0) if fe is not None else None
        wr = wr * 1024.0 * scale.get("WRITE_SIZE", 1.0) if wr is not None else None
        hbm = fe + wr if fe is not None and wr is not None else None
        act, gui = mean(k, "SQ_ACTIVE_INST_VALU"), mean(k, "GRBM_GUI_ACTIVE")
        busy = act / NUM_CU / gui if act is not None and gui else None
        # VALU issue fraction: wave64 VALU instructions x 4 cycles (16-lane
        # SIMD) over the 1024 SIMDs' cycles of the launch at 2.4 GHz
        vi = mean(k, "SQ_INSTS_VALU")
        issue = vi * 4.0 / (us * 1e-6 * CLOCK_HZ * 4 * NUM_CU) if vi is not None else None
        f64 = [mean(k, "SQ_INSTS_VALU_" + x + "_F64") for x in ("ADD", "MUL", "FMA", "TRANS")]
        f32 = [mean(k, "SQ_INSTS_VALU_" + x + "_F32") for x in ("ADD", "MUL", "FMA", "TRANS")]
        fl64 = 64.0 * (f64[0] + f64[1] + 2.0 * f64[2] + f64[3]) if None not in f64 else None
        fl32 = 64.0 * (f32[0] + f32[1] + 2.0 * f32[2] + f32[3]) if None not in f32 else None

        def fmt(x, spec):
            return "-" if x is None else format(x, spec)
        row = [short(k), fmt(n / nsteps, ".2f"), fmt(us, ".1f"), fmt(fe and fe / 1e6, ".1f"), fmt(wr and wr / 1e6, ".1f"),
               fmt(hbm and hbm / 1e6, ".1f"), fmt(hbm and hbm / (us * 1e-6) / (HBM_PEAK_GBS * 1e9), ".3f"),
               fmt(mean(k, "SQ_INSTS_VALU"), ".4g"), fmt(busy, ".3f"), fmt(issue, ".3f"),
               fmt(fl64 and fl64 / 1e9, ".2f"),
               fmt(fl64 and fl64 / (us * 1e-6) / (FP64_PEAK_TFS * 1e12), ".3f"), fmt(fl32 and fl32 / 1e9, ".2f"),
               fmt(mean(k, "SQ_INSTS_LDS"), ".4g"), fmt(mean(k, "SQ_INSTS_SALU"), ".4g"),
               fmt(mean(k, "SQ_WAVES"), ".4g")]
        lines.append(" | ".join(row))
    if scale:
        lines.append("# FETCH_SIZE / WRITE_SIZE scaled by true bytes / counter bytes of a 1 GiB elementwise "
                     "read + write (torch.neg): " + ", ".join(f"{c} x{v:.3f}" for c, v in sorted(scale.items())) +
                     "; the runtime's blit copy of 1 GiB reads as " +
                     ", ".join(f"{c} {v / true_bytes:.3f}" for c, v in sorted(blit.items())) + " of its bytes")
    print("\n".join(lines))


def read_counter_summary(workload, form):
    """The committed per-kernel counter table of one form of a workload
    (written by --summarize), as {kernel: {column: value}}, or None."""
    path = ROOT / "profiles" / f"{workload}_{form}_counters.txt"
    if not path.exists():
        return None
    cols, out = None, {}
    for ln in path.read_text().splitlines():
        if ln.startswith("#") or not ln.strip():
            continue
        parts = [x.strip() for x in ln.split("|")]
        if cols is None:
            cols = parts
            continue
        out[parts[0]] = {c: (None if v == "-" else float(v)) for c, v in zip(cols[1:], parts[1:])}
    return {"rows": out, "source": str(path.relative_to(ROOT))}


def profile_form(args, prob, forms):
    """Each listed form alone, bracketed by marker dispatches (see
    --profile-form)."""
    import torch
    from sparkhip import capi
    capi.init(0)
    out = []
    for form in forms.split(","):
        b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
        b.pack(prob.arrays)
        if prob.ug is not None:
            b.set_uniform_grid(*prob.ug)
        if "origins" in prob.info:
            b.set_amr_grid(prob.info["nb0"], prob.info["origins"], prob.info["bc"])
        b.advance(max(1, args.warmup))
        b.synchronize()
        torch.zeros(1, device="cuda")  # marker: the profiled steps follow
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.advance(args.steps)
        b.synchronize()
        el = time.perf_counter() - t0
        torch.zeros(1, device="cuda")  # marker: the profiled steps end
        torch.cuda.synchronize()
        b.close()
        out.append({"form": form, "ms_per_step": round(el / args.steps * 1e3, 3)})
        print(f"[bench] profiled {form}: {out[-1]['ms_per_step']} ms/step", file=sys.stderr, flush=True)
    print(json.dumps({"profile_forms": out, "workload": args.workload, "steps": args.steps}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the block set is a single-GPU "
              f"shard and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.summarize:
        summarize(args)
        return
    from sparkhip import capi

    prob = make_problem(args.workload)
    if args.profile_form:
        profile_form(args, prob, args.profile_form)
        return
    capi.init(0)

    def batch(form):
        b = capi.Batch(dict(prob.cfg, form=form), prob.blocks)
        b.pack(prob.arrays)
        if prob.ug is not None:
            b.set_uniform_grid(*prob.ug)
        if "origins" in prob.info:
            b.set_amr_grid(prob.info["nb0"], prob.info["origins"], prob.info["bc"])
        return b

    def timed(form):
        b = batch(form)
        if args.warmup:
            b.advance(args.warmup)   # the batched forms tune their tiles here
        b.synchronize()
        t0 = time.perf_counter()
        b.advance(args.steps)
        b.synchronize()
        el = time.perf_counter() - t0
        return b, el

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    note(f"{args.workload}: {len(prob.blocks)} blocks, {prob.zones} zones")
    # headline: the plain FP64 form
    b, el = timed("plain_fp64")
    value = prob.zones * args.steps / el
    note(f"plain_fp64 {el / args.steps * 1e3:.3f} ms/step")

    # per-kernel HIP-event timing over a separate pass (same stream)
    b.profile(True)
    b.advance(args.profile_steps)
    stats = b.profile_read()
    b.profile(False)
    b.close()

    forms = {"plain_fp64": {"zone_updates_per_s": round(value, 1), "ms_per_step": round(el / args.steps * 1e3, 4),
                            "kernel_ms_per_step": {x["name"]: round(x["ms"] / args.profile_steps, 4)
                                                   for x in sorted(stats, key=lambda x: -x["ms"])}}}
    batched_ok = prob.cfg.get("nspecies", 0) == 0 and prob.cfg.get("eos", "gamma") in ("gamma", 0)
    for form in [f for f in args.forms.split(",") if f and batched_ok]:
        bf, elf = timed(form)
        bf.profile(True)
        bf.advance(args.profile_steps)
        st = bf.profile_read()
        bf.profile(False)
        tiles = bf.tiles()
        bf.close()
        fk = [x for x in st if x["name"] == "k_fused_stage"]
        entry_k = {x["name"]: round(x["ms"] / args.profile_steps, 4) for x in sorted(st, key=lambda x: -x["ms"])}
        entry = {"zone_updates_per_s": round(prob.zones * args.steps / elf, 1),
                 "ms_per_step": round(elf / args.steps * 1e3, 4),
                 "speedup_vs_plain_fp64": round(el / elf, 3),
                 "tiles": [list(t) for t in tiles],
                 "kernel_ms_per_step": entry_k}
        if fk:
            k = fk[0]
            entry["fused_stage_avg_ms"] = round(k["ms"] / k["launches"], 4)
            entry["fused_stage_valu_tflops"] = round(k["flops"] / (k["ms"] * 1e-3) / 1e12, 2)
        forms[form] = entry
        note(f"{form} {entry['ms_per_step']} ms/step")

    accuracy = None
    if args.accuracy_steps > 0 and batched_ok:
        # FP32-flux against the plain FP64 form (bit-exact with the FP64
        # oracle) after the same number of steps from the same state
        res = {}
        for form in ("plain_fp64", "fp32_flux"):
            bf = batch(form)
            dts = bf.advance(args.accuracy_steps)
            out = [a.copy() for a in prob.arrays]
            bf.unpack(out)
            bf.close()
            res[form] = (dts, out)
            note(f"accuracy run {form}: {args.accuracy_steps} steps, last dt {dts[-1]:.4e}")
        accuracy = {"steps": args.accuracy_steps, "reference": "plain_fp64", "form": "fp32_flux",
                    "dt_final_rel_diff": float(abs(res["fp32_flux"][0][-1] - res["plain_fp64"][0][-1]) /
                                               res["plain_fp64"][0][-1]),
                    **norms(prob, res["fp32_flux"][1], res["plain_fp64"][1])}
        del res

    tot = sum(s["ms"] for s in stats)
    sweeps = ("k_flux_x", "k_flux_y", "k_flux_z")  # one kernel template (not k_flux_exchange)
    flux = [s for s in stats if s["name"] in sweeps]
    fam = {}
    for s_ in stats:
        k = "k_flux" if s_["name"] in sweeps else s_["name"]
        fam[k] = fam.get(k, 0.0) + s_["ms"]
    dom = max(fam, key=fam.get)
    fl_ms = sum(s["ms"] for s in flux)
    fl_bytes = sum(s["bytes"] for s in flux)
    fl_flops = sum(s["flops"] for s in flux)
    fl_launch = sum(s["launches"] for s in flux)
    avg_s = fl_ms * 1e-3 / max(1, fl_launch)
    alg_gbs = fl_bytes / (fl_ms * 1e-3) / 1e9
    alg_tfs = fl_flops / (fl_ms * 1e-3) / 1e12
    # counters of the timed plain-form dispatches (committed summary, see
    # --summarize): HBM bytes per launch and VALU busy of the k_flux family
    cs = read_counter_summary(args.workload, "plain_fp64")
    frows = [v for k, v in cs["rows"].items() if k.startswith("k_flux<")] if cs else []
    traffic = (sum(r["hbm_MB"] for r in frows) / len(frows) * 1e6
               if frows and all(r.get("hbm_MB") is not None for r in frows) else None)
    # VALU issue fraction (wave64 VALU instructions x 4 cycles over the SIMD
    # cycles of the launch), else the SQ_ACTIVE_INST_VALU busy fraction
    col = "valu_issue" if frows and all(r.get("valu_issue") is not None for r in frows) else "valu_busy"
    busy = (sum(r[col] for r in frows) / len(frows)
            if frows and all(r.get(col) is not None for r in frows) else None)
    hbm_counter_frac = traffic / avg_s / (HBM_PEAK_GBS * 1e9) if traffic else None
    # the binding roof: the counters' VALU busy against their HBM fraction;
    # without counters, the algorithmic intensity against the FP64 ridge
    if busy is not None and hbm_counter_frac is not None:
        bound = "valu" if busy > hbm_counter_frac else "hbm"
    else:
        bound = "valu" if fl_flops / max(fl_bytes, 1.0) > FP64_PEAK_TFS * 1e12 / (HBM_PEAK_GBS * 1e9) else "hbm"
    roofline = {
        "bound": bound,
        "kernel": "k_flux (x, y, z sweeps: WENO5-Z + HLLC face fluxes; the last sweep also applies the "
                  "conserved update + EOS)",
        "achieved": round(alg_tfs, 2) if bound == "valu" else round(alg_gbs, 1),
        "peak": FP64_PEAK_TFS if bound == "valu" else HBM_PEAK_GBS,
        "unit": "TFLOP/s" if bound == "valu" else "GB/s",
        "frac": round(alg_tfs / FP64_PEAK_TFS if bound == "valu" else alg_gbs / HBM_PEAK_GBS, 4),
        "traffic": round(traffic) if traffic else None,
        "traffic_source": cs["source"] if cs else None,
        "avg_launch_ms": round(avg_s * 1e3, 4),
        "alg_bytes_per_launch": round(fl_bytes / max(1, fl_launch)),
        "alg_flops_per_launch": round(fl_flops / max(1, fl_launch)),
        "hbm": {"alg_gbs": round(alg_gbs, 1), "alg_frac": round(alg_gbs / HBM_PEAK_GBS, 4),
                "counter_gbs": round(traffic / avg_s / 1e9, 1) if traffic else None,
                "counter_frac": round(hbm_counter_frac, 4) if hbm_counter_frac else None},
        "valu_fp64": {"alg_tflops": round(alg_tfs, 2), "alg_frac": round(alg_tfs / FP64_PEAK_TFS, 4),
                      "counter_valu_frac": round(busy, 4) if busy is not None else None,
                      "counter_valu_column": col if busy is not None else None},
    }
    # algorithmic FP64 state bytes per zone-update: 8 (48 + 8 N_s)
    b_alg = 8.0 * (48 + 8 * prob.cfg.get("nscalars", 0))
    # the plain path's counter HBM fraction: every kernel's committed bytes per
    # dispatch times its dispatches per step, over this run's time per step
    path_counter = None
    if cs and all(r.get("hbm_MB") is not None for r in cs["rows"].values()):
        step_hbm = sum(r["hbm_MB"] * 1e6 * r["per_step"] for r in cs["rows"].values())
        path_counter = step_hbm / (el / args.steps) / (HBM_PEAK_GBS * 1e9)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload] + ", plain FP64 form",
            "name": args.workload, "zones": prob.zones, "blocks": len(prob.blocks), "form": "plain_fp64",
        },
        "roofline": roofline,
        "path_hbm_frac_algorithmic": round(value * b_alg / 1e9 / HBM_PEAK_GBS, 5),
        "alg_bytes_per_zone_update": b_alg,
        # the plain form's HBM target (>= 60 % of 8 TB/s), read on counters
        "path_hbm_frac_counters": round(path_counter, 4) if path_counter is not None else None,
        "kernel_share": {s["name"]: round(s["ms"] / tot, 4) for s in sorted(stats, key=lambda s: -s["ms"])},
        "dominant_kernel": dom,
        "forms": forms,
    }
    if accuracy is not None:
        line["accuracy"] = accuracy
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
