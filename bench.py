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
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counter
    readings give the scale of those counters for the traffic figure."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
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


def counter_traffic(workload, kernel_prefix):
    """HBM bytes per launch of a kernel family from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (already scaled by the
    copy-kernel calibration), or None when no summary exists."""
    path = ROOT / "profiles" / f"{workload}_hbm_traffic.txt"
    if not path.exists():
        return None
    tot = []
    for ln in path.read_text().splitlines():
        if ln.startswith(kernel_prefix):
            tot.append(float(ln.split()[-1]) * 1e6)
    if not tot:
        return None
    return {"bytes_per_launch": round(sum(tot) / len(tot)), "source": str(path.relative_to(ROOT))}


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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the block set is a single-GPU "
              f"shard and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    from sparkhip import capi

    prob = make_problem(args.workload)
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
    achieved = fl_bytes / (fl_ms * 1e-3) / 1e9
    traffic = counter_traffic(args.workload, "k_flux<")
    roofline = {
        "bound": "hbm",
        "kernel": "k_flux (x, y, z sweeps: WENO5-Z + HLLC face fluxes)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic["bytes_per_launch"] if traffic else None,
        "traffic_source": traffic["source"] if traffic else None,
        "avg_launch_ms": round(fl_ms / max(1, fl_launch), 4),
        "alg_bytes_per_launch": round(fl_bytes / max(1, fl_launch)),
        "valu_fp64": {"achieved_tflops": round(fl_flops / (fl_ms * 1e-3) / 1e12, 2), "peak_tflops": FP64_PEAK_TFS,
                      "frac": round(fl_flops / (fl_ms * 1e-3) / 1e12 / FP64_PEAK_TFS, 4)},
    }
    step_bytes = 384.0 * prob.zones  # algorithmic FP64 state traffic per zone-update
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
        "path_hbm_frac_algorithmic": round(value * step_bytes / prob.zones / 1e9 / HBM_PEAK_GBS, 5),
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
