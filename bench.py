#!/usr/bin/env python3
"""Benchmark: packets/s of one nonlinear iteration (transport pass + cell
update) and time to a converged model, FP32-mixed vs plain FP64, on one
MI355X.

A step is one iteration of src/main.f90:128-185 of the reference: a full
transport pass of all packets (launch, walk, tallies, rates) followed by the
per-leaf cell update.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  Extra fields:
  forms        per-form packets/s, walk-kernel time, time to converged model
  roofline     walk kernel: algorithmic bytes (segments x bytes/segment)
               over the kernel's event-timed duration, against 8 TB/s
  cpu_baseline the FP64 CPU oracle timed on the host cores: whole passes of
               the same workload on the converged state, repeated for
               about 10 s

  python bench.py --trace-summary KERNEL_TRACE.csv [--steps K]
    condenses a rocprofv3 --kernel-trace CSV of a bench run into per-kernel
    statistics over all dispatches and over the last K dispatches of each
    kernel (the timed steps of the form that launches it).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per scored path segment, H/He on a Cartesian grid
# (d = dust terms, n_tr = metal transitions, both 0 here):
#   FP64 form:  8*(4 + d + n_tr) state reads + 16*(7 + d + 2 n_tr) tally RMWs
#   mixed form: 4*(4 + d + n_tr) state reads (FP32 opacity record) + same RMWs
BYTES_PER_SEGMENT = {"fp64": 8 * 4 + 16 * 7, "mixed": 4 * 4 + 16 * 7}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--packets", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--trace-summary", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    return ap.parse_args()


def converge(make_model, log):
    """Time to converged model: the full iteration to conv_crit='vol' plus the
    final consistency pass (src/main.f90:124-227)."""
    from mochii_amd import run_model
    m = make_model()
    t0 = time.perf_counter()
    res = run_model(m)
    dt = time.perf_counter() - t0
    log(f"  converged={res['converged']} after {res['iterations']} iterations in {dt:.3f} s")
    return m, res, dt


def trace_summary(path, last):
    """Per-kernel duration statistics from a rocprofv3 kernel-trace CSV."""
    import csv
    import numpy as np
    per = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            d = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
            per.setdefault(row["Kernel_Name"], []).append((int(row["Start_Timestamp"]), d))
    total = sum(d for v in per.values() for _, d in v)
    rows = sorted(per.items(), key=lambda kv: -sum(d for _, d in kv[1]))
    print(f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'pct':>6s} {'last':>5s} {'avg_us_last':>12s}")
    for name, v in rows:
        v.sort()
        d = np.array([x for _, x in v], dtype=np.float64)
        tail = d[-last:]
        print(f"{name[:70]:70s} {d.size:6d} {d.mean() / 1e3:10.1f} {100 * d.sum() / total:6.2f} "
              f"{tail.size:5d} {tail.mean() / 1e3:12.1f}")


def main():
    args = parse()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.steps)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the pass would shard packets over GPUs and all-reduce the tallies)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import mochii_amd as mc
    from mochii_amd.workloads import stromgren

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    n, N = args.grid, args.packets
    params, nH = stromgren(n=n, nphotons=N)

    def make_model(prec=mc.MC_MIXED):
        m = mc.Model(dict(params, precision=prec))
        m.gas_state_setup(nH)
        return m

    forms = {}
    conv_state = {}
    for name, prec in (("fp64", mc.MC_FP64), ("mixed", mc.MC_MIXED)):
        log(f"[{name}] converging the model")
        if not args.no_converge:
            m, res, t_conv = converge(lambda: make_model(prec), log)
            conv_state[name] = m.state("xHI")["xHI"]
        else:
            m, res, t_conv = make_model(prec), {"iterations": 0, "converged": False}, None
            for _ in range(10):
                m.ionizing_field_and_rates()
                m.gas_equilibrium_update()
        # timed steps on the (converged) model: transport pass + cell update
        for _ in range(args.warmup):
            m.ionizing_field_and_rates()
            m.gas_equilibrium_update()
        walk_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m.ionizing_field_and_rates()
            t = m.timing()
            walk_ms.append(t["walk_ms"])
            m.gas_equilibrium_update()
        dt = time.perf_counter() - t0
        tal = m.tallies()
        nseg = tal["seg_direct"] + tal["seg_scattered"]
        wms = float(np.mean(walk_ms))
        forms[name] = dict(packets_per_s=N * args.steps / dt, ms_per_step=1e3 * dt / args.steps,
                           walk_ms=wms, segments_per_pass=nseg,
                           time_to_converged_s=t_conv, iterations=res["iterations"],
                           converged=res["converged"],
                           walk_GBs=nseg * BYTES_PER_SEGMENT[name] / (wms * 1e-3) / 1e9)
        log(f"[{name}] {forms[name]}")
        state_for_cpu = m.state()
        m.close()

    err = None
    if "fp64" in conv_state and "mixed" in conv_state:
        a, b = 1 - conv_state["mixed"], 1 - conv_state["fp64"]
        ion = b > 1e-3
        err = dict(max_rel_xHII_ionised=float(np.max(np.abs(a[ion] / b[ion] - 1))),
                   max_abs_xHII_neutral=float(np.max(np.abs(a[~ion] - b[~ion]))) if (~ion).any() else 0.0)

    cpu = None
    if not args.no_cpu:
        from tests.cases import oracle_model
        ncpu = min(os.cpu_count() or 1, 16)
        orc = oracle_model(params, nH)
        orc.set_state(xHI=state_for_cpu["xHI"], xHeI=state_for_cpu["xHeI"],
                      xHeII=state_for_cpu["xHeII"])
        npass = 0
        t0 = time.perf_counter()
        while True:
            orc.transport_pass(1, N + 1, ncpu)
            orc.rates_compute()
            orc.equilibrium_update()
            npass += 1
            dtc = time.perf_counter() - t0
            if dtc >= args.cpu_seconds or npass >= 50:
                break
        cpu = {"value": npass * N / dtc, "unit": "packets/s", "cores": ncpu, "kind": "port",
               "sample": f"FP64 C oracle, {npass} whole steps (pass of {N} packets + rates + "
                         f"cell update) on the GPU-converged state, {ncpu} threads, "
                         f"{dtc:.1f} s"}
        log(f"[cpu] {cpu}")

    fm = forms["mixed"]
    line = {
        "metric": "packets/s + time-to-converged model, surrogate+FP32 vs plain FP64, "
                  "<=1% T_e/x_HII",
        "value": fm["packets_per_s"],
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"Stromgren sphere, pure H, uniform car {n}^3, box 8 pc, "
                               f"nH=100, 40000 K Q(H)=1e49, case B, T=1e4 K, "
                               f"{N:.0e} Sobol packets/iteration",
                   "grid": f"{n}^3", "packets_per_iteration": N},
        "forms": forms,
        "error_mixed_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": fm["walk_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fm["walk_GBs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_walk<MC_MIXED, kCarDDA>",
                     "bytes_per_segment": BYTES_PER_SEGMENT["mixed"],
                     "segments_per_launch": fm["segments_per_pass"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
