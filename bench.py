#!/usr/bin/env python3
"""Benchmark: packets/s and time to a converged model on one MI355X, the
surrogate + FP32 form against the plain all-FP64 form.

Default workload: H + He H II region with thermal balance on an adaptive
octree (n_H lognormal, mean 100 cm^-3, sigma_ln 0.5, from a seeded k^-11/3
Gaussian random field on 128^3; at most 7 levels, refined where the density
varies by more than 0.3 dex; He/H = 0.1 and the trace metals C N O Ne S Ar;
one 45000 K star with Q(H) = 3e49 s^-1; explicit diffuse recombination
packets, case A; 1e7 stellar packets per iteration).
`--workload stromgren` runs the pure-H Stromgren sphere on a 64^3 car grid
(fixed T, 1e6 packets per iteration) instead.

A step is one iteration of src/main.f90:128-185 of the reference: a full
transport pass (stellar + diffuse packets: launch, walk, tallies, rates) and
the per-leaf cell update (thermal balance, or fixed-T balance for the
Stromgren workload).  Inputs are resident in HBM before the timed region.

Forms: "fp64"      all-FP64 transport, the reference's bisection solve;
       "mixed_mlp" FP32 cell-local geometry and opacity (FP64 tallies and
                   solves) and the learned warm start of the thermal solve
                   (network trained on the GPU in an untimed set-up phase on
                   cell states of smaller models of the same physics).

Prints ONE JSON line.  Extra fields:
  forms        per form: packets/s, walk-kernel and solve times, time to the
               converged model, iterations, solve work
  error        per-cell error of the mixed_mlp converged model against the
               fp64 one: T_e relative; x_HII relative where x_HII > 1e-3,
               absolute elsewhere
  roofline     the walk kernel of the value's form: algorithmic bytes per
               scored segment x segments over its event-timed duration
  cpu_baseline the FP64 CPU oracle on the host cores on a bounded sample of
               the same pass (~10 s)
  ir_band      the dust thermal-infrared band on the converged model with
               grains added (dust-to-gas 0.01, power-law extinction to the
               far infrared, scattering on): Heat_dust of one ionizing pass,
               then transport_dust_emission, 2 passes of 1e7 packets, per
               form: packets/s, walk time, segments, reabsorbed fraction,
               closure, and the infrared walk's algorithmic bandwidth

  python bench.py --trace-summary KERNEL_TRACE.csv [--steps K] [--skip-last S]
    condenses a rocprofv3 --kernel-trace CSV of a bench run into per-kernel
    statistics over all dispatches and over the last K dispatches of each
    kernel before its final S (for the walk of a default run, K = 5 x steps
    and S = 10 select the timed steps of the form that launches it).

  python bench.py --copy-calibration ... runs a plain device copy of 1 GiB
    (3 times) before the benchmark, so that a rocprofv3 --pmc run of
    FETCH_SIZE / WRITE_SIZE can be scaled by what the counter reads for a
    copy of known size; --pmc-summary COUNTER_CSV [--kernel NAME] then
    prints the per-dispatch counter means of the copy and of the kernel.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per scored path segment (d = dust, 0 here; n_tr = metal
# transitions below the packet's energy, counted per segment by the walk):
# state reads + 7 FP64 tally read-modify-writes + 2 per metal transition
#   car FP64:  8 (4 + d + n_tr) + 16 (7 + d + 2 n_tr)           = 144 + 40 n_tr
#   car mixed: 4 (4 + d + n_tr) (FP32 opacity record) + 112      = 128 + 36 n_tr
#   octree:    + leaf geometry per segment: FP64 centre + half + neighbour
#              and child links (44 B), FP32 centre + half + links (24 B)
BYTES_PER_SEGMENT = {("car", "fp64"): 144, ("car", "mixed_mlp"): 128,
                     ("octree", "fp64"): 188, ("octree", "mixed_mlp"): 152}
BYTES_PER_METAL_TR = {"fp64": 40, "mixed_mlp": 36}
# the mixed-form metal walk's own tally bytes: MOMENTS FP64 energy moments of
# the packet's sub-bin (read-modify-write 16 B each) replace the 7 H/He + 2
# per metal transition coefficient sums; its reads are those of the record
MOMENTS = 4
# the infrared walk per scored segment: leaf geometry (octree FP64 44 B /
# FP32 24 B; car 0) + rhokap 8 B + one FP64 path-tally read-modify-write 16 B
IR_BYTES_PER_SEGMENT = {("car", "fp64"): 24, ("car", "mixed"): 24,
                        ("octree", "fp64"): 68, ("octree", "mixed"): 48}
PC2CM = 3.0856776e18
WALK_KERNEL = {("car", "fp64"): "k_walk<MC_FP64, kCarDDA, M>",
               ("car", "mixed_mlp"): "k_walk<MC_MIXED, kCarDDA, M>",
               ("octree", "fp64"): "k_walk<MC_FP64, kFaceExit, M>",
               ("octree", "mixed_mlp"): "k_walk<MC_MIXED, kFaceExit, M>"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("hii_octree", "stromgren"), default="hii_octree")
    ap.add_argument("--packets", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--trace-summary", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--ir-packets", type=int, default=10_000_000)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--pmc-summary", default=None)
    ap.add_argument("--kernel", default="k_walk<1, 1, 1>")
    ap.add_argument("--per-pass", type=int, default=5)
    ap.add_argument("--skip-last", type=int, default=0)
    ap.add_argument("--no-ir", action="store_true")
    return ap.parse_args()


def trace_summary(path, last, skip=0):
    """Per-kernel duration statistics from a rocprofv3 kernel-trace CSV; the
    'last' columns are the last K dispatches of each kernel before its final
    `skip` ones (a default bench run ends with one pass of each form from a
    common state and one ionizing pass of each form for the infrared leg,
    5 walk dispatches each, after the timed steps)."""
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
        tail = d[:d.size - skip][-last:] if d.size > skip else d[-last:]
        print(f"{name[:70]:70s} {d.size:6d} {d.mean() / 1e3:10.1f} {100 * d.sum() / total:6.2f} "
              f"{tail.size:5d} {tail.mean() / 1e3:12.1f}")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


COPY_BYTES = 1 << 30


def copy_calibration():
    """A plain device-to-device copy of COPY_BYTES, three times: the
    reference reading for the HBM byte counters."""
    import torch
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def pmc_summary(path, kernel, per_pass=5):
    """HBM bytes of the named kernel from a rocprofv3 --pmc run of
    FETCH_SIZE or WRITE_SIZE (KiB per dispatch): the counter of the three
    calibration copies of COPY_BYTES gives the scale, the last per_pass
    dispatches of the kernel are one transport pass."""
    import csv
    import numpy as np
    rows = list(csv.DictReader(open(path)))
    ctr = rows[0]["Counter_Name"]
    cp = sorted((float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]),
                reverse=True)[:3]
    expect = COPY_BYTES / 1024.0                      # read once, written once
    scale = expect / float(np.mean(cp))
    v = [float(r["Counter_Value"]) for r in rows if kernel in r["Kernel_Name"]]
    last = v[-per_pass:]
    total = sum(last) * 1024.0 * scale
    print(f"counter {ctr}: calibration copy of {COPY_BYTES} B reads {np.round(cp, 1)} KiB "
          f"(expected {expect:.0f}), scale {scale:.4f}")
    print(f"{kernel}: {len(v)} dispatches; last {per_pass} (one pass) KiB {np.round(last)}")
    print(f"{ctr.lower()}_bytes_per_pass: {total:.6e}")


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_walk_hii_octree_metals_1e7_traffic.md")


def measured_traffic(workload):
    """HBM bytes per pass of the walk kernel, from the committed summary of
    the two counter runs (FETCH_SIZE, WRITE_SIZE) of this workload."""
    import re
    if workload != "hii_octree" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    m = re.search(r"^traffic_bytes_per_pass:\s*([0-9.eE+-]+)", open(TRAFFIC_PROFILE).read(), re.M)
    return float(m.group(1)) if m else None


def build_workload(name, packets):
    from mochii_amd.workloads import hii_octree, stromgren
    if name == "hii_octree":
        N = packets or 10_000_000
        params, nH, octree = hii_octree(nphotons=N)
        desc = (f"H+He H II region with thermal balance, octree of {nH.size} leaves (<=7 levels, "
                f"0.3 dex) over a lognormal n_H field (mean 100, sigma_ln 0.5, k^-11/3, 128^3), "
                f"box 16 pc, 45000 K Q(H)=3e49, He/H=0.1, trace metals C N O Ne S Ar "
                f"(15 transitions), explicit diffuse recombination packets (case A), "
                f"{N:.0e} stellar packets/iteration")
        return params, nH, octree, "octree", desc
    N = packets or 1_000_000
    params, nH = stromgren(n=64, nphotons=N)
    desc = (f"Stromgren sphere, pure H, uniform car 64^3, box 8 pc, nH=100, 40000 K Q(H)=1e49, "
            f"case B, T=1e4 K, {N:.0e} Sobol packets/iteration")
    return params, nH, None, "car", desc


def make_model(params, nH, octree, precision, warm=None, mlp=None):
    import mochii_amd as mc
    from mochii_amd.workloads import HI_COOLING_FIT
    from mochii_amd.workloads import elements_of
    m = mc.Model(dict(params, precision=precision))
    if octree is not None:
        m.grid_octree(*octree)
    els = elements_of(params)
    if els:
        m.species_setup(els)
    m.gas_state_setup(nH)
    if params.get("solve_te"):
        m.cooling_setup(HI_COOLING_FIT)
        if params.get("cooling_model") == mc.lib.MC_COOL_LOCAL_NE:
            m.nlevel_setup()
        if mlp is not None:
            m.mlp_set(mlp)
        if warm is not None:
            m.set_warm_start(warm)
    return m


def train_surrogate():
    """Train the warm-start network on cell states of smaller models of the
    same physics (a 32^3 car H II region and a 64^3-field octree)."""
    import numpy as np
    import mochii_amd as mc
    from mochii_amd import surrogate
    from mochii_amd.workloads import HI_COOLING_FIT, hii_octree, hii_thermal
    t0 = time.perf_counter()
    Xs, Ys = [], []
    p, nH = hii_thermal(n=32, nphotons=300_000, diffuse=True, seed=99, metals=True)
    m = make_model(p, nH, None, mc.MC_FP64)
    X, Y = surrogate.collect(m, 40)
    Xs.append(X)
    Ys.append(Y)
    m.close()
    p, nH, octree = hii_octree(n_field=64, nphotons=1_000_000, field_seed=7, seed=98, lmax=6)
    m = make_model(p, nH, octree, mc.MC_MIXED)
    X, Y = surrogate.collect(m, 40)
    Xs.append(X)
    Ys.append(Y)
    m.close()
    t_collect = time.perf_counter() - t0
    X, Y = surrogate.balance(np.concatenate(Xs), np.concatenate(Ys), max_rows=400_000)
    tm = mc.Model(dict(p))
    P, hist = surrogate.train(tm, X, Y, schedule=((3000, 2e-3), (2000, 4e-4)), batch=2048,
                              log=log)
    tm.close()
    t_total = time.perf_counter() - t0
    return P, dict(rows=int(X.shape[0]), collect_s=t_collect, train_s=t_total - t_collect,
                   final_loss=float(hist[-200:].mean()))


def run_form(name, params, nH, octree, args, mlp=None):
    import numpy as np
    import mochii_amd as mc
    from mochii_amd import run_model
    prec = mc.MC_FP64 if name == "fp64" else mc.MC_MIXED
    warm = (mc.lib.MC_WARM_NONE if name == "fp64" else
            mc.lib.MC_WARM_MLP if mlp is not None else mc.lib.MC_WARM_NEWTON)
    thermal = bool(params.get("solve_te"))
    m = make_model(params, nH, octree, prec, warm if thermal else None, mlp if thermal else None)

    def update():
        return m.gas_thermal_update() if thermal else m.gas_equilibrium_update()

    res, t_conv = {"iterations": 0, "converged": False}, None
    if not args.no_converge:
        log(f"[{name}] converging the model")
        t0 = time.perf_counter()
        res = run_model(m, log=lambda msg: log('  ' + msg))
        t_conv = time.perf_counter() - t0
        log(f"  converged={res['converged']} after {res['iterations']} iterations in {t_conv:.2f} s")
    else:
        for _ in range(10):
            m.ionizing_field_and_rates()
            update()
    state = m.state()
    for _ in range(args.warmup):
        m.ionizing_field_and_rates()
        update()
    walk_ms, solve_ms, npk, nseg, nmet, evals, brk = [], [], [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.ionizing_field_and_rates()
        tm = m.timing()
        walk_ms.append(tm["walk_ms"])
        npk.append(params["nphotons"] + m.diffuse_info()[1])
        nseg.append(sum(m.segment_counts()))
        nmet.append(m.metal_segment_transitions())
        update()
        solve_ms.append(m.timing()["solve_ms"])
        if thermal:
            st = m.thermal_stats()
            evals.append(st["evaluations"])
            brk.append(st["bracketed"])
    dt = time.perf_counter() - t0
    out = dict(packets_per_s=float(np.sum(npk) / dt), ms_per_step=1e3 * dt / args.steps,
               stellar_packets=int(params["nphotons"]),
               diffuse_packets=int(np.mean(npk)) - int(params["nphotons"]),
               walk_ms=float(np.mean(walk_ms)), solve_ms=float(np.mean(solve_ms)),
               segments_per_pass=int(np.mean(nseg)),
               metal_transitions_per_segment=float(np.sum(nmet) / max(np.sum(nseg), 1)),
               time_to_converged_s=t_conv,
               iterations=res["iterations"], converged=res["converged"])
    if thermal:
        out["net_rate_evals_per_bracketed_leaf"] = float(np.sum(evals) / max(np.sum(brk), 1))
    if prec == mc.MC_MIXED and params.get("metal_eth") is not None:
        nsub, ferr = m.tally_moment_info()
        out["moment_sub_bins"], out["moment_fit_err"] = int(nsub), float(ferr)
    log(f"[{name}] {out}")
    m.close()
    return out, state


def cell_errors(a, b):
    """Per-cell error of state a against state b."""
    import numpy as np
    xa, xb = 1 - a["xHI"], 1 - b["xHI"]
    ion = xb > 1e-3
    rel_x = np.abs(xa[ion] / xb[ion] - 1)
    abs_x = np.abs(xa[~ion] - xb[~ion])
    out = dict(max_rel_xHII_ionised=float(rel_x.max()) if rel_x.size else 0.0,
               p99_rel_xHII_ionised=float(np.percentile(rel_x, 99)) if rel_x.size else 0.0,
               max_abs_xHII_neutral=float(abs_x.max()) if abs_x.size else 0.0,
               frac_cells_xHII_within_1pct=float((np.sum(rel_x <= 0.01) + np.sum(abs_x <= 0.01))
                                                 / xa.size))
    rel_T = np.abs(a["Te"] / b["Te"] - 1)
    out.update(max_rel_Te=float(rel_T.max()), p99_rel_Te=float(np.percentile(rel_T, 99)),
               frac_cells_Te_within_1pct=float(np.mean(rel_T <= 0.01)))
    return out


def one_step(name, params, nH, octree, state, mlp):
    """One iteration (pass + cell update) of a form from a given state."""
    import mochii_amd as mc
    prec = mc.MC_FP64 if name == "fp64" else mc.MC_MIXED
    warm = (mc.lib.MC_WARM_NONE if name == "fp64" else
            mc.lib.MC_WARM_MLP if mlp is not None else mc.lib.MC_WARM_NEWTON)
    thermal = bool(params.get("solve_te"))
    m = make_model(params, nH, octree, prec, warm if thermal else None, mlp if thermal else None)
    m.set_state(**{k: state[k] for k in ("xHI", "xHeI", "xHeII", "ne", "Te")})
    m.gas_opacity_fill()
    m.ionizing_field_and_rates()
    m.gas_thermal_update() if thermal else m.gas_equilibrium_update()
    out = m.state()
    m.close()
    return out


def ir_band(params, nH, octree, grid, state, packets):
    """The dust thermal-infrared band (transport_dust_emission) on the
    converged model with grains added, in both precision forms."""
    import numpy as np
    import mochii_amd as mc
    from mochii_amd.workloads import elements_of, synthetic_kext
    tab = synthetic_kext(lam_max=1.0e4)
    out = {"workload": f"converged model + grains: dust-to-gas 0.01 (rhokap = n_H 4.868e-22 "
                       f"cm^2 d2cm), C_ext ~ lambda^-1.2 to 1e4 um, albedo 0.4, g 0.6, "
                       f"scattering on; Heat_dust of one ionizing pass; 2 passes of "
                       f"{packets:.0e} packets",
           "packets_per_pass": int(packets)}
    for name, prec in (("fp64", mc.MC_FP64), ("mixed", mc.MC_MIXED)):
        p = dict(params, precision=prec, ion_add_dust=1, ion_dust_scatter=1, lambda_ref=0.55,
                 dust_no_photons=float(packets), dust_niter=2, dust_tol=1e-12)
        m = mc.Model(p)
        if octree is not None:
            m.grid_octree(*octree)
        m.ion_dust_table(*tab)
        els = elements_of(params)
        if els:
            m.species_setup(els)
        m.gas_state_setup(nH)
        m.set_state(rhokap=nH * 4.868e-22 * PC2CM,
                    **{k: state[k] for k in ("xHI", "xHeI", "xHeII", "ne", "Te")})
        m.ionizing_field_and_rates()
        m.dust_temp_setup()
        m.dust_temp_compute()
        m.dust_emis_transport_setup()
        m.transport_dust_emission()                  # warm-up (same work)
        t0 = time.perf_counter()
        rep = m.transport_dust_emission()
        dt = time.perf_counter() - t0
        tm = m.dust_ir_timing()
        nseg = tm["seg_direct"] + tm["seg_scattered"]     # last pass
        walk_pass = tm["walk_ms"] / rep["passes"]
        bps = IR_BYTES_PER_SEGMENT[(grid, name)]
        out[name] = dict(packets_per_s=rep["passes"] * packets / dt, s_per_call=dt,
                         walk_ms_per_pass=walk_pass,
                         launch_ms_per_pass=tm["launch_ms"] / rep["passes"],
                         other_ms=tm["other_ms"], segments_last_pass=int(nseg),
                         scattered_segments=int(tm["seg_scattered"]),
                         reabsorbed_fraction=rep["L_abs_path"] / rep["L_emit"],
                         closure=rep["closure"], drel=rep["drel"],
                         t_dust_max=float(np.max(m.dust("t_dust"))),
                         walk_GBps=nseg * bps / (walk_pass * 1e-3) / 1e9, bytes_per_segment=bps)
        log(f"[ir {name}] {out[name]}")
        m.close()
    return out


def cpu_baseline(params, nH, octree, state, seconds):
    """The FP64 oracle's transport pass on the host cores over a bounded
    sample of the stellar packets of one pass on the converged state."""
    from tests.cases import oracle_model
    ncpu = min(os.cpu_count() or 1, 16)
    orc = oracle_model(dict(params, diffuse_field=0), nH, octree=octree)
    orc.set_state(**{k: state[k] for k in ("xHI", "xHeI", "xHeII", "ne", "Te")})
    orc.opacity_fill()                   # stage cache of the metal opacity
    N = params["nphotons"]
    n, dt = min(20_000, N), 0.0
    while True:                          # grow the sample until it takes ~seconds
        t0 = time.perf_counter()
        orc.transport_pass(1, n + 1, ncpu)
        dt = time.perf_counter() - t0
        if dt >= 0.7 * seconds or n >= N:
            break
        n = int(min(N, max(2 * n, n * seconds / max(dt, 1e-3))))
    return {"value": n / dt, "unit": "packets/s", "cores": ncpu, "kind": "port",
            "sample": f"FP64 C oracle transport pass over the first {n} of the {N} stellar "
                      f"packet ids on the GPU-converged state (diffuse packets off), {ncpu} "
                      f"threads, {dt:.1f} s"}


def main():
    args = parse()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.steps, args.skip_last)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.kernel, args.per_pass)
        return
    if args.copy_calibration:
        copy_calibration()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the pass would shard packets over GPUs and all-reduce the tallies)",
              file=sys.stderr)
        sys.exit(2)

    params, nH, octree, grid, desc = build_workload(args.workload, args.packets)
    thermal = bool(params.get("solve_te"))
    mlp, train_info = None, None
    if thermal and not args.no_surrogate:
        log("[surrogate] collecting cell states and training the warm-start network")
        mlp, train_info = train_surrogate()
        log(f"  {train_info}")

    forms, states = {}, {}
    for name in ("fp64", "mixed_mlp"):
        forms[name], states[name] = run_form(name, params, nH, octree, args, mlp)
    # form error: one iteration of each form from the same (fp64-converged)
    # state, so that only the form differs; the converged models themselves
    # also differ by the iteration's cell-level Monte Carlo noise (front
    # leaves that flip between neutral and ionised), reported beside it
    err = None
    if not args.no_converge:
        s64 = one_step("fp64", params, nH, octree, states["fp64"], mlp)
        smx = one_step("mixed_mlp", params, nH, octree, states["fp64"], mlp)
        err = {"one_step_from_common_state": cell_errors(smx, s64),
               "converged_models": cell_errors(states["mixed_mlp"], states["fp64"])}
    fm, f64 = forms["mixed_mlp"], forms["fp64"]
    speed = None
    if f64["time_to_converged_s"] and fm["time_to_converged_s"]:
        speed = f64["time_to_converged_s"] / fm["time_to_converged_s"]

    ir = None
    if not args.no_ir:
        log("[ir] dust thermal-infrared band on the converged model")
        ir = ir_band(params, nH, octree, grid, states["fp64"], args.ir_packets)
    cpu = None if args.no_cpu else cpu_baseline(params, nH, octree, states["mixed_mlp"],
                                                args.cpu_seconds)
    if cpu:
        log(f"[cpu] {cpu}")

    # algorithmic bytes per segment: the H/He record plus the metal weights
    # and raw tallies of the transitions below each packet's energy
    bps = (BYTES_PER_SEGMENT[(grid, "mixed_mlp")] +
           BYTES_PER_METAL_TR["mixed_mlp"] * fm["metal_transitions_per_segment"])
    achieved = fm["segments_per_pass"] * bps / (fm["walk_ms"] * 1e-3) / 1e9
    form_bps = None
    if fm.get("moment_sub_bins"):
        form_bps = (BYTES_PER_SEGMENT[(grid, "mixed_mlp")] - 16 * 7 + 16 * MOMENTS +
                    4 * fm["metal_transitions_per_segment"])
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
        "config": {"workload": desc, "leaves": int(nH.size),
                   "packets_per_iteration": int(params["nphotons"])},
        "forms": forms,
        "time_to_converged_speedup_vs_fp64": speed,
        "surrogate_training": train_info,
        "error_mixed_mlp_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload),
                     "traffic_basis": "HBM bytes per pass (all 5 walk launches), rocprofv3 "
                                      "FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy "
                                      "(profiles/k_walk_hii_octree_metals_1e7_traffic.md)",
                     "kernel": WALK_KERNEL[(grid, "mixed_mlp")].replace(
                         "M>", "1>" if params.get("metal_eth") else "0>"),
                     "bytes_per_segment": bps,
                     "form_bytes_per_segment": form_bps,
                     "form_achieved": (None if form_bps is None else
                                       fm["segments_per_pass"] * form_bps /
                                       (fm["walk_ms"] * 1e-3) / 1e9),
                     "metal_transitions_per_segment": fm["metal_transitions_per_segment"],
                     "segments_per_pass": fm["segments_per_pass"],
                     "walk_ms_per_pass": fm["walk_ms"]},
        "cpu_baseline": cpu,
        "ir_band": ir,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
