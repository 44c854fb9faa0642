#!/usr/bin/env python3
"""Benchmark: packets/s and time to a converged model on one MI355X, the
surrogate + FP32 form against the plain all-FP64 form.

Headline workload (the value): the large AMR model -- a 10-level octree of
~2.0e7 leaves refined on a seeded multifractal density (3-level lognormal
cascade, mean n_H 10 cm^-3, dynamic range ~1e5) over a 100 pc box, 16 point
sources (30000-55000 K) at seeded positions plus an isotropic external field
through the bounding sphere, H + He with thermal balance, dust with HG
scattering and the dust thermal-infrared band, 1e9 packets per iteration.
A step is one iteration of src/main.f90:128-185 of the reference: the full
transport pass (launch, walk, tallies, rates) and the per-leaf thermal
update.  W untimed warm-up steps, then K timed steps per form; inputs are
resident in HBM before the timed region.

Forms: "fp64"      all-FP64 transport, the reference's bisection solve;
       "mixed_mlp" FP32 cell-local geometry and opacity (FP64 tallies and
                   solves) and the learned warm start of the thermal solve
                   (network trained on the GPU, untimed, on cell states of
                   smaller models of the same physics).

The time to a converged model of the large model does not fit one default
run (~50 iterations of 1e9 packets per form); it is measured by
`--converge-large multifractal` into profiles/ and read from there.  The
default run converges the octree H II region with metals (7 levels, 3e5
leaves, 1e7 packets) in both forms, with a second FP64 run at another Sobol
seed as the Monte Carlo noise floor and ablation rows, and times the 256^3
turbulent cloud with metals (1e8 packets) in both forms.

Prints ONE JSON line (extra fields: forms, converge, turbulent, roofline,
cpu_baseline, ir_band).

  python bench.py --trace-summary KERNEL_TRACE.csv [--steps K] [--skip-last S]
  python bench.py --copy-calibration ...; --pmc-summary COUNTER_CSV [--kernel NAME]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PC2CM = 3.0856776e18


def bytes_per_segment(form, octree, dust, ntr):
    """Algorithmic bytes of one scored segment: state reads 8 (4 + d + n_tr)
    (FP64) or 4 (4 + d + n_tr) (FP32 record of the mixed form), 7 + d + 2 n_tr
    FP64 tally read-modify-writes of 16 B, and 44 B of octree geometry."""
    d = 1 if dust else 0
    reads = (8 if form == "fp64" else 4) * (4 + d + ntr)
    return reads + 16 * (7 + d + 2 * ntr) + (44 if octree else 0)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=1e9)
    ap.add_argument("--ir-packets", type=float, default=1e8)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-turbulent", action="store_true")
    ap.add_argument("--no-ir", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed_mlp")
    ap.add_argument("--converge-large", default=None,
                    help="converge the large model ('multifractal') or the turbulent cloud "
                         "('turbulent') in both forms and print the result")
    ap.add_argument("--trace-summary", default=None)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--pmc-summary", default=None)
    ap.add_argument("--kernel", default="k_walk<1, 1, 0,")
    ap.add_argument("--state", choices=("start", "ionised"), default="start",
                    help="state of the timed steps: the workload's own start, or every leaf "
                         "ionised (x_HI 1e-3, He singly ionised, 1e4 K: the walk's regime late "
                         "in a converging run)")
    ap.add_argument("--row-issue", choices=("auto", "lane", "together"), default="auto",
                    help="issue form of the mixed walk's moment rows (mc_set_row_issue)")
    ap.add_argument("--per-pass", type=int, default=0,
                    help="dispatches of the kernel in one pass (0: all of them)")
    ap.add_argument("--skip-last", type=int, default=0)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---- profile helpers ----------------------------------------------------------
def trace_summary(path, last, skip=0):
    """Per-kernel duration statistics from a rocprofv3 kernel-trace CSV; the
    'last' columns are the last K dispatches of each kernel before its final
    `skip` ones."""
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


def pmc_summary(path, kernel, per_pass):
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
    expect = COPY_BYTES / 1024.0
    scale = expect / float(np.mean(cp))
    v = [float(r["Counter_Value"]) for r in rows if kernel in r["Kernel_Name"]]
    last = v[-per_pass:] if per_pass > 0 else v
    total = sum(last) * 1024.0 * scale
    print(f"counter {ctr}: calibration copy of {COPY_BYTES} B reads {np.round(cp, 1)} KiB "
          f"(expected {expect:.0f}), scale {scale:.4f}")
    print(f"{kernel}: {len(v)} dispatches; last {len(last)} (one pass)")
    print(f"{ctr.lower()}_bytes_per_pass: {total:.6e}")


def profile_value(name, key):
    """A number recorded in a committed profiles/ summary ('key: value')."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{re.escape(key)}:\s*([0-9.eE+-]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


# ---- models -------------------------------------------------------------------
def make_model(params, nH, octree=None, precision=None, warm=None, mlp=None, rhokap=None,
               dust_table=None):
    import mochii_amd as mc
    from mochii_amd.workloads import HI_COOLING_FIT, elements_of
    p = dict(params) if precision is None else dict(params, precision=precision)
    m = mc.Model(p)
    if octree is not None:
        m.grid_octree(*octree)
    if dust_table is not None:
        m.ion_dust_table(*dust_table)
    els = elements_of(params)
    if els:
        m.species_setup(els)
    m.gas_state_setup(nH)
    if rhokap is not None:
        m.set_state(rhokap=rhokap)
    if params.get("solve_te"):
        m.cooling_setup(HI_COOLING_FIT)
        if params.get("cooling_model") == mc.lib.MC_COOL_LOCAL_NE:
            m.nlevel_setup()
        if mlp is not None:
            m.mlp_set(mlp)
        if warm is not None:
            m.set_warm_start(warm)
    return m


def form_setup(name, mlp):
    import mochii_amd as mc
    prec = mc.MC_FP64 if name.startswith("fp64") else mc.MC_MIXED
    warm = {"fp64": mc.lib.MC_WARM_NONE, "mixed_bisect": mc.lib.MC_WARM_NONE,
            "mixed_newton": mc.lib.MC_WARM_NEWTON}.get(
        name, mc.lib.MC_WARM_MLP if mlp is not None else mc.lib.MC_WARM_NEWTON)
    return prec, warm


def train_surrogate():
    """Train the warm-start network on cell states of smaller models of the
    same physics (a 32^3 car H II region and a 64^3-field octree)."""
    import numpy as np
    import mochii_amd as mc
    from mochii_amd import surrogate
    from mochii_amd.workloads import hii_octree, hii_thermal
    t0 = time.perf_counter()
    Xs, Ys = [], []
    p, nH = hii_thermal(n=32, nphotons=300_000, diffuse=True, seed=99, metals=True)
    m = make_model(p, nH, precision=mc.MC_FP64)
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


# ---- the headline: timed steps of the large model ------------------------------
def timed_steps(name, params, nH, octree, rhokap, tab, args, mlp):
    """W warm-up and K timed steps (pass + thermal update) of one form."""
    import numpy as np
    prec, warm = form_setup(name, mlp)
    t0 = time.perf_counter()
    m = make_model(params, nH, octree, prec, warm, mlp, rhokap, tab)
    t_setup = time.perf_counter() - t0
    log(f"[{name}] set up in {t_setup:.1f} s ({nH.size} leaves)")
    from mochii_amd.lib import MC_ROWS_AUTO, MC_ROWS_LANE, MC_ROWS_TOGETHER
    m.set_row_issue({"auto": MC_ROWS_AUTO, "lane": MC_ROWS_LANE,
                     "together": MC_ROWS_TOGETHER}[args.row_issue])
    if args.state == "ionised":
        n = nH.size
        m.set_state(xHI=np.full(n, 1e-3), xHeI=np.full(n, 1e-3), xHeII=np.full(n, 0.9),
                    ne=nH * (1.0 - 1e-3 + 0.1 * (0.9 + 2 * 0.099)), Te=np.full(n, 1e4))
        m.gas_opacity_fill()
    for k in range(args.warmup):
        m.ionizing_field_and_rates()
        m.gas_thermal_update()
        log(f"  warm-up step {k + 1}: {m.timing()}")
    walk, solve, segs, evals, brk, step_s, clos, natom, rows = [], [], [], [], [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        m.ionizing_field_and_rates()
        tm = m.timing()
        walk.append(tm["walk_ms"])
        segs.append(m.segment_counts())
        natom.append(m.walk_atomics())
        rows.append(m.row_issue()[1])
        led = m.energy_ledger()
        clos.append(abs(led[1] + led[2] - led[0]) / led[0])
        m.gas_thermal_update()
        solve.append(m.timing()["solve_ms"])
        st = m.thermal_stats()
        evals.append(st["evaluations"])
        brk.append(st["bracketed"])
        step_s.append(time.perf_counter() - ts)
        log(f"  step {k + 1}: {step_s[-1]:.2f} s, walk {walk[-1]:.0f} ms, thermal "
            f"{solve[-1]:.0f} ms, segments {segs[-1]}")
    dt = time.perf_counter() - t0
    N = int(params["nphotons"])
    out = dict(packets_per_s=N * args.steps / dt, ms_per_step=1e3 * dt / args.steps,
               packets_per_step=N, steps=args.steps, warmup=args.warmup,
               walk_ms=float(np.mean(walk)), solve_ms=float(np.mean(solve)),
               segments_direct=int(np.mean([s[0] for s in segs])),
               segments_scattered=int(np.mean([s[1] for s in segs])),
               walk_atomics=int(np.mean(natom)),
               walk_atomics_per_s=float(np.sum(natom) / (np.sum(walk) * 1e-3)),
               max_energy_closure=float(np.max(clos)),
               net_rate_evals_per_bracketed_leaf=float(np.sum(evals) / max(np.sum(brk), 1)),
               setup_s=t_setup, moment_sub_bins=int(m.tally_moment_info()[0]),
               step_ms=[1e3 * x for x in step_s],
               # issue form of the moment rows per timed step (1 transposed,
               # 0 per lane, -1 none)
               row_issue=rows)
    return m, out


def scattered_rmw_probe(rows, width, n_ops, together=False):
    """The device's scattered FP64 read-modify-write rate: n_ops updates of
    `width` doubles at hashed rows of a rows x width array, one lane adding
    a row's doubles in turn (mc_probe_scattered_rmw: `width` atomic requests
    per row) or `width` lanes adding them in one instruction
    (mc_probe_row_atomics: one request per row)."""
    import ctypes as C
    from mochii_amd.lib import check, library
    ms = C.c_double()
    fn = library().mc_probe_row_atomics if together else library().mc_probe_scattered_rmw
    check(fn(int(rows), int(width), int(n_ops), C.byref(ms), None), "scattered-RMW probe")
    t = ms.value * 1e-3
    req = n_ops if together else n_ops * width
    return dict(rows=int(rows), width=int(width), n_ops=int(n_ops), together=bool(together),
                ms=ms.value, rows_per_s=n_ops / t, atomics_per_s=req / t,
                doubles_per_s=n_ops * width / t, bytes_per_s=n_ops * width * 8 / t)


def ir_band(m, name, packets):
    """transport_dust_emission on the model's state: two passes of packets."""
    m.dust_temp_setup()
    m.dust_temp_compute()
    m.dust_emis_transport_setup()
    t0 = time.perf_counter()
    rep = m.transport_dust_emission()
    dt = time.perf_counter() - t0
    tm = m.dust_ir_timing()
    out = dict(packets_per_s=rep["passes"] * packets / dt, s_per_call=dt, passes=rep["passes"],
               walk_ms_per_pass=tm["walk_ms"] / max(rep["passes"], 1),
               launch_ms_per_pass=tm["launch_ms"] / max(rep["passes"], 1),
               segments_last_pass=int(tm["seg_direct"] + tm["seg_scattered"]),
               reabsorbed_fraction=rep["L_abs_path"] / max(rep["L_emit"], 1e-300),
               closure=rep["closure"], drel=rep["drel"])
    log(f"[ir {name}] {out}")
    return out


# ---- the converged comparison --------------------------------------------------
def cell_errors(a, b):
    """Per-cell error of state a against state b: T_e relative; x_HII
    relative where x_HII > 1e-3 in b, absolute elsewhere; x_HeII the same."""
    import numpy as np
    out = {}
    for key, xa, xb in (("xHII", 1 - a["xHI"], 1 - b["xHI"]), ("xHeII", a["xHeII"], b["xHeII"])):
        ion = xb > 1e-3
        rel = np.abs(xa[ion] / xb[ion] - 1)
        ab = np.abs(xa[~ion] - xb[~ion])
        out[f"max_rel_{key}_ionised"] = float(rel.max()) if rel.size else 0.0
        out[f"p99_rel_{key}_ionised"] = float(np.percentile(rel, 99)) if rel.size else 0.0
        out[f"max_abs_{key}_neutral"] = float(ab.max()) if ab.size else 0.0
        out[f"frac_cells_{key}_within_1pct"] = float((np.sum(rel <= 0.01) + np.sum(ab <= 0.01))
                                                    / xa.size)
    rel_T = np.abs(a["Te"] / b["Te"] - 1)
    out.update(max_rel_Te=float(rel_T.max()), p99_rel_Te=float(np.percentile(rel_T, 99)),
               frac_cells_Te_within_1pct=float(np.mean(rel_T <= 0.01)))
    return out


def converge(name, params, nH, octree, mlp, seed=None, rhokap=None, tab=None, ir_packets=0):
    """run_model (src/main.f90:124-227) of one form to convergence; with
    ir_packets, the dust thermal-infrared band after it (main.f90:244-258),
    timed on its own."""
    import mochii_amd as mc
    from mochii_amd import run_model
    prec, warm = form_setup(name, mlp)
    p = dict(params) if seed is None else dict(params, qmc_seed=seed)
    m = make_model(p, nH, octree, prec, warm, mlp if warm == mc.lib.MC_WARM_MLP else None,
                   rhokap, tab)
    t0 = time.perf_counter()
    res = run_model(m, log=lambda s: log("  " + s) if "iteration" in s and
                    int(s.split()[1].rstrip(":")) % 10 == 0 else None)
    t = time.perf_counter() - t0
    st = m.state()
    log(f"[converge {name}{'' if seed is None else f' seed {seed}'}] {res['iterations']} "
        f"iterations, converged={res['converged']}, {t:.2f} s")
    out = dict(time_to_converged_s=t, iterations=res["iterations"],
               converged=bool(res["converged"]))
    if ir_packets:
        out["ir_band"] = ir_band(m, name, ir_packets)
        out["time_to_converged_with_ir_s"] = t + out["ir_band"]["s_per_call"]
    m.close()
    return out, st


def converge_leg(mlp, train_info):
    """The octree H II region with metals converged in each form; the noise
    floor (FP64 at another Sobol seed) and ablation rows."""
    from mochii_amd.workloads import hii_octree
    params, nH, octree = hii_octree(nphotons=10_000_000)
    rows, states = {}, {}
    for name in ("fp64", "mixed_mlp", "mixed_newton", "mixed_bisect"):
        if name == "mixed_mlp" and mlp is None:
            continue
        rows[name], states[name] = converge(name, params, nH, octree, mlp)
    rows["fp64_seed2"], states["fp64_seed2"] = converge("fp64", params, nH, octree, mlp,
                                                         seed=params["qmc_seed"] + 1)
    f64 = rows["fp64"]["time_to_converged_s"]
    best = "mixed_mlp" if "mixed_mlp" in rows else "mixed_newton"
    floor = cell_errors(states["fp64_seed2"], states["fp64"])
    err = cell_errors(states[best], states["fp64"])
    speed = f64 / rows[best]["time_to_converged_s"]
    speed_tr = None
    if train_info and best == "mixed_mlp":
        speed_tr = f64 / (rows[best]["time_to_converged_s"] + train_info["collect_s"]
                          + train_info["train_s"])
    within = (err["frac_cells_Te_within_1pct"] >= 0.99 and
              err["frac_cells_xHII_within_1pct"] >= 0.99)
    at_floor = (err["frac_cells_Te_within_1pct"] >= floor["frac_cells_Te_within_1pct"] and
                err["frac_cells_xHII_within_1pct"] >= floor["frac_cells_xHII_within_1pct"])
    return dict(workload=f"octree H II region with metals: {nH.size} leaves (<=7 levels over a "
                         f"128^3 lognormal field), C N O Ne S Ar, diffuse packets, 1e7 "
                         f"packets/iteration, cooling local_ne",
                forms=rows, speedup_vs_fp64=speed, speedup_vs_fp64_with_training=speed_tr,
                error_vs_fp64=err, noise_floor_fp64_seed2=floor,
                meets_1pct_target=bool(within), within_noise_floor=bool(at_floor),
                form_used=best)


def turbulent_leg(args, mlp):
    """The 256^3 turbulent cloud with metals: timed steps of each form."""
    from mochii_amd.workloads import turbulent_cloud
    params, nH = turbulent_cloud()
    out = {}
    a2 = argparse.Namespace(**vars(args))
    a2.steps, a2.warmup = 2, 1
    for name in ("fp64", "mixed_mlp"):
        m, out[name] = timed_steps(name, params, nH, None, None, None, a2, mlp)
        nsub, _ = m.tally_moment_info()
        out[name]["moment_sub_bins"] = int(nsub)
        m.close()
    out["step_speedup"] = out["fp64"]["ms_per_step"] / out["mixed_mlp"]["ms_per_step"]
    out["workload"] = ("turbulent cloud: car 256^3, lognormal n_H (mean 50, sigma_ln 1.0, "
                       "k^-11/3), 4 sources 35000-50000 K, Q(H) 1e48-1e49, H/He + C N O Ne S Ar, "
                       "1e8 packets/iteration")
    return out


# ---- the CPU baseline -------------------------------------------------------------
def cpu_baseline(params, nH, octree, rhokap, tab, state, seconds):
    """The FP64 oracle on the host cores, a whole step in the value's unit:
    an empty pass gives the pass's fixed cost (per-thread tally set-up and
    reduction), a sized sample of the same pass, less that fixed cost, gives
    the per-packet transport rate; the step is the full packet count at that
    rate, plus the fixed cost once, plus one thermal update over every leaf."""
    from tests.cases import oracle_model
    ncpu = os.cpu_count() or 1
    nthr = min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu)))
    t0 = time.perf_counter()
    orc = oracle_model(dict(params), nH, rhokap=rhokap, dust_table=tab, octree=octree)
    orc.set_state(**{k: state[k] for k in ("xHI", "xHeI", "xHeII", "ne", "Te")})
    from mochii_amd.workloads import HI_COOLING_FIT
    orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
    t_setup = time.perf_counter() - t0
    log(f"[cpu] oracle set up in {t_setup:.1f} s, {nthr} threads")
    N = int(params["nphotons"])
    # the pass's fixed cost (per-thread tally set-up and reduction) from an
    # empty pass, a first rate from 20000 packets, then a sample sized to
    # ~0.6 `seconds` of packet work, whose rate is the value's
    t0 = time.perf_counter()
    orc.transport_pass(1, 1, nthr)
    fixed = time.perf_counter() - t0
    n1 = 20_000
    t0 = time.perf_counter()
    orc.transport_pass(1, n1 + 1, nthr)
    t1 = time.perf_counter() - t0
    rate1 = n1 / max(t1 - fixed, 0.05)
    n2 = int(min(N, max(4 * n1, rate1 * 0.6 * seconds)))
    t0 = time.perf_counter()
    orc.transport_pass(1, n2 + 1, nthr)
    t2 = time.perf_counter() - t0
    rate = n2 / max(t2 - fixed, 1e-9)
    orc.rates_compute()
    t0 = time.perf_counter()
    orc.thermal_update(nthreads=nthr)
    t_th = time.perf_counter() - t0
    step = N / rate + fixed + t_th
    log(f"[cpu] empty pass {fixed:.1f} s, {n1} packets {t1:.1f} s, {n2} packets {t2:.1f} s -> "
        f"{rate:.3e} packets/s, thermal {t_th:.1f} s")
    return {"value": N / step, "unit": "packets/s", "cores": nthr, "kind": "port",
            "nproc": ncpu, "threads_used": nthr,
            "transport_packets_per_s": rate, "step_s_projected": step,
            "sample": f"FP64 C oracle on the same model and state: an empty transport pass "
                      f"(its fixed tally set-up and reduction, {fixed:.1f} s), a pass of the "
                      f"first {n2} of the {N} packet ids ({t2:.1f} s; less the fixed cost, the "
                      f"per-packet rate) and one thermal update of all {nH.size} leaves "
                      f"({t_th:.1f} s); a whole step projected to {N:.0e} packets, {nthr} "
                      f"threads (the box's CPU share)"}


def converge_large(which, mlp, train_info):
    """Time to a converged model of a large workload in both forms."""
    from mochii_amd.workloads import multifractal_cloud, turbulent_cloud
    if which == "turbulent":
        params, nH = turbulent_cloud()
        octree = rhokap = tab = None
    else:
        params, nH, octree, rhokap, tab = multifractal_cloud()
    rows, states = {}, {}
    irp = int(params.get("dust_no_photons", 0)) if params.get("ion_add_dust") else 0
    for name in ("mixed_mlp", "fp64"):
        rows[name], states[name] = converge(name, params, nH, octree, mlp, rhokap=rhokap, tab=tab,
                                            ir_packets=irp)
    out = dict(workload=which, leaves=int(nH.size), packets_per_iteration=int(params["nphotons"]),
               forms=rows,
               speedup_vs_fp64=rows["fp64"]["time_to_converged_s"] /
               rows["mixed_mlp"]["time_to_converged_s"],
               error_vs_fp64=cell_errors(states["mixed_mlp"], states["fp64"]),
               surrogate_training=train_info)
    if train_info:
        out["speedup_vs_fp64_with_training"] = rows["fp64"]["time_to_converged_s"] / (
            rows["mixed_mlp"]["time_to_converged_s"] + train_info["collect_s"] +
            train_info["train_s"])
    print(json.dumps(out))
    for k in ("fp64", "mixed_mlp"):
        print(f"time_to_converged_{k}_s: {rows[k]['time_to_converged_s']:.6e}")
        print(f"iterations_{k}: {rows[k]['iterations']}")
        if "time_to_converged_with_ir_s" in rows[k]:
            print(f"time_to_converged_with_ir_{k}_s: {rows[k]['time_to_converged_with_ir_s']:.6e}")
    print(f"speedup_vs_fp64: {out['speedup_vs_fp64']:.6e}")
    if "speedup_vs_fp64_with_training" in out:
        print(f"speedup_vs_fp64_with_training: {out['speedup_vs_fp64_with_training']:.6e}")
    for k, v in out["error_vs_fp64"].items():
        print(f"{k}: {v:.6e}")


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
    mlp, train_info = None, None
    if not args.no_surrogate:
        log("[surrogate] collecting cell states and training the warm-start network")
        mlp, train_info = train_surrogate()
        log(f"  {train_info}")
    if args.converge_large:
        converge_large(args.converge_large, mlp, train_info)
        return

    from mochii_amd.workloads import multifractal_cloud
    t0 = time.perf_counter()
    params, nH, octree, rhokap, tab = multifractal_cloud(nphotons=int(args.packets),
                                                         ir_packets=args.ir_packets)
    log(f"[large model] {nH.size} leaves generated in {time.perf_counter() - t0:.1f} s")
    forms, ir, state = {}, {}, None
    names = args.forms.split(",")
    head_name = "mixed_mlp" if "mixed_mlp" in names else names[0]
    for name in names:
        # the value is the head form's K timed steps; a comparison form runs
        # the same W warm-up steps and the first min(K, 5) of the timed
        # ones, which bounds the run when K is large; the forms are compared
        # on the same step indices (the model's state, and with it the step's
        # cost, changes from step to step)
        a2 = args
        if name != head_name:
            a2 = argparse.Namespace(**vars(args))
            a2.steps = min(args.steps, 5)
        m, forms[name] = timed_steps(name, params, nH, octree, rhokap, tab, a2, mlp)
        if not args.no_ir:
            ir[name] = ir_band(m, name, args.ir_packets)
        if name == "mixed_mlp" or state is None:
            state = m.state()
        m.close()
    head = forms[head_name]
    f64 = forms.get("fp64")
    # the walk's tally pattern on the device without the walk: one 4-double
    # moment row per update over the model's moment array (leaves x
    # sub-bins), single doubles over the same bytes, and 4-double rows over
    # 3.2 MB (inside one XCD's L2) and 128 MB (inside the MALL): the FP64
    # atomic rate does not depend on where the rows live
    nsub = max(head.get("moment_sub_bins", 0), 1)
    probe = [scattered_rmw_probe(nH.size * nsub, 4, 4e9),
             scattered_rmw_probe(nH.size * nsub * 4, 1, 8e9),
             scattered_rmw_probe(100_000, 4, 4e9),
             scattered_rmw_probe(4_000_000, 4, 4e9),
             scattered_rmw_probe(nH.size * nsub, 4, 4e9, together=True)]
    log(f"[probe] scattered RMW: {probe}")
    conv = None if args.no_converge else converge_leg(mlp, train_info)
    turb = None if args.no_turbulent else turbulent_leg(args, mlp)
    cpu = None if args.no_cpu else cpu_baseline(params, nH, octree, rhokap, tab, state,
                                                args.cpu_seconds)
    nseg = head["segments_direct"] + head["segments_scattered"]
    bps = bytes_per_segment("mixed", True, True, 0)
    achieved = nseg * bps / (head["walk_ms"] * 1e-3) / 1e9
    # HBM bytes per launch of the walk from the committed counter runs of
    # the same workload (one pass = dispatches_per_pass launches)
    prof = "k_walk_multifractal_1e9_traffic.md"
    fetch = profile_value(prof, "fetch_size_bytes_per_pass")
    write = profile_value(prof, "write_size_bytes_per_pass")
    ndisp = profile_value(prof, "dispatches_per_pass")
    pseg = profile_value(prof, "segments_per_pass")
    traffic = traffic_seg = None
    if None not in (fetch, write, ndisp, pseg):
        traffic = (fetch + write) / ndisp
        traffic_seg = (fetch + write) / pseg
    conv_keys = ("time_to_converged_fp64_s", "time_to_converged_mixed_mlp_s", "iterations_fp64",
                 "iterations_mixed_mlp", "speedup_vs_fp64", "speedup_vs_fp64_with_training",
                 "frac_cells_xHII_within_1pct", "frac_cells_Te_within_1pct", "p99_rel_Te",
                 "p99_rel_xHII_ionised")
    large_conv = {k: profile_value("converge_multifractal.md", k) for k in conv_keys}
    turb_conv = {k: profile_value("converge_turbulent.md", k) for k in conv_keys}
    for cv in (large_conv, turb_conv):
        if cv["frac_cells_Te_within_1pct"] is not None:
            cv["meets_1pct_target"] = bool(cv["frac_cells_Te_within_1pct"] >= 0.99 and
                                           cv["frac_cells_xHII_within_1pct"] >= 0.99)
        cv["source"] = "committed summary of python3 bench.py --converge-large (profiles/)"
    line = {
        "metric": "packets/s + time-to-converged model, surrogate+FP32 vs plain FP64, "
                  "<=1% T_e/x_HII",
        "value": head["packets_per_s"],
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "large AMR model: 10-level octree over a seeded multifractal "
                               "density (3-level lognormal cascade, mean n_H 10, dynamic range "
                               "1.2e5), 100 pc box, 16 blackbody sources 30000-55000 K + an "
                               "isotropic external field through the bounding sphere, H+He "
                               "thermal balance (local_ne cooling), dust with HG scattering and "
                               "the dust-IR band, synthetic data",
                   "leaves": int(nH.size), "packets_per_iteration": int(params["nphotons"]),
                   "ir_packets_per_pass": int(args.ir_packets),
                   "state": ("every leaf ionised" if args.state == "ionised" else
                             "the workload's start (neutral gas)"),
                   "row_issue": args.row_issue},
        "forms": forms,
        "step_speedup_vs_fp64": (sum(f64["step_ms"]) / sum(head["step_ms"][:len(f64["step_ms"])]))
                                if f64 else None,
        "step_speedup_basis": "FP64 form's timed steps against the same step indices of the "
                              "head form (both after the same warm-up)",
        "time_to_converged_large_model": large_conv,
        "time_to_converged_turbulent_cloud": turb_conv,
        "converge": conv,
        "turbulent": turb,
        "surrogate_training": train_info,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_walk<MC_MIXED, kFaceExit, 0, %s>" % (
                         "kRowsTogether" if head.get("row_issue") and
                         min(head["row_issue"]) == 1 else
                         "kRowsLane" if head.get("row_issue") and
                         max(head["row_issue"]) == 0 else "kRowsLane/kRowsTogether"),
                     "bytes_per_segment": bps, "segments_per_pass": nseg,
                     "walk_ms_per_pass": head["walk_ms"],
                     "traffic_bytes_per_segment": traffic_seg,
                     "atomics": {
                         "walk_atomics_per_pass": head["walk_atomics"],
                         "walk_atomics_per_segment": head["walk_atomics"] / max(nseg, 1),
                         "walk_atomics_per_s": head["walk_atomics_per_s"],
                         "ceiling_atomics_per_s": max(p["atomics_per_s"] for p in probe),
                         "frac_of_ceiling": head["walk_atomics_per_s"] /
                                            max(p["atomics_per_s"] for p in probe),
                         "basis": "global FP64 tally atomic requests (one per 64-byte line "
                                  "an atomic instruction touches) counted by the walk "
                                  "(mc_walk_atomics) over its event-timed duration, against "
                                  "the scattered-RMW probes' highest request rate"},
                     "scattered_rmw": {
                         "probe_rows_per_s": probe[0]["rows_per_s"],
                         "probe_atomics_per_s_single": probe[1]["atomics_per_s"],
                         "probe_atomics_per_s_3MB": probe[2]["atomics_per_s"],
                         "probe_atomics_per_s_128MB": probe[3]["atomics_per_s"],
                         "probe_rows_per_s_four_lanes": probe[4]["rows_per_s"],
                         "walk_segments_per_s": nseg / (head["walk_ms"] * 1e-3),
                         "walk_frac_of_probe": nseg / (head["walk_ms"] * 1e-3) /
                                               probe[0]["rows_per_s"],
                         "probes": probe},
                     "traffic_basis": "HBM bytes per walk launch (one pass = "
                                      f"{ndisp and int(ndisp)} launches), rocprofv3 FETCH_SIZE + "
                                      "WRITE_SIZE scaled by a 1 GiB copy "
                                      "(profiles/k_walk_multifractal_1e9_traffic.md)"},
        "cpu_baseline": cpu,
        "ir_band": ir or None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
