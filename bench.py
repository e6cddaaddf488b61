"""Throughput of the HW1F swap exposure path on one MI355X.

One step = one full exposure job on resident data: Philox4x32-10 + exact HW1F simulation of all
paths, per-path mark-to-market of every swap at every monitoring date, netting-set EE / PFE95 /
CVA, per-trade EE / CVA, plus the PV / DV01 / bucketed-DV01 ladder of every trade. The workload is
the one-netting-set portfolio: 1000 seeded synthetic swaps, 100k paths x 360 monthly dates, FP64.

Prints one JSON line (rank 0). ``--gpus`` above 1 is refused: the bench measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "trade-paths-dates/sec for HW1F swap EE/PFE/CVA (GPU vs CPU); CVA err vs FP64"
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
EXP_FLOPS = 20            # FP64 exp as range reduction + degree-~10 FMA polynomial + scaling


def algorithmic_work(port, dates, n_paths, n_sets, factors=1):
    """Algorithmic work of one value_kernel launch (the dominant kernel).

    The roofline figure is the per-unit work of the swap valuation itself: per live (trade, path,
    date) 2 m + 6 FP64 FLOPs (m fixed-leg FMAs over the trade's m remaining flows, the two floating-leg
    terms, max, accumulate), plus one exp per live cash-flow date per (path, date), counted apart.

    Returned beside it, the work this kernel executes after sharing every discount factor and annuity
    across the trades of a schedule class:
      * live (trade, path, date):        3 FLOPs   x = s F - (s K) A (one FMA; s F once per work item),
                                                   max (not a FLOP), add
      * live (class, path, date):        2 m + 2   annuity FMAs over the m remaining flows,
                                                   floating leg (sub, + div mid-period)
      * live (set, class, path, date):   4 FLOPs   V_s += alpha F - beta A (2 FMAs)
      * live (cash-flow date, path, date): exp + 2 FLOPs   P = A exp(-B r)
                                           (G2++: exp + 4, P = exp((c - B1 x) - B2 y))
    Bytes: state read (8 B per factor and path-date) + exposure write (8 B per set-path-date).

    Returns (valuation FLOPs, executed FLOPs, bytes, live trade-path-dates, exps).
    """
    tol = 1e-9
    off = port.flow_off
    # schedule classes
    keys = {}
    cls = np.empty(port.n_trades, dtype=np.int64)
    for j in range(port.n_trades):
        k = (float(port.start[j]), tuple(port.pay[off[j]:off[j + 1]]), tuple(port.acc[off[j]:off[j + 1]]))
        cls[j] = keys.setdefault(k, len(keys))
    classes = list(keys.keys())
    mats = np.array([c[1][-1] for c in classes])
    cflows = [np.array(c[1]) for c in classes]
    n_cls_trades = np.bincount(cls, minlength=len(classes))
    set_cls = len({(int(port.set_id[j]), int(cls[j])) for j in range(port.n_trades)})
    cf_dates = np.unique(np.concatenate([port.pay, port.start]))
    flops = 0.0
    u_live = 0.0
    val_flops = 0.0
    n_exp = 0.0
    for t in dates:
        live = mats > t + tol
        rem = np.array([np.sum(f > t + tol) for f in cflows])
        u_live += float(np.sum(n_cls_trades[live]))
        val_flops += float(np.sum(n_cls_trades[live] * (2.0 * rem[live] + 6.0)))
        n_exp += float(np.sum(cf_dates > t + tol))
        per_path = (3.0 * np.sum(n_cls_trades[live]) + np.sum(2.0 * rem[live] + 2.0)
                    + 4.0 * set_cls * np.mean(live) + (EXP_FLOPS + 2.0 * factors) * np.sum(cf_dates > t + tol))
        flops += per_path
    flops *= n_paths
    u_live *= n_paths
    bytes_ = 8.0 * n_paths * len(dates) * (factors + n_sets)
    return val_flops * n_paths, flops, bytes_, u_live, n_exp * n_paths


def schedule_families(port):
    """Schedule classes (distinct start / payments / accruals) grouped into families whose schedules
    are prefixes of one representative, longest first (the plan builder's grouping). Returns
    (classes as (start, pay array, acc array), class of each trade, family of each class)."""
    off = port.flow_off
    keys = {}
    cls = np.empty(port.n_trades, dtype=np.int64)
    for j in range(port.n_trades):
        k = (float(port.start[j]), tuple(port.pay[off[j]:off[j + 1]]), tuple(port.acc[off[j]:off[j + 1]]))
        cls[j] = keys.setdefault(k, len(keys))
    classes = [(k[0], np.array(k[1]), np.array(k[2])) for k in keys]
    order = sorted(range(len(classes)), key=lambda c: -len(classes[c][1]))
    reps, fam = [], np.empty(len(classes), dtype=np.int64)
    for c in order:
        s0, pay, acc = classes[c]
        for f, r in enumerate(reps):
            rs, rp, ra = classes[r]
            if rs == s0 and np.array_equal(rp[:len(pay)], pay) and np.array_equal(ra[:len(acc)], acc):
                fam[c] = f
                break
        else:
            fam[c] = len(reps)
            reps.append(c)
    return classes, cls, fam


def path_mode_work(port, dates, n_paths, factors=1):
    """Arithmetic of the class-shared valuation as value_path_kernel performs it, per (path, date):
      * each live cash-flow date: one bond, exp + 2 FLOPs (G2++: exp + 4);
      * each live family flow: one annuity FMA (2 FLOPs);
      * each (netting set, live class) entry at the class's last flow: alpha (st - P) - beta A,
        a subtraction and two FMAs (5 FLOPs);
      * each family start off the date: a forward start's bond (exp + 2); a period in progress adds
        its fixing factor (exp + 2) and the product (1).
    Returns (FLOPs, exps) of one launch. Families with no flow at a cash-flow date cost the kernel a
    zero-accrual FMA each; those are executed but not counted here."""
    tol = 1e-9
    classes, cls, fam = schedule_families(port)
    n_fam = int(fam.max()) + 1 if len(fam) else 0
    reps = {}
    for c in range(len(classes)):
        if fam[c] not in reps or len(classes[c][1]) > len(classes[reps[fam[c]]][1]):
            reps[fam[c]] = c
    entries = np.zeros(len(classes))
    for s, c in {(int(port.set_id[j]), int(cls[j])) for j in range(port.n_trades)}:
        entries[c] += 1
    mats = np.array([c[1][-1] for c in classes])
    flops = exps = 0.0
    for t in dates:
        cf = set()
        n_flows = 0
        for f in range(n_fam):
            s0, pay, _ = classes[reps[f]]
            if not (t < pay[-1] - tol):
                continue
            live = pay[pay > t + tol]
            cf.update(live.tolist())
            n_flows += len(live)
            prev = pay[pay <= t + tol]
            prev = prev[-1] if len(prev) else s0
            if t < s0 - tol:
                flops += EXP_FLOPS + 2.0
                exps += 1
            elif abs(t - s0) > tol and abs(t - prev) > tol:
                flops += 2.0 * EXP_FLOPS + 5.0
                exps += 2
        flops += len(cf) * (EXP_FLOPS + 2.0 * factors) + 2.0 * n_flows + 5.0 * entries[mats > t + tol].sum()
        exps += len(cf)
    return flops * n_paths, exps * n_paths


PROFILE = os.path.join(ROOT, "profiles", "value_kernel_hw1f_1000swaps_100kpaths_360dates.md")
PROFILE2 = os.path.join(ROOT, "profiles", "value2_kernel_hw2f_1000swaps_100kpaths_360dates.md")
PROFILEX = os.path.join(ROOT, "profiles", "xva_reducer_1000swaps_100kpaths_360dates.md")
PROFILEB = os.path.join(ROOT, "profiles", "value_path_kernel_batch_64sets_4Mpaths_120dates.md")


def profile_counters(profile=PROFILE, section="## Issue counters"):
    """The per-launch counter table of a committed rocprofv3 summary (name -> value)."""
    out = {}
    try:
        with open(profile) as f:
            inside = False
            for line in f:
                if line.startswith("## "):
                    inside = line.startswith(section)
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if inside and len(cells) == 2 and cells[0].isupper():
                    try:
                        out[cells[0]] = float(cells[1])
                    except ValueError:
                        pass
    except OSError:
        pass
    return out


def counter_fp64_flops(counters):
    """FP64 FLOPs one launch executed, from the SQ counters (wave instructions x 64 lanes; an FMA is two)."""
    keys = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64")
    if not all(k in counters for k in keys):
        return None
    return 64.0 * (2.0 * counters[keys[0]] + counters[keys[1]] + counters[keys[2]])


def measured_traffic(kernel="value_kernel", profile=PROFILE):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of the default workload (each counter scaled by what it reads for a plain 1 GiB copy)."""
    try:
        with open(profile) as f:
            in_traffic = False
            for line in f:
                if line.startswith("## "):
                    in_traffic = line.startswith("## HBM traffic")
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                name = cells[0].replace("`void ", "`")
                if in_traffic and len(cells) == 6 and (name.startswith(f"`eqhip::{kernel}(")
                                                       or name.startswith(f"`eqhip::{kernel}<")):
                    return float(cells[5]) * 1e6
    except OSError:
        pass
    return None


ORACLE_FLAGS = "gcc -O2 -fPIC -std=c99 -ffp-contract=off -fopenmp (oracle/Makefile), scalar FP64, no -march"


def host_cpu():
    """The host the CPU leg runs on: logical CPUs, the CPUs this process may use (affinity and cgroup
    quota) and the model name."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                usable = min(usable, max(1, int(int(q) // int(period))))
    except (OSError, ValueError):
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "model": model, "compiler": ORACLE_FLAGS}


def cpu_threads():
    return min(16, host_cpu()["usable_cpus"])


def _median_rate(run, units, repeats=5):
    run()  # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return units / med, med


def cpu_baseline(port, dates, n_steps, model, seed, disc, target_s=15.0):
    """The CPU FP64 restatement timed on this host's cores on bounded samples of the same job (the same
    portfolio, dates and Philox paths, drawn inside the run):

    * fused, schedule classes shared (the baseline): one simulation, every distinct zero-coupon bond once
      per (path, date), trades of one schedule sharing their floating leg and annuity, OpenMP over paths,
      PFE95 by selection; median of five runs after a warm-up;
    * fused, bonds shared only: the same without class sharing (its set values are the per-trade
      restatement's bit for bit);
    * reference structure: every trade simulated on its own with a sort per date, one thread (as the
      reference's per-calendar-group re-simulation and per-trade profiles do).
    Returns (baseline dict, oracle results of the bonds-shared sample and its path count)."""
    import oracle

    threads = cpu_threads()
    pd_ = port.as_dict()
    units_per_path = port.n_trades * len(dates)
    kw = dict(scheme=disc, hazard_rate=0.02, recovery_rate=0.4, want_trade=True, n_threads=threads, seed=seed)
    # size the six runs of the baseline to about target_s in all, from two probes (the difference leaves
    # out the run's fixed set-up)
    times = []
    for probe in (32 * threads, 256 * threads):
        t0 = time.perf_counter()
        oracle.portfolio_exposure_fused(model, pd_, dates, n_steps, n_paths=probe, share_classes=True, **kw)
        times.append(time.perf_counter() - t0)
    per_path = max(times[1] - times[0], 1e-6) / (224 * threads)
    n_cls = int(min(1 << 20, max(probe, (target_s / 6.0) / per_path)))
    rate_cls, s_cls = _median_rate(lambda: oracle.portfolio_exposure_fused(
        model, pd_, dates, n_steps, n_paths=n_cls, share_classes=True, **kw), units_per_path * n_cls)
    n_zcb = max(threads, n_cls // 4)
    t0 = time.perf_counter()
    ref = oracle.portfolio_exposure_fused(model, pd_, dates, n_steps, n_paths=n_zcb, **kw)
    s_zcb = time.perf_counter() - t0
    rate_zcb = units_per_path * n_zcb / s_zcb
    # reference structure on the first trades
    normals = oracle.philox_normals
    n_ref_trades, n_ref_paths = 16, 512
    zr = normals(seed, n_ref_paths, n_steps)
    off = port.flow_off
    t0 = time.perf_counter()
    for j in range(n_ref_trades):
        f0, f1 = off[j], off[j + 1]
        single = {"set_id": np.zeros(1, np.int32), "notional": port.notional[j:j + 1],
                  "fixed_rate": port.fixed_rate[j:j + 1], "start": port.start[j:j + 1],
                  "flow_off": np.array([0, f1 - f0], np.int32), "pay": port.pay[f0:f1], "acc": port.acc[f0:f1]}
        oracle.portfolio_exposure(model, single, dates, n_steps, zr, scheme=disc, hazard_rate=0.02,
                                  recovery_rate=0.4, n_threads=1)
    s_ref = time.perf_counter() - t0
    baseline = {
        "value": rate_cls, "unit": "trade-paths-dates/s", "cores": threads, "kind": "port",
        "sample": f"same portfolio and dates, first {n_cls} Philox paths (drawn in the run), one simulation, "
                  f"zero-coupon bonds once per (path, date), schedule classes sharing legs, OpenMP over paths, "
                  f"PFE95 by selection; median of 5 runs after a warm-up ({s_cls:.2f} s each)",
        "host": host_cpu(),
        "bonds_shared_only": {"value": rate_zcb, "cores": threads,
                              "sample": f"the same without class sharing, first {n_zcb} paths, one run "
                                        f"({s_zcb:.2f} s)"},
        "reference_structure": {"value": n_ref_trades * n_ref_paths * len(dates) / s_ref, "cores": 1,
                                "sample": f"first {n_ref_trades} trades x {n_ref_paths} paths, each trade simulated "
                                          f"on its own with a sort per date, one thread ({s_ref:.1f} s)"},
    }
    return baseline, ref, n_zcb


def cpu_baseline_sets(jobs, dates, n_steps, model, seed, disc, target_s=15.0):
    """Batched sets on the CPU: the fused restatement with schedule classes shared (one simulation per
    set, OpenMP over paths), every set on the same fraction of its paths, sized to about target_s."""
    import oracle

    threads = cpu_threads()
    probe_port, _ = max(jobs, key=lambda j: j[0].n_trades)
    kw = dict(scheme=disc, hazard_rate=0.02, recovery_rate=0.4, want_trade=False, share_classes=True,
              n_threads=threads, seed=seed)
    times = []
    for probe in (32 * threads, 256 * threads):
        t0 = time.perf_counter()
        oracle.portfolio_exposure_fused(model, probe_port.as_dict(), dates, n_steps, n_paths=probe, **kw)
        times.append(time.perf_counter() - t0)
    per_unit = max(times[1] - times[0], 1e-9) / (224 * threads * probe_port.n_trades * len(dates))
    total_units = sum(p.n_trades * n for p, n in jobs) * len(dates)
    frac = min(1.0, target_s / (per_unit * total_units))
    units, elapsed = 0, 0.0
    for p, n in jobs:
        m = max(threads, int(frac * n))
        t0 = time.perf_counter()
        oracle.portfolio_exposure_fused(model, p.as_dict(), dates, n_steps, n_paths=m, **kw)
        elapsed += time.perf_counter() - t0
        units += p.n_trades * m * len(dates)
    return {"value": units / elapsed, "unit": "trade-paths-dates/s", "cores": threads, "kind": "port",
            "sample": f"every set on its first {100 * frac:.2f}% of paths (at least {threads}), the fused "
                      f"restatement with schedule classes shared, one simulation per set, OpenMP over paths "
                      f"({elapsed:.1f} s in all)", "host": host_cpu()}


def cpu_greeks_by_differences(port, dates, n_steps, model, seed, bumps, target_s=15.0):
    """CVA and its model sensitivities on the CPU as a fast CPU engine gets them: the fused restatement
    (schedule classes shared, OpenMP over paths) at the base and at the eight central bumps with common
    random numbers (the same Philox paths); credit sensitivities follow from the base EE in closed form.
    The rate counts the job's trade-paths-dates once, as the GPU line does."""
    import oracle

    threads = cpu_threads()
    pd_ = port.as_dict()
    kw = dict(scheme=1, hazard_rate=0.02, recovery_rate=0.4, share_classes=True, n_threads=threads, seed=seed)
    times = []
    for probe in (32 * threads, 256 * threads):
        t0 = time.perf_counter()
        oracle.portfolio_exposure_fused(model, pd_, dates, n_steps, n_paths=probe, **kw)
        times.append(time.perf_counter() - t0)
    per_path_run = max(times[1] - times[0], 1e-9) / (224 * threads)
    n = max(threads, int(target_s / (9 * per_path_run)))
    names = ("a", "b", "sigma", "r0")
    t0 = time.perf_counter()
    oracle.portfolio_exposure_fused(model, pd_, dates, n_steps, n_paths=n, **kw)
    for i, nm in enumerate(names):
        for sgn in (1, -1):
            m = list(model)
            m[i] += sgn * bumps[nm]
            oracle.portfolio_exposure_fused(tuple(m), pd_, dates, n_steps, n_paths=n, **kw)
    elapsed = time.perf_counter() - t0
    return {"value": port.n_trades * n * len(dates) / elapsed, "unit": "trade-paths-dates/s", "cores": threads,
            "kind": "port",
            "sample": f"CVA + a, b, sigma, r0 by central differences (9 fused runs, schedule classes shared, "
                      f"common Philox paths) on the first {n} paths, credit sensitivities in closed form "
                      f"({elapsed:.1f} s)", "host": host_cpu()}


def cpu_baseline_2f(port, dates, n_steps, model, seed, disc, target_s=15.0):
    """G2++: the per-trade restatement (one simulation for the portfolio, OpenMP over paths) on the first
    Philox paths; median of five runs after a warm-up."""
    import oracle

    threads = cpu_threads()
    pd_ = port.as_dict()
    probe = 2 * threads
    z = oracle.philox_normals2(seed, probe, n_steps)
    t0 = time.perf_counter()
    oracle.portfolio_exposure(model, pd_, dates, n_steps, z, scheme=disc, want_trade=True, n_threads=threads)
    per_path = max(time.perf_counter() - t0, 1e-6) / probe
    n = max(threads, int((target_s / 6.0) / per_path) // threads * threads)
    z = oracle.philox_normals2(seed, n, n_steps)
    holder = {}

    def run():
        holder["ref"] = oracle.portfolio_exposure(model, pd_, dates, n_steps, z, scheme=disc, hazard_rate=0.02,
                                                  recovery_rate=0.4, want_trade=True, n_threads=threads)

    rate, med = _median_rate(run, port.n_trades * n * len(dates))
    return ({"value": rate, "unit": "trade-paths-dates/s", "cores": threads, "kind": "port",
             "sample": f"G2++ restatement, first {n} Philox paths, one simulation for the portfolio, OpenMP over "
                       f"paths; median of 5 runs after a warm-up ({med:.2f} s each)", "host": host_cpu()},
            holder["ref"], n)


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix
# network FLOPs per evaluated sample: 3->64 (+bias), 64->64, 64->64, 64->1
MLP_FLOPS = 2 * (3 * 64 + 64 * 64 + 64 * 64 + 64)


def run_bermudan(args):
    """Learned continuation value on the Bermudan book: 200 payer swaptions, LSM policy on 2^18
    training paths, polynomial judge and BF16 networks trained on the GPU, network exposure on 2^20
    fresh paths x 120 quarterly dates. The timed step is the network's exposure pass: each trade's
    network evaluated on a per-date (r, log-numeraire) node lattice by MFMA, every alive path
    interpolated from it; the per-path network pass is timed once beside it."""
    from engine_quant_amd import eqhip
    from engine_quant_amd.berm import BermudanBook, MlpConfig
    from engine_quant_amd.portfolios import MODEL, bermudan_book

    spy, hz = 4, 30
    nr, nl = (int(v) for v in args.berm_lattice.split("x"))
    bk = bermudan_book(args.berm_trades, seed=args.seed)
    ctx = eqhip.Context(0)
    B = BermudanBook(ctx, MODEL, bk["notional"], bk["strike"], bk["first_ex"], bk["n_ex"], args.berm_train,
                     args.berm_infer, spy, hz, seed_train=args.seed * 101, seed_infer=args.seed * 211)
    pol = B.fit_policy()
    B.exposure_lsm(1)  # warm-up of the judge's kernels
    judge = []
    for _ in range(3):
        t0 = time.perf_counter()
        ee_lsm, _ = B.exposure_lsm(1)
        judge.append(1000.0 * (time.perf_counter() - t0))
    judge_ms = float(np.median(judge))
    B.exposure_lsm(0)  # out-of-sample polynomials for the accuracy comparison
    cfg = MlpConfig.for_model(MODEL, hz, steps=args.berm_steps, batch=args.berm_batch, lr=3e-3, lr_final=1e-5)
    curve = B.train(cfg)
    train_ms = B.timing()["train_ms"]
    B.exposure_mlp(1)
    t0 = time.perf_counter()
    ee_dir = B.exposure_mlp(1)
    direct_ms = 1000.0 * (time.perf_counter() - t0)
    for _ in range(args.warmup):
        B.exposure_mlp(1, lattice=(nr, nl))
    infer, net_ms, interp_ms = [], [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ee_net = B.exposure_mlp(1, lattice=(nr, nl))
        infer.append(1000.0 * (time.perf_counter() - t0))
        tm = B.timing()
        net_ms.append(tm["lattice_net_ms"])
        interp_ms.append(tm["lattice_interp_ms"])
        print(f"# step done: network exposure {infer[-1]:.1f} ms", file=sys.stderr, flush=True)
    mse_net, mse_poly = B.mse()
    tau, _ = B.pathwise(1, want_y0=False)
    kmax = (bk["first_ex"] + bk["n_ex"] - 1) * spy
    alive = 0
    for j in range(B.n_trades):
        t = tau[j]
        exercised_step = np.where(t < 0, 10 ** 9, t * spy)
        counts = np.bincount(np.minimum(exercised_step, kmax[j] + 1), minlength=kmax[j] + 2)
        # alive at k <=> exercise step >= k: paths with step >= k
        ge = np.cumsum(counts[::-1])[::-1]
        alive += int(ge[1:kmax[j] + 1].sum())
    del tau
    units = int(kmax.sum()) * args.berm_infer
    ms = float(np.median(infer))
    dates = np.arange(1, spy * hz + 1) / spy
    cva_net = ctx.unilateral_cva_batch(*MODEL, dates, ee_net, 0.02, 0.4)
    cva_lsm = ctx.unilateral_cva_batch(*MODEL, dates, ee_lsm, 0.02, 0.4)
    port_net, port_lsm = ee_net.sum(0), ee_lsm.sum(0)
    n_nodes = int(kmax.sum()) * nr * nl
    net_flops = n_nodes * MLP_FLOPS
    t_net, t_interp = float(np.median(net_ms)), float(np.median(interp_ms))
    # the interpolation pass's least traffic: r and log-numeraire once per (path, date), the exercise
    # year once per (trade, path), the node tables once
    interp_bytes = 16.0 * args.berm_infer * int(kmax.max()) + 1.0 * B.n_trades * args.berm_infer + 4.0 * n_nodes
    achieved_bw = interp_bytes / (t_interp / 1000.0) / 1e9

    def err(ee):
        return float(np.abs(ee - ee_lsm).sum() / ee_lsm.sum())

    out = {
        "metric": "trade-paths-dates/sec of the learned continuation-value exposure (Bermudan book, BF16 MFMA)",
        "value": units / (ms / 1000.0), "unit": "trade-paths-dates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate); fp64 paths", "data": "synthetic",
        "config": {"workload": "bermudan-surrogate", "n_trades": B.n_trades, "n_train": args.berm_train,
                   "n_infer": args.berm_infer, "dates": "quarterly x 120", "network": "3-64-64-64-1 tanh, one per trade",
                   "train_steps": args.berm_steps, "batch": args.berm_batch, "seed": args.seed,
                   "lattice": f"{nr} short-rate x {nl} log-numeraire nodes per (trade, date)"},
        "roofline": {"kernel": "mlp_lattice_ee_kernel", "bound": "hbm", "achieved": achieved_bw, "peak": 8000.0,
                     "unit": "GB/s", "frac": achieved_bw / 8000.0, "traffic": None,
                     "algorithmic_bytes": interp_bytes, "kernel_ms": t_interp,
                     "note": "per-path interpolation of the node tables from LDS; limited by LDS gathers and "
                             "load latency, not by HBM bandwidth"},
        "network_roofline": {"kernel": "mlp_lattice_kernel", "bound": "mfma", "achieved": net_flops / (t_net / 1000.0) / 1e12,
                             "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": net_flops / (t_net / 1000.0) / 1e12 / BF16_PEAK_TFLOPS,
                             "algorithmic_flops": net_flops, "nodes": n_nodes, "kernel_ms": t_net},
        "accuracy": {
            "notional_weighted_trade_ee_abs_err": err(ee_net),
            "notional_weighted_trade_ee_abs_err_per_path_network": err(ee_dir),
            "lattice_vs_per_path_network": float(np.abs(ee_net - ee_dir).sum() / ee_dir.sum()),
            "portfolio_ee_l1_rel_err": float(np.abs(port_net - port_lsm).sum() / port_lsm.sum()),
            "portfolio_ee_max_rel_err_of_peak": float(np.abs(port_net - port_lsm).max() / port_lsm.max()),
            "portfolio_cva_rel_err": float(abs(cva_net.sum() - cva_lsm.sum()) / cva_lsm.sum()),
            "oos_mse_per_unit_notional": {"network": float(mse_net.mean()), "cubic": float(mse_poly.mean()),
                                          "trades_where_network_better": float((mse_net < mse_poly).mean())}},
        "timing_ms": {"simulate": B.timing()["simulate_ms"], "lsm_policy": B.timing()["policy_ms"],
                      "policy_apply": B.timing()["apply_ms"], "lsm_judge_exposure": judge_ms, "train": train_ms,
                      "network_exposure": ms, "network_exposure_per_path": direct_ms,
                      "lattice_network_kernel": t_net, "lattice_interpolation_kernel": t_interp},
        "speedup_vs_lsm_judge": judge_ms / ms,
        "alive_samples": alive,
        "lsm_judge_rate": units / (judge_ms / 1000.0),
        "price_mean": float(pol["price_infer"].mean()), "loss_first_last": [float(curve[:50].mean()),
                                                                            float(curve[-200:].mean())],
    }
    if not args.no_cpu:
        import oracle
        n_cpu, trades = 1 << 17, list(range(32))
        z = oracle.philox_normals(args.seed, n_cpu, spy * hz)
        r, lb = oracle.berm_simulate(MODEL, 1.0 / spy, z)
        sub = {k: np.asarray(v)[trades] for k, v in bk.items()}
        t0 = time.perf_counter()
        t_c, y_c = oracle.berm_apply(MODEL, sub, spy, r, lb, pol["ex_coef"][trades])
        cpu_units = 0
        for i, j in enumerate(trades):
            for k in range(1, int(kmax[j]) + 1):
                oracle.berm_judge(MODEL, spy, r, lb, t_c[i], y_c[i], k)
            cpu_units += int(kmax[j]) * n_cpu
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_units / cpu_s, "unit": "trade-paths-dates/s", "cores": 1, "kind": "port",
                               "sample": f"plain LSM exposure (policy application + per-date cubic regression) of "
                                         f"trades 0-31 on {n_cpu} paths, {cpu_s:.1f} s, one thread"}
        out["speedup_vs_cpu_lsm"] = out["value"] / out["cpu_baseline"]["value"]
    B.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_batch(args):
    """64 netting sets (trades log-uniform 10-2000, paths log-uniform 10k-200k), 120 quarterly dates,
    one simulation + one autotuned valuation launch; per-set EE/PFE95/CVA."""
    from engine_quant_amd import eqhip
    from engine_quant_amd.portfolios import MODEL, batched_sets, quarterly_dates

    jobs = batched_sets(64, seed=args.seed)
    dates = quarterly_dates(30)
    ctx = eqhip.Context(0)
    batch = eqhip.Batch(ctx, MODEL, jobs, dates, len(dates), seed=args.seed, disc=eqhip.DISC_EXACT)
    for _ in range(max(1, args.warmup)):
        batch.execute(0.02, 0.4)  # the first execute runs the autotuner
    tm0 = batch.timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = batch.execute(0.02, 0.4)
        print(f"# step done: batch value kernel {batch.timing()['value_ms']:.2f} ms", file=sys.stderr, flush=True)
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    tm = batch.timing()
    units = sum(p.n_trades * n for p, n in jobs) * len(dates)
    v_ms = tm["value_ms"]
    # Every job is valued in path mode (no per-trade outputs, two schedule families): the roofline takes
    # the class-shared arithmetic that kernel performs; the per-trade count (each swap valued on its own)
    # is reported beside it.
    per_trade = sum(algorithmic_work(p, dates, n, p.n_sets)[0] for p, n in jobs)
    flops = 0.0
    for p, n in jobs:
        flops += path_mode_work(p, dates, n)[0]
    achieved = flops / (v_ms / 1000.0) / 1e12
    default_workload = args.seed == 7
    traffic = measured_traffic("value_path_kernel", PROFILEB) if default_workload else None
    counted = counter_fp64_flops(profile_counters(PROFILEB)) if default_workload else None
    out = {"metric": METRIC, "value": units / (ms / 1000.0), "unit": "trade-paths-dates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "hw1f-irs-batched-netting-sets", "n_sets": len(jobs),
                      "n_trades_total": int(sum(p.n_trades for p, _ in jobs)),
                      "n_paths_total": int(sum(n for _, n in jobs)), "n_dates": len(dates), "dates": "quarterly",
                      "rng": "philox4x32-10", "discretisation": "exact", "seed": args.seed},
           "kernel_ms": {"simulate": tm["simulate_ms"], "value": tm["value_ms"], "reduce": tm["reduce_ms"],
                         "cva": tm["cva_ms"], "total": tm["total_ms"]},
           "autotuner": tm0["tune_log"].strip().splitlines()[-1] if tm0["tune_log"].strip() else "",
           "tune_log": tm0["tune_log"].strip().splitlines(),
           "roofline": {"kernel": "value_path_kernel (one launch for all 64 sets)", "bound": "fp64-valu",
                        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                        "traffic_source": os.path.relpath(PROFILEB, ROOT) if traffic else None,
                        "algorithmic_flops": flops,
                        "algorithmic_unit": "per (path, date): exp + 2 per live cash-flow date, 2 per live family "
                                            "flow, 5 per (set, live class) entry, exp + 2 per start off the date "
                                            "(2 exp + 5 mid-period)",
                        "per_trade_flops": per_trade,
                        "per_trade_unit": "2m+6 FLOPs per live (trade, path, date), m = remaining fixed flows",
                        "executed_flops_counters": counted,
                        "executed_tflops_counters": counted / (v_ms / 1000.0) / 1e12 if counted else None,
                        "counters_source": os.path.relpath(PROFILEB, ROOT) if counted else None,
                        "value_kernel_ms": v_ms},
           "cva_total": float(sum(r["set_cva"].sum() for r in res))}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sets(jobs, dates, len(dates), MODEL, args.seed, eqhip.DISC_EXACT,
                                                args.cpu_seconds)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    batch.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_adjoint(args):
    """Pathwise CVA adjoint of the 1000-swap netting set on 200k paths x 120 quarterly dates, with
    short-rate checkpoints every 12 steps, against central finite differences with common random
    numbers; plus the analytic PV pillar / model sensitivities of every trade."""
    from engine_quant_amd import eqhip
    from engine_quant_amd.portfolios import MODEL, quarterly_dates, swap_portfolio, zero_curve

    pillars, zeros = zero_curve(args.seed)
    port = swap_portfolio(args.trades, seed=args.seed, pillars=pillars, zeros=zeros)
    dates = quarterly_dates(30)
    n_paths, n_steps = args.adj_paths, len(dates)
    ctx = eqhip.Context(0)
    plan = ctx.plan(MODEL, port, dates, n_paths, n_steps, seed=args.seed, rng=eqhip.RNG_PHILOX,
                    disc=eqhip.DISC_EXACT)
    for _ in range(args.warmup):
        plan.cva_adjoint(0, 0.02, 0.4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cva, grad, _ = plan.cva_adjoint(0, 0.02, 0.4)
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    plan.execute(None, 0.02, 0.4)
    fwd_ms = 1000.0 * (time.perf_counter() - t0)
    v_ms = plan.timing()["value_ms"]
    # the forward run has no per-trade outputs and two schedule families: path mode. The roofline takes the
    # class-shared arithmetic that kernel performs and reports the per-trade count beside it
    flops = algorithmic_work(port, dates, n_paths, port.n_sets)[0]
    exec_flops = path_mode_work(port, dates, n_paths)[0]
    achieved = exec_flops / (v_ms / 1000.0) / 1e12
    bumps = {"a": 1e-4, "b": 1e-4, "sigma": 1e-4, "r0": 1e-4}
    names = ("a", "b", "sigma", "r0")
    fd = {}
    for i, nm in enumerate(names):
        vals = []
        for sgn in (1, -1):
            m = list(MODEL)
            m[i] += sgn * bumps[nm]
            r = ctx.exposure_run(tuple(m), port, dates, n_paths, n_steps, seed=args.seed, rng=eqhip.RNG_PHILOX,
                                 disc=eqhip.DISC_EXACT, hazard_rate=0.02, recovery_rate=0.4)
            vals.append(r["set_cva"][0])
        fd[nm] = (vals[0] - vals[1]) / (2 * bumps[nm])
    t0 = time.perf_counter()
    dz, npv, greeks = ctx.pv_sensitivities(pillars, zeros, port, MODEL)
    sens_ms = 1000.0 * (time.perf_counter() - t0)
    _, dv01, bk = ctx.pv_dv01(pillars, zeros, port, bump=1e-4, bucketed=True)
    units = port.n_trades * n_paths * n_steps
    out = {"metric": "trade-paths-dates/sec of the pathwise CVA adjoint (value + reverse sweep)",
           "value": units / (ms / 1000.0), "unit": "trade-paths-dates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "hw1f-irs-cva-adjoint", "n_trades": port.n_trades, "n_paths": n_paths,
                      "n_dates": n_steps, "dates": "quarterly", "checkpoint_every": 12, "seed": args.seed},
           "cva": cva, "grad": {k: grad[k] for k in grad},
           "fd_central_crn_1e-4": fd,
           "adjoint_vs_fd_rel_err": {k: abs(grad[k] - fd[k]) / max(abs(fd[k]), 1e-300) for k in names},
           "forward_exposure_ms": fwd_ms, "adjoint_overhead_vs_forward": ms / fwd_ms,
           "fd_cost_in_forward_runs": 2 * len(names),
           "pv_pillar_sensitivities": {"ms": sens_ms, "n_pillars": len(pillars),
                                       "max_rel_err_vs_bucketed_dv01_over_bump": float(
                                           np.max(np.abs(bk - 1e-4 * dz) / np.maximum(np.abs(bk), 1.0)))},
           "cva_pillar_sensitivities": "identically zero: the CVA discounts with the model bond P(0,t|r0)",
           "roofline": {"kernel": "value_path_kernel (the adjoint's forward run)", "bound": "fp64-valu",
                        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "executed_flops_model": exec_flops,
                        "executed_unit": "path_mode_work: per (path, date) exp + 2 per live cash-flow date, 2 per "
                                         "live family flow, 5 per (set, live class) entry, family starts",
                        "per_trade_flops": flops,
                        "per_trade_tflops": flops / (v_ms / 1000.0) / 1e12,
                        "per_trade_unit": "2m+6 FLOPs per live (trade, path, date), m = remaining fixed flows",
                        "value_kernel_ms": v_ms,
                        "note": "the reverse sweep (adjoint_kernel) is timed in the committed adjoint profile"}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_greeks_by_differences(port, dates, n_steps, MODEL, args.seed, bumps,
                                                        args.cpu_seconds)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    plan.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_xva(args):
    """The XVA reducer (quant-engine xva::calculate) on the headline job's Monte Carlo cube, resident in
    HBM: one step = eqhip_plan_xva over the 360 x 100k signed netting-set cube (CSA variation margin,
    positive / negative parts, EE, ENE, the 95% quantile, CVA/DVA/FVA/MVA/KVA)."""
    from engine_quant_amd import eqhip, xva
    from engine_quant_amd.portfolios import MODEL, monthly_dates, swap_portfolio, zero_curve
    import oracle

    pillars, zeros = zero_curve(args.seed)
    port = swap_portfolio(args.trades, seed=args.seed, pillars=pillars, zeros=zeros)
    dates = monthly_dates(args.years)
    ctx = eqhip.Context(0)
    plan = ctx.plan(MODEL, port, dates, args.paths, len(dates), seed=args.seed, rng=eqhip.RNG_PHILOX,
                    disc=eqhip.DISC_EXACT, keep_cube=True)
    plan.execute(None, 0.02, 0.4)
    req = xva.XvaRequest(operation_id="bench", counterparty_default=xva.DefaultSpec(0.02, 0.4),
                         own_default=xva.DefaultSpec(0.01, 0.4), funding=xva.FundingSpec(0.004, 0.006, 0.0, 0.08),
                         collateral_agreements=[xva.CollateralAgreement("csa", threshold=5e5,
                                                                        minimum_transfer_amount=1e5,
                                                                        independent_amount=1e4,
                                                                        initial_margin=2e6)])
    params = xva._params(req)
    for _ in range(args.warmup):
        plan.xva(params)
    rows_ms, sel_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ee, pfe, m = plan.xva(params)
        tm = plan.timing()
        rows_ms.append(tm["xva_rows_ms"])
        sel_ms.append(tm["xva_select_ms"])
        print(f"# step done: xva rows {tm['xva_rows_ms']:.3f} ms, select {tm['xva_select_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    cells = len(dates) * args.paths
    r_ms = float(np.mean(rows_ms))
    bytes_rows = 16.0 * cells  # read the signed cube, write the positive parts
    achieved = bytes_rows / (r_ms / 1000.0) / 1e9
    default_workload = (args.trades, args.paths, args.years, args.seed) == (1000, 100_000, 30, 7)
    traffic = measured_traffic("xva_rows_kernel", PROFILEX) if default_workload else None
    out = {
        "metric": "exposure-cube cells/sec reduced by the XVA reducer (CVA/DVA/FVA/MVA/KVA, EE, 95% quantile)",
        "value": cells / (ms_per_step / 1000.0), "unit": "path-dates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "xva-reducer-on-resident-mc-cube", "n_trades": args.trades, "n_paths": args.paths,
                   "n_dates": len(dates), "cube": "signed netting-set values of the HW1F job, resident in HBM",
                   "csa": "threshold 5e5, MTA 1e5, IA 1e4, IM 2e6"},
        "roofline": {"kernel": "xva_rows_kernel", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": os.path.relpath(PROFILEX, ROOT) if traffic else None,
                     "algorithmic_bytes": bytes_rows, "xva_rows_ms": r_ms, "xva_select_ms": float(np.mean(sel_ms))},
        "measures": {k: float(v) for k, v in zip(("cva", "dva", "fva", "mva", "kva"), m)},
    }
    if not args.no_cpu:
        cube = xva.plan_cube(plan)
        c = req.collateral_agreements[0]
        t0 = time.perf_counter()
        ee_o, pfe_o, m_o = oracle.xva_calculate(
            cube.times, cube.paths, df=cube.discount_factors, counterparty=(0.02, 0.4, ()), own=(0.01, 0.4, ()),
            funding_spread=0.004, borrowing_spread=0.006, capital_factor=0.08, initial_margin_total=2e6,
            csa={"threshold": c.threshold, "minimum_transfer_amount": c.minimum_transfer_amount,
                 "independent_amount": c.independent_amount, "variation_margin": True}, confidence=0.95)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells / cpu_s, "unit": "path-dates/s", "cores": 1, "kind": "port",
                               "sample": f"the whole {len(dates)} x {args.paths} cube, one thread ({cpu_s:.1f} s)"}
        out["cva_rel_err_vs_oracle"] = abs(m[0] - m_o["cva"]) / abs(m_o["cva"])
        out["pfe_bitwise_equal"] = bool(np.array_equal(pfe, pfe_o))
    print(json.dumps(out), flush=True)


# The reference's own CPU-vs-GPU benchmark (rust/crates/engine-core/examples/gpu_vs_cpu_bench.rs:29-87):
# wall time of expected_exposure_profile for the 5y annual par payer IRS, 260 Euler steps, monitoring dates
# {0,1,2,3,4}, seed 42, after one warm-up call; published GPU (burn-wgpu) times, PLAN.md:1066-1069.
REFERENCE_GPU_MS = {1_000: 53.8, 10_000: 76.9, 100_000: 90.7, 1_000_000: 2170.0}


def run_reference_case(args):
    """The reference's published benchmark case through the reference-shaped entry point
    (eqhip_irs_hull_white_exposure_profile): each measurement is one whole call (plan set-up, Philox
    draws, Euler simulation, valuation, EE/PFE, copies back), as the reference timed one call."""
    from engine_quant_amd import eqhip
    import oracle

    a, b, sigma, r0 = 0.1, 0.03, 0.01, 0.02
    pay, acc, times = [1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5, [0.0, 1.0, 2.0, 3.0, 4.0]
    n_steps, seed = 260, 42
    ctx = eqhip.Context(0)

    def call(n):
        return ctx.exposure_profile(a, b, sigma, r0, 1e6, 0.0, True, 0.0, pay, acc, times, n_steps, n, seed,
                                    rng=eqhip.RNG_PHILOX, disc=eqhip.DISC_EULER)

    sizes = {}
    for n in sorted(REFERENCE_GPU_MS):
        call(min(n, 1_000))  # warm-up, as the reference discards its first call
        call(n)
        runs = []
        for _ in range(max(5, args.steps // 4)):
            t0 = time.perf_counter()
            call(n)
            runs.append(1000.0 * (time.perf_counter() - t0))
        ms = float(np.median(runs))
        sizes[n] = {"ms_per_call": ms, "reference_gpu_ms": REFERENCE_GPU_MS[n],
                    "speedup_vs_reference_gpu": REFERENCE_GPU_MS[n] / ms}
        print(f"# {n} paths: {ms:.3f} ms per call (reference GPU {REFERENCE_GPU_MS[n]} ms)", file=sys.stderr,
              flush=True)
    n = 100_000
    units = 1 * n * len(times)
    value = units / (sizes[n]["ms_per_call"] / 1000.0)
    ref_value = units / (REFERENCE_GPU_MS[n] / 1000.0)
    # kernel times of the same case from a plan's HIP events
    bf = lambda t: (1 - np.exp(-a * t)) / a  # noqa: E731  (the model bond, hull_white.rs:50-80)
    zcb = lambda t: np.exp((bf(t) - t) * (a * a * b - sigma * sigma / 2) / (a * a) - sigma * sigma * bf(t) ** 2
                           / (4 * a)) * np.exp(-bf(t) * r0)  # noqa: E731
    k_par = (1.0 - zcb(5.0)) / sum(zcb(t) for t in pay)
    port = eqhip.singleton_portfolio([1e6], [k_par], 0.0, pay, acc)
    plan = ctx.plan((a, b, sigma, r0), port, times, n, n_steps, seed=seed, rng=eqhip.RNG_PHILOX,
                    disc=eqhip.DISC_EULER)
    plan.execute()
    tm = plan.timing()
    plan.close()
    out = {
        "metric": METRIC, "value": value, "unit": "trade-paths-dates/s", "n_gpus": 1,
        "steps": max(5, args.steps // 4), "warmup": 1, "ms_per_step": sizes[n]["ms_per_call"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": value / ref_value, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "reference-gpu-vs-cpu-bench-case", "n_trades": 1, "n_paths": n, "n_dates": 5,
                   "n_steps": n_steps, "rng": "philox4x32-10", "discretisation": "euler", "seed": seed,
                   "swap": "5y annual par payer, N 1e6", "baseline": "reference burn-wgpu GPU time at 100k paths, "
                   "PLAN.md:1068 (unspecified development machine)", "sizes": sizes},
        "kernels_ms": {"simulate": tm["simulate_ms"], "value": tm["value_ms"], "reduce": tm["reduce_ms"]},
        "roofline": None,
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        z = oracle.philox_normals(seed, n, n_steps)
        _, ee_o, pfe_o = oracle.exposure_profile(a, b, sigma, r0, 1e6, 0.0, True, 0.0, pay, acc, times, n_steps, z)
        cpu_s = time.perf_counter() - t0
        _, ee, pfe = call(n)
        out["cpu_baseline"] = {"value": units / cpu_s, "unit": "trade-paths-dates/s", "cores": 1, "kind": "port",
                               "sample": f"the whole case ({n} paths), Philox draws and the reference-structured "
                                         f"profile, one thread ({cpu_s:.2f} s)"}
        out["ee_max_rel_err_vs_oracle"] = float(np.max(np.abs(ee - ee_o)) / max(np.max(np.abs(ee_o)), 1e-300))
    ctx.close()
    print(json.dumps(out), flush=True)


def run_caller(args):
    """The portfolio caller path: Engine.price_many over the headline portfolio through the C++ host (all
    five measures), one simulation per calendar group (the reference's price_many, price.cpp:534-578)
    against one simulation for every group (ExecutionContext simulation="shared")."""
    import engine_quant_amd as engine
    from engine_quant_amd.portfolios import MODEL, swap_portfolio, zero_curve

    pillars, zeros = zero_curve(args.seed)
    port = swap_portfolio(args.trades, seed=args.seed, pillars=pillars, zeros=zeros)
    eng = engine.Engine()
    a, b, sigma, r0 = MODEL
    model = eng.create_model("HullWhite1F", {"a": a, "b": b, "sigma": sigma, "r0": r0})
    off = port.flow_off
    trades = [eng.create_product("IRSwap", {"notional": float(port.notional[j]), "fixed_rate": float(port.fixed_rate[j]),
                                            "start": float(port.start[j]), "payment_times": list(port.pay[off[j]:off[j + 1]]),
                                            "accruals": list(port.acc[off[j]:off[j + 1]])})
              for j in range(port.n_trades)]
    market = engine.MarketSnapshot(pillars=list(pillars), zero_rates=list(zeros), hazard_rate=0.02, recovery_rate=0.4)
    n_paths, n_steps = args.paths, 360
    pricing = engine.PricingContext({"n_paths": float(n_paths), "n_steps": float(n_steps), "seed": float(args.seed)})
    measures = ["PV", "DV01", "ExpectedExposure", "PFE95", "UnilateralCVA"]
    own_dates = sum(1 + int(off[j + 1] - off[j]) - 1 for j in range(port.n_trades))  # start + all but the last payment
    units = n_paths * own_dates
    groups = len({(float(port.start[j]), tuple(port.pay[off[j]:off[j + 1]])) for j in range(port.n_trades)})
    res = {}
    for mode in ("per_group", "shared"):
        ex = engine.ExecutionContext({"backend": "gpu", "discretisation": "exact", "simulation": mode})
        eng.price_many(trades, measures, model, market, pricing, ex)  # warm-up
        runs = []
        for _ in range(max(3, args.steps // 5)):
            t0 = time.perf_counter()
            out = eng.price_many(trades, measures, model, market, pricing, ex)
            runs.append(time.perf_counter() - t0)
            print(f"# {mode}: {1000 * runs[-1]:.1f} ms", file=sys.stderr, flush=True)
        cva = float(sum(r.measures["UnilateralCVA"].scalar for r in out))
        res[mode] = {"ms": 1000.0 * float(np.median(runs)), "cva_sum": cva}
    value = units / (res["shared"]["ms"] / 1000.0)
    out = {
        "metric": METRIC, "value": value, "unit": "trade-paths-dates/s", "n_gpus": 1,
        "steps": max(3, args.steps // 5), "warmup": 1, "ms_per_step": res["shared"]["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "price-many-caller-path", "n_trades": port.n_trades, "calendar_groups": groups,
                   "n_paths": n_paths, "n_steps": n_steps, "monitoring": "each trade's reset dates",
                   "measures": measures, "discretisation": "exact"},
        "per_group": {"ms_per_call": res["per_group"]["ms"],
                      "value": units / (res["per_group"]["ms"] / 1000.0), "cva_sum": res["per_group"]["cva_sum"]},
        "shared": {"ms_per_call": res["shared"]["ms"], "cva_sum": res["shared"]["cva_sum"]},
        "speedup_shared_vs_per_group": res["per_group"]["ms"] / res["shared"]["ms"],
        "roofline": None,
    }
    print(json.dumps(out), flush=True)


def run_greeks(args):
    """The Greek measure on the GPU measures (cpp/engine/src/greeks.cpp): the CVA Greek sweep of one swap
    through the host engine (Engine.all_greeks: model, curve-pillar and credit factors by central
    bump-and-revalue with common random numbers), and the pathwise route (the CVA adjoint) for the four
    model parameters beside the bumped ones."""
    import engine_quant_amd as engine

    eng = engine.Engine()
    a, b, sigma, r0 = 0.1, 0.03, 0.01, 0.02
    model = eng.create_model("HullWhite1F", {"a": a, "b": b, "sigma": sigma, "r0": r0})
    swap = eng.create_product("IRSwap", {"notional": 1e6, "fixed_rate": 0.021, "payment_times": [1, 2, 3, 4, 5.0],
                                         "accruals": [1.0] * 5})
    market = engine.MarketSnapshot(pillars=[1.0, 2.0, 5.0], zero_rates=[0.02, 0.022, 0.025], hazard_rate=0.02,
                                   recovery_rate=0.4)
    n_paths, n_steps = args.paths, 208
    pricing = engine.PricingContext({"n_paths": float(n_paths), "n_steps": float(n_steps), "seed": float(args.seed)})
    gpu = engine.ExecutionContext({"backend": "gpu", "discretisation": "exact"})
    eng.all_greeks(swap, "UnilateralCVA", model, market, pricing, gpu, include_curve_buckets=True)  # warm-up
    runs = []
    for _ in range(max(3, args.steps // 5)):
        t0 = time.perf_counter()
        rep = eng.all_greeks(swap, "UnilateralCVA", model, market, pricing, gpu, include_curve_buckets=True)
        runs.append(1000.0 * (time.perf_counter() - t0))
    sweep_ms = float(np.median(runs))
    greeks = {g.risk_factor: g.value for g in rep.greeks}
    route = {}
    for name in ("a", "b", "sigma", "r0"):
        for method in ("bump_and_reval", "pathwise"):
            spec = [("Greek", {"metric": "UnilateralCVA", "risk_factor": f"model.{name}", "method": method})]
            eng.price(swap, spec, model, market, pricing, gpu)
            t0 = time.perf_counter()
            g = eng.price(swap, spec, model, market, pricing, gpu)["Greek"]
            route[f"{name}/{method}"] = {"ms": 1000.0 * (time.perf_counter() - t0), "value": g.scalar}
    n_revals = sum(2 if not k.startswith("time.") else 1 for k in greeks)  # central stencils
    units = n_revals * n_paths * 5  # revaluations x paths x monitoring dates
    out = {
        "metric": "Greek sweep of UnilateralCVA through the host engine (bump-and-revalue with CRN on the GPU)",
        "value": units / (sweep_ms / 1000.0), "unit": "revalued trade-paths-dates/s", "n_gpus": 1,
        "steps": len(runs), "warmup": 1, "ms_per_step": sweep_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "greek-sweep-unilateral-cva", "swap": "5y annual payer, N 1e6, K 2.1%",
                   "n_paths": n_paths, "n_steps": n_steps, "risk_factors": len(greeks), "revaluations": n_revals},
        "greeks": greeks, "skipped": list(rep.skipped), "routes": route, "roofline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--trades", type=int, default=1000)
    ap.add_argument("--paths", type=int, default=100_000)
    ap.add_argument("--years", type=int, default=30)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--disc", choices=["exact", "euler"], default="exact")
    ap.add_argument("--serial-pv", action="store_true",
                    help="run the PV/DV01 ladder after the exposure job on the same stream instead of beside it")
    ap.add_argument("--model", choices=["hw1f", "hw2f"], default="hw1f",
                    help="short-rate model of the swaps workload: HullWhite1F (default) or HullWhite2F (G2++)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-trade-pfe", action="store_true", help="skip the per-trade PFE95 measurement")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["swaps", "batch", "adjoint", "bermudan", "xva", "reference-case",
                                           "caller", "greeks"],
                    default="swaps",
                    help="swaps: the one-netting-set IRS exposure job (default); batch: 64 netting sets in one "
                         "launch; adjoint: pathwise CVA Greeks; bermudan: the learned continuation-value "
                         "exposure of the Bermudan book; xva: the XVA reducer on the resident Monte Carlo cube; "
                         "reference-case: the reference's own published GPU benchmark case; caller: price_many "
                         "through the host engine, per calendar group and shared; greeks: the CVA Greek sweep "
                         "(each prints its own JSON line)")
    ap.add_argument("--adj-paths", type=int, default=200_000)
    ap.add_argument("--berm-trades", type=int, default=200)
    ap.add_argument("--berm-train", type=int, default=1 << 18)
    ap.add_argument("--berm-infer", type=int, default=1 << 20)
    ap.add_argument("--berm-steps", type=int, default=20000, help="Adam steps of the network training")
    ap.add_argument("--berm-batch", type=int, default=4096, help="samples per network per training step")
    ap.add_argument("--berm-lattice", default="128x8", help="network nodes per (trade, date): short rate x log-numeraire")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (hardware-counter scale factor)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library runs on one GPU "
              "(paths would shard by global path index across GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "bermudan":
        run_bermudan(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "adjoint":
        run_adjoint(args)
        return
    if args.workload == "xva":
        run_xva(args)
        return
    if args.workload == "reference-case":
        run_reference_case(args)
        return
    if args.workload == "caller":
        run_caller(args)
        return
    if args.workload == "greeks":
        run_greeks(args)
        return

    from engine_quant_amd import eqhip
    from engine_quant_amd.portfolios import MODEL, MODEL2, monthly_dates, swap_portfolio, zero_curve

    pillars, zeros = zero_curve(args.seed)
    port = swap_portfolio(args.trades, seed=args.seed, pillars=pillars, zeros=zeros)
    dates = monthly_dates(args.years)
    n_steps = len(dates)
    disc = eqhip.DISC_EXACT if args.disc == "exact" else eqhip.DISC_EULER
    hazard, recovery = 0.02, 0.4
    model = MODEL2 if args.model == "hw2f" else MODEL
    factors = 2 if args.model == "hw2f" else 1

    ctx = eqhip.Context(0)
    plan = ctx.plan(model, port, dates, args.paths, n_steps, seed=args.seed, rng=eqhip.RNG_PHILOX, disc=disc,
                    want_trade_ee=True)
    # the PV / DV01 ladder does not depend on the exposure job: a second context (its own stream) runs
    # it from a worker thread while the exposure kernels run (ctypes releases the GIL in foreign calls)
    from concurrent.futures import ThreadPoolExecutor
    ctx_pv = eqhip.Context(0)
    pool = ThreadPoolExecutor(max_workers=1)

    def step():
        if args.serial_pv:
            res = plan.execute(None, hazard, recovery, want_cva=True)
            pv, dv, bk = ctx.pv_dv01(pillars, zeros, port, bump=1e-4, bucketed=True)
            return res, pv, dv
        ladder = pool.submit(ctx_pv.pv_dv01, pillars, zeros, port, 1e-4, True)
        res = plan.execute(None, hazard, recovery, want_cva=True)
        pv, dv, bk = ladder.result()
        return res, pv, dv

    for _ in range(args.warmup):
        step()
    value_ms, total_kernel_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, pv, dv = step()
        tm = plan.timing()
        value_ms.append(tm["value_ms"])
        total_kernel_ms.append(tm["total_ms"])
        print(f"# step done: value_kernel {tm['value_ms']:.3f} ms, exposure kernels {tm['total_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0  # every step ends with a stream synchronisation
    ms_per_step = 1000.0 * elapsed / args.steps
    units = port.n_trades * args.paths * len(dates)
    value = units / (ms_per_step / 1000.0)

    flops, exec_flops, bytes_, u_live, n_exp = algorithmic_work(port, dates, args.paths, port.n_sets, factors)
    v_ms = float(np.mean(value_ms))
    achieved = flops / (v_ms / 1000.0) / 1e12
    executed = exec_flops / (v_ms / 1000.0) / 1e12
    tm = plan.timing()
    default_workload = (args.trades, args.paths, args.years, args.seed, args.disc) == (1000, 100_000, 30, 7, "exact")
    # one path per lane (the default variant) runs the kernel of its own
    kernel = "value2_kernel" if factors == 2 else ("value_p1_kernel" if tm["paths_per_tile"] == 64 else "value_kernel")
    profile = PROFILE2 if factors == 2 else PROFILE
    traffic = measured_traffic(kernel, profile) if default_workload else None
    counted_flops = counter_fp64_flops(profile_counters(profile)) if default_workload else None
    counted_tflops = counted_flops / (v_ms / 1000.0) / 1e12 if counted_flops else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "trade-paths-dates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.model}-irs-one-netting-set", "n_trades": args.trades, "n_paths": args.paths,
                   "n_dates": len(dates), "n_steps": n_steps, "dates": "monthly", "rng": "philox4x32-10",
                   "discretisation": args.disc, "seed": args.seed, "n_sets": port.n_sets,
                   "outputs": "set EE/PFE95/CVA, trade EE/CVA, PV/DV01/bucketed DV01 (20 pillars)",
                   "pv_ladder": "serial on the job's stream" if args.serial_pv else
                                "second context and stream, concurrent with the exposure job",
                   "live_units": u_live},
        "roofline": {"kernel": kernel, "bound": "fp64-valu", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "compute": "fp64 on the vector ALUs (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "traffic_source": os.path.relpath(profile, ROOT) if traffic else None,
                     "algorithmic_flops": flops, "algorithmic_exps": n_exp,
                     "algorithmic_unit": "2m+6 FLOPs per live (trade, path, date), m = remaining fixed flows",
                     "frac_note": "achieved is the per-trade algorithm's work over the kernel's time; the "
                                  "kernel shares each schedule class's bonds and annuities across its trades "
                                  "and executes about a fifth of that work, so frac can exceed 1; "
                                  "executed_frac_counters is the fraction of the FP64 peak the kernel executes",
                     "executed_flops_model": exec_flops, "executed_tflops_model": executed,
                     "executed_flops_counters": counted_flops,
                     "executed_tflops_counters": counted_tflops,
                     "executed_frac_counters": (counted_tflops / FP64_PEAK_TFLOPS) if counted_tflops else None,
                     "counters_source": os.path.relpath(profile, ROOT) if counted_flops else None,
                     "algorithmic_bytes": bytes_,
                     "value_kernel_ms": v_ms, "exposure_kernels_ms": float(np.mean(total_kernel_ms)),
                     "paths_per_tile": tm["paths_per_tile"], "tiles_per_block": tm["tiles_per_block"],
                     "blocks": tm["n_value_blocks"]},
        "set_cva": float(res["set_cva"][0]),
    }
    if factors == 1 and not args.no_trade_pfe:
        # per-trade PFE95 beside the headline outputs, outside its timed region: the same job with the
        # class legs written by the valuation and one exact select per live (trade, date)
        plan_pfe = ctx.plan(model, port, dates, args.paths, n_steps, seed=args.seed, rng=eqhip.RNG_PHILOX, disc=disc,
                            want_trade_ee=True, want_trade_pfe=True)
        runs = []
        for i in range(4):
            t1 = time.perf_counter()
            r_pfe = plan_pfe.execute(None, hazard, recovery, want_cva=True)
            wall = time.perf_counter() - t1
            if i > 0:
                runs.append((plan_pfe.timing(), wall))
        plan_pfe.close()
        assert np.array_equal(r_pfe["set_ee"], res["set_ee"]) and np.array_equal(r_pfe["set_pfe"], res["set_pfe"])
        live_rows = int(u_live // args.paths)
        pfe_ms = float(np.mean([t["trade_pfe_ms"] for t, _ in runs]))
        out["trade_pfe"] = {
            "trade_select_ms": pfe_ms,
            "value_kernel_with_class_legs_ms": float(np.mean([t["value_ms"] for t, _ in runs])),
            "exposure_kernels_ms": float(np.mean([t["total_ms"] + t["trade_pfe_ms"] for t, _ in runs])),
            "wall_ms_per_execute": 1000.0 * float(np.mean([w for _, w in runs])),
            "date_groups": int(runs[-1][0]["trade_pfe_groups"]),
            "live_trade_date_rows": live_rows,
            "note": "exact rank-ceil(0.95 n)-1 order statistic of every live (trade, date) over all paths; "
                    "runs 1-3 after one warm-up execute",
        }
    if not args.no_cpu:
        if factors == 2:
            out["cpu_baseline"], ref, n_ref = cpu_baseline_2f(port, dates, n_steps, model, args.seed, disc,
                                                              args.cpu_seconds)
        else:
            out["cpu_baseline"], ref, n_ref = cpu_baseline(port, dates, n_steps, model, args.seed, disc,
                                                           args.cpu_seconds)
        # the same paths on the GPU (counter-based stream: path p's draws do not depend on n_paths)
        small = ctx.exposure_run(model, port, dates, n_ref, n_steps, seed=args.seed, rng=eqhip.RNG_PHILOX,
                                 disc=disc, hazard_rate=hazard, recovery_rate=recovery, want_trade_ee=True)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
        out["cva_rel_err_vs_fp64_oracle"] = abs(small["set_cva"][0] - ref["set_cva"][0]) / max(
            abs(ref["set_cva"][0]), 1e-300)
        out["accuracy_sample_paths"] = n_ref
        out["trade_ee_max_rel_err_vs_oracle"] = float(
            np.max(np.abs(small["trade_ee"] - ref["trade_ee"]) / np.maximum(np.abs(ref["trade_ee"]).max(), 1e-300)))
    pool.shutdown()
    plan.close()
    ctx_pv.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
