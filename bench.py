"""Benchmark of the boiling path on one MI355X.

Workload (default): 4096 independent cases of uneven size in one launch per
kernel -- base grid N ~ U{128..4096}, superheat U(2,20) K, film thickness
U(0.05,0.5) mm, refinement r=2, drawn from SplitMix64 seed 20261202.  A
"step" is one full RK3 time step of every case (1d.f:196-295); the value is
completed instance-steps/s with the state resident in HBM, every Poisson
solve taken by V-cycles until its normwise backward error and its last
correction are both within 1e-10.

Timed runs of the same batch from the same initial state, each over the same
window of steps after the same untimed warmup:
  mixed+learned    FP32 V-cycle corrections in FP64 defect correction,
                   started from the history projection corrected by the
                   warm-start network (the value);
  mixed+projected  the same without the network (what the network adds);
  mixed from zero  FP32 V-cycle corrections from a zero guess (what the
                   history projection adds);
  plain FP64       FP64 V-cycles from a zero initial guess (IOLDV=0);
  thomas           the reference's own direct solve (TRDIAG3 order, bit for
                   bit the FP64 oracle): the yardstick for which cases stop.
Before them the network is trained on the GPU on pairs the solver generates
from a disjoint batch (seed 20261203); the training is not timed.

Every run reports its stopped cases (the reference's STOP branches) and how
they compare with the Thomas run's, case by case (status and step), and the
interface position's largest relative difference from the Thomas run over
every case alive there (a case alive there but stopped here counts apart);
over the cases alive in both, each field's (W, P, T, PHI, and the scalars
QWALL and PLACE) relative difference from the Thomas run, case by case
(max|f - f_Thomas| / max|f_Thomas|), its largest and median value.  A
second Thomas run with every case's z0 one ulp up gives the same figures
for rounding alone (the differences a one-ulp change of the input makes).

The JSON line also carries
  roofline      the dominant kernel of the headline mode: its algorithmic
                bytes per launch (distinct arrays x elements x size, twice
                when read and written) over its exclusive launch duration,
                taken with HIP events on a one-lane run of the same batch
                and steps (one launch per kernel and substep, no concurrent
                launches), against 8 TB/s; `traffic` is the HBM bytes per
                launch that rocprofv3's FETCH_SIZE/WRITE_SIZE counters gave
                for that kernel on this workload, read from the committed
                profile summary (null when there is none);
  step_roofline the whole step: the algorithmic bytes of every kernel of
                every step completed in the timed window (the same per-kernel
                rule) over the window's time, against 8 TB/s;
  cpu_baseline  the FP64 oracle (the serial restatement, built -O3 -fopenmp)
                on the GPU box's host cores: one case per OpenMP thread over
                a bounded sample of the same batch, set-up excluded.
``--workload large`` / ``deck`` and ``--mode`` select the single large case,
the shipped deck and the other Poisson modes.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Lanes run on streams of their own, and streams beyond the process's
# hardware queues share a queue, which serialises them; HIP reads the queue
# count once, when it starts, and defaults to 4.  Four lanes, the batch's
# own stream and the host-side copies need more than that: ask for 8 before
# anything starts HIP.
LANES_BATCH = 4
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def mg_rows_total(n, coarse=32):
    """rows of every level of the V-cycle hierarchy (levels halve down to 32 rows)"""
    tot, m = n, n
    while m > coarse:
        m = (m + 1) // 2
        tot += m
    return tot


def kernel_bytes(case, kernel, mode):
    """Algorithmic HBM bytes of one launch for one case: every distinct
    global array the kernel touches, elements touched x element size, twice
    when read and written (ghost cells, reuse through caches and re-reads
    not counted).  Band-sized work (a few dozen refined planes) is counted
    with its own arrays."""
    n = int(case.n)
    # the direct and Thomas modes pass the energy and momentum rows through
    # four tridiagonal work vectors (written, read back); the multigrid modes
    # move them from LDS into the solve's registers
    trid = 4 * 16 if mode < 2 else 0
    per_cell = {
        # PSI_CN read, RK3ZO written
        "dt": 16,
        # W,PSI_ZN,PSI_CN,VVDZ,AKUV,CKUV read; T r/w; DF_Z,VF_Z,DF_C,PSI_Z,PSI_C,RHSPS
        # written
        "cont_energy": 6 * 8 + 16 + 6 * 8 + trid,
        # W,PSI_Z,PSI_C,RK3ZO r/w; PSI_CN,DF_Z,VF_Z,SSDZ,VVDZ,P,PSI_ZN,AKW,CKW read;
        # IMASK (int); KCOU1/KCOU2 (int); Poisson matrix written; in the
        # multigrid modes the solution through TG (written, read back)
        "momentum": 4 * 16 + 9 * 8 + 4 + 8 + trid + 4 * 8 + (16 if mode >= 2 else 0),
        # W,P,WO r/w; PSI_Z,VVDZ,SSDZ read; KCOU1/KCOU2; WT written
        "ucalc": 3 * 16 + 3 * 8 + 8 + 8,
    }
    if kernel in per_cell:
        return per_cell[kernel] * n
    if kernel == "poisson":
        if mode == 0:   # Thomas: a r/w (pivots), b,c read, g r/w, P written
            return (16 + 8 + 8 + 16 + 8) * n
        if mode == 1:   # partitioned: a,b,c,g r/w, P written
            return (4 * 16 + 8) * n
        # V-cycles: a,c,g read (b = -(a + c) is evaluated), P written, the
        # two iterate buffers r/w, the coarse levels' setup rows (4 doubles)
        # r/w, the weights of every level's pairs of rows (2 for the C row, 3
        # for the F row, in the cycle's precision) r/w, the warm-start guess
        # read
        tot = mg_rows_total(n)
        w = 8 if mode == 2 else 4
        return (3 * 8 + 8 + 2 * 16) * n + 2 * 32 * (tot - n) + 2 * 5 * w * (tot - n) + (8 * n if mode in (4, 5) else 0)
    # band kernels: 11 layers of <=16 planes, refined window of ~200 planes of
    # PSIF/SFGAS/geometry, NS window of ~100 cells of T/PSI_CN/geometry
    r = int(case.r)
    band = 11 * 16 * 8 * 4
    if kernel == "transport":
        return band + 200 * 8 * 4 + 100 * 8 * 6 + 12 * r
    if kernel == "temp_grad":
        return band + 200 * 8 * 6 + 100 * 8 * 8
    if kernel == "warm_start":
        # the projection: a, c, g and the history's 4 solutions + P read in
        # each of its 2 passes, the guess written (the network's window of
        # 256 cells adds a few KB per case, not counted)
        return (2 * 8 * (3 + 5) + 8) * n if mode in (4, 5) else 0
    return 0


STEP_KERNELS = ("cont_energy", "temp_grad", "momentum", "warm_start", "poisson", "ucalc", "transport")


def step_bytes(case, mode):
    """Algorithmic HBM bytes of one full step of one case: DETERMINE_DT once,
    the substep's kernels three times."""
    return kernel_bytes(case, "dt", mode) + 3 * sum(kernel_bytes(case, k, mode) for k in STEP_KERNELS)


def step_roofline(cases, mode, res):
    """The whole step against the HBM roofline: the algorithmic bytes of
    every step the cases completed in the timed window, over the window's
    time (all lanes at once)."""
    tot = sum(step_bytes(c, mode) * int(d) for c, d in zip(cases, res["done_per_case"]))
    ach = tot / res["el"] / 1e9
    return {"bound": "hbm", "algorithmic_bytes": tot, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "timing": "the timed steps, all lanes"}


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(cases, workload, steps=3):
    """The FP64 oracle on the host cores (gcc -O3 -fopenmp -ffp-contract=off).
    Batch: mode (ii) of the reference-style comparison -- one case per OpenMP
    thread over every case of the batch for `steps` steps, set-up excluded.
    A single case: its steps on one core (its loops are serial)."""
    from oracle.binding import run_cases
    info = host_info()
    if len(cases) == 1:
        st, done, t_set, t_step, thr = run_cases(cases * 1, 200 if cases[0].n <= 1024 else 20)
        v = float(done.sum()) / t_step
        return {"value": v, "unit": "instance-steps/s", "cores": 1, "kind": "port",
                "sample": f"{int(done.sum())} steps of the {workload} (N={cases[0].n}, r={cases[0].r}) by the "
                          f"FP64 oracle on one host core, {t_step:.1f} s, set-up excluded",
                **info}
    st, done, t_set, t_step, thr = run_cases(cases, steps)
    v = float(done.sum()) / t_step
    out = {"value": v, "unit": "instance-steps/s", "cores": thr, "kind": "port",
           "sample": f"all {len(cases)} cases of the same batch, {steps} steps each ({int(done.sum())} "
                     f"instance-steps, {int((st != 0).sum())} cases stopped) by the FP64 oracle, one case per "
                     f"OpenMP thread on {thr} threads, {t_step:.1f} s of stepping; set-up ({t_set:.1f} s) "
                     "excluded",
           "mode": "instance-parallel", **info}
    # mode (i), the reference's own style: one case at a time, its per-cell
    # extrapolation loops shared by the threads; on a sample of the batch
    # (every 64th case), with one thread and with all of them
    from oracle.binding import run_single
    sample = cases[::64]
    mode_i = {}
    for t in (1, thr):
        st1, done1, _, t1, used = run_single(sample, steps, t)
        mode_i[f"threads_{used}"] = float(done1.sum()) / t1 if t1 > 0 else None
    out["mode_i"] = {"unit": "instance-steps/s", **mode_i,
                     "sample": f"{len(sample)} cases (every 64th of the batch), {steps} steps each, one case at a "
                               "time with the pseudo-time extrapolation's per-cell loops shared by the OpenMP "
                               "threads (the reference's !$OMP PARALLEL DO style)"}
    return out


def lanes_of(cases, args):
    """Case sets run side by side on as many streams (2..8) or one sequence (1)."""
    return args.lanes if args.lanes else (LANES_BATCH if len(cases) >= 256 else 1)


# fields compared, case by case, with the Thomas run of the same batch
FIELD_CMP = ("W", "P", "T", "PHI")


def timed_run(L, cases, mode, args, net=None, keep_fields=False):
    """Warmup steps, then args.steps timed steps (host clock, synchronised
    on both sides); only the steps the cases completed are counted (a case
    that reaches one of the reference's STOP conditions keeps its status and
    is left behind).  keep_fields: the FIELD_CMP fields of the running cases
    after the timed steps (read after the clock stops)."""
    import numpy as np
    b = L.Batch(cases)
    if net is not None:
        b.set_net(net)
    b.set_poisson(mode, args.rtol, 200)
    b.set_block(args.threads)
    b.set_lanes(lanes_of(cases, args))
    b.set_lane_split(args.split == "binned")
    if getattr(args, "graph", False):
        b.use_graph(True)
    if args.warmup:
        b.run_all(args.warmup)
    tuning = b.autotune(args.tune) if args.tune > 0 else b.tuning()
    b.sync()
    s0 = np.array([b.scalar_array(i) for i in range(len(cases))])
    _, steps0 = b.progress()
    b.sync()
    t0 = time.perf_counter()
    nbad = b.run_all(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    st, steps1 = b.progress()
    s1 = np.array([b.scalar_array(i) for i in range(len(cases))])
    ok = (st == 0)
    fields = None
    if keep_fields:
        fields = {f: [b.field(f, i) if ok[i] else None for i in range(len(cases))] for f in FIELD_CMP}
    b.close()
    done = steps1 - steps0
    res = {"el": el, "done": int(done.sum()), "done_per_case": done, "value": float(done.sum()) / el,
           "cells": float(sum(int(c.n) * int(d) for c, d in zip(cases, done))) / el,
           "ok": ok, "place": s1[:, L.SCALARS.index("PLACE")], "qwall": s1[:, L.SCALARS.index("QWALL")],
           "nbad": nbad, "status": st, "steps": steps1, "tuning": tuning, "fields": fields}
    if mode >= L.POISSON_MG64 and ok.any():
        ci, si, ri = (L.SCALARS.index(k) for k in ("POISSON_CYCLES", "POISSON_SOLVES", "POISSON_RESID"))
        res["vcycles_per_solve"] = float((s1[ok, ci] - s0[ok, ci]).sum() / max((s1[ok, si] - s0[ok, si]).sum(), 1.0))
        res["resid_max"] = float(s1[ok, ri].max())
    return res


def kernel_profile(L, cases, mode, args, net=None):
    """Exclusive GPU ms per launch of every kernel class: the same batch and
    steps on one lane with HIP events between the launches (one launch per
    kernel and substep, nothing running beside it)."""
    b = L.Batch(cases)
    if net is not None:
        b.set_net(net)
    b.set_poisson(mode, args.rtol, 200)
    b.set_block(args.threads)
    b.set_lanes(1)
    if args.warmup:
        b.run_all(args.warmup)
    if args.tune > 0:
        b.autotune(args.tune)
    b.sync()
    b.set_timing(True)
    b.run_all(args.steps)
    b.sync()
    tim = b.timings()
    st, _ = b.progress()
    b.close()
    launches = {"dt": 1, "cont_energy": 3, "temp_grad": 3, "momentum": 3, "poisson": 3, "poisson_mg": 3, "ucalc": 3,
                "transport": 3, "warm_start": 3}
    per = {k: tim.get(k, 0.0) / (args.steps * nl) for k, nl in launches.items()}
    if mode >= L.POISSON_MG64:
        per["poisson"] = per.pop("poisson_mg")
    else:
        per.pop("poisson_mg")
    return per, (st == 0)


def profile_traffic(kernel, workload_key):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 counter summary (profiles/<kernel>_<workload>_hbm_traffic.md,
    line 'traffic_bytes_per_launch: <bytes>'), or None."""
    path = os.path.join(PROFILES, f"{kernel}_{workload_key}_hbm_traffic.md")
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def roofline(L, cases, mode, per, alive, workload_key):
    """Dominant kernel by exclusive launch time: algorithmic bytes of one
    launch (the cases still running) over that time."""
    dom = max((k for k in per if k != "warm_start"), key=lambda k: per[k])
    live = [c for c, g in zip(cases, alive) if g]
    bytes_launch = sum(kernel_bytes(c, dom, mode) for c in live)
    ach = bytes_launch / (per[dom] * 1e-3) / 1e9
    kname = "k_" + ("poisson_mg" if dom == "poisson" and mode >= L.POISSON_MG64 else
                    "poisson_direct" if dom == "poisson" else dom)
    return {"bound": "hbm", "kernel": kname, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": profile_traffic(kname, workload_key),
            "ms_per_launch": per[dom], "bytes_per_launch": bytes_launch,
            "timing": "exclusive launch duration, HIP events, one lane"}


def stopped(L, res):
    import numpy as np
    st = res["status"]
    codes, counts = np.unique(st[st != 0], return_counts=True)
    return {L.status_string(int(k)): int(v) for k, v in zip(codes, counts)}


def versus_thomas(res, ref):
    """Stops and interface of a run against the Thomas run of the same batch."""
    import numpy as np
    same = (res["status"] == ref["status"]) & (res["steps"] == ref["steps"])
    alive_ref = ref["status"] == 0
    both = alive_ref & (res["status"] == 0)
    rel = np.abs(res["place"][both] - ref["place"][both]) / np.abs(ref["place"][both])
    diff = np.where(~same)[0]
    out = {"cases_with_other_stop_than_thomas": int(diff.size),
           "those_cases": [int(i) for i in diff[:20]],
           "alive_in_thomas_but_stopped_here": int((alive_ref & (res["status"] != 0)).sum()),
           "interface_rel_diff_max_over_cases_alive_in_both": float(rel.max()) if rel.size else None}
    if res.get("fields") is not None and ref.get("fields") is not None and both.any():
        # per case max|f - f_Thomas| / max|f_Thomas| over the field, then the
        # largest and the median over the cases alive in both runs
        errs = {}
        for f in FIELD_CMP:
            e = []
            for i in np.where(both)[0]:
                a, r = res["fields"][f][i], ref["fields"][f][i]
                den = float(np.max(np.abs(r))) if r.size else 0.0
                d = float(np.max(np.abs(a - r))) if r.size else 0.0
                e.append(d / den if den > 0.0 else d)
            errs[f] = e
        qr = np.abs(ref["qwall"][both])
        errs["QWALL"] = list(np.abs(res["qwall"][both] - ref["qwall"][both]) / np.where(qr > 0.0, qr, 1.0))
        errs["PLACE"] = list(rel)
        out["field_rel_diff_vs_thomas"] = {
            "cases": int(both.sum()),
            "max": {k: float(np.max(v)) for k, v in errs.items()},
            "median": {k: float(np.median(v)) for k, v in errs.items()}}
    return out


def train_net(L, args):
    """Pairs from a disjoint batch, then Adam on the GPU (not timed)."""
    from lica_boiling1d_amd.learned import Net, collect_pairs
    t0 = time.perf_counter()
    net = Net(seed=args.seed)
    tcases = L.batched_cases(args.train_cases, seed=20261203)
    nwin = collect_pairs(net, tcases, steps=args.train_steps, windows_per_case=args.train_windows, window=256,
                         warmup=args.warmup)
    t1 = time.perf_counter()
    loss = net.train(args.train_iters, batch=64, lr=1e-3, seed=1)
    t2 = time.perf_counter()
    return net, {"pairs": nwin, "window_cells": 256, "cases": args.train_cases, "seed": 20261203,
                 "steps": args.train_steps, "iters": args.train_iters, "minibatch": 64, "final_loss": loss,
                 "collect_s": t1 - t0, "train_s": t2 - t1}


def run_summary(L, res, args):
    out = {"value": res["value"], "ms_per_step": res["el"] * 1000.0 / args.steps}
    if "vcycles_per_solve" in res:
        out["vcycles_per_solve"] = res["vcycles_per_solve"]
        out["poisson_backward_error_max"] = res["resid_max"]
    if res["nbad"]:
        out["stopped_cases"] = stopped(L, res)
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of `mib` MiB (torch's copy kernel), five
    times: each dispatch reads and writes exactly mib MiB, so the ratio of
    a counter's reading to that size scales the same counter elsewhere."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 8, "dispatches": 5}), flush=True)


# algorithmic bytes per row of a line: TRDIAG1 reads A, B, C, R and writes UU
# and the multiplier on the way down, reads both and writes UU on the way up;
# TRDIAG1P reads A, B, C, F and writes Q, S, F down, reads Q, S, F and writes
# S, QE on the backward pass, reads S, QE and writes F at the end; PTDIAG1A
# reads A..F and writes O, Q, R, then reads O, Q, R and writes UU
LINE_BYTES = {"trdiag1": 72, "trdiag1p": 120, "ptdiag1a": 104}


def bench_lines(args):
    """The line solvers of the 3-D code on device arrays: one plane of
    --line-count lines x --line-rows rows (column-major, lines contiguous),
    each solver launched W + K times on the null stream and timed with HIP
    events around each launch; the FP64 oracle on 4096 of the lines as the
    CPU baseline.  Device memory through the HIP runtime the library uses."""
    import ctypes
    import numpy as np
    from lica_boiling1d_amd import _abi as abi
    lib = abi.load()
    rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rt.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rt.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    rt.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def say(m):
        print("bench: " + m, file=sys.stderr, flush=True)

    m1, m2 = args.line_count, args.line_rows
    steps = args.steps if args.steps is not None else 20
    warm = args.warmup if args.warmup is not None else 3
    rng = np.random.default_rng(20261203)
    held = []

    def put(x):
        x = np.asfortranarray(x, dtype=np.float64)
        p = ctypes.c_void_p()
        if rt.hipMalloc(ctypes.byref(p), x.nbytes) != 0 or rt.hipMemcpy(p, x.ctypes.data, x.nbytes, 1) != 0:
            raise RuntimeError("device memory")
        held.append(p)
        return p

    def rnd(lo=-1.0, hi=1.0):
        return rng.uniform(lo, hi, (m1, m2))

    r_h = rnd()
    a, c, r = put(rnd()), put(rnd()), put(r_h)
    b, d, e, a5 = put(rnd(3.0, 4.0)), put(rnd()), put(rnd()), put(rnd(5.0, 6.0))
    zero = np.zeros((m1, m2))
    uu, o, q, rr, f = (put(zero) for _ in range(5))
    work = put(np.zeros((3 * m1, m2)))
    say(f"line solvers on {m1} lines x {m2} rows")
    calls = {
        "trdiag1": lambda: lib.lica_trdiag_lines_dev(a, b, c, r, uu, work, m1, m2, 1, m2, 1, m1, 0, None),
        "trdiag1p": lambda: lib.lica_trdiag_periodic_lines_dev(a, b, c, f, work, m1, m2, 1, m2, 1, m1, None),
        "ptdiag1a": lambda: lib.lica_ptdiag_lines_dev(a5, b, c, d, e, r, uu, m1, m2, 1, m1, 1, m2, o, q, rr, None),
    }
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for x in ev:
        rt.hipEventCreate(ctypes.byref(x))
    res = {}
    for name, call in calls.items():
        ms = 0.0
        for it in range(warm + steps):
            if name == "trdiag1p":
                # the periodic solve overwrites its right-hand side
                rt.hipMemcpy(f, r, m1 * m2 * 8, 3)
            rt.hipEventRecord(ev[0], None)
            if call() != 0:
                raise RuntimeError(name)
            rt.hipEventRecord(ev[1], None)
            rt.hipEventSynchronize(ev[1])
            t = ctypes.c_float()
            rt.hipEventElapsedTime(ctypes.byref(t), ev[0], ev[1])
            if it >= warm:
                ms += t.value
        per = ms / steps
        ach = LINE_BYTES[name] * m1 * m2 / (per * 1e-3) / 1e9
        res[name] = {"ms_per_launch": per, "rows_per_s": m1 * m2 / (per * 1e-3),
                     "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "traffic": None,
                                  "bytes_per_row": LINE_BYTES[name]}}
        say(f"{name}: {per:.3f} ms per launch, {ach:.0f} GB/s")
    for x in ev:
        rt.hipEventDestroy(x)
    for p in held:
        rt.hipFree(p)
    # the oracle on a sample of the plane's lines
    from oracle import binding as O
    ms_n = min(4096, m1)
    A, B, C = (rng.uniform(-1.0, 1.0, (ms_n, m2)) for _ in range(3))
    B += 4.0
    R = r_h[:ms_n]
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 2.0:
        O.trdiag_lines(A, B, C, R, 1, m2, 1, ms_n)
        reps += 1
    cpu = ms_n * m2 * reps / (time.perf_counter() - t0)
    t1 = res["trdiag1"]
    out = {"metric": "tridiagonal rows solved per second (TRDIAG1 lines, 1 GPU)", "value": t1["rows_per_s"],
           "unit": "rows/s", "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": t1["ms_per_launch"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "line solvers", "lines": m1, "rows": m2},
           "roofline": dict(t1["roofline"], kernel="k_trdiag_lines"),
           "cpu_baseline": {"value": cpu, "unit": "rows/s", "cores": 1, "kind": "port",
                            "sample": f"TRDIAG1 on {ms_n} lines x {m2} rows by the FP64 oracle, repeated for 2 s on "
                                      f"one core (NumPy copies in and out included)"},
           "periodic": res["trdiag1p"], "pentadiagonal": res["ptdiag1a"]}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=["large", "batch", "deck", "lines"], default="batch",
                    help="lines: the 3-D code's line solvers (TRDIAG1, TRDIAG1P, PTDIAG1A) on one large plane")
    ap.add_argument("--line-count", type=int, default=262144, help="lines of the plane (--workload lines)")
    ap.add_argument("--line-rows", type=int, default=128, help="rows per line (--workload lines)")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--mode", choices=["learned", "projected", "direct", "thomas", "mg64", "mixed"], default="learned")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=0, help="case sets on as many streams (1..8); 0: 4 for >= 256 cases")
    ap.add_argument("--graph", action="store_true", help="replay each step as a captured hipGraph")
    ap.add_argument("--split", choices=("binned", "interleaved"), default="binned",
                    help="lanes as size bins of equal work (own launch shapes) or every L-th case")
    ap.add_argument("--tune", type=int, default=2,
                    help="trial steps per launch shape of the autotuner, after the warmup (0: default shapes)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--train-cases", type=int, default=4096)
    ap.add_argument("--train-steps", type=int, default=25)
    ap.add_argument("--train-windows", type=int, default=1)
    ap.add_argument("--train-iters", type=int, default=3000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the comparison runs")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (the known-size kernel that scales the HBM "
                         "byte counters of a profile) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the cases of a batch are "
              "independent, so more GPUs would shard the batch by case with no exchange.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 20 if args.workload == "batch" else 200
    if args.workload == "lines":
        bench_lines(args)
        return
    if args.warmup is None:
        args.warmup = 5

    import lica_boiling1d_amd as L
    if args.workload == "large":
        cases = [L.single_large_case(seed=1)]
        wl = {"workload": "single large case", "n": 8192, "r": 4, "superheat_K": 10.0, "film_mm": 0.2,
              "perturbation": "1% sin, seed 1", "instances": 1}
        wkey = "large"
    elif args.workload == "deck":
        cases = [L.deck_case()]
        wl = {"workload": "shipped deck case", "n": 256, "r": 2, "instances": 1}
        wkey = "deck"
    else:
        cases = L.batched_cases(args.batch)
        wl = {"workload": "batched cases", "instances": args.batch, "n": "U{128..4096}", "r": 2,
              "superheat_K": "U(2,20)", "film_mm": "U(0.05,0.5)", "seed": 20261202}
        wkey = f"batch{args.batch}"
    mode = {"learned": L.POISSON_LEARNED, "projected": L.POISSON_PROJECTED, "direct": L.POISSON_DIRECT,
            "thomas": L.POISSON_THOMAS, "mg64": L.POISSON_MG64, "mixed": L.POISSON_MIXED}[args.mode]
    wl["poisson"] = args.mode
    wl["lanes"] = lanes_of(cases, args)
    wl["lane_split"] = args.split
    wl["autotune_steps"] = args.tune
    wl["graph"] = bool(args.graph)
    if mode >= L.POISSON_MG64:
        wl["poisson_stop"] = f"backward error <= {args.rtol:g} and last correction <= {args.rtol:g} max|p|"

    def say(m):
        print("bench: " + m, file=sys.stderr, flush=True)

    say(f"{wl['workload']} x{len(cases)}, mode {args.mode}, warmup {args.warmup}, steps {args.steps}")
    net, train = None, None
    if mode == L.POISSON_LEARNED:
        args.train_cases = min(args.train_cases, max(len(cases), 64))
        say(f"training the warm-start network: {args.train_cases} cases x {args.train_steps} steps of pairs")
        net, train = train_net(L, args)
        say(f"trained: {train['pairs']} windows, loss {train['final_loss']:.3e}, "
            f"{train['collect_s']:.1f} s + {train['train_s']:.1f} s")
    say(f"timed run, mode {args.mode}")
    res = timed_run(L, cases, mode, args, net, keep_fields=not args.no_fp64)
    if res["done"] == 0:
        print("bench: no case advanced", file=sys.stderr)
        sys.exit(1)
    say(f"{res['value']:.1f} instance-steps/s")
    dtype = ("f64+f32+bf16" if mode == L.POISSON_LEARNED else
             "f64+f32" if mode in (L.POISSON_MIXED, L.POISSON_PROJECTED) else "f64")
    out = {"metric": "instance-steps/s at Poisson resid 1e-10: mixed+learned vs plain FP64 (1 GPU)",
           "value": res["value"], "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["el"] * 1000.0 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": wl,
           "cell_updates_per_s": res["cells"]}
    if "vcycles_per_solve" in res:
        out["vcycles_per_solve"] = res["vcycles_per_solve"]
        out["poisson_backward_error_max"] = res["resid_max"]
    if res["nbad"]:
        out["stopped_cases"] = stopped(L, res)
    out["launch_shapes"] = res["tuning"]
    out["step_roofline"] = step_roofline(cases, mode, res)
    say("one-lane kernel profile")
    per, alive = kernel_profile(L, cases, mode, args, net)
    out["roofline"] = roofline(L, cases, mode, per, alive, wkey)
    out["kernel_ms_per_launch"] = per
    if train is not None:
        out["training"] = train
    if not args.no_fp64:
        runs = {}
        if mode == L.POISSON_LEARNED:
            say("timed run, mixed+projected (no network)")
            runs["mixed_projected"] = timed_run(L, cases, L.POISSON_PROJECTED, args, keep_fields=True)
        if mode in (L.POISSON_LEARNED, L.POISSON_PROJECTED):
            say("timed run, mixed from zero (no warm start)")
            runs["mixed_from_zero"] = timed_run(L, cases, L.POISSON_MIXED, args, keep_fields=True)
        if mode != L.POISSON_MG64:
            say("timed run, plain FP64 V-cycles from zero")
            runs["plain_fp64"] = timed_run(L, cases, L.POISSON_MG64, args, keep_fields=True)
        ulp = None
        if mode != L.POISSON_THOMAS:
            say("timed run, Thomas (the reference's direct solve)")
            runs["thomas"] = timed_run(L, cases, L.POISSON_THOMAS, args, keep_fields=True)
            # the same with every case's film thickness z0 one ulp up: how far
            # rounding alone moves each field (the yardstick of the differences
            # of the multigrid runs from the Thomas run)
            say("timed run, Thomas with every case's z0 one ulp up")
            import numpy as np
            ulp_cases = [c.copy(z0=float(np.nextafter(c.z0, 1.0))) for c in cases]
            ulp = timed_run(L, ulp_cases, L.POISSON_THOMAS, args, keep_fields=True)
        ref = runs.get("thomas", res)
        for k, r in runs.items():
            out[k] = run_summary(L, r, args)
        if "mixed_projected" in runs:
            pr = runs["mixed_projected"]
            out["network_gain_vs_projected"] = res["value"] / pr["value"] if pr["value"] > 0 else None
            out["vcycles_saved_by_network_per_solve"] = pr.get("vcycles_per_solve", 0.0) - res.get(
                "vcycles_per_solve", 0.0)
        if "mixed_from_zero" in runs:
            mz = runs["mixed_from_zero"]
            pr = runs.get("mixed_projected", res)
            out["projection_gain_vs_mixed_from_zero"] = pr["value"] / mz["value"] if mz["value"] > 0 else None
            out["vcycles_saved_by_projection_per_solve"] = mz.get("vcycles_per_solve", 0.0) - pr.get(
                "vcycles_per_solve", 0.0)
        if "plain_fp64" in runs:
            pf = runs["plain_fp64"]
            out["speedup_vs_plain_fp64"] = res["value"] / pf["value"] if pf["value"] > 0 else None
            out["vcycles_saved_per_solve"] = pf.get("vcycles_per_solve", 0.0) - res.get("vcycles_per_solve", 0.0)
        out["versus_thomas"] = {"this_run": versus_thomas(res, ref)}
        for k, r in runs.items():
            if k != "thomas":
                out["versus_thomas"][k] = versus_thomas(r, ref)
        if ulp is not None:
            out["versus_thomas"]["thomas_z0_one_ulp_up"] = versus_thomas(ulp, ref)
    if not args.no_cpu:
        say("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(cases, wl["workload"], args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
