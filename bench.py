"""Benchmark of the boiling path on one MI355X.

Workload (default): 4096 independent cases of uneven size in one launch per
kernel -- base grid N ~ U{128..4096}, superheat U(2,20) K, film thickness
U(0.05,0.5) mm, refinement r=2, drawn from SplitMix64 seed 20261202.  A
"step" is one full RK3 time step of every case (1d.f:196-295); the value is
completed instance-steps/s with the state resident in HBM, every Poisson
solve taken by V-cycles to a normwise backward error of 1e-10.

Three timed runs of the same batch from the same initial state, each over
the same window of steps after the same untimed warmup (past the start-up
transient in which the pressure field builds up from zero):
  mixed+learned    FP32 V-cycle corrections in FP64 defect correction,
                   started from the history projection corrected by the
                   warm-start network (the value);
  mixed+projected  the same without the network (what the network adds);
  plain FP64       FP64 V-cycles from a zero initial guess (IOLDV=0).
Before them the network is trained on the GPU on pairs the solver generates
from a disjoint batch (seed 20261203); the training is not timed.

The JSON line also carries
  roofline      the dominant kernel's algorithmic bytes per launch over its
                mean launch time (HIP events between the launches on each
                lane's stream, over the timed steps), against 8 TB/s; with
                four lanes (the cases split into four sets whose launch
                sequences run on four streams at once) a launch covers a
                quarter of the cases, and bytes and times are both per
                launch;
  cpu_baseline  the oracle (serial FP64 restatement, one core) on a bounded
                sample of the same cases.
``--workload large`` / ``deck`` and ``--mode`` select the single large case,
the shipped deck and the other Poisson modes.
"""
import argparse
import json
import os
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


def kernel_bytes(case, kernel, mode, vcycles=0.0):
    """Algorithmic HBM bytes of one launch for one case: every distinct
    global array the kernel touches, elements touched x 8 B (4 B for int
    arrays), twice when read and written.  Band-sized work (a few dozen
    refined planes) is counted with its own arrays."""
    n = int(case.n)
    per_cell = {
        # PSI_CN read, RK3ZO written
        "dt": 16,
        # W,PSI_ZN,PSI_CN,VVDZ,AKUV,CKUV read; T r/w; DF_Z,VF_Z,DF_C,PSI_Z,PSI_C,RHSPS
        # written; four tridiagonal work vectors written and read back
        "cont_energy": 6 * 8 + 16 + 6 * 8 + 4 * 16,
        # W,PSI_Z,PSI_C,RK3ZO r/w; PSI_CN,DF_Z,VF_Z,SSDZ,VVDZ,P,PSI_ZN,AKW,CKW read;
        # IMASK (int); KCOU1/KCOU2 (int); tridiagonal work r/w; Poisson matrix written
        "momentum": 4 * 16 + 9 * 8 + 4 + 8 + 4 * 16 + 4 * 8,
        # Thomas: a r/w (pivots), b,c read, g r/w, P written; partitioned: a,b,c,g r/w;
        # multigrid: read a,b,c,g and write P, build the level hierarchy (about
        # 2n rows of a,b,c,x,f,r), then per V-cycle 384 B per fine cell in FP64
        # (2+2 smoothing sweeps, residual, transfers over all levels), half of
        # that for the FP32 cycles plus the FP64 defect correction (32 B)
        "poisson": ((16 + 8 + 8 + 16 + 8) if mode == 0 else
                    (4 * 16 + 8) if mode == 1 else
                    int(40 + 2 * 48 + vcycles * (384 if mode == 2 else 192 + 32))),
        # W,P,WO r/w; PSI_Z,VVDZ,SSDZ read; KCOU1/KCOU2; WT written
        "ucalc": 3 * 16 + 3 * 8 + 8 + 8,
    }
    if kernel in per_cell:
        return per_cell[kernel] * n
    # band kernels: 11 layers of <=16 planes, refined window of ~200 planes of
    # PSIF/SFGAS/geometry, NS window of ~100 cells of T/PSI_CN/geometry
    r = int(case.r)
    band = 11 * 16 * 8 * 4
    if kernel == "transport":
        return band + 200 * 8 * 4 + 100 * 8 * 6 + 12 * r
    if kernel == "temp_grad":
        return band + 200 * 8 * 6 + 100 * 8 * 8
    return 0


def cpu_baseline(cases, seconds=15.0):
    """Oracle (serial FP64 restatement, one core) on a bounded sample: the
    cases in order, a few steps each (every step of a single case), until
    `seconds` have passed."""
    from oracle import Oracle
    per_case = 3 if len(cases) > 1 else 100000
    t0 = time.perf_counter()
    steps, ncase, cells = 0, 0, 0
    for c in cases:
        o = Oracle(c)
        ncase += 1
        for _ in range(per_case):
            st = o.step(1)
            steps += 1
            cells += int(c.n)
            if st != 0 or time.perf_counter() - t0 >= seconds:
                break
        o.close()
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    if len(cases) == 1:
        sample = f"{steps} steps of the same case (N={cases[0].n}, r={cases[0].r}) on one host core, {el:.1f} s"
    else:
        sample = (f"first {ncase} cases of the same batch, up to {per_case} steps each ({steps} instance-steps, "
                  f"mean N={cells / max(steps, 1):.0f}) on one host core, {el:.1f} s, case set-up included")
    return {"value": steps / el, "unit": "instance-steps/s", "cores": 1, "kind": "port", "sample": sample}


def lanes_of(cases, args):
    """Case sets run side by side on as many streams (2..4) or one sequence (1)."""
    return args.lanes if args.lanes else (LANES_BATCH if len(cases) >= 256 else 1)


def timed_run(L, cases, mode, args, net=None):
    """One timed run: warmup steps, then args.steps steps with per-kernel HIP
    events; only the steps the cases completed are counted (a case that
    reaches one of the reference's STOP conditions keeps its status and is
    left behind)."""
    import numpy as np
    b = L.Batch(cases)
    if net is not None:
        b.set_net(net)
    b.set_poisson(mode, args.rtol, 200)
    b.set_block(args.threads)
    b.set_lanes(lanes_of(cases, args))
    if args.warmup:
        b.run_all(args.warmup)
    b.sync()
    s0 = [b.scalar_array(i) for i in range(len(cases))]
    _, steps0 = b.progress()
    b.set_timing(True)
    b.sync()
    t0 = time.perf_counter()
    nbad = b.run_all(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    tim = b.timings()
    st, steps1 = b.progress()
    s1 = [b.scalar_array(i) for i in range(len(cases))]
    b.close()
    done = steps1 - steps0
    ok = (st == 0)
    res = {"el": el, "done": int(done.sum()), "value": float(done.sum()) / el, "timings": tim,
           "cells": float(sum(int(c.n) * int(d) for c, d in zip(cases, done))) / el,
           "ok": ok, "place": np.array([s[3] for s in s1]), "nbad": nbad, "status": st}
    if mode >= L.POISSON_MG64:
        ci, si, ri = (L.SCALARS.index(k) for k in ("POISSON_CYCLES", "POISSON_SOLVES", "POISSON_RESID"))
        cyc = sum(a[ci] - z[ci] for a, z, g in zip(s1, s0, ok) if g)
        sol = sum(a[si] - z[si] for a, z, g in zip(s1, s0, ok) if g)
        res["vcycles_per_solve"] = cyc / max(sol, 1.0)
        res["resid_max"] = max((a[ri] for a, g in zip(s1, ok) if g), default=0.0)
    return res


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


def roofline(L, cases, mode, args, res):
    """Dominant kernel of a timed run: algorithmic bytes per launch (the cases
    still running) over its mean launch time."""
    launches = {"dt": 1, "cont_energy": 3, "temp_grad": 3, "momentum": 3, "poisson": 3, "ucalc": 3, "transport": 3,
                "warm_start": 3}
    tim = dict(res["timings"])
    if mode >= L.POISSON_MG64:
        tim["poisson"] = tim.get("poisson_mg", 0.0)
    # with L lanes every kernel is launched L times per substep, each launch
    # on 1/L of the cases: ms per launch and bytes per launch both divide by L
    nla = lanes_of(cases, args)
    per = {k: tim.get(k, 0.0) / (args.steps * nl * nla) for k, nl in launches.items()}   # ms per launch
    # the warm-start slot (projection, network) is reported, not rated
    dom = max((k for k in per if k != "warm_start"), key=lambda k: per[k])
    vc = res.get("vcycles_per_solve", 0.0)
    live = [c for c, g in zip(cases, res["ok"]) if g]
    bytes_launch = sum(kernel_bytes(c, dom, mode, vc) for c in live) / nla
    ach = bytes_launch / (per[dom] * 1e-3) / 1e9
    return ({"bound": "hbm", "kernel": "k_" + ("poisson_mg" if dom == "poisson" and mode >= L.POISSON_MG64 else dom),
             "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
             "ms_per_launch": per[dom], "bytes_per_launch": bytes_launch}, per)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=["large", "batch", "deck"], default="batch")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--mode", choices=["learned", "projected", "direct", "thomas", "mg64", "mixed"], default="learned")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=0, help="case sets on as many streams (1..4); 0: 4 for >= 256 cases")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--train-cases", type=int, default=4096)
    ap.add_argument("--train-steps", type=int, default=25)
    ap.add_argument("--train-windows", type=int, default=1)
    ap.add_argument("--train-iters", type=int, default=3000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain-FP64 comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the cases of a batch are "
              "independent, so more GPUs would shard the batch by case with no exchange.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 60 if args.workload == "batch" else 200
    if args.warmup is None:
        args.warmup = 20

    import numpy as np
    import lica_boiling1d_amd as L
    if args.workload == "large":
        cases = [L.single_large_case(seed=1)]
        wl = {"workload": "single large case", "n": 8192, "r": 4, "superheat_K": 10.0, "film_mm": 0.2,
              "perturbation": "1% sin, seed 1", "instances": 1}
    elif args.workload == "deck":
        cases = [L.deck_case()]
        wl = {"workload": "shipped deck case", "n": 256, "r": 2, "instances": 1}
    else:
        cases = L.batched_cases(args.batch)
        wl = {"workload": "batched cases", "instances": args.batch, "n": "U{128..4096}", "r": 2,
              "superheat_K": "U(2,20)", "film_mm": "U(0.05,0.5)", "seed": 20261202}
    mode = {"learned": L.POISSON_LEARNED, "projected": L.POISSON_PROJECTED, "direct": L.POISSON_DIRECT,
            "thomas": L.POISSON_THOMAS, "mg64": L.POISSON_MG64, "mixed": L.POISSON_MIXED}[args.mode]
    wl["poisson"] = args.mode
    wl["lanes"] = lanes_of(cases, args)
    if mode >= L.POISSON_MG64:
        wl["poisson_backward_error"] = args.rtol

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
    res = timed_run(L, cases, mode, args, net)
    if res["done"] == 0:
        print("bench: no case advanced", file=sys.stderr)
        sys.exit(1)
    say(f"{res['value']:.1f} instance-steps/s")
    out = {"metric": "instance-steps/s at Poisson resid 1e-10: mixed+learned vs plain FP64 (1 GPU)",
           "value": res["value"], "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["el"] * 1000.0 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": ("f64+f32+bf16" if mode == L.POISSON_LEARNED else
                     "f64+f32" if mode in (L.POISSON_MIXED, L.POISSON_PROJECTED) else "f64"),
           "data": "synthetic", "config": wl, "cell_updates_per_s": res["cells"]}
    if "vcycles_per_solve" in res:
        out["vcycles_per_solve"] = res["vcycles_per_solve"]
        out["poisson_backward_error_max"] = res["resid_max"]
    out["roofline"], out["kernel_ms_per_launch"] = roofline(L, cases, mode, args, res)
    if res["nbad"]:
        st = res["status"]
        codes, counts = np.unique(st[st != 0], return_counts=True)
        out["stopped_cases"] = {L.status_string(int(k)): int(v) for k, v in zip(codes, counts)}
    if train is not None:
        out["training"] = train
    if mode == L.POISSON_LEARNED and not args.no_fp64:
        say("timed run, mixed+projected (no network)")
        pr = timed_run(L, cases, L.POISSON_PROJECTED, args)
        out["mixed_projected"] = {"value": pr["value"], "ms_per_step": pr["el"] * 1000.0 / args.steps,
                                  "vcycles_per_solve": pr.get("vcycles_per_solve"),
                                  "poisson_backward_error_max": pr.get("resid_max"),
                                  "kernel_ms_per_launch": roofline(L, cases, L.POISSON_PROJECTED, args, pr)[1]}
        if pr["nbad"]:
            st = pr["status"]
            codes, counts = np.unique(st[st != 0], return_counts=True)
            out["mixed_projected"]["stopped_cases"] = {L.status_string(int(k)): int(v) for k, v in zip(codes, counts)}
        out["network_gain_vs_projected"] = res["value"] / pr["value"] if pr["value"] > 0 else None
        out["vcycles_saved_by_network_per_solve"] = pr.get("vcycles_per_solve", 0.0) - res.get("vcycles_per_solve", 0.0)
    if mode >= L.POISSON_MIXED and not args.no_fp64:
        say("timed run, plain FP64 V-cycles from zero")
        ref = timed_run(L, cases, L.POISSON_MG64, args)
        both = res["ok"] & ref["ok"]
        rel = np.abs(res["place"][both] - ref["place"][both]) / np.abs(ref["place"][both])
        out["plain_fp64"] = {"value": ref["value"], "ms_per_step": ref["el"] * 1000.0 / args.steps,
                             "vcycles_per_solve": ref.get("vcycles_per_solve"),
                             "poisson_backward_error_max": ref.get("resid_max"),
                             "kernel_ms_per_launch": roofline(L, cases, L.POISSON_MG64, args, ref)[1]}
        if ref["nbad"]:
            st = ref["status"]
            codes, counts = np.unique(st[st != 0], return_counts=True)
            out["plain_fp64"]["stopped_cases"] = {L.status_string(int(k)): int(v) for k, v in zip(codes, counts)}
        out["speedup_vs_plain_fp64"] = res["value"] / ref["value"] if ref["value"] > 0 else None
        out["vcycles_saved_per_solve"] = ref.get("vcycles_per_solve", 0.0) - res.get("vcycles_per_solve", 0.0)
        out["interface_rel_diff_vs_plain_fp64_max"] = float(rel.max()) if rel.size else None
    if not args.no_cpu:
        say("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(cases, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
