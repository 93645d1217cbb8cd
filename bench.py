"""Benchmark of the boiling path on one MI355X.

Workload (default): the single large case -- base grid N=8192 cells,
level-set refinement r=4, 10 K superheat (Ja=0.009), a 0.2 mm film
(z0=0.0607) with a seeded 1 % sinusoidal height perturbation -- advanced by
the plain-FP64 GPU path, Poisson solved directly (residual at FP64 rounding,
far below 1e-10).  A "step" is one full RK3 time step of the case
(1d.f:196-295).  The value is instance-steps/s with the state resident in
HBM.  ``--workload batch`` runs independent cases of uneven size in one
launch per kernel instead.

The JSON line also carries
  roofline      the dominant kernel's algorithmic bytes per launch over its
                mean launch time (HIP events between the launches on the
                library's stream, over the timed steps), against 8 TB/s;
  cpu_baseline  the oracle (serial FP64 restatement, one core) on a bounded
                number of steps of the same case.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ["init", "dt", "cont_energy", "temp_grad", "momentum", "poisson", "ucalc", "transport"]


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


def cpu_baseline(case, seconds=15.0):
    """Oracle (serial FP64 restatement, one core) on the same case."""
    from oracle import Oracle
    o = Oracle(case)
    t0 = time.perf_counter()
    steps = 0
    while True:
        st = o.step(1)
        steps += 1
        if st != 0:
            break
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 100000:
            break
    el = time.perf_counter() - t0
    o.close()
    return {"value": steps / el, "unit": "instance-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of the same case (N={case.n}, r={case.r}) on one host core, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["large", "batch", "deck"], default="large")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--mode", choices=["direct", "thomas", "mg64", "mixed"], default="direct")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the cases of a batch are "
              "independent, so more GPUs would shard the batch by case with no exchange.", file=sys.stderr)
        sys.exit(2)

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
    mode = {"direct": L.POISSON_DIRECT, "thomas": L.POISSON_THOMAS, "mg64": L.POISSON_MG64,
            "mixed": L.POISSON_MIXED}[args.mode]
    wl["poisson"] = args.mode
    print(f"bench: {wl['workload']} x{len(cases)}, mode {args.mode}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    b = L.Batch(cases)
    b.set_poisson(mode, args.rtol, 200)
    b.set_block(args.threads)
    if args.graph:
        b.use_graph(True)
    # a case that reaches one of the reference's STOP conditions keeps its
    # status and is left behind; only steps the cases completed are counted
    if args.warmup:
        b.run_all(args.warmup)
    b.sync()
    _, steps0 = b.progress()
    # timed region: per-kernel HIP events between the launches (no graph)
    timed_graph = args.graph
    if not timed_graph:
        b.set_timing(True)
    b.sync()
    t0 = time.perf_counter()
    nbad = b.run_all(args.steps)
    b.sync()
    el = time.perf_counter() - t0
    tim = b.timings() if not timed_graph else None
    st, steps1 = b.progress()
    done = int((steps1 - steps0).sum())
    if done == 0:
        print("bench: no case advanced", file=sys.stderr)
        sys.exit(1)
    ms_step = el * 1000.0 / args.steps
    value = done / el
    out = {"metric": "instance-steps/s at Poisson resid 1e-10: mixed+learned vs plain FP64 (1 GPU)",
           "value": value, "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic", "config": wl}
    out["cell_updates_per_s"] = float(sum(int(c.n) * int(d) for c, d in zip(cases, steps1 - steps0))) / el
    if mode >= L.POISSON_MG64:
        cyc = sum(b.scalars(i)["POISSON_CYCLES"] for i in range(min(len(cases), 64)))
        sol = sum(b.scalars(i)["POISSON_SOLVES"] for i in range(min(len(cases), 64)))
        out["vcycles_per_solve"] = cyc / max(sol, 1.0)
        out["poisson_backward_error_max"] = max(b.scalars(i)["POISSON_RESID"] for i in range(min(len(cases), 64)))
    if nbad:
        import numpy as np
        codes, counts = np.unique(st[st != 0], return_counts=True)
        out["stopped_cases"] = {L.status_string(int(k)): int(v) for k, v in zip(codes, counts)}
    if tim is not None:
        launches = {"dt": 1, "cont_energy": 3, "temp_grad": 3, "momentum": 3, "poisson": 3, "ucalc": 3, "transport": 3}
        if mode >= L.POISSON_MG64:
            tim["poisson"] = tim.get("poisson_mg", 0.0)
        per = {}
        for k, nl in launches.items():
            per[k] = tim.get(k, 0.0) / (args.steps * nl)   # ms per launch
        dom = max(per, key=lambda k: per[k])
        vc = out.get("vcycles_per_solve", 0.0)
        bytes_launch = sum(kernel_bytes(c, dom, mode, vc) for c in cases)
        ach = bytes_launch / (per[dom] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "k_" + dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                           "ms_per_launch": per[dom], "bytes_per_launch": bytes_launch}
        out["kernel_ms_per_launch"] = per
    if not args.no_cpu:
        big = max(cases, key=lambda c: int(c.n))
        cb = cpu_baseline(big if len(cases) == 1 else cases[0], args.cpu_seconds)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
