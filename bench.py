"""Throughput of the pseudo-time inner iteration on one MI355X.

Workload (default): turbulent channel 129x129x129, SA/DES eddy viscosity
(itk = 2), 3-level FAS V-cycle (129 -> 65 -> 33, iter = 1, 2, 3 smoother
passes), BDF2 dual time stepping, FP64 plain.  A timed step is one physical
time step: one desalm_mg_driver call of `--iters` pseudo-iterations (no early
exit), inputs resident in HBM.  value = fine nodes x pseudo-iterations /
wall second (Mcell-iters/s).

The roofline object is computed live for the dominant kernel family from
HIP events recorded by the library on its stream during the timed steps:
achieved GB/s = algorithmic bytes per launch / mean launch time.  Algorithmic
bytes count each compulsory FP64 array read or written once per node of the
level (scratch excluded).  The CPU baseline times the C oracle (a
single-core restatement of the reference algorithm) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "des-alm_amd", "python"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "Mcell-iters/s + wall s/physical step to 1e-8 resid, FP64 plain vs mixed+learned"
HBM_PEAK_GBS = 8000.0

# compulsory FP64 words per node of the level, per launch (reads + writes)
WORDS = {
    "dtau": 9 + 1 + 1,                        # metrics, xnut -> dtau
    "residual": 4 + 9 + 1 + 1 + 3 + 3 + 1 + 4 + 1,   # q,metrics,aj,xnut,qn,qnm1,dtau -> rh,diss1 (fine)
    "residual_recalc": 4 + 9 + 1 + 1 + 3 + 3 + 1 + 4,  # ... diss1 -> rh
    "daf_i": 4 + 1 + 1 + 3 + 4,               # rh,dtau,aj,csi -> rh
    "daf_j": 4 + 1 + 3 + 3 + 4,               # rh,dtau,csi,eta -> rh
    "daf_k": 4 + 1 + 3 + 3 + 4 + 4,           # rh,dtau,eta,zet,q -> q
}


def level_words(name, n):
    w = WORDS[name]
    if name == "residual" and n > 1:
        w += 4 + 4  # pk read and (first pass) written
    elif name == "residual_recalc" and n > 1:
        w += 4
    return w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=10, help="pseudo-iterations per physical step")
    ap.add_argument("--n", type=int, default=129, help="nodes per direction")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=6)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    from desalm import cases
    from desalm.lib import Solver

    case = cases.channel(args.n, args.n, args.n, ng=3, itm=args.iters, iters=(1, 2, 3))
    p = case.params
    n1 = case.n1
    s = Solver(p, device=0).load_case(case)
    for _ in range(args.warmup):
        s.mg_driver()
    s.profile(True)
    iters_done = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itc, hist = s.mg_driver()
        iters_done += itc
    t1 = time.perf_counter()
    stats = s.profile_stats()
    s.profile(False)
    wall = t1 - t0
    value = n1 * iters_done / wall / 1e6
    q = s.get_state()[0]
    finite = bool(np.all(np.isfinite(q)))

    # dominant kernel family by total device time
    fam = {}
    for (nm, n), (cnt, ms) in stats.items():
        fam.setdefault(nm, [0, 0.0])
        fam[nm][0] += cnt
        fam[nm][1] += ms
    dominant = max((k for k in fam if k in WORDS), key=lambda k: fam[k][1])
    dims = p.level_dims()
    cnt1, ms1 = stats[(dominant, 1)]
    bytes_per_launch = level_words(dominant, 1) * 8 * dims[0][0] * dims[0][1] * dims[0][2]
    achieved = bytes_per_launch / (ms1 / cnt1 * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": f"{dominant} (fine level)", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": bytes_per_launch,
                "mean_launch_ms": round(ms1 / cnt1, 4)}

    cpu = None
    if not args.no_cpu:
        import pyoracle as po
        cc = cases.channel(args.n, args.n, args.n, ng=3, itm=args.cpu_iters, iters=(1, 2, 3))
        o = po.Oracle(cc.params).load_case(cc)
        tc0 = time.perf_counter()
        itc_c, _ = o.mg_driver()
        tc1 = time.perf_counter()
        cpu = {"value": round(n1 * itc_c / (tc1 - tc0) / 1e6, 4), "unit": "Mcell-iters/s", "cores": 1,
               "kind": "port",
               "sample": f"{itc_c} pseudo-iterations of the same {args.n}^3 channel V-cycle, C oracle, 1 thread "
                         f"({tc1 - tc0:.1f} s)"}

    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mcell-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"channel {args.n}^3 SA/DES, 3-level FAS V-cycle (iter 1,2,3), "
                                  f"BDF2 dual time, FP64 plain, {args.iters} pseudo-iterations per step",
                      "fine_nodes": n1, "pseudo_iterations": iters_done, "parallelism": "single GPU"},
           "roofline": roofline, "cpu_baseline": cpu,
           "kernels_ms": {f"{nm}@L{n}": [cnt, round(ms, 3)] for (nm, n), (cnt, ms) in sorted(stats.items())},
           "state_finite": finite, "pivot_events": s.pivot_events,
           "final_log10_update_norms": [round(float(v), 4) for v in hist[-1]]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
