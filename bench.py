"""Throughput of the pseudo-time inner iteration on one MI355X.

Workload (default, the one the metric is quoted on): SA-DES dual time
stepping of a uniform inflow U = 1 past one actuator line (3 blades, radius
0.5, tip-speed ratio 6) on a 193x129x129 stretched box, dt = 0.01, 3-level
FAS V-cycle (iter = 1, 2, 3 smoother passes), SA itk = 2, FP64 plain.  A
timed step is one physical time step (desalm_step: the blades advance, then
`--iters` pseudo-iterations, no early exit, each pseudo-iteration one HIP
graph replay), inputs resident in HBM.
value = fine nodes x pseudo-iterations / wall second (Mcell-iters/s).

The roofline object is computed live for the dominant kernel family from
HIP events recorded by the library on its stream around every launch during
a second, profiled pass of the same timed steps (launches then run
eagerly): achieved GB/s = algorithmic bytes per launch / mean launch time.
Algorithmic bytes count each compulsory FP64 array read or written once per
node of the level (scratch excluded).  The CPU baseline times the
reference's own Fortran (built from its sources, when oracle/_ref is
present) or else the C restatement, on a bounded sample of the same
workload.
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

# compulsory FP64 words per node of the level, per launch (reads + writes):
# each live input array read once and each output written once, the
# library's scratch arrays excluded
WORDS = {
    "dtau": 9 + 1 + 1,                        # metrics, xnut -> dtau
    "residual": 4 + 9 + 1 + 1 + 3 + 3 + 1 + 4 + 1,   # q,metrics,aj,xnut,qn,qnm1,dtau -> rh,diss1 (fine)
    "residual_recalc": 4 + 9 + 1 + 1 + 3 + 3 + 1 + 4,  # q,metrics,aj,xnut,qn,qnm1,diss1 -> rh
    "daf_i": 4 + 1 + 1 + 3 + 4,               # rh,dtau,aj,csi -> rh
    "daf_j": 4 + 1 + 3 + 3 + 4,               # rh,dtau,csi,eta -> rh
    "daf_k": 4 + 1 + 3 + 3 + 4 + 4,           # rh,dtau,eta,zet,q -> q
    # des_eddy, per call: itk x (explicit SA rhs + update 22 words, three ADI
    # sweeps 11 words each) + vorticity source once (13 words); itk applied below
    "des_eddy": None,
    "prolong": 33,                            # coarse q, qold, q, dq of the three smoothing sweeps
    "resid_norm": 16,                         # q, qold -> qold (L1 norms), p (pinned-pressure shift)
}


def level_words(name, n, itk=2):
    if name == "des_eddy":
        return itk * (22 + 3 * 11) + 13
    w = WORDS[name]
    if name == "residual" and n > 1:
        w += 4 + 4  # pk read and (first pass) written
    elif name == "residual_recalc" and n > 1:
        w += 4
    return w


# kernel-name prefix -> family, for the fine-level dispatches of a counter run
# (k_prep, k_residual and k_visc_pk are told apart by their place in the pass)
FAMILY_OF = [("k_sa_", "des_eddy"), ("k_daf_sweep<0", "daf_i"), ("k_daf_sweep<1", "daf_j"),
             ("k_daf_sweep<2", "daf_k"), ("k_prolong_fill", "prolong"), ("k_rs", "prolong"),
             ("k_resid_", "resid_norm"), ("k_bc_fm", "bcond_fm"), ("k_pfix_", "bcond_fm")]
# one dispatch per family call
FAMILY_MARK = {"des_eddy": "k_sa_prep", "daf_i": "k_daf_sweep<0", "daf_j": "k_daf_sweep<1", "daf_k": "k_daf_sweep<2",
               "prolong": "k_prolong_fill", "resid_norm": "k_resid_part", "bcond_fm": "k_bc_fm<2>",
               "dtau": "k_prep", "residual": "k_residual", "residual_recalc": "k_residual"}

# warm-start samples per recorded training step and the sample buffer: every
# one of the 200 training steps stays in the buffer
DT_SAMPLES = 32768
DT_CAPACITY = 1 << 23
HEADLINE_DIMS = (193, 129, 129)
T_START = time.perf_counter()


def log(msg):
    """progress line on stderr (stdout carries only the JSON result)"""
    print(f"bench.py [{time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def workload_name(dims):
    return "actuator-line box %dx%dx%d" % tuple(dims)


def kernel_short(name):
    return name.replace("desalm::(anonymous namespace)::", "").replace("void ", "").split("(")[0].strip()


def classify_dispatches(seq):
    """Fine-level kernel family (or None) of every dispatch of a profiled
    run, given (short kernel name, grid size) in launch order.  Dispatches
    before the first k_copy_qold (the tile autotuner's trial launches) are
    skipped.  A dispatch is fine-level when its grid is the largest of its
    kernel and first template argument (so the i, j and k sweeps and their
    instantiations for short lines are told apart); k_prep, k_residual and
    k_visc_pk belong to the recalc branch when they follow a fine k sweep."""
    first = next((i for i, (nm, _) in enumerate(seq) if nm == "k_copy_qold"), 0)
    key = lambda nm: nm.split(",")[0]  # noqa: E731
    fine_grid = {}
    for nm, g in seq[first:]:
        fine_grid[key(nm)] = max(fine_grid.get(key(nm), 0), g)
    out = [None] * len(seq)
    after_daf = False
    for i in range(first, len(seq)):
        nm, g = seq[i]
        fine = g == fine_grid[key(nm)]
        f = next((fm for pre, fm in FAMILY_OF if nm.startswith(pre)), None)
        if nm in ("k_prep", "k_residual", "k_visc_pk"):
            f = ("residual_recalc" if after_daf else ("dtau" if nm == "k_prep" else "residual")) if fine else None
        elif f not in ("bcond_fm",) and not nm.startswith("__amd") and nm != "k_alm_hub":
            # (the hub zeroing closes bcond_fm on the actuator-line box)
            after_daf = False
        if nm.startswith("k_daf_sweep<2") and fine:
            after_daf = True
        out[i] = f if (f is not None and (fine or nm == "k_resid_final")) else None
    return out


def pmc_summary(fetch_dir, write_dir, dims):
    """HBM bytes per fine-level launch of every kernel family from two
    rocprofv3 counter runs of the same command (FETCH_SIZE, WRITE_SIZE, KiB
    per dispatch).  Each counter is scaled by what it reads for k_copy_qold,
    which moves exactly 5 n^3 doubles each way.  Writes profiles/traffic.json
    and prints a markdown table.  Runs on the CPU."""
    import csv
    import glob

    def short(name):
        return name.replace("desalm::(anonymous namespace)::", "").replace("void ", "").split("(")[0].strip()

    def load(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if r["Counter_Name"] == counter:
                        rows.append((int(r["Dispatch_Id"]), short(r["Kernel_Name"]), int(r["Grid_Size"]),
                                     float(r["Counter_Value"])))
        if not rows:
            raise SystemExit(f"no {counter} rows under {d}")
        return sorted(rows)

    fe, wr = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    if [(a[1], a[2]) for a in fe] != [(b[1], b[2]) for b in wr]:
        raise SystemExit("the two counter runs launched different kernel sequences")
    n1 = dims[0] * dims[1] * dims[2]
    copy_kib = 5 * n1 * 8 / 1024.0
    cf = [v for _, nm, _, v in fe if nm == "k_copy_qold"]
    cw = [v for _, nm, _, v in wr if nm == "k_copy_qold"]
    if not cf:
        raise SystemExit("no k_copy_qold dispatch to calibrate on")
    sf, sw = copy_kib / (sum(cf) / len(cf)), copy_kib / (sum(cw) / len(cw))
    fam = {}
    kern = {}
    fams_of = classify_dispatches([(nm, g) for _, nm, g, _ in fe])
    for (_, nm, g, vf), (_, _, _, vw), f in zip(fe, wr, fams_of):
        rd, wt = vf * sf * 1024.0, vw * sw * 1024.0
        if f is None:
            continue
        e = fam.setdefault(f, {"calls": 0, "read": 0.0, "write": 0.0})
        if nm.startswith(FAMILY_MARK[f]):
            e["calls"] += 1
        e["read"] += rd
        e["write"] += wt
        k = kern.setdefault(f"{nm}@{f}", [0, 0.0, 0.0])
        k[0] += 1
        k[1] += rd
        k[2] += wt
    fams = {}
    for f, e in fam.items():
        c = max(e["calls"], 1)
        fams[f"{f}@1"] = {"hbm_bytes_per_launch": int(round((e["read"] + e["write"]) / c)),
                          "read_bytes": int(round(e["read"] / c)), "write_bytes": int(round(e["write"] / c)),
                          "calls": e["calls"]}
    out = {"workload": workload_name(dims), "calibration": {"kernel": "k_copy_qold", "bytes_each_way": 5 * n1 * 8,
                                            "fetch_scale": round(sf, 4), "write_scale": round(sw, 4)},
           "families": fams,
           "kernels": {k: {"dispatches": c, "read_bytes": int(r / c), "write_bytes": int(w / c)}
                       for k, (c, r, w) in kern.items()}}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"calibration on k_copy_qold ({5 * n1 * 8 / 1e6:.1f} MB each way): FETCH_SIZE x {sf:.4f}, "
          f"WRITE_SIZE x {sw:.4f}\n")
    print("| family (fine level) | calls | read MB | write MB | HBM MB / launch | algorithmic MB | ratio |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for key, e in sorted(fams.items(), key=lambda kv: -kv[1]["hbm_bytes_per_launch"]):
        f = key.split("@")[0]
        hbm = e["hbm_bytes_per_launch"]
        alg = f"{level_words(f, 1) * 8 * n1 / 1e6:.1f}" if f in WORDS else ""
        ratio = f"{hbm / (level_words(f, 1) * 8 * n1):.2f}" if f in WORDS else ""
        print(f"| {f} | {e['calls']} | {e['read_bytes'] / 1e6:.1f} | {e['write_bytes'] / 1e6:.1f} | "
              f"{hbm / 1e6:.1f} | {alg} | {ratio} |")
    print("\n| kernel @ family (fine level) | dispatches | read MB | write MB |")
    print("|---|---:|---:|---:|")
    for k, v in sorted(out["kernels"].items(), key=lambda kv: -(kv[1]["read_bytes"] + kv[1]["write_bytes"])):
        print(f"| `{k}` | {v['dispatches']} | {v['read_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} |")


def trace_summary(trace_dir):
    """Per-kernel and per-family device times from a rocprofv3
    --kernel-trace run (its results database) of the headline workload:
    dispatches, mean and total duration per kernel and grid size, and the
    mean time per call of every fine-level kernel family (the figure the
    bench's roofline divides by).  Prints markdown.  Runs on the CPU."""
    import glob
    import sqlite3

    def short(name):
        return name.replace("desalm::(anonymous namespace)::", "").replace("void ", "").split("(")[0].strip()

    dbs = glob.glob(os.path.join(trace_dir, "**", "*results.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no rocprofv3 results database under {trace_dir}")
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select name, grid_x * grid_y * grid_z, start, duration from kernels order by start").fetchall()
    kern = {}
    for nm, g, _, d in rows:
        e = kern.setdefault((short(nm), g), [0, 0.0])
        e[0] += 1
        e[1] += d
    total = sum(e[1] for e in kern.values())
    fam = {}
    seq = [(short(nm), g) for nm, g, _, _ in rows]
    for (nm, _), (_, _, _, d), f in zip(seq, rows, classify_dispatches(seq)):
        if f is None:
            continue
        e = fam.setdefault(f, [0, 0.0])
        if nm.startswith(FAMILY_MARK[f]):
            e[0] += 1
        e[1] += d
    print(f"{len(rows)} dispatches, {total / 1e6:.3f} ms of kernel time\n")
    print("| fine-level family | calls | mean ms per call | total ms |")
    print("|---|---:|---:|---:|")
    for f, (c, d) in sorted(fam.items(), key=lambda kv: -kv[1][1]):
        print(f"| {f} | {c} | {d / max(c, 1) / 1e6:.4f} | {d / 1e6:.3f} |")
    print("\n| kernel | grid (work-items) | dispatches | mean us | total ms | share |")
    print("|---|---:|---:|---:|---:|---:|")
    for (nm, g), (c, d) in sorted(kern.items(), key=lambda kv: -kv[1][1])[:40]:
        print(f"| `{nm}` | {g} | {c} | {d / c / 1e3:.2f} | {d / 1e6:.3f} | {100.0 * d / total:.1f}% |")


# the overset workload: six curvilinear blocks of uneven sizes over a
# sinusoidal hill wall (amplitude 0.2), SA at Re 10595
OVERSET_BLOCKS = [(97, 65, 49), (113, 81, 57), (129, 65, 65), (145, 97, 73), (161, 97, 81), (81, 49, 41)]


def overset_run(args):
    """The six-block overset group, FP64, graph path: Mcell-iters/s with one
    launch per kernel over all blocks and zone by zone, from the same state."""
    import numpy as np
    from desalm import overset
    from desalm.lib import Group, Solver

    cs, tab = overset.hill_blocks(OVERSET_BLOCKS, [0.15, 0.32, 0.5, 0.68, 0.86], itm=args.iters, amp=0.2,
                                  ren=10595.0, ng=3)
    n = sum(a * b * c for a, b, c in OVERSET_BLOCKS)
    out = {"workload": "overset: 6 blocks " + ", ".join(f"{a}x{b}x{c}" for a, b, c in OVERSET_BLOCKS) +
                       f" ({n} nodes) over a sinusoidal hill wall, SA Re 10595, 3-level V-cycle, "
                       f"{args.iters} pseudo-iterations per call", "unit": "Mcell-iters/s"}
    for key, batched in (("value_one_launch", True), ("value_zone_by_zone", False)):
        svs = [Solver(c.params, device=0).load_case(c) for c in cs]
        g = Group(svs)
        g.set_batched(batched)
        g.set_overset(tab)
        g.mg_driver()
        done = 0
        ta = time.perf_counter()
        for _ in range(args.steps):
            itc, _ = g.mg_driver()
            done += itc
        tb = time.perf_counter()
        out[key] = round(n * done / (tb - ta) / 1e6, 3)
        out["state_finite"] = out.get("state_finite", True) and all(
            bool(np.all(np.isfinite(sv.get_state()[0]))) for sv in svs)
        g.close()
        for sv in svs:
            sv.close()
    return out


def dual_time_leg(args):
    """Physical time steps of SA-DES dual time stepping on the actuator-line
    box: the plain FP64 method (every step starts from q^n, as the
    reference) against extrapolation and the learned warm start, FP64 and
    mixed precision.  For each of a ladder of tolerances -- d decades below
    the plain start's first update norm (max over u, v, w of log10 mean
    |dq|, mg_ck_conver's absolute test eo_min), d in --dt-drops -- every form
    iterates from the same saved state to that tolerance, at most
    `conv_itm` pseudo-iterations, and its wall time per physical step is
    recorded.  The 1e-8 drop of the workload is the ladder's 8-decade rung."""
    import numpy as np
    from desalm import cases
    from desalm.lib import Solver
    from desalm.params import PREC_MIXED

    ni, nj, nk = (int(v) for v in args.dt_grid.split(","))
    cap = args.conv_itm
    drops = [float(d) for d in args.dt_drops.split(",")]
    t_build = time.perf_counter()
    case = cases.alm_box(ni, nj, nk, ng=3, itm=cap, it_min=1, iters=(1, 2, 3))
    case_mixed = cases.alm_box(ni, nj, nk, ng=3, itm=cap, it_min=1, iters=(1, 2, 3), precision=PREC_MIXED)
    t_build = time.perf_counter() - t_build
    sp = Solver(case.params, device=0).load_case(case)
    sm = Solver(case_mixed.params, device=0).load_case(case_mixed)
    for sv, cc in ((sp, case), (sm, case_mixed)):
        if args.autotune:
            sv.autotune(reps=3)
            sv.set_state(*cc.state())
        sv.ws_configure(hidden=32, samples_per_record=DT_SAMPLES, capacity=DT_CAPACITY, seed=20260505)
    # warm-up history (3 plain steps of 30 pseudo-iterations), then the
    # training steps, each recorded after it converged to the first rung
    sp.set_convergence(30, 30, -99.0, -99.0)
    for _ in range(3):
        sp.step()
    itc0, h0 = sp.step()
    r0 = float(h0[0, 1:4].max())
    tol_train = r0 - args.dt_drop
    sp.set_convergence(cap, 1, -99.0, tol_train)
    t_gen = time.perf_counter()
    gen_itc = []
    for _ in range(args.dt_train_steps):
        itc, _ = sp.step()
        sp.ws_record()
        gen_itc.append(itc)
    t_gen = time.perf_counter() - t_gen
    log(f"dual-time leg: {args.dt_train_steps} training steps generated")
    t_train = time.perf_counter()
    loss = sp.ws_train(args.dt_train_iters, batch=16384, lr=2e-3)
    t_train = time.perf_counter() - t_train
    sm.ws_params(sp.ws_params())
    sm.ws_stats(sp.ws_stats())

    forms = [("fp64_plain", sp, Solver.WS_NONE), ("fp64_extrap", sp, Solver.WS_EXTRAP),
             ("fp64_learned", sp, Solver.WS_LEARNED), ("mixed_learned", sm, Solver.WS_LEARNED)]
    res = {d: {f: {"itc": [], "wall_s": [], "first": [], "reached": [], "last": []} for f, _, _ in forms}
           for d in drops}
    fallbacks = {f: 0 for f, _, _ in forms}
    tail = []
    for _ in range(args.dt_eval_steps):
        snap = sp.snapshot()
        after = None
        for d in drops:
            tol = r0 - d
            for name, sv, mode in forms:
                sv.restore(snap)
                sv.ws_mode(mode)
                sv.set_convergence(cap, 1, -99.0, tol)
                fb0 = sv.last_status()["fallbacks"]
                t0 = time.perf_counter()
                itc, h = sv.step()
                wall = time.perf_counter() - t0
                fallbacks[name] += sv.last_status()["fallbacks"] - fb0
                dd = h[:itc, 1:4].max(axis=1)
                r = res[d][name]
                r["itc"].append(int(itc))
                r["wall_s"].append(wall)
                r["first"].append(float(dd[0]))
                r["last"].append(float(dd[-1]))
                r["reached"].append(bool(dd[-1] < tol))
                if name == "fp64_plain" and d == drops[0]:
                    after = sp.snapshot()
                if name == "fp64_plain" and d == drops[-1] and itc >= 20:
                    half = itc // 2
                    tail.append(float((dd[-1] - dd[half]) / (itc - 1 - half)))
        log(f"dual-time leg: evaluation step {len(res[drops[0]]['fp64_plain']['itc'])} done")
        # the trajectory continues from the plain step at the first rung
        sp.restore(after)
        sp.ws_mode(Solver.WS_NONE)

    def summary(r):
        return {"pseudo_iterations_mean": round(float(np.mean(r["itc"])), 2),
                "wall_s_per_step": round(float(np.mean(r["wall_s"])), 4),
                "first_log10_norm_mean": round(float(np.mean(r["first"])), 3),
                "last_log10_norm_mean": round(float(np.mean(r["last"])), 3),
                "all_reached": all(r["reached"]), "pseudo_iterations": r["itc"]}

    ladder = []
    for d in drops:
        fs = {f: summary(r) for f, r in res[d].items()}
        wp, wm = fs["fp64_plain"]["wall_s_per_step"], fs["mixed_learned"]["wall_s_per_step"]
        ladder.append({"decades_below_plain_start": d, "log10_norm": round(r0 - d, 3), "forms": fs,
                       "speedup_mixed_learned_vs_fp64_plain": round(wp / wm, 3) if wm > 0 else None,
                       "speedup_valid": fs["fp64_plain"]["all_reached"] and fs["mixed_learned"]["all_reached"]})
    reached_all = [e["decades_below_plain_start"] for e in ladder
                   if all(f["all_reached"] for f in e["forms"].values())]
    rate = float(np.mean(tail)) if tail else None
    rung8 = next((e for e in ladder if e["decades_below_plain_start"] >= 8.0), None)
    out = {"workload": f"actuator-line box {ni}x{nj}x{nk}, SA-DES, dt {case.params.delti}, 3 levels, "
                       f"rotor 3 blades R 0.5 TSR 6",
           "plain_start_log10_norm": round(r0, 3), "cap": cap,
           "test": "max over u, v, w of log10 mean |dq| < tolerance (eo_min of mg_ck_conver)",
           "ladder": ladder,
           "deepest_drop_all_forms_reached": max(reached_all) if reached_all else None,
           "fallbacks_to_plain": fallbacks,
           "target_1e-8_drop": {
               "reached_by": None if rung8 is None else
               [f for f, v in rung8["forms"].items() if v["all_reached"]],
               "plain_tail_rate_decades_per_iter": None if rate is None else round(rate, 5),
               "plain_last_log10_norm_at_cap": None if rung8 is None else
               rung8["forms"]["fp64_plain"]["last_log10_norm_mean"]}}
    out["training"] = {"steps": args.dt_train_steps, "mean_pseudo_iterations": round(float(np.mean(gen_itc)), 2),
                       "samples": int(min(args.dt_train_steps * DT_SAMPLES, DT_CAPACITY)),
                       "stencil": "node, 6 face neighbours, 2nd neighbours along i; last 3 steps",
                       "adam_steps": args.dt_train_iters,
                       "loss_first": round(float(loss[:20].mean()), 5), "loss_last": round(float(loss[-20:].mean()), 5),
                       "wall_s": {"case_build": round(t_build, 2), "generate": round(t_gen, 2),
                                  "train": round(t_train, 3)}}
    sp.close()
    sm.close()
    return out


def large_leg(args):
    """The large-grid DES workload: 257x193x193 actuator-line box, 4-level
    V-cycle, random-flow-generator inflow (5% intensity, integral length
    0.1), FP32-DAF mixed precision with the learned warm start (trained on
    the GPU on plain steps of this run), autotuned tiles, one HIP graph per
    pseudo-iteration.  Reports Mcell-iters/s, the device memory the context
    holds, and per kernel family on the fine level its algorithmic GB/s
    (bench.py's words per node over the profiled launch time)."""
    from desalm import cases, rfg
    from desalm.lib import Solver
    from desalm.params import PREC_MIXED

    ni, nj, nk = 257, 193, 193
    case = cases.alm_box(ni, nj, nk, ng=4, itm=args.iters, iters=(1, 2, 3, 3), precision=PREC_MIXED)
    sv = Solver(case.params, device=0).load_case(case)
    tiles = sv.autotune(reps=3)
    sv.set_state(*case.state())
    sv.set_rfg(rfg.modes(1000, 20260606), rfg.inflow_tables(nj, nk, intensity=0.05, length=0.1))
    sv.ws_configure(hidden=32, samples_per_record=65536, capacity=1 << 21, seed=20260707)
    # history, then recorded plain steps of 20 pseudo-iterations, training
    sv.set_convergence(20, 20, -99.0, -99.0)
    for _ in range(3):
        sv.step()
    for _ in range(args.large_train_steps):
        sv.step()
        sv.ws_record()
    loss = sv.ws_train(400, batch=16384, lr=2e-3)
    log("large leg: warm start trained")
    sv.ws_mode(Solver.WS_LEARNED)
    sv.set_convergence(args.iters, args.iters, -99.0, -99.0)
    sv.step()
    done = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itc, _ = sv.step()
        done += itc
    t1 = time.perf_counter()
    sv.profile(True)
    sv.step()
    st = sv.profile_stats()
    sv.profile(False)
    n1 = ni * nj * nk
    per_family = {}
    for (nm, lv), (cnt, ms) in st.items():
        if lv == 1 and nm in WORDS:
            b = level_words(nm, 1, case.params.itk) * 8 * n1
            per_family[nm] = {"mean_launch_ms": round(ms / cnt, 4),
                              "algorithmic_GBps": round(b / (ms / cnt * 1e-3) / 1e9, 1)}
    out = {"workload": f"actuator-line box {ni}x{nj}x{nk}, SA-DES, 4 levels, RFG inflow, FP32-DAF mixed, "
                       f"learned warm start",
           "value": round(n1 * done / (t1 - t0) / 1e6, 3), "unit": "Mcell-iters/s",
           "ms_per_step": round((t1 - t0) / args.steps * 1e3, 3), "pseudo_iterations": done,
           "device_bytes": sv.device_bytes, "fine_nodes": n1, "nodes_all_levels": case.params.nodes_all(),
           "warm_start": {"train_steps": args.large_train_steps, "adam_steps": 400,
                          "loss_first": round(float(loss[:20].mean()), 5),
                          "loss_last": round(float(loss[-20:].mean()), 5),
                          "fallbacks_to_plain": sv.last_status()["fallbacks"]},
           "tiles": {"visc_by_level": tiles[0], "sa_rhs": tiles[1]}, "fine_families": per_family}
    sv.close()
    return out


def cpu_baseline(args, dims):
    """The reference's own mg_driver (oracle/_ref, its Fortran built FP64 from
    its sources, one rank, timed inside the program) and the C restatement
    (oracle, one thread) on a bounded sample of the headline workload.  The
    reference is the baseline when it is present; a missing build is said
    on stderr and in the record."""
    import pyoracle as po
    from desalm import cases

    ni, nj, nk = dims
    n1 = ni * nj * nk
    cc = cases.alm_box(ni, nj, nk, ng=3, itm=args.cpu_iters, it_min=args.cpu_iters, iters=(1, 2, 3))
    log("CPU baseline: C restatement")
    o = po.Oracle(cc.params).load_case(cc)
    tc0 = time.perf_counter()
    itc_c, _ = o.mg_driver()
    tc1 = time.perf_counter()
    del o
    port = {"value": round(n1 * itc_c / (tc1 - tc0) / 1e6, 4), "unit": "Mcell-iters/s", "cores": 1, "kind": "port",
            "sample": f"{itc_c} pseudo-iterations of the same {workload_name(dims)} V-cycle, C restatement, "
                      f"1 thread ({tc1 - tc0:.1f} s)"}
    why = None
    if args.ref_iters <= 0:
        why = "--ref-iters 0"
    elif not po.ref_available():
        why = "oracle/_ref/desalm_ref_harness is not built (oracle/ref/build_ref.sh needs the reference sources)"
    else:
        cr = cases.alm_box(ni, nj, nk, ng=3, itm=args.ref_iters, it_min=args.ref_iters, iters=(1, 2, 3))
        log("CPU baseline: the reference's Fortran")
        try:
            r = po.run_reference(cr, mode=1, nsteps=1, timeout=900)
        except Exception as e:
            r = {"seconds": None}
            why = "the reference run failed: " + str(e)[:200]
        if r.get("seconds"):
            return {"value": round(n1 * args.ref_iters / r["seconds"] / 1e6, 4), "unit": "Mcell-iters/s",
                    "cores": 1, "kind": "reference",
                    "sample": f"{args.ref_iters} pseudo-iterations of the same {workload_name(dims)} V-cycle, the "
                              f"reference's mg_driver (its Fortran sources built FP64 with amdflang -O2), one rank "
                              f"({r['seconds']:.1f} s inside the program)",
                    "port": port, "reference_over_port_time": round(
                        (r["seconds"] / args.ref_iters) / ((tc1 - tc0) / itc_c), 3)}
        why = why or "the reference run reported no time"
    print(f"bench.py: CPU baseline is the C restatement, not the reference: {why}", file=sys.stderr)
    port["why_not_reference"] = why
    return port


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=10, help="pseudo-iterations per physical step")
    ap.add_argument("--channel-n", type=int, default=129, help="nodes per direction of the channel extra")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-large", dest="large", action="store_false",
                    help="skip the 257x193x193 large-grid DES workload")
    ap.add_argument("--large-train-steps", type=int, default=12,
                    help="plain steps recorded for the large leg's warm start")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes of the LDS-tiled kernels")
    ap.add_argument("--tiles", default=None, metavar="V1,V2,V3:S",
                    help="fix the launch shapes instead of autotuning: k planes per workgroup of k_visc_pk "
                         "per level and of k_sa_rhs (e.g. 8,4,1:8, as a previous run's 'tiles' reported)")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--ref-iters", type=int, default=2,
                    help="pseudo-iterations of the reference Fortran timed for the CPU baseline (0: port only)")
    ap.add_argument("--no-extras", action="store_true", help="skip the mixed-precision and channel extras")
    ap.add_argument("--conv-itm", type=int, default=300,
                    help="pseudo-iteration cap per physical step of the dual-time leg (0 skips it)")
    ap.add_argument("--dt-grid", default="193,129,129", help="dual-time leg grid")
    ap.add_argument("--dt-drops", default="2,3,4,8",
                    help="tolerance ladder of the dual-time leg: decades below the plain start's first norm")
    ap.add_argument("--dt-train-steps", type=int, default=200,
                    help="plain physical steps recorded as training samples for the warm start")
    ap.add_argument("--dt-eval-steps", type=int, default=3, help="physical steps compared after training")
    ap.add_argument("--dt-drop", type=float, default=2.0,
                    help="decades below the plain start's first norm the training steps converge to")
    ap.add_argument("--dt-train-iters", type=int, default=1500, help="Adam steps of the warm-start network")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no run: summarise two rocprofv3 counter runs of the headline workload into "
                         "profiles/traffic.json")
    ap.add_argument("--trace-summary", metavar="TRACE_DIR",
                    help="no run: summarise a rocprofv3 --kernel-trace run of the headline workload")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], HEADLINE_DIMS)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    from desalm import cases
    from desalm.lib import Solver

    from desalm.params import PREC_FP64, PREC_MIXED

    fixed_tiles = None
    if args.tiles:
        v, _, sa = args.tiles.partition(":")
        fixed_tiles = ([int(t) for t in v.split(",") if t], int(sa or 0))
        args.autotune = False

    def timed_run(make_case, precision, profile, physical, tiles_from=None):
        """warmup + timed physical steps of --iters pseudo-iterations each;
        returns a dict (Mcell-iters/s, wall s, iterations, profile stats, last
        norm history, state finite, pivot events, params, tiles).  tiles_from:
        the launch shapes of an earlier run of the same workload, reused."""
        cs = make_case(precision)
        sv = Solver(cs.params, device=0).load_case(cs)
        if tiles_from is not None:
            sv.set_tiles(*tiles_from)
            tiles = sv.get_tiles()
        elif args.autotune:
            # launch shapes of the LDS-tiled kernels, timed on this state
            # (scratch only), then the state is reloaded
            tiles = sv.autotune(reps=3)
            sv.set_state(*cs.state())
        else:
            if fixed_tiles:
                sv.set_tiles(*fixed_tiles)
            tiles = sv.get_tiles()
        run = sv.step if physical else sv.mg_driver
        for _ in range(args.warmup):
            run()
        if profile:
            sv.profile(True)
        done = 0
        ta = time.perf_counter()
        for _ in range(args.steps):
            itc, hist = run()
            done += itc
        tb = time.perf_counter()
        st = sv.profile_stats() if profile else None
        if profile:
            sv.profile(False)
        r = {"value": cs.n1 * done / (tb - ta) / 1e6, "wall": tb - ta, "done": done, "stats": st, "hist": hist,
             "finite": bool(np.all(np.isfinite(sv.get_state()[0]))), "pivots": sv.pivot_events,
             "params": cs.params, "tiles": tiles}
        sv.close()
        return r

    ni, nj, nk = HEADLINE_DIMS

    def alm(precision):
        return cases.alm_box(ni, nj, nk, ng=3, itm=args.iters, it_min=args.iters, iters=(1, 2, 3),
                             precision=precision)

    def channel(precision):
        n = args.channel_n
        return cases.channel(n, n, n, ng=3, itm=args.iters, it_min=args.iters, iters=(1, 2, 3), precision=precision)

    # production path: every pseudo-iteration is one HIP graph replay
    log("headline: actuator-line box, FP64 plain, graph replays")
    head = timed_run(alm, PREC_FP64, False, True)
    log(f"headline: {head['value']:.1f} Mcell-iters/s")
    p = head["params"]
    n1 = p.img * p.jmg * p.kmg
    # the same workload with the library's per-launch HIP-event profiler
    # (launches then run eagerly): per-kernel-family device times
    prof = timed_run(alm, PREC_FP64, True, True, tiles_from=head["tiles"])
    log("profiled pass done")
    stats = prof["stats"]
    extras = None
    if not args.no_extras:
        mx = timed_run(alm, PREC_MIXED, False, True)
        log(f"mixed: {mx['value']:.1f} Mcell-iters/s")
        ch = timed_run(channel, PREC_FP64, False, False)
        chm = timed_run(channel, PREC_MIXED, False, False)
        extras = {"mixed": {"value": round(mx["value"], 3), "unit": "Mcell-iters/s",
                            "ms_per_step": round(mx["wall"] / args.steps * 1e3, 3), "state_finite": mx["finite"],
                            "dtype": "f64 residual, f32 DAF sweeps"},
                  "channel": {"workload": f"channel {args.channel_n}^3 SA at Re_tau 395, 3-level V-cycle, "
                                          f"streamwise/spanwise periodic (wrap) with the mean pressure gradient "
                                          f"as a body force, {args.iters} pseudo-iterations per step",
                              "value_fp64": round(ch["value"], 3), "value_mixed": round(chm["value"], 3),
                              "unit": "Mcell-iters/s", "ms_per_step_fp64": round(ch["wall"] / args.steps * 1e3, 3),
                              "state_finite": ch["finite"] and chm["finite"]},
                  "overset": overset_run(args)}
        log(f"overset: {extras['overset']['value_one_launch']:.1f} Mcell-iters/s in one launch per kernel, "
            f"{extras['overset']['value_zone_by_zone']:.1f} zone by zone")

    # dominant kernel family by total device time
    fam = {}
    for (nm, n), (cnt, ms) in stats.items():
        fam.setdefault(nm, [0, 0.0])
        fam[nm][0] += cnt
        fam[nm][1] += ms
    dominant = max((k for k in fam if k in WORDS and (k, 1) in stats), key=lambda k: fam[k][1])
    cnt1, ms1 = stats[(dominant, 1)]
    bytes_per_launch = level_words(dominant, 1, p.itk) * 8 * n1
    achieved = bytes_per_launch / (ms1 / cnt1 * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE,
    # WRITE_SIZE, scaled by a plain copy of known size), if recorded for
    # this family and workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        ent = tr.get("families", {}).get(f"{dominant}@1")
        if ent and tr.get("workload") == workload_name(HEADLINE_DIMS):
            traffic = ent["hbm_bytes_per_launch"]
    roofline = {"bound": "hbm", "kernel": f"{dominant} (fine level)", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "bytes_per_launch": bytes_per_launch,
                "mean_launch_ms": round(ms1 / cnt1, 4),
                "timing": "HIP events around each launch of the family on the library stream, profiled pass "
                          f"of the same workload ({round(prof['value'], 3)} Mcell-iters/s with eager launches)"}
    # every fine-level family beside it (algorithmic GB/s and roofline fraction)
    families = {}
    for (nm, lv), (cnt, ms) in sorted(stats.items()):
        if lv == 1 and nm in WORDS:
            b = level_words(nm, 1, p.itk) * 8 * n1
            gbs = b / (ms / cnt * 1e-3) / 1e9
            families[nm] = {"mean_launch_ms": round(ms / cnt, 4), "algorithmic_GBps": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 3)}

    conv = dual_time_leg(args) if args.conv_itm > 0 else None
    large = large_leg(args) if args.large else None
    if large:
        log(f"large leg: {large['value']:.1f} Mcell-iters/s")
    cpu = None if args.no_cpu else cpu_baseline(args, HEADLINE_DIMS)
    if cpu:
        log(f"CPU baseline ({cpu['kind']}): {cpu['value']} Mcell-iters/s")

    out = {"metric": METRIC, "value": round(head["value"], 3), "unit": "Mcell-iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["wall"] / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{workload_name(HEADLINE_DIMS)} SA-DES, actuator line 3 blades R 0.5 TSR 6, "
                                  f"dt 0.01, 3-level FAS V-cycle (iter 1,2,3), FP64 plain, {args.iters} "
                                  f"pseudo-iterations per physical step",
                      "fine_nodes": n1, "pseudo_iterations": head["done"], "parallelism": "single GPU"},
           "roofline": roofline, "fine_families": families, "cpu_baseline": cpu, "extras": extras,
           "converge": conv, "large": large,
           "kernels_ms": {f"{nm}@L{n}": [cnt, round(ms, 3)] for (nm, n), (cnt, ms) in sorted(stats.items())},
           "tiles": {"k_planes_per_workgroup_visc_by_level": head["tiles"][0],
                     "k_planes_per_workgroup_sa_rhs": head["tiles"][1],
                     "chosen_by": "desalm_autotune" if args.autotune else ("--tiles" if fixed_tiles else "default")},
           "state_finite": head["finite"], "pivot_events": head["pivots"],
           "final_log10_update_norms": [round(float(v), 4) for v in head["hist"][-1]]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
