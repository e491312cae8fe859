"""LP solves/sec on one MI355X: Mehrotra IPM to a 1e-8 relative duality gap.

Workload (default): the block-angular LP -- 64 diagonal blocks of 1,500 x 4,000 (8 nonzeros per
column, N(0,1)) plus 50 linking rows at 5% density (U(-1,1)); m = 96,050, n = 256,000,
nnz = 2.69M; b and c from a strictly feasible primal/dual pair (seed 0).  A step is one complete
solve (x = s = 1, y = 0 start, to rel. gap, primal and dual residuals <= 1e-8) with the solver
set up once (Ruiz scaling, AMD ordering, supernodal plan and data resident in HBM before the
timed region).  Every timed solve is checked: status Optimal, the 1e-8 criteria, the FP64
certificate, and the objective against the HiGHS golden value when one is recorded.

roofline: the dominant kernel, k_step_dag (one launch = every 64-column block-column STEP tile
and its split-off partial products for one etree level of the supernodal FP64 factorization,
run in dependency order; MFMA-bound).  Its algorithmic flops per factorization are
F_chol = sum_j c_j^2 (exact column counts of L) minus the panel flops of the other launch kinds
(first diagonal blocks, ancestor updates); achieved = those flops x factorizations timed / the
summed device time of its launches, measured live over the timed solves with HIP events on the
solver's stream at every change of launch kind.  traffic: HBM bytes per k_step_dag launch from
the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/), scaled by the same counters
on a copy kernel of known size.
cpu_baseline: the reference's own CPU route -- solve_with_phase1 (Phase I + Phase II primal
simplex over the sparse LU basis factorization), compiled from the reference sources into
oracle/_ref/libref.so -- timed on one pinned core on the largest member of the 2,000 x 5,000
random-LP family it solves within the sample budget (m = 500, n = 1,250); on the workload itself
its Phase I needs >= m = 96,050 pivots and stops at the reference's 10,000-pivot limit.  The
C oracle IPM (same algorithm as the device, dense FP64 Cholesky, 1 thread) on one block of the
workload, scaled to 64 blocks x the GPU's iteration count, is reported beside it ("port_ipm").

chosen_forms: each chosen form against plain FP64 on the same instances, in the same run --
mixed precision (FP32 MFMA factor + FP64 PCG) on this workload and on the 700 x 700 transport LP
with 20,000 side rows; the batched persistent solver on the 4,096-LP batch (m ~ U{50..800}),
FP64 and mixed, against a sequential loop of the plain single-LP device IPM on a sample of the
same LPs; the learned warm start on the parametric family at 400 x 1,000 (the 20,000 x 50,000
family scaled down so its training set can be generated in the run); and the reference's simplex
route itself run on the device (lk_simplex) on the LP the CPU baseline times, beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "soverign-gpu-accelerator-_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

METRIC = "LP solves/sec & IPM iters to 1e-8 gap: chosen form vs plain FP64, 1 MI355X"
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
CPU_ITERS = 8
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_k_step_block_angular_64x1500x4000_fp64.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def golden(name):
    p = os.path.join(ROOT, "tests", "golden", "objectives.json")
    if os.path.exists(p):
        with open(p) as f:
            g = json.load(f).get(name)
        if g and g.get("objective") is not None:
            return g
    return None


def build_workload(args):
    from lakshya import gen
    if args.workload == "block_angular":
        lp = gen.block_angular(args.blocks, 1500, 4000, 8, 50, 0.05, seed=args.seed)
        name = f"ba_{args.blocks}x1500x4000_L50_s{args.seed}"
    elif args.workload == "random_sparse":
        lp = gen.random_sparse(2000, 5000, 6, seed=args.seed)
        name = f"rs_2000_5000_6_s{args.seed}"
    elif args.workload == "parametric":     # the base LP of the parametric family, 20,000 x 50,000
        lp = gen.parametric_base(seed=args.seed)
        name = f"rs_20000_50000_8_s{args.seed}"
    else:
        lp = gen.transport(seed=args.seed)
        name = f"tr_700x700_side20000_s{args.seed}"
    return lp, name


def check_solve(native, lp, r, gold):
    ok = (r["status_name"] == "Optimal" and r["rel_gap"] <= 1e-8 and
          r["rel_primal_residual"] <= 1e-8 and r["rel_dual_residual"] <= 1e-8)
    cert8, _ = native.verify_optimal(lp, r["x"], r["y"], tol=1e-8)
    cert6, _ = native.verify_optimal(lp, r["x"], r["y"], tol=1e-6)
    obj_err = None
    if gold is not None:
        obj_err = abs(r["objective"] - gold) / max(1.0, abs(gold))
        ok = ok and obj_err <= 1e-8
    return ok, cert8, cert6, obj_err


def cpu_baseline(gpu_iters, blocks, seed):
    """Oracle IPM on one block of the workload, 2 iterations, scaled (see module docstring)."""
    import pyoracle
    from lakshya import gen
    lp = gen.block_angular(1, 1500, 4000, 8, 50, 0.05, seed=seed)
    csr = lp.csr()
    t = time.perf_counter()
    pyoracle.ipm_solve(csr, lp.b, lp.c, lp.n, stop_rule=1, max_iter=CPU_ITERS)
    dt = time.perf_counter() - t
    per_iter = dt / CPU_ITERS
    solve_s = per_iter * blocks * max(gpu_iters, 1)
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"C oracle (same IPM, dense FP64 Cholesky, 1 thread) on 1 of {blocks} "
                       f"blocks (m=1550, n=4000) for {CPU_ITERS} iterations = {dt:.1f} s; scaled x{blocks} "
                       f"blocks x {gpu_iters} iterations = {solve_s:.0f} s per solve")}


def _ref_route_child(q, cpu, m, n, seed):
    """Child process: the reference CPU route on one random LP, pinned to one core."""
    import ctypes as C
    try:
        os.sched_setaffinity(0, {cpu})
    except (AttributeError, OSError):
        pass
    from lakshya import gen
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.ref_phase1_solve.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), dp, C.POINTER(C.c_double)]
    lp = gen.random_sparse(m, n, 6, seed=seed)
    st, obj, ms = C.c_int(), C.c_double(), C.c_double()
    x = np.zeros(lp.n)
    rc = L.ref_phase1_solve(lp.m, lp.n, lp.indptr.astype(np.int32), lp.indices.astype(np.int32),
                            lp.data.astype(np.float64), lp.c.astype(np.float64),
                            lp.b.astype(np.float64), C.byref(st), C.byref(obj), x, C.byref(ms))
    q.put({"rc": rc, "status": st.value, "objective": obj.value, "ms": ms.value})


def cpu_reference_route(m=500, n=1250, seed=1, timeout_s=120):
    """The reference's CPU simplex route on one pinned core (see the module docstring)."""
    import multiprocessing as mp
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so")):
        return None
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    cpu = sorted(os.sched_getaffinity(0))[-1] if hasattr(os, "sched_getaffinity") else 0
    p = ctx.Process(target=_ref_route_child, args=(q, cpu, m, n, seed))
    p.start()
    p.join(timeout_s)
    if p.is_alive():
        p.kill()
        p.join()
        return {"value": 0.0, "unit": "solves/s", "cores": 1, "kind": "reference",
                "sample": f"reference solve_with_phase1 + SparseLU on rs_{m}_{n}_6_s{seed}: "
                          f"no result within {timeout_s} s"}
    r = q.get()
    names = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "IterationLimit"}
    solved = r["rc"] == 0 and r["status"] == 0
    return {"value": (1e3 / r["ms"]) if solved else 0.0, "unit": "solves/s", "cores": 1,
            "kind": "reference", "route_seconds": r["ms"] / 1e3, "route_objective": r["objective"],
            "route_status": names.get(r["status"], r["status"]),
            "sample": (f"reference CPU route (solve_with_phase1: Phase I + Phase II primal simplex, "
                       f"SparseLUFactorization; oracle/_ref built from the reference sources, -O2, "
                       f"no fast-math) on one pinned core, random LP m={m} n={n} (6 nnz/col, seed "
                       f"{seed}): {names.get(r['status'], r['status'])} in {r['ms'] / 1e3:.1f} s, "
                       f"objective {r['objective']:.10g}.  The workload (m=96,050) needs >= 96,050 "
                       f"Phase-I pivots against the route's 10,000-pivot limit: IterationLimit, "
                       f"no solve")}


def step_roofline(live, st, fp32, headline):
    """Roofline of the factorization's block-column launches: k_step_dag (STEPS), and where a
    level is factored with one launch per block column, also its k_part / k_psum / k_step
    launches (their products are most of that level's flops, so they are priced together).
    The HBM traffic figure is the profiled k_step_dag of the block-angular headline LP only."""
    peak = PEAK_FP32 if fp32 else PEAK_FP64
    nfact = max(live["factorization"]["launches"], 1)
    step, part = live["step"], live["part"]
    ms = step["ms"] + part["ms"]
    launches = step["launches"] + part["launches"]
    # algorithmic flops of these launches: the exact Cholesky flops minus the (stored) work of
    # the DIAG and UPDATE launches, so their share is not overstated
    f_step = st["flops_exact"] - (live["diag"]["flops"] + live["update"]["flops"]) / nfact
    achieved = f_step * nfact / (ms * 1e-3)
    traffic = None
    if headline and part["launches"] == 0 and os.path.exists(TRAFFIC):
        with open(TRAFFIC) as f:
            tr = json.load(f)
        traffic = tr["bytes_per_launch"]
    kernel = "k_step_dag (all block-column steps of a level, dependency-driven: MFMA GEMM + solve + next diagonal)"
    if part["launches"]:
        kernel += ", with k_part / k_psum / k_step (one launch each per block column of a level one supernode dominates)"
    fac = live["factorization"]
    return {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel": kernel,
            "launches": launches, "avg_launch_ms": ms / max(launches, 1),
            "algorithmic_flops_per_launch": f_step * nfact / max(launches, 1),
            "stored_flops_per_launch": (step["flops"] + part["flops"]) / max(launches, 1),
            "factorizations": nfact,
            "part_launches": part["launches"], "part_ms": part["ms"],
            "factorization": {"ms_avg": fac["ms"] / nfact, "flops_exact": st["flops_exact"],
                              "achieved_tflops": st["flops_exact"] / (fac["ms"] / nfact * 1e-3) / 1e12}}


def copy_calibration(mib, reps):
    """Streaming copy kernel over `mib` MiB, `reps` launches: the known-size reference for the HBM
    traffic counters (FETCH_SIZE / WRITE_SIZE of one launch read / write exactly mib MiB)."""
    import ctypes as C
    from lakshya import native
    L = native.lib()
    vp = C.c_void_p
    for name, args in (("lkd_copy_f64", [vp, vp, C.c_int64, vp]), ("lkd_event_create", [C.POINTER(vp)]),
                       ("lkd_event_record", [vp, vp]), ("lkd_event_elapsed_ms", [C.POINTER(C.c_float), vp, vp]),
                       ("lkd_stream_sync", [vp])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = C.c_int
    nbytes = mib << 20
    src = native.DeviceBuffer(nbytes)
    dst = native.DeviceBuffer(nbytes)
    n = nbytes // 8
    e0, e1 = vp(), vp()
    assert L.lkd_event_create(C.byref(e0)) == 0 and L.lkd_event_create(C.byref(e1)) == 0
    assert L.lkd_copy_f64(dst.ptr, src.ptr, n, None) == 0 and L.lkd_stream_sync(None) == 0
    assert L.lkd_event_record(e0, None) == 0
    for _ in range(reps):
        assert L.lkd_copy_f64(dst.ptr, src.ptr, n, None) == 0
    assert L.lkd_event_record(e1, None) == 0 and L.lkd_stream_sync(None) == 0
    ms = C.c_float()
    assert L.lkd_event_elapsed_ms(C.byref(ms), e0, e1) == 0
    per = ms.value / reps
    print(json.dumps({"copy_kernel": "k_copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "launches": reps + 1, "avg_ms": per, "GB_per_s": 2 * nbytes / (per * 1e-3) / 1e9}))


# ------------------------------------------------------------------ chosen forms
def _certified(r):
    return (r["status_name"] == "Optimal" and r["rel_gap"] <= 1e-8 and
            r["rel_primal_residual"] <= 1e-8 and r["rel_dual_residual"] <= 1e-8)


def _timed_solves(native, lp, gold, steps, **opts):
    """Set up one solver, one untimed solve, `steps` timed solves; returns a summary and the
    last result."""
    t = time.perf_counter()
    solver = native.IPM(lp, relative=True, **opts)
    setup = time.perf_counter() - t
    solver.run()
    rs = []
    t = time.perf_counter()
    for _ in range(steps):
        rs.append(solver.run())
    dt = time.perf_counter() - t
    solver.close()
    ok = all(_certified(r) for r in rs)
    errs = [abs(r["objective"] - gold) / max(1.0, abs(gold)) for r in rs] if gold is not None else None
    return {"solves_per_s": steps / dt, "ms_per_solve": dt * 1e3 / steps, "setup_s": setup,
            "iterations": [r["iterations"] for r in rs],
            "pcg_steps": [r["refinement_steps"] for r in rs],
            "certified": ok, "max_obj_err_vs_golden": max(errs) if errs else None}, rs[-1]


def _vs(a, b):
    """Agreement of two certified solutions of one LP: relative objective difference and the
    largest relative difference of x."""
    return {"rel_objective_diff": abs(a["objective"] - b["objective"]) / max(1.0, abs(b["objective"])),
            "max_rel_x_diff": float(np.max(np.abs(a["x"] - b["x"]) / (1.0 + np.abs(b["x"]))))}


def form_mixed(native, lp, gold, steps, fp64_summary, fp64_last):
    mx, last = _timed_solves(native, lp, gold, steps, precision=1)
    mx["vs_fp64"] = _vs(last, fp64_last)
    return {"fp64": fp64_summary, "mixed": mx,
            "mixed_over_fp64_solves_per_s": mx["solves_per_s"] / fp64_summary["solves_per_s"]}


def form_transport(native, steps):
    from lakshya import gen
    lp = gen.transport(seed=0)
    g = golden("tr_700x700_side20000_s0")
    gold = g["objective"] if g else None
    f64, last64 = _timed_solves(native, lp, gold, steps)
    out = form_mixed(native, lp, gold, steps, f64, last64)
    out["workload"] = {"name": "tr_700x700_side20000_s0", "m": lp.m, "n": lp.n, "nnz": lp.nnz}
    return out


def _batch_lp(args):
    from lakshya import gen
    return gen.batch_lp(*args)


def form_batch(native, count=4096, sample=48, workers=8):
    """The 4,096-LP batch: batched FP64 and mixed against a sequential loop of the plain
    single-LP device IPM (set-up + solve per LP, as a caller with distinct LPs pays it) on
    `sample` LPs spread over the batch, extrapolated to the batch."""
    import concurrent.futures as cf
    import multiprocessing as mp
    from lakshya import gen
    sizes = gen.batch_sizes(count)
    t = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        lps = list(ex.map(_batch_lp, [(int(sizes[i]), i) for i in range(count)], chunksize=64))
    gen_s = time.perf_counter() - t
    out = {"workload": {"count": count, "m_range": [int(sizes.min()), int(sizes.max())],
                        "generate_s": gen_s}}
    for name, prec in (("batched_fp64", 0), ("batched_mixed", 1)):
        bt = native.Batch(lps, precision=prec)
        bt.solve()                                     # first solve: graph capture, warm caches
        sm = bt.solve()
        res = bt.results()
        if name == "batched_fp64":
            t = time.perf_counter()
            xo = bt.crossover()                        # launch + host copies of every output
            xs = time.perf_counter() - t
            out["crossover_fp64"] = {"lps_per_s": count / xs, "seconds": xs,
                                     "optimal_vertices": sum(1 for o in xo if o["status"] == 0),
                                     "pivots_mean": float(np.mean([o["pivots"] for o in xo]))}
        bt.close()
        it = np.array([r["iterations"] for r in res])
        cert = sum(1 for r in res if r["status"] == 0 and r["rel_gap"] <= 1e-8 and
                   r["rel_primal"] <= 1e-8 and r["rel_dual"] <= 1e-8)
        out[name] = {"lps_per_s": count / (sm["ms"] / 1e3), "ms": sm["ms"], "certified": cert,
                     "factor_tflops": sm["flops_chol"] / (sm["ms"] * 1e-3) / 1e12,
                     "iterations_mean": float(it.mean()), "iterations_max": int(it.max()),
                     "wg_per_cu": sm["wg_per_cu"], "graph_replays": sm["graph_replays"]}
        if name == "batched_fp64":
            objs = np.array([r["objective"] for r in res])
    idx = np.linspace(0, count - 1, sample).astype(int)
    t = time.perf_counter()
    seq = [native.ipm_solve(lps[i], relative=True) for i in idx]
    dt = time.perf_counter() - t
    scert = sum(1 for r in seq if _certified(r))
    sobj = np.array([r["objective"] for r in seq])
    seq_rate = sample / dt
    out["sequential_plain_fp64"] = {"lps_per_s": seq_rate, "sample": int(sample),
                                    "certified": scert, "seconds": dt,
                                    "max_rel_obj_diff_vs_batched":
                                        float(np.max(np.abs(sobj - objs[idx]) / np.maximum(1.0, np.abs(objs[idx]))))}
    out["batched_fp64_over_sequential"] = out["batched_fp64"]["lps_per_s"] / seq_rate
    out["batched_mixed_over_sequential"] = out["batched_mixed"]["lps_per_s"] / seq_rate
    return out


def form_simplex(native, cpu, m=500, n=1250, seed=1):
    """The reference's CPU route (Phase I + Phase II simplex) run on the device (lk_simplex) on the
    LP the CPU baseline times: wall time of the call (upload, launch, read-back), status and
    objective, beside the reference build's time and objective on one pinned core."""
    from lakshya import gen
    lp = gen.random_sparse(m, n, 6, seed=seed)
    native.simplex([lp])                               # first call: module and arena set-up
    t = time.perf_counter()
    r = native.simplex([lp])[0]
    dt = time.perf_counter() - t
    out = {"workload": f"rs_{m}_{n}_6_s{seed}", "device_seconds": dt, "status": r["status_name"],
           "objective": r["objective"], "iterations": list(r["iterations"])}
    if cpu and cpu.get("kind") == "reference" and cpu.get("route_status") == "Optimal":
        out["reference_cpu_seconds"] = cpu["route_seconds"]
        out["device_over_reference_cpu"] = cpu["route_seconds"] / dt
        out["rel_objective_diff_vs_reference"] = abs(r["objective"] - cpu["route_objective"]) / \
            max(1.0, abs(cpu["route_objective"]))
    return out


def form_warm(ntrain=20_000, ntest=1000):
    from lakshya import warmstart
    return warmstart.experiment(400, 1000, 0.1, ntrain, ntest, target_iter=2, epochs=20)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="block_angular",
                    choices=["block_angular", "random_sparse", "transport", "parametric"])
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="all",
                    help="comma list of chosen-form comparisons: mixed,transport,batch,warm,simplex (all / none)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the streaming copy kernel over MIB MiB (traffic-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (--gpus {args.gpus} requested): this build runs on a "
              "single MI355X; multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    from lakshya import native
    if native.device_count() < 1:
        log("no GPU visible")
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(args.steps, 1))
        return
    lp, name = build_workload(args)
    g = golden(name)
    gold = g["objective"] if g else None
    log(f"workload {name}: m={lp.m} n={lp.n} nnz={lp.nnz} golden={gold}")

    t = time.perf_counter()
    solver = native.IPM(lp, relative=True, precision=1 if args.precision == "mixed" else 0)
    setup_ms = (time.perf_counter() - t) * 1e3
    st = solver.kkt_stats()
    log(f"setup {setup_ms:.0f} ms; supernodes={int(st['nsuper'])} nnz(L)={st['nnz_L']:.3e} "
        f"F_chol={st['flops_exact']:.3e} flop; levels={int(st['levels'])}")

    for w in range(args.warmup):
        r = solver.run()
        log(f"warmup {w}: {r['status_name']} iters={r['iterations']} obj={r['objective']:.12g} "
            f"gap={r['rel_gap']:.2e} {r['solve_ms']:.0f} ms")

    solver.set_launch_timing(True)
    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = solver.run()
        results.append(r)
    elapsed = time.perf_counter() - t0
    all_ok = True
    for k, r in enumerate(results):
        ok, c8, c6, oe = check_solve(native, lp, r, gold)
        all_ok &= ok
        log(f"step {k}: {r['status_name']} iters={r['iterations']} obj={r['objective']:.12g} "
            f"gap={r['rel_gap']:.2e} rp={r['rel_primal_residual']:.2e} "
            f"rd={r['rel_dual_residual']:.2e} cert1e-8={c8} cert1e-6={c6} obj_err={oe} "
            f"factor={r['factor_ms']:.0f}/{r['solve_ms']:.0f} ms fact={r['factorizations']} "
            f"pcg={r['refinement_steps']} fail={r.get('fail_code')} ok={ok}")
    iters = [r["iterations"] for r in results]
    value = args.steps / elapsed if all_ok else 0.0
    live = solver.launch_timing()
    solver.set_launch_timing(False)
    solver.close()
    log("launch timing " + json.dumps(live))
    roofline = step_roofline(live, st, args.precision == "mixed",
                             args.workload == "block_angular" and args.blocks == 64)

    forms = {}
    want = {"mixed", "transport", "batch", "warm", "simplex"} if args.forms == "all" else \
        set() if args.forms == "none" else set(args.forms.split(","))
    if "mixed" in want and args.precision == "fp64":
        log("chosen form: mixed precision on the workload")
        f64 = {"solves_per_s": value, "ms_per_solve": elapsed * 1e3 / args.steps,
               "setup_s": setup_ms / 1e3, "iterations": iters,
               "pcg_steps": [r["refinement_steps"] for r in results], "certified": bool(all_ok),
               "max_obj_err_vs_golden": None}
        forms["workload_mixed"] = form_mixed(native, lp, gold, args.steps, f64, results[-1])
        log(json.dumps(forms["workload_mixed"]))
    if "transport" in want:
        log("chosen form: mixed precision on the transport LP")
        forms["transport_mixed"] = form_transport(native, 1)
        log(json.dumps(forms["transport_mixed"]))
    if "batch" in want:
        log("chosen form: batched solver on the 4,096-LP batch")
        forms["batch_4096"] = form_batch(native)
        log(json.dumps(forms["batch_4096"]))
    if "warm" in want:
        log("chosen form: learned warm start (parametric family 400 x 1,000)")
        forms["warm_start"] = form_warm()
        log(json.dumps(forms["warm_start"]))

    cpu = None
    if not args.no_cpu:
        port = cpu_baseline(int(np.median(iters)), args.blocks, args.seed)
        cpu = cpu_reference_route() or port
        if cpu is not port:
            cpu["port_ipm"] = port
    if "simplex" in want:
        log("the reference's simplex route on the device, beside its CPU build")
        forms["simplex_route"] = form_simplex(native, cpu)
        log(json.dumps(forms["simplex_route"]))

    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if args.precision == "fp64" else "f32-factor+f64",
           "data": "synthetic (seeded generator, lakshya.gen)",
           "config": {"workload": name, "m": lp.m, "n": lp.n, "nnz": lp.nnz,
                      "form": "plain FP64 IPM" if args.precision == "fp64" else "mixed FP32/FP64",
                      "stop": "rel gap, primal, dual residual <= 1e-8",
                      "iterations": iters, "setup_ms": round(setup_ms, 1),
                      "all_checks_passed": bool(all_ok), "golden_objective": gold},
           "roofline": roofline, "cpu_baseline": cpu, "chosen_forms": forms}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
