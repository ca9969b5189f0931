"""Benchmark: Monte-Carlo DC operating points of the 64x64 resistor-diode mesh.

One step = one batched DC solve of all instances to convergence (19-step
GMIN plan, solveDC semantics), with netlist parameters already resident in
HBM.  Prints ONE JSON line:

  value        converged solves/s (FP64 KCL <= 1e-9 A), whole job, 1 GPU,
               plain FP64 path (the reference's method and arithmetic)
  roofline     the kernel with the largest device time in the timed steps:
               algorithmic bytes per launch / its mean HIP-event launch time
  fp32_refine  the same solves on the FP32 path (FP32 Jacobian and LU, FP64
               Newton residual and update), with its error against the FP64
               results
  autotune     the launch class per front-order bucket: replayed from the
               committed table karam_amd/tuning/rdmesh64_b4096.json (written
               by `--retune`, which runs the autotuner on this batch), so that
               every run and every profile of it launches the same kernels
  seeded       learned-seeder starts on the FP64 and on the FP32 path; the
               FP32 one's speedup over plain FP64 is `target_speedup`
  cpu_baseline the oracle (FP64 C restatement, OpenMP over instances) timed
               on this host's cores over a bounded sample of the instances,
               whose voltages are also compared node by node with the GPU's

    python bench.py [--steps K] [--warmup W] [--instances B] [--cpu-seconds S]
    python bench.py --copy-probe      # plain device copy for counter scaling

Other workloads (--workload): logic (random CMOS logic, DC and transient),
mixed (uneven circuit sizes), ac (AC sweeps of the mesh with capacitors),
noise (noise analysis of the same circuit), hightw (high-treewidth logic).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC circuit solves/sec at KCL resid ≤1e-9 A; NR iters/solve, seeded vs plain"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, public spec
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_rdmesh64_b4096.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(gib=2):
    """A plain device-to-device copy of a known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import torch
    n = gib * (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}), flush=True)


def seeder_probe():
    """The seeder's products at the learned-seeder configuration's shape,
    for profiling: an MLP 512 -> 3 x 1024 -> 20,000, 8192 synthetic
    training rows, two epochs (batch 1024), then three predictions of the
    8192 rows."""
    from karam_amd import nr
    ctx = nr.Context(0)
    rng = np.random.default_rng(0)
    F, O, rows = SEEDER_FEATURES, 20000, 8192
    s = nr.Seeder(ctx, F, O, hidden=SEEDER_HIDDEN, n_hidden=SEEDER_LAYERS)
    X = rng.standard_normal((rows, F)).astype(np.float32)
    s.record_solution(X, rng.standard_normal((rows, O)) * 0.1 + 0.9)
    t0 = time.perf_counter()
    loss = s.train(epochs=2, batch=1024)
    t_train = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(3):
        s.predict_seed(X)
    t_pred = (time.perf_counter() - t0) / 3
    print(json.dumps({"seeder_probe": {"train_2_epochs_ms": t_train * 1e3, "loss": loss,
                                       "predict_8192_ms_incl_host_copies": t_pred * 1e3}}), flush=True)
    return 0


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (lines `traffic_bytes_per_launch | <kernel> | <bytes>`), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for ln in open(TRAFFIC_FILE):
        f = [x.strip() for x in ln.split("|")]
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def seeded_parity(nl, fields, seeds, leg):
    """The FP64 oracle started from the GPU's seeds on the first instances,
    against one seeded leg's voltages and iteration counts."""
    from oracle import oracle as O
    k = len(seeds)
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), {f: v[:k] for f, v in fields.items()}, k, threads=8,
                              seeds=seeds, start_step=18, polish=1)
    return {"instances": k, "max_err_over_tol": err_over_tol(leg["_v_sample"], Vo),
            "iterations_equal": int((leg["_it_sample"] == so["iterations"]).sum()),
            "reference": "FP64 oracle started from the same seed at the last GMIN step, one polish "
                         "iteration"}


def cpu_baseline(nl, fields, B, V_gpu, it_gpu, seconds, threads, topo=None):
    """Oracle over a bounded sample of the same instances on `threads` host
    threads; returns (baseline dict, parity dict).  parity["_v"] holds the
    oracle's voltages of the sample (callers drop it before printing)."""
    from oracle import oracle as O
    T = topo if topo is not None else O.Topology(nl)
    n0 = min(B, threads)
    t = time.perf_counter()
    Vo, so = O.solve_dc_batch(nl, T, {k: v[:n0] for k, v in fields.items()}, n0, threads=threads)
    dt0 = time.perf_counter() - t
    n = int(min(B, max(n0, (seconds / max(dt0, 1e-3)) * n0) // threads * threads))
    n = max(n, n0)
    if n > n0:
        t = time.perf_counter()
        Vr, sr = O.solve_dc_batch(nl, T, {k: v[n0:n] for k, v in fields.items()}, n - n0, threads=threads)
        dt1 = time.perf_counter() - t
        Vo = np.concatenate([Vo, Vr])
        so = {k: np.concatenate([so[k], sr[k]]) for k in so}
        rate = (n - n0) / dt1
    else:
        rate = n0 / dt0
    err = np.abs(V_gpu[:n] - Vo) / (1e-9 * np.abs(Vo) + 1e-12)
    base = {"value": rate, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of {B} instances (same netlist and parameters), FP64 oracle "
                      f"(multifrontal LU on the reference fill pattern), OpenMP {threads} threads"}
    parity = {"instances": n, "max_err_over_tol": float(err.max()),
              "within_tol": int((err.max(axis=1) <= 1.0).sum()),
              "iterations_equal": int((it_gpu[:n] == so["iterations"]).sum()),
              "tolerance": "1e-9 rel + 1e-12 V abs", "_v": Vo}
    return base, parity


def ref_dense_baseline(nl, fields, it_gpu, threads, topo):
    """The linear solve the reference actually runs, solveLU_Pivoted
    (math/linalg.h:742-854, reached from solvers/circuitsim_solver.cpp:
    291-295), timed on the first Newton system (seed start, gmin 1e-3) of
    sampled instances and extrapolated to a solve as (mean Newton iterations
    per solve, which the GPU's equal) x (time per system); the stamps and the
    audit are O(n) beside it.  Two timings: the reference's own code built
    from its sources (oracle/_ref, one system on one thread: it takes tens of
    seconds at n = 4096) and the restatement the oracle holds (bit-identical
    answers), one system per host thread at once."""
    from concurrent.futures import ThreadPoolExecutor
    from karam_amd import _abi, netlists
    from oracle import oracle as O
    from oracle import ref as R
    n = nl.num_nodes
    rp, ci = topo.export(_abi.NR_EXPORT_ROW_PTR), topo.export(_abi.NR_EXPORT_COL_IDX)
    k = min(threads, len(it_gpu))
    systems = []
    for b in range(k):
        nb = netlists.instance(nl, fields, b)
        _, av, rhs = O.stamp(nb, topo, O.seed(nb), 1e-3)
        systems.append((av, rhs))
    iters = float(np.mean(it_gpu))
    out = {"unit": "solves/s", "newton_iterations_per_solve": iters, "extrapolated": True,
           "sample": f"first Newton system of instances 0..{k - 1}, n = {n}"}
    t = time.perf_counter()
    with ThreadPoolExecutor(k) as ex:
        xs = list(ex.map(lambda sy: O.lu_dense(n, rp, ci, *sy), systems))
    dt = time.perf_counter() - t
    out["restatement"] = {"value": k / (iters * dt), "cores": k, "s_per_system": dt, "kind": "port"}
    if R.available():
        t = time.perf_counter()
        rc, x = R.lu_pivoted(n, rp, ci, *systems[0])
        dt1 = time.perf_counter() - t
        out["reference_code"] = {"value": 1.0 / (iters * dt1), "cores": 1, "s_per_system": dt1, "kind": "reference",
                                 "same_answer_as_restatement": bool(np.array_equal(x, xs[0][1])),
                                 "build": "oracle/Makefile ref: -O2, no AVX2 (the SIMD units do not compile here)"}
    else:
        out["reference_code"] = None
    log(f"ref-faithful dense LU: {out}")
    return out


SEEDER_FEATURES, SEEDER_HIDDEN, SEEDER_LAYERS = 512, 1024, 3


def timed_solves(bt, steps, **kw):
    """`steps` batched DC solves timed without per-kernel events (as every
    leg is), then one more solve with a HIP event pair around every launch for
    the kernel rows; returns the converged-solve rate, ms per solve, the
    kernel rows and the last timed solve's statistics and result."""
    from karam_amd import _abi
    conv = 0
    t0 = time.perf_counter()
    for k in range(steps):
        r = bt.solve_dc(**kw)
        st = bt.stats()
        conv += int(((st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)).sum())
        log(f"step {k} {kw}: {r['solve_ms']:.1f} ms, converged {r['n_converged']}, sweeps {r['batched_iterations']}")
    total = time.perf_counter() - t0
    rp = bt.solve_dc(record_timing=2, **{**kw, "use_graph": 0})     # events need eager launches
    kst = {name: {f: x[f] for f in ("launches", "ms", "bytes", "flops")} for name, x in bt.kernel_stats().items()}
    log(f"profiled solve {kw}: {rp['solve_ms']:.1f} ms")
    return conv / total, total * 1e3 / steps, kst, st, r


def roofline(kst, traffic=True):
    dom = max(kst, key=lambda n: kst[n]["ms"])
    d = kst[dom]
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(dom) if traffic else None, "kernel": dom,
            "algorithmic_bytes_per_launch": d["bytes"] / d["launches"],
            "mean_launch_us": d["ms"] * 1e3 / d["launches"], "launches": d["launches"],
            "device_ms_share": d["ms"] / sum(x["ms"] for x in kst.values())}


FP64_PEAK_TFS = 78.6           # MI355X FP64, public spec


def roofline_either(kst):
    """The dominant kernel against whichever of HBM bandwidth and the FP64
    peak it is closer to (its algorithmic bytes and flops per launch)."""
    dom = max(kst, key=lambda n: kst[n]["ms"])
    d = kst[dom]
    gbs = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    tfs = d["flops"] / (d["ms"] * 1e-3) / 1e12
    mem = gbs / HBM_PEAK_GBS >= tfs / FP64_PEAK_TFS
    ach, peak = (gbs, HBM_PEAK_GBS) if mem else (tfs, FP64_PEAK_TFS)
    return {"bound": "hbm" if mem else "fp64", "achieved": ach, "peak": peak, "unit": "GB/s" if mem else "TFLOP/s",
            "frac": ach / peak, "traffic": None, "kernel": dom, "algorithmic_bytes_per_launch": d["bytes"] / d["launches"],
            "algorithmic_flops_per_launch": d["flops"] / d["launches"], "mean_launch_us": d["ms"] * 1e3 / d["launches"],
            "launches": d["launches"], "device_ms_share": d["ms"] / sum(x["ms"] for x in kst.values())}


def kernel_table(kst):
    return {n: {"launches": x["launches"], "ms": round(x["ms"], 3),
                "GBps": round(x["bytes"] / max(x["ms"], 1e-9) / 1e6, 1)} for n, x in
            sorted(kst.items(), key=lambda kv: -kv[1]["ms"])}


def err_over_tol(v, ref):
    """max |v - ref| / (1e-9 |ref| + 1e-12 V)"""
    return float((np.abs(v - ref) / (1e-9 * np.abs(ref) + 1e-12)).max())


def iter_stats(it):
    return {"mean": float(it.mean()), "p50": float(np.percentile(it, 50)), "p99": float(np.percentile(it, 99))}


def seeded_path(a, ctx, nl, T, bt, fields, B, V64):
    """The learned-seeder path on the same instances: train on GPU-solved
    instances drawn from a disjoint seed range (untimed, reported), then time
    `steps` solves that each predict every instance's seed on the device and
    start Newton at the last GMIN step, with the plain start as fallback, on
    the FP64 and on the FP32 path.  A seeded solve takes one Newton iteration
    after its convergence test passes (polish = 1): the test accepts KCL <
    1e-9 A, which leaves nodes up to ~1e-6 V from the root, while the plain
    start's last GMIN step ends at the root.  A solve counts when it is
    KCL-converged and within the voltage tolerance (1e-9 rel + 1e-12 V) of
    the plain FP64 answer.  Two seeds without learning start at the same step
    for comparison: the mean of the training solutions, and the operating
    point of the nominal netlist."""
    from karam_amd import _abi, netlists, nr
    n = T.info["n"]
    t0 = time.perf_counter()
    s = nr.Seeder(ctx, SEEDER_FEATURES, n, hidden=SEEDER_HIDDEN, n_hidden=SEEDER_LAYERS, init_seed=1)
    gen = nr.Batch(ctx, T, B)
    done, vsum, vcnt = 0, np.zeros(n), 0
    while done < a.seeder_train:
        m = min(B, a.seeder_train - done)
        ftr = netlists.mc_fields(nl, B, seed=500000 + done)
        gen.set_fields(ftr)
        gen.solve_dc()
        st = gen.stats()
        ok = st["status"] == _abi.NR_STATUS_CONVERGED
        ok[m:] = False
        Vg = gen.voltages()[ok]
        s.record_solution(netlists.mc_features(nl, ftr, SEEDER_FEATURES)[ok], Vg)
        vsum += Vg.sum(axis=0)
        vcnt += len(Vg)
        done += m
    gen.close()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = s.train(epochs=a.seeder_epochs, batch=256, lr=1e-3)
    t_train = time.perf_counter() - t0
    log(f"seeder: {s.history()} solutions in {t_data:.1f} s, trained in {t_train:.1f} s, loss {loss:.4f}")
    X = netlists.mc_features(nl, fields, SEEDER_FEATURES)
    seed_v = s.predict_seed(X[:8])
    dims = [SEEDER_FEATURES] + [SEEDER_HIDDEN] * SEEDER_LAYERS + [(n + 63) // 64 * 64]
    flops = 2.0 * B * sum(x * y for x, y in zip(dims[:-1], dims[1:]))
    mean_seed = np.tile(vsum / max(1, vcnt), (B, 1))
    nom = nr.Batch(ctx, T, 1)
    nom.solve_dc()
    nominal_seed = np.tile(nom.voltages()[0], (B, 1))
    nom.close()

    def counted(st, V):
        ok = (st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)
        within = (np.abs(V - V64) <= 1e-9 * np.abs(V64) + 1e-12).all(axis=1)
        return ok & within

    def leg(prec, set_seed, label):
        set_seed()
        bt.solve_dc(seeded_start_step=18, precision=prec, polish=1, use_graph=a.graph)         # warm-up
        t_set, sweeps = 0.0, 0
        t0 = time.perf_counter()
        for k in range(a.steps):
            tp = time.perf_counter()
            set_seed()
            t_set += time.perf_counter() - tp
            r = bt.solve_dc(seeded_start_step=18, precision=prec, polish=1, use_graph=a.graph)
            st = bt.stats()
            sweeps = r["batched_iterations"]
            log(f"{label} step {k}: {r['solve_ms']:.1f} ms, converged {r['n_converged']}, sweeps {sweeps}")
        total = time.perf_counter() - t0
        # every step solves the same instances from the same seeds, so the last
        # step's count is every step's
        V = bt.voltages()
        conv = int(counted(st, V).sum()) * a.steps
        return {
            "value": conv / total, "unit": "solves/s", "ms_per_step": total * 1e3 / a.steps,
            "speedup_vs_plain": None, "batched_sweeps_per_step": sweeps,
            "fallback_instances": int((st["seeded"] == 0).sum()),
            "counted_instances": int(counted(st, V).sum()),
            "iterations": iter_stats(st["iterations"]),
            "max_err_over_tol_vs_plain_fp64": err_over_tol(V, V64),
            "seed_ms_per_step": t_set * 1e3 / a.steps,
            "_v_sample": V[:8], "_it_sample": st["iterations"][:8],
        }

    out = {}
    for key, prec in (("fp64", _abi.NR_PREC_FP64), ("fp32_refine", _abi.NR_PREC_FP32_REFINE)):
        out[key] = leg(prec, lambda: bt.set_seed_from(s, X), f"seeded {key}")
    for key, seed in (("mean_seed_fp32_refine", mean_seed), ("nominal_seed_fp32_refine", nominal_seed)):
        out[key] = leg(_abi.NR_PREC_FP32_REFINE, lambda sd=seed: bt.set_seed(sd), key)
        for k in [k for k in out[key] if k.startswith("_")]:
            del out[key][k]
    out["seeder"] = {"features": SEEDER_FEATURES, "hidden": SEEDER_HIDDEN, "layers": SEEDER_LAYERS, "outputs": n,
                     "train_instances": int(s.history()), "train_data_s": t_data, "train_s": t_train,
                     "final_loss_normalised": loss, "polish": 1,
                     "predict_tflops_wall": flops / (out["fp64"]["seed_ms_per_step"] * 1e-3) / 1e12,
                     "mfma_peak_tflops_bf16": 2500.0}
    out["_seed_sample"] = seed_v
    return out


# The random-logic configuration: NAND2 / NAND3 / INV gates (fan-in 1-3,
# fan-out <= 4), 20,000 nodes and 39,934 level-1 MOSFETs (every gate brings
# two MOSFETs per node), about 17 gate levels of 640 gates; as
# tests/golden/make_logic20k.py, whose fixture holds its Min-Fill ordering.
LOGIC20K = dict(n_target=20000, seed=1, width=640, p_inv=0.3, p_nand3=0.1)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def host_info(threads):
    """The host the CPU baseline ran on: CPU model, logical CPUs, the CPUs
    this process may use, and the threads the oracle used."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "threads_used": threads}


def cpu_threads(a):
    """Oracle threads: every CPU this process may use, at most 16 (the host
    share of one GPU on the benchmark machines)."""
    return a.cpu_threads or min(16, len(os.sched_getaffinity(0)))


def logic_workload(a):
    """Random CMOS logic at the configured size (LOGIC20K: 20,000 nodes,
    39,934 level-1 MOSFETs), 2048 Monte-Carlo instances (per-device VTO
    sigma 30 mV, KP sigma 3 %): DC operating points on the plain FP64 path and
    on the FP32 + FP64-refinement path, then a 500-step backward-Euler
    transient of 0.5 ns from the operating points on both paths, with the
    FP32 path's per-node error against FP64 and the GPU's against the oracle
    on a bounded sample.  value = converged FP64 DC solves/s."""
    from karam_amd import _abi, netlists, nr
    B = a.instances if a.instances != 4096 else 2048
    kw = dict(LOGIC20K) if a.logic_nodes == LOGIC20K["n_target"] else \
        dict(n_target=a.logic_nodes, seed=1, width=max(32, a.logic_nodes // 31), p_inv=0.3, p_nand3=0.1)
    nl = netlists.random_logic(**kw)
    f = netlists.mc_fields(nl, B, seed=1000, sigma_vto=0.03, sigma_kp=0.03)
    ctx = nr.Context(0)
    t = time.perf_counter()
    T = nr.Topology(nl, ctx)
    symbolic_ms = (time.perf_counter() - t) * 1e3
    Tt = nr.Topology(nl, ctx, transient=True)
    log(f"logic n={T.info['n']} mos={T.info['n_mos']} nnz_f={T.info['nnz_f']} tw={T.info['treewidth']} "
        f"levels={T.info['n_levels']} max_front={T.info['max_front']} symbolic {symbolic_ms:.0f} ms")
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    tune = {}
    if not a.no_tune:                 # untimed: launch shapes and the whole-phase kernels, per precision
        for key, prec in (("fp64", _abi.NR_PREC_FP64), ("fp32_refine", _abi.NR_PREC_FP32_REFINE)):
            t0 = time.perf_counter()
            tu = bt.autotune(precision=prec, reps=2)
            tune[key] = {"s": time.perf_counter() - t0, "factor_solve_ms_before": tu["ms_before"],
                         "factor_solve_ms_levels": tu["ms_levels"], "factor_solve_ms_persistent": tu["ms_persistent"],
                         "persistent": tu["persistent"], "factor_solve_ms_after": tu["ms_after"],
                         "factor_solve_ms_graph": tu["ms_graph"], "factor_solve_ms_graph_fanout": tu["ms_graph_fanout"],
                         "fanout": tu["fanout"], "few_instances_ms_before": tu["ms_few_before"],
                         "few_instances_ms_after": tu["ms_few_after"], "few_instances_map_up_to": tu["few_max"]}
            log(f"autotune {key}: {tu['ms_before']:.2f} -> {tu['ms_after']:.2f} ms per factor + back substitution "
                f"(per-level {tu['ms_levels']:.2f}, whole-phase {tu['ms_persistent']:.2f}; graph {tu['ms_graph']:.2f}, "
                f"fanned out {tu['ms_graph_fanout']:.2f})")
    for _ in range(a.warmup):
        bt.solve_dc(use_graph=a.graph)
    rate, ms, kst, st, r = timed_solves(bt, a.steps, use_graph=a.graph)
    V, it = bt.voltages(), st["iterations"]
    ok64 = (st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)
    bt.solve_dc(precision=_abi.NR_PREC_FP32_REFINE, use_graph=a.graph)            # warm-up
    rate32, ms32, kst32, st32, r32 = timed_solves(bt, a.steps, precision=_abi.NR_PREC_FP32_REFINE,
                                                  use_graph=a.graph)
    V32 = bt.voltages()
    ok32 = (st32["status"] == _abi.NR_STATUS_CONVERGED) & (st32["kcl"] <= 1e-9)
    both = ok64 & ok32
    bt.close()
    probes = [2, 3, T.info["n"] // 2, T.info["n"]]
    # the transient with the inputs' pulses (the configuration), and with the
    # inputs held at their DC values (every step a quiescent Newton solve)
    quiet = dataclasses.replace(nl, vsrc_wave=np.zeros_like(nl.vsrc_wave))
    Tq = nr.Topology(quiet, ctx, transient=True)
    tran = {}
    for key, topo, prec in (("fp64", Tt, _abi.NR_PREC_FP64), ("fp32_refine", Tt, _abi.NR_PREC_FP32_REFINE),
                            ("quiescent_fp64", Tq, _abi.NR_PREC_FP64),
                            ("quiescent_fp32_refine", Tq, _abi.NR_PREC_FP32_REFINE)):
        tr = nr.Batch(ctx, topo, B)
        tr.set_fields(f)
        tr.set_seed(V)
        t0 = time.perf_counter()
        rt = tr.transient(h=0.5e-9, steps=a.tran_steps, start=_abi.NR_TRAN_START_SEED, probes=probes,
                          use_graph=a.tran_graph, precision=prec)
        dt = time.perf_counter() - t0
        stt, ts = tr.stats(), tr.tran_stats()
        done = stt["status"] == _abi.NR_STATUS_CONVERGED
        acc = ts["accepted"]
        tran[key] = {"wall_s": dt, "instances_completed": int(rt["n_converged"]),
                     "accepted_steps_mean": float(acc.mean()), "accepted_steps_min": int(acc.min()),
                     "instance_steps_per_s": float(acc.sum()) / dt,
                     "newton_iterations_per_accepted_step": float(stt["iterations"].sum() / max(1, acc.sum())),
                     "batched_sweeps": rt["batched_iterations"],
                     "_v": tr.voltages(), "_done": done, "_wave": tr.waveform(), "_rec": tr.wave_recorded(),
                     "_st": stt}
        tr.close()
        log(f"transient {key}: {dt:.1f} s, completed {rt['n_converged']}/{B}, accepted steps mean {acc.mean():.1f}")
    Tq.close()
    td = tran["fp64"]["_done"] & tran["fp32_refine"]["_done"]
    tq = tran["quiescent_fp64"]["_done"] & tran["quiescent_fp32_refine"]["_done"]
    out = {
        "metric": METRIC, "value": rate, "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded levelised random NAND2/NAND3/INV logic, per-device VTO sigma 30 mV and KP "
                "sigma 3 %",
        "config": {"workload": f"logic{T.info['n']} Monte-Carlo DC operating point + {a.tran_steps}-step 0.5 ns "
                               f"backward-Euler transient, FP64 and FP32 + FP64 refinement",
                   "netlist": kw, "instances": B, "nodes": T.info["n"], "mosfets": T.info["n_mos"],
                   "nnz_a": T.info["nnz_a"], "nnz_f": T.info["nnz_f"], "treewidth": T.info["treewidth"],
                   "symbolic_ms": symbolic_ms, "sweeps_from_hip_graphs": bool(a.graph),
                   "batched_sweeps_per_step": r["batched_iterations"], "supernodes": T.info["n_supernodes"], "levels": T.info["n_levels"]},
        "autotune": tune or None,
        "converged_fp64": int(ok64.sum()),
        "nr_iters_per_solve": {"plain": iter_stats(it[ok64]) if ok64.any() else None,
                               "plain_fp32_refine": iter_stats(st32["iterations"][ok32]) if ok32.any() else None},
        "roofline": roofline(kst, traffic=False), "kernels": kernel_table(kst),
        "fp32_refine": {"value": rate32, "unit": "solves/s", "ms_per_step": ms32,
                        "speedup_vs_plain": rate32 / rate if rate else None, "converged": int(ok32.sum()),
                        "finished_on_fp64": int((st32["lu_fallbacks"] > 0).sum()),
                        "max_err_over_tol_vs_plain_fp64": err_over_tol(V32[both], V[both]) if both.any() else None,
                        "roofline": roofline(kst32, traffic=False)},
        "transient": {"steps": a.tran_steps, "h_s": 0.5e-9, "start": "DC operating point (FP64)",
                      "note": "the reference's transient Newton (no MOSFET limiting, no damping, 50 iterations) "
                              "stops at the first input edge on most instances of this logic; the GPU stops "
                              "where the oracle does. The quiescent legs hold the inputs at their DC values.",
                      **{k: {x: y for x, y in tran[k].items() if not x.startswith("_")} for k in tran},
                      "fp32_max_err_over_tol_vs_fp64_final": err_over_tol(tran["fp32_refine"]["_v"][td],
                                                                          tran["fp64"]["_v"][td])
                      if td.any() else None,
                      "quiescent_fp32_max_err_over_tol_vs_fp64_final": err_over_tol(
                          tran["quiescent_fp32_refine"]["_v"][tq], tran["quiescent_fp64"]["_v"][tq])
                      if tq.any() else None},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads(a)
        To = oracle_topology(nl, kw, T)
        base, parity = cpu_baseline(nl, f, B, V, it, a.cpu_seconds, threads, topo=To)
        base["host"] = host_info(threads)
        out["cpu_baseline"] = base
        out["parity_sample"] = parity
        # the FP32 path's per-node error against the FP64 oracle on the sample
        k = parity["instances"]
        Vo = parity.pop("_v")
        sel = ok32[:k]
        parity["fp32_refine_max_err_over_tol_vs_oracle"] = err_over_tol(V32[:k][sel], Vo[sel]) if sel.any() else None
        # where the FP32 answer differs from FP64, the difference as a current
        # at each node (G_ii |dv_i|, G = the DC Jacobian's diagonal at the FP64
        # answer) against the KCL acceptance of 1e-9 A: NAND stack nodes held
        # by off transistors and gmin (G ~ 3e-12 S) take volts of freedom
        bb = np.flatnonzero(both)[:8]
        out["fp32_refine"]["max_node_current_diff_A_vs_plain_fp64"] = max(
            float((O.node_conductance(netlists.instance(nl, f, b), To, V[b]) * np.abs(V32[b] - V[b])).max())
            for b in bb) if len(bb) else None
        out["fp32_refine"]["node_current_diff_instances"] = len(bb)
        # the transient's first steps of two instances against the oracle
        m, s = min(2, B), min(a.tran_steps, a.tran_check_steps)
        To_t = oracle_topology(nl, kw, Tt, transient=True)
        _, Wo, so = O.solve_transient_batch(nl, To_t, {x: y[:m] for x, y in f.items()}, m, h=0.5e-9, steps=s,
                                            start=1, v0=V[:m], probes=probes, threads=m)
        # the same steps on the GPU: statuses, accepted steps, iterations and
        # the probe waveform up to where each instance stopped
        tr = nr.Batch(ctx, Tt, m)
        tr.set_fields({x: y[:m] for x, y in f.items()})
        tr.set_seed(V[:m])
        tr.transient(h=0.5e-9, steps=s, start=_abi.NR_TRAN_START_SEED, probes=probes)
        W, sg, rec = tr.waveform(), tr.stats(), tr.wave_recorded()
        tr.close()
        same = [bool(sg["status"][b] == so["status"][b] and sg["iterations"][b] == so["iterations"][b]
                     and rec[b] == so["recorded"][b] and np.array_equal(W[b, :rec[b]], Wo[b, :rec[b]]))
                for b in range(m)]
        out["transient"]["parity_sample"] = {
            "instances": m, "steps": s, "bit_identical": int(sum(same)),
            "steps_accepted": [int(so["steps_done"][b]) for b in range(m)],
            "reference": "FP64 oracle (same start, same steps): status, iterations, probe waveform"}
        log(f"cpu baseline {base['value']:.2f} solves/s on {threads} threads; parity {parity}")
    print(json.dumps(out), flush=True)
    return 0


def oracle_topology(nl, kw, T, transient=False):
    """The oracle's topology of `nl`.  For the configured netlist its literal
    Min-Fill ordering comes from the fixture (recomputing it takes about a
    quarter of an hour), checked against the library's; otherwise the oracle
    runs Min-Fill itself."""
    from karam_amd import _abi
    from oracle import oracle as O
    fx = os.path.join(GOLDEN, "logic20k_ordering.npz")
    if kw == LOGIC20K and os.path.exists(fx):
        with np.load(fx, allow_pickle=False) as g:
            o = g["ordering_raw"]
        if not np.array_equal(o, T.export(_abi.NR_EXPORT_ORDERING_RAW)):
            raise RuntimeError("the library's Min-Fill ordering differs from the oracle's fixture")
        return O.Topology(nl, transient=transient, ordering=o)
    return O.Topology(nl, transient=transient)


def small_signal_batch(a):
    """The AC / noise workloads' circuit: the benchmark mesh (R-diode mesh with
    a 2 pF capacitor at every node), its Monte-Carlo instances and their DC
    operating points, and a batch on the transient topology seeded with
    them."""
    from karam_amd import netlists, nr
    from karam_amd.netlists import _f64, _i32
    B = a.instances if a.instances != 4096 else 1024
    nl = netlists.rd_mesh(a.side, seed=1)
    n = nl.num_nodes
    nl.cap_n1, nl.cap_n2, nl.cap_farads = _i32(np.arange(1, n + 1)), _i32(np.zeros(n)), _f64(np.full(n, 2e-12))
    f = netlists.mc_fields(nl, B, seed=1000)
    ctx = nr.Context(0)
    T = nr.Topology(nl, ctx)
    D = nr.Batch(ctx, T, B)
    D.set_fields(f)
    D.solve_dc()
    V = D.voltages()
    D.close()
    Tt = nr.Topology(nl, ctx, transient=True)
    bt = nr.Batch(ctx, Tt, B)
    bt.set_fields(f)
    bt.set_seed(V)
    return nl, f, V, Tt, bt, B, n


def ac_workload(a):
    """AC small-signal sweeps of the benchmark mesh (64 x 64 R-diode mesh with
    a 2 pF capacitor at every node): DC operating points of the Monte-Carlo
    instances, then a.ac_points log-spaced frequencies from 1 Hz to 10 GHz,
    excited at node 1 as solveAC excites it.  One frequency point of one
    instance is one complex sparse factorisation and solve."""
    nl, f, V, Tt, bt, B, n = small_signal_batch(a)
    probes = [1, 2, n // 2, n]
    kw = dict(f_start=1.0, f_stop=10e9, input_node=1, amplitude=1.0, probes=probes)
    for _ in range(a.warmup):
        bt.ac(n_points=2, **kw)
    t0 = time.perf_counter()
    freqs, H, st = bt.ac(n_points=a.ac_points, record_timing=2, **kw)
    wall = time.perf_counter() - t0
    kst = bt.kernel_stats()
    points = B * a.ac_points
    log(f"ac: {points} instance-frequency points in {wall:.2f} s, singular {(st != 0).sum()}")
    out = {
        "metric": "AC instance-frequency points/sec", "value": points / wall, "unit": "points/s", "n_gpus": 1,
        "steps": a.ac_points, "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.ac_points, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex f64",
        "data": "synthetic: seeded R-diode mesh with 2 pF per node, per-instance R and IS sigma 5 %",
        "config": {"workload": f"rdmesh{a.side} + C, AC sweep 1 Hz - 10 GHz, {a.ac_points} points",
                   "instances": B, "nodes": n, "nnz_f": Tt.info["nnz_f"], "max_front": Tt.info["max_front"]},
        "kernels": kernel_table(kst), "roofline": None, "cpu_baseline": None,
    }
    if not a.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads(a)
        m = min(B, threads)
        To = O.Topology(nl, transient=True)
        t0 = time.perf_counter()
        fo, Ho, so = O.solve_ac_batch(nl, To, {k: v[:m] for k, v in f.items()}, m, V[:m], n_points=a.ac_points,
                                      threads=threads, **kw)
        t_cpu = time.perf_counter() - t0
        err = np.abs(H[:m] - Ho) / (1e-9 * np.abs(Ho) + 1e-12 * np.abs(Ho).max())
        out["cpu_baseline"] = {"value": m * a.ac_points / t_cpu, "unit": "points/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} of {B} instances, all {a.ac_points} points, oracle (complex "
                                         f"multifrontal LU on the reference fill pattern), one instance per thread"}
        out["parity_sample"] = {"instances": m, "max_err_over_tol": float(err.max()),
                                "tolerance": "1e-9 rel + 1e-12 x max |H|"}
        log(f"ac cpu baseline {m * a.ac_points / t_cpu:.1f} points/s; parity {out['parity_sample']}")
    print(json.dumps(out), flush=True)
    return 0


def noise_workload(a):
    """Noise analysis of the benchmark mesh (the AC workload's circuit and
    operating points): output noise PSD at the middle node over a.ac_points
    log-spaced frequencies from 1 Hz to 10 GHz, and its integral.  Sources:
    every resistor's thermal noise and every diode's shot noise.  One
    frequency point of one instance is one complex sparse factorisation of
    Y^T, one solve and the sum over the sources."""
    nl, f, V, Tt, bt, B, n = small_signal_batch(a)
    out_node = n // 2 + a.side // 2
    kw = dict(f_start=1.0, f_stop=10e9, output_node=out_node, temp_K=300.15)
    for _ in range(a.warmup):
        bt.noise(n_points=2, **kw)
    t0 = time.perf_counter()
    freqs, psd, vrms, st = bt.noise(n_points=a.ac_points, record_timing=2, **kw)
    wall = time.perf_counter() - t0
    kst = bt.kernel_stats()
    points = B * a.ac_points
    n_src = len(nl.res_n1) + len(nl.dio_anode) + len(nl.mos_d)
    log(f"noise: {points} instance-frequency points in {wall:.2f} s, singular {(st != 0).sum()}, "
        f"vrms median {np.median(vrms):.3e} V")
    out = {
        "metric": "noise instance-frequency points/sec", "value": points / wall, "unit": "points/s", "n_gpus": 1,
        "steps": a.ac_points, "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.ac_points, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex f64",
        "data": "synthetic: seeded R-diode mesh with 2 pF per node, per-instance R and IS sigma 5 %",
        "config": {"workload": f"rdmesh{a.side} + C, noise at node {out_node}, 1 Hz - 10 GHz, {a.ac_points} points",
                   "instances": B, "nodes": n, "noise_sources": n_src, "nnz_f": Tt.info["nnz_f"]},
        "kernels": kernel_table(kst), "roofline": None, "cpu_baseline": None,
    }
    if not a.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads(a)
        m = min(B, threads)
        To = O.Topology(nl, transient=True)
        t0 = time.perf_counter()
        fo, po, vo, so = O.solve_noise_batch(nl, To, {k: v[:m] for k, v in f.items()}, m, V[:m],
                                             n_points=a.ac_points, mode=O.NOISE_ADJOINT, threads=threads, **kw)
        t_cpu = time.perf_counter() - t0
        err = np.abs(psd[:m] - po) / (1e-9 * np.abs(po))
        out["cpu_baseline"] = {"value": m * a.ac_points / t_cpu, "unit": "points/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} of {B} instances, all {a.ac_points} points, oracle (one "
                                         f"transposed complex multifrontal solve per point, as the GPU), one "
                                         f"instance per thread; the reference's formulation solves once per "
                                         f"source ({n_src} per point)"}
        out["parity_sample"] = {"instances": m, "max_err_over_tol": float(err.max()), "tolerance": "1e-9 rel"}
        log(f"noise cpu baseline {m * a.ac_points / t_cpu:.1f} points/s; parity {out['parity_sample']}")
    print(json.dumps(out), flush=True)
    return 0


def hightw_workload(a):
    """DC operating points of high-treewidth random logic: NAND2 / INV gates
    whose inputs come from 65 slots of the previous six levels (no locality),
    a.hightw_nodes nodes (2000: treewidth 214, fronts up to order 215 with
    their panels in HBM), Monte-Carlo instances with VTO sigma 30 mV and KP
    sigma 3 %.  The reference sends circuits of treewidth above 50 away from
    its sparse LU (to tensor contraction, Schur decomposition or PCG with a
    dense fallback); here they run on the same multifrontal factorisation.
    Launch shapes autotuned first (untimed)."""
    from karam_amd import _abi, netlists, nr
    B = a.instances if a.instances != 4096 else 1024
    nl = netlists.random_logic(a.hightw_nodes, seed=3, spread=32, window=6, width=64)
    f = netlists.mc_fields(nl, B, seed=1000, sigma_vto=0.03, sigma_kp=0.03)
    ctx = nr.Context(0)
    t = time.perf_counter()
    T = nr.Topology(nl, ctx)
    symbolic_ms = (time.perf_counter() - t) * 1e3
    info = {k: T.info[k] for k in ("n", "nnz_f", "treewidth", "max_front", "n_supernodes")}
    log(f"hightw {info} symbolic {symbolic_ms:.0f} ms")
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    tune = None
    if not a.no_tune:
        tu = bt.autotune(precision=_abi.NR_PREC_FP64, reps=2)
        tune = {"factor_solve_ms_before": tu["ms_before"], "factor_solve_ms_after": tu["ms_after"]}
        log(f"autotune: {tu['ms_before']:.2f} -> {tu['ms_after']:.2f} ms per factor + back substitution")
    for _ in range(a.warmup):
        bt.solve_dc()
    rate, ms, kst, st, r = timed_solves(bt, a.steps)
    V = bt.voltages()
    it = st["iterations"]
    rate32, ms32, kst32, st32, _ = timed_solves(bt, a.steps, precision=_abi.NR_PREC_FP32_REFINE)
    out = {
        "metric": "MC circuit solves/sec at KCL resid <=1e-9 A (high-treewidth logic)", "value": rate,
        "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded random NAND2/INV logic without locality, per-instance VTO / KP",
        "config": {"workload": f"random logic, {info['n']} nodes, treewidth {info['treewidth']}", "instances": B,
                   **info, "symbolic_ms": symbolic_ms, "autotune": tune},
        "iterations_mean": float(it.mean()), "converged": int((st["status"] == 0).sum()),
        "roofline": roofline_either(kst), "kernels": kernel_table(kst),
        "fp32_refine": {"value": rate32, "ms_per_step": ms32, "roofline": roofline_either(kst32),
                        "max_err_vs_fp64": float(np.max(np.abs(bt.voltages() - V)))},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        threads = cpu_threads(a)
        base, parity = cpu_baseline(nl, f, B, V, it, a.cpu_seconds, threads)
        parity.pop("_v")
        base["host"] = host_info(threads)
        out["cpu_baseline"], out["parity_sample"] = base, parity
        log(f"hightw cpu baseline {base['value']:.1f} solves/s; parity {parity}")
    print(json.dumps(out), flush=True)
    return 0


def mixed_workload(a):
    """Uneven sizes: a.mixed_instances instances whose node counts are
    log-uniform in [200, 50000], the family of instance i being i mod 3 (the
    R-diode mesh, the inverter chain, random NAND2/NAND3/INV logic at the
    logic configuration's depth, about n / 31 gates per level).  Node counts
    are bucketed into 8 log-spaced sizes per family; each bucket is one
    circuit with its own sparsity pattern and ordering, solved as one batch of
    Monte-Carlo variants: DC operating points on the FP32 path with hipGraph
    sweeps, launch shapes autotuned per circuit (untimed).  The circuits are
    dealt to a.mixed_streams contexts (one HIP stream each), largest first to
    the least loaded, and solved concurrently from as many host threads, so
    that small circuits run beside the large ones.  Symbolic analysis and
    tuning happen before the timed region and are reported separately.
    After the timed run, one eager solve per circuit with per-kernel events
    gives the roofline of the dominant kernel over the whole workload, and
    the oracle times one instance of each circuit on its own host thread for
    the CPU baseline."""
    from concurrent.futures import ThreadPoolExecutor
    from karam_amd import _abi, netlists, nr
    rng = np.random.default_rng(12345)
    N = a.mixed_instances
    n_of = np.exp(rng.uniform(np.log(200.0), np.log(50000.0), N))
    fam_of = np.arange(N) % 3
    edges = np.exp(np.linspace(np.log(200.0), np.log(50000.0), 9))
    bk = np.clip(np.searchsorted(edges, n_of, side="right") - 1, 0, 7)
    S = max(1, a.mixed_streams)
    ctxs = [nr.Context(0) for _ in range(S)]
    prec = _abi.NR_PREC_FP32_REFINE if a.mixed_precision == "fp32" else _abi.NR_PREC_FP64
    fams = ("rdmesh", "chain", "logic")
    jobs = []
    for fam in range(3):
        for b in range(8):
            cnt = int(((fam_of == fam) & (bk == b)).sum())
            if cnt == 0:
                continue
            nrep = float(np.sqrt(edges[b] * edges[b + 1]))
            if fam == 0:
                side = max(2, int(round(np.sqrt(nrep))))
                nl = netlists.rd_mesh(side, seed=100 + b)
                kw = {}
            elif fam == 1:
                nl = netlists.inverter_chain(max(1, int(round(nrep)) - 2))
                kw = dict(sigma_vto=0.03, sigma_kp=0.03)
            else:
                nn = int(round(nrep))
                nl = netlists.random_logic(nn, seed=200 + b, width=max(32, nn // 31), p_inv=0.3, p_nand3=0.1)
                kw = dict(sigma_vto=0.03, sigma_kp=0.03)
            jobs.append({"fam": fam, "b": b, "cnt": cnt, "nl": nl, "kw": kw, "nodes": nl.num_nodes})
    # largest first to the least loaded context (load ~ nodes x instances)
    load = [0.0] * S
    for j in sorted(jobs, key=lambda j: -j["nodes"] * j["cnt"]):
        w = int(np.argmin(load))
        j["w"] = w
        load[w] += j["nodes"] * j["cnt"]
    symb_s = tune_s = 0.0
    for j in jobs:
        ctx = ctxs[j["w"]]
        t0 = time.perf_counter()
        j["T"] = nr.Topology(j["nl"], ctx)
        j["symbolic_s"] = time.perf_counter() - t0
        symb_s += j["symbolic_s"]
        j["f"] = netlists.mc_fields(j["nl"], j["cnt"], seed=1000 + 37 * j["b"] + j["fam"], **j["kw"])
        j["bt"] = nr.Batch(ctx, j["T"], j["cnt"])
        j["bt"].set_fields(j["f"])
        t0 = time.perf_counter()
        if not a.no_tune:
            j["bt"].autotune(precision=prec, reps=2)
        # one stream per level: four contexts already solve concurrently, and
        # with the level fan-out on as well the workload ran slower
        j["bt"].set_fanout(False, precision=prec)
        tune_s += time.perf_counter() - t0
        log(f"set up {fams[j['fam']]} n={j['T'].info['n']} B={j['cnt']}: symbolic {j['symbolic_s'] * 1e3:.0f} ms, "
            f"tuned in {time.perf_counter() - t0:.1f} s")

    def run(w):
        for j in jobs:
            if j["w"] == w:
                t0 = time.perf_counter()
                j["r"] = j["bt"].solve_dc(precision=prec, use_graph=1)
                j["solve_s"] = time.perf_counter() - t0
                j["st"] = j["bt"].stats()
                log(f"solved {fams[j['fam']]} n={j['nodes']} B={j['cnt']} on stream {w}: "
                    f"{j['solve_s'] * 1e3:.0f} ms")
        return w

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=S) as ex:
        list(ex.map(run, range(S)))
    solve_s = time.perf_counter() - t0
    # per-kernel events: one eager solve per circuit (untimed), summed
    kst = {}
    for j in jobs:
        j["bt"].solve_dc(precision=prec, record_timing=2)
        log(f"kernel events {fams[j['fam']]} n={j['nodes']}")
        for k, x in j["bt"].kernel_stats().items():
            y = kst.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0})
            for q in y:
                y[q] += x[q]
    rows, conv_total = [], 0
    for j in jobs:
        st, T = j["st"], j["T"]
        ok = (st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)
        j["ok"] = ok
        j["Vfirst"] = j["bt"].voltages()[0]
        j["bt"].close()
        conv = int(ok.sum())
        conv_total += conv
        rows.append({"family": fams[j["fam"]], "nodes": T.info["n"], "instances": j["cnt"], "converged": conv,
                     "nnz_f": T.info["nnz_f"], "treewidth": T.info["treewidth"], "stream": j["w"],
                     "solve_ms": j["solve_s"] * 1e3, "symbolic_ms": j["symbolic_s"] * 1e3,
                     "sweeps": j["r"]["batched_iterations"], "iterations_mean": float(st["iterations"].mean()),
                     "finished_on_fp64": int((st["lu_fallbacks"] > 0).sum()),
                     "status_counts": np.bincount(st["status"], minlength=3).tolist()})
        log(f"{fams[j['fam']]} n={T.info['n']} B={j['cnt']} stream {j['w']}: {j['solve_s'] * 1e3:.0f} ms, "
            f"converged {conv}")
    per_fam = {}
    for r_ in rows:
        x = per_fam.setdefault(r_["family"], {"instances": 0, "converged": 0, "solve_ms": 0.0})
        x["instances"] += r_["instances"]
        x["converged"] += r_["converged"]
        x["solve_ms"] += r_["solve_ms"]
    log(f"mixed: {conv_total} converged in {solve_s:.2f} s on {S} streams (per-circuit solve time summed: "
        f"{sum(r_['solve_ms'] for r_ in rows) / 1e3:.2f} s)")
    out = {
        "metric": METRIC, "value": conv_total / solve_s, "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": solve_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == _abi.NR_PREC_FP32_REFINE else "f64",
        "data": "synthetic: node counts log-uniform in [200, 50000] (seeded), families i mod 3, 8 size buckets "
                "per family, Monte-Carlo variants per bucket",
        "config": {"workload": f"mixed{N}: uneven-size Monte-Carlo DC operating points, three families",
                   "instances": N, "circuits": len(rows), "precision": a.mixed_precision, "hipgraph": True,
                   "autotuned": not a.no_tune, "streams": S},
        "converged": conv_total, "symbolic_s": symb_s, "tune_s": tune_s,
        "value_with_symbolic": conv_total / (solve_s + symb_s),
        "circuit_solve_s_summed": sum(r_["solve_ms"] for r_ in rows) / 1e3,
        "families": per_fam, "circuits": rows,
        "roofline": roofline_either(kst),
        "roofline_multifrontal": roofline({k: v for k, v in kst.items()
                                           if k.startswith(("k_factor", "k_backsolve"))}, traffic=False),
        "kernels": kernel_table(kst),
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"], out["parity_sample"] = mixed_cpu_baseline(a, jobs, prec)
    print(json.dumps(out), flush=True)
    return 0


def mixed_cpu_baseline(a, jobs, prec):
    """The oracle on one instance of each circuit, each circuit on its own
    host thread (all at once), its FP32 path as the GPU ran; the workload's
    CPU time on C cores is sum(instances x seconds per instance) / C.  The
    sampled instances are also the parity sample."""
    from concurrent.futures import ThreadPoolExecutor
    from karam_amd import _abi
    from oracle import oracle as O
    threads = cpu_threads(a)
    oprec = O.PREC_FP32_REFINE if prec == _abi.NR_PREC_FP32_REFINE else O.PREC_FP64
    run = list(jobs)

    def one(j):
        To = O.Topology(j["nl"], ordering=j["T"].export(_abi.NR_EXPORT_ORDERING_RAW))
        t0 = time.perf_counter()
        Vo, so = O.solve_dc_batch(j["nl"], To, {k: v[:1] for k, v in j["f"].items()}, 1, threads=1, precision=oprec)
        dt = time.perf_counter() - t0
        log(f"oracle n={j['nodes']}: {dt:.1f} s")
        return dt, Vo[0], so

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, run))
    wall = time.perf_counter() - t0
    per = {}
    for j, (dt, Vo, so) in zip(run, res):
        per[id(j)] = dt
    total = sum(j["cnt"] * per[id(j)] for j in jobs)
    N = sum(j["cnt"] for j in jobs)
    parity = []
    for j, (dt, Vo, so) in zip(run, res):
        Vg = j["Vfirst"]
        err = float((np.abs(Vg - Vo) / (1e-9 * np.abs(Vo) + 1e-12)).max())
        parity.append({"family": ("rdmesh", "chain", "logic")[j["fam"]], "nodes": j["nodes"],
                       "oracle_s": dt, "err_over_tol": err, "bit_identical": bool(np.array_equal(Vg, Vo)),
                       "iterations_gpu": int(j["st"]["iterations"][0]), "iterations_oracle": int(so["iterations"][0]),
                       "status_equal": bool(int(j["st"]["status"][0]) == int(so["status"][0]))})
    base = {"value": N / (total / threads), "unit": "solves/s", "cores": threads, "kind": "port",
            "extrapolated": True,
            "sample": f"instance 0 of each of the {len(run)} circuits (oracle FP32 path, AUTO route, one host "
                      f"thread per circuit, {wall:.1f} s wall), each circuit's time scaled by its instance count",
            "host": host_info(threads)}
    log(f"mixed cpu baseline {base['value']:.2f} solves/s ({threads} threads, extrapolated)")
    return base, parity


def pipelined_legs(a, nl, n, s, make, mean_seed):
    """Throughput over a stream of Monte-Carlo batches: a.seeder_pipe_batches
    batches of 2048 test instances (seeds disjoint from the training set and
    the single-batch legs), dealt round robin to a.seeder_pipe_streams
    contexts (one HIP stream and host thread each, every context with its
    own topology and tuned launch classes), so that one batch's tail of slow
    instances overlaps the next batch's bulk.  The plain FP64 start and the
    MLP-seeded FP64 start (seed predicted inside the timed region, plain
    start as fallback) run the same batches the same way; value = counted
    (KCL-converged) solves over the wall time of all batches.  The level
    fan-out stays off: sweeps are captured from several threads."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from karam_amd import _abi, nr
    B, S, K = 2048, a.seeder_pipe_streams, a.seeder_pipe_batches
    data = [make(2_000_000 + 10_000 * k) for k in range(K)]
    ctxs = [nr.Context(0) for _ in range(S)]
    bts = []
    for c in ctxs:
        T = nr.Topology(nl, c)
        bt = nr.Batch(c, T, B)
        bt.set_fields(data[0][0])
        bt.autotune(reps=2)
        bt.set_fanout(False)
        bts.append(bt)
    lock = threading.Lock()

    def run(seeded, warm=False):
        conv = [0] * S

        def work(w):
            bt = bts[w]
            for k in (range(w, w + 1) if warm else range(w, K, S)):
                f, X = data[k % K]
                bt.set_fields(f)
                if seeded:
                    with lock:                 # one seeder, its workspace shared
                        V0 = s.predict_seed(X)
                    bt.set_seed(V0)
                    bt.solve_dc(seeded_start_step=18, polish=1, use_graph=1)
                else:
                    bt.clear_seed()
                    bt.solve_dc(use_graph=1)
                st = bt.stats()
                conv[w] += int(((st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)).sum())
            return w

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=S) as ex:
            list(ex.map(work, range(S)))
        wall = time.perf_counter() - t0
        return {"value": sum(conv) / wall, "unit": "solves/s", "wall_s": wall, "counted_instances": sum(conv)}

    run(False, warm=True)                      # untimed: every bucket's sweep graph of each leg
    plain = run(False)
    run(True, warm=True)
    seeded = run(True)
    log(f"pipelined over {K} batches on {S} streams: plain {plain['value']:.0f} solves/s, "
        f"seeded {seeded['value']:.0f} solves/s")
    for bt in bts:
        bt.close()
    return {"batches": K, "instances_per_batch": B, "streams": S, "plain_fp64": plain, "seeded_fp64": seeded,
            "speedup_seeded_over_plain": seeded["value"] / plain["value"] if plain["value"] else None}


def seeder_workload(a):
    """The learned seeder on the random-logic configuration (LOGIC20K,
    20,000 nodes): an MLP 512 -> 3 x 1024 (GELU) -> 20,000 node voltages,
    BF16 products with FP32 accumulation, trained on a.seeder_train FP64
    operating points that the GPU solver itself produces from seeded
    Monte-Carlo perturbations (per-device VTO sigma 30 mV and KP sigma 3 %,
    supply and input sources sigma 2 %).  The first batch of training points
    is solved from the plain start; later batches start at the last GMIN
    step from the mean of the first batch's solutions, without the plain
    fallback, and only their converged operating points (KCL <= 1e-9 A at
    the final gmin) are kept.

    Test set: 2048 instances from a disjoint seed range.  Reference answer:
    the plain FP64 solve followed by two more Newton iterations (polish), the
    root the 1e-9 A acceptance leaves a neighbourhood of (instances whose
    plain solve fails have none and are not compared); on this logic a
    node held only by off transistors and gmin (G ~ 3e-12 S) may sit volts
    from it at KCL 1e-9 A, so answers are compared with the root.  Legs: the
    plain FP64 solve (19-step GMIN start), seeded at the last GMIN step on
    FP64 and on the FP32 path with the plain start as fallback, and two
    seeds without learning at the same step (the training mean, the nominal
    operating point).  A solve counts when KCL-converged (<= 1e-9 A, the
    metric's definition); its distance from the root is reported as the
    largest node current G_ii |v - root| (G = the DC Jacobian's diagonal at
    the root, from the oracle) on a sample of instances, and as the voltage
    error over the 1e-9 rel + 1e-12 V tolerance.  value = counted seeded
    FP32-path solves/s."""
    from concurrent.futures import ThreadPoolExecutor
    from karam_amd import _abi, netlists, nr
    B = 2048
    nl = netlists.random_logic(**LOGIC20K)
    n = nl.num_nodes
    mc = dict(sigma_vto=0.03, sigma_kp=0.03, sigma_vsrc=0.02)
    ctx = nr.Context(0)
    T = nr.Topology(nl, ctx)
    gen = nr.Batch(ctx, T, B)
    tune = gen.autotune(reps=2)
    log(f"seeder workload: logic n={n}, autotune persistent={tune['persistent']} fanout={tune['fanout']} "
        f"few-instance map up to {tune['few_max']} ({tune['ms_few_before']:.2f} -> {tune['ms_few_after']:.2f} ms)")
    s = nr.Seeder(ctx, SEEDER_FEATURES, n, hidden=SEEDER_HIDDEN, n_hidden=SEEDER_LAYERS, init_seed=1)

    def make(seed):
        f = netlists.mc_fields(nl, B, seed=seed, **mc)
        return f, netlists.mc_features(nl, f, SEEDER_FEATURES)

    t0 = time.perf_counter()
    done, vsum, vcnt, conv_gen, gpu_s, batches = 0, np.zeros(n), 0, 0, 0.0, 0
    mean0 = None
    with ThreadPoolExecutor(1) as ex:
        nxt = ex.submit(make, 1_000_000)
        while done < a.seeder_train:
            f, X = nxt.result()
            m = min(B, a.seeder_train - done)
            if done + m < a.seeder_train:
                nxt = ex.submit(make, 1_000_000 + done + B)
            gen.set_fields(f)
            if mean0 is None:
                gen.clear_seed()
                r = gen.solve_dc(use_graph=1)
            else:
                gen.set_seed(mean0)
                r = gen.solve_dc(use_graph=1, seeded_start_step=18, fallback_plain=0)
            gpu_s += r["solve_ms"] * 1e-3
            st = gen.stats()
            ok = (st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)
            ok[m:] = False
            Vg = gen.voltages()[ok]
            s.record_solution(X[ok], Vg)
            vsum += Vg.sum(axis=0)
            vcnt += len(Vg)
            if mean0 is None:
                mean0 = np.tile(vsum / max(1, vcnt), (B, 1))
            conv_gen += int(ok.sum())
            done += m
            batches += 1
            log(f"training data: {done} solved, {conv_gen} kept, {time.perf_counter() - t0:.0f} s")
    gen.close()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = s.train(epochs=a.seeder_epochs, batch=1024, lr=1e-3)
    t_train = time.perf_counter() - t0
    dims = [SEEDER_FEATURES] + [SEEDER_HIDDEN] * SEEDER_LAYERS + [(n + 63) // 64 * 64]
    fwd = 2.0 * sum(x * y for x, y in zip(dims[:-1], dims[1:]))     # flops per sample
    n_train = int(s.history())
    train_tfs = 3.0 * fwd * (n_train // 1024 * 1024) * a.seeder_epochs / t_train / 1e12
    log(f"seeder: {n_train} solutions in {t_data:.0f} s, trained {a.seeder_epochs} epochs in {t_train:.1f} s "
        f"({train_tfs:.0f} TFLOP/s wall), loss {loss:.4f}")
    # the test instances and their root
    f, X = make(1000)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    bt.solve_dc(use_graph=1, polish=2)
    root = bt.voltages()
    sr = bt.stats()
    # the root exists where that solve converged; elsewhere no answer is
    # compared with it
    root_ok = (sr["status"] == _abi.NR_STATUS_CONVERGED) & (sr["kcl"] <= 1e-9)
    rate, ms, kst, st, r = timed_solves(bt, a.steps, use_graph=1)
    V64 = bt.voltages()
    ok64 = (st["status"] == _abi.NR_STATUS_CONVERGED) & (st["kcl"] <= 1e-9)
    it64 = st["iterations"]

    from oracle import oracle as O
    To = oracle_topology(nl, LOGIC20K, T)
    with_root = np.flatnonzero(root_ok)
    sample = [int(b) for b in with_root[::max(1, len(with_root) // 8)][:8]]
    G = {b: O.node_conductance(netlists.instance(nl, f, b), To, root[b]) for b in sample}

    def node_current(V, ok):
        """largest G_ii |v_i - root_i| over the sample, with where it is"""
        best = None
        for b in sample:
            if not ok[b]:
                continue
            c = G[b] * np.abs(V[b] - root[b])
            i = int(np.argmax(c))
            if best is None or c[i] > best["A"]:
                best = {"A": float(c[i]), "instance": b, "node": i + 1, "v": float(V[b][i]),
                        "v_root": float(root[b][i]), "G_S": float(G[b][i])}
        return best

    plain_counted = ok64
    mean_seed = np.tile(vsum / max(1, vcnt), (B, 1))
    nom = nr.Batch(ctx, T, 1)
    nom.solve_dc()
    nominal_seed = np.tile(nom.voltages()[0], (B, 1))
    nom.close()

    def leg(prec, set_seed, label):
        set_seed()
        bt.solve_dc(seeded_start_step=18, precision=prec, polish=1, use_graph=1)      # warm-up
        t_set = 0.0
        t1 = time.perf_counter()
        for k in range(a.steps):
            tp = time.perf_counter()
            set_seed()
            t_set += time.perf_counter() - tp
            rr = bt.solve_dc(seeded_start_step=18, precision=prec, polish=1, use_graph=1)
        total = time.perf_counter() - t1
        sx, V = bt.stats(), bt.voltages()
        ok = (sx["status"] == _abi.NR_STATUS_CONVERGED) & (sx["kcl"] <= 1e-9)
        counted = ok
        log(f"{label}: {total * 1e3 / a.steps:.0f} ms per solve of {B}, counted {int(counted.sum())}, "
            f"fallback {int((sx['seeded'] == 0).sum())}, sweeps {rr['batched_iterations']}")
        return {"value": int(counted.sum()) * a.steps / total, "unit": "solves/s", "ms_per_step": total * 1e3 / a.steps,
                "counted_instances": int(counted.sum()), "kcl_converged_instances": int(ok.sum()),
                "fallback_instances": int((sx["seeded"] == 0).sum()),
                "iterations": iter_stats(sx["iterations"][ok]) if ok.any() else None,
                "batched_sweeps_per_step": rr["batched_iterations"],
                "max_err_over_tol_vs_root": err_over_tol(V[ok & root_ok], root[ok & root_ok])
                if (ok & root_ok).any() else None,
                "compared_with_root": int((ok & root_ok).sum()),
                "max_node_current_vs_root_A": node_current(V, ok),
                "seed_ms_per_step": t_set * 1e3 / a.steps}

    legs = {"seeded_fp64": leg(_abi.NR_PREC_FP64, lambda: bt.set_seed_from(s, X), "seeded fp64"),
            "seeded_fp32_refine": leg(_abi.NR_PREC_FP32_REFINE, lambda: bt.set_seed_from(s, X), "seeded fp32"),
            "mean_seed_fp64": leg(_abi.NR_PREC_FP64, lambda: bt.set_seed(mean_seed), "mean seed fp64"),
            "nominal_seed_fp64": leg(_abi.NR_PREC_FP64, lambda: bt.set_seed(nominal_seed), "nominal seed fp64")}
    plain_rate = int(plain_counted.sum()) / (ms * 1e-3)
    for v in legs.values():
        v["speedup_vs_plain_fp64"] = v["value"] / plain_rate if plain_rate else None
    # seed prediction alone: the 2048 instances' forward pass (host copies included)
    t1 = time.perf_counter()
    for _ in range(5):
        s.predict_seed(X)
    pred_s = (time.perf_counter() - t1) / 5
    pipelined = pipelined_legs(a, nl, n, s, make, mean_seed) if a.seeder_pipe_batches > 0 else None
    out = {
        "metric": METRIC, "value": legs["seeded_fp32_refine"]["value"], "unit": "solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": 1, "ms_per_step": legs["seeded_fp32_refine"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64 (solve), bf16 (MLP)",
        "data": "synthetic: LOGIC20K random logic, Monte-Carlo per-device VTO/KP and supply sigma 2 %",
        "config": {"workload": f"learned seeder on logic{n}: MLP {SEEDER_FEATURES} -> {SEEDER_LAYERS} x "
                               f"{SEEDER_HIDDEN} -> {n}, trained on {n_train} FP64 operating points",
                   "instances": B, "nodes": n, "training_points": n_train, "epochs": a.seeder_epochs},
        "nr_iters_per_solve": {"plain": iter_stats(it64[ok64]) if ok64.any() else None,
                               **{k: v["iterations"] for k, v in legs.items()}},
        "plain_fp64": {"value": plain_rate, "ms_per_step": ms, "kcl_converged": int(ok64.sum()),
                       "counted_instances": int(plain_counted.sum()),
                       "max_err_over_tol_vs_root": err_over_tol(V64[ok64 & root_ok], root[ok64 & root_ok])
                       if (ok64 & root_ok).any() else None,
                       "compared_with_root": int((ok64 & root_ok).sum()),
                       "max_node_current_vs_root_A": node_current(V64, ok64), "node_current_sample": sample,
                       "root": "the plain FP64 solve with two Newton iterations after convergence (where it "
                               "converges: root_instances)", "root_instances": int(root_ok.sum())},
        **legs,
        "target_speedup": {"seeded_fp32_refine_over_plain_fp64": legs["seeded_fp32_refine"]["speedup_vs_plain_fp64"],
                           "seeded_fp64_over_plain_fp64": legs["seeded_fp64"]["speedup_vs_plain_fp64"], "goal": 2.0},
        "seeder": {"training_data_s": t_data, "training_data_gpu_solve_s": gpu_s, "training_batches": batches,
                   "training_points_kept": conv_gen, "train_s": t_train, "train_tflops_wall": train_tfs,
                   "final_loss_normalised": loss, "predict_ms_2048": pred_s * 1e3,
                   "predict_tflops_wall_incl_host_copies": fwd * B / pred_s / 1e12, "mfma_peak_tflops_bf16": 2500.0},
        "roofline": roofline(kst, traffic=False), "kernels": kernel_table(kst), "cpu_baseline": None,
        "autotune": {k: tune[k] for k in ("persistent", "fanout", "ms_graph", "ms_graph_fanout", "few_max",
                                          "ms_few_before", "ms_few_after")},
        "pipelined": pipelined,
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref-dense", action="store_true",
                    help="skip timing the dense pivoting LU the reference runs (about a minute of host time)")
    ap.add_argument("--no-seeder", action="store_true")
    ap.add_argument("--seeder-train", type=int, default=8192)
    ap.add_argument("--seeder-epochs", type=int, default=20)
    ap.add_argument("--seeder-pipe-batches", type=int, default=0,
                    help="seeder workload: Monte-Carlo batches of the pipelined legs (0: none)")
    ap.add_argument("--seeder-pipe-streams", type=int, default=2,
                    help="seeder workload: contexts (HIP streams) of the pipelined legs")
    ap.add_argument("--workload", choices=["rdmesh64", "logic", "mixed", "ac", "noise", "hightw", "seeder"],
                    default="rdmesh64")
    ap.add_argument("--ac-points", type=int, default=50)
    ap.add_argument("--mixed-instances", type=int, default=10000)
    ap.add_argument("--mixed-precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--mixed-streams", type=int, default=4, help="contexts (HIP streams) solving circuits at once")
    ap.add_argument("--logic-nodes", type=int, default=20000)
    ap.add_argument("--tran-check-steps", type=int, default=40,
                    help="logic workload: transient steps of the oracle's sample")
    ap.add_argument("--hightw-nodes", type=int, default=2000)
    ap.add_argument("--tran-steps", type=int, default=500)
    ap.add_argument("--tran-graph", type=int, default=1, help="replay transient sweeps from hipGraphs")
    ap.add_argument("--graph", type=int, default=1, help="mesh and logic workloads: replay DC sweeps from hipGraphs")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--seeder-probe", action="store_true", help="the seeder's products at the learned-seeder shape")
    ap.add_argument("--no-tune", action="store_true", help="default launch classes (no table, no tuning)")
    ap.add_argument("--retune", action="store_true", help="run the autotuner and write its table to --tune-table")
    ap.add_argument("--tune-table", default=os.path.join(ROOT, "karam_amd", "tuning", "rdmesh64_b4096.json"),
                    help="launch-class table replayed by the default benchmark")
    a = ap.parse_args()
    if a.copy_probe:
        return copy_probe()
    if a.seeder_probe:
        return seeder_probe()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: Monte-Carlo instances shard by contiguous ranges with no exchange, "
              "and multi-GPU runs are not part of this build (use --gpus 1)", file=sys.stderr)
        return 2

    if a.workload == "logic":
        return logic_workload(a)
    if a.workload == "mixed":
        return mixed_workload(a)
    if a.workload == "ac":
        return ac_workload(a)
    if a.workload == "noise":
        return noise_workload(a)
    if a.workload == "seeder":
        return seeder_workload(a)
    if a.workload == "hightw":
        return hightw_workload(a)

    from karam_amd import _abi, netlists, nr

    B = a.instances
    nl = netlists.rd_mesh(a.side, seed=1)
    fields = netlists.mc_fields(nl, B, seed=1000)
    ctx = nr.Context(0)
    t = time.perf_counter()
    T = nr.Topology(nl, ctx)
    symbolic_ms = (time.perf_counter() - t) * 1e3
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(fields)
    log(f"workload rdmesh{a.side} B={B} n={T.info['n']} nnz_f={T.info['nnz_f']} symbolic {symbolic_ms:.1f} ms")
    tune = {}
    precs = (("fp64", _abi.NR_PREC_FP64), ("fp32_refine", _abi.NR_PREC_FP32_REFINE))
    if a.retune:                      # untimed: launch shapes per front-order bucket, per precision
        table, persistent, fanout = {}, {}, {}
        for key, prec in precs:
            t0 = time.perf_counter()
            tu = bt.autotune(precision=prec, reps=3)
            tune[key] = {"s": time.perf_counter() - t0, "factor_solve_ms_before": tu["ms_before"],
                         "factor_solve_ms_after": tu["ms_after"],
                         "buckets": [[b["rmax"] if b["rmax"] < 1 << 20 else None, b["team"], b["block"], b["fronts"],
                                      "full" if b["full"] else "panel"]
                                     for b in tu["buckets"]]}
            table[key] = [[b["team"], b["block"], b["full"]] for b in tu["buckets"]]
            persistent[key] = int(tu["persistent"])
            fanout[key] = int(tu["fanout"])
            tune[key].update({"factor_solve_ms_levels": tu["ms_levels"],
                              "factor_solve_ms_persistent": tu["ms_persistent"], "persistent": tu["persistent"],
                              "factor_solve_ms_graph": tu["ms_graph"],
                              "factor_solve_ms_graph_fanout": tu["ms_graph_fanout"], "fanout": tu["fanout"]})
            log(f"autotune {key}: {tu['ms_before']:.2f} -> {tu['ms_after']:.2f} ms per factor + back substitution "
                f"(per-level {tu['ms_levels']:.2f}, whole-phase {tu['ms_persistent']:.2f}; graph "
                f"{tu['ms_graph']:.2f}, fanned out {tu['ms_graph_fanout']:.2f})")
        with open(a.tune_table, "w") as fo:
            json.dump({"workload": f"rdmesh{a.side}, {B} instances", "classes": table, "persistent": persistent,
                       "fanout": fanout}, fo, indent=1)
        tune["table_written"] = os.path.relpath(a.tune_table, ROOT)
    elif not a.no_tune and os.path.exists(a.tune_table):
        # the committed table from a tuning run on MI355X: every run (and every
        # profile of it) uses the same launch class per front-order bucket
        saved = json.load(open(a.tune_table))
        table, persistent, fanout = saved["classes"], saved.get("persistent", {}), saved.get("fanout", {})
        for key, prec in precs:
            for bucket, (team, block, full) in enumerate(table[key]):
                bt.set_launch_class(bucket, team, block, full=full, precision=prec)
            bt.set_persistent(bool(persistent.get(key, 0)), precision=prec)
            bt.set_fanout(bool(fanout.get(key, 0)), precision=prec)
            tune[key] = {"buckets": table[key], "persistent": int(persistent.get(key, 0)),
                         "fanout": int(fanout.get(key, 0))}
        tune["table"] = os.path.relpath(a.tune_table, ROOT)
        log(f"launch classes from {tune['table']}")
    for w in range(a.warmup):
        r = bt.solve_dc(use_graph=a.graph)
        log(f"warmup {w}: {r['solve_ms']:.1f} ms, converged {r['n_converged']}")

    rate, ms, kst, st, r = timed_solves(bt, a.steps, use_graph=a.graph)
    V = bt.voltages()
    evals = r["device_evals"]
    it = st["iterations"]
    out = {
        "metric": METRIC, "value": rate, "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded 64x64 R-diode mesh, per-instance Gaussian sigma 5% on every R and IS",
        "config": {"workload": f"rdmesh{a.side} Monte-Carlo DC operating point, plain FP64 batched GPU path",
                   "instances": B, "nodes": T.info["n"], "resistors": len(nl.res_ohms),
                   "diodes": len(nl.dio_is), "current_sources": len(nl.isrc_amps),
                   "nnz_a": T.info["nnz_a"], "nnz_f": T.info["nnz_f"], "treewidth": T.info["treewidth"],
                   "gmin_steps": 19, "symbolic_ms": symbolic_ms, "sweeps_from_hip_graphs": bool(a.graph),
                   "batched_sweeps_per_step": r["batched_iterations"]},
        "nr_iters_per_solve": {"plain": iter_stats(it), "plain_fp32_refine": None, "seeded": None,
                               "seeded_fp32_refine": None},
        "device_evals_per_s": evals * 1e3 / ms,
        "roofline": roofline(kst),
        "kernels": kernel_table(kst),
        "cpu_baseline": None,
        "autotune": tune or None,
    }
    # the FP32 path on the same instances, its error against the FP64 results
    bt.solve_dc(precision=_abi.NR_PREC_FP32_REFINE, use_graph=a.graph)                # warm-up
    rate32, ms32, kst32, st32, r32 = timed_solves(bt, a.steps, precision=_abi.NR_PREC_FP32_REFINE,
                                                  use_graph=a.graph)
    V32 = bt.voltages()
    out["nr_iters_per_solve"]["plain_fp32_refine"] = iter_stats(st32["iterations"])
    out["fp32_refine"] = {"value": rate32, "unit": "solves/s", "ms_per_step": ms32,
                          "speedup_vs_plain": rate32 / rate if rate else None,
                          "max_err_over_tol_vs_plain_fp64": err_over_tol(V32, V),
                          "max_kcl": float(st32["kcl"].max()),
                          "batched_sweeps_per_step": r32["batched_iterations"],
                          "roofline": roofline(kst32, traffic=False), "kernels": kernel_table(kst32)}
    seeds = None
    if not a.no_seeder:
        sd = seeded_path(a, ctx, nl, T, bt, fields, B, V)
        seeds = sd.pop("_seed_sample")
        out["seeded"] = sd
        for key in ("fp64", "fp32_refine", "mean_seed_fp32_refine", "nominal_seed_fp32_refine"):
            sd[key]["speedup_vs_plain"] = sd[key]["value"] / rate if rate else None
        out["nr_iters_per_solve"]["seeded"] = sd["fp64"].pop("iterations")
        out["nr_iters_per_solve"]["seeded_fp32_refine"] = sd["fp32_refine"].pop("iterations")
        out["target_speedup"] = {"seeded_fp32_refine_over_plain_fp64": sd["fp32_refine"]["speedup_vs_plain"],
                                 "goal": 2.0}
    if not a.no_cpu:
        threads = cpu_threads(a)
        base, parity = cpu_baseline(nl, fields, B, V, it, a.cpu_seconds, threads)
        parity.pop("_v")
        base["host"] = host_info(threads)
        out["cpu_baseline"] = base
        out["parity_sample"] = parity
        if not a.no_ref_dense:
            from oracle import oracle as O
            out["cpu_baseline_ref_dense"] = ref_dense_baseline(nl, fields, it, threads, O.Topology(nl))
        if seeds is not None:
            for key in ("fp64", "fp32_refine"):
                out["seeded"][key]["parity_sample"] = seeded_parity(nl, fields, seeds, out["seeded"][key])
        # every path against a tight FP64 root: the oracle's plain solve and two
        # more Newton iterations, on the first 8 instances
        from oracle import oracle as O
        k = min(8, B)
        root, _ = O.solve_dc_batch(nl, O.Topology(nl), {x: y[:k] for x, y in fields.items()}, k, threads=threads,
                                   polish=2)
        acc = {"instances": k, "root": "FP64 oracle, plain start, two Newton iterations after convergence",
               "plain_fp64": err_over_tol(V[:k], root), "plain_fp32_refine": err_over_tol(V32[:k], root)}
        if "seeded" in out:
            for key in ("fp64", "fp32_refine"):
                acc["seeded_" + key] = err_over_tol(out["seeded"][key]["_v_sample"][:k], root)
        out["accuracy_vs_root_over_tol"] = acc
        log(f"cpu baseline {base['value']:.1f} solves/s on {threads} threads; parity {parity}; vs root {acc}")
    if "seeded" in out:
        for key in ("fp64", "fp32_refine"):
            for k in [k for k in out["seeded"][key] if k.startswith("_")]:
                del out["seeded"][key][k]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
