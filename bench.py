"""Benchmark: sigma builds/s (FP64 vs mixed) and Davidson solves on one MI355X.

Workload (default, --workload sparse): the batched TPSCI sigma build of
fermicg_amd.workloads.tpsci_sparse -- 32 orbitals, 32 electrons, 8 clusters
of 4, 16 states per Fock sector, the 104,777 Fock configurations within 3
transferred electrons, 2,000,089 rank <= 1 configurations selected by
splitmix64 hash, 4 roots.  Synthetic seeded integrals and random start
vectors; nothing is read from disk.  --workload bst measures the Tucker/BST
sigma build (24 orbitals, 6 clusters, ranks 1..8, 464,878 variables).

One step = one sigma build H c over the whole space for all roots, inputs
resident in HBM (device pointers, no PCIe in the timed region).  The headline
`value` is the mixed-precision sigma of --precision (default for the sparse
workload: FP32 Hamiltonian blocks applied to the trial vectors rounded to
bf16, FP32 accumulation; FP32 blocks and vectors for the others); the FP64,
FP32, split-BF16, BF16 and FP32-with-bf16-vector rates, Davidson solves (FP64
vs mixed with FP64 residual refresh at restarts, with and without the
one-time plan build) and their |dE| against FP64 are reported beside it.

Prints ONE JSON line.  --gpus N with N > 1 exits: the library runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Threads the CPU baseline uses -- every core this job is allotted
    (OMP_NUM_THREADS when the host sets it, else all of nproc) -- and what
    the host reports about its processors."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = nproc
    env = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(int(env) if env and env.isdigit() else allowed, allowed))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": nproc, "affinity": allowed, "omp_num_threads": env,
                     "model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 30; 5 for pt2 and 1 for pt2_12c, whose steps "
                         "are whole PT2 energies)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 3; 1 for pt2 and pt2_12c)")
    ap.add_argument("--workload", default="sparse",
                    choices=["bst", "tpsci", "sparse", "serve", "pt2", "pt2_12c"])
    ap.add_argument("--requests", type=int, default=32, help="serve workload: requests")
    ap.add_argument("--rate", type=float, default=50.0, help="serve workload: arrivals per s")
    ap.add_argument("--slots", type=int, default=4, help="serve workload: concurrent solves")
    ap.add_argument("--order", default="fifo", choices=("fifo", "size"),
                    help="serve workload: which ready request a free slot takes")
    ap.add_argument("--precision", default=None,
                    choices=["fp64", "fp32", "bf16x2", "bf16", "fp32_bf16c"],
                    help="the headline sigma precision (default fp32_bf16c for the sparse "
                         "workload, whose one-row apply gathers c rounded to bf16; fp32 for "
                         "the others, whose multi-row boxes stage c in FP32)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the Davidson solves")
    ap.add_argument("--ss-vecs", type=int, default=12,
                    help="Davidson subspace vectors per root before a restart")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows (sparse) or bra blocks (bst) the CPU baseline times; 0: about "
                         "10 s of work (100,000 sparse rows, 2,000 BST blocks)")
    ap.add_argument("--clusters", type=int, default=6, help="bst workload: number of clusters")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (known bytes) to scale PMC counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library runs the sigma build on one "
              "device (bra Fock blocks would be sharded across GPUs with an all-gather of "
              "c before each sigma); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.precision is None:
        args.precision = "fp32_bf16c" if args.workload == "sparse" else "fp32"
    if args.steps is None:
        args.steps = {"pt2": 5, "pt2_12c": 1}.get(args.workload, 30)
    if args.warmup is None:
        args.warmup = 1 if args.workload in ("pt2", "pt2_12c") else 3
    if args.cpu_rows <= 0:
        args.cpu_rows = 100000 if args.workload == "sparse" else 2000
    if args.workload == "serve":
        return serve(args)
    if args.workload == "pt2":
        return pt2(args)
    if args.workload == "pt2_12c":
        return pt2_12c(args)
    import torch
    from fermicg_amd import _lib
    from fermicg_amd.marshal import flatten_ham, flatten_ops
    from fermicg_amd.tpsci import Engine, cluster_sizes_of

    t0 = time.time()
    if args.workload == "bst":
        from fermicg_amd.bst import cache_hamiltonian, ci_solve
        from fermicg_amd.workloads import bst_medium
        p, st = bst_medium(n_clusters=args.clusters)
    elif args.workload == "sparse":
        from fermicg_amd.workloads import tpsci_sparse
        p, st = tpsci_sparse()
    else:
        from fermicg_amd.workloads import tpsci_small
        p, st = tpsci_small(R=1)
    meta = dict(p.meta)
    log("workload", meta, "built in %.1fs" % (time.time() - t0))

    eng = Engine(0)
    t1 = time.time()
    precs = ("fp64", "fp32", "bf16x2", "bf16", "fp32_bf16c")
    if args.workload == "bst":
        pl = cache_hamiltonian(st, st, p.cluster_ops, p.clustered_ham, precisions=precs, engine=eng)
    else:
        pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, precisions=precs)
    ps = pl.stats()
    log("plan", ps, "total %.1fs" % (time.time() - t1))

    n = pl.nvar
    R = st.coeffs.shape[1]
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        # reads 1 GiB and writes 1 GiB per dispatch: the reference point that
        # FETCH_SIZE / WRITE_SIZE of the sigma kernels are scaled against
        xa = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        xb = torch.empty_like(xa)
        for _ in range(3):
            xb.copy_(xa)
        torch.cuda.synchronize()
        del xa, xb
    c_dev = torch.from_numpy(np.ascontiguousarray(st.coeffs.T)).to(dev)
    s_dev = torch.empty_like(c_dev)
    torch.cuda.synchronize()

    def run_steps(prec, k):
        for _ in range(k):
            pl.sigma_device(c_dev.data_ptr(), s_dev.data_ptr(), R, prec)

    rates = {}
    kernel_ms = {}
    for prec in precs:
        run_steps(prec, args.warmup)
        eng.ctx.synchronize()
        torch.cuda.synchronize()
        ta = time.perf_counter()
        eng.ctx.event(0)
        run_steps(prec, args.steps)
        eng.ctx.event(1)
        eng.ctx.synchronize()
        torch.cuda.synchronize()
        tb = time.perf_counter()
        ms_wall = (tb - ta) * 1e3 / args.steps
        ms_dev = eng.ctx.elapsed_ms(0, 1) / args.steps
        # apply-kernel-only duration (events around the apply launch)
        pl.sigma_device(c_dev.data_ptr(), s_dev.data_ptr(), R, prec)
        k_ms = pl.last_sigma_ms()
        rates[prec] = 1e3 / ms_wall
        kernel_ms[prec] = {"step_ms": ms_wall, "stream_ms_per_step": ms_dev, "apply_kernel_ms": k_ms,
                           "apply_variant": pl.apply_variant(prec, R)}
        log(prec, kernel_ms[prec])

    # parity on sampled bra blocks against the CPU oracle (and CPU baseline timing)
    parity, cpu = {}, None
    sig = {}
    for prec in precs:
        pl.sigma_device(c_dev.data_ptr(), s_dev.data_ptr(), R, prec)
        eng.ctx.synchronize()
        sig[prec] = s_dev.cpu().numpy().T.copy()
    if not args.no_cpu:
        from oracle.pyoracle import BstOracle, Oracle
        threads, host = host_cpu()
        fo = flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops))
        fh = flatten_ham(p.clustered_ham)
        rng = np.random.Generator(np.random.PCG64(12345))
        if args.workload == "bst":
            orc = BstOracle(fo, fh, st)
            nb = len(st.blocks)
            rows = np.sort(rng.choice(nb, min(args.cpu_rows, nb), replace=False))
            ta = time.perf_counter()
            ref, idx, (tbuild, tcon) = orc.sigma_rows(st.coeffs, rows, nthreads=threads, timings=True)
            wall = time.perf_counter() - ta
            frac = len(rows) / nb
            # reference ci_solve iterations read cached ops (cache_hamiltonian):
            # the per-iteration CPU cost is the contraction share of the work
            con_wall = wall * tcon / max(tbuild + tcon, 1e-30)
            cpu = {"value": frac / con_wall, "unit": "sigma builds/s", "cores": threads,
                   "kind": "port",
                   "sample": "C restatement of build_sigma! (src/tucker_inner.jl:165) on %d of %d "
                             "bra Tucker blocks (seeded), %d OpenMP threads, %.2f s wall; value "
                             "uses the contraction share (cached ops, as in ci_solve iterations); "
                             "with per-term op builds: %.4f sigma/s"
                             % (len(rows), nb, threads, wall, frac / wall)}
        elif args.workload == "sparse":
            orc = Oracle(fo, fh, st)
            idx = np.sort(rng.choice(n, min(args.cpu_rows, n), replace=False))
            ta = time.perf_counter()
            ref = orc.sigma_rows(st.coeffs, idx, nthreads=threads)
            wall = time.perf_counter() - ta
            cpu = {"value": len(idx) / n / wall, "unit": "sigma builds/s", "cores": threads,
                   "kind": "port",
                   "sample": "C restatement of tps_ci_matvec (src/tpsci_outer.jl:462) on %d of %d "
                             "bra configurations (seeded), %d OpenMP threads, %.2f s wall"
                             % (len(idx), n, threads, wall)}
        else:
            orc = Oracle(fo, fh, st)
            idx = np.arange(n)
            ta = time.perf_counter()
            ref = orc.sigma(st.coeffs, nthreads=threads)
            wall = time.perf_counter() - ta
            cpu = {"value": 1.0 / wall, "unit": "sigma builds/s", "cores": threads, "kind": "port",
                   "sample": "C restatement of tps_ci_matvec, full sigma, %d threads" % threads}
        cpu["host"] = host
        scale = np.abs(ref).max()
        for prec in precs:
            parity[prec] = float(np.abs(sig[prec][idx] - ref).max() / scale)
        log("parity", parity, "cpu", cpu)

    # Davidson solves: FP64 vs mixed (FP64 residual refresh at restarts)
    solves, dEs = {}, {}
    if not args.no_solve:
        for prec in dict.fromkeys(("fp64", "fp32", "bf16", "fp32_bf16c", args.precision)):
            ta = time.perf_counter()
            e, v, stt = pl.davidson(st.coeffs, conv_thresh=1e-8, max_iter=600,
                                    max_ss_vecs=args.ss_vecs,
                                    lindep_thresh=1e-12, precision=prec, refresh_k=0)
            dt = time.perf_counter() - ta
            solves[prec] = {"seconds": dt, "energies": [float(x) for x in e], **stt}
            log("solve", prec, solves[prec])
        for lo in solves:
            if lo != "fp64":
                dEs[lo] = float(np.max(np.abs(np.array(solves[lo]["energies"])
                                              - np.array(solves["fp64"]["energies"]))))
    dE = dEs.get(args.precision)

    # roofline of the dominant kernel (the sigma apply of the headline precision)
    prec = args.precision
    esz = {"fp64": 8, "fp32": 4, "bf16x2": 4, "bf16": 2, "fp32_bf16c": 4}[prec]
    # c as the apply reads it: bf16 for fp32_bf16c's one-row boxes (every
    # box of the sparse workload), FP32 for multi-row boxes
    csz = 8 if prec == "fp64" else 2 if (prec == "fp32_bf16c" and args.workload == "sparse") else 4
    bytes_alg = ps["block_elems"] * esz + 4.0 * ps["panel_cols"] + n * R * (csz + 8)
    k_ms = kernel_ms[prec]["apply_kernel_ms"]
    achieved = bytes_alg / (k_ms * 1e-3) / 1e9
    # HBM bytes per launch measured by the rocprofv3 --pmc passes of this
    # workload (committed under profiles/, copy-calibrated); null without one
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "%s_apply_traffic.json" % meta["workload"])
    if os.path.exists(tpath):
        with open(tpath) as fh_:
            tj = json.load(fh_)
        if prec in tj.get("bytes_per_launch", {}):
            traffic = tj["bytes_per_launch"][prec]
            traffic_src = os.path.relpath(tpath, ROOT)
    out = {
        "metric": "sigma builds/s (mixed-precision sigma vs FP64 GPU at |dE|<=1 uEh)",
        "value": rates[prec],
        "unit": "sigma builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / rates[prec],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x2": "bf16x2/f32", "bf16": "bf16/f32",
                  "fp32_bf16c": "f32 (c as bf16)"}[prec],
        "data": {"bst": "synthetic (seeded integrals, random Tucker cores and factors)",
                 "sparse": "synthetic (seeded integrals, hash-selected rank-1 configurations, "
                           "random start vectors)",
                 "tpsci": "synthetic (seeded integrals, random start vector)"}[args.workload],
        "config": {"workload": meta["workload"], "n_var": int(n), "roots": int(R),
                   "n_fock": meta.get("n_fock"), "n_blocks": meta.get("n_blocks"),
                   "block_pairs": ps["pairs"], "terms_x_pairs": ps["tasks"],
                   "hamiltonian_block_elems": ps["block_elems"],
                   "plan_build_s": ps["build_s"], "parallelism": "single GPU"},
        "sigma_per_s": rates,
        "kernel_ms": kernel_ms,
        "solves": solves,
        "solves_per_s": {k: 1.0 / v["seconds"] for k, v in solves.items()},
        # a solve from scratch also builds the plan once (the reference's
        # ci_solve includes cache_hamiltonian, src/tucker_outer.jl:131-137)
        "solves_per_s_with_plan": {k: 1.0 / (v["seconds"] + ps["build_s"])
                                   for k, v in solves.items()},
        "abs_dE_mixed_vs_fp64": dE,
        "abs_dE_vs_fp64": dEs,
        "parity_rel_vs_cpu_oracle": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "bytes_algorithmic": bytes_alg,
                     "kernel": "k_apply + k_apply_rows<%s> (block-sparse H c)" % prec},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def pt2(args):
    """PT2 of a 60,025-configuration TPSCI state (fermicg_amd.workloads.
    pt2_medium): one step = compute_pt2_energy at thresh_foi = 1e-9 (target
    space, screened pair plan, sigma, H0 diagonal, compensated sum), inputs
    on the host as the reference's call takes them.  CPU baseline: the
    oracle's restatement of compute_pt2_energy on the same state."""
    from fermicg_amd.cluster_ops import OP_NAMES
    from fermicg_amd.marshal import flatten_ham, flatten_ops
    from fermicg_amd.tpsci import (Engine, cluster_sizes_of, compute_expectation_value,
                                   compute_pt2_energy)
    from fermicg_amd.workloads import pt2_medium
    thresh = 1e-9
    t0 = time.time()
    p, st = pt2_medium()
    meta = dict(p.meta)
    log("workload", meta, "built in %.1fs" % (time.time() - t0))
    eng = Engine(0)
    E0 = compute_expectation_value(st, p.cluster_ops, eng.local_operator(p.clustered_ham, "Hcmf"),
                                   engine=eng)
    times, e2, stats = [], None, None
    for k in range(args.warmup + args.steps):
        ta = time.perf_counter()
        e2 = compute_pt2_energy(st, p.cluster_ops, p.clustered_ham, E0=E0, thresh_foi=thresh,
                                verbose=0, engine=eng)
        dt = time.perf_counter() - ta
        stats = compute_pt2_energy.last_stats
        if k >= args.warmup:
            times.append(dt)
        log("pt2 step %d: %.3f s" % (k, dt), stats)
    sec = float(np.mean(times))
    cpu, parity = None, None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle
        threads, host = host_cpu()
        orc = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                     flatten_ham(p.clustered_ham), st)
        ta = time.perf_counter()
        e2_ref, n_jobs = orc.pt2(st.coeffs, E0, h0_op=OP_NAMES.index("Hcmf"), thresh=thresh,
                                 nthreads=threads)
        wall = time.perf_counter() - ta
        parity = float(np.abs(e2 - e2_ref).max() / np.abs(e2_ref).max())
        cpu = {"value": 1.0 / wall, "unit": "PT2 energies/s", "cores": threads, "kind": "port",
               "sample": "C restatement of compute_pt2_energy (src/tpsci_pt2_energy.jl:10) on the "
                         "whole workload (%d target Fock jobs), %d OpenMP threads, %.1f s wall"
                         % (n_jobs, threads, wall)}
        cpu["host"] = host
        log("parity", parity, "cpu", cpu)
    bf, bs = stats["build_flops"], stats["build_seconds"]
    out = {
        "metric": "PT2 energies/s (compute_pt2_energy, thresh_foi 1e-9, FP64)",
        "value": 1.0 / sec, "unit": "PT2 energies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded integrals, seeded orthonormal state vectors)",
        "config": {"workload": meta["workload"], "n_var": meta["n_cfg"], "roots": int(st.R),
                   "n_fock": meta["n_fock"], "n_target_fock": stats["n_target_fock"],
                   "n_first_order": stats["n_first_order"], "block_pairs": stats["block_pairs"],
                   "hamiltonian_block_elems": stats["block_elems"], "thresh_foi": thresh,
                   "parallelism": "single GPU"},
        "e2": [float(x) for x in e2], "E0": [float(x) for x in E0],
        "parity_rel_vs_cpu_oracle": parity,
        "roofline": {"bound": "mfma", "achieved": bf / bs / 1e12, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": bf / bs / 1e12 / 78.6, "traffic": None,
                     "kernel": "k_build_blocks of the first-order pair plan (FP64 VALU, whose peak "
                               "equals the FP64 MFMA peak on MI355X); time includes the host join",
                     "flops_algorithmic": bf, "seconds": bs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def pt2_12c(args):
    """Second-order energy of the 48-orbital, 12-cluster, M=32 workload
    (fermicg_amd.workloads.pt2_12c_M32: 9,020,329 rank <= 1 variational
    configurations over the 26,533 Fock configurations within 2 transferred
    electrons) at thresh_foi = 1e-5, over the sparse first-order space
    (fcg_tpsci_pt2_sparse).  State: the CMF reference plus its regularised
    first-order amplitudes above 0.1 (cmf_first_order_state, select=0.1:
    1,923 nonzero coefficients), E0 = <c|Hcmf|c>.  One step = one
    compute_pt2_energy call (items, screening, contraction count and fill,
    sort, per-configuration sums), inputs on the host as the reference's call
    takes them.  CPU baseline: the oracle's restatement over the
    configurations reached (orc_tpsci_pt2_sparse) on the same state cut to
    its largest coefficients, in kept contributions per second."""
    import torch
    from fermicg_amd import _lib  # noqa: F401  (fails loudly when the library is missing)
    from fermicg_amd.tpsci import Engine, compute_pt2_energy
    from fermicg_amd.workloads import cmf_first_order_state, pt2_12c_M32
    thresh, select = 1e-5, 0.1
    t0 = time.time()
    p, st = pt2_12c_M32(M=32)
    meta = dict(p.meta)
    log("workload", meta, "built in %.1fs" % (time.time() - t0))
    eng = Engine(0)
    t1 = time.time()
    st.set_vector(cmf_first_order_state(p, st, eng, select=select))
    h0 = eng.local_operator(p.clustered_ham, "Hcmf")
    E0 = np.einsum("ir,ir->r", st.coeffs,
                   eng.tpsci_plan(st, p.cluster_ops, h0, matrix_free=True).sigma(st.coeffs))
    eng.drop_plans()
    nnz = int(np.count_nonzero(st.coeffs))
    log("state: %d nonzero, E0 %s, %.1fs" % (nnz, E0, time.time() - t1))
    def run(prec):
        times, e2, stats = [], None, None
        for k in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            e2 = compute_pt2_energy(st, p.cluster_ops, p.clustered_ham, E0=E0, thresh_foi=thresh,
                                    verbose=0, method="sparse", engine=eng, precision=prec)
            dt = time.perf_counter() - ta
            stats = dict(compute_pt2_energy.last_stats)
            if k >= args.warmup:
                times.append(dt)
            log("pt2 %s step %d: %.3f s" % (prec, k, dt), stats)
        return float(np.mean(times)), e2, stats

    sec64, e2_64, stats64 = run("fp64")
    sec, e2, stats = run("fp32")   # the headline: the mixed contraction
    E = stats["n_contributions"]
    dE = float(np.abs(np.asarray(e2) - np.asarray(e2_64)).max())
    cpu = None
    if not args.no_cpu:
        from fermicg_amd.marshal import flatten_ham, flatten_ops
        from fermicg_amd.tpsci import cluster_sizes_of
        from oracle.pyoracle import Oracle
        threads, host = host_cpu()
        keep = 12
        c = st.coeffs.copy()
        order = np.argsort(-np.abs(c[:, 0]))
        c[order[keep:]] = 0.0
        orc = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                     flatten_ham(p.clustered_ham), st)
        ta = time.perf_counter()
        _, n_jobs, n_x = orc.pt2_sparse(c, E0, thresh=thresh, nthreads=threads)
        wall = time.perf_counter() - ta
        nc = orc.last_contributions
        cpu = {"value": nc / wall, "unit": "kept contributions/s", "cores": threads,
               "kind": "port",
               "sample": "orc_tpsci_pt2_sparse (C restatement of compute_pt2_energy, "
                         "src/tpsci_pt2_energy.jl:10, over the configurations reached) on the "
                         "state cut to its %d largest coefficients: %d contributions, %d "
                         "first-order configurations, %d target Fock jobs, %d OpenMP threads, "
                         "%.1f s wall" % (keep, nc, n_x, n_jobs, threads, wall),
               "host": host}
        log("cpu", cpu)
    write_bytes = E * (16 + 8 * int(st.R))   # key + value per kept contribution
    cs = stats["contraction_seconds"]
    out = {
        "metric": "PT2 kept contributions/s (compute_pt2_energy over the sparse first-order "
                  "space, thresh_foi 1e-5; mixed FP32 contraction with FP64 sums, FP64 beside)",
        "value": E / sec, "unit": "contributions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "fp64": {"contributions_per_s": stats64["n_contributions"] / sec64,
                 "ms_per_step": 1e3 * sec64, "e2": [float(x) for x in e2_64],
                 "contraction_seconds": stats64["contraction_seconds"]},
        "mixed_over_fp64": sec64 / sec, "abs_dE_pt2_vs_fp64": dE,
        "rel_dE_pt2_vs_fp64": dE / float(np.abs(e2_64).max()),
        "data": "synthetic (seeded integrals; CMF reference plus selected first-order amplitudes)",
        "config": {"workload": meta["workload"], "n_var": meta["n_cfg"], "n_fock": meta["n_fock"],
                   "nonzero_coefficients": nnz, "roots": int(st.R), "thresh_foi": thresh,
                   "select": select, "n_items": stats["n_items"],
                   "n_kets_kept": stats["n_kets_kept"], "n_contributions": E,
                   "n_configs": stats["n_configs"], "n_first_order": stats["n_first_order"],
                   "n_chunks": stats["n_chunks"], "parallelism": "single GPU"},
        "pt2_energies_per_s": 1.0 / sec,
        "e2": [float(x) for x in e2], "E0": [float(x) for x in E0],
        "phase_seconds": {"contraction": cs, "sort_and_sum": stats["sort_seconds"],
                          "total": stats["seconds"]},
        "roofline": {"bound": "hbm", "achieved": write_bytes / cs / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": write_bytes / cs / 1e9 / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": "k_pt2_emit<float> (mixed), count and fill passes together: "
                               "algorithmic bytes = kept contributions x (16-byte key + 8 R-byte "
                               "value) written",
                     "bytes_algorithmic": write_bytes, "seconds": cs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def serve(args):
    """Served mix: a Poisson stream of independent Davidson solves of
    2-8 clusters, M in {4, 8, 16}, 1-4 roots (fermicg_amd.serve), FP64 then
    mixed precision over the same stream."""
    from fermicg_amd.serve import SolveServer, request_stream, summarize
    reqs = request_stream(args.requests, rate=args.rate, seed=0)
    # producer processes start before anything in this process touches the GPU
    srv = SolveServer(slots=args.slots, producers=min(16, os.cpu_count() or 1), hbm_budget=200e9,
                      conv_thresh=1e-7, order=args.order).start()
    runs, res_all = {}, {}
    for prec in ("fp64", "fp32"):
        srv.set_precision(prec)
        res = srv.run(reqs, log=log)
        res_all[prec] = res
        runs[prec] = dict(summarize(res), peak_hbm_bytes=srv.peak_bytes)
        log(prec, runs[prec])
    srv.close()
    dE = max((max(abs(a - b) for a, b in zip(r64["energies"], r32["energies"]))
              for r64, r32 in zip(res_all["fp64"], res_all["fp32"])
              if r64["energies"] and r32["energies"]), default=None)
    kinds = {}
    for r in res_all["fp64"]:
        kinds[r["kind"]] = kinds.get(r["kind"], 0) + 1
    out = {
        "metric": "solves/s of the served mix (mixed FP32 vs FP64 GPU at |dE|<=1 uEh); p50/p99 latency",
        "value": runs["fp32"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.requests, "warmup": 0,
        "ms_per_step": 1e3 / runs["fp32"]["solves_per_s"] if runs["fp32"]["solves_per_s"] else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded integrals per request, Poisson arrivals)",
        "config": {"workload": "serve_mix_poisson%d" % int(args.rate), "requests": args.requests,
                   "rate_per_s": args.rate, "slots": args.slots, "order": args.order, "spaces": kinds,
                   "parallelism": "single GPU"},
        "serve": runs, "abs_dE_mixed_vs_fp64": dE,
        "latency_s": {p: {"p50": runs[p]["p50_latency_s"], "p99": runs[p]["p99_latency_s"]}
                      for p in runs},
        "requests": [{k: r[k] for k in ("rid", "n_clusters", "M", "R", "kind", "n_var", "arrival",
                                        "ready", "start", "latency", "plan_s", "solve_s",
                                        "footprint", "iterations")}
                     for r in res_all["fp32"]],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
