#!/usr/bin/env python3
"""Time-to-1e-6 of the per-tick structural solve on one MI355X.

Workload: the 128^3 single island (60 % occupied box union on an anchored
z=0 slab, kappa lognormal sigma 0.5, FP32 inputs).  One step = solve
A phi = b to ||b - A phi|| / ||b|| <= 1e-6 (FP32 MG-preconditioned PCG inside
FP64 iterative refinement, cold start), then the phase-field history update
H = max(H, psi_e) and the failure scan, all on device arrays already resident
in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pfsf-core_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "time-to-1e-6 residual at 128^3 + batched islands/s vs plain, % of HBM peak"

# Algorithmic bytes per ACTIVE voxel of each profiled finest-level kernel (each
# array element counted once per launch, neighbour reuse free; kappa is the
# anchor-coded array, so no anchor-mass array is read):
#   0 ks_spmv   read z, p_in, kappa 12; write p_out, q 8                     = 20
#   1 ks_down of an odd iteration (moves the FP64 iterate by the directions
#               of both iterations of the pair): read r_in, q, kappa, diag 16,
#               p, p_prev 8, x (f64) 8; write x (f64) 8, r_out 4,
#               coarse rhs 0.5                                               = 44.5
#   5 ks_down of an even iteration (iterate update left pending): read r_in,
#               q, kappa, diag 16; write r_out 4, coarse rhs 0.5             = 20.5
#   3 ks_up     read r, kappa, diag 12, coarse correction 0.5; write z 4     = 16.5
ALG_BYTES = {0: 20.0, 1: 44.5, 5: 20.5, 3: 16.5}
KERNEL_NAMES = {0: "ks_spmv", 1: "ks_down(odd)", 5: "ks_down(even)", 3: "ks_up(finest)"}
# finest-level bytes of an average iteration (the two ks_down kinds alternate)
ITER_ALG_BYTES = ALG_BYTES[0] + 0.5 * (ALG_BYTES[1] + ALG_BYTES[5]) + ALG_BYTES[3]
# HBM bytes per launch of these kernels on this workload, from the committed
# FETCH_SIZE / WRITE_SIZE counter summary (copy-kernel calibrated).
TRAFFIC_FILE = ROOT / "profiles" / "solve_128_hbm_traffic.md"


def measured_traffic():
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("bench-traffic:"):
                return {k: int(v) for k, v in (t.split("=") for t in line.split()[1:])}
    except OSError:
        pass
    return {}
# Whole-solve algorithmic bytes per active voxel: a fused PCG iteration
# (59 B) plus a 2-level V-cycle application (67 B) per iteration, and the
# initial residual (17 B).
SOLVE_ITER_BYTES = 126
SOLVE_INIT_BYTES = 17


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the all-core CPU legs: OMP_NUM_THREADS when set (this
    process's share of the host), else every CPU the OS reports."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return n if n > 0 else (os.cpu_count() or 1)


def host_info():
    """nproc and the CPU model (from /proc/cpuinfo, as lscpu reports it)."""
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(isl, sample_iters=40):
    """The oracle's FP64 Jacobi-PCG (C) to 1e-6 on the same island.

    All host threads: the whole solve, timed (its per-voxel operator loop in
    parallel; the dot products serial, so the iterates and the iteration
    count are the same bits on any thread count).  One thread: a bounded
    sample of `sample_iters` iterations timed and scaled by the iteration
    count the all-thread solve measured in this run."""
    from oracle import oracle as O

    used = O.set_threads(host_threads())
    t = time.perf_counter()
    _, iters, rel = O.pcg(isl, tol=1e-6, maxit=20000)
    full = time.perf_counter() - t
    O.set_threads(1)
    t = time.perf_counter()
    _, it1, _ = O.pcg(isl, tol=0.0, maxit=sample_iters)
    dt1 = time.perf_counter() - t
    per_iter1 = dt1 / max(it1, 1)
    hi = host_info()
    all_cores = {"value": round(full * 1e3, 1), "unit": "ms", "cores": used, "kind": "port",
                 "sample": f"the whole FP64 Jacobi-PCG solve of the C oracle on the 128^3 "
                           f"island to 1e-6 ({iters} iterations, true residual {rel:.3e}) on "
                           f"{used} threads", **hi}
    one_core = {"value": round(per_iter1 * iters * 1e3, 1), "unit": "ms", "cores": 1,
                "kind": "port",
                "sample": f"{it1} iterations of the same solver on one thread "
                          f"({dt1:.1f} s, {per_iter1 * 1e3:.1f} ms/iter) times the {iters} "
                          f"iterations the solve above took", **hi}
    return all_cores, one_core


def batch_bench(torch, D, W, steps=3, plain_stride=16, coarse_sweeps=0):
    """Batched solve of the 4,096-island batch against plain per-island solves.

    Islands/s = 4096 / (batched solve time); the plain rate comes from solving
    every plain_stride-th island (by size) one at a time with its own solver,
    the summed solve times scaled by plain_stride."""
    import numpy as np

    t0 = time.perf_counter()
    isls = W.island_batch(4096)
    gen_s = time.perf_counter() - t0
    log(f"batch: generated {len(isls)} islands in {gen_s:.1f} s")
    dev = torch.device("cuda:0")
    vt = [torch.from_numpy(i.vtype).to(dev) for i in isls]
    kp = [torch.from_numpy(i.kappa).to(dev) for i in isls]
    bs = [torch.from_numpy(i.source).to(dev) for i in isls]
    phis = [torch.empty(i.vtype.shape, dtype=torch.float32, device=dev) for i in isls]
    n_act = int(sum(i.n_active for i in isls))
    from pfsf_amd import capi

    def run(shape, precision, label):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bat = D.BatchSolver(list(zip(vt, kp)), tol=1e-6, max_iters=2000,
                            coarse_sweeps=coarse_sweeps, shape=shape, precision=precision)
        setup = time.perf_counter() - t0
        log(f"batch {label}: hierarchies built in {setup:.1f} s")
        classes = bat.classes()
        tuning = bat.tuning()
        bat.solve(bs, None, phis)  # warm-up (graph capture)
        ts, inf = [], []
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            i, per = bat.solve(bs, None, phis, per_island=True)
            inf.append(i)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            log(f"batch {label}: solve {ts[-1] * 1e3:.1f} ms, max iterations {i.max_iterations}")
        bat.close()
        return setup, (classes, tuning), ts, inf, per

    # the headline batch: tile shapes picked per island-size class by the tuner
    setup_s, (classes, tuning), times, infos, per = run(capi.BATCH_AUTOTUNE, 0, "tuned")
    # beside it: every island on the default shape, and BF16 coarse kappa
    # on the tuned shapes
    _, _, t_def, i_def, _ = run(0, 0, "default shape")
    s_bf, c_bf, t_bf, i_bf, _ = run(capi.BATCH_AUTOTUNE, 1, "bf16 coarse")
    t_batch = float(np.median(times))
    info = infos[-1]
    dev_ms = float(np.median([i.solve_ms for i in infos]))
    # whole-solve algorithmic bytes, the survey's per-voxel count per island:
    # (fused PCG iteration + 2-level V-cycle) per iteration plus the initial
    # residual, summed over islands with each island's own iteration count
    alg = float(sum((SOLVE_ITER_BYTES * p.iterations + SOLVE_INIT_BYTES) * i.n_active
                    for p, i in zip(per, isls)))
    roof = {"bound": "hbm", "scope": "whole batched solve",
            "achieved": round(alg / (dev_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(alg / (dev_ms * 1e-3) / HBM_PEAK, 4),
            "alg_bytes": round(alg), "device_ms": round(dev_ms, 3),
            "per_active_voxel": f"{SOLVE_ITER_BYTES} B x the island's PCG iterations + "
                                f"{SOLVE_INIT_BYTES} B, summed over the islands",
            "traffic": None}
    # plain: one solver per island, islands one after another
    order = np.argsort([i.n_active for i in isls], kind="stable")
    sample = order[::plain_stride]
    t_plain, it_plain = 0.0, 0
    for k in sample:
        s = D.Solver(vt[k], kp[k], tol=1e-6, max_iters=2000, coarse_sweeps=coarse_sweeps)
        x = torch.zeros(isls[k].vtype.shape, dtype=torch.float64, device=dev)
        s.solve(bs[k], x)  # warm-up (graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x.zero_()
        inf = s.solve(bs[k], x)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        it_plain += inf.iterations
        s.close()
    log(f"plain: {len(sample)} islands in {t_plain * 1e3:.1f} ms")
    t_plain_all = t_plain * plain_stride
    return {
        "islands": len(isls), "active_voxels": n_act, "generate_s": round(gen_s, 2),
        "setup_s": round(setup_s, 2), "batch_ms": round(t_batch * 1e3, 2),
        "islands_per_s": round(len(isls) / t_batch, 1),
        "plain_islands_per_s": round(len(isls) / t_plain_all, 1),
        "speedup_vs_plain": round(t_plain_all / t_batch, 2),
        "voxels_per_s": round(n_act / t_batch, 1),
        "converged": info.converged, "max_iterations": info.max_iterations,
        "mean_iterations": round(info.total_iterations / info.islands, 1),
        "graph_sweeps": info.sweeps, "max_rel_residual": info.max_rel_residual,
        "device_ms": round(info.solve_ms, 2), "roofline": roof,
        "shapes": {"tuned_setup_s": round(setup_s, 2), "classes": classes, "decision": tuning,
                   "candidates": "fine column depth x coarse tile depth; ms = the class "
                                 "sample's 16-iteration solve with that shape",
                   "default_8x2_batch_ms": round(float(np.median(t_def)) * 1e3, 2),
                   "default_8x2_max_iterations": i_def[-1].max_iterations},
        "bf16_coarse": {"batch_ms": round(float(np.median(t_bf)) * 1e3, 2),
                        "islands_per_s": round(len(isls) / float(np.median(t_bf)), 1),
                        "speedup_vs_plain_fp32": round(t_plain_all / float(np.median(t_bf)), 2),
                        "converged": i_bf[-1].converged,
                        "max_iterations": i_bf[-1].max_iterations,
                        "max_rel_residual": i_bf[-1].max_rel_residual,
                        "setup_s": round(s_bf, 2),
                        "tuned": [[c["tz_fine"], c["tz_coarse"]] for c in c_bf[0]]},
        "plain_sample": f"{len(sample)} islands (every {plain_stride}th by voxel count), "
                        f"{t_plain * 1e3:.1f} ms summed, {it_plain} PCG iterations, scaled "
                        f"x{plain_stride}",
        "tol": 1e-6, "workload": "4096 grown box-union islands, voxel counts log-uniform "
                                 "64..262144, 1-8 anchors each, cold start"}


def mixed_precision_bench(torch, D, W, steps=3, coarse_sweeps=0):
    """The 256^3 island (kappa lognormal sigma 1.5, ~70 % occupied): time to
    1e-6 with FP32 coarse levels and with BF16 coarse kappa; phi against a
    1e-11 solve of the same system, and the failure codes of both."""
    import numpy as np

    isl = W.workload(3)
    dev = torch.device("cuda:0")
    vt = torch.from_numpy(isl.vtype).to(dev)
    kp = torch.from_numpy(isl.kappa).to(dev)
    b = torch.from_numpy(isl.source).to(dev)
    rc = torch.from_numpy(isl.rcomp).to(dev)
    rt = torch.from_numpy(isl.rtens).to(dev)
    ref = D.Solver(vt, kp, tol=1e-11, max_iters=4000)
    xr = torch.zeros(isl.vtype.shape, dtype=torch.float64, device=dev)
    pr = torch.empty(isl.vtype.shape, dtype=torch.float32, device=dev)
    iref = ref.solve(b, xr, pr)
    ref.close()
    code_ref = torch.empty_like(vt)
    D.failure_scan(vt, kp, pr, rc, rt, code_ref)
    out = {"grid": [isl.lx, isl.ly, isl.lz], "active_voxels": isl.n_active, "tol": 1e-6,
           "reference": f"same system solved to {iref.rel_residual:.2e} (FP64 residual)"}
    for prec, name in ((0, "fp32"), (1, "bf16_coarse")):
        s = D.Solver(vt, kp, precision=prec, tol=1e-6, max_iters=2000,
                     coarse_sweeps=coarse_sweeps)
        x = torch.zeros_like(xr)
        ph = torch.empty_like(pr)
        s.solve(b, x, ph)  # warm-up (graph capture)
        ms, info = [], None
        for _ in range(steps):
            x.zero_()
            info = s.solve(b, x, ph)
            ms.append(info.solve_ms)
        # the finest-level kernels at this size (launch events, as for 128^3)
        kern = {k: s.profile(k, reps=20) for k in KERNEL_NAMES}
        s.close()
        code = torch.empty_like(vt)
        D.failure_scan(vt, kp, ph, rc, rt, code)
        err = float(torch.linalg.norm((ph - pr).double()) / torch.linalg.norm(pr.double()))
        dn = kern[1]
        out[name] = {"solve_ms": round(float(np.median(ms)), 3), "iterations": info.iterations,
                     "passes": info.refinements, "levels": info.levels,
                     "rel_residual": info.rel_residual, "phi_rel_l2_vs_reference": err,
                     "failure_codes_differing": int((code != code_ref).sum()),
                     "kernel_ms": {KERNEL_NAMES[k]: round(v, 5) for k, v in kern.items()},
                     "ks_down_alg_GBps": round(ALG_BYTES[1] * isl.n_active / (dn * 1e-3) / 1e9, 1),
                     "ks_down_frac": round(ALG_BYTES[1] * isl.n_active / (dn * 1e-3) / HBM_PEAK,
                                           4)}
        log(f"256^3 {name}: {out[name]}")
    return out


def adjoint_bench(torch, W, steps=2):
    """The 96^3 island: 50-step load path, loss (forward) and dL/dkappa
    (forward + checkpointed adjoint), solves to 1e-6; then the same with
    damage coupling the steps (each state's operator rebuilt from its
    degraded kappa)."""
    import numpy as np

    from pfsf_amd import adjoint as A

    isl = W.workload(4)
    dev = torch.device("cuda:0")
    vt, kp, b, rc = (torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                     for a in (isl.vtype, isl.kappa, isl.source, isl.rcomp))
    A.load_path_loss(vt, kp, b, rc)  # warm-up
    fw = [A.load_path_loss(vt, kp, b, rc) for _ in range(steps)]
    gr = [A.load_path_grad(vt, kp, b, rc)[0] for _ in range(steps)]
    f, g = fw[-1], gr[-1]
    out = {"grid": [isl.lx, isl.ly, isl.lz], "active_voxels": isl.n_active, "load_steps": 50,
           "checkpoint_every": 10, "tol": 1e-6, "loss": f.loss,
           "forward_ms": round(float(np.median([i.ms for i in fw])), 2),
           "gradient_ms": round(float(np.median([i.ms for i in gr])), 2),
           "gradient_solves": {"forward": g.forward_solves, "recompute": g.recompute_solves,
                               "adjoint": g.adjoint_solves},
           "gradient_pcg_iterations": g.pcg_iterations,
           "check": "finite differences on 64 voxels and an exact FP64 gradient: "
                    "tests/test_gpu_adjoint.py"}
    fw = [A.load_path_loss(vt, kp, b, rc, damage=True) for _ in range(steps)]
    gr = [A.load_path_grad(vt, kp, b, rc, damage=True)[0] for _ in range(steps)]
    f, g = fw[-1], gr[-1]
    out["damage"] = {"loss": f.loss,
                     "forward_ms": round(float(np.median([i.ms for i in fw])), 2),
                     "gradient_ms": round(float(np.median([i.ms for i in gr])), 2),
                     "gradient_pcg_iterations": g.pcg_iterations,
                     "check": "exact FP64 damage-path gradient: tests/test_gpu_adjoint.py"}
    log(f"96^3 adjoint: {out}")
    return out


def label_bench(torch, D, W, isl, steps=10):
    """Connected-component labelling (pfsf_hip_label_components, 26-connected,
    64 records) of the 128^3 island and of a fragmented 256^3 grid (random
    voxels at 12 % density: ~10^5 components), and the solver set-up that
    builds on the same union-find (anchored mask + hierarchy), median ms."""
    import numpy as np

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0x9F5F0000 + 5)
    r = rng.random((256, 256, 256))
    frag = np.where(r < 0.11, 1, np.where(r < 0.12, 2, 0)).astype(np.uint8)
    out = {}
    for name, t in (("island_128", isl.vtype), ("fragmented_256", frag)):
        tg = torch.from_numpy(np.ascontiguousarray(t)).to(dev)
        ids = torch.empty(tg.shape, dtype=torch.int32, device=dev)
        rec = torch.empty((64, 12), dtype=torch.int32, device=dev)
        D.label_components(tg, 26, 64, ids, rec)  # warm-up
        ts = []
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, num, ov = D.label_components(tg, 26, 64, ids, rec)
            ts.append(time.perf_counter() - t0)
        roots = int(torch.count_nonzero(ids == torch.arange(1, tg.numel() + 1, device=dev,
                                                            dtype=torch.int32).view(tg.shape)))
        ms = float(np.median(ts)) * 1e3
        out[name] = {"ms": round(ms, 3), "voxels": int(tg.numel()),
                     "voxels_per_s": round(tg.numel() / (ms * 1e-3), 1), "components": roots,
                     "records": num, "overflow": ov}
        del tg, ids
    vt = torch.from_numpy(isl.vtype).to(dev)
    kp = torch.from_numpy(isl.kappa).to(dev)
    ts = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = D.Solver(vt, kp, tol=1e-6, max_iters=2000)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        s.close()
    out["solver_setup_128_ms"] = round(float(np.median(ts)) * 1e3, 2)
    out["note"] = ("synchronous C-ABI calls timed on the host: scratch allocation and the "
                   "count readback included")
    log(f"labels: {out}")
    return out


def rbgs_bench(torch, D, W, sweeps=200, steps=5):
    """The plain method on the dense 32^3 block (all voxels occupied, z=0
    anchored): 200 eight-colour RBGS sweeps from phi = 0, then the failure
    scan, on device arrays; the same run of the C oracle (one host thread) is
    timed beside it and both outputs are compared bit for bit."""
    import numpy as np

    from oracle import oracle as O

    isl = W.workload(0)
    dev = torch.device("cuda:0")
    vt, kp, b, rc, rt = (torch.from_numpy(a).to(dev)
                         for a in (isl.vtype, isl.kappa, isl.source, isl.rcomp, isl.rtens))
    phi = torch.zeros(isl.vtype.shape, dtype=torch.float32, device=dev)
    code = torch.zeros_like(vt)
    ts = []
    for _ in range(steps + 1):
        phi.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        D.rbgs(vt, kp, b, phi, sweeps)
        D.failure_scan(vt, kp, phi, rc, rt, code)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    gpu_ms = float(np.median(ts[1:])) * 1e3
    t0 = time.perf_counter()
    ref_phi = O.rbgs(isl, sweeps=sweeps)
    ref_code = O.failure_scan(isl, ref_phi)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    nthr = O.set_threads(host_threads())  # the same run on all host threads (same bits)
    t0 = time.perf_counter()
    all_phi = O.rbgs(isl, sweeps=sweeps)
    all_code = O.failure_scan(isl, all_phi)
    cpu_all_ms = (time.perf_counter() - t0) * 1e3
    O.set_threads(1)
    assert np.array_equal(all_phi, ref_phi) and np.array_equal(all_code, ref_code)
    n = isl.n_active
    # algorithmic bytes of a sweep: read phi, kappa, b, type, write phi = 17 B
    # per active voxel (each array element once per sweep)
    alg = 17.0 * n * sweeps
    out = {"grid": [isl.lx, isl.ly, isl.lz], "active_voxels": n, "sweeps": sweeps,
           "gpu_ms": round(gpu_ms, 3),
           "roofline": {"bound": "hbm", "achieved": round(alg / (gpu_ms * 1e-3) / 1e9, 1),
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(alg / (gpu_ms * 1e-3) / HBM_PEAK, 5),
                        "per_active_voxel": "17 B per sweep"},
           "gpu_voxel_updates_per_s": round(n * sweeps / (gpu_ms * 1e-3), 1),
           "cpu_ms": round(cpu_ms, 1), "cpu_cores": 1, "cpu_kind": "port",
           "cpu_all_cores_ms": round(cpu_all_ms, 1), "cpu_all_cores": nthr,
           "cpu_voxel_updates_per_s": round(n * sweeps / (cpu_ms * 1e-3), 1),
           "phi_bit_exact": bool(np.array_equal(phi.cpu().numpy(), ref_phi)),
           "codes_bit_exact": bool(np.array_equal(code.cpu().numpy(), ref_code)),
           "note": "synchronous C-ABI calls timed on the host: the 200 sweeps in one launch "
                   "(work-group z keeps plane z and its couplings in LDS; neighbouring planes "
                   "pass as tagged 8-byte words, two dependent hand-offs per sweep) plus the "
                   "failure scan.  The bound is the chain of 400 hand-offs and 1600 colour "
                   "passes, not HBM: the HBM fraction is given for the record only"}
    log(f"rbgs 32^3: {out}")
    return out


def data_prep_bench(torch, D, isl, steps=10):
    """Island data prep on the device (pfsf_hip_build_island_data) for the
    128^3 island's voxels with 16 seeded materials (one indestructible) and
    random curing, against the C oracle's restatement of the Java loop on the
    same inputs (one host thread), all outputs compared bit for bit."""
    import numpy as np

    from oracle import oracle as O
    from pfsf_amd import capi

    rng = np.random.default_rng(0x9F5F0000 + 6)
    nm = 16
    rows = np.stack([rng.uniform(1500, 7800, nm), rng.uniform(5, 60, nm),
                     rng.uniform(0.5, 6, nm), (np.arange(nm) == 0).astype(float)], 1)
    table = np.zeros(nm, dtype=capi.MATERIAL_DTYPE)
    table["density"], table["rcomp"], table["rtens"] = rows[:, 0], rows[:, 1], rows[:, 2]
    ind = rows[:, 3].astype(np.uint8)
    mi = np.where(isl.vtype != 0, rng.integers(0, nm, isl.vtype.shape), -1).astype(np.int32)
    anchor = (isl.vtype == 2).astype(np.uint8)
    curing = rng.uniform(0.5, 1.0, isl.vtype.shape).astype(np.float32)
    dev = torch.device("cuda:0")
    gm, ga, gc = (torch.from_numpy(a).to(dev) for a in (mi, anchor, curing))
    D.build_island_data(gm, table, ind, ga, gc)  # warm-up
    ts = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        got = D.build_island_data(gm, table, ind, ga, gc)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    gpu_ms = float(np.median(ts)) * 1e3
    t0 = time.perf_counter()
    want = O.build_island_data(mi, rows.astype(np.float32), anchor, curing)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    exact = all(np.array_equal(got[k].cpu().numpy().view(np.uint8), want[k].view(np.uint8))
                for k in ("type", "source", "cond", "rcomp", "rtens", "max_phi"))
    out = {"grid": [isl.lx, isl.ly, isl.lz], "materials": nm, "gpu_ms": round(gpu_ms, 3),
           "gpu_voxels_per_s": round(isl.n / (gpu_ms * 1e-3), 1), "cpu_ms": round(cpu_ms, 1),
           "cpu_cores": 1, "cpu_kind": "port", "bit_exact": bool(exact),
           "note": "synchronous C-ABI call timed on the host, material table upload included"}
    log(f"data prep 128^3: {out}")
    return out


def host_tick_bench(isl, steps=3):
    """The same 128^3 island through the host-buffer C ABI, as the Java runtime
    drives it: pfsf_register_island_buffers + pfsf_tick_dbb per tick (PCIe
    upload of every registered array, hierarchy rebuild when type or kappa
    changed, solve, history, scan, failure compaction, phi readback).  Cold: phi zeroed before each tick;
    warm: the previous tick's phi is the start (an unchanged island)."""
    import numpy as np

    from pfsf_amd import capi
    from pfsf_amd.engine import Engine

    out = {}
    with Engine(max_island_size=isl.n, tick_budget_ms=1 << 30) as e:
        if e.init() != capi.OK or e.add_island(1, isl.lx, isl.ly, isl.lz) != capi.OK:
            return {"error": "engine init"}
        e.register(1, isl, kappa=isl.kappa)
        e.tick_dbb([1], capacity=1 << 16)  # warm-up
        for mode in ("cold", "warm"):
            ts = []
            for _ in range(steps):
                if mode == "cold":
                    e.phi(1)[...] = 0
                e.lib.pfsf_mark_full_rebuild(e.h, 1)
                t0 = time.perf_counter()
                r, _ = e.tick_dbb([1], capacity=1 << 16)
                ts.append(time.perf_counter() - t0)
            out[f"{mode}_tick_ms"] = round(float(np.median(ts)) * 1e3, 3)
    out["note"] = ("pfsf_tick_dbb on host buffers: PCIe traffic included, so this is not the "
                   "bench value")
    log(f"host-buffer tick 128^3: {out}")
    return out


def multi_island_tick_bench(W, count=256, steps=3):
    """A tick of `count` dirty islands through the host-buffer C ABI
    (register + pfsf_tick_dbb of all of them, PCIe upload and readback
    included), islands drawn as the batch workload's (log-uniform 64..262,144
    voxels).  kappa: each island registers its per-voxel kappa, and the
    islands are solved in one batched solve; soa6: faces only (kappa on all
    six faces), the islands' SoA-6 solves run side by side on the engine's
    host worker threads.  Cold: every phi zeroed before the tick."""
    import numpy as np

    from pfsf_amd import capi
    from pfsf_amd.engine import Engine

    isls = W.island_batch(count, index=7)
    out = {"islands": count, "active_voxels": int(sum(i.n_active for i in isls))}
    for mode in ("kappa", "soa6"):
        with Engine(max_island_size=1 << 22, tick_budget_ms=1 << 30,
                    vram_budget=64 << 30) as e:
            if e.init() != capi.OK:
                return {"error": "engine init"}
            ids = list(range(1, count + 1))
            for k, isl in zip(ids, isls):
                if e.add_island(k, isl.lx, isl.ly, isl.lz, origin=(80 * (k % 64), 0, 80 * (k // 64))) \
                        != capi.OK:
                    return {"error": f"add_island {k}"}
                e.register(k, isl, kappa=isl.kappa if mode == "kappa" else None)
            e.tick_dbb(ids, capacity=1 << 16)  # warm-up: hierarchies built
            ts = []
            for _ in range(steps):
                for k in ids:
                    e.phi(k)[...] = 0
                    e.lib.pfsf_mark_full_rebuild(e.h, k)
                t0 = time.perf_counter()
                r, ev = e.tick_dbb(ids, capacity=1 << 16)
                ts.append(time.perf_counter() - t0)
            conv = sum(e.status(k)[1].last_solve.converged for k in ids)
            out[mode] = {"tick_ms": round(float(np.median(ts)) * 1e3, 2),
                         "islands_per_s": round(count / float(np.median(ts)), 1),
                         "converged": int(conv), "failure_records": int(len(ev))}
    out["note"] = ("pfsf_tick_dbb on host buffers, cold start: PCIe traffic and the per-island "
                   "uploads included")
    log(f"multi-island tick: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--levels", type=int, default=0, help="MG levels (0 = auto hierarchy)")
    ap.add_argument("--precision", type=int, default=0, help="0 FP32, 1 BF16 coarse kappa")
    ap.add_argument("--coarse-sweeps", type=int, default=0,
                    help="Chebyshev steps on the coarsest level of the 128^3, 256^3 and "
                         "batched solves (0 = solver default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-island batch")
    ap.add_argument("--no-256", action="store_true", help="skip the 256^3 mixed-precision record")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 96^3 adjoint record")
    ap.add_argument("--no-labels", action="store_true", help="skip the labelling record")
    ap.add_argument("--no-rbgs", action="store_true", help="skip the dense 32^3 RBGS record")
    ap.add_argument("--no-prep", action="store_true", help="skip the island data-prep record")
    ap.add_argument("--no-host-tick", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) tick records")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the 4096-island batch and print its record")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="finish with 3 plain 256 MiB device copies (reference kernel for "
                         "hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (islands would "
              "shard across GPUs by bin-packing them by voxel count).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pfsf_amd import device as D
    from pfsf_amd import workloads as W

    if args.batch_only:
        print(json.dumps({"batched": batch_bench(torch, D, W, steps=args.steps,
                                                 coarse_sweeps=args.coarse_sweeps)}))
        return

    isl = W.workload(1)
    n_act = isl.n_active
    dev = torch.device("cuda:0")
    vt = torch.from_numpy(isl.vtype).to(dev)
    kp = torch.from_numpy(isl.kappa).to(dev)
    b = torch.from_numpy(isl.source).to(dev)
    rc = torch.from_numpy(isl.rcomp).to(dev)
    rt = torch.from_numpy(isl.rtens).to(dev)
    x64 = torch.zeros(isl.vtype.shape, dtype=torch.float64, device=dev)
    phi = torch.zeros(isl.vtype.shape, dtype=torch.float32, device=dev)
    H = torch.zeros_like(phi)
    code = torch.zeros_like(vt)
    solver = D.Solver(vt, kp, mg_levels=args.levels, precision=args.precision, tol=1e-6,
                      max_iters=2000, coarse_sweeps=args.coarse_sweeps)

    def step():
        x64.zero_()
        info = solver.solve(b, x64, phi)
        D.history_update(vt, kp, phi, H)
        D.failure_scan(vt, kp, phi, rc, rt, code)
        return info

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    solve_ms = float(np.median([i.solve_ms for i in infos]))
    last = infos[-1]

    # live per-kernel timing (HIP events on the solver stream)
    prof = {}
    for k in KERNEL_NAMES:
        if k in (3, 5) and last.levels < 2:
            continue
        prof[k] = solver.profile(k, reps=100)
    iter_ms = solver.profile(4, reps=30)
    dom = max(prof, key=prof.get)
    ach = ALG_BYTES[dom] * n_act / (prof[dom] * 1e-3)
    traffic = measured_traffic().get(KERNEL_NAMES[dom])
    # whole-iteration algorithmic rate
    it_bytes = ITER_ALG_BYTES * n_act
    # whole solve: per active voxel, a fused PCG iteration (SpMV + dot 13,
    # update + z + dot 29, direction 17 = 59 B) plus one 2-level V-cycle
    # (67 B) per iteration, and 17 B for the initial residual
    solve_bytes = (SOLVE_ITER_BYTES * last.iterations + SOLVE_INIT_BYTES) * n_act
    hist = np.bincount(code.cpu().numpy().ravel(), minlength=5).tolist()
    # accuracy of the timed phi: against the same system solved to an FP64
    # residual of 1e-11 (tests/test_gpu_solver.py checks the 1e-6 phi against
    # the CPU oracle's FP64 solution too)
    ref = D.Solver(vt, kp, tol=1e-11, max_iters=4000)
    xr = torch.zeros_like(x64)
    pr = torch.empty_like(phi)
    iref = ref.solve(b, xr, pr)
    ref.close()
    phi_err = float(torch.linalg.norm((phi - pr).double()) / torch.linalg.norm(pr.double()))
    code_ref = torch.empty_like(code)
    D.failure_scan(vt, kp, pr, rc, rt, code_ref)

    out = {
        "metric": METRIC,
        "value": round(ms_step, 4),
        "unit": "ms per 128^3 solve to 1e-6 (+ history + scan)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "single island 128^3 box union, PCG + V-cycle to 1e-6, "
                               "then psi/H update and failure scan",
                   "grid": [isl.lx, isl.ly, isl.lz], "active_voxels": n_act,
                   "solid_voxels": isl.n_solid, "mg_levels": last.levels,
                   "precision": ["fp32", "bf16-coarse"][args.precision],
                   "inner": "FP32 SpMV/smoothing, FP64 dots + FP64 refinement residual"},
        "solve_ms_median": round(solve_ms, 4),
        "pcg_iterations": last.iterations, "refinement_passes": last.refinements,
        "rel_residual": last.rel_residual,
        "phi_rel_l2_vs_fp64_reference": phi_err,
        "failure_codes_differing_vs_reference": int((code != code_ref).sum()),
        "reference": f"same system solved to an FP64 residual of {iref.rel_residual:.2e}",
        "pcg_iteration_ms": round(iter_ms, 5),
        "iteration_alg_GBps": round(it_bytes / (iter_ms * 1e-3) / 1e9, 1),
        "voxel_updates_per_s": round(n_act * last.iterations / (solve_ms * 1e-3), 1),
        "solve_alg_bandwidth": {
            "bytes": round(solve_bytes), "GBps": round(solve_bytes / (solve_ms * 1e-3) / 1e9, 1),
            "frac": round(solve_bytes / (solve_ms * 1e-3) / HBM_PEAK, 4),
            "per_active_voxel": f"{SOLVE_ITER_BYTES} B x {last.iterations} iterations + "
                                f"{SOLVE_INIT_BYTES} B"},
        "kernel_ms": {KERNEL_NAMES[k]: round(v, 5) for k, v in prof.items()},
        "failure_histogram": hist,
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom],
                     "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK, 4), "traffic": traffic,
                     "traffic_source": TRAFFIC_FILE.name if traffic else None,
                     "hbm_GBps_measured_traffic": (round(traffic / (prof[dom] * 1e-3) / 1e9, 1)
                                                   if traffic else None),
                     "alg_bytes_per_active_voxel": ALG_BYTES[dom],
                     "alg_bytes_per_launch": round(ALG_BYTES[dom] * n_act),
                     "kernel_us": round(prof[dom] * 1e3, 3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["cpu_baseline_one_core"] = cpu_baseline(isl)
    solver.close()
    if not args.no_host_tick:
        out["host_buffer_tick_128"] = host_tick_bench(isl)
        out["host_buffer_tick_256_islands"] = multi_island_tick_bench(W)
    if not args.no_batch:
        out["batched"] = batch_bench(torch, D, W, coarse_sweeps=args.coarse_sweeps)
    if not args.no_256:
        out["mixed_precision_256"] = mixed_precision_bench(torch, D, W,
                                                           coarse_sweeps=args.coarse_sweeps)
    if not args.no_adjoint:
        out["adjoint_96"] = adjoint_bench(torch, W)
    if not args.no_labels:
        out["labels"] = label_bench(torch, D, W, isl)
    if not args.no_rbgs:
        out["rbgs_32"] = rbgs_bench(torch, D, W)
    if not args.no_prep:
        out["data_prep_128"] = data_prep_bench(torch, D, isl)
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
