#!/usr/bin/env python3
"""Benchmark of the HSS ULV solve (and matvec) on one MI355X.

Workload (the metric's configuration): n = 65536 points sorted uniform on
[-1, 1] (numpy default_rng(0)), kernel log|x_i - x_j| with unit diagonal,
bisection leaf size 128, direct compression at rtol = atol = 1e-9 on the device
(hss_compress_dev; 512 leaves, depth 9), 64 right-hand sides N(0,1)
(default_rng(1)).  The batched and serving sections build their compressed
Cauchy and log-kernel instances on the device the same way (untimed).

A "step" is one ULV solve of the 64-column block with a factorization resident
in HBM (inputs and outputs already on the device):
  * fp64  : Float64 factor + Float64 solve (backward stable, no refinement);
  * mixed : Float32 factor + Float32 solve, Float64 HSS-matvec residual and
            iterative refinement until max_j ||b_j - A x_j|| / ||b_j|| <= 1e-10.
The headline ``value`` is the fp64 path on this workload (solves/s; its
relative residual, computed with the library's Float64 matvec, is reported
and must be <= 1e-10; the CPU baseline's solution is the parity reference).  The mixed-precision path on the same matrix (config
C3), the batched uneven instances (C4: one launch per tree level for 256
matrices, mixed vs plain fp64), the serving loop (C5: p50/p99 latency under
Poisson arrivals), the factor time, the matvec bandwidth, parity against the
CPU oracle and the CPU baseline are reported beside it.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
Prints ONE JSON line on stdout (progress goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector = matrix)
FP32_MFMA_PEAK_TFLOPS = 157.3
N_POINTS, LEAF, TOL, NRHS = 65536, 128, 1e-9, 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_matrix(hip, W):
    """The n = 65536 log-kernel matrix, compressed on the device (untimed), and its
    generators read back for the CPU baseline and the byte models."""
    t = time.time()
    A = W.log_kernel(hip, N_POINTS, LEAF, TOL, 0)
    nodes, blob = A.generators()
    log(f"[bench] n={N_POINTS} log kernel compressed on the device in {time.time() - t:.1f} s (untimed)")
    return A, nodes, blob


# ---------------------------------------------------------------------------
# algorithmic byte / flop models (per launch): every operator, input and output counted once
# ---------------------------------------------------------------------------
def matvec_leaf_bytes(nodes, nrhs, beta_nonzero=False):
    """k_mv_leaf*: D, U, B_seg, F read once, C written once (8 B each)."""
    b = 0
    for r in nodes[nodes["left"] < 0]:
        m, n, ku = int(r["m"]), int(r["n"]), int(r["ku"])
        b += m * n + m * ku + (n + ku + m * (2 if beta_nonzero else 1)) * nrhs
    return 8 * b


def matvec_bytes(nodes, nrhs):
    """Whole matvec: every generator once + B in + C out (algorithmic)."""
    G = 0
    for r in nodes:
        if r["left"] < 0:
            G += int(r["m"]) * int(r["n"]) + int(r["m"]) * int(r["ku"]) + int(r["n"]) * int(r["kv"])
        else:
            L, R = nodes[r["left"]], nodes[r["right"]]
            G += int(L["ku"]) * int(R["kv"]) + int(R["ku"]) * int(L["kv"])
        if r["parent"] >= 0:
            p = nodes[r["parent"]]
            G += int(r["ku"]) * int(p["ku"]) + int(r["kv"]) * int(p["kv"])
    n, m = int(nodes["n"][0]), int(nodes["m"][0])
    return 8 * G + 8 * nrhs * (n + m), 2 * nrhs * G


def factor_flops(nodes, fact):
    """LAPACK flop counts of one ULV factorization (the reference's per-node calls,
    ulvfactor.jl:39-54, 74-86): per non-pass node with an m x n block, k = rank eliminated
    by the QL and nk = m - k LQ reflectors: geqlf(m, k) = 2 k^2 (m - k/3); ormql on D
    (m x n, k reflectors) = 4 m n k - 2 n k^2; gelqf(nk, n) = 2 nk^2 (n - nk/3); the LQ
    reflectors applied to the k bottom rows and to V' (kv rows) = 4 (k + kv) n nk -
    2 (k + kv) nk^2; the merge of a branch 2 m n k; the root LU (2/3) m^3.  The explicit
    solve operators this library forms on top are not counted."""
    total = 0.0
    for i in range(1, len(nodes)):
        f = fact.node_info(i)
        m, n, k, nk = f["m"], f["n"], f["k"], f["nk"]
        if f["pass"]:
            continue
        kv = int(nodes["kv"][i])
        total += 2.0 * k * k * (m - k / 3.0)
        total += 4.0 * m * n * k - 2.0 * n * k * k
        total += 2.0 * nk * nk * (n - nk / 3.0)
        total += 4.0 * (k + kv) * n * nk - 2.0 * (k + kv) * nk * nk
        if nodes["left"][i] >= 0:
            total += 2.0 * m * n * k
    m0 = fact.node_info(0)["m"]
    total += 2.0 / 3.0 * m0 ** 3
    return total


def solve_kernel_model(nodes, fact, kernel, level, nrhs, wf, wv):
    """Algorithmic bytes and flops of one launch of a solve kernel at one level.

    k_ulv_gup (height ``level``): per node zloc = Gz x (Gz nk x Kx), then
      [reduced b; u] = Gr [x; zloc] (Gr R x (Kx + nk), R = My - nk): reads Gz and Gr once
      and the input x (Kx rows: the leaf's rows of B in Float64, or the children's y
      blocks), writes y (My rows) per column.
    k_ulv_gdown (depth ``level``): per node z = Qt [zloc; z_parent] with Qt n x n
      (identity for a pass-through node): reads Qt once and n input rows, writes n rows
      (the solution X in Float64 at a leaf) per column.
    wf / wv = bytes per factor / workspace element.
    """
    byts = 0
    flops = 0
    heights = nodes["height"]
    depths = nodes["depth"]
    up = kernel.startswith("k_ulv_gup")
    down = kernel.startswith("k_ulv_gdown")
    for i in range(1, len(nodes)):
        leaf = nodes["left"][i] < 0
        if up and heights[i] == level:
            f = fact.node_info(i)
            My, Kx, nk = f["My"], f["Kx"], f["nk"]
            g = nk * Kx + (My - nk) * (Kx + nk)
            byts += wf * g + nrhs * (Kx * (8 if leaf else wv) + My * wv)
            flops += 2 * g * nrhs
        elif down and depths[i] == level:
            f = fact.node_info(i)
            n = f["n"]
            q = 0 if f["pass"] else n * n
            byts += wf * q + nrhs * (n * wv + n * (8 if leaf else wv))
            flops += 2 * q * nrhs
    return byts, flops


def _ref_node_entries(nodes, fact, i):
    """Factor entries of node i in the reference's own (reflector) form, split by sweep
    (ulvfactor.jl:22-57, 74-89; the solve-unit count F of the measurement plan):
      up:   QL reflectors (m k), the L1 triangle (nk (nk+1)/2), L2a (k nk), V_top (nk kv),
            and at a branch the merge's couplings U1 B12 (mr1 kv2), U2 B21 (mr2 kv1) and
            translators W1 (kv1 kv), W2 (kv2 kv);
      down: the LQ reflectors below the diagonal of the nk x n block, nk n - nk (nk+1)/2
            (with the L1 triangle, the (m - k) n of the packed LQ block).
    Returns (up_entries, up_reflector_entries, down_entries)."""
    f = fact.node_info(i)
    m, n, k, nk = f["m"], f["n"], f["k"], f["nk"]
    kv = int(nodes["kv"][i])
    up = refl = down = 0
    if not f["pass"]:
        refl = m * k
        up = m * k + nk * (nk + 1) // 2 + k * nk + nk * kv
        down = nk * n - nk * (nk + 1) // 2
    if nodes["left"][i] >= 0:
        a, b = int(nodes["left"][i]), int(nodes["right"][i])
        fa, fb = fact.node_info(a), fact.node_info(b)
        kva, kvb = int(nodes["kv"][a]), int(nodes["kv"][b])
        up += fa["mr"] * kvb + fb["mr"] * kva + (kva + kvb) * kv
    return up, refl, down


def solve_ref_model(nodes, fact, kernel, level, nrhs, wf, wv):
    """Algorithmic bytes and flops of one solve launch counted on the reference's factor
    form (``_ref_node_entries``), not on this library's explicit node operators.
    k_ulv_gup (height ``level``): the node's up entries, its input (Kx rows: B in Float64
    at a leaf, else the children's reduced b and u) read and y = [zloc; b; u] (My rows)
    written per column.  k_ulv_gdown (depth ``level``): the node's LQ reflectors, n rows in
    ([zloc; parent z]) and n rows out (X in Float64 at a leaf) per column.  Flops: 4 per
    reflector entry and 2 per other entry, per column."""
    byts = flops = 0
    up = kernel.startswith("k_ulv_gup")
    for i in range(1, len(nodes)):
        leaf = nodes["left"][i] < 0
        if up and nodes["height"][i] == level:
            f = fact.node_info(i)
            e, refl, _ = _ref_node_entries(nodes, fact, i)
            byts += wf * e + nrhs * (f["Kx"] * (8 if leaf else wv) + f["My"] * wv)
            flops += nrhs * (4 * refl + 2 * (e - refl))
        elif not up and nodes["depth"][i] == level:
            f = fact.node_info(i)
            _, _, e = _ref_node_entries(nodes, fact, i)
            byts += wf * e + nrhs * f["n"] * (wv + (8 if leaf else wv))
            flops += 4 * nrhs * e
    return byts, flops


def solve_ref_total(nodes, fact, nrhs, wf=8):
    """Whole-solve algorithmic bytes of the measurement plan: w_f F + 8 nrhs 2n, F = every
    node's reference-form factor entries plus the root's LU (m^2) and couplings."""
    F = 0
    for i in range(1, len(nodes)):
        up, _, down = _ref_node_entries(nodes, fact, i)
        F += up + down
    r = fact.node_info(0)
    F += r["m"] ** 2
    if nodes["left"][0] >= 0:
        a, b = int(nodes["left"][0]), int(nodes["right"][0])
        F += fact.node_info(a)["mr"] * int(nodes["kv"][b]) + fact.node_info(b)["mr"] * int(nodes["kv"][a])
    n = int(nodes["n"][0])
    return wf * F + 8 * nrhs * 2 * n, F


# ---------------------------------------------------------------------------
# batched (C4) and serving (C5) measurements on kernel matrices compressed on the device
# (hssmatrices.jl_amd.workloads; generation is untimed)
# ---------------------------------------------------------------------------
def _dev_rhs(torch, rng, n, r, dev):
    """B (r x n torch tensor = column-major n x r) and a zero X; synchronized, because torch
    fills on its own stream and the library works on its context stream."""
    B = torch.from_numpy(np.ascontiguousarray(rng.standard_normal((r, n)))).to(dev)
    X = torch.zeros_like(B)
    torch.cuda.synchronize()
    return B, X


def _timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def batched(hip, W, torch, steps=10, warmup=2, count=256, seed=3, log=print):
    """C4: `count` compressed Cauchy matrices (n log-uniform in [1024, 16384], leaf 64,
    rtol 1e-8), 1-32 RHS each, every solve to relative residual <= 1e-10.  One batched plan
    (one launch per tree level for all instances; Float32 factors + Float64 refinement in
    lockstep, or Float64 factors + refinement) against the plain path: one Float64 solve
    call per instance, refined to the same 1e-10 where its first solve is not there."""
    dev = torch.device("cuda:0")
    t = time.time()
    mats, ns, shifts, rs = W.cauchy_instances(hip, count, seed=seed)
    ks = [hip_summary(m) for m in mats]
    log(f"[bench] batch: {count} compressed Cauchy instances built on the device in {time.time() - t:.1f} s "
        f"(n {ns.min()}-{ns.max()}, ranks {min(k[0] for k in ks)}-{max(k[1] for k in ks)}, total RHS {int(rs.sum())})")
    rng = np.random.default_rng(seed + 100)
    Bs, Xs = [], []
    for n, r in zip(ns, rs):
        B, X = _dev_rhs(torch, rng, int(n), int(r), dev)
        Bs.append(B)
        Xs.append(X)
    out = {"workload": f"{count} compressed Cauchy matrices 1/(x_i-x_j) (unit diagonal), n {int(ns.min())}-"
                       f"{int(ns.max())} log-uniform, leaf 64, rtol 1e-8, ranks {min(k[0] for k in ks)}-"
                       f"{max(k[1] for k in ks)}, 1-32 RHS each (seed {seed})",
           "instances": count, "n_range": [int(ns.min()), int(ns.max())], "rhs_total": int(rs.sum())}
    for prec in ("fp64", "fp32"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        facts = hip.ulvfactor_batch(mats, prec)
        torch.cuda.synchronize()
        fms = (time.perf_counter() - t0) * 1e3
        plan = hip.Plan(facts, [int(r) for r in rs], [b.data_ptr() for b in Bs], [int(b.shape[1]) for b in Bs],
                        [x.data_ptr() for x in Xs], [int(x.shape[1]) for x in Xs], refinement=True)
        try:
            ms, rep = _timed(torch, lambda: plan.execute(tol=1e-10, max_iter=30), steps, warmup)
        except hip.NotConverged as e:
            log(f"[bench] batch {prec}: {e}")
            out[prec] = {"error": str(e)}
            plan.close()
            continue
        recs = profile_launches(hip, lambda: plan.execute(tol=1e-10, max_iter=30), 1)
        _, byk = summarize(recs, 1)
        out[prec] = {"batch_ms": round(ms, 3), "solves_per_s": round(count * 1e3 / ms, 1),
                     "factor_batch_ms": round(fms, 2), "refine_steps": rep["iterations"],
                     "worst_rel_residual": rep["rel_residual"],
                     "worst_rel_residual_fp64_check": _batch_residual(mats, Bs, Xs, rs, torch),
                     "per_kernel_ms": {k: round(v, 4) for k, v in byk.items()}}
        log(f"[bench] batch {prec}: {ms:.3f} ms per batch -> {count * 1e3 / ms:.1f} solves/s "
            f"({rep['iterations']} refinement steps, factor {fms:.1f} ms); per kernel {out[prec]['per_kernel_ms']}")
        # the same plan after the autotuner chose each level's launch (hss_plan_tune)
        t0 = time.perf_counter()
        tune = plan.tune(reps=3)
        tune_wall = (time.perf_counter() - t0) * 1e3
        tms, _ = _timed(torch, lambda: plan.execute(tol=1e-10, max_iter=30), steps, warmup)
        out[prec]["tuned"] = {"batch_ms": round(tms, 3), "solves_per_s": round(count * 1e3 / tms, 1),
                              "tuning_wall_ms": round(tune_wall, 1),
                              "one_solve_ms_before_after": [round(tune["ms_before"], 3), round(tune["ms_after"], 3)],
                              "worst_rel_residual_fp64_check": _batch_residual(mats, Bs, Xs, rs, torch)}
        out[prec]["best_solves_per_s"] = max(out[prec]["solves_per_s"], out[prec]["tuned"]["solves_per_s"])
        log(f"[bench] batch {prec} tuned (in {tune_wall:.0f} ms): {tms:.3f} ms per batch -> "
            f"{count * 1e3 / tms:.1f} solves/s")
        plan.close()
        if prec == "fp64":
            # plain path: one resident-factor FP64 solve call per instance; then the same calls
            # with refinement to 1e-10 (the FP64 factor's own IR), which the metric's residual
            # asks of every solve
            def plain():
                for f, b, x, r in zip(facts, Bs, Xs, rs):
                    f.solve_dev(int(r), b.data_ptr(), b.shape[1], x.data_ptr(), x.shape[1])
            pms, _ = _timed(torch, plain, max(1, steps // 2), warmup)
            res_plain = _batch_residual(mats, Bs, Xs, rs, torch)

            def plain_ir():
                for f, b, x, r in zip(facts, Bs, Xs, rs):
                    f.solve_dev(int(r), b.data_ptr(), b.shape[1], x.data_ptr(), x.shape[1], tol=1e-10, max_iter=10)
            pims, _ = _timed(torch, plain_ir, max(1, steps // 2), warmup)
            out["plain_fp64"] = {"ms": round(pms, 3), "solves_per_s": round(count * 1e3 / pms, 1),
                                 "worst_rel_residual": res_plain,
                                 "refined_to_1e-10": {"ms": round(pims, 3), "solves_per_s": round(count * 1e3 / pims, 1),
                                                      "worst_rel_residual": _batch_residual(mats, Bs, Xs, rs, torch)}}
            log(f"[bench] plain per-instance fp64: {pms:.3f} ms -> {count * 1e3 / pms:.1f} solves/s (worst residual "
                f"{res_plain:.2e}); refined to 1e-10: {pims:.3f} ms -> {count * 1e3 / pims:.1f} solves/s")
        for f in facts:
            f.close()
    base = out["plain_fp64"]["refined_to_1e-10"]["solves_per_s"]
    for prec, key in (("fp32", "mixed_batched_vs_plain_fp64"), ("fp64", "fp64_batched_vs_plain_fp64")):
        if "solves_per_s" in out[prec]:
            out[key] = round(out[prec]["best_solves_per_s"] / base, 2)
    return out


def construction(hip, W, torch, log=print, n_div=16384):
    """Building HSS matrices on the device: the n = 65536 log-kernel matrix compressed directly
    (hss_compress_dev) and by randomized adaptive compression (hss_randcompress_dev, the
    reference's defaults kest 10, noversampling 10, stepsize 20), each timed from the dense
    matrix in HBM (its formation by torch excluded) and checked on 4 random vectors against the
    dense matrix; and the HSS divide A \\ A (hss_mat_ldiv) of an n = n_div log-kernel matrix,
    which should be the identity."""
    from hssmatrices_jl_amd.compression import KernelMatrix
    from hssmatrices_jl_amd.hssmatrix import bisection_cluster
    out = {}
    x = W.log_points(N_POINTS, 0)
    K = KernelMatrix(x, "log", 1.0)
    T = K.device_dense()
    cl = bisection_cluster(N_POINTS, LEAF)
    g = torch.Generator(device="cuda:0").manual_seed(5)
    Xv = torch.randn((4, N_POINTS), dtype=torch.float64, device="cuda:0", generator=g)
    Yref = (Xv @ T).cpu().numpy().T  # T holds A column-major as a row-major (n, n) tensor: x' A' = (A x)'
    Xh = Xv.cpu().numpy().T
    hssrank = lambda nd: int(max(nd["ku"][1:].max(), nd["kv"][1:].max()))
    for name, fn in (("direct", lambda: hip.compress_dev(T.data_ptr(), N_POINTS, cl, cl, TOL, TOL)),
                     ("randomized", lambda: hip.randcompress_dev(T.data_ptr(), N_POINTS, cl, cl, TOL, TOL,
                                                                 report=True))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        M, smp = (r if isinstance(r, tuple) else (r, None))
        nd, _ = M.generators()
        err = float(np.linalg.norm(M @ Xh - Yref) / np.linalg.norm(Yref))
        out[name] = {"ms": round(ms, 1), "hssrank": hssrank(nd), "rel_err_on_4_vectors": err}
        if smp is not None:
            out[name]["samples"] = smp
        log(f"[bench] {name} compression n={N_POINTS}: {ms:.0f} ms, hssrank {hssrank(nd)}, rel err {err:.2e}"
            + (f", samples {smp}" if smp is not None else ""))
        M.close()
    del T, Xv
    torch.cuda.empty_cache()
    Ad = W.log_kernel(hip, n_div, LEAF, TOL, 7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Q = Ad.hss_ldiv(Ad, TOL, TOL)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    Xd = np.random.default_rng(8).standard_normal((n_div, 4))
    e = float(np.linalg.norm(Q @ Xd - Xd) / np.linalg.norm(Xd))
    out["divide"] = {"case": f"A \\ A, log kernel n={n_div} (the identity)", "ms": round(ms, 1),
                     "hssrank": hssrank(Q.generators()[0]), "rel_err_vs_identity": e}
    log(f"[bench] divide {out['divide']}")
    Q.close()
    Ad.close()
    return out


def hip_summary(m):
    k = np.concatenate([m.nodes["ku"][1:], m.nodes["kv"][1:]])
    return int(k.min()), int(k.max())


def _batch_residual(mats, Bs, Xs, rs, torch):
    """Largest relative residual max_j ||B_j - A X_j|| / ||B_j|| over every instance of the batch,
    computed with the library's FP64 matvec (untimed check of the timed results)."""
    worst = 0.0
    for A, B, X, r in zip(mats, Bs, Xs, rs):
        R = B.clone()
        torch.cuda.synchronize()  # torch's stream and the library's stream are distinct
        A.matmul_dev(int(r), 1.0, X.data_ptr(), X.shape[1], -1.0, R.data_ptr(), R.shape[1])
        torch.cuda.synchronize()
        rel = (torch.linalg.vector_norm(R, dim=1) / torch.linalg.vector_norm(B, dim=1)).max().item()
        worst = max(worst, rel)
    return worst


def _serve_run(hip, torch, srv, ids, ns, arrivals, rhs, which, bufs, log):
    """Submits request k at arrivals[k] seconds (all at once when arrivals is None), steps the
    server until every request is done; returns latencies (ms), refinement steps, worst
    residual, failures and the wall time."""
    tickets = {}
    lat, iters = [], []
    worst, failed = 0.0, 0
    requests = len(rhs)
    t0 = time.perf_counter()
    nxt = done = 0
    while done < requests:
        now = time.perf_counter() - t0
        while nxt < requests and (arrivals is None or arrivals[nxt] <= now):
            B, X = bufs[nxt]
            tk = srv.submit(ids[which[nxt]], int(rhs[nxt]), B.data_ptr(), B.shape[1], X.data_ptr(), X.shape[1])
            tickets[tk] = nxt
            nxt += 1
        if srv.step() == 0 and nxt < requests and arrivals is not None:
            time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.001)))
        for tk in list(tickets):
            r = srv.poll(tk)
            if r is not None:
                lat.append(r["latency_s"] * 1e3)
                iters.append(r["iterations"])
                worst = max(worst, r["rel_residual"])
                failed += int(r["status"] != 1)
                del tickets[tk]
                done += 1
    return lat, iters, worst, failed, time.perf_counter() - t0


def serving(hip, W, torch, count=32, requests=2000, seed=4, log=print, budget=200 * 2 ** 30):
    """C5: `count` resident FP32 factorizations of compressed log-kernel matrices (n
    log-uniform in [4096, 65536], leaf 128, rtol 1e-9); requests pick a matrix uniformly,
    RHS ~ geometric(mean 8) capped at 128, refinement to 1e-10.  First the saturation
    throughput (every request submitted at once), then Poisson arrivals at 70 % of it:
    p50/p99 latency and solves/s."""
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    t = time.time()
    mats, ns = W.log_kernel_instances(hip, count, seed=seed)
    ks = [hip_summary(m) for m in mats]
    srv = hip.Server(budget_bytes=budget, prec="fp32", tol=1e-10, max_iter=20, max_inflight=64)
    ids = [srv.admit(m) for m in mats]
    log(f"[bench] serve: {count} compressed log-kernel matrices built and their FP32 factors resident "
        f"({srv.used_bytes() / 2**20:.1f} MiB) in {time.time() - t:.1f} s")
    rhs = np.minimum(rng.geometric(1.0 / 8.0, requests), 128)
    which = rng.integers(0, count, requests)
    bufs = [_dev_rhs(torch, rng, int(ns[w]), int(r), dev) for w, r in zip(which, rhs)]
    # saturation: all requests at once
    lat_s, it_s, worst_s, fail_s, wall_s = _serve_run(hip, torch, srv, ids, ns, None, rhs, which, bufs, log)
    sat = requests / wall_s
    rate = 0.7 * sat
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    lat, iters, worst, failed, wall = _serve_run(hip, torch, srv, ids, ns, arrivals, rhs, which, bufs, log)
    srv.close()
    out = {"workload": f"{count} compressed log-kernel matrices, n {int(ns.min())}-{int(ns.max())} log-uniform, "
                       f"leaf 128, rtol 1e-9, ranks {min(k[0] for k in ks)}-{max(k[1] for k in ks)}; FP32 factors, "
                       f"refinement to 1e-10 (seed {seed})",
           "factors": count, "n_range": [int(ns.min()), int(ns.max())], "requests": requests,
           "rhs_mean": round(float(rhs.mean()), 2),
           "saturation": {"solves_per_s": round(sat, 1), "wall_s": round(wall_s, 3),
                          "latency_ms": {"p50": round(float(np.percentile(lat_s, 50)), 3),
                                         "p99": round(float(np.percentile(lat_s, 99)), 3)},
                          "worst_rel_residual": worst_s, "not_converged": fail_s},
           "arrival_rate_per_s": round(rate, 1), "solves_per_s": round(requests / wall, 1),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3)},
           "refine_steps": [int(min(iters)), int(max(iters))],
           "worst_rel_residual": worst, "not_converged": failed}
    log(f"[bench] serve: {out}")
    return out


PROFILE_COUNTERS = ROOT / "profiles" / "hss_c2_n65536_hbm_counters.md"


def sweep_launch(nodes, fact, kernel, level, nwork, nrhs):
    """rocprofv3 name, template arguments and launch size (threads) of one FP64 solve sweep
    launch recorded as ``kernel``, as the library launches it (padded to a multiple of 8
    workgroups).
    k_ulv_gup_s / k_ulv_gdown_s (operator staged in LDS) and k_ulv_gup_x / k_ulv_gdown_x
    (operator streamed; the same kernels with the second template argument false): one
    workgroup per node and tw-column tile, tw = 16 for narrow levels (fewer than 128 tiles
    of 32 columns) else 32, 256 threads on wide levels (2048 or more such tiles) else 512; k_ulv_gup: one 256-thread workgroup per node
    and 32 columns; k_ulv_gdown: 256-thread tiles of 64 x 64 outputs (128 x 32 for at
    most 32 right-hand sides)."""
    if kernel.endswith(("_s", "_x")):
        tw = 16 if nrhs <= 16 or nwork * (-(-nrhs // 32)) < 128 else 32
        total = nwork * (-(-nrhs // tw))
        nt = 256 if nwork * (-(-nrhs // 32)) >= 2048 else 512  # wide levels: 256-thread workgroups
        name = kernel[:-2] + "_s"
        staged = "true" if kernel.endswith("_s") else "false"
        return name, f"<double, {staged}, {nt}>", 8 * (-(-total // 8)) * nt
    if kernel == "k_ulv_gup":
        total = nwork * (-(-nrhs // 32))
        return kernel, "<double>", 8 * (-(-total // 8)) * 256
    tm_rows, tn_cols = (64, 64) if nrhs > 32 else (128, 32)
    sel = [i for i in range(1, len(nodes)) if nodes["depth"][i] == level]
    max_m = max(fact.node_info(i)["n"] for i in sel)
    total = nwork * (-(-max_m // tm_rows)) * (-(-nrhs // tn_cols))
    return kernel, "<double, 2, 2>" if nrhs > 32 else "<double, 4, 1>", 8 * (-(-total // 8)) * 256


def profiled_traffic(kernel: str, type_args: str, grid_x: int, grid_y: int):
    """HBM bytes of one launch as measured by the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs (copy-calibrated, last column of the table), or None when that launch shape was not
    profiled."""
    if not PROFILE_COUNTERS.exists():
        return None
    want = f"`hss::{kernel}{type_args}`"
    for line in PROFILE_COUNTERS.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == want and cells[1] == str(grid_x) and cells[2] == str(grid_y):
            return int(float(cells[5]) * 1e6)
    return None


def profile_launches(hip, fn, reps):
    hip.prof_reset()
    hip.prof_enable(True)
    for _ in range(reps):
        fn()
    recs = hip.prof_records()
    hip.prof_enable(False)
    hip.prof_reset()
    return recs


def summarize(recs, reps):
    tot = {}
    for r in recs:
        key = (r["kernel"], r["level"])
        t = tot.setdefault(key, [0.0, 0, r["nwork"]])
        t[0] += r["ms"]
        t[1] += 1
    by_kernel = {}
    for (k, lev), (ms, cnt, nw) in tot.items():
        by_kernel.setdefault(k, 0.0)
        by_kernel[k] += ms / reps
    return tot, by_kernel


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=NRHS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=8, help="1-thread CPU oracle calls in the baseline sample")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances (C4) measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (C5) measurement")
    ap.add_argument("--no-construct", action="store_true",
                    help="skip the construction measurements (device compression, randomized compression, divide)")
    ap.add_argument("--profile-only", action="store_true",
                    help="for rocprofv3: a few fp64 solves, matvecs and one plain 256 MiB device copy "
                         "(the copy calibrates FETCH_SIZE/WRITE_SIZE), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              f"a run over several GPUs would shard the tree by subtrees, one per GPU", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as G
    pkg = G.load_package()
    import importlib
    hip = importlib.import_module("hssmatrices_jl_amd.hip")
    import torch

    W = importlib.import_module("hssmatrices_jl_amd.workloads")
    torch.cuda.init()
    ctx = hip.default_context()
    A, nodes, blob = build_matrix(hip, W)
    summ = pkg.tree_summary(nodes)
    n = int(nodes["n"][0])
    nrhs = args.nrhs
    log(f"[bench] HSS n={n} leaves={summ['leaves']} depth={summ['max_depth']} ranks {summ['rank_min']}-"
        f"{summ['rank_max']} G={summ['G']}")
    dev = torch.device("cuda:0")

    # one untimed factorization of each precision first: the first launch of a kernel in a
    # process pays for loading its code object
    A.ulvfactor("fp64").close()
    A.ulvfactor("fp32").close()
    t0 = time.perf_counter()
    F64 = A.ulvfactor("fp64")
    f64_wall = time.perf_counter() - t0
    f64_ms = F64.factor_ms
    t0 = time.perf_counter()
    F32 = A.ulvfactor("fp32")
    f32_wall = time.perf_counter() - t0
    f32_ms = F32.factor_ms
    fflops = factor_flops(nodes, F64)
    log(f"[bench] factor fp64 {f64_ms:.2f} ms (wall {f64_wall*1e3:.1f}), fp32 {f32_ms:.2f} ms; "
        f"factor bytes fp64 {F64.device_bytes()/1e6:.1f} MB; {fflops/1e9:.2f} GFLOP (LAPACK counts) -> "
        f"{fflops / f64_ms / 1e9:.2f} TFLOP/s")

    rng = np.random.default_rng(1)
    B_host = rng.standard_normal((n, nrhs))
    Bt = torch.from_numpy(np.ascontiguousarray(B_host.T)).to(dev)  # (nrhs, n): column-major n x nrhs
    Xt = torch.zeros_like(Bt)
    X2t = torch.zeros_like(Bt)
    Ct = torch.zeros_like(Bt)
    torch.cuda.synchronize()  # torch's fills run on its own stream, the library on its context stream
    ld = n

    def solve64():
        F64.solve_dev(nrhs, Bt.data_ptr(), ld, Xt.data_ptr(), ld)

    def solve_mixed():
        return F32.solve_dev(nrhs, Bt.data_ptr(), ld, X2t.data_ptr(), ld, tol=1e-10, max_iter=10)

    def matvec():
        A.matmul_dev(nrhs, 1.0, Bt.data_ptr(), ld, 0.0, Ct.data_ptr(), ld)

    if args.profile_only:
        for _ in range(max(args.steps, 1)):
            solve64()
        for _ in range(max(args.steps, 1)):
            matvec()
        src = torch.empty(32 * 2 ** 20, dtype=torch.float64, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] profile-only run done (copy of 268435456 bytes)")
        return

    # ---- fp64: throughput (back to back) and latency (each step synchronized) ----
    for _ in range(args.warmup):
        solve64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve64()
    torch.cuda.synchronize()
    fp64_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    lat64 = []
    for _ in range(args.steps):
        s = time.perf_counter()
        solve64()
        torch.cuda.synchronize()
        lat64.append((time.perf_counter() - s) * 1e3)
    log(f"[bench] fp64 solve: {fp64_ms:.3f} ms/step, p50 {np.percentile(lat64, 50):.3f} ms")

    # ---- mixed: every step refines to 1e-10 (synchronous) ----
    reps = []
    for _ in range(args.warmup):
        solve_mixed()
    torch.cuda.synchronize()
    latm = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        reps.append(solve_mixed())
        torch.cuda.synchronize()
        latm.append((time.perf_counter() - s) * 1e3)
    mixed_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    iters = [r["iterations"] for r in reps]
    worst_res = max(r["rel_residual"] for r in reps)
    log(f"[bench] mixed solve: {mixed_ms:.3f} ms/step, refinement steps {min(iters)}-{max(iters)}, "
        f"worst residual {worst_res:.2e}")

    # ---- the same mixed solve with GMRES-IR (each correction by Float64 GMRES, FP32 factor as
    # preconditioner): a robustness path for matrices where plain refinement stalls ----
    def solve_gmres():
        return F32.solve_dev(nrhs, Bt.data_ptr(), ld, X2t.data_ptr(), ld, tol=1e-10, max_iter=10, method="gmres")

    for _ in range(args.warmup):
        solve_gmres()
    torch.cuda.synchronize()
    greps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        greps.append(solve_gmres())
    torch.cuda.synchronize()
    gmres_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gmres_c3 = {"ms_per_step": round(gmres_ms, 4), "solves_per_s": round(1e3 / gmres_ms, 3),
                "outer_steps": [min(r["iterations"] for r in greps), max(r["iterations"] for r in greps)],
                "worst_rel_residual": max(r["rel_residual"] for r in greps)}
    log(f"[bench] mixed solve, GMRES-IR: {gmres_ms:.3f} ms/step, outer steps {gmres_c3['outer_steps']}")

    # ---- out of the Infinity Cache: 4 distinct factorizations of the same matrix (4 x 190 MB of
    # operators, more than the 256 MB cache) solved round robin, so every solve streams from HBM ----
    rot = [F64] + [A.ulvfactor("fp64") for _ in range(3)]
    Xrot = [torch.zeros_like(Bt) for _ in rot]
    torch.cuda.synchronize()

    def solve_rot(i):
        rot[i % 4].solve_dev(nrhs, Bt.data_ptr(), ld, Xrot[i % 4].data_ptr(), ld)

    for i in range(4 * max(args.warmup, 1)):
        solve_rot(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(4 * args.steps):
        solve_rot(i)
    torch.cuda.synchronize()
    rot_ms = (time.perf_counter() - t0) * 1e3 / (4 * args.steps)
    rot_bytes = sum(solve_kernel_model(nodes, F64, kname, lev, nrhs, 8, 8)[0]
                    for kname, levels in (("k_ulv_gup", range(summ["max_depth"])), ("k_ulv_gdown", range(1, summ["max_depth"] + 1)))
                    for lev in levels)
    log(f"[bench] fp64 solve over 4 rotating factorizations: {rot_ms:.3f} ms/solve, "
        f"{rot_bytes / rot_ms / 1e6:.1f} GB/s algorithmic (sweeps, root excluded)")
    for f in rot[1:]:
        f.close()
    del Xrot

    # ---- C3's right-hand-side counts: 1 and 128 columns, FP64 and mixed ----
    rhs_counts = {}
    for r in (1, 128):
        Br = torch.from_numpy(np.ascontiguousarray(np.random.default_rng(2).standard_normal((r, n)))).to(dev)
        Xr64 = torch.zeros_like(Br)
        Xrm = torch.zeros_like(Br)
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            F64.solve_dev(r, Br.data_ptr(), ld, Xr64.data_ptr(), ld)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            F64.solve_dev(r, Br.data_ptr(), ld, Xr64.data_ptr(), ld)
        torch.cuda.synchronize()
        r64 = (time.perf_counter() - t0) * 1e3 / args.steps
        # the same FP64 solve as a one-request plan after the autotuner (hss_plan_tune)
        tplan = hip.Plan([F64], [r], [Br.data_ptr()], [ld], [Xr64.data_ptr()], [ld])
        tplan.tune(reps=3)
        for _ in range(args.warmup):
            tplan.execute()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tplan.execute()
        torch.cuda.synchronize()
        r64t = (time.perf_counter() - t0) * 1e3 / args.steps
        tplan.close()
        its = []
        t0 = time.perf_counter()
        for _ in range(max(args.steps // 2, 1)):
            its.append(F32.solve_dev(r, Br.data_ptr(), ld, Xrm.data_ptr(), ld, tol=1e-10, max_iter=10)["iterations"])
        torch.cuda.synchronize()
        rm = (time.perf_counter() - t0) * 1e3 / max(args.steps // 2, 1)
        Bh = Br.cpu().numpy().T
        Xm = Xrm.cpu().numpy().T
        res = (A @ Xm) - Bh  # the library's Float64 matvec
        worst = float(np.max(np.linalg.norm(res, axis=0) / np.linalg.norm(Bh, axis=0)))
        rhs_counts[str(r)] = {"fp64_ms": round(r64, 4), "fp64_tuned_plan_ms": round(r64t, 4), "mixed_ms": round(rm, 4),
                              "mixed_refine_steps": [int(min(its)), int(max(its))],
                              "mixed_worst_rel_residual": worst}
        log(f"[bench] {r} RHS: fp64 {r64:.3f} ms (tuned plan {r64t:.3f}), mixed {rm:.3f} ms ({min(its)}-{max(its)} steps, "
            f"residual {worst:.2e})")
        del Br, Xr64, Xrm

    # ---- matvec ----
    for _ in range(args.warmup):
        matvec()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        matvec()
    torch.cuda.synchronize()
    mv_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    mv_bytes, mv_flops = matvec_bytes(nodes, nrhs)
    log(f"[bench] matvec: {mv_ms:.3f} ms, {mv_bytes / mv_ms / 1e6:.1f} GB/s algorithmic")

    # ---- adjoint (hssmatrix.jl:165-171): A' x and A' \ b, which X A (matmul.jl:14) and
    # X / A (hssmatrix.jl:236) run, on the same kernels and buffers ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    At = A.adjoint()
    adj_create_ms = (time.perf_counter() - t0) * 1e3

    def matvec_t():
        At.matmul_dev(nrhs, 1.0, Bt.data_ptr(), ld, 0.0, Ct.data_ptr(), ld)

    for _ in range(args.warmup):
        matvec_t()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        matvec_t()
    torch.cuda.synchronize()
    adj_mv_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    adj_mv_bytes, _ = matvec_bytes(At.nodes, nrhs)
    Ft = At.ulvfactor("fp64")
    Xa = torch.zeros_like(Bt)
    torch.cuda.synchronize()

    def solve_t():
        Ft.solve_dev(nrhs, Bt.data_ptr(), ld, Xa.data_ptr(), ld)

    for _ in range(args.warmup):
        solve_t()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve_t()
    torch.cuda.synchronize()
    adj_solve_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    Xa_h = Xa.cpu().numpy().T
    Ra = (At @ Xa_h) - B_host
    adj_res = float(np.max(np.linalg.norm(Ra, axis=0) / np.linalg.norm(B_host, axis=0)))
    Ft.close()
    At.close()
    del Xa
    log(f"[bench] adjoint: created in {adj_create_ms:.1f} ms; matvec {adj_mv_ms:.3f} ms "
        f"({adj_mv_bytes / adj_mv_ms / 1e6:.1f} GB/s algorithmic); fp64 solve {adj_solve_ms:.3f} ms, "
        f"residual {adj_res:.2e}")
    # ---- generator file (hss_mat_save / hss_mat_load) of the same matrix ----
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        gpath = pathlib.Path(td) / "a.hssg"
        t0 = time.perf_counter()
        A.save(gpath)
        gen_save_ms = (time.perf_counter() - t0) * 1e3
        gen_bytes = gpath.stat().st_size
        t0 = time.perf_counter()
        Ag = hip.HipHssMatrix.load(gpath)
        gen_load_ms = (time.perf_counter() - t0) * 1e3
        Cg = torch.zeros_like(Bt)
        torch.cuda.synchronize()
        Ag.matmul_dev(nrhs, 1.0, Bt.data_ptr(), ld, 0.0, Cg.data_ptr(), ld)
        matvec()
        torch.cuda.synchronize()
        gen_same = bool(torch.equal(Cg, Ct))
        Ag.close()
        del Cg
    log(f"[bench] generator file: {gen_bytes / 1e6:.1f} MB, save {gen_save_ms:.1f} ms, load {gen_load_ms:.1f} ms, "
        f"loaded matrix's matvec identical: {gen_same}")

    # ---- HSS x HSS product (matmul.jl:65-118): C = A A, ranks doubled ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C = A @ A
    prod_ms = (time.perf_counter() - t0) * 1e3
    hip.prof_enable(True)
    hip.prof_reset()
    C2 = A @ A
    prod_dev_ms = sum(r["ms"] for r in hip.prof_records() if r["kernel"] == "k_gemm_tasks")
    hip.prof_enable(False)
    C2.close()
    Cp = torch.zeros_like(Bt)
    torch.cuda.synchronize()

    def matvec_c():
        C.matmul_dev(nrhs, 1.0, Bt.data_ptr(), ld, 0.0, Cp.data_ptr(), ld)

    for _ in range(args.warmup):
        matvec_c()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        matvec_c()
    torch.cuda.synchronize()
    prod_mv_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    Yc = Cp.cpu().numpy().T
    Yaa = A @ (A @ B_host)
    prod_rel = float(np.linalg.norm(Yc - Yaa) / np.linalg.norm(Yaa))
    # recompression of the product (ranks doubled) at the compression tolerance
    hssrank = lambda nd: int(max(nd["ku"][1:].max(), nd["kv"][1:].max()))
    t0 = time.perf_counter()
    Crec = C.recompress(TOL, TOL)
    rec_ms = (time.perf_counter() - t0) * 1e3
    rec = {"ms": round(rec_ms, 1), "hssrank_before": hssrank(C.nodes), "hssrank_after": hssrank(Crec.nodes),
           "hssrank_A": hssrank(nodes),
           "rel_diff_vs_A_A_x": float(np.linalg.norm(Crec @ B_host[:, :4] - Yaa[:, :4]) /
                                      np.linalg.norm(Yaa[:, :4]))}
    Crec.close()
    C.close()
    del Cp
    log(f"[bench] HSS product A A: {prod_ms:.1f} ms wall ({prod_dev_ms:.2f} ms of device tasks); its matvec "
        f"{prod_mv_ms:.3f} ms; vs A (A X): {prod_rel:.1e}; recompressed {rec}")

    # ---- per-kernel timing (HIP events around each launch on the library stream) ----
    prof_reps = 3
    recs_solve = profile_launches(hip, solve64, prof_reps)
    tot_s, byk_s = summarize(recs_solve, prof_reps)
    recs_mv = profile_launches(hip, matvec, prof_reps)
    tot_m, byk_m = summarize(recs_mv, prof_reps)
    recs_mx = profile_launches(hip, solve_mixed, 1)
    _, byk_x = summarize(recs_mx, 1)
    fact_recs = profile_launches(hip, lambda: A.ulvfactor("fp64").close(), 1)
    _, byk_f = summarize(fact_recs, 1)
    fact_levels = {f"{r['kernel']}@{r['level']}": round(r["ms"], 4) for r in fact_recs}
    log(f"[bench] fp64 factor per kernel (ms): { {k: round(v, 4) for k, v in byk_f.items()} }")
    log(f"[bench] fp64 factor per launch (ms): {fact_levels}")
    log(f"[bench] fp64 solve per kernel (ms/solve): { {k: round(v, 4) for k, v in byk_s.items()} }")
    log(f"[bench] matvec per kernel (ms): { {k: round(v, 4) for k, v in byk_m.items()} }")
    log(f"[bench] mixed solve per kernel (ms/solve): { {k: round(v, 4) for k, v in byk_x.items()} }")
    # the busiest launch of the fp64 solve sweeps (largest time per solve)
    cand = [(key, v) for key, v in tot_s.items() if key[0].startswith(("k_ulv_gup", "k_ulv_gdown"))]
    (dk, dlev), (dms, dcnt, dnw) = max(cand, key=lambda kv: kv[1][0])
    avg_ms = dms / dcnt
    # roofline on the reference-form bytes of the launch (QL/LQ reflectors, L1, L2a, V_top,
    # couplings, inputs and outputs); the explicit operators this library streams instead
    # are reported beside them as operator_bytes
    dbytes, dflops = solve_ref_model(nodes, F64, dk, dlev, nrhs, 8, 8)
    obytes, _ = solve_kernel_model(nodes, F64, dk, dlev, nrhs, 8, 8)
    achieved = dbytes / avg_ms / 1e6
    sv_bytes, sv_F = solve_ref_total(nodes, F64, nrhs)
    log(f"[bench] dominant solve kernel {dk} level {dlev}: {avg_ms:.4f} ms/launch, "
        f"{dbytes/1e6:.2f} MB (reference form; {obytes/1e6:.2f} MB of explicit operators) -> {achieved:.1f} GB/s, "
        f"{dflops/avg_ms/1e9:.2f} TFLOP/s; whole solve {sv_bytes/1e6:.1f} MB (F = {sv_F} entries) in "
        f"{fp64_ms:.4f} ms -> {sv_bytes / fp64_ms / 1e6:.1f} GB/s")
    pname, targs, pgrid = sweep_launch(nodes, F64, dk, dlev, int(dnw), nrhs)
    traffic = profiled_traffic(pname, targs, pgrid, 1)
    mvleaf = [(key, v) for key, v in tot_m.items() if key[0].startswith("k_mv_leaf")]
    mv_leaf_ms = mvleaf[0][1][0] / mvleaf[0][1][1]
    mv_leaf_bytes = matvec_leaf_bytes(nodes, nrhs)

    # ---- residuals (the library's Float64 matvec), then the CPU baseline, whose solution is
    # also the parity reference ----
    X64 = Xt.cpu().numpy().T
    Xmx = X2t.cpu().numpy().T
    parity = {}
    cpu = None
    R = (A @ X64) - B_host
    parity["fp64_rel_residual"] = float(np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B_host, axis=0)))
    R = (A @ Xmx) - B_host
    parity["mixed_rel_residual"] = float(np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B_host, axis=0)))
    if not args.no_cpu:
        from oracle import oracle as O
        from threadpoolctl import threadpool_info, threadpool_limits
        with threadpool_limits(1):
            t0 = time.perf_counter()
            Xo = O.ulvfactsolve(nodes, blob, B_host)
            t1 = time.perf_counter() - t0
        parity["fp64_vs_oracle_rel"] = float(np.linalg.norm(X64 - Xo) / np.linalg.norm(Xo))
        parity["mixed_vs_oracle_rel"] = float(np.linalg.norm(Xmx - Xo) / np.linalg.norm(Xo))
        times = [t1]
        with threadpool_limits(1):
            for _ in range(max(args.cpu_reps - 1, 0)):
                t0 = time.perf_counter()
                O.ulvfactsolve(nodes, blob, B_host)
                times.append(time.perf_counter() - t0)
        ncores = os.cpu_count() or 1
        with threadpool_limits(ncores):
            t0 = time.perf_counter()
            O.ulvfactsolve(nodes, blob, B_host)
            tall = time.perf_counter() - t0
        blas = [i for i in threadpool_info() if i.get("internal_api") == "openblas"]
        model = "?"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu_fs = float(np.mean(times))
        cpu = {
            "value": round(1.0 / float(np.mean(times)), 4),
            "unit": "solves/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"{len(times)} calls of the oracle's ulvfactsolve (C restatement of ulvfactor.jl:10-107 "
                       f"over OpenBLAS {blas[0].get('version') if blas else '?'} via scipy), n={n}, {nrhs} RHS, "
                       f"1 BLAS thread; each call factors and solves as the reference's hssA\\B does"),
            "all_threads": {"threads": ncores, "seconds_per_solve": round(tall, 3)},
            "seconds_per_solve": round(cpu_fs, 3),
            "cpu_model": model,
            # like for like: the reference's hssA\B factors and solves on every call, so the GPU's
            # factor + solve is what the CPU's call compares with
            "gpu_factor_plus_solve_per_s": round(1e3 / (f64_ms + fp64_ms), 3),
            "gpu_vs_cpu_factor_plus_solve": round(cpu_fs * 1e3 / (f64_ms + fp64_ms), 1),
        }
        log(f"[bench] cpu baseline {cpu}")
    log(f"[bench] parity {parity}")

    construct = None if args.no_construct else construction(hip, W, torch, log=log)
    batch = None if args.no_batch else batched(hip, W, torch, log=log)
    serve = None if args.no_serve else serving(hip, W, torch, log=log)

    value = 1e3 / fp64_ms
    if parity["fp64_rel_residual"] > 1e-10:
        log(f"[bench] WARNING: fp64 residual {parity['fp64_rel_residual']:.2e} above 1e-10")
    line = {
        "metric": "ULV solves/sec + p50/p99 latency, mixed-prec vs FP64 GPU at 1e-10 residual",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp64_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: kernel matrices on seeded uniform points, compressed on the device (log|x_i-x_j| "
                "n=65536 seeds 0/1; 256 Cauchy instances seed 3; 32 log-kernel instances seed 4)",
        "config": {
            "workload": f"HSS n={n} log kernel, leaf {LEAF}, rtol {TOL:g}, {summ['leaves']} leaves, ranks "
                        f"{summ['rank_min']}-{summ['rank_max']}, {nrhs} RHS, resident Float64 ULV factor "
                        f"(relative residual <= 1e-10)",
            "nrhs": nrhs,
            "n": n,
        },
        "latency_ms": {"p50": round(float(np.percentile(lat64, 50)), 4),
                       "p99": round(float(np.percentile(lat64, 99)), 4)},
        "mixed_c3": {"solves_per_s": round(1e3 / mixed_ms, 3), "ms_per_step": round(mixed_ms, 4),
                     "latency_ms": {"p50": round(float(np.percentile(latm, 50)), 4),
                                    "p99": round(float(np.percentile(latm, 99)), 4)},
                     "refine_steps": [min(iters), max(iters)],
                     "worst_rel_residual": worst_res, "factor_ms": round(f32_ms, 3),
                     "gmres_ir": gmres_c3},
        "fp64": {"solves_per_s": round(1e3 / fp64_ms, 3), "ms_per_step": round(fp64_ms, 4),
                 "latency_ms": {"p50": round(float(np.percentile(lat64, 50)), 4),
                                "p99": round(float(np.percentile(lat64, 99)), 4)},
                 "factor_ms": round(f64_ms, 3), "factor_plus_solve_per_s": round(1e3 / (f64_ms + fp64_ms), 3),
                 "factor_gflop_lapack": round(fflops / 1e9, 3),
                 "factor_tflops": round(fflops / f64_ms / 1e9, 3),
                 "factor_frac_of_fp64_peak": round(fflops / f64_ms / 1e9 / FP64_PEAK_TFLOPS, 4)},
        "fp64_rotating_4": {"note": "4 distinct resident factorizations (760 MB of operators, more than the 256 MB "
                                    "Infinity Cache) solved round robin: every solve streams its operators from HBM",
                            "ms_per_solve": round(rot_ms, 4), "solves_per_s": round(1e3 / rot_ms, 3),
                            "gbps_algorithmic": round(rot_bytes / rot_ms / 1e6, 1),
                            "frac_of_hbm_peak": round(rot_bytes / rot_ms / 1e6 / HBM_PEAK_GBS, 4)},
        "rhs_counts_c3": rhs_counts,
        "mixed_vs_fp64_solves": round(fp64_ms / mixed_ms, 3),
        "matvec": {"ms": round(mv_ms, 4), "gbps_algorithmic": round(mv_bytes / mv_ms / 1e6, 1),
                   "tflops": round(mv_flops / mv_ms / 1e9, 3),
                   "leaf_kernel": {"ms": round(mv_leaf_ms, 4),
                                   "gbps": round(mv_leaf_bytes / mv_leaf_ms / 1e6, 1)}},
        "adjoint": {"create_ms": round(adj_create_ms, 2), "matvec_ms": round(adj_mv_ms, 4),
                    "matvec_gbps_algorithmic": round(adj_mv_bytes / adj_mv_ms / 1e6, 1),
                    "fp64_solve_ms": round(adj_solve_ms, 4), "fp64_solve_rel_residual": adj_res},
        "hss_product": {"case": "A A (ranks doubled)", "wall_ms": round(prod_ms, 1),
                        "device_task_ms": round(prod_dev_ms, 3), "product_matvec_ms": round(prod_mv_ms, 4),
                        "rel_diff_vs_A_A_x": prod_rel, "recompressed": rec},
        "construction": construct,
        "generator_file": {"bytes": int(gen_bytes), "save_ms": round(gen_save_ms, 1),
                           "load_ms": round(gen_load_ms, 1), "loaded_matvec_identical": gen_same},
        "roofline": {"kernel": f"{dk} (level {dlev}, {dnw} nodes)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/hss_c2_n65536_hbm_counters.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated; looked up for this launch shape, not measured in this run)",
                     "algorithmic_bytes": int(dbytes), "operator_bytes": int(obytes),
                     "avg_launch_ms": round(avg_ms, 5),
                     "whole_solve": {"algorithmic_bytes": int(sv_bytes), "factor_entries": int(sv_F),
                                     "ms": round(fp64_ms, 4),
                                     "achieved": round(sv_bytes / fp64_ms / 1e6, 1),
                                     "frac": round(sv_bytes / fp64_ms / 1e6 / HBM_PEAK_GBS, 4)}},
        "per_kernel_ms": {"fp64_factor": {k: round(v, 4) for k, v in byk_f.items()},
                          "fp64_solve": {k: round(v, 4) for k, v in byk_s.items()},
                          "mixed_solve": {k: round(v, 4) for k, v in byk_x.items()},
                          "matvec": {k: round(v, 4) for k, v in byk_m.items()}},
        "parity": parity,
        "batched_c4": batch,
        "serving_c5": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
