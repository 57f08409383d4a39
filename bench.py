#!/usr/bin/env python3
"""Benchmark of the exact mod-p matrix engine on one MI355X.

Headline workload (one "step"): C = A * B mod 2147483629 with A, B 8192 x 8192 uint32
residues (the 8192^3 GEMM configuration), computed by the library's default route (the residue
route: ten int8 MFMA products per multiply-add modulo small coprime moduli and a CRT pass, in place
of the limb kernels' sixteen; every limb kernel is timed beside it on the same inputs); inputs are
generated on the device (SplitMix64 / SeedFor, salts 0x01 / 0x02) and stay resident in
HBM.  The metric is exact mod-p multiply-adds per second (n^3 per step).  The same run
also times the plain VALU kernel on the same inputs (limb vs plain), checks the two
results are identical (and spot-checks entries against exact integer arithmetic), and
times the batched echelon workload (4096 sparse jobs mod 1000003) as one persistent
launch versus one launch per job (jobs / s).

Prints ONE JSON line.  `--gpus N` with N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_C3 = 8192
PEAK_INT8_TOPS = 5000.0      # MI355X dense INT8 MFMA, tera-ops/s
LIMB_KERNELS = {1: "limb_kernel<%d>", 2: "limb2_kernel<%d>", 3: "limb3_kernel<%d>", 4: "limb4_kernel<%d>",
                5: "limb5_kernel<%d>", 6: "res_gemm_kernel<4, 2>"}
RESIDUE_TILE = 6             # the residue route: one int8 product per small modulus, CRT into C


def kernel_name(tile, limbs):
    k = LIMB_KERNELS[tile]
    return k % limbs if "%d" in k else k


def int8_products(E, tile, p, k, limbs):
    """int8 MFMA products per multiply-add: limbs^2 for the limb kernels, the moduli count for the residues."""
    return E.residue_moduli(p, k) if tile == RESIDUE_TILE else limbs * limbs
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_C3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--tile", type=int, default=0,
                    help="limb kernel: 0 default, 1 register-staged LDS copy, 2 direct-to-LDS ring, "
                         "3 ring with pipelined LDS reads, 4 the same on eight waves, 5 96 x 128 tiles on twelve waves")
    ap.add_argument("--no-cpu", action="store_true", help="do not time the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="do not time the batched echelon workload")
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--no-lu", action="store_true", help="do not time the 4096^2 rank-3000 RREF")
    ap.add_argument("--no-lu-cpu", action="store_true", help="do not time the reference's CPU RREF on it")
    ap.add_argument("--no-mixed", action="store_true", help="do not run the 20000-job mixed stream")
    ap.add_argument("--mixed-jobs", type=int, default=20000)
    ap.add_argument("--no-howell", action="store_true", help="do not time the Howell batch")
    ap.add_argument("--howell-jobs", type=int, default=6666)
    ap.add_argument("--no-ring2k", action="store_true", help="do not time the Z/2^32 LU on the device session")
    ap.add_argument("--no-u64", action="store_true", help="do not time the 64-bit-residue GEMM")
    ap.add_argument("--no-level1", action="store_true", help="do not time the HBM-bound level-1/2 kernels")
    ap.add_argument("--no-f2", action="store_true", help="do not time the F_2 RREF and product")
    ap.add_argument("--no-splitk", action="store_true", help="do not time the split-K product")
    ap.add_argument("--no-smith", action="store_true", help="do not time the Smith valuations over Z/2^32")
    ap.add_argument("--no-howell-ell", action="store_true", help="do not time the Howell batch over Z/3^10")
    ap.add_argument("--no-f3", action="store_true", help="do not time the F_3 RREF and product")
    ap.add_argument("--no-smith-cpu", action="store_true", help="do not time the oracle's Smith sample on the host")
    ap.add_argument("--only", choices=["batch", "lu", "mixed", "cpu", "howell", "ring2k", "level1", "u64", "f2",
                                       "splitk", "smith", "howell_ell", "f3", "mfma", "operands"],
                    help="run one secondary workload alone (for profiling) and print its record")
    ap.add_argument("--pmc-mode", action="store_true",
                    help="for counter collection: one 1 GiB device copy (the traffic calibration), then one limb "
                         "GEMM of the headline shape; no timing, no JSON")
    ap.add_argument("--pmc-reps", type=int, default=1, help="--pmc-mode: the number of limb GEMMs (later ones run "
                                                            "at the clock of a warm GPU)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(n, limbs, kernel):
    """HBM bytes per main-kernel launch from the committed counter passes (profiles/*_traffic.json), if one
    covers this shape and kernel: (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor) in bytes."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
        with open(path) as f:
            t = json.load(f)
        if t.get("n") == n and t.get("limbs") == limbs and t.get("kernel") == kernel:
            return t["traffic_bytes"]
    return None


def gemm_timed(E, ctx, A, B, Cm, variant, steps, warmup, ws, tile=0):
    """Times `steps` GEMMs on one stream; returns (wall ms per step, main-kernel ms per step)."""
    import torch
    from hacs_amd import _lib
    s = torch.cuda.current_stream()
    h = s.cuda_stream
    L = _lib.lib()
    evs = []
    for _ in range(steps):
        a, b = C.c_void_p(), C.c_void_p()
        _lib.check(L.hx_event_create(C.byref(a), 1))
        _lib.check(L.hx_event_create(C.byref(b), 1))
        evs.append((a, b))
    o = _lib.hx_gemm_opts()
    o.variant = variant
    o.tile = tile
    if ws is not None:
        o.workspace = C.c_void_p(ws.data_ptr())
        o.workspace_bytes = ws.numel()
    for _ in range(warmup):
        _lib.check(L.hx_gemm_u32(C.byref(ctx.c), E.view(A), E.view(B), E.view(Cm), 0, h, C.byref(o)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        o.main_begin, o.main_end = evs[i]
        _lib.check(L.hx_gemm_u32(C.byref(ctx.c), E.view(A), E.view(B), E.view(Cm), 0, h, C.byref(o)))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    tot = 0.0
    for a, b in evs:
        ms = C.c_float()
        _lib.check(L.hx_event_elapsed_ms(a, b, C.byref(ms)))
        tot += ms.value
        L.hx_event_destroy(a)
        L.hx_event_destroy(b)
    return wall, tot / steps


def spot_check(C_host, A_dev, B_dev, p, count=256, seed=7):
    """count entries of C against exact Python-integer dot products."""
    import torch
    rng = np.random.default_rng(seed)
    n = C_host.shape[0]
    rows = rng.integers(0, n, count)
    cols = rng.integers(0, n, count)
    Ar = A_dev[torch.as_tensor(rows, device="cuda")].cpu().numpy().view(np.uint32).astype(object)
    Bc = B_dev[:, torch.as_tensor(cols, device="cuda")].cpu().numpy().view(np.uint32).astype(object)
    for t in range(count):
        exact = int(np.dot(Ar[t], Bc[:, t])) % p
        if exact != int(C_host[rows[t], cols[t]]):
            return False
    return True


def batched_echelon(E, njobs):
    """C4: jobs / s of one persistent launch vs one launch per job."""
    import torch
    from hacs_amd import workloads as W
    p = W.P_C4
    ctx = E.ModCtx(p)
    shapes = W.c4_shapes(njobs)
    src = [E.to_device(W.sparse_job(j, int(m), int(n), p)) for j, (m, n) in enumerate(shapes)]
    work = [s.clone() for s in src]
    batch = E.RrefBatch(ctx, work)
    batch.run()
    batch.status()
    torch.cuda.synchronize()
    reps = 3
    t_batch = []
    for _ in range(reps):
        for w_, s_ in zip(work, src):
            w_.copy_(s_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run()
        torch.cuda.synchronize()
        t_batch.append(time.perf_counter() - t0)
    eager = [w_.clone() for w_ in work]
    ranks_eager = [r.rank for r in batch.results()]
    # the same launch captured in a hipGraph on a side stream, replayed from resident inputs
    import ctypes as C
    from hacs_amd import _lib
    L = _lib.lib()
    gs = torch.cuda.Stream()
    g = C.c_void_p()
    t_graph = []
    with torch.cuda.stream(gs):
        h = gs.cuda_stream
        _lib.check(L.hx_graph_begin(h))
        batch.run(stream=h)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for _ in range(reps):
            for w_, s_ in zip(work, src):
                w_.copy_(s_)
            gs.synchronize()
            t0 = time.perf_counter()
            _lib.check(L.hx_graph_launch(g, h))
            gs.synchronize()
            t_graph.append(time.perf_counter() - t0)
        _lib.check(L.hx_graph_destroy(g))
    graph_same = all(torch.equal(a, b) for a, b in zip(work, eager)) and \
        [r.rank for r in batch.results()] == ranks_eager
    del eager
    singles = [E.RrefBatch(ctx, [w_]) for w_ in work]
    for w_, s_ in zip(work, src):
        w_.copy_(s_)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sb in singles:
        sb.run()
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    # the same per-job launches dealt round-robin over 8 streams, so independent jobs can overlap
    streams = [torch.cuda.Stream() for _ in range(8)]
    for w_, s_ in zip(work, src):
        w_.copy_(s_)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, sb in enumerate(singles):
        sb.run(stream=streams[i % 8].cuda_stream)
    torch.cuda.synchronize()
    t_single8 = time.perf_counter() - t0
    del singles
    ranks = [r.rank for r in batch.results()]
    # the other launch shape (the automatic rule picks one; both give the same bytes)
    from hacs_amd import _lib
    info = batch.info()
    other = _lib.HX_RREF_BATCH_TWO_PER_CU if info["workgroups_per_cu"] == 1 else _lib.HX_RREF_BATCH_ONE_PER_CU
    alt = E.RrefBatch(ctx, work, mode=other)
    t_alt = []
    for _ in range(reps):
        for w_, s_ in zip(work, src):
            w_.copy_(s_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        alt.run()
        torch.cuda.synchronize()
        t_alt.append(time.perf_counter() - t0)
    alt_same = [r.rank for r in alt.results()] == ranks
    alt_info = alt.info()
    del alt
    work_macs = float(sum((int(m) - 1) * (r * int(n) - r * (r - 1) // 2) for (m, n), r in zip(shapes, ranks)))
    return {
        "jobs": njobs,
        "jobs_per_s": njobs / min(t_batch),
        "per_job_launch_jobs_per_s": njobs / t_single,
        "speedup_vs_per_job": t_single / min(t_batch),
        "per_job_launch_8_streams_jobs_per_s": njobs / t_single8,
        "speedup_vs_per_job_8_streams": t_single8 / min(t_batch),
        "launch": info,
        "other_launch": {"launch": alt_info, "jobs_per_s": njobs / min(t_alt), "ms": min(t_alt) * 1e3,
                         "same_ranks": alt_same},
        "graph_jobs_per_s": njobs / min(t_graph),
        "graph_equals_eager": graph_same,
        "elim_work_macs": work_macs,
        "ms": min(t_batch) * 1e3,
    }


def howell_batch(E, njobs):
    """C5(c): Howell forms over Z/2^16 of the stream's Howell jobs (dense uniform [0, 2^16), n
    log-uniform in [32, 512]) -- jobs/s of one persistent launch vs one launch per job, and the
    oracle's Howell on the host's cores over the same jobs, whose outputs are compared with the
    device's."""
    import torch
    from hacs_amd import workloads as W
    k = 16
    shapes = [n for kind, n in W.c5_stream(3 * njobs + 3, salt=0x51) if kind == "howell"][:njobs]
    seeds = [W.seed_for(n, n, 1 << k, 0x94) ^ j for j, n in enumerate(shapes)]
    mats = [E.random_entries(n, n, 1 << k, sd) for n, sd in zip(shapes, seeds)]
    b = E.HowellBatch(k, mats)
    b.run()
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(3):
        t0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    singles = [E.HowellBatch(k, [M]) for M in mats]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sb in singles:
        sb.run()
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    out = {"jobs": len(shapes), "k": k, "jobs_per_s": len(shapes) / best, "ms": best * 1e3,
           "per_job_launch_jobs_per_s": len(shapes) / t_single, "speedup_vs_per_job": t_single / best,
           "mean_n": float(np.mean(shapes))}
    # the oracle's Howell (hxo_howell_fast, restating Ring2k.h's blocked echelon and Sec. 13.5 pass) on
    # the host's threads, one job per thread, over all jobs; outputs compared with the device's
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    threads, nproc, model = host_cpu()
    inputs = [E.from_device(M) for M in mats]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        forms = list(ex.map(lambda A: O.howell_fast(k, A), inputs))
    t_cpu = time.perf_counter() - t0
    got = b.results()
    same = all(np.array_equal(g.H, H) and np.array_equal(g.pivots, piv) and np.array_equal(g.pivot_v, pv)
               for g, (H, piv, pv) in zip(got, forms))
    out["cpu_port"] = {"jobs": len(shapes), "seconds": t_cpu, "cores": threads, "nproc": nproc, "cpu_model": model,
                       "jobs_per_s": len(shapes) / t_cpu, "device_equals_port": bool(same),
                       "routine": "hxo_howell_fast, one job per thread"}
    return out


def howell_ell_leg(E, njobs=1000, ell=3, k=10):
    """Howell forms over Z/3^10 (elim::howell with a prime-power context) of the first C5(c) Howell shapes,
    entries uniform in [0, 3^10): jobs/s of one persistent launch, and the oracle's restatement of the
    reference's pass on the host's threads over the same jobs, compared with the device's."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from hacs_amd import workloads as W
    from oracle import oracle as O
    q = ell ** k
    shapes = [n for kind, n in W.c5_stream(3 * njobs + 3, salt=0x51) if kind == "howell"][:njobs]
    seeds = [W.seed_for(n, n, q, 0x95) ^ j for j, n in enumerate(shapes)]
    mats = [E.random_entries(n, n, q, sd) for n, sd in zip(shapes, seeds)]
    b = E.HowellBatch(k, mats, ell=ell)
    b.run()
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(3):
        t0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    out = {"jobs": len(shapes), "ell": ell, "k": k, "jobs_per_s": len(shapes) / best, "ms": best * 1e3,
           "mean_n": float(np.mean(shapes))}
    threads, nproc, model = host_cpu()
    inputs = [E.from_device(M) for M in mats]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        forms = list(ex.map(lambda A: O.howell_ell(ell, k, A), inputs))
    t_cpu = time.perf_counter() - t0
    same = all(np.array_equal(g.H, H) and np.array_equal(g.pivots, piv) and np.array_equal(g.pivot_v, pv)
               for g, (H, piv, pv) in zip(b.results(), forms))
    out["cpu_port"] = {"jobs": len(shapes), "seconds": t_cpu, "cores": threads, "nproc": nproc, "cpu_model": model,
                       "jobs_per_s": len(shapes) / t_cpu, "device_equals_port": bool(same),
                       "routine": "hxo_howell_ell, one job per thread"}
    return out


def ring2k_lu(E, m=2048, panel=64, k=32):
    """The blocked LU over Z/2^32 (the device route of ring2k::echelon_blocked, hx_ring2k_echelon): an
    m x m matrix of uniform words with rows scaled by powers of two, the panels factored on the host,
    swaps / TRSM / trailing products on the device with the next panel fetched by lookahead.  Reports
    the seconds, the trailing wraparound multiply-adds per second and the route's phase split; the
    oracle's serial restatement of the host code times the same LU on one core for reference (its
    result is compared entrywise)."""
    import ctypes as C
    from hacs_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(5)
    A0 = rng.integers(0, 1 << 32, (m, m), dtype=np.uint64).astype(np.uint32)
    A0[::3] = (A0[::3] << np.uint32(3)).astype(np.uint32)
    piv = np.zeros(m, np.uint32)
    pv = np.zeros(m, np.uint32)
    rank = C.c_uint32(0)
    best, A = 1e30, None
    for rep in range(2):
        A = A0.copy()
        _lib.check(L.hx_ring2k_stats_reset())
        t0 = time.perf_counter()
        _lib.check(L.hx_ring2k_echelon(4, k, m, m, A.ctypes.data_as(C.c_void_p), m, panel, m, 0, 0, C.byref(rank),
                                       piv.ctypes.data_as(_lib._U32P), pv.ctypes.data_as(_lib._U32P)))
        best = min(best, time.perf_counter() - t0)
    st = _lib.hx_ring2k_stats()
    _lib.check(L.hx_ring2k_stats_get(C.byref(st)))
    r = rank.value
    macs = float(sum((m - min(c + panel, m)) * (m - min(c + panel, m)) * min(panel, m - c) for c in range(0, r, panel)))
    out = {"m": m, "n": m, "panel": panel, "k": k, "seconds": best, "rank": r, "trailing_wraparound_macs": macs,
           "macs_per_s": macs / best, "panels": int(st.panels), "lookaheads": int(st.lookaheads),
           "phase_s": {f: getattr(st, "ns_" + f) * 1e-9 for f in ("open", "fetch", "panel", "step", "host", "close")}}
    from oracle import oracle as O
    t0 = time.perf_counter()
    R_o, r_o, piv_o, pv_o = O.r2_echelon(k, A0, panel)
    t_o = time.perf_counter() - t0
    out["cpu_port"] = {"seconds": t_o, "cores": 1, "routine": "hxo_r2_echelon32 (the host code's serial route)",
                       "device_equals_port": bool(r_o == r and np.array_equal(piv_o, piv[:r]) and
                                                  np.array_equal(pv_o, pv[:r]) and np.array_equal(R_o, A)),
                       "speedup": t_o / best}
    return out


def smith_leg(E, n=4096, k=32, cpu=True, n_cpu=2048):
    """The Smith valuations over Z/2^32 of an n x n matrix of uniform words (the workload of the reference's
    ring2k timing rung, Ring2kLadderValidation.cpp:1024-1033) on the device (hx_ring2k_smith: F_2 rank profile,
    odd-pivot LU on the session, halved Schur complement per level; host buffer in, PCIe included), min of 2.
    Beside it a bounded sample: the same on an n_cpu x n_cpu matrix, on the device and by the oracle's
    restatement of smith_naive on one host core, whose sigma is compared."""
    def matrix(size):
        rng = np.random.default_rng(0x7173 + size + k)
        return rng.integers(0, 1 << 32, (size, size), dtype=np.uint64).astype(np.uint32)

    def timed(A):
        best, res = 1e30, None
        for _ in range(2):
            t0 = time.perf_counter()
            res = E.ring2k_smith(A, k)
            best = min(best, time.perf_counter() - t0)
        return best, res

    best, res = timed(matrix(n))
    out = {"n": n, "k": k, "seconds": best, "rank": int(len(res.sigma)), "levels": int(len(res.level_ranks)),
           "sigma_tail": [int(x) for x in res.sigma[-8:]]}
    if cpu:
        from oracle import oracle as O
        A = matrix(n_cpu)
        t_dev, r_dev = timed(A)
        t0 = time.perf_counter()
        sig = O.smith(k, A)
        t_o = time.perf_counter() - t0
        out["cpu_port"] = {"n": n_cpu, "seconds": t_o, "device_seconds": t_dev, "cores": 1, "kind": "port",
                           "routine": "hxo_smith_naive (global minimal-valuation pivoting, one core)",
                           "device_equals_port": bool(np.array_equal(sig, r_dev.sigma)), "speedup": t_o / t_dev}
    return out


def gemm_u64_leg(E, n=4096, p=2305843009213693951, reps=3):
    """C = A B mod 2^61 - 1 on 64-bit residues (hx_gemm_u64), n^3 mod-p multiply-adds per second."""
    import torch
    from hacs_amd import workloads as W
    ctx = E.ModCtx64(p)
    rng = np.random.default_rng(9)
    A = E.to_device64((rng.integers(0, 1 << 62, (n, n), dtype=np.uint64) % np.uint64(p)))
    B = E.to_device64((rng.integers(0, 1 << 62, (n, n), dtype=np.uint64) % np.uint64(p)))
    Cm = torch.empty((n, n), dtype=torch.int64, device="cuda")
    E.gemm_u64(ctx, A, B, Cm)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        E.gemm_u64(ctx, A, B, Cm)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"n": n, "modulus": p, "ms": ms, "macs_per_s": float(n) ** 3 / (ms * 1e-3)}


def f2_leg(E, n=8192, rank=6000, reps=3):
    """F_2 on bit-packed rows: the in-place RREF of an n x n matrix of planted rank (rank random rows, the
    others sums of two of them, shuffled) and the product C = A B at n^3 bits; ms and bit
    multiply-adds per second (n^3 / t for the product)."""
    import torch
    rng = np.random.default_rng(82)
    A = np.empty((n, n), np.uint32)
    A[:rank] = rng.integers(0, 2, (rank, n))
    pairs = rng.integers(0, rank, (n - rank, 2))
    A[rank:] = A[pairs[:, 0]] ^ A[pairs[:, 1]]
    A = A[rng.permutation(n)]
    W = torch.from_numpy(E.pack_bits(A).view(np.int64).copy()).cuda()
    work = torch.empty_like(W)
    times = []
    for _ in range(reps + 1):
        work.copy_(W)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r, _ = E.f2_rref(work, n)
        times.append(time.perf_counter() - t)
    Bw = torch.from_numpy(E.pack_bits(rng.integers(0, 2, (n, n)).astype(np.uint32)).view(np.int64).copy()).cuda()
    Cw = torch.empty_like(W)
    E.f2_gemm(W, Bw, Cw, n, n)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        E.f2_gemm(W, Bw, Cw, n, n)
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / reps
    rref_s = min(times[1:])
    return {"n": n, "rank": int(r), "planted_rank": rank, "rref_ms": rref_s * 1e3,
            "rref_bit_ops_per_s": float(n) * n * rank / rref_s, "gemm_ms": gms,
            "gemm_bit_macs_per_s": float(n) ** 3 / (gms * 1e-3)}


def f3_leg(E, n=8192, rank=6000, reps=3):
    """F_3 on bit-sliced rows: the in-place RREF of an n x n matrix of planted rank (rank random rows, the
    others A[a] + 2 A[b] of two of them, shuffled) and the product C = A B at n^3 trits; ms and trit
    multiply-adds per second."""
    import torch
    rng = np.random.default_rng(84)
    A = np.empty((n, n), np.uint32)
    A[:rank] = rng.integers(0, 3, (rank, n))
    pairs = rng.integers(0, rank, (n - rank, 2))
    A[rank:] = (A[pairs[:, 0]] + 2 * A[pairs[:, 1]]) % 3
    A = A[rng.permutation(n)]
    W = torch.from_numpy(E.pack_trits(A).view(np.int64).copy()).cuda()
    work = torch.empty_like(W)
    times = []
    for _ in range(reps + 1):
        work.copy_(W)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r, _ = E.f3_rref(work, n)
        times.append(time.perf_counter() - t)
    Bw = torch.from_numpy(E.pack_trits(rng.integers(0, 3, (n, n)).astype(np.uint32)).view(np.int64).copy()).cuda()
    Cw = torch.empty_like(W)
    E.f3_gemm(W, Bw, Cw, n, n)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        E.f3_gemm(W, Bw, Cw, n, n)
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / reps
    rref_s = min(times[1:])
    return {"n": n, "rank": int(r), "planted_rank": rank, "rref_ms": rref_s * 1e3,
            "rref_trit_ops_per_s": float(n) * n * rank / rref_s, "gemm_ms": gms,
            "gemm_trit_macs_per_s": float(n) ** 3 / (gms * 1e-3)}


def splitk_leg(E, m=256, n=256, k=65536, reps=5):
    """A thin deep product mod 2147483629 (m x n x k = 256 x 256 x 65536): the plain kernel unsplit, the
    plain kernel split K ways by the automatic policy, and the limb route; ms each, same bytes."""
    import torch
    from hacs_amd import workloads as W
    p = W.P_C3
    ctx = E.ModCtx(p)
    A = E.random_entries(m, k, p, 0x51)
    B = E.random_entries(k, n, p, 0x52)
    out = {}
    ref = None
    for name, kw in [("plain", {"variant": E.HX_GEMM_PLAIN, "splits": 1}),
                     ("plain_split", {"variant": E.HX_GEMM_PLAIN, "splits": 0}),
                     ("limb", {"variant": E.HX_GEMM_LIMB})]:
        Cm = torch.empty((m, n), dtype=torch.int32, device="cuda")
        E.gemm(ctx, A, B, Cm, E.GemmOp.Set, **kw)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            E.gemm(ctx, A, B, Cm, E.GemmOp.Set, **kw)
        e1.record()
        torch.cuda.synchronize()
        out[name + "_ms"] = e0.elapsed_time(e1) / reps
        ref = Cm if ref is None else ref
        out[name + "_same_bytes"] = bool(torch.equal(Cm, ref))
    out.update({"shape": [m, n, k], "modulus": p, "splits": E.suggest_splits(m, n, k)})
    return out


def lu_c2(E, reps=3, cpu=True):
    """C2: RREF of the 4096 x 4096 rank-3000 matrix mod 2^31-1 (rank, pivots, pivot_rows, RREF)."""
    import torch
    from hacs_amd import workloads as W
    p = W.P_C2
    ctx = E.ModCtx(p)
    A = W.c2_device()
    R = torch.empty_like(A)
    res = E.rref(ctx, A, R)      # warm-up (also the parity-checked call in the tests)
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        res = E.rref(ctx, A, R)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    m, n = A.shape
    work = float((m - 1) * (res.rank * n - res.rank * (res.rank - 1) // 2))
    out = {"seconds": best, "rank": res.rank, "elim_work_macs": work, "elim_macs_per_s": work / best,
           "shape": [m, n], "modulus": p}
    # the same elimination through the capturable entry (caller workspace, results on the device), captured
    # once in a hipGraph and replayed: the chain's launches come from the graph, not from the host
    import ctypes as C
    from hacs_amd import _lib
    L = _lib.lib()
    mn = min(m, n)
    d_out = torch.zeros(1 + 2 * mn, dtype=torch.int32, device="cuda")
    ws = torch.empty(E.rref_workspace_bytes(m, n), dtype=torch.uint8, device="cuda")
    R2 = torch.empty_like(A)
    s = torch.cuda.Stream()
    g = C.c_void_p()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        E.rref_device(ctx, A, R2, d_out, ws, stream=h)          # eager once (creates the side streams)
        s.synchronize()
        _lib.check(L.hx_graph_begin(h))
        E.rref_device(ctx, A, R2, d_out, ws, stream=h)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        gbest = 1e30
        for _ in range(reps + 1):
            R2.fill_(0)
            s.synchronize()
            t0 = time.perf_counter()
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            gbest = min(gbest, time.perf_counter() - t0)
        _lib.check(L.hx_graph_destroy(g))
    gres = E.rref_device_result(d_out, m, n)
    out["graph_seconds"] = gbest
    out["graph_equals_eager"] = bool(gres.rank == res.rank and np.array_equal(gres.pivots, res.pivots) and
                                     np.array_equal(gres.pivot_rows, res.pivot_rows) and torch.equal(R2, R))
    if cpu:
        from oracle import oracle as O
        threads, nproc, model = host_cpu()
        Ah = E.from_device(A)
        t0 = time.perf_counter()
        r_o, piv_o, prow_o, R_o = O.rref_rule(p, Ah, True, threads)
        t_o = time.perf_counter() - t0
        Rd = E.from_device(R)
        same = r_o == res.rank and np.array_equal(piv_o, res.pivots) and np.array_equal(prow_o, res.pivot_rows) \
            and np.array_equal(R_o, Rd[:r_o]) and not Rd[r_o:].any()
        out["cpu_port"] = {"seconds": t_o, "cores": threads, "nproc": nproc, "cpu_model": model,
                           "routine": "hxo_rref_rule (the oracle's smallest-row-rule elimination, restating "
                                      "DenseElim.cpp's RREF conventions)",
                           "device_equals_port": bool(same), "speedup": t_o / best}
    return out


def mixed_stream(njobs):
    """C5: autotune a training sample exhaustively, train the MLP selector on the GPU, compare it with
    the autotuner's best and with the static policy on held-out jobs, then run the whole stream with
    the selector's choices (batched per kind and choice)."""
    import torch
    from hacs_amd import autotune as AT
    t0 = time.perf_counter()
    train = [AT.Job(k, n, i) for i, (k, n) in enumerate(AT.sample_jobs(240, 0x71))]
    held = [AT.Job(k, n, i) for i, (k, n) in enumerate(AT.sample_jobs(120, 0x72))]
    tuner = AT.Autotuner(reps=3)
    Tt, Th = tuner.tune(train), tuner.tune(held)
    for jb in train + held:
        jb.free()
    t_tune = time.perf_counter() - t0
    sel = AT.Selector(seed=0)
    t1 = time.perf_counter()
    loss = sel.fit([j.kind for j in train], [j.n for j in train], Tt, steps=3000)
    t_fit = time.perf_counter() - t1
    slots_h = sel.choose([j.kind for j in held], [j.n for j in held])
    stat_h = [AT.static_choice(j.kind, j.n) for j in held]
    best = float(np.nanmin(np.where(np.isfinite(Th), Th, np.nan), axis=1).sum())
    t_sel = float(sum(Th[i, s] for i, s in enumerate(slots_h)))
    t_stat = float(sum(Th[i, s] for i, s in enumerate(stat_h)))
    agree = float(np.mean([s == int(np.argmin(Th[i])) for i, s in enumerate(slots_h)]))
    # the whole stream, in memory-bounded chunks, with the selector's and with the static choices
    stream = AT.sample_jobs(njobs, 0x51)
    kinds = [k for k, _ in stream]
    ns = [n for _, n in stream]
    t2 = time.perf_counter()
    slots = sel.choose(kinds, ns)
    t_choose = time.perf_counter() - t2
    statics = [AT.static_choice(k, n) for k, n in stream]
    t_run = {"selector": 0.0, "static": 0.0}
    t_kind = {}
    # batched per kind (as a scheduler would), in chunks of at most 40 GB of inputs and outputs: the
    # GEMMs (~160 GB) take several chunks, the echelon and Howell jobs one each
    budget = 40 << 30
    chunks = []
    for kind in AT.KINDS:
        idx = [j for j in range(njobs) if stream[j][0] == kind]
        i = 0
        while i < len(idx):
            chunk, used = [], 0
            while i < len(idx) and (not chunk or used < budget):
                n = stream[idx[i]][1]
                used += (3 if kind == "gemm" else 2) * 4 * n * n
                chunk.append(idx[i])
                i += 1
            chunks.append(chunk)
    for chunk in chunks:
        jobs = [AT.Job(stream[j][0], stream[j][1], j, salt=0x90) for j in chunk]
        t_run["selector"] += AT.run_stream(jobs, [slots[j] for j in chunk])
        AT.run_stream(jobs, [slots[j] for j in chunk], per_kind=t_kind)
        t_run["static"] += AT.run_stream(jobs, [statics[j] for j in chunk])
        for jb in jobs:
            jb.free()
        torch.cuda.empty_cache()
    counts = {k: kinds.count(k) for k in AT.KINDS}
    # the selector study (moduli of 2, 3 and 4 limbs, n in [48, 640]): where the static rule errs
    study = AT.selector_study(n_train=160, n_held=80, reps=3, steps=3000)
    return {"jobs": njobs, "kinds": counts, "selector_study": study, "seconds": t_run["selector"], "jobs_per_s": njobs / t_run["selector"],
            "static_seconds": t_run["static"], "selector_gain_vs_static": t_run["static"] / t_run["selector"],
            "heldout_jobs": len(held), "heldout_selector_vs_best": t_sel / best,
            "heldout_static_vs_best": t_stat / best, "heldout_choice_agreement": agree,
            "train_jobs": len(train), "train_loss": loss, "tune_seconds": t_tune, "fit_seconds": t_fit,
            "seconds_by_kind": t_kind,
            "exhaustive_tune_seconds_per_job": t_tune / (len(train) + len(held)),
            "selector_choose_seconds": t_choose}


def operand_dependence(E, n=N_C3, reps=8):
    """The headline product on random residues and on all-equal ones (every entry 1: the MFMA inputs barely
    toggle), on the residue route and on limb kernel 4: main-kernel ms each (HIP events, mean of `reps`).
    The difference is what operand data costs the chip's clock."""
    import torch
    from hacs_amd import workloads as W
    p = W.P_C3
    ctx = E.ModCtx(p)
    rnd = (E.random_entries(n, n, p, W.seed_for(n, n, p, 0x01)), E.random_entries(n, n, p, W.seed_for(n, n, p, 0x02)))
    one = (torch.ones((n, n), dtype=torch.int32, device="cuda"), torch.ones((n, n), dtype=torch.int32, device="cuda"))
    ws = torch.empty(max(E.gemm_workspace_bytes(ctx, n, n, n, E.HX_GEMM_LIMB, t) for t in (4, RESIDUE_TILE)),
                     dtype=torch.uint8, device="cuda")
    C_ = torch.empty((n, n), dtype=torch.int32, device="cuda")
    out = {"n": n, "modulus": p}
    for name, (A, B) in (("random", rnd), ("all_ones", one)):
        for tile in (RESIDUE_TILE, 4):
            _, k = gemm_timed(E, ctx, A, B, C_, E.HX_GEMM_LIMB, reps, 2, ws, tile)
            out[f"{name}_tile{tile}_kernel_ms"] = k
    return out


def mfma_calibration(E, blocks=1024, iters=30000):
    """The limb kernel's MFMA instruction from registers only (hx_mfma_rate, about 40 ms per run): the
    sustained rate with uniformly random operands (what the limb planes of random residues present; the
    ceiling of the limb GEMM on this chip) and with near-constant operands (the clock the published peak
    assumes), best of three runs each."""
    rnd = max(E.mfma_rate(blocks, iters, True) for _ in range(3))
    cst = max(E.mfma_rate(blocks, iters, False) for _ in range(3))
    return {"sustained_random_tops": rnd, "sustained_constant_tops": cst,
            "sustained_note": f"{blocks} x 256 threads, {iters} rounds of 16 v_mfma_i32_32x32x32_i8 from registers"}


def level1_kernels(E, n=8192, reps=20):
    """The HBM-bound level-1/2 kernels on n x n residues (256 MiB per matrix at n = 8192): average
    kernel time over `reps` launches (torch events on the launching stream) and the algorithmic bytes
    each moves (every input read once, every output written once), against the 8 TB/s HBM peak;
    torch's own device copy of the same bytes is timed beside them."""
    import torch
    from hacs_amd import workloads as W
    p = W.P_C3
    ctx = E.ModCtx(p)
    X = E.random_entries(n, n, p, 0x51)
    Y = E.random_entries(n, n, p, 0x52)
    Z = torch.empty_like(X)
    cols = torch.from_numpy(np.random.default_rng(1).permutation(n).astype(np.int32)).cuda().reshape(-1, 1)
    xv = X[:, 0:1].contiguous()
    yv = torch.empty((n, 1), dtype=torch.int32, device="cuda")
    word = 4.0 * n * n
    cases = {
        "copy": (lambda: E.copy(X, Z), 2 * word),
        "axpy": (lambda: E.axpy(ctx, 12345, X, Y), 3 * word),
        "scal": (lambda: E.scal(ctx, 3, Y), 2 * word),
        "transpose": (lambda: E.transpose(X, Z), 2 * word),
        "gather": (lambda: E.gather(ctx, 1, 7, X, 0, cols, Z, False), 2 * word),
        "gemv": (lambda: E.gemv(ctx, X, xv, yv), word),
        "torch_copy": (lambda: Z.copy_(X), 2 * word),
    }
    out = {}
    for name, (fn, nbytes) in cases.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "bytes": nbytes, "GB_per_s": gbs, "frac_hbm": gbs / PEAK_HBM_GBS}
    return {"n": n, "kernels": out}


def host_cpu():
    """(threads to use, nproc, CPU model): the oracle runs on at most 16 host threads (the share of one
    GPU's host cores)."""
    nproc = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(16, nproc), nproc, model


def cpu_baseline(p):
    """The oracle's restatement of the reference CPU GEMM (hxo_gemm_fast: Kernels.cpp:680-710's u32
    route, threaded over row tiles as Executor.cpp:69-100) on the host's cores, on a bounded sample of
    the headline product (1024 x 1024 outputs at the full 8192 depth: about 30 thread-seconds of work over
    the three runs on 16 threads), min of 3."""
    from oracle import oracle as O
    from hacs_amd import workloads as W
    threads, nproc, model = host_cpu()
    m, k, n = 1024, 8192, 1024
    A = W.random_entries(m * k, p, W.seed_for(k, k, p, 0x01)).reshape(m, k)              # rows 0..m-1 of A
    B = np.ascontiguousarray(W.random_entries(k * k, p, W.seed_for(k, k, p, 0x02)).reshape(k, k)[:, :n])
    best, total = 1e30, 0.0
    for _ in range(3):
        t0 = time.perf_counter()
        O.gemm_fast(p, A, B, None, 0, threads)
        dt = time.perf_counter() - t0
        best, total = min(best, dt), total + dt
    return {"value": m * k * n / best, "unit": "mod-p mult-adds/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model, "seconds": best, "thread_seconds": total * threads,
            "sample": f"C = A B mod {p}, A {m}x{k}, B {k}x{n} (the 8192-deep product of the headline "
                      f"config on {m}x{n} outputs, {m * k * n:.3g} mult-adds), hxo_gemm_fast on {threads} threads, "
                      f"min of 3"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import hacs_amd
    from hacs_amd import engine as E
    from hacs_amd import workloads as W
    if not hacs_amd.available():
        raise SystemExit("no gfx950 device: " + hacs_amd.lib().hx_unavailable_reason().decode())
    n, p = args.n, W.P_C3
    ctx = E.ModCtx(p)
    if args.pmc_mode:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)                  # 1 GiB read + 1 GiB written
        torch.cuda.synchronize()
        A = E.random_entries(n, n, p, W.seed_for(n, n, p, 0x01))
        B = E.random_entries(n, n, p, W.seed_for(n, n, p, 0x02))
        Cm = torch.empty((n, n), dtype=torch.int32, device="cuda")
        for _ in range(args.pmc_reps):
            E.gemm(ctx, A, B, Cm, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=args.tile)
        torch.cuda.synchronize()
        log("pmc mode done")
        return
    if args.only:
        leg = {"batch": lambda: batched_echelon(E, args.batch_jobs), "lu": lambda: lu_c2(E, cpu=not args.no_lu_cpu),
               "mixed": lambda: mixed_stream(args.mixed_jobs),
               "howell": lambda: howell_batch(E, args.howell_jobs),
               "ring2k": lambda: ring2k_lu(E),
               "u64": lambda: gemm_u64_leg(E),
               "f2": lambda: f2_leg(E),
               "splitk": lambda: splitk_leg(E),
               "smith": lambda: smith_leg(E, cpu=not args.no_smith_cpu),
               "howell_ell": lambda: howell_ell_leg(E),
               "f3": lambda: f3_leg(E),
               "mfma": lambda: mfma_calibration(E),
               "operands": lambda: operand_dependence(E),
               "level1": lambda: level1_kernels(E),
               "cpu": lambda: cpu_baseline(p)}[args.only]()
        print(json.dumps({"only": args.only, **leg}), flush=True)
        return
    log(f"device {E.device_info()}; C3 n={n} p={p}")
    A = E.random_entries(n, n, p, W.seed_for(n, n, p, 0x01))
    B = E.random_entries(n, n, p, W.seed_for(n, n, p, 0x02))
    C_limb = torch.empty((n, n), dtype=torch.int32, device="cuda")
    C_plain = torch.empty((n, n), dtype=torch.int32, device="cuda")
    ws = torch.empty(max(E.gemm_workspace_bytes(ctx, n, n, n, E.HX_GEMM_LIMB, t) for t in LIMB_KERNELS),
                     dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    wall_l, main_l = gemm_timed(E, ctx, A, B, C_limb, E.HX_GEMM_LIMB, args.steps, args.warmup, ws, args.tile)
    log(f"limb: {wall_l:.3f} ms/step (main kernel {main_l:.3f} ms)")
    kver = E.gemm_route(ctx, n, n, n, E.HX_GEMM_LIMB, args.tile)      # the kernel the headline ran
    others, same_alt = [], True
    C_alt = torch.empty((n, n), dtype=torch.int32, device="cuda")
    for alt in LIMB_KERNELS:
        if alt == kver:
            continue
        _, main_alt = gemm_timed(E, ctx, A, B, C_alt, E.HX_GEMM_LIMB, 5, 1, ws, alt)
        same = bool(torch.equal(C_limb, C_alt))
        same_alt = same_alt and same
        others.append({"tile": alt, "kernel": kernel_name(alt, ctx.limbs), "kernel_ms": main_alt,
                       "int8_products_per_mac": int8_products(E, alt, p, n, ctx.limbs)})
        log(f"other limb kernel (tile {alt}): main kernel {main_alt:.3f} ms; same bytes: {same}")
    del C_alt
    wall_p, main_p = gemm_timed(E, ctx, A, B, C_plain, E.HX_GEMM_PLAIN, args.plain_steps, 1, None)
    log(f"plain: {wall_p:.3f} ms/step")
    same = bool(torch.equal(C_limb, C_plain))
    spot = spot_check(E.from_device(C_limb), A, B, p)
    log(f"limb == plain: {same}; spot entries exact: {spot}")
    macs = float(n) ** 3
    value = macs / (wall_l * 1e-3)
    kname = kernel_name(kver, ctx.limbs)
    prods = int8_products(E, kver, p, n, ctx.limbs)
    achieved = 2.0 * prods * macs / (main_l * 1e-3) / 1e12
    res = {
        "metric": "exact mod-p GEMM mult-adds/s at n=8192 (limb-MFMA vs plain); batched LU jobs/s",
        "value": value,
        "unit": "mod-p mult-adds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_l,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"GEMM mod {p}, m=n=k={n}, uint32 residues (SplitMix64 salts 0x01/0x02), "
                               + ("residue route (%d int8 MFMA products per multiply-add modulo small coprime "
                                  "moduli, CRT into C)" % prods if kver == RESIDUE_TILE else
                                  "limb-MFMA kernel (%d int8 products per multiply-add)" % prods)
                               + ", inputs resident in HBM",
                   "modulus": p, "n": n, "limbs": ctx.limbs, "parallelism": "single GPU"},
        "plain_value": macs / (wall_p * 1e-3),
        "plain_ms_per_step": wall_p,
        "limb_vs_plain": wall_p / wall_l,
        "parity": {"limb_equals_plain": same, "spot_entries_exact": spot, "limb_kernels_agree": same_alt},
        "limb_other_kernels": others,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT8_TOPS, "traffic": profiled_traffic(n, ctx.limbs, kname),
                     "kernel": kname,
                     "kernel_ms": main_l,
                     "algorithmic_ops_per_launch": 2.0 * prods * macs,
                     "int8_products_per_mac": prods},
    }
    cal = mfma_calibration(E)
    log(f"MFMA calibration: {cal}")
    res["roofline"].update(cal)
    res["roofline"]["frac_of_sustained_random"] = achieved / cal["sustained_random_tops"]
    if not args.no_batch:
        b = batched_echelon(E, args.batch_jobs)
        log(f"batched echelon: {b}")
        res["batched_lu"] = b
    if not args.no_lu:
        lu = lu_c2(E, cpu=not args.no_lu_cpu)
        log(f"C2 rref: {lu}")
        res["lu_c2"] = lu
    if not args.no_howell:
        hw = howell_batch(E, args.howell_jobs)
        log(f"howell batch: {hw}")
        res["howell_batch"] = hw
    if not args.no_ring2k:
        r2 = ring2k_lu(E)
        log(f"ring2k LU: {r2}")
        res["ring2k_lu"] = r2
    if not args.no_smith:
        sm = smith_leg(E, cpu=not args.no_smith_cpu)
        log(f"ring2k smith: {sm}")
        res["ring2k_smith"] = sm
    if not args.no_howell_ell:
        he = howell_ell_leg(E)
        log(f"howell over Z/3^10: {he}")
        res["howell_ell"] = he
    if not args.no_f3:
        f3 = f3_leg(E)
        log(f"F_3: {f3}")
        res["f3"] = f3
    if not args.no_u64:
        g64 = gemm_u64_leg(E)
        log(f"u64 GEMM: {g64}")
        res["gemm_u64"] = g64
    if not args.no_f2:
        f2 = f2_leg(E)
        log(f"F_2: {f2}")
        res["f2"] = f2
    if not args.no_splitk:
        sk = splitk_leg(E)
        log(f"split-K: {sk}")
        res["splitk"] = sk
    if not args.no_level1:
        l1 = level1_kernels(E)
        log(f"level-1/2 kernels: {l1}")
        res["level1"] = l1
    if not args.no_mixed:
        mx = mixed_stream(args.mixed_jobs)
        log(f"mixed stream: {mx}")
        res["mixed_stream"] = mx
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(p)
        log(f"cpu baseline: {res['cpu_baseline']}")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
