"""Exhaustive autotuner, the learned selector, and the mixed job stream.

Kinds and their choices (every choice computes the same bytes; only time differs):
  gemm     C = A B mod 2147483629 (square n)      : kernel variant  plain (VALU) | limb (INT8 MFMA)
  echelon  RREF mod 1000003 of a sparse n x n job : panel width     8 | 16 | 32
  howell   Howell form over Z/2^16 (dense n x n)  : one kernel (a single choice)

The autotuner times every choice of every job it is given (one job per launch, HIP events on
the job's stream, median of repeats) and sets the labels: its fastest choice is the label.  The
selector is a 2-layer MLP (hx_mlp_* kernels; 64 hidden units, FP32) mapping job features
(log sizes, modulus bits, density, kind) to predicted log2-times of the choices; it is trained
on the autotuner's timings with the library's own backward and Adam kernels, and picks the
cheapest predicted valid choice.  A job the selector has not been trained for (unknown kind)
falls back to the library's static policy.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import engine as E
from . import workloads as W
from .selector import MLP

KINDS = ("gemm", "echelon", "howell")
CHOICES = {"gemm": [E.HX_GEMM_PLAIN, E.HX_GEMM_LIMB], "echelon": [8, 16, 32], "howell": [0]}
NOUT = 3
P_GEMM, P_ECH, K_HOWELL = W.P_C3, W.P_C4, 16
DENSITY = {"gemm": 1.0, "echelon": 0.05, "howell": 1.0}
BITS = {"gemm": 31, "echelon": 20, "howell": 16}


def features(kind: str, n: int) -> list[float]:
    ln = math.log2(max(n, 1)) / 12.0
    oh = [1.0 if kind == k else 0.0 for k in KINDS]
    return [ln, ln, ln * ln, BITS[kind] / 32.0, DENSITY[kind]] + oh


def static_choice(kind: str, n: int) -> int:
    """The choice without the network: the library's automatic GEMM policy, the widest panel."""
    if kind == "gemm":
        return 1 if n >= 64 and float(n) ** 3 >= 4.0e6 else 0
    return len(CHOICES[kind]) - 1 if kind == "echelon" else 0


class Job:
    """One job of the stream with its inputs in HBM (generated from its seed)."""

    def __init__(self, kind: str, n: int, j: int, salt: int = 0x60):
        self.kind, self.n, self.j = kind, n, j
        if kind == "gemm":
            self.A = E.random_entries(n, n, P_GEMM, W.seed_for(n, n, P_GEMM, salt + 1) ^ j)
            self.B = E.random_entries(n, n, P_GEMM, W.seed_for(n, n, P_GEMM, salt + 2) ^ j)
            self.C = torch.empty((n, n), dtype=torch.int32, device="cuda")
        elif kind == "echelon":     # == workloads.sparse_job(j, n, n, P_ECH, salt + 3), generated in HBM
            self.src = E.random_sparse(n, n, P_ECH, W.seed_for(n, n, P_ECH, salt + 3) ^ j)
            self.M = self.src.clone()
        else:                        # == workloads.dense_job(j, n, n, 2^16, salt + 4)
            self.src = E.random_entries(n, n, 1 << K_HOWELL, W.seed_for(n, n, 1 << K_HOWELL, salt + 4) ^ j)

    def free(self):
        for a in ("A", "B", "C", "src", "M"):
            if hasattr(self, a):
                delattr(self, a)


def _time_call(fn, reps: int = 3) -> float:
    s = torch.cuda.current_stream()
    ts = []
    fn()                                    # warm-up (and a check that it runs)
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(ts))


def time_choice(job: Job, slot: int, reps: int = 3) -> float:
    if job.kind == "gemm":
        ctx = E.ModCtx(P_GEMM)
        v = CHOICES["gemm"][slot]
        return _time_call(lambda: E.gemm(ctx, job.A, job.B, job.C, E.GemmOp.Set, variant=v), reps)
    if job.kind == "echelon":
        b = E.RrefBatch(E.ModCtx(P_ECH), [job.M], panels=[CHOICES["echelon"][slot]])

        def run():
            job.M.copy_(job.src)
            b.run()
        t_copy = _time_call(lambda: job.M.copy_(job.src), reps)
        return max(_time_call(run, reps) - t_copy, 1e-7)
    b = E.HowellBatch(K_HOWELL, [job.src])
    return _time_call(b.run, reps)


class Autotuner:
    """Exhaustive: every valid choice of every job, timed."""

    def __init__(self, reps: int = 3):
        self.reps = reps

    def tune(self, jobs: list[Job]) -> np.ndarray:
        T = np.full((len(jobs), NOUT), np.inf)
        for i, jb in enumerate(jobs):
            for s in range(len(CHOICES[jb.kind])):
                T[i, s] = time_choice(jb, s, self.reps)
        return T


class Selector:
    """The learned selector (MLP on the GPU); predictions are log2-seconds per choice slot."""

    def __init__(self, seed: int = 0):
        self.net = MLP(len(features("gemm", 64)), 64, NOUT, seed=seed)
        self.mu = 0.0
        self.trained = False
        self.final_loss = None

    def fit(self, kinds: list[str], ns: list[int], T: np.ndarray, steps: int = 3000, lr: float = 3e-3) -> float:
        X = torch.as_tensor(np.array([features(k, n) for k, n in zip(kinds, ns)], np.float32), device="cuda")
        mask = np.isfinite(T).astype(np.float32)
        Y = np.where(np.isfinite(T), np.log2(np.where(np.isfinite(T), T, 1.0)), 0.0)
        self.mu = float((Y * mask).sum() / mask.sum())
        Yd = torch.as_tensor((Y - self.mu).astype(np.float32), device="cuda")
        Md = torch.as_tensor(mask, device="cuda")
        loss = None
        for t in range(steps):
            loss = self.net.train_step(X, Yd, Md, lr=lr * (0.3 if t > steps * 2 // 3 else 1.0))
        self.trained = True
        self.final_loss = loss
        return loss

    def choose(self, kinds: list[str], ns: list[int]) -> list[int]:
        if not self.trained:
            return [static_choice(k, n) for k, n in zip(kinds, ns)]
        X = torch.as_tensor(np.array([features(k, n) for k, n in zip(kinds, ns)], np.float32), device="cuda")
        pred = self.net.predict(X).cpu().numpy()
        out = []
        for k, row in zip(kinds, pred):
            nv = len(CHOICES[k])
            out.append(int(np.argmin(row[:nv])))
        return out


def run_stream(jobs: list[Job], slots: list[int], per_kind: dict | None = None) -> float:
    """Executes every job with its chosen slot, one batch per kind carrying the per-job choices
    (GEMM: per-job kernel variant in one batch; echelon: per-job panel width in one persistent
    launch; Howell: one persistent launch), so a choice never splits a launch.  Returns the wall
    seconds of the launches (inputs already in HBM; batch tables built beforehand); with
    `per_kind`, also adds each kind's own seconds to it (each kind then timed on its own)."""
    ctx_g, ctx_e = E.ModCtx(P_GEMM), E.ModCtx(P_ECH)
    by = {k: [(jb, s) for jb, s in zip(jobs, slots) if jb.kind == k] for k in KINDS}
    runners = []
    if by["gemm"]:
        runners.append(("gemm", E.GemmBatch(ctx_g, [(jb.A, jb.B, jb.C, E.GemmOp.Set, CHOICES["gemm"][s])
                                                    for jb, s in by["gemm"]])))
    if by["echelon"]:
        for jb, _ in by["echelon"]:
            jb.M.copy_(jb.src)
        runners.append(("echelon", E.RrefBatch(ctx_e, [jb.M for jb, _ in by["echelon"]],
                                               panels=[CHOICES["echelon"][s] for _, s in by["echelon"]])))
    if by["howell"]:
        runners.append(("howell", E.HowellBatch(K_HOWELL, [jb.src for jb, _ in by["howell"]])))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for kind, r in runners:
        if per_kind is not None:
            t1 = time.perf_counter()
            r.run()
            torch.cuda.synchronize()
            per_kind[kind] = per_kind.get(kind, 0.0) + time.perf_counter() - t1
        else:
            r.run()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def sample_jobs(count: int, salt: int, n_caps=(4096, 1024, 512)) -> list[tuple[str, int]]:
    """(kind, n) pairs from the stream's distribution with its own salt (training / held-out sets)."""
    return [(k, min(n, n_caps[KINDS.index(k)])) for k, n in W.c5_stream(count, salt=salt)]
