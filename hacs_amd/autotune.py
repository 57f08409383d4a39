"""Exhaustive autotuner, the learned selector, and the mixed job stream.

Kinds and their choices (every choice computes the same bytes; only time differs):
  gemm     C = A B mod p (square n)               : plain (VALU) | limb (INT8 MFMA) kernel 1 | 2 | 3 | 4
  echelon  RREF mod p of a sparse n x n job       : panel width 8 | 16 | 32
  howell   Howell form over Z/2^16 (dense n x n)  : one kernel (a single choice)
The stream's moduli are the configuration's (2147483629 for GEMM, 1000003 for echelon); the selector
study draws them from STUDY_P, so the limb count (a feature) changes which choice is fastest.

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
# the choices per kind (slots): GEMM (variant, limb kernel); echelon panel width; Howell one kernel
CHOICES = {"gemm": [(E.HX_GEMM_PLAIN, 0), (E.HX_GEMM_LIMB, 1), (E.HX_GEMM_LIMB, 2), (E.HX_GEMM_LIMB, 3),
                    (E.HX_GEMM_LIMB, 4)],
           "echelon": [8, 16, 32], "howell": [0]}
NOUT = 5
P_GEMM, P_ECH, K_HOWELL = W.P_C3, W.P_C4, 16
DENSITY = {"gemm": 1.0, "echelon": 0.05, "howell": 1.0}
# moduli of the selector study: the limb count of the GEMM route (L = 2, 3, 4) changes where the matrix
# cores pay off, which the library's static rule (a size threshold) does not see
STUDY_P = {"gemm": (65521, 16777213, 2147483629), "echelon": (1000003, 65521, 2147483647), "howell": (1 << 16,)}


def limbs_of(p: int) -> int:
    half = (p - 1) // 2
    return next((l for l in (1, 2, 3, 4) if half <= 127 * (256 ** l - 1) // 255), 0)


def default_p(kind: str) -> int:
    return {"gemm": P_GEMM, "echelon": P_ECH, "howell": 1 << K_HOWELL}[kind]


def features(kind: str, n: int, p: int | None = None) -> list[float]:
    """log size, its square, the limb count and modulus bits (GEMM / echelon), density, kind one-hot."""
    p = default_p(kind) if p is None else p
    ln = math.log2(max(n, 1)) / 12.0
    oh = [1.0 if kind == k else 0.0 for k in KINDS]
    bits = math.log2(p) / 32.0
    return [ln, ln * ln, limbs_of(p) / 4.0 if kind != "howell" else 0.0, bits, DENSITY[kind]] + oh


def static_choice(kind: str, n: int, p: int | None = None) -> int:
    """The choice without the network: the library's automatic GEMM rule (limb route with its default
    kernel once n^3 >= 4e6 and n >= 64, else plain, whatever the modulus), the widest panel."""
    if kind == "gemm":
        return 4 if n >= 64 and float(n) ** 3 >= 4.0e6 else 0
    return len(CHOICES[kind]) - 1 if kind == "echelon" else 0


class Job:
    """One job of the stream with its inputs in HBM (generated from its seed); p = its modulus."""

    def __init__(self, kind: str, n: int, j: int, salt: int = 0x60, p: int | None = None):
        self.kind, self.n, self.j = kind, n, j
        self.p = default_p(kind) if p is None else p
        if kind == "gemm":
            self.A = E.random_entries(n, n, self.p, W.seed_for(n, n, self.p, salt + 1) ^ j)
            self.B = E.random_entries(n, n, self.p, W.seed_for(n, n, self.p, salt + 2) ^ j)
            self.C = torch.empty((n, n), dtype=torch.int32, device="cuda")
        elif kind == "echelon":     # == workloads.sparse_job(j, n, n, p, salt + 3), generated in HBM
            self.src = E.random_sparse(n, n, self.p, W.seed_for(n, n, self.p, salt + 3) ^ j)
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
        ctx = E.ModCtx(job.p)
        v, tile = CHOICES["gemm"][slot]
        return _time_call(lambda: E.gemm(ctx, job.A, job.B, job.C, E.GemmOp.Set, variant=v, tile=tile), reps)
    if job.kind == "echelon":
        b = E.RrefBatch(E.ModCtx(job.p), [job.M], panels=[CHOICES["echelon"][slot]])

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

    def fit(self, kinds: list[str], ns: list[int], T: np.ndarray, steps: int = 3000, lr: float = 3e-3,
            ps: list[int] | None = None) -> float:
        ps = [None] * len(kinds) if ps is None else ps
        X = torch.as_tensor(np.array([features(k, n, p) for k, n, p in zip(kinds, ns, ps)], np.float32), device="cuda")
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

    def choose(self, kinds: list[str], ns: list[int], ps: list[int] | None = None) -> list[int]:
        ps = [None] * len(kinds) if ps is None else ps
        if not self.trained:
            return [static_choice(k, n, p) for k, n, p in zip(kinds, ns, ps)]
        X = torch.as_tensor(np.array([features(k, n, p) for k, n, p in zip(kinds, ns, ps)], np.float32), device="cuda")
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
    by = {k: [(jb, s) for jb, s in zip(jobs, slots) if jb.kind == k] for k in KINDS}
    runners = []
    for p in sorted({jb.p for jb, _ in by["gemm"]}):          # one batch per modulus
        group = [(jb, s) for jb, s in by["gemm"] if jb.p == p]
        runners.append(("gemm", E.GemmBatch(E.ModCtx(p), [(jb.A, jb.B, jb.C, E.GemmOp.Set) + tuple(CHOICES["gemm"][s])
                                                          for jb, s in group])))
    for p in sorted({jb.p for jb, _ in by["echelon"]}):
        group = [(jb, s) for jb, s in by["echelon"] if jb.p == p]
        for jb, _ in group:
            jb.M.copy_(jb.src)
        runners.append(("echelon", E.RrefBatch(E.ModCtx(p), [jb.M for jb, _ in group],
                                               panels=[CHOICES["echelon"][s] for _, s in group])))
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


def study_jobs(count: int, salt: int, lo: int = 48, hi: int = 640) -> list[tuple[str, int, int]]:
    """(kind, n, p) for the selector study: GEMM and echelon jobs with n log-uniform in [lo, hi] (where
    the choices trade places) and moduli drawn from STUDY_P (the limb count varies), seeded."""
    d = W.splitmix_draws(W.seed_for(count, 7, 0, salt), 0, 3 * count)
    out = []
    for j in range(count):
        kind = ("gemm", "echelon")[int(d[3 * j] % np.uint64(2))]
        n = int(W._log_uniform(d[3 * j + 1:3 * j + 2], lo, hi)[0])
        ps = STUDY_P[kind]
        out.append((kind, n, int(ps[int(d[3 * j + 2] % np.uint64(len(ps)))])))
    return out


def selector_study(n_train: int = 160, n_held: int = 80, reps: int = 3, steps: int = 3000, seed: int = 0) -> dict:
    """Train the selector on autotuned study jobs (study_jobs: GEMM and echelon, n in [48, 640], moduli
    with 2, 3 and 4 limbs) and score it on held-out ones against the autotuner's best and against the
    static rule.  Returns the totals, the per-kind split and the agreement with the autotuner."""
    train_spec = study_jobs(n_train, 0x73)
    held_spec = study_jobs(n_held, 0x74)
    tuner = Autotuner(reps=reps)
    tr = [Job(k, n, i, p=p) for i, (k, n, p) in enumerate(train_spec)]
    Tt = tuner.tune(tr)
    for jb in tr:
        jb.free()
    hd = [Job(k, n, i, salt=0x70, p=p) for i, (k, n, p) in enumerate(held_spec)]
    Th = tuner.tune(hd)
    for jb in hd:
        jb.free()
    sel = Selector(seed=seed)
    loss = sel.fit([k for k, _, _ in train_spec], [n for _, n, _ in train_spec], Tt, steps=steps,
                   ps=[p for _, _, p in train_spec])
    slots = sel.choose([k for k, _, _ in held_spec], [n for _, n, _ in held_spec], [p for _, _, p in held_spec])
    stat = [static_choice(k, n, p) for k, n, p in held_spec]
    best_slot = [int(np.argmin(np.where(np.isfinite(Th[i]), Th[i], np.inf))) for i in range(len(held_spec))]
    out = {"train_jobs": n_train, "heldout_jobs": n_held, "train_loss": loss}
    for name, rows in (("all", range(len(held_spec))),
                       ("gemm", [i for i, s in enumerate(held_spec) if s[0] == "gemm"]),
                       ("echelon", [i for i, s in enumerate(held_spec) if s[0] == "echelon"])):
        rows = list(rows)
        best = float(sum(Th[i, best_slot[i]] for i in rows))
        t_sel = float(sum(Th[i, slots[i]] for i in rows))
        t_stat = float(sum(Th[i, stat[i]] for i in rows))
        out[name] = {"best_s": best, "selector_s": t_sel, "static_s": t_stat, "selector_vs_best": t_sel / best,
                     "static_vs_best": t_stat / best, "static_vs_selector": t_stat / t_sel,
                     "selector_agrees": float(np.mean([slots[i] == best_slot[i] for i in rows])),
                     "static_agrees": float(np.mean([stat[i] == best_slot[i] for i in rows]))}
    return out
