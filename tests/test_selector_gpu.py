"""The kernel choices of the mixed job stream, the exhaustive autotuner and the learned selector.

Every choice of every kind must give the oracle's bytes (a misprediction can only cost time);
the selector, trained on the autotuner's timings by the library's own MLP kernels, must pick
choices whose measured total time stays close to the autotuner's best on held-out jobs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_every_choice_is_bit_identical(hx_gpu, O):
    from hacs_amd import autotune as AT
    from hacs_amd import engine as E
    from hacs_amd import workloads as W
    for kind, n, j in [("gemm", 96, 1), ("gemm", 300, 2), ("echelon", 150, 3), ("echelon", 37, 4), ("howell", 70, 5)]:
        jb = AT.Job(kind, n, j)
        outs = []
        for slot in range(len(AT.CHOICES[kind])):
            AT.time_choice(jb, slot, reps=1)
            torch.cuda.synchronize()
            if kind == "gemm":
                outs.append(E.from_device(jb.C).copy())
            elif kind == "echelon":
                outs.append(E.from_device(jb.M).copy())
            else:
                b = E.HowellBatch(AT.K_HOWELL, [jb.src])
                b.run()
                torch.cuda.synchronize()
                outs.append(b.results()[0].H)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), f"{kind} n={n}: choices differ"
        if kind == "gemm":
            want = O.gemm_fast(AT.P_GEMM, E.from_device(jb.A), E.from_device(jb.B))
            assert np.array_equal(outs[0], want)
        elif kind == "echelon":
            A = W.sparse_job(j, n, n, AT.P_ECH, 0x60 + 3)
            r, piv, prow, R = O.rref_rule(AT.P_ECH, A)
            assert np.array_equal(outs[0][:r], R) and not outs[0][r:].any()
        else:
            A = W.dense_job(j, n, n, 1 << 16, 0x60 + 4)
            assert np.array_equal(outs[0], O.howell_fast(16, A)[0])


def test_selector_trains_and_selects(hx_gpu):
    from hacs_amd import autotune as AT
    caps = (1024, 256, 128)
    train = [AT.Job(k, n, i) for i, (k, n) in enumerate(AT.sample_jobs(45, 0x71, caps))]
    held = [AT.Job(k, n, i) for i, (k, n) in enumerate(AT.sample_jobs(24, 0x72, caps))]
    tuner = AT.Autotuner(reps=3)
    Tt = tuner.tune(train)
    Th = tuner.tune(held)
    sel = AT.Selector(seed=1)
    loss = sel.fit([j.kind for j in train], [j.n for j in train], Tt, steps=1500)
    assert np.isfinite(loss)
    slots = sel.choose([j.kind for j in held], [j.n for j in held])
    best = float(np.nanmin(np.where(np.isfinite(Th), Th, np.nan), axis=1).sum())
    t_sel = float(sum(Th[i, s] for i, s in enumerate(slots)))
    print(f"held-out: best {best * 1e3:.3f} ms, selector {t_sel * 1e3:.3f} ms ({t_sel / best:.3f}x)")
    assert t_sel <= 1.25 * best


def test_selector_gains_over_static_rule(hx_gpu):
    """The study the selector is for: GEMM and echelon jobs with n in [48, 640] and moduli of 2, 3 and 4
    limbs, where the library's static rule (a size threshold that ignores the limb count, the default
    limb kernel, the widest panel) is wrong on part of the jobs.  Held out, the trained network stays
    within 5 % of the exhaustive autotuner's best and costs less than the static rule."""
    from hacs_amd import autotune as AT
    r = AT.selector_study(n_train=120, n_held=60, reps=3, steps=2500)
    print("selector study:", r)
    assert r["all"]["selector_vs_best"] <= 1.05, r
    assert r["all"]["static_vs_selector"] > 1.0, r


def test_device_generators_match_workloads(hx_gpu):
    from hacs_amd import engine as E
    from hacs_amd import workloads as W
    for (m, n, j) in [(37, 51, 3), (200, 130, 11)]:
        seed = W.seed_for(m, n, W.P_C4, 0x63) ^ j
        assert np.array_equal(E.from_device(E.random_sparse(m, n, W.P_C4, seed)), W.sparse_job(j, m, n, W.P_C4, 0x63))
        seed2 = W.seed_for(m, n, 1 << 16, 0x64) ^ j
        assert np.array_equal(E.from_device(E.random_entries(m, n, 1 << 16, seed2)), W.dense_job(j, m, n, 1 << 16, 0x64))
