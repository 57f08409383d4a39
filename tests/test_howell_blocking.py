"""The panel-blocked order of the Howell pass (the algorithm behind the batched Howell kernel,
hacs_amd/csrc/hx_howell.hip) restated in Python on small cases and compared with the oracle.

It pins the reformulation the kernel relies on: per panel of w columns the pass runs on the panel
columns only, recording every active row's multipliers; the trailing columns are then brought up to
date in one sweep (pivot rows in step order, T_t = u_t^-1 (row_t - sum_{s<t} q[row_t][s] T_s);
every other active row row -= sum_t q[row][t] T_t), with a closure row made at step s treated as a
row that starts from zero with multiplier -2^(k-v_s) at step s.  Reduce-above as in the pass.
"""
import numpy as np
import pytest

M32 = 0xffffffff


def _ctz(x):
    return (x & -x).bit_length() - 1


def _inv_odd(u):
    x = u
    for _ in range(5):
        x = (x * (2 - u * x)) & M32
    return x


def howell_panel_blocked(k, A, w):
    m, n = A.shape
    mask = M32 if k >= 32 else (1 << k) - 1
    W = np.zeros((m + n, n), np.uint64)
    W[:m] = A.astype(np.uint64) & mask
    act = list(range(m))
    piv_ids, pivots, pvs = [], [], []
    ncl = 0
    for c0 in range(0, n, w):
        c1 = min(n, c0 + w)
        wc = c1 - c0
        pan = {i: [int(W[i, c]) for c in range(c0, c1)] for i in act if i is not None}
        Q, steps = {}, []
        for t in range(wc):
            best = None
            for pos, i in enumerate(act):
                if i is not None and pan[i][t] & mask:
                    key = (_ctz(pan[i][t] & mask), pos)
                    best = key if best is None or key < best else best
            if best is None:
                steps.append(None)
                continue
            v, bpos = best
            pid = act[bpos]
            P = pan[pid]
            ui = _inv_odd((P[t] & mask) >> v)
            for tt in range(t, wc):
                P[tt] = (P[tt] * ui) & M32
            act[bpos] = None
            for i in act:
                if i is None:
                    continue
                q = (pan[i][t] & mask) >> v
                Q.setdefault(i, [0] * wc)[t] = q
                for tt in range(t, wc):
                    pan[i][tt] = (pan[i][tt] - q * P[tt]) & M32
            if v > 0:
                cid, sh = m + ncl, k - v
                ncl += 1
                pan[cid] = [0 if tt <= t else (P[tt] << sh) & M32 for tt in range(wc)]
                Q[cid] = [((-(1 << sh)) & M32) if tt == t else 0 for tt in range(wc)]
                W[cid, :] = 0
                act.append(cid)
            steps.append((pid, ui))
            piv_ids.append(pid)
            pivots.append(c0 + t)
            pvs.append(v)
        for i, vals in pan.items():
            W[i, c0:c1] = vals
        if c1 == n:
            continue
        T = {}
        for t, st in enumerate(steps):
            if st is None:
                continue
            pid, ui = st
            row = W[pid, c1:].copy()
            for s in range(t):
                if s in T and Q.get(pid, [0] * wc)[s]:
                    row = (row - np.uint64(Q[pid][s]) * T[s]) & M32
            T[t] = (row * np.uint64(ui)) & M32
            W[pid, c1:] = T[t]
        for i in act:
            if i is not None and i in Q:
                row = W[i, c1:].copy()
                for t in T:
                    if Q[i][t]:
                        row = (row - np.uint64(Q[i][t]) * T[t]) & M32
                W[i, c1:] = row
    h = len(piv_ids)
    H = np.array([[0 if j < pivots[a] else int(W[piv_ids[a], j]) & mask for j in range(n)] for a in range(h)],
                 np.uint64).reshape(h, n)
    for i in range(1, h):
        for ip in range(i):
            q = (int(H[ip, pivots[i]]) & mask) >> pvs[i]
            if q:
                H[ip] = (H[ip] - np.uint64(q) * H[i]) & mask
    return H.astype(np.uint32), np.array(pivots, np.uint32), np.array(pvs, np.uint32)


@pytest.mark.parametrize("k", [1, 2, 3, 8, 16, 31, 32])
def test_panel_blocked_order_equals_oracle(O, k):
    rng = np.random.default_rng(900 + k)
    top = 1 << min(k, 32)
    for trial in range(24):
        m, n = (int(x) for x in rng.integers(1, 28, 2))
        if trial % 3 == 0:
            A = rng.integers(0, top, (m, n), dtype=np.uint64)
        elif trial % 3 == 1:       # planted valuations: non-unit pivots and closure rows
            A = rng.integers(0, 4, (m, n), dtype=np.uint64) << rng.integers(0, k, (m, n)).astype(np.uint64)
        else:                      # one valuation per row
            A = rng.integers(0, 3, (m, n), dtype=np.uint64) * np.uint64(1 << int(rng.integers(0, k)))
        A = (A & np.uint64(top - 1)).astype(np.uint32)
        w = int(rng.choice([1, 2, 3, 8, 16]))
        H, piv, pv = howell_panel_blocked(k, A, w)
        Hr, pr, vr = O.howell_fast(k, A)
        assert np.array_equal(piv, pr) and np.array_equal(pv, vr) and np.array_equal(H, Hr), (m, n, k, w)
