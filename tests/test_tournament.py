"""CPU check of the pivot tournament used by the large RREF (hx_rref_large.hip): the row rank profile
of a panel found on 256-row chunks (level 0), then on groups of up to 16 candidate lists (merge
levels, repeated until one list is left), re-running the smallest-row rule on the raw candidate rows
at every merge, must give the same pivot rows, in the same pivot order, as the rule on the whole
panel (the oracle's hxo_rref_rule).  The claim holds for any merge tree; a pairwise tree over 64-row
leaves is checked as well.  Planted ranks, zero rows and a negative control where a chunk drops its
candidates.  The speculative first step (the rule on the first 64 rows alone, accepted only when it
finds a pivot in every panel column) must give the rule's pivots too, with the tournament deciding
otherwise."""
import numpy as np
import pytest

P = 101


def _rule(rows, A):
    """The smallest-row rule on the listed rows (ascending) of panel A, division-free: the chosen rows
    (ascending) and the pivot order [(row, column)]."""
    M = A[rows].astype(np.int64) % P
    done = np.zeros(len(rows), bool)
    order = []
    for q in range(A.shape[1]):
        cand = [i for i in range(len(rows)) if not done[i] and M[i, q] != 0]
        if not cand:
            continue
        b = cand[0]
        done[b] = True
        for i in range(len(rows)):
            if i != b and not done[i] and M[i, q] != 0:
                M[i] = (M[b, q] * M[i] - M[i, q] * M[b]) % P
        order.append((rows[b], q))
    return [r for r, d in zip(rows, done) if d], order


def _pairwise(lists, A):
    while len(lists) > 1:
        lists = [_rule(sorted(lists[i] + (lists[i + 1] if i + 1 < len(lists) else [])), A)[0]
                 for i in range(0, len(lists), 2)]
    return lists[0] if lists else []


def _tournament(A, drop_leaf=None):
    """The kernels' tree: 256-row chunks, then merges of up to 16 lists until one is left."""
    m = A.shape[0]
    lists = [_rule(list(range(r, min(r + 256, m))), A)[0] for r in range(0, m, 256)]
    if drop_leaf is not None and drop_leaf < len(lists):
        lists[drop_leaf] = []
    while len(lists) > 1:
        lists = [_rule(sorted(sum(lists[i:i + 16], [])), A)[0] for i in range(0, len(lists), 16)]
    final = lists[0] if lists else []
    return _rule(final, A)[1]


def _pairwise_tree(A):
    m = A.shape[0]
    leaves = [_rule(list(range(r, min(r + 64, m))), A)[0] for r in range(0, m, 64)]
    return _rule(_pairwise(leaves, A), A)[1]


@pytest.mark.parametrize("seed", range(12))
def test_tree_tournament_equals_rule(O, seed):
    rng = np.random.default_rng(seed)
    m, w = int(rng.integers(1, 1400)), 32
    rank = int(rng.integers(0, 33))
    A = ((rng.integers(0, P, (m, rank)) @ rng.integers(0, P, (rank, w))) % P).astype(np.uint32)
    A[rng.random(m) < 0.25] = 0
    r, piv, prow, _ = O.rref_rule(P, A, want_rref=False)
    want = list(zip(prow.tolist(), piv.tolist()))
    assert _tournament(A) == want
    assert _pairwise_tree(A) == want


def test_tree_tournament_negative_control(O):
    """Dropping one leaf's candidates (a plausible indexing mistake) changes the pivots."""
    rng = np.random.default_rng(99)
    A = ((rng.integers(0, P, (300, 6)) @ rng.integers(0, P, (6, 32))) % P).astype(np.uint32)
    A = np.concatenate([A, A[:40]])
    A[:256] = 0
    A[256:262] = ((rng.integers(0, P, (6, 6)) @ rng.integers(0, P, (6, 32))) % P).astype(np.uint32)
    r, piv, prow, _ = O.rref_rule(P, A, want_rref=False)
    want = list(zip(prow.tolist(), piv.tolist()))
    assert _tournament(A) == want
    assert _tournament(A, drop_leaf=1) != want


def _speculative(A, spr=64, accept_partial=False):
    """rrp_spec_kernel, then the tournament when it does not settle the panel: (pivot order, settled)."""
    rows = list(range(min(spr, A.shape[0])))
    order = _rule(rows, A)[1]
    if len(order) == A.shape[1] or (accept_partial and order):
        return order, True
    return _tournament(A), False


@pytest.mark.parametrize("seed", range(16))
def test_speculative_first_rows_equal_rule(O, seed):
    rng = np.random.default_rng(1000 + seed)
    m, w = int(rng.integers(1, 900)), 32
    full = seed % 2 == 0
    rank = 32 if full else int(rng.integers(0, 33))
    A = ((rng.integers(0, P, (m, rank)) @ rng.integers(0, P, (rank, w))) % P).astype(np.uint32)
    if seed % 4 == 1:
        A[: min(m, 64), int(rng.integers(0, w))] = 0       # a column whose pivot lies below row 64
    if seed % 4 == 2 and m > 10:
        A[1:10] = A[0]                                     # dependent rows among the first 64
    r, piv, prow, _ = O.rref_rule(P, A, want_rref=False)
    want = list(zip(prow.tolist(), piv.tolist()))
    assert _speculative(A)[0] == want


def test_speculative_settles_and_defers():
    """The shortcut settles a dense full-rank panel and defers a panel whose column is zero in the
    first 64 rows; accepting its partial result (a plausible mistake) changes the pivots."""
    rng = np.random.default_rng(5)
    A = rng.integers(0, P, (400, 32)).astype(np.uint32)
    assert _speculative(A)[1]
    B = A.copy()
    B[:64, 7] = 0
    order, settled = _speculative(B)
    assert not settled
    assert _speculative(B, accept_partial=True)[0] != order


PW = 2147483647
_RINV = pow(1 << 32, -1, PW)


def _dfree(v, x, fneg, y):
    return (v * x + fneg * y) * _RINV % PW


def _spec_w(A, stale_read=False):
    """rrp_spec_kernel's Gauss-Jordan with coefficient tracking on the first rows of a panel, as the
    kernel runs it (Montgomery steps, column q left stale, the pivot row read as zero left of q):
    (pivot order, W = S[P, J]^-1) or None when a column has no pivot there."""
    rows, wc = len(A), len(A[0])
    pan = [list(r) for r in A]
    sig = [1] * rows
    co = [[0] * wc for _ in range(rows)]
    chosen, order = [False] * rows, []
    for q in range(wc):
        cand = [u for u in range(rows) if not chosen[u] and pan[u][q]]
        if not cand:
            return None
        bu, t = cand[0], len(order)
        v, prow, pco, sb = pan[bu][q], list(pan[bu]), list(co[bu]), sig[bu]
        for u in range(rows):
            f = pan[u][q]
            if u == bu or f == 0:
                continue
            fn = PW - f
            for j in range(wc):
                if j != q:
                    y = prow[j] if (stale_read or j > q) else 0
                    pan[u][j] = _dfree(v, pan[u][j], fn, y)
            sig[u] = _dfree(v, sig[u], fn, 0)
            for t2 in range(t):
                co[u][t2] = _dfree(v, co[u][t2], fn, pco[t2])
            co[u][t] = _dfree(v, 0, fn, sb)
        chosen[bu] = True
        order.append((bu, q))
    W = []
    for t, (u, q) in enumerate(order):
        di = pow(pan[u][q], -1, PW)
        W.append([(co[u][t2] + (sig[u] if t2 == t else 0)) * di % PW for t2 in range(len(order))])
    return order, W


def _inverse(K):
    k = len(K)
    M = [list(K[i]) + [int(i == j) for j in range(k)] for i in range(k)]
    for c in range(k):
        piv = next(i for i in range(c, k) if M[i][c])
        M[c], M[piv] = M[piv], M[c]
        iv = pow(M[c][c], -1, PW)
        M[c] = [x * iv % PW for x in M[c]]
        for i in range(k):
            if i != c and M[i][c]:
                f = M[i][c]
                M[i] = [(a - f * b) % PW for a, b in zip(M[i], M[c])]
    return [r[k:] for r in M]


@pytest.mark.parametrize("seed", range(6))
def test_spec_gauss_jordan_gives_inverse(seed):
    """The spec search's pivots equal the smallest-row rule's and its W is exactly S[P, J]^-1; the
    stale pivot-column read (a plausible mistake) breaks W."""
    rng = np.random.default_rng(2000 + seed)
    rows, wc = 24, [8, 12, 5, 8, 10, 6][seed]
    A = rng.integers(0, PW, (rows, wc)).astype(object)
    if seed % 2:
        A[2:5] = (A[0] * 3 + A[1] * 7) % PW          # dependent rows among the candidates
    A = [[int(x) for x in r] for r in A]
    res = _spec_w(A)
    assert res is not None
    order, W = res
    M = [r[:] for r in A]
    done, want = [False] * rows, []
    for q in range(wc):                              # the rule, plain modular arithmetic
        b = next(i for i in range(rows) if not done[i] and M[i][q])
        done[b] = True
        for i in range(rows):
            if i != b and not done[i] and M[i][q]:
                f, iv = M[i][q], pow(M[b][q], -1, PW)
                M[i] = [(a - f * iv * c) % PW for a, c in zip(M[i], M[b])]
        want.append((b, q))
    assert order == want
    K = [[A[u][q2] for _, q2 in order] for u, _ in order]
    assert W == _inverse(K)
    assert _spec_w(A, stale_read=True)[1] != W
