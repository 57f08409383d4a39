"""The oracle, pinned independently of its own formulas: published test vectors, closed forms,
exact integer arithmetic, uniqueness properties, brute force on tiny inputs, the compiled
reference (oracle/_ref) where it was built, and negative controls that plant a mistake in the
oracle's inputs or rule and must be caught."""
import itertools

import numpy as np
import pytest

from tests.matrix_kinds import KINDS, make


def test_splitmix_and_fnv_vectors(O):
    # SplitMix64 seeded 0: the published first outputs
    assert O.splitmix64_stream(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # FNV-1a 64 of the empty input is the offset basis; of one zero byte is basis ^ 0 * prime
    M = np.zeros((0, 0), np.uint32)
    h0 = O.fnv_matrix(M)
    basis, prime = 0xCBF29CE484222325, 0x100000001B3
    want = basis
    for _ in range(16):          # rows = 0 and cols = 0 fed as 8 bytes each
        want = ((want ^ 0) * prime) & ((1 << 64) - 1)
    assert h0 == want


def test_modular_primitives(O):
    for p in [2, 3, 65521, 1000003, 2147483629, 2147483647, 4294967291]:
        assert O.barrett_m(p) == (1 << 64) // p
        k = O.k_max(p)
        assert k * (p - 1) ** 2 <= (1 << 64) - 1 < (k + 1) * (p - 1) ** 2     # tightness
        for x in [0, 1, p - 1, p, p * p - 1, (1 << 64) - 1, (1 << 63) + 12345]:
            assert O.reduce_u64(p, x) == x % p
        rng = np.random.default_rng(p)
        for a in rng.integers(1, p, 20, dtype=np.uint64):
            assert (int(a) * O.inv(int(a), p)) % p == 1
    assert O.inv(3, 15) == 0 and O.inv(0, 7) == 0


@pytest.mark.parametrize("p", [2, 251, 65521, 1000003, 2147483629, 4294967291])
def test_gemm_exact(O, p):
    rng = np.random.default_rng(p)
    for (m, n, k) in [(1, 1, 1), (7, 5, 300), (33, 17, 1025)]:
        A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
        Cc = rng.integers(0, p, (m, n), dtype=np.uint64).astype(np.uint32)
        exact = A.astype(object).dot(B.astype(object))
        for op in (0, 1, 2):
            base = 0 if op == 0 else Cc.astype(object)
            want = ((base + exact) if op < 2 else (base - exact)) % p
            want = np.asarray(want, dtype=object).astype(np.uint64).astype(np.uint32)
            assert np.array_equal(O.gemm_naive(p, A, B, Cc, op), want)
            assert np.array_equal(O.gemm_fast(p, A, B, Cc, op, 3), want)
    # rank-1 x all-(p-1): closed form sum_k a_i b_k (p-1) = -a_i sum_k b_k
    u = rng.integers(0, p, (9, 1), dtype=np.uint64)
    v = rng.integers(0, p, (1, 700), dtype=np.uint64)
    A = ((u * v) % p).astype(np.uint32)
    B = np.full((700, 4), p - 1, np.uint32)
    row = [(-(int(u[i, 0]) * sum(int(x) for x in v[0])) % p) for i in range(9)]
    assert np.array_equal(O.gemm_fast(p, A, B), np.array([[r] * 4 for r in row], np.uint32))


def test_ranked_matrix_rank(O):
    p = 2147483647
    A = O.ranked_matrix(80, 70, 25, p)
    assert O.rref_naive(p, A)[0] == 25
    B = O.overwrite_rows(A, p, 10)
    assert O.rref_naive(p, B)[0] == 25          # rows replaced by combinations of other rows


@pytest.mark.parametrize("p", [2, 3, 13, 65521, 2147483647])
def test_rref_oracles_agree_and_rules(O, p):
    rng = np.random.default_rng(p + 1)
    for kind in KINDS:
        for (m, n) in [(1, 1), (12, 9), (9, 12), (40, 40)]:
            A = make(rng, m, n, p, kind, rank=max(1, min(m, n) // 2))
            r, piv, R = O.rref_naive(p, A)
            r2, piv2, prow, R2 = O.rref_rule(p, A, threads=3)
            assert r == r2 and np.array_equal(piv, piv2) and np.array_equal(R[:r], R2) and not R[r:].any()
            # the pivot rows are the row rank profile: the pivot columns of RREF(A^T)
            assert np.array_equal(np.sort(prow), O.rref_naive(p, A.T.copy())[1])
            # uniqueness: a row permutation changes neither the RREF nor its pivots
            perm = rng.permutation(m)
            r3, piv3, R3 = O.rref_naive(p, A[perm])
            assert r3 == r and np.array_equal(piv3, piv) and np.array_equal(R3, R)


def test_rref_rule_negative_control(O):
    """A planted order where the largest-row rule would disagree: pivot_rows must be [0, 2], not [1, 2]
    (rows 0 and 1 are equal; the smallest original row wins)."""
    p = 13
    A = np.array([[1, 2, 3], [1, 2, 3], [0, 1, 5]], np.uint32)
    r, piv, prow, _ = O.rref_rule(p, A)
    assert r == 2 and piv.tolist() == [0, 1] and prow.tolist() == [0, 2]


def test_rref_brute_force_rank_tiny(O):
    """Rank over F_3 of every 2x3 matrix by counting the row space's size."""
    p = 3
    for vals in itertools.product(range(3), repeat=6):
        A = np.array(vals, np.uint32).reshape(2, 3)
        span = {tuple((a * A[0] + b * A[1]) % 3) for a in range(3) for b in range(3)}
        assert O.rref_naive(p, A)[0] == {1: 0, 3: 1, 9: 2}[len(span)]


@pytest.mark.parametrize("k", [1, 2, 3, 7, 16, 31])
def test_howell_oracles(O, k):
    rng = np.random.default_rng(k)
    for (m, n) in [(1, 1), (4, 6), (7, 3), (20, 20), (30, 45)]:
        A = rng.integers(0, 1 << k, (m, n), dtype=np.uint64).astype(np.uint32)
        A2 = ((A.astype(np.uint64) << rng.integers(0, k, (m, 1)).astype(np.uint64)) % (1 << k)).astype(np.uint32)
        for X in (A, A2):
            H1, p1, v1 = O.howell_naive(k, X)
            H2, p2, v2 = O.howell_fast(k, X)
            assert np.array_equal(H1, H2) and np.array_equal(p1, p2) and np.array_equal(v1, v2)
            # uniqueness: row permutations and appended zero rows give the same Howell form
            perm = rng.permutation(m)
            Y = np.vstack([X[perm], np.zeros((2, n), np.uint32)])
            H3, p3, v3 = O.howell_fast(k, Y)
            assert np.array_equal(H1, H3) and np.array_equal(p1, p3) and np.array_equal(v1, v3)
    if k == 2:
        H, piv, pv = O.howell_fast(2, np.array([[2, 1]], np.uint32))
        assert H.tolist() == [[2, 1], [0, 2]] and piv.tolist() == [0, 1] and pv.tolist() == [1, 1]


def test_howell_k1_equals_rref(O):
    rng = np.random.default_rng(9)
    A = rng.integers(0, 2, (25, 30), dtype=np.uint64).astype(np.uint32)
    H, piv, pv = O.howell_fast(1, A)
    r, piv2, R = O.rref_naive(2, A)
    assert np.array_equal(H, R[:r]) and np.array_equal(piv, piv2) and not pv.any()


def test_mlp_oracle_finite_differences():
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(2)
    X = rng.normal(size=(16, 5))
    W1, b1 = rng.normal(size=(7, 5)), rng.normal(size=7)
    W2, b2 = rng.normal(size=(3, 7)), rng.normal(size=3)
    T = rng.normal(size=(16, 3))
    mask = (rng.random((16, 3)) < 0.6).astype(float)
    _, dW1, db1, dW2, db2 = MO.backward(X, W1, b1, W2, b2, T, mask)
    eps = 1e-6
    for P, G in ((W1, dW1), (b1, db1), (W2, dW2), (b2, db2)):
        for idx in [tuple(rng.integers(0, s) for s in P.shape) for _ in range(5)]:
            P[idx] += eps
            lp = MO.loss(X, W1, b1, W2, b2, T, mask)
            P[idx] -= 2 * eps
            lm = MO.loss(X, W1, b1, W2, b2, T, mask)
            P[idx] += eps
            assert abs((lp - lm) / (2 * eps) - G[idx]) < 1e-5 * max(1.0, abs(G[idx]))


@pytest.mark.parametrize("p", [2, 65521, 2147483647, 4294967291])
def test_level1_oracles_through_gemm(O, p):
    """axpy / scal / gather / axpy_identity of the oracle, each re-derived as a k <= 2 product through
    the oracle's C GEMM (a separate code path with its own reduction) and checked against Python
    integers; plus canonicalise on the whole Lazy2p range and the 64-bit reduction."""
    rng = np.random.default_rng(p)
    x = rng.integers(0, p, (7, 9), dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, p, (7, 9), dtype=np.uint64).astype(np.uint32)
    for alpha in {0, 1, p - 1, int(rng.integers(0, p))}:
        for sub in (False, True):
            got = O.axpy(p, alpha, x, y, subtract=sub)
            coef = np.array([[(p - alpha) % p if sub else alpha], [1]], np.uint32)
            want = O.gemm_naive(p, np.stack([x.ravel(), y.ravel()], 1), coef).reshape(x.shape)
            assert np.array_equal(got, want)
            i, j = 3, 5
            assert int(got[i, j]) == (int(y[i, j]) + (-1 if sub else 1) * alpha * int(x[i, j])) % p
        assert np.array_equal(O.scal(p, alpha, x), O.gemm_naive(p, x.reshape(-1, 1), np.array([[alpha]], np.uint32)).reshape(x.shape))
    # canonicalise: every v < 2p maps to v mod p
    v = np.concatenate([np.arange(0, min(2 * p, 1 << 32), max(1, (2 * p) // 97), dtype=np.uint64),
                        np.array([p - 1, p, min(2 * p - 1, (1 << 32) - 1)], np.uint64)]).astype(np.uint32)
    assert [int(t) for t in O.canonicalise(p, v)] == [int(t) % p for t in v]
    acc = np.array([[0, 1, (1 << 64) - 1], [p, p * p - 1, (1 << 63) + 7]], dtype=np.uint64)
    assert [int(t) for t in O.canonicalise_acc64(p, acc).ravel()] == [int(t) % p for t in acc.ravel()]
    # gather: the selected entries times beta plus alpha on the shifted identity, as one k = 2 product
    a = rng.integers(0, p, (11, 13), dtype=np.uint64).astype(np.uint32)
    rows = np.array([10, 0, 4, 4, 7], np.uint32)
    cols = np.array([12, 3, 0, 9, 1, 5], np.uint32)
    c = rng.integers(0, p, (5, 6), dtype=np.uint64).astype(np.uint32)
    alpha, beta, off = int(rng.integers(1, p)) if p > 2 else 1, int(rng.integers(1, p)) if p > 2 else 1, 1
    E = np.array([[1 if off + rows[i] == cols[t] else 0 for t in range(6)] for i in range(5)], np.uint32)
    assert E.sum() > 0
    S = a[rows][:, cols]
    for acc_ in (False, True):
        got = O.gather(p, alpha, beta, a, off, cols, c, accumulate=acc_, rows=rows)
        Aop = np.stack([E.ravel(), S.ravel()] + ([c.ravel()] if acc_ else []), 1)
        coef = np.array([[alpha], [beta]] + ([[1]] if acc_ else []), np.uint32)
        assert np.array_equal(got, O.gemm_naive(p, Aop, coef).reshape(c.shape))
    # identity source (a absent): alpha on the cells with row_off + i == t
    got = O.gather(p, alpha, beta, None, 2, None, c)
    assert np.array_equal(got, ((np.eye(5, 6, 2, dtype=np.uint64) * alpha) % p).astype(np.uint32))
    # axpy_identity == axpy with an identity x on the window's diagonal
    y2 = rng.integers(0, p, (4, 6), dtype=np.uint64).astype(np.uint32)
    for sub in (False, True):
        assert np.array_equal(O.axpy_identity(p, alpha, y2, sub),
                              O.axpy(p, alpha, np.eye(4, 6, dtype=np.uint32), y2, subtract=sub))


def test_level1_oracle_negative_control(O):
    """A planted mistake (the identity tested against r(i) without row_off) must be caught by the
    product form above."""
    p = 65521
    c = np.zeros((4, 4), np.uint32)
    good = O.gather(p, 5, 0, None, 1, None, c)
    planted = O.gather(p, 5, 0, None, 0, None, c)
    assert not np.array_equal(good, planted)
    assert np.array_equal(good, (np.eye(4, 4, 1, dtype=np.uint32) * 5))


def test_fnv_published_vectors_and_word_order(O):
    """FNV-1a-64's published vectors pin the byte step (offset basis, prime, xor-then-multiply); the
    matrix checksum equals that byte step over (rows, cols, entries) as little-endian 64-bit words,
    which pins the word width and byte order; FNV-1 (multiply-then-xor) and big-endian words differ."""
    assert O.fnv1a_bytes(b"") == 0xCBF29CE484222325
    assert O.fnv1a_bytes(b"a") == 0xAF63DC4C8601EC8C
    assert O.fnv1a_bytes(b"foobar") == 0x85944171F73967E8
    M = O.engine_matrix(5, 9, 2147483629, 0x07)
    words = np.concatenate([np.array([5, 9], np.uint64), M.astype(np.uint64).ravel()])
    assert O.fnv_matrix(M) == O.fnv1a_bytes(words.astype("<u8").tobytes())
    assert O.fnv_matrix(M) != O.fnv1a_bytes(words.astype(">u8").tobytes())
    h = 0xCBF29CE484222325
    for byte in words.astype("<u8").tobytes():      # FNV-1: multiply, then xor
        h = ((h * 0x100000001B3) & ((1 << 64) - 1)) ^ byte
    assert O.fnv_matrix(M) != h


def test_fnv_against_reference_hasher(O):
    """The oracle's checksum equals the reference's own engine::FnvHasher / FnvBytes on the fixture
    written with oracle/_ref (tests/golden/make_fnv_golden.py)."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fnv_reference.npz"),
                allow_pickle=False)
    M = G["matrix"]
    assert O.fnv_matrix(M) == int(G["words_fnv"][0])
    assert np.array_equal(G["words"][2:], M.astype(np.uint64).ravel())
    for s, n, want in zip(G["strings"], G["string_lens"], G["strings_fnv"]):
        assert O.fnv1a_bytes(bytes(s)[:int(n)]) == int(want)


def test_r2_echelon_vs_reference(O):
    """The oracle's blocked LU over Z/2^k (u32 / u64 rungs) equals the reference's own echelon_blocked
    entrywise on the fixture (matrix, rank, pivots, valuations); the Howell form mod 2^k of the result
    equals the input's (row operations only, all unimodular)."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "r2_reference.npz"),
                allow_pickle=False)
    for i in range(int(G["count"][0])):
        elem, k, m, n, panel, climit, r = (int(x) for x in G[f"case{i}_meta"])
        A = G[f"case{i}_A"]
        R, rank, piv, pv = O.r2_echelon(k, A, panel, climit)
        assert rank == r and np.array_equal(piv, G[f"case{i}_piv"]) and np.array_equal(pv, G[f"case{i}_pv"]), i
        assert np.array_equal(R, G[f"case{i}_R"]), f"case {i}: matrix"
        if k <= 32:
            mk = np.uint64((1 << k) - 1)
            Ain = (A.astype(np.uint64) & mk).astype(np.uint32)
            Rout = (R.astype(np.uint64) & mk).astype(np.uint32)
            H0, p0, v0 = O.howell_fast(k, Ain)
            H1, p1, v1 = O.howell_fast(k, Rout)
            assert np.array_equal(H0, H1) and np.array_equal(p0, p1) and np.array_equal(v0, v1), f"case {i}: rowspan"


def test_r2_echelon_negative_control(O):
    """The comparison is entrywise on the native words, not only mod 2^k: adding 2^k to one input entry
    leaves the pivots and valuations alone (they read values mod 2^k) but must change the matrix."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "r2_reference.npz"),
                allow_pickle=False)
    elem, k, m, n, panel, climit, r = (int(x) for x in G["case0_meta"])
    assert elem == 4 and k < 32
    A = G["case0_A"].copy()
    A[20, 5] = np.uint32((int(A[20, 5]) + (1 << k)) & 0xFFFFFFFF)
    R, rank, piv, pv = O.r2_echelon(k, A, panel, climit)
    assert rank == r and np.array_equal(piv, G["case0_piv"]) and np.array_equal(pv, G["case0_pv"])
    assert not np.array_equal(R, G["case0_R"])


def test_gemm_u64_exact_and_reference(O):
    """The u64 GEMM oracle: exact Python integers (every op, ragged shapes, k = 0, even and odd moduli
    up to 2^64 - 2), the all-(p-1) closed form K (p-1)^2 = K mod p, and the reference's own
    kernels::gemm<uint64_t> on the committed fixture."""
    import os
    rng = np.random.default_rng(64)
    for p in (2305843009213693951, 18446744073709551557, (1 << 64) - 2, 4294967311, 7, 1 << 40):
        for (m, n, k) in ((5, 7, 9), (17, 3, 200), (3, 4, 0)):
            def draw(shape):
                return ((rng.integers(0, 1 << 32, shape, dtype=np.uint64) << np.uint64(32)) |
                        rng.integers(0, 1 << 32, shape, dtype=np.uint64)) % np.uint64(p)
            A, B, C0 = draw((m, k)), draw((k, n)), draw((m, n))
            ex = A.astype(object).dot(B.astype(object)) if k else np.zeros((m, n), dtype=object)
            for op in (0, 1, 2):
                want = (ex if op == 0 else (C0.astype(object) + ex if op == 1 else C0.astype(object) - ex)) % p
                assert (O.gemm_u64(p, A, B, C0, op).astype(object) == want).all(), (p, m, n, k, op)
    p, K = 18446744073709551557, 3000
    A = np.full((2, K), p - 1, np.uint64)
    assert (O.gemm_u64(p, A, A.T.copy()) == np.uint64(K % p)).all()
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "u64_reference.npz"),
                allow_pickle=False)
    for i in range(int(G["count"][0])):
        p = int(G[f"case{i}_p"][0])
        A, B, C0 = G[f"case{i}_A"], G[f"case{i}_B"], G[f"case{i}_C0"]
        assert np.array_equal(O.gemm_u64(p, A, B), G[f"case{i}_set"]), i
        assert np.array_equal(O.gemm_u64(p, A, B, C0, 2), G[f"case{i}_sub"]), i


def _smith_levels(sigma, nlev):
    return np.bincount(np.asarray(sigma, np.int64), minlength=nlev)[:nlev] if nlev else np.zeros(0, np.int64)


def test_smith_oracle_vs_reference_fixture(O):
    """The oracle's Smith valuations (a restatement of smith_naive) equal the reference's own ring2k::smith
    by levels on the committed fixture (sigma; level_ranks = the count of each valuation over the levels run)."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "smith_reference.npz"),
                allow_pickle=False)
    for i in range(int(G["count"][0])):
        k = int(G[f"case{i}_k"][0])
        sig, lev = G[f"case{i}_sigma"], G[f"case{i}_levels"]
        assert np.array_equal(O.smith(k, G[f"case{i}_A"]), sig), f"case {i}"
        assert np.array_equal(_smith_levels(sig, len(lev)), lev), f"case {i}"


def test_smith_oracle_closed_forms(O):
    """What the mathematics fixes: a diagonal 2^v u (u odd) has the sorted v < k as valuations, also under
    permutations and unit-triangular row / column mixing (unimodular); the zero matrix has none; at k = 1 the
    count is the F_2 rank; multiplying A by 2^t adds t to every valuation and drops those reaching k."""
    rng = np.random.default_rng(0x5311)
    for (m, n, k, dt) in ((30, 40, 32, np.uint32), (45, 25, 61, np.uint64), (20, 20, 7, np.uint32)):
        r = min(m, n) - 3
        vals = np.sort(rng.integers(0, k, r))
        D = np.zeros((m, n), dt)
        for i, v in enumerate(vals):
            D[i, i] = dt(((int(rng.integers(0, 1 << 30)) | 1) << int(v)) & int(np.iinfo(dt).max))
        assert np.array_equal(O.smith(k, D), vals)
        A = D[rng.permutation(m)][:, rng.permutation(n)]
        with np.errstate(over="ignore"):
            for d in (1, 2, 4, 7):
                A[d:] += rng.integers(0, 1 << 31, (m - d, 1)).astype(dt) * A[:-d]
                A[:, d:] += A[:, :-d] * rng.integers(0, 1 << 31, (1, n - d)).astype(dt)
        assert np.array_equal(O.smith(k, A), vals)
        t = 3
        with np.errstate(over="ignore"):
            A2 = (A << dt(t)).astype(dt)
        assert np.array_equal(O.smith(k, A2), (vals + t)[vals + t < k])
    assert len(O.smith(40, np.zeros((6, 9), np.uint64))) == 0
    B = rng.integers(0, 2, (37, 53)).astype(np.uint32)
    B[5] = B[3] ^ B[4]
    assert len(O.smith(1, B)) == O.rref_rule(2, B)[0]


def test_smith_oracle_brute_force_tiny(O):
    """On tiny matrices over Z/4 and Z/8 the valuations follow from the column span itself: im A = (+) Z/2^(k - v_i),
    so the number of its elements killed by 2^t is 2^(sum_i min(t, k - v_i)); counted by enumerating every x."""
    rng = np.random.default_rng(0x5312)
    for k in (2, 3):
        q = 1 << k
        for (m, n) in ((2, 2), (2, 3), (3, 2), (3, 3)):
            for _ in range(12):
                A = rng.integers(0, q, (m, n)).astype(np.int64)
                if rng.random() < 0.3:
                    A[rng.integers(0, m)] *= 2
                xs = np.array(list(itertools.product(range(q), repeat=n)), np.int64)
                img = np.unique((xs @ A.T) % q, axis=0)
                c = [0]
                for t in range(1, k + 1):
                    killed = int(((img * (1 << t)) % q == 0).all(axis=1).sum())
                    c.append(int(round(np.log2(killed))))
                # #{i : k - v_i >= t} = c_t - c_{t-1}
                want = []
                for t in range(k, 0, -1):
                    cnt = (c[t] - c[t - 1]) - (sum(1 for v in want if k - v >= t + 1) if t < k else 0)
                    want += [k - t] * cnt
                want.sort()
                assert np.array_equal(O.smith(k, A.astype(np.uint32)), np.array(want, np.uint32)), (k, A.tolist())


def test_smith_oracle_negative_control(O):
    """A planted mistake (one entry of a unit diagonal made even) must change the valuations."""
    D = np.eye(6, dtype=np.uint32) * 5
    assert np.array_equal(O.smith(8, D), np.zeros(6, np.uint32))
    D[2, 2] = 6
    assert not np.array_equal(O.smith(8, D), np.zeros(6, np.uint32))


def test_howell_ell_oracle_vs_reference_fixture(O):
    """The oracle's Howell form over Z/ell^k equals the reference's own elim::howell (prime_power contexts)
    on the committed fixture: H, pivots and valuations, entry for entry."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "howell_ell_reference.npz"),
                allow_pickle=False)
    for i in range(int(G["count"][0])):
        ell, k = (int(x) for x in G[f"case{i}_meta"])
        H, piv, pv = O.howell_ell(ell, k, G[f"case{i}_A"])
        assert np.array_equal(H, G[f"case{i}_H"]), f"case {i}"
        assert np.array_equal(piv, G[f"case{i}_piv"]) and np.array_equal(pv, G[f"case{i}_pv"]), f"case {i}"


def test_howell_ell_oracle_special_cases(O):
    """What fixes the Howell form independently: at ell = 2 it is the Z/2^k form (the oracles of
    ring2k::howell); at k = 1 the RREF over F_ell; [[ell, 1]] over Z/ell^2 = [[ell, 1], [0, ell]]; the form
    of a row permutation, or with a zero row or a combination of rows appended, is the same."""
    rng = np.random.default_rng(0x4E12)
    for k in (3, 8, 16):
        A = rng.integers(0, 1 << k, (20, 14)).astype(np.uint32)
        A[::3] = (A[::3] * 4) % (1 << k)
        got = O.howell_ell(2, k, A)
        want = O.howell_fast(k, A)
        assert all(np.array_equal(x, y) for x, y in zip(got, want)), k
    for ell in (3, 7, 65521):
        A = rng.integers(0, ell, (15, 22)).astype(np.uint32)
        A[4] = (A[1].astype(np.uint64) * 2 % ell).astype(np.uint32)
        H, piv, pv = O.howell_ell(ell, 1, A)
        r, rpiv, _, R = O.rref_rule(ell, A)
        assert len(piv) == r and np.array_equal(piv, rpiv) and np.array_equal(H, R) and not pv.any()
    for ell in (3, 5, 251):
        H, piv, pv = O.howell_ell(ell, 2, np.array([[ell, 1]], np.uint32))
        assert np.array_equal(H, np.array([[ell, 1], [0, ell]], np.uint32))
        assert list(piv) == [0, 1] and list(pv) == [1, 1]
    ell, k = 5, 6
    q = ell ** k
    A = rng.integers(0, q, (12, 10), dtype=np.uint64)
    A[::2] = A[::2] * ell % q
    A = A.astype(np.uint32)
    base = O.howell_ell(ell, k, A)
    comb = ((A[3].astype(np.uint64) * 7 + A[8].astype(np.uint64) * 11) % q).astype(np.uint32)
    for B in (A[rng.permutation(12)], np.vstack([A, np.zeros((1, 10), np.uint32)]), np.vstack([A, comb[None, :]])):
        got = O.howell_ell(ell, k, B)
        assert all(np.array_equal(x, y) for x, y in zip(got, base))


def test_howell_ell_oracle_row_span_tiny(O):
    """Over Z/9 on 2 x 2 and 3 x 2 matrices: the row span of H, enumerated, equals that of A; H leads each
    row with exactly 3^v at strictly increasing columns, and the entries above a pivot are below 3^v."""
    import itertools
    rng = np.random.default_rng(0x4E13)
    ell, k, q = 3, 2, 9
    for (m, n) in ((2, 2), (3, 2), (2, 3)):
        for _ in range(15):
            A = rng.integers(0, q, (m, n)).astype(np.int64)
            if rng.random() < 0.4:
                A[0] = A[0] * 3 % q
            H, piv, pv = O.howell_ell(ell, k, A.astype(np.uint32))
            H = H.astype(np.int64)

            def span(M):
                if len(M) == 0:
                    return {tuple([0] * n)}
                return {tuple((np.array(c) @ M) % q) for c in itertools.product(range(q), repeat=len(M))}
            assert span(A) == span(H), A.tolist()
            assert all(piv[i] < piv[i + 1] for i in range(len(piv) - 1))
            for i, (c, v) in enumerate(zip(piv, pv)):
                assert H[i, c] == 3 ** v and not H[i, :c].any()
                assert all(H[ip, c] < 3 ** v for ip in range(i))


def test_howell_ell_oracle_negative_control(O):
    """A planted mistake (one entry of a non-pivot column changed by ell^(k-1)) must change the form."""
    rng = np.random.default_rng(0x4E14)
    ell, k = 7, 3
    q = ell ** k
    A = rng.integers(0, q, (3, 8)).astype(np.uint32)
    base = O.howell_ell(ell, k, A)
    B = A.copy()
    B[2, 7] = (int(B[2, 7]) + ell ** (k - 1)) % q
    got = O.howell_ell(ell, k, B)
    assert not all(np.array_equal(x, y) for x, y in zip(got, base))
