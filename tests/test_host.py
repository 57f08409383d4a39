"""CPU-only checks: the C-ABI library loads and exports every symbol the header declares (no compute
call without a GPU), the modulus context matches ModCtx.h's constants, and the synthetic workload
generators agree with the oracle's restatement of the reference generators."""
import math
import ctypes as C

import numpy as np

from hacs_amd import _lib


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_modctx_constants(O):
    L = _lib.lib()
    for p in [2, 3, 127, 251, 255, 256, 65279, 65280, 65521, 1000003, 16711423, 16711424, 2147483629,
              2147483647, 4278124287, 4278124288, 4294967291]:
        c = _lib.hx_modctx()
        _lib.check(L.hx_modctx_init(p, C.byref(c)))
        assert c.p == p and c.barrett_m == O.barrett_m(p) and c.k_max == O.k_max(p)
        assert c.two64_mod == (1 << 64) % p
        half = (p - 1) // 2
        want = next((l for l in (1, 2, 3, 4) if half <= 127 * (256 ** l - 1) // 255), 0)
        assert c.limbs == want, (p, c.limbs, want)
    c = _lib.hx_modctx()
    assert L.hx_modctx_init(1, C.byref(c)) == _lib.HX_E_ARGUMENT


def test_library_reports_no_device_without_gpu():
    L = _lib.lib()
    if L.hx_available():
        return
    assert L.hx_unavailable_reason()
    c = _lib.hx_device_info()
    assert L.hx_device_info_get(C.byref(c)) == _lib.HX_E_UNAVAILABLE


def test_workload_generators_match_oracle(O):
    from hacs_amd import workloads as W
    for (r, c, ell, salt) in [(64, 48, 65521, 1), (8192, 8192, 2147483629, 2), (4096, 0, 1000003, 0x41)]:
        assert W.seed_for(r, c, ell, salt) == O.seed_for(r, c, ell, salt)
    seed = W.seed_for(100, 30, 2147483629, 1)
    assert np.array_equal(W.random_entries(3000, 2147483629, seed), O.random_entries(3000, 2147483629, seed))
    assert np.array_equal(W.engine_matrix(64, 48, 65521, 1), O.engine_matrix(64, 48, 65521, 1))
    p = 2147483647
    A = O.ranked_matrix(96, 80, 30, p)
    assert np.array_equal(W.overwrite_rows(A, p, 8), O.overwrite_rows(A, p, 8))
    sh = W.c4_shapes(4096)
    assert sh.min() >= 16 and sh.max() <= 512 and abs(np.log(sh).mean() - np.log(16 * 32 ** 0.5)) < 0.05
    J = W.sparse_job(3, 40, 50, W.P_C4)
    nz = J[J != 0]
    assert set(np.unique(nz)) <= {1, 2, W.P_C4 - 1, W.P_C4 - 2}
    assert 0.02 < (J != 0).mean() < 0.09


def test_limb_k_chunk_contract():
    """hx_gemm_limb_k_chunk (host arithmetic, imma_k_chunk's role): the chosen K chunk's limb planes
    fit the workspace, the next 64 more do not (maximality), it is a multiple of 64 unless it is the
    whole of K, never above the int32-exact chunk, 0 exactly when one 64-deep chunk does not fit, and
    the full workspace runs K in one piece."""
    from hacs_amd import engine as E
    for p in (65521, 16777213, 2147483629):
        ctx = E.ModCtx(p)
        L, kmax = ctx.limbs, 2147483647 // (ctx.limbs * 16384) // 64 * 64
        for (m, n, k) in [(1, 1, 1), (300, 260, 1000), (8192, 8192, 8192), (64, 128, 40000), (5, 700, 63)]:
            per_k = L * (-(-m // 64) * 64 + -(-n // 128) * 128)
            full = E.gemm_workspace_bytes(ctx, m, n, k)
            assert E.limb_k_chunk(ctx, m, n, k, full) == min(k, kmax)
            for nbytes in [0, per_k * 64 - 1, per_k * 64, per_k * 200, per_k * 1000 + 5, full - 1, full, 1 << 40]:
                kc = E.limb_k_chunk(ctx, m, n, k, nbytes)
                if kc == 0:
                    assert nbytes < per_k * 64 and nbytes < full
                    continue
                assert kc <= min(k, kmax)
                assert kc == min(k, kmax) or kc % 64 == 0
                assert per_k * (-(-kc // 64) * 64) <= nbytes, (p, m, n, k, nbytes, kc)
                if kc < min(k, kmax):
                    assert per_k * (kc + 64) > nbytes


def test_arena_stats_without_device():
    """The arena's counters are readable without a device (all zero until something is allocated)."""
    from hacs_amd import engine as E
    st = E.arena_stats()
    assert set(st) >= {"reserved_bytes", "live_bytes", "driver_mallocs", "reuses", "oom_retries"}
    assert st["live_bytes"] <= st["reserved_bytes"]


def test_split_k_policy():
    """The split-K count the automatic policy takes (host arithmetic): 1 while the output tiles alone
    fill the GPU (256 of 64 x 64) or k < 1024, else enough K slices of >= 256 for about 1024 workgroups,
    at most 64, never an empty slice; and the workspace a split product needs (tile counters, then
    one 64 x 64 block of u32 partials per tile and slice)."""
    L = _lib.lib()
    assert L.hx_gemm_suggest_splits(8192, 8192, 8192) == 1
    assert L.hx_gemm_suggest_splits(1024, 1024, 100000) == 1          # 256 tiles
    assert L.hx_gemm_suggest_splits(64, 64, 1000) == 1                # k < 1024
    assert L.hx_gemm_suggest_splits(256, 256, 65536) == 64            # 16 tiles: min(64, 65536 / 256)
    assert L.hx_gemm_suggest_splits(500, 500, 4096) == 16             # 64 tiles: min(16, 4096 / 256)
    assert L.hx_gemm_suggest_splits(64, 64, 1100) == 4                # 1 tile: min(1024, 4) slices of 288
    assert L.hx_gemm_suggest_splits(0, 64, 5000) == 1
    for (m, n, k, s) in [(256, 256, 65536, 64), (100, 70, 3333, 7), (1, 1, 40000, 64)]:
        tiles = -(-m // 64) * -(-n // 64)
        kc = -(-(-(-k // s)) // 32) * 32
        eff = -(-k // kc)
        want = -(-tiles * 4 // 256) * 256 + tiles * eff * 64 * 64 * 4
        assert L.hx_gemm_split_workspace_bytes(m, n, k, s) == want, (m, n, k, s)
    # a count above what k allows shrinks (no empty slice): k = 40 in slices of 32 is 2 slices
    assert L.hx_gemm_split_workspace_bytes(64, 64, 40, 64) == 256 + 2 * 64 * 64 * 4
    assert L.hx_gemm_split_workspace_bytes(64, 64, 5000, 1) == 0


def test_residue_moduli_count():
    """hx_gemm_residue_moduli (host arithmetic): the least N whose moduli product exceeds 2.05 k (p/2)^2, checked
    against that product computed here in Python integers, at the workload's modulus, a small prime and the
    largest 32-bit one, with the route's k limit (k 127^2 < 2^31) and k = 0 declined."""
    from hacs_amd import engine as E
    qs = [255, 254, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211]
    for a in range(len(qs)):
        for b in range(a + 1, len(qs)):
            assert math.gcd(qs[a], qs[b]) == 1

    def least(p, k):
        bound, M = 41 * k * (p // 2) ** 2, 1
        for i, q in enumerate(qs):
            M *= q
            if 20 * M > bound:
                return i + 1
        return 0

    for p in (2, 3, 251, 65521, 16777213, 2147483629, 4294967291):
        for k in (1, 64, 700, 8192, 32768, 65536, 133144):
            assert E.residue_moduli(p, k) == least(p, k), (p, k)
    assert E.residue_moduli(2147483629, 8192) == 10          # the 8192^3 workload: 10 int8 products, not 16
    assert E.residue_moduli(2147483629, 133144) == 10        # 11 under a margin of 8 instead of 2.05
    assert E.residue_moduli(2147483629, 133145) == 0         # int32 sums would pass 2^31
    assert E.residue_moduli(2147483629, 0) == 0
    assert 133144 * 127 * 127 <= 2 ** 31 - 1 < 133145 * 127 * 127


def test_gemm_route_policy():
    """hx_gemm_route (host arithmetic): the default takes the residue route (6) for four-limb moduli at
    m, n, k >= 4096 and the limb kernel 4 below that or for fewer limbs; explicit tiles are kept (6 only where
    the route applies); PLAIN and AUTO's small products run the plain kernel (0)."""
    from hacs_amd import engine as E
    c4, c4s, c2 = E.ModCtx(2147483629), E.ModCtx(16777213), E.ModCtx(65521)
    assert (c4.limbs, c4s.limbs, c2.limbs) == (4, 4, 3)
    assert E.gemm_route(c4, 8192, 8192, 8192) == 6 and E.gemm_route(c4, 4096, 4096, 4096) == 6
    assert E.gemm_route(c4, 4096, 4096, 2048) == 4 and E.gemm_route(c4, 2048, 8192, 8192) == 4
    assert E.gemm_route(c4s, 8192, 8192, 8192) == 6 and E.residue_moduli(16777213, 8192) == 8
    assert E.gemm_route(c2, 8192, 8192, 8192) == 4                       # 6 moduli against 9 limb products
    assert E.gemm_route(c4, 8192, 8192, 200000) == 4                     # k past the int32 bound
    for t in (1, 2, 3, 4, 5):
        assert E.gemm_route(c4, 8192, 8192, 8192, E.HX_GEMM_LIMB, t) == t
    assert E.gemm_route(c4, 64, 64, 64, E.HX_GEMM_LIMB, 6) == 6
    assert E.gemm_route(c4, 64, 64, 200000, E.HX_GEMM_LIMB, 6) == 4
    assert E.gemm_route(c4, 8192, 8192, 8192, E.HX_GEMM_PLAIN) == 0 and E.gemm_route(c4, 8, 8, 8, E.HX_GEMM_AUTO) == 0
    assert E.gemm_workspace_bytes(c4, 8192, 8192, 8192) == 10 * 3 * 8192 * 8192
    assert E.gemm_workspace_bytes(c4, 2048, 2048, 2048) == 4 * (2048 + 2048) * 2048


def test_residue_split_arithmetic_exhaustive():
    """The residue route's operand split (res_bytes4 in hx_gemm.hip), restated in numpy for every t a byte dot
    product can give (t < 4 * 255^2 + 256) and every modulus: qt = trunc(fma(t, fl32(1/q), 1/2)) -- the
    fma's single rounding emulated by rounding the exact float64 value t * fl32(1/q) + 1/2 (42 significant
    bits at most) to float32 -- leaves r = t - q qt congruent to t with |r| <= 127, so the byte is a balanced
    residue and the products' int32 bound k 127^2 holds.  The byte dot product itself: t = sum_j byte_j(w)
    (2^8j mod q) + q - (2^31 mod q) is congruent to w - 2^31 modulo q for random words w."""
    qs = np.array([255, 254, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211], np.int64)
    t = np.arange(4 * 255 * 255 + 256, dtype=np.int64)
    for q in qs:
        iq = np.float64(np.float32(1.0) / np.float32(q))
        f = (t.astype(np.float64) * iq + 0.5).astype(np.float32)
        qt = np.trunc(f).astype(np.int64)
        r = t - q * qt
        assert np.abs(r).max() <= 127 and not ((t - r) % q).any(), int(q)
    rng = np.random.default_rng(5)
    w = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.int64)
    for q in qs:
        tp = [1, 256 % q, 65536 % q, (1 << 24) % q]
        t2 = sum(((w >> (8 * j)) & 255) * tp[j] for j in range(4)) + q - (1 << 31) % q
        assert t2.max() < 4 * 255 * 255 + 256 and not ((t2 - (w - (1 << 31))) % q).any(), int(q)


def test_residue_crt_reconstruction():
    """The residue route's CRT (the product epilogue's s_i = v y_i mod q_i, then crt_kernel in hx_gemm.hip),
    restated with Python integers: for sums x at the route's bound |x| <= k (p/2)^2 (both extremes, zero,
    +-1, random), with N = hx_gemm_residue_moduli(p, k), the 16-bit-split sums of s_i (M / q_i mod p), the
    fixed-point quotient T = round(sum_i s_i round(2^20 / q_i) / 2^20) and x = sum - T (M mod p) give x mod p
    exactly; with only as many moduli as keep their product below the bound, the extreme fails (a bound is
    needed; the route's factor 2.05 over the 2 exactness needs is margin for the fixed-point quotient)."""
    from hacs_amd import engine as E
    qs = [255, 254, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211]

    def crt(x, p, N):
        M = 1
        for q in qs[:N]:
            M *= q
        hi = lo = fr = 0
        for q in qs[:N]:
            Mi = M // q
            y = pow(Mi % q, -1, q)
            s = ((x % q) * y) % q                      # the epilogue's byte (v is congruent to x)
            c = Mi % p
            hi += s * (c >> 16)
            lo += s * (c & 0xFFFF)
            fr += s * ((2**20 + q // 2) // q)
        T = (fr + 2**19) >> 20
        return ((hi << 16) + lo + T * (p - M % p)) % p

    rng = np.random.default_rng(7)
    for p in (2147483629, 4294967291, 16777213, 65521):
        for k in (64, 8192, 133144):
            N = E.residue_moduli(p, k)
            assert N >= 1
            bound = k * (p // 2) ** 2
            xs = [bound, -bound, 0, 1, -1] + [int(v) for v in rng.integers(-2**62, 2**62, 20)]
            xs += [bound * int(a) // 1000 for a in rng.integers(-1000, 1001, 20)]
            for x in xs:
                x = max(-bound, min(bound, x))
                assert crt(x, p, N) == x % p, (p, k, x)
            M, nf = 1, 0                              # the most moduli whose product is still below the bound
            while nf < len(qs) and M * qs[nf] < bound:
                M *= qs[nf]
                nf += 1
            if nf >= 1:
                assert crt(bound, p, nf) != bound % p, (p, k, nf)
