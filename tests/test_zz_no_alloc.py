"""hx_gemm_u32_noalloc, the GEMM entry point that allocates nothing (its limb planes live in a fixed
file-scope device array, K chunked to fit): called through the C-ABI, its output equals hx_gemm_u32's
byte for byte on two calls in a row (the second smaller), for every variant; the library's arena
counters do not move; a shape whose one 64-deep K chunk exceeds the array is refused (LIMB) or runs
on the plain kernel (AUTO)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noalloc():
    from hacs_amd import _lib
    L = _lib.lib()
    return L.hx_gemm_u32_noalloc, int(L.hx_gemm_noalloc_scratch_bytes())


def _call(f, E, ctx, A, B, Cm, op, variant, tile=0):
    from hacs_amd import _lib
    o = _lib.hx_gemm_opts()
    o.variant, o.tile = variant, tile
    return f(C.byref(ctx.c), E.view(A), E.view(B), E.view(Cm), op, E.current_stream(), C.byref(o))


def test_noalloc_equals_gemm_two_calls(hx_gpu, O):
    from hacs_amd import engine as E
    from hacs_amd import _lib
    f, cap = _noalloc()
    assert cap == 64 << 20
    p = 2147483629
    ctx = E.ModCtx(p)
    # first call: K runs in three chunks of the array (4 limbs x (4096 + 4096) x 2048 bytes = 64 MiB);
    # second call: a smaller product in one chunk
    for (m, n, k) in [(4096, 4096, 6000), (300, 260, 1000)]:
        print(f"noalloc GEMM {m}x{n}x{k} mod {p}", flush=True)
        A = E.random_entries(m, k, p, 0x11 + m)
        B = E.random_entries(k, n, p, 0x12 + n)
        C0 = E.random_entries(m, n, p, 0x13 + k)
        for variant in (E.HX_GEMM_LIMB, E.HX_GEMM_AUTO, E.HX_GEMM_PLAIN):
            for op in (E.GemmOp.Set, E.GemmOp.Sub):
                want = C0.clone()
                E.gemm(ctx, A, B, want, op, variant=variant)
                torch.cuda.synchronize()
                st0 = E.arena_stats()
                got = C0.clone()
                torch.cuda.synchronize()
                _lib.check(_call(f, E, ctx, A, B, got, op, variant))
                torch.cuda.synchronize()
                st1 = E.arena_stats()
                assert st1["driver_mallocs"] == st0["driver_mallocs"] and st1["live_bytes"] == st0["live_bytes"]
                assert torch.equal(got, want), f"{(m, n, k)} variant {variant} op {op}"
        if m * n * k <= 1e8:
            got = torch.empty((m, n), dtype=torch.int32, device="cuda")
            _lib.check(_call(f, E, ctx, A, B, got, E.GemmOp.Set, E.HX_GEMM_LIMB))
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(got), O.gemm_fast(p, E.from_device(A), E.from_device(B)))


def test_noalloc_limit(hx_gpu, O):
    """limbs x (round_up(m, 64) + round_up(n, 128)) x 64 bytes above the array: LIMB fails with
    HX_E_ARGUMENT before launching anything, AUTO runs the plain kernel and matches hx_gemm_u32."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    f, cap = _noalloc()
    p = 2147483629
    ctx = E.ModCtx(p)
    m, n, k = 270000, 1, 64                # 4 x (270016 + 128) x 64 > 64 MiB
    assert 4 * (-(-m // 64) * 64 + 128) * 64 > cap
    A = E.random_entries(m, k, p, 0x21)
    B = E.random_entries(k, n, p, 0x22)
    got = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    assert _call(f, E, ctx, A, B, got, E.GemmOp.Set, E.HX_GEMM_LIMB) == _lib.HX_E_ARGUMENT
    assert "scratch" in _lib.lib().hx_last_error().decode()
    _lib.check(_call(f, E, ctx, A, B, got, E.GemmOp.Set, E.HX_GEMM_AUTO))
    want = torch.zeros_like(got)
    E.gemm(ctx, A, B, want, E.GemmOp.Set, variant=E.HX_GEMM_PLAIN)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
