// A C++ caller of the engine facade (include/hacs_mi355x_engine.hpp), written the way the reference's
// device callers use engine::cuda: views, GemmOp, gemm / GemmBatch / trsm_unit, imma_rref's row-slab
// form with source / sink callbacks, CudaError on a contract breach.  Checks every result against
// exact __int128 arithmetic; prints one line per check and exits 0 only if all pass.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/hacs_mi355x_engine.hpp"

static int g_fail = 0;
static void expect(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "[PASS]" : "[FAIL]", what);
    if (!ok) ++g_fail;
}

static std::vector<uint32_t> gemm_ref(uint32_t p, const std::vector<uint32_t>& A, const std::vector<uint32_t>& B,
                                      uint32_t m, uint32_t n, uint32_t k) {
    std::vector<uint32_t> C((size_t)m * n);
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            unsigned __int128 s = 0;
            for (uint32_t t = 0; t < k; ++t) s += (unsigned __int128)A[(size_t)i * k + t] * B[(size_t)t * n + j];
            C[(size_t)i * n + j] = (uint32_t)(s % p);
        }
    return C;
}

int main() {
    if (!hxe::available()) {
        std::printf("[FAIL] no device: %s\n", hxe::unavailable_reason());
        return 1;
    }
    std::mt19937_64 rng(42);
    hxe::Stream s = hxe::create_stream();
    const uint32_t p = 2147483629u, m = 130, n = 70, k = 300;
    hxe::ModCtx32 ctx(p);
    std::vector<uint32_t> A((size_t)m * k), B((size_t)k * n), C((size_t)m * n);
    for (auto& x : A) x = (uint32_t)(rng() % p);
    for (auto& x : B) x = (uint32_t)(rng() % p);
    const std::vector<uint32_t> want = gemm_ref(p, A, B, m, n, k);
    hxe::DeviceBuf dA(m, k), dB(k, n), dC(m, n);
    dA.upload(A.data(), s);
    dB.upload(B.data(), s);
    for (int variant : {HX_GEMM_PLAIN, HX_GEMM_LIMB}) {
        hxe::GemmOpts o;
        o.variant = variant;
        hxe::gemm(ctx, dA.view(), dB.view(), dC.view(), hxe::GemmOp::Set, s, o);
        dC.download(C.data(), s);
        expect(C == want, variant == HX_GEMM_PLAIN ? "gemm (plain) == exact" : "gemm (limb) == exact");
    }
    // GemmBatch of two uneven jobs sharing A
    hxe::DeviceBuf dC2(m, n);
    hxe::GemmBatch batch;
    batch.push(dA.view(), dB.view(), dC2.view(), hxe::GemmOp::Set);
    batch.push(dA.view().sub(0, 50, 0, k), dB.view(), dC.view().sub(0, 50, 0, n), hxe::GemmOp::Sub);
    batch.upload(ctx);
    batch.run(s);
    std::vector<uint32_t> C2((size_t)m * n);
    dC2.download(C2.data(), s);
    dC.download(C.data(), s);
    bool zero = true;
    for (uint32_t i = 0; i < 50 * n; ++i) zero = zero && C[i] == 0;   // C held A B: C - A B = 0 on those rows
    expect(C2 == want && zero, "GemmBatch (uneven jobs, Set and Sub)");
    // level-1 / tape support: C2 = 3 A B + C2 (axpy), its transpose, a gathered column block; launch counters
    hxe::reset_launch_stats();
    hxe::axpy(ctx, 3u, dC.view(), dC2.view(), false, s);
    hxe::DeviceBuf dT(n, m);
    hxe::transpose(dC2.view(), dT.view(), s);
    std::vector<uint32_t> T((size_t)n * m);
    dT.download(T.data(), s);
    bool tr_ok = true;
    for (uint32_t i = 0; i < m && tr_ok; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t c_ij = i < 50 ? 0u : want[(size_t)i * n + j];     // dC: A B with rows < 50 cleared
            const uint64_t e = ((uint64_t)want[(size_t)i * n + j] + 3ull * c_ij) % p;
            if (T[(size_t)j * m + i] != (uint32_t)e) { tr_ok = false; break; }
        }
    expect(tr_ok, "axpy then transpose == exact");
    const hxe::LaunchStats ls = hxe::launch_stats();
    expect(ls.kernel_launches == 2 && ls.gemm_launches == 0, "launch_stats counts the two launches");
    // contract breach -> CudaError
    bool threw = false;
    try {
        hxe::gemm(ctx, dA.view(), dA.view(), dC.view(), hxe::GemmOp::Set, s);
    } catch (const hxe::CudaError& e) {
        threw = e.code() == HX_E_ARGUMENT;
    }
    expect(threw, "shape mismatch throws CudaError(HX_E_ARGUMENT)");
    // imma_rref's row-slab form: a rank-2 matrix over F_p
    const uint32_t rm = 5, rn = 4;
    std::vector<uint32_t> M = {1, 2, 3, 4, 2, 4, 6, 8, 0, 1, 1, 1, 1, 3, 4, 5, 0, 0, 0, 0};
    std::vector<uint32_t> R((size_t)rm * rn, 7);
    hxe::RrefResult res = hxe::rref(
        ctx, rm, rn,
        [&](uint32_t r0, uint32_t, int64_t* ld) -> const uint32_t* { *ld = rn; return M.data() + (size_t)r0 * rn; },
        [&](uint32_t, int64_t* ld) -> uint32_t* { *ld = rn; return R.data(); }, 2);
    const bool rows_ok = R[0] == 1 && R[1] == 0 && R[2] == 1 && R[3] == 2 && R[4] == 0 && R[5] == 1 && R[6] == 1 &&
                         R[7] == 1;
    expect(res.rank == 2 && res.pivots == std::vector<uint32_t>({0, 1}) &&
               res.pivot_rows == std::vector<uint32_t>({0, 2}) && rows_ok,
           "rref (row-slab source / sink): rank, pivots, pivot_rows, rows");
    hxe::destroy_stream(s);
    std::printf("%s\n", g_fail ? "FAILED" : "ALL PASSED");
    return g_fail ? 1 : 0;
}
