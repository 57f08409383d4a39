// binding_probe.cpp -- links hacs_amd/binding/EngineHip.cpp with the reference's own ModArithmetic
// objects (oracle/_ref) and libhacs_mi355x.so, installs the hooks and runs the reference's
// elim::rref / ring2k::echelon_blocked through them.  Without a device the hooks decline and the
// reference's host code runs (same values); with one and ENGINE_DEVICE=gpu they take the device.
// Prints one line per check; exit status 0 when every check holds.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CudaCtx.h"
#include "CudaKernels.h"
#include "DenseBuf.h"
#include "DenseElim.h"
#include "ModCtx.h"
#include "Binary/Ring2k.h"

namespace engine { namespace cuda { namespace hip { void install_hooks(); } } }

int main() {
    using namespace engine;
    int bad = 0;
    auto check = [&](bool ok, const char* what) {
        std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
        if (!ok) ++bad;
    };
    cuda::hip::install_hooks();
    check(elim::device_rref_hook() != nullptr, "dense RREF hook installed");
    check(ring2k::device::ops() != nullptr, "ring2k device Ops installed");
    std::printf("device: %s\n", cuda::available() ? "available" : cuda::unavailable_reason());
    // a rank-deficient 300 x 260 matrix mod 2^31 - 1 through the reference's elim::rref
    const uint32_t p = 2147483647u, m = 300, n = 260;
    ModCtx<uint32_t> ctx(p);
    DenseBuf<uint32_t> A(m, n);
    uint64_t x = 12345;
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            A.at(i, j) = (i % 3 == 2) ? 0u : (uint32_t)((x >> 33) % p);
        }
    const elim::RrefResult r = elim::rref<uint32_t>(ctx, view(A));
    check(r.rank == 200, "elim::rref rank of the 300 x 260 test matrix is 200");
    bool lead = true;
    for (uint32_t t = 0; t < r.rank; ++t) lead = lead && A.at(t, r.pivots[t]) == 1u;
    check(lead, "RREF rows lead with 1 at their pivots");
    // the blocked LU over Z/2^32 through echelon_blocked (the device Ops are asked, and decline off-device)
    ring2k::Mat<uint32_t> M(120, 90);
    for (uint32_t i = 0; i < 120; ++i)
        for (uint32_t j = 0; j < 90; ++j) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            M.at(i, j) = (uint32_t)(x >> 32) << (i % 5);
        }
    const ring2k::EchelonInfo e = ring2k::echelon_blocked<ring2k::Traits<uint32_t>>(M, 32, ring2k::PivotRule::MinValuation, 32);
    check(e.rank == 90 && e.pivots.size() == 90, "ring2k::echelon_blocked rank 90");
    const cuda::LaunchStats ls = cuda::launch_stats();
    std::printf("launches: %llu\n", (unsigned long long)ls.kernel_launches);
    return bad == 0 ? 0 : 1;
}
