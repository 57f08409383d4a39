// hx_howell.hip -- the Howell form over Z/2^k (1 <= k <= 32) of many uneven matrices in ONE
// persistent launch (ring2k::howell, Binary/Ring2k.h:959-969; ENGINE_API.md Sec. 13.5).
//
// One 256-thread workgroup per job, jobs pulled from an atomic counter (largest first).  Per job the
// Sec. 13.5 pass runs on the rows of A (copied into the job's work rows), in native u32 wraparound
// (Z/2^k is the quotient of Z/2^32, Ring2k.h:17-22; only valuations, zero tests and the output look
// at the value mod 2^k):
//   per column c: the active row of MINIMAL valuation, ties to the topmost active row (active order =
//   input rows, then closure rows in creation order; a chosen row leaves the list); the pivot row is
//   scaled by the inverse of its unit part (Newton, to the full 32-bit width) so it leads with 2^v;
//   every other active row r gets r -= ((r[c] mod 2^k) >> v) * P (exact: v is minimal); when v > 0 the
//   closure row 2^(k-v) * P joins the active list at the bottom.
// Then reduce above, pivots left to right: H[i'] -= ((H[i'][c_i] mod 2^k) >> v_i) * H[i] for i' < i.
// The Howell form is unique (ENGINE_API.md Sec. 13.5), so this equals the reference's echelon-first
// route (echelon_blocked + howell_pass) entry for entry.
#include <algorithm>
#include <numeric>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int HT = 256;

struct HJob {
    uint32_t m, n;
    const uint32_t* A;
    int64_t lda;
    uint32_t* H;
    uint32_t* pivots;
    uint32_t* pivot_v;
    uint32_t* h;
    uint32_t* work;       // (m + n) x n rows, then (m + n) active ids, then n pivot row ids
};

struct HArgs {
    const HJob* jobs;
    const uint32_t* order;
    uint32_t njobs;
    uint32_t* next;
    uint32_t k;
};

__device__ __forceinline__ uint32_t inv_odd(uint32_t u) {
    uint32_t x = u;   // u * u == 1 mod 8: 3 correct bits, each Newton step doubles them
#pragma unroll
    for (int i = 0; i < 4; ++i) x = x * (2u - u * x);
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off));
    return x;
}

__device__ void howell_one(const HJob& job, uint32_t k) {
    __shared__ uint32_t s_red[2][HT / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m = job.m, n = job.n;
    const uint32_t mask = k >= 32 ? 0xffffffffu : ((1u << k) - 1u);
    if (n == 0) {
        if (tid == 0) *job.h = 0;
        return;
    }
    uint32_t* rows = job.work;                               // row id i at rows + i * n
    uint32_t* act = job.work + (size_t)(m + n) * n;          // active order: row ids, or ~0u once chosen
    uint32_t* prow = act + (m + n);                          // pivot row ids in pivot order
    for (uint64_t idx = tid; idx < (uint64_t)m * n; idx += HT) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        rows[(size_t)i * n + j] = job.A[(int64_t)i * job.lda + j] & mask;
    }
    for (uint32_t i = tid; i < m; i += HT) act[i] = i;
    __syncthreads();
    uint32_t na = m, h = 0, ncl = 0;
    for (uint32_t c = 0; c < n; ++c) {
        // 1. minimal valuation, ties to the topmost active row: key = v << 24 | position
        uint32_t best = 0xffffffffu;
        for (uint32_t pos = tid; pos < na; pos += HT) {
            const uint32_t id = act[pos];
            if (id == 0xffffffffu) continue;
            const uint32_t x = rows[(size_t)id * n + c] & mask;
            if (x == 0) continue;
            const uint32_t v = (uint32_t)__builtin_ctz(x);
            best = min(best, (v << 24) | pos);
        }
        best = wave_min_u32(best);
        if (lane == 0) s_red[c & 1][wave] = best;
        __syncthreads();
        best = min(min(s_red[c & 1][0], s_red[c & 1][1]), min(s_red[c & 1][2], s_red[c & 1][3]));
        if (best == 0xffffffffu) continue;
        const uint32_t bpos = best & 0xffffffu, v = best >> 24;
        const uint32_t pid = act[bpos];
        uint32_t* P = rows + (size_t)pid * n;
        // 2. scale by the inverse of the unit part (P[c] == 2^v afterwards)
        const uint32_t ui = inv_odd(P[c] >> v);
        __syncthreads();                                     // everyone has read act[bpos] and P[c]
        for (uint32_t j = c + tid; j < n; j += HT) P[j] = (P[j] * ui) & mask;
        if (tid == 0) act[bpos] = 0xffffffffu;
        __syncthreads();
        // 3. eliminate every other active row (one wave per row, lanes along the row)
        for (uint32_t pos = wave; pos < na; pos += HT / 64) {
            const uint32_t id = act[pos];
            if (id == 0xffffffffu) continue;
            uint32_t* R = rows + (size_t)id * n;
            const uint32_t x = R[c] & mask;
            if (x == 0) continue;
            const uint32_t q = x >> v;
            for (uint32_t j = c + lane; j < n; j += 64) R[j] -= q * P[j];
        }
        // 4. the closure row 2^(k-v) P at the bottom of the active list
        if (v > 0) {
            const uint32_t cid = m + ncl;
            uint32_t* Z = rows + (size_t)cid * n;
            const uint32_t sh = k - v;
            for (uint32_t j = tid; j < n; j += HT) Z[j] = j <= c ? 0u : (sh >= 32 ? 0u : (P[j] << sh));
            if (tid == 0) act[na] = cid;
            ++na;
            ++ncl;
        }
        if (tid == 0) {
            prow[h] = pid;
            job.pivots[h] = c;
            job.pivot_v[h] = v;
        }
        ++h;
        __syncthreads();
    }
    // H rows in pivot order (masked), then reduce above, pivots left to right
    for (uint64_t idx = tid; idx < (uint64_t)h * n; idx += HT) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        job.H[(size_t)i * n + j] = rows[(size_t)prow[i] * n + j] & mask;
    }
    __syncthreads();
    for (uint32_t i = 1; i < h; ++i) {
        const uint32_t ci = job.pivots[i], vi = job.pivot_v[i];
        const uint32_t* Hi = job.H + (size_t)i * n;
        for (uint32_t ip = wave; ip < i; ip += HT / 64) {
            uint32_t* Hp = job.H + (size_t)ip * n;
            const uint32_t q = (Hp[ci] & mask) >> vi;
            if (q == 0) continue;
            for (uint32_t j = ci + lane; j < n; j += 64) Hp[j] = (Hp[j] - q * Hi[j]) & mask;
        }
        __syncthreads();
    }
    if (tid == 0) *job.h = h;
    __syncthreads();
}

__global__ __launch_bounds__(HT) void howell_batch_kernel(HArgs a) {
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t jj = s_job;
        __syncthreads();
        if (jj >= a.njobs) break;
        howell_one(a.jobs[a.order[jj]], a.k);
    }
}

__global__ void zero_counter_kernel(uint32_t* x) {
    if (threadIdx.x == 0) x[0] = 0;
}

struct HBatch {
    uint32_t k = 0, njobs = 0, grid = 0;
    HJob* d_jobs = nullptr;
    uint32_t* d_order = nullptr;
    uint32_t* d_next = nullptr;
};

void destroy(HBatch* b) {
    if (!b) return;
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->d_next) (void)hipFree(b->d_next);
    delete b;
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API size_t hx_howell_work_words(uint32_t m, uint32_t n) {
    return (size_t)(m + n) * n + (m + n) + n;
}

HX_API int hx_howell_batch_create(uint32_t k, const hx_howell_job* jobs, uint32_t njobs, hx_howell_batch* out) {
    return guard("hx_howell_batch_create", [&] {
        HX_CHECK(k >= 1 && k <= 32, "howell: need 1 <= k <= 32");
        HX_CHECK(out != nullptr && (jobs != nullptr || njobs == 0), "NULL argument");
        HBatch* b = new HBatch();
        try {
            b->k = k;
            b->njobs = njobs;
            std::vector<HJob> hj(njobs);
            for (uint32_t i = 0; i < njobs; ++i) {
                const hx_howell_job& j = jobs[i];
                HX_CHECK(j.n == 0 || (j.H && j.pivots && j.pivot_v && j.work), "howell job: NULL pointer");
                HX_CHECK(j.h != nullptr, "howell job: h is NULL");
                HX_CHECK(j.m == 0 || j.A != nullptr, "howell job: A is NULL");
                HX_CHECK(j.lda >= (int64_t)j.n, "howell job: lda < n");
                HX_CHECK((uint64_t)j.m + j.n < (1u << 24), "howell job: too many rows");
                hj[i] = HJob{j.m, j.n, j.A, j.lda, j.H, j.pivots, j.pivot_v, j.h, j.work};
            }
            std::vector<uint32_t> order(njobs);
            std::iota(order.begin(), order.end(), 0u);
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
                return (double)hj[x].m * hj[x].n * hj[x].n > (double)hj[y].m * hj[y].n * hj[y].n;
            });
            HX_HIP(hipMalloc(&b->d_jobs, std::max<size_t>(njobs, 1) * sizeof(HJob)));
            HX_HIP(hipMalloc(&b->d_order, std::max<size_t>(njobs, 1) * sizeof(uint32_t)));
            HX_HIP(hipMalloc(&b->d_next, sizeof(uint32_t)));
            if (njobs) {
                HX_HIP(hipMemcpy(b->d_jobs, hj.data(), njobs * sizeof(HJob), hipMemcpyHostToDevice));
                HX_HIP(hipMemcpy(b->d_order, order.data(), njobs * sizeof(uint32_t), hipMemcpyHostToDevice));
            }
            int cus = 256;
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, 0) == hipSuccess) cus = prop.multiProcessorCount;
            b->grid = std::max<uint32_t>(1, std::min<uint32_t>(njobs, (uint32_t)cus * 4));
        } catch (...) {
            destroy(b);
            throw;
        }
        *out = b;
    });
}

HX_API int hx_howell_batch_run(hx_howell_batch vb, hx_stream s) {
    return guard("hx_howell_batch_run", [&] {
        HX_CHECK(vb != nullptr, "NULL batch");
        HBatch* b = (HBatch*)vb;
        if (b->njobs == 0) return;
        zero_counter_kernel<<<1, 64, 0, (hipStream_t)s>>>(b->d_next);
        after_launch("zero_counter_kernel");
        HArgs a{b->d_jobs, b->d_order, b->njobs, b->d_next, b->k};
        howell_batch_kernel<<<b->grid, HT, 0, (hipStream_t)s>>>(a);
        after_launch("howell_batch_kernel");
    });
}

HX_API int hx_howell_batch_destroy(hx_howell_batch vb) {
    return guard("hx_howell_batch_destroy", [&] { destroy((HBatch*)vb); });
}

}  // extern "C"
