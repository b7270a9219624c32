// mesh_parts.hip -- the points of a partition grouped by part: perm lists the
// point indices of part 0, then part 1, ..., each in increasing index order,
// and offsets[p] .. offsets[p + 1] is part p's range.  This is the order
// coupe's working set ends in after rcb() (its Items are permuted in place,
// coupe/src/algorithms/recursive_bisection.rs:122-181, :682-683), made stable
// so that it is deterministic.  Off the partition's hot path: a stable radix
// sort of (id, index) pairs (rocPRIM) and one boundary pass.
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "rcb_engine.h"

namespace coupe_hip {

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) k_max_id(const uint64_t *__restrict__ ids, uint64_t n, uint64_t *out) {
    __shared__ uint64_t s[kThreads / 64];
    uint64_t m = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) m = ids[i] > m ? ids[i] : m;
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t x = __shfl_xor(m, o, 64);
        m = x > m ? x : m;
    }
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < kThreads / 64; v++) m = s[v] > m ? s[v] : m;
        atomicMax((unsigned long long *)out, (unsigned long long)m);
    }
}

__global__ void __launch_bounds__(kThreads)
k_pairs(const uint64_t *__restrict__ ids, uint64_t n, uint32_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        keys[i] = (uint32_t)ids[i];
        idx[i] = (uint32_t)i;
    }
}

// offsets[p] = the first position whose key is >= p, for p in [0, num_parts].
__global__ void __launch_bounds__(kThreads)
k_bounds(const uint32_t *__restrict__ sorted, uint64_t n, uint64_t num_parts, uint64_t *__restrict__ off) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i <= n; i += stride) {
        const int64_t prev = i == 0 ? -1 : (int64_t)sorted[i - 1];
        const int64_t cur = i == n ? (int64_t)num_parts : (int64_t)sorted[i];
        for (int64_t p = prev + 1; p <= cur; p++) off[p] = i;
    }
}

unsigned grid_of(uint64_t n) { return (unsigned)std::min<uint64_t>((n + kThreads - 1) / kThreads + 1, 4096); }

}  // namespace

// perm: n entries; offsets is sized here (num_parts + 1 entries).
enum coupe_err group_by_part(hipStream_t st, const uint64_t *ids, uint64_t n, uint32_t *perm, DevBuf &offsets,
                             uint64_t *num_parts, DevBuf &scratch) {
    try {
        uint64_t top = 0;
        if (n) {
            scratch.ensure(256);
            if (hipMemsetAsync(scratch.ptr, 0, sizeof(uint64_t), st) != hipSuccess) throw HipFailure{hipErrorUnknown};
            k_max_id<<<grid_of(n), kThreads, 0, st>>>(ids, n, (uint64_t *)scratch.ptr);
            hipError_t r = hipMemcpyAsync(&top, scratch.ptr, sizeof top, hipMemcpyDeviceToHost, st);
            if (r == hipSuccess) r = hipStreamSynchronize(st);
            if (r != hipSuccess) throw HipFailure{r};
            top += 1;
        }
        if (top > 0xffffffffull || n > 0xffffffffull) return COUPE_ERR_ALLOC;   // 32-bit keys and indices
        *num_parts = top;
        offsets.ensure((top + 1) * sizeof(uint64_t));
        if (n == 0) {
            const uint64_t zero = 0;
            hipError_t r = hipMemcpyAsync(offsets.ptr, &zero, sizeof zero, hipMemcpyHostToDevice, st);
            if (r == hipSuccess) r = hipStreamSynchronize(st);
            return r == hipSuccess ? COUPE_ERR_OK : map_hip_error(r);
        }
        const int bits = top <= 1 ? 1 : 64 - __builtin_clzll(top - 1);
        size_t temp = 0;
        if (rocprim::radix_sort_pairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, bits, st) !=
            hipSuccess)
            throw HipFailure{hipErrorUnknown};
        const size_t arr = ((n * sizeof(uint32_t) + 255) / 256) * 256;
        scratch.ensure(3 * arr + temp + 256);
        char *base = (char *)scratch.ptr;
        uint32_t *keys = (uint32_t *)base, *keys_out = (uint32_t *)(base + arr), *idx = (uint32_t *)(base + 2 * arr);
        void *tmp = base + 3 * arr;
        k_pairs<<<grid_of(n), kThreads, 0, st>>>(ids, n, keys, idx);
        if (rocprim::radix_sort_pairs(tmp, temp, (const uint32_t *)keys, keys_out, (const uint32_t *)idx, perm,
                                      (size_t)n, 0, bits, st) != hipSuccess)
            throw HipFailure{hipErrorUnknown};
        k_bounds<<<grid_of(n + 1), kThreads, 0, st>>>(keys_out, n, top, (uint64_t *)offsets.ptr);
        hipError_t r = hipGetLastError();
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        if (r != hipSuccess) throw HipFailure{r};
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(st);
        return map_hip_error(f.err);
    }
    return COUPE_ERR_OK;
}

}  // namespace coupe_hip
