// imbalance.hip -- per-part loads and imbalance of a partition, on device
// (coupe/src/imbalance.rs:14-105).  Entry points in include/coupe_hip.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "rcb_engine.h"
#include "rcb_device.h"

namespace coupe_hip {

// ---------------------------------------------------------------------------
// Per-part loads (coupe::imbalance::compute_parts_load, imbalance.rs:14-39):
// a weight histogram over part ids, privatised in LDS when the parts fit.
// Integer weights add in 64-bit two's complement (i32 wrapping is recovered
// by truncation); f64 weights add in f64.  Ids >= num_parts are counted as
// errors (coupe's `acc[part]` would panic).
// ---------------------------------------------------------------------------
constexpr uint32_t kLoadLds = 4096;   // parts privatised in LDS (32 KB)

template <typename WIN, typename Acc>
__global__ void __launch_bounds__(kBlock)
k_parts_load(const uint64_t *__restrict__ ids, const WIN *__restrict__ w, uint64_t n, uint64_t num_parts,
             Acc *__restrict__ loads, unsigned long long *bad) {
    __shared__ Acc sl[kLoadLds];
    const bool lds = num_parts <= kLoadLds;
    if (lds)
        for (uint32_t j = threadIdx.x; j < num_parts; j += kBlock) sl[j] = (Acc)0;
    __syncthreads();
    uint32_t nbad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint64_t p = ids[i];
        if (p >= num_parts) {
            nbad++;
            continue;
        }
        const Acc v = AccOps<Acc>::of(w[i]);
        if (lds) atomic_add_acc(&sl[p], v);
        else atomic_add_acc(&loads[p], v);
    }
    nbad = wave_sum(nbad);
    if (lane_id() == 0 && nbad) atomicAdd(bad, (unsigned long long)nbad);
    __syncthreads();
    if (lds)
        for (uint32_t j = threadIdx.x; j < num_parts; j += kBlock)
            if (sl[j] != (Acc)0) atomic_add_acc(&loads[j], sl[j]);
}


namespace {

void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFailure{e};
}

// Loads in the accumulator type, copied to the host.
template <typename WIN, typename Acc>
void loads_typed(Context &cx, hipStream_t st, const uint64_t *ids, uint64_t n, const WIN *w, uint64_t num_parts,
                 std::vector<Acc> &out, uint64_t *bad) {
    cx.loads.ensure(std::max<uint64_t>(num_parts, 1) * sizeof(Acc) + 64);
    Acc *d = (Acc *)cx.loads.ptr;
    unsigned long long *d_bad = (unsigned long long *)((char *)cx.loads.ptr + num_parts * sizeof(Acc));
    d_bad = (unsigned long long *)(((uintptr_t)d_bad + 7) & ~(uintptr_t)7);
    ck(hipMemsetAsync(cx.loads.ptr, 0, cx.loads.bytes, st));
    if (n) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_parts_load<WIN, Acc>, kBlock, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 2;
        int cus = 0;
        ck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cx.device));
        uint64_t g = std::min<uint64_t>((uint64_t)cus * per_cu, (n + kBlock - 1) / kBlock);
        k_parts_load<WIN, Acc><<<(unsigned)std::max<uint64_t>(g, 1), kBlock, 0, st>>>(ids, w, n, num_parts, d, d_bad);
        ck(hipGetLastError());
    }
    out.resize(num_parts);
    unsigned long long hb = 0;
    if (num_parts) ck(hipMemcpyAsync(out.data(), d, num_parts * sizeof(Acc), hipMemcpyDeviceToHost, st));
    ck(hipMemcpyAsync(&hb, d_bad, sizeof hb, hipMemcpyDeviceToHost, st));
    ck(hipStreamSynchronize(st));
    *bad = hb;
}

}  // namespace

// Loads converted to the weight type (i32 wraps like Rust's release `+=`).
enum coupe_err parts_load(Context &cx, hipStream_t st, const uint64_t *ids, uint64_t n, int wkind, const void *w,
                          uint64_t num_parts, void *host_loads) {
    if (wkind != COUPE_INT && wkind != COUPE_INT64 && wkind != COUPE_DOUBLE) return COUPE_ERR_BAD_TYPE;
    if (n && (!ids || !w)) return COUPE_ERR_CRASH;
    try {
        ck(hipSetDevice(cx.device));
        uint64_t bad = 0;
        if (wkind == COUPE_DOUBLE) {
            std::vector<double> v;
            loads_typed<double, double>(cx, st, ids, n, (const double *)w, num_parts, v, &bad);
            if (bad) return COUPE_ERR_CRASH;
            memcpy(host_loads, v.data(), num_parts * sizeof(double));
        } else {
            std::vector<uint64_t> v;
            if (wkind == COUPE_INT)
                loads_typed<int32_t, uint64_t>(cx, st, ids, n, (const int32_t *)w, num_parts, v, &bad);
            else
                loads_typed<int64_t, uint64_t>(cx, st, ids, n, (const int64_t *)w, num_parts, v, &bad);
            if (bad) return COUPE_ERR_CRASH;
            for (uint64_t j = 0; j < num_parts; j++) {
                if (wkind == COUPE_INT) ((int32_t *)host_loads)[j] = (int32_t)(uint32_t)v[j];
                else ((int64_t *)host_loads)[j] = (int64_t)v[j];
            }
        }
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        return map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
    return COUPE_ERR_OK;
}

// imbalance.rs:41-77: max over parts of (load - ideal) / ideal, with
// ideal = (sum of the loads, in W) / num_parts; 0 when num_parts or ideal is 0.
enum coupe_err imbalance(Context &cx, hipStream_t st, const uint64_t *ids, uint64_t n, int wkind, const void *w,
                         uint64_t num_parts, double *out) {
    *out = 0.0;
    if (num_parts == 0) return COUPE_ERR_OK;
    std::vector<unsigned char> buf(num_parts * 8);
    enum coupe_err e = parts_load(cx, st, ids, n, wkind, w, num_parts, buf.data());
    if (e != COUPE_ERR_OK) return e;
    std::vector<double> f(num_parts);
    double total;
    if (wkind == COUPE_DOUBLE) {
        const double *l = (const double *)buf.data();
        double s = 0.0;
        for (uint64_t j = 0; j < num_parts; j++) { s += l[j]; f[j] = l[j]; }
        total = s;
    } else if (wkind == COUPE_INT) {
        const int32_t *l = (const int32_t *)buf.data();
        uint32_t s = 0;
        for (uint64_t j = 0; j < num_parts; j++) { s += (uint32_t)l[j]; f[j] = (double)l[j]; }
        total = (double)(int32_t)s;
    } else {
        const int64_t *l = (const int64_t *)buf.data();
        uint64_t s = 0;
        for (uint64_t j = 0; j < num_parts; j++) { s += (uint64_t)l[j]; f[j] = (double)l[j]; }
        total = (double)(int64_t)s;
    }
    const double ideal = total / (double)num_parts;
    if (ideal == 0.0) return COUPE_ERR_OK;
    double mx = -1.0 / 0.0;
    for (uint64_t j = 0; j < num_parts; j++) {
        const double r = (f[j] - ideal) / ideal;
        if (r > mx) mx = r;
    }
    *out = mx;
    return COUPE_ERR_OK;
}

}  // namespace coupe_hip
