// mesh.hip -- element barycentres on device, the points the coupe tools
// partition (tools/lib/lib.rs:511-539).  Entry point in include/coupe_hip.h.
#include <hip/hip_runtime.h>

#include <new>

#include "rcb_engine.h"
#include "rcb_device.h"

namespace coupe_hip {

namespace {

void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFailure{e};
}

// One thread per element: the sum of its nodes' coordinates in node order,
// starting from 0.0, then divided by the node count -- the reference's f64
// operations in the reference's order (no contraction).  A node index out
// of range is reported through `bad` and the element is skipped.
template <int D>
__global__ void __launch_bounds__(kBlock)
k_barycentres(const double *__restrict__ coords, uint64_t n_nodes, const int64_t *__restrict__ nodes, uint32_t k,
              uint64_t n_el, double *__restrict__ out, unsigned int *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n_el; e += stride) {
        double bc[D];
#pragma unroll
        for (int a = 0; a < D; a++) bc[a] = 0.0;
        bool ok = true;
        for (uint32_t j = 0; j < k; j++) {
            const int64_t v = nodes[e * k + j];
            if (v < 0 || (uint64_t)v >= n_nodes) {
                ok = false;
                break;
            }
#pragma unroll
            for (int a = 0; a < D; a++) bc[a] = bc[a] + coords[(uint64_t)v * D + a];
        }
        if (!ok) {
            atomicOr(bad, 1u);
            continue;
        }
#pragma unroll
        for (int a = 0; a < D; a++) out[e * D + a] = bc[a] / (double)k;
    }
}

}  // namespace

enum coupe_err barycentres(Context &cx, hipStream_t st, const double *coords, uint64_t n_nodes, size_t dim,
                           const int64_t *nodes, uint64_t n_el, size_t k, double *out) {
    if (dim != 2 && dim != 3) return COUPE_ERR_BAD_DIMENSION;
    if (k == 0 || k > 64) return COUPE_ERR_BAD_TYPE;
    if (n_el == 0) return COUPE_ERR_OK;
    if (!coords || !nodes || !out) return COUPE_ERR_CRASH;
    try {
        ck(hipSetDevice(cx.device));
        cx.loads.ensure(64);
        unsigned int *d_bad = (unsigned int *)cx.loads.ptr;
        ck(hipMemsetAsync(d_bad, 0, sizeof(unsigned int), st));
        const unsigned g = (unsigned)std::min<uint64_t>((n_el + kBlock - 1) / kBlock, 65536);
        if (dim == 2) k_barycentres<2><<<g, kBlock, 0, st>>>(coords, n_nodes, nodes, (uint32_t)k, n_el, out, d_bad);
        else k_barycentres<3><<<g, kBlock, 0, st>>>(coords, n_nodes, nodes, (uint32_t)k, n_el, out, d_bad);
        ck(hipGetLastError());
        unsigned int h_bad = 0;
        ck(hipMemcpyAsync(&h_bad, d_bad, sizeof h_bad, hipMemcpyDeviceToHost, st));
        ck(hipStreamSynchronize(st));
        if (h_bad) return COUPE_ERR_CRASH;   // coupe indexes the node array and panics
    } catch (const HipFailure &f) {
        (void)hipGetLastError();
        return map_hip_error(f.err);
    } catch (const std::bad_alloc &) {
        return COUPE_ERR_ALLOC;
    }
    return COUPE_ERR_OK;
}

}  // namespace coupe_hip
