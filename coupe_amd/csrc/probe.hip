// probe.hip -- STREAM-style bandwidth probe of the device (measurement only,
// built as libcoupe_probe.so and used by bench.py).  Four f32 arrays are read
// and four written, the shape of one partition of the RCB engine: it gives
// the practical HBM ceiling that the partition's achieved bandwidth is set
// against, next to the 8 TB/s data-sheet peak.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads)
k_copy4(const float *__restrict__ a0, const float *__restrict__ a1, const float *__restrict__ a2,
        const float *__restrict__ a3, float *__restrict__ b0, float *__restrict__ b1, float *__restrict__ b2,
        float *__restrict__ b3, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        b0[i] = a0[i];
        b1[i] = a1[i];
        b2[i] = a2[i];
        b3[i] = a3[i];
    }
}

__global__ void __launch_bounds__(kThreads)
k_read4(const float *__restrict__ a0, const float *__restrict__ a1, const float *__restrict__ a2,
        const float *__restrict__ a3, uint64_t n, float *__restrict__ sink) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    float s = 0.0f;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
        s += a0[i] + a1[i] + a2[i] + a3[i];
    if (s == -1.0f) sink[0] = s;   // keeps the loads; never true for the zeroed inputs
}

__global__ void __launch_bounds__(kThreads)
k_write4(float *__restrict__ b0, float *__restrict__ b1, float *__restrict__ b2, float *__restrict__ b3,
         uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const float v = (float)(i & 1023);
        b0[i] = v;
        b1[i] = v;
        b2[i] = v;
        b3[i] = v;
    }
}

}  // namespace

// Best bandwidth (GB/s) over grids of 4 and 8 workgroups per CU, `reps`
// timed launches each, for 4 x n floats copied / read / written.  Returns 0,
// or a HIP error code.
extern "C" int coupe_probe_stream(int device, uint64_t n, int reps, double *copy_gbps, double *read_gbps,
                                  double *write_gbps) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    float *buf = nullptr;
    e = hipMalloc(&buf, 8 * n * sizeof(float) + 256);
    if (e != hipSuccess) return (int)e;
    float *a[4], *b[4];
    for (int k = 0; k < 4; k++) {
        a[k] = buf + (uint64_t)k * n;
        b[k] = buf + (uint64_t)(4 + k) * n;
    }
    float *sink = buf + 8 * n;
    int cus = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best[3] = {0, 0, 0};
    if ((e = hipMemset(buf, 0, 8 * n * sizeof(float))) != hipSuccess) goto done;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) goto done;
    for (int per_cu : {4, 8}) {
        const unsigned g = (unsigned)(cus * per_cu);
        for (int v = 0; v < 3; v++) {
            auto launch = [&]() {
                if (v == 0) k_copy4<<<g, kThreads>>>(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], n);
                else if (v == 1) k_read4<<<g, kThreads>>>(a[0], a[1], a[2], a[3], n, sink);
                else k_write4<<<g, kThreads>>>(b[0], b[1], b[2], b[3], n);
            };
            launch();
            if ((e = hipDeviceSynchronize()) != hipSuccess) goto done;
            if ((e = hipEventRecord(e0, nullptr)) != hipSuccess) goto done;
            for (int r = 0; r < reps; r++) launch();
            if ((e = hipEventRecord(e1, nullptr)) != hipSuccess) goto done;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) goto done;
            float ms = 0.0f;
            if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) goto done;
            const double bytes = (v == 0 ? 32.0 : 16.0) * (double)n * reps;
            const double gbps = bytes / (ms * 1e-3) / 1e9;
            if (gbps > best[v]) best[v] = gbps;
        }
    }
    *copy_gbps = best[0];
    *read_gbps = best[1];
    *write_gbps = best[2];
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return (int)e;
}
