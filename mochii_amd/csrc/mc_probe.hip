// Device rate of scattered FP64 read-modify-writes, the access pattern of
// the walk's per-leaf tallies: the measured ceiling the walk's tally traffic
// is compared with (a streaming-bandwidth peak does not apply to updates of
// rows chosen at random in a tens-of-GB array).  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "mc_internal.h"

namespace {

constexpr int kProbeBlock = 256;

__device__ __forceinline__ uint64_t probe_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// op i adds 1 to the `width` doubles of row hash(i) mod rows (grid-stride;
// every lane exits once i passes n_ops)
__global__ __launch_bounds__(kProbeBlock) void k_probe_rmw(double *a, int64_t rows, int width,
                                                           int64_t n_ops)
{
    const int64_t stride = (int64_t)gridDim.x * kProbeBlock;
    for (int64_t i = (int64_t)blockIdx.x * kProbeBlock + threadIdx.x; i < n_ops; i += stride) {
        const uint64_t r = probe_mix((uint64_t)i) % (uint64_t)rows;
        double *row = a + r * (uint64_t)width;
        for (int k = 0; k < width; ++k) unsafeAtomicAdd(row + k, 1.0);
    }
}

// the transposed form: lane-op i adds 1 to double (i mod width) of row
// hash(i / width), so that the `width` lanes of a row issue its adds in one
// instruction (width divides 64 and 8: a row never straddles a 64-byte line)
__global__ __launch_bounds__(kProbeBlock) void k_probe_rows(double *a, int64_t rows, int width,
                                                            int64_t n_lane_ops)
{
    const int64_t stride = (int64_t)gridDim.x * kProbeBlock;
    for (int64_t i = (int64_t)blockIdx.x * kProbeBlock + threadIdx.x; i < n_lane_ops; i += stride) {
        const uint64_t r = probe_mix((uint64_t)(i / width)) % (uint64_t)rows;
        unsafeAtomicAdd(a + r * (uint64_t)width + (uint64_t)(i % width), 1.0);
    }
}

__global__ __launch_bounds__(kProbeBlock) void k_probe_fill(double *a, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * kProbeBlock;
    for (int64_t i = (int64_t)blockIdx.x * kProbeBlock + threadIdx.x; i < n; i += stride)
        a[i] = 0.0;
}

}  // namespace

static int probe_run(int64_t rows, int width, int64_t n_ops, bool together, double *ms,
                     double *checksum)
{
    const size_t n = (size_t)rows * (size_t)width;
    double *a = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * sizeof(double));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        mc_set_error("scattered-RMW probe: hipMalloc of " + std::to_string(n * 8) +
                     " bytes failed");
        return -1;
    }
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned grid = (unsigned)(ncu > 0 ? 8 * ncu : 2048);
    k_probe_fill<<<grid, kProbeBlock>>>(a, (int64_t)n);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // one untimed run touches the pages, then the timed one
    if (e == hipSuccess) {
        if (together) k_probe_rows<<<grid, kProbeBlock>>>(a, rows, width, n_ops * width);
        else k_probe_rmw<<<grid, kProbeBlock>>>(a, rows, width, n_ops);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
        if (together) k_probe_rows<<<grid, kProbeBlock>>>(a, rows, width, n_ops * width);
        else k_probe_rmw<<<grid, kProbeBlock>>>(a, rows, width, n_ops);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess && checksum) {
        // two runs of n_ops adds of 1.0 to width doubles each: the total
        // is exact in FP64 for any order (integers below 2^53)
        double s = 0.0;
        double *h = nullptr;
        const size_t chunk = (size_t)1 << 24;
        e = hipHostMalloc(&h, chunk * sizeof(double), hipHostMallocDefault);
        for (size_t o = 0; e == hipSuccess && o < n; o += chunk) {
            const size_t m = n - o < chunk ? n - o : chunk;
            e = hipMemcpy(h, a + o, m * sizeof(double), hipMemcpyDeviceToHost);
            for (size_t k = 0; e == hipSuccess && k < m; ++k) s += h[k];
        }
        if (h) (void)hipHostFree(h);
        *checksum = s;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (e != hipSuccess) {
        mc_set_error(std::string("scattered-RMW probe: ") + hipGetErrorString(e));
        return -1;
    }
    *ms = (double)t;
    return 0;
}

extern "C" int mc_probe_scattered_rmw(int64_t rows, int width, int64_t n_ops, double *ms,
                                      double *checksum)
{
    if (rows <= 0 || width < 1 || width > 16 || n_ops <= 0 || !ms) {
        mc_set_error("mc_probe_scattered_rmw: rows > 0, 1 <= width <= 16, n_ops > 0");
        return -1;
    }
    return probe_run(rows, width, n_ops, false, ms, checksum);
}

extern "C" int mc_probe_row_atomics(int64_t rows, int width, int64_t n_rows, double *ms,
                                    double *checksum)
{
    if (rows <= 0 || (width != 1 && width != 2 && width != 4 && width != 8) || n_rows <= 0 ||
        n_rows > ((int64_t)1 << 40) || !ms) {
        mc_set_error("mc_probe_row_atomics: rows > 0, width 1, 2, 4 or 8, 0 < n_rows <= 2^40");
        return -1;
    }
    return probe_run(rows, width, n_rows, true, ms, checksum);
}
