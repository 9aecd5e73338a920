// Streaming probe for the benchmark's bandwidth ceilings (not part of liblmgx): R input arrays
// read and W output arrays written per element, 16-byte accesses, grid-stride loops over n
// elements, best of `reps` timed launches.  The mixes of interest are those of the hot kernels:
// a read stream carrying a small share of writes runs well below a pure read stream or a copy
// on HBM3E, so each kernel is compared with the ceiling of its own read/write mix.  NT = true
// issues every load and store with the non-temporal hint, as the line kernels do for the
// once-read streams of a level larger than the Infinity Cache.
#include <hip/hip_runtime.h>

typedef double dv2 __attribute__((ext_vector_type(2)));

template <int R, int W, bool NT>
__global__ __launch_bounds__(256) void k_stream_mix(const dv2 *const *in, dv2 *const *out, long n,
                                                    double *sink)
{
    dv2 acc = (dv2)(0.0);
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        dv2 s = (dv2)(0.0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (NT) s += __builtin_nontemporal_load(in[r] + i);
            else s += in[r][i];
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const dv2 v = s + (dv2)((double)w);
            if constexpr (NT) __builtin_nontemporal_store(v, out[w] + i);
            else out[w][i] = v;
        }
        acc += s;
    }
    if (acc.x == 12345.678) sink[0] = acc.y;     // keeps the reads of a write-free mix
}

template <int R, int W, bool NT>
static hipError_t time_mix(const dv2 *const *din, dv2 *const *dout, long n, int reps,
                           double *sink, double *gbps)
{
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) return e;
    const int grid = 256 * 32;
    hipLaunchKernelGGL((k_stream_mix<R, W, NT>), dim3(grid), dim3(256), 0, 0, din, dout, n, sink);
    float best = 1e30f;
    for (int q = 0; q < reps && e == hipSuccess; ++q) {
        hipEventRecord(a, 0);
        hipLaunchKernelGGL((k_stream_mix<R, W, NT>), dim3(grid), dim3(256), 0, 0, din, dout, n,
                           sink);
        hipEventRecord(b, 0);
        if ((e = hipEventSynchronize(b)) != hipSuccess) break;
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        best = ms < best ? ms : best;
    }
    if (e == hipSuccess) e = hipGetLastError();
    hipEventDestroy(a);
    hipEventDestroy(b);
    *gbps = (double)(R + W) * (double)n * 16.0 / (best * 1e-3) / 1e9;
    return e;
}

// GB/s of the mix (reads, writes) over arrays of n 16-byte elements, every access with the
// non-temporal hint when `nontemporal` is nonzero; 0 on success, -1 for a mix without an
// instance, else the HIP error code.
extern "C" int lmgx_stream_probe_hint(int reads, int writes, int nontemporal, long n, int reps,
                                      double *gbps)
{
    if (!gbps || n <= 0 || reps <= 0) return -1;
    const int na = reads + writes;
    dv2 *arr[16] = {nullptr};
    const dv2 **din = nullptr;
    dv2 **dout = nullptr;
    double *sink = nullptr;
    hipError_t e = hipSuccess;
    int rc = 0;
    for (int q = 0; q < na && e == hipSuccess; ++q)
        if ((e = hipMalloc(&arr[q], n * sizeof(dv2))) == hipSuccess)
            e = hipMemset(arr[q], 0, n * sizeof(dv2));
    if (e == hipSuccess) e = hipMalloc(&din, 16 * sizeof(void *));
    if (e == hipSuccess) e = hipMalloc(&dout, 16 * sizeof(void *));
    if (e == hipSuccess) e = hipMalloc(&sink, sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(din, arr, reads * sizeof(void *), hipMemcpyHostToDevice);
    if (e == hipSuccess && writes)
        e = hipMemcpy(dout, arr + reads, writes * sizeof(void *), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
#define MIX(R, W, NT)                                                                      \
    if (reads == R && writes == W && (nontemporal != 0) == NT)                                \
        e = time_mix<R, W, NT>(din, dout, n, reps, sink, gbps);                               \
    else
        MIX(10, 0, false) MIX(10, 1, false) MIX(9, 1, false) MIX(6, 1, false) MIX(5, 1, false)
        MIX(5, 0, false) MIX(1, 1, false) MIX(10, 1, true) MIX(9, 1, true) MIX(6, 1, true)
        MIX(1, 1, true) rc = -1;
#undef MIX
    }
    for (int q = 0; q < na; ++q) hipFree(arr[q]);
    hipFree(din);
    hipFree(dout);
    hipFree(sink);
    return rc ? rc : (int)e;
}

// The same with ordinary loads and stores.
extern "C" int lmgx_stream_probe(int reads, int writes, long n, int reps, double *gbps)
{
    return lmgx_stream_probe_hint(reads, writes, 0, n, reps, gbps);
}
