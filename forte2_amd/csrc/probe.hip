// Peak probes of the device: the MFMA rate that independent chains of matrix
// instructions reach on every SIMD, and the HBM read bandwidth of one streaming
// pass over a buffer.  They give the achievable ceilings beside the published
// ones (HBM 8 TB/s; dense MFMA bf16 2.5 PFLOP/s, fp32 157.3 TFLOP/s, fp64 78.6
// TFLOP/s), measured under the same clock and power limits as the J/K kernels.
// Times come from HIP events recorded on the caller's stream.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "f2probe.h"

namespace {

typedef float fx4 __attribute__((ext_vector_type(4)));
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

thread_local char g_err[256] = "";

int fail(int rc, const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return rc;
}

constexpr int kThreads = 256;  // four waves per workgroup
constexpr int kChains = 4;     // independent accumulators per wave (c0..c3)

// Every wave runs kChains independent MFMA chains for iters rounds; operands
// stay in registers, so the rate is the matrix pipes' own.  The result feeds a
// store that the host never asks for (sentinel), which keeps every MFMA live.
__global__ __launch_bounds__(kThreads) void mfma_bf16_probe(int iters, float sentinel, float* __restrict__ sink) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)(1e-3f * (float)((threadIdx.x + i) & 7));
        b[i] = (__bf16)(1e-3f * (float)((threadIdx.x * 3 + i) & 7));
    }
    fx4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    int it = iters;  // >= 1
    do {
        c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c3, 0, 0, 0);
    } while (--it > 0);
    const fx4 t = c0 + c1 + c2 + c3;
    const float s = t[0] + t[1] + t[2] + t[3];
    if (s == sentinel) sink[(int64_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void mfma_f32_probe(int iters, float sentinel, float* __restrict__ sink) {
    const float a = 1e-3f * (float)(threadIdx.x & 7), b = 1e-3f * (float)((threadIdx.x >> 3) & 7);
    fx4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    int it = iters;  // >= 1
    do {
        c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c3, 0, 0, 0);
    } while (--it > 0);
    const fx4 t = c0 + c1 + c2 + c3;
    const float s = t[0] + t[1] + t[2] + t[3];
    if (s == sentinel) sink[(int64_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void mfma_f64_probe(int iters, float sentinel, float* __restrict__ sink) {
    const double a = 1e-3 * (double)(threadIdx.x & 7), b = 1e-3 * (double)((threadIdx.x >> 3) & 7);
    dx4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    int it = iters;  // >= 1
    do {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
    } while (--it > 0);
    const dx4 t = c0 + c1 + c2 + c3;
    const double s = t[0] + t[1] + t[2] + t[3];
    if ((float)s == sentinel) sink[(int64_t)blockIdx.x * kThreads + threadIdx.x] = (float)s;
}

// One pass over n4 16-byte words: eight nontemporal loads in flight per thread
// per round (the J kernels stream B with nontemporal loads as well).
__global__ __launch_bounds__(kThreads) void read_probe(const fx4* __restrict__ p, int64_t n4, float sentinel,
                                                       float* __restrict__ sink) {
    constexpr int kDepth = 8;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    fx4 acc[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) acc[u] = fx4{0, 0, 0, 0};
    for (; i + (kDepth - 1) * stride < n4; i += kDepth * stride) {
        fx4 v[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; ++u) v[u] = __builtin_nontemporal_load(p + i + u * stride);
#pragma unroll
        for (int u = 0; u < kDepth; ++u) acc[u] += v[u];
    }
    for (; i < n4; i += stride) acc[0] += __builtin_nontemporal_load(p + i);
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < kDepth; ++u) s += (acc[u][0] + acc[u][1]) + (acc[u][2] + acc[u][3]);
    if (s == sentinel) sink[(int64_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

int cu_count(int* n) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(-2, "hipGetDevice failed");
    if (hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || *n <= 0)
        return fail(-2, "cannot read the compute-unit count");
    return 0;
}

// Runs launch() once untimed, then reps times between two events; *ms is the
// mean per launch.
template <class F>
int timed(hipStream_t st, int reps, F launch, double* ms) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return fail(-2, "hipEventCreate failed");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(-2, "hipEventCreate failed");
    }
    launch();
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, st);
    int rc = 0;
    if (hipGetLastError() != hipSuccess) rc = fail(-2, "probe launch failed");
    if (rc == 0 && hipEventSynchronize(e1) != hipSuccess) rc = fail(-2, "probe did not complete");
    float t = 0.f;
    if (rc == 0 && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(-2, "hipEventElapsedTime failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == 0) *ms = (double)t / reps;
    return rc;
}

}  // namespace

extern "C" {

int f2probe_sink_len(int64_t* len) {
    if (!len) return fail(-1, "null output");
    int cus = 0;
    if (int rc = cu_count(&cus)) return rc;
    *len = (int64_t)cus * 16 * kThreads;
    return 0;
}

int f2probe_mfma(int kind, int iters, float* sink, int64_t sink_len, void* stream, double* flops, double* ms) {
    if (!sink || !flops || !ms) return fail(-1, "null argument");
    if (iters <= 0) return fail(-1, "iters must be positive");
    int cus = 0;
    if (int rc = cu_count(&cus)) return rc;
    const int blocks = cus * 8;  // eight waves per SIMD
    if (sink_len < (int64_t)blocks * kThreads) return fail(-1, "sink shorter than f2probe_sink_len");
    hipStream_t st = (hipStream_t)stream;
    double per;  // flops of one MFMA instruction of one wave
    int rc;
    if (kind == F2PROBE_BF16) {
        per = 2.0 * 16 * 16 * 32;
        rc = timed(st, 3, [&] { hipLaunchKernelGGL(mfma_bf16_probe, dim3(blocks), dim3(kThreads), 0, st, iters, -1.f, sink); }, ms);
    } else if (kind == F2PROBE_F32) {
        per = 2.0 * 16 * 16 * 4;
        rc = timed(st, 3, [&] { hipLaunchKernelGGL(mfma_f32_probe, dim3(blocks), dim3(kThreads), 0, st, iters, -1.f, sink); }, ms);
    } else if (kind == F2PROBE_F64) {
        per = 2.0 * 16 * 16 * 4;
        rc = timed(st, 3, [&] { hipLaunchKernelGGL(mfma_f64_probe, dim3(blocks), dim3(kThreads), 0, st, iters, -1.f, sink); }, ms);
    } else {
        return fail(-1, "unknown probe kind");
    }
    if (rc) return rc;
    *flops = per * kChains * (double)iters * blocks * (kThreads / 64);
    return 0;
}

int f2probe_read(const void* buf, int64_t bytes, int reps, float* sink, int64_t sink_len, void* stream, double* ms) {
    if (!buf || !sink || !ms) return fail(-1, "null argument");
    if (bytes <= 0 || bytes % 16 != 0 || ((uintptr_t)buf & 15) != 0)
        return fail(-1, "buffer must be 16-byte aligned and a positive multiple of 16 bytes");
    if (reps <= 0) return fail(-1, "reps must be positive");
    int cus = 0;
    if (int rc = cu_count(&cus)) return rc;
    const int blocks = cus * 16;
    if (sink_len < (int64_t)blocks * kThreads) return fail(-1, "sink shorter than f2probe_sink_len");
    hipStream_t st = (hipStream_t)stream;
    const fx4* p = (const fx4*)buf;
    const int64_t n4 = bytes / 16;
    return timed(st, reps, [&] { hipLaunchKernelGGL(read_probe, dim3(blocks), dim3(kThreads), 0, st, p, n4, -1.f, sink); }, ms);
}

const char* f2probe_last_error(void) { return g_err; }

}  // extern "C"
