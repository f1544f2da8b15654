/* libf2probe.so: peak probes of the device, for the measurement of the J/K
 * kernels against achievable ceilings.  No counterpart in the reference (its
 * J/K path runs on host BLAS); these entry points are measurement only and no
 * J/K entry point depends on them.
 *
 * All entry points return 0 on success and a negative status otherwise
 * (-1 invalid argument, -2 HIP error); f2probe_last_error() describes the last
 * failure of the calling thread.  Pointers are device pointers on the current
 * HIP device; `stream` is a hipStream_t (NULL: the default stream).  Each call
 * runs one untimed launch, then times the others with HIP events recorded on
 * `stream`, and returns when they have finished. */
#ifndef F2PROBE_H
#define F2PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum f2probe_kind { F2PROBE_BF16 = 0, F2PROBE_F32 = 1, F2PROBE_F64 = 2 };

/* Floats of device memory that `sink` must hold for either probe (the probes
 * write to it only when a result equals a sentinel no input reaches). */
int f2probe_sink_len(int64_t* len);

/* Independent MFMA chains from registers on every SIMD, eight waves per SIMD:
 * v_mfma_f32_16x16x32_bf16 (F2PROBE_BF16, the instruction of the BF16x3
 * kernels), v_mfma_f32_16x16x4_f32 or v_mfma_f64_16x16x4_f64, `iters` rounds.
 * *flops = flops of one launch, *ms = its mean duration. */
int f2probe_mfma(int kind, int iters, float* sink, int64_t sink_len, void* stream, double* flops, double* ms);

/* One streaming read of `bytes` (16-byte aligned, a multiple of 16) per launch,
 * `reps` timed launches; *ms = mean duration of one launch. */
int f2probe_read(const void* buf, int64_t bytes, int reps, float* sink, int64_t sink_len, void* stream, double* ms);

const char* f2probe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
