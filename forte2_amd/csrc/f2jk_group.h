// Grouped (multi-job) J/K launches for the serving layer.
//
// One persistent launch per phase covers the blocks of every job in the
// group.  A block walks the concatenated block index space with a stride of
// gridDim.x, finds its job from the per-phase block counts (a short uniform
// scan, at most kMaxGroup jobs) and runs the same device body as the
// single-job kernels with the job's own plan.  The decomposition of a job
// (element groups, P chunks, K splits) is its own plan, independent of the
// other jobs; the tile width of the BF16x3 half-transform is the group's
// (the widest job's), which changes which block computes an element of X but
// not the order of its k-sum.  So a job's J/K are bitwise the same alone and
// in any mix.
//
// Precisions: FP64 (J and K from the FP64 B), FP32 and BF16x3 (J from the
// FP32 B with FP64 accumulation, K on MFMA from the FP32 B), one precision
// per group.  A job computes J and K, or K only.
//
// The job table travels by value in the kernel arguments (kMaxGroup jobs fit
// the 4 KiB argument limit), so a launch needs no descriptor upload.
// Restrictions of the grouped path: nbf even, one X chunk per job.
#pragma once
#include <type_traits>

#include "f2jk.h"
#include "f2jk_kernels.h"

namespace f2jk {

constexpr int kMaxGroup = 12;

enum GroupPhase {
    GP_PAD = 0,      // Cp[n][ioff+i] = C[n][i] (FP64; also the FP32 copy, or the
                     // transposed hi/lo bf16 copies, of the precision forms)
    GP_DKS = 1,      // Dpart[d] = Cp_d Cp_d^T (lower tiles)
    GP_DRED = 2,     // D[d] mirrored from Dpart[d]
    GP_GAMMA = 3,    // J pass 1
    GP_GRED = 4,     // gamma reduction
    GP_JACC = 5,     // J pass 2
    GP_JRED = 6,     // J reduction + mirror
    GP_KX = 7,       // half-transform (whole B: one chunk)
    GP_KS = 8,       // contraction, splits of P
    GP_KRED = 9,     // K reduction + mirror
    GP_NPHASE = 10
};

struct GJob {
    const void* B;        // FP64 B (FP64 groups) or the FP32 copy (FP32 / BF16x3 groups)
    const int* rowoff;
    const double* C[2];
    double *Cp, *D, *Dpart, *gpart, *gamma, *jpart, *Kpart;
    float* Cp32;          // FP32 coefficients (FP32 K)
    u16 *Cth, *Ctl;       // transposed hi/lo coefficients (BF16x3 K)
    void* X;              // half-transformed chunk: FP64, FP32, or hi then lo bf16
    double* J[2];
    double* K[2];
    long long ldc[2];
    int N, Q, nd, nocc[2], width[2], ioff[2], OPAD, KP;
    int nvec, neg, npc1, npc2, ppw2;
    int nlower, nsplit, ncolt;
    int want_j;
};

struct GArgs {
    int njobs;
    GJob job[kMaxGroup];
};
static_assert(sizeof(GArgs) <= 4096, "the job table must fit the kernel argument limit");

__host__ __device__ inline int gcdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// blocks of one job in one phase (256 threads per block everywhere)
template <int PREC>
__host__ __device__ inline int phase_blocks(const GJob& g, int phase) {
    const int j = g.want_j ? 1 : 0;
    switch (phase) {
        case GP_PAD: return gcdiv((long long)(PREC == F2JK_PREC_BF16X3 ? g.KP : g.N) * g.OPAD, 256);
        case GP_DKS: return j * g.nlower * g.nd;
        case GP_DRED: return j * g.N * gcdiv(g.N, 256) * g.nd;
        case GP_GAMMA: return j * g.neg * g.npc1;
        case GP_GRED: return j * gcdiv((long long)g.nd * g.Q, 256);
        case GP_JACC: return j * g.neg * g.npc2;
        case GP_JRED: return j * gcdiv(g.nvec, 256);
        case GP_KX: return gcdiv((long long)g.Q * g.N, kTile) * g.ncolt;
        case GP_KS: return g.nlower * g.nsplit * g.nd;
        case GP_KRED: return g.N * gcdiv(g.N, 256) * g.nd;
    }
    return 0;
}

// mirrored reduction of split partials: out[m][n] = out[n][m] = sum_s part[s][m][n]
__device__ __forceinline__ void mirror_reduce(const double* part, int nsplit, int N, double* out, int m, int n) {
    if (n > m || m >= N) return;
    const long long NN = (long long)N * N;
    double acc = 0.0;
    for (int s = 0; s < nsplit; ++s) acc += part[s * NN + (long long)m * N + n];
    out[(long long)m * N + n] = acc;
    out[(long long)n * N + m] = acc;
}

// PREC: the group's precision; NBW, VA: tile width and load width of the
// BF16x3 half-transform (unused otherwise).
template <int PHASE, int PREC, int NBW = 2, int VA = 2>
__global__ __launch_bounds__(256, PREC != F2JK_PREC_BF16X3 ? (PHASE == GP_KX || PHASE == GP_KS ? 2 : 1)
                                  : PHASE == GP_KX ? kx_bf16_waves_per_eu<NBW>()
                                  : PHASE == GP_KS ? 2 : 1)
void grouped_kernel(const GArgs a) {
    using TB = typename std::conditional<PREC == F2JK_PREC_FP64, double, float>::type;
    int total = 0;
    for (int j = 0; j < a.njobs; ++j) total += phase_blocks<PREC>(a.job[j], PHASE);
    for (int lb = blockIdx.x; lb < total; lb += gridDim.x) {
        int j = 0, base = 0;
        for (;;) {
            const int nb = phase_blocks<PREC>(a.job[j], PHASE);
            if (lb < base + nb) break;
            base += nb;
            ++j;
        }
        const GJob& g = a.job[j];
        const int local = lb - base;
        const int tid = threadIdx.x;
        if constexpr (PHASE == GP_PAD) {
            const long long t = (long long)local * 256 + tid;
            if constexpr (PREC == F2JK_PREC_BF16X3) {
                // one thread per (row n, column c) of Cp and per (c, k) of the
                // transposed hi/lo copy, k < KP (zero past N)
                if (t < (long long)g.KP * g.OPAD) {
                    const int c = (int)(t / g.KP), k = (int)(t % g.KP);
                    const int d = (g.nd == 2 && c >= g.ioff[1]) ? 1 : 0;
                    const int i = c - g.ioff[d];
                    const double x = (i < g.nocc[d] && k < g.N) ? g.C[d][(long long)k * g.ldc[d] + i] : 0.0;
                    __bf16 h, l;
                    split_bf16_from_double(x, h, l);
                    g.Cth[t] = __builtin_bit_cast(u16, h);
                    g.Ctl[t] = __builtin_bit_cast(u16, l);
                    if (k < g.N) g.Cp[(long long)k * g.OPAD + c] = x;
                }
            } else if (t < (long long)g.N * g.OPAD) {
                const int n = (int)(t / g.OPAD), c = (int)(t % g.OPAD);
                const int d = (g.nd == 2 && c >= g.ioff[1]) ? 1 : 0;
                const int i = c - g.ioff[d];
                const double x = (i < g.nocc[d]) ? g.C[d][(long long)n * g.ldc[d] + i] : 0.0;
                g.Cp[t] = x;
                if constexpr (PREC == F2JK_PREC_FP32) g.Cp32[t] = (float)x;
            }
        } else if constexpr (PHASE == GP_DKS) {
            const int d = local / g.nlower, b = local % g.nlower;
            ks_kernel_body<double>(g.Cp, g.N, g.OPAD, g.ioff[d], g.width[d], 1, 1, g.nlower,
                                   g.Dpart + (long long)d * g.N * g.N, 0, b, g.nlower);
        } else if constexpr (PHASE == GP_DRED) {
            const int per = g.N * gcdiv(g.N, 256);
            const int d = local / per, r = local % per;
            const int m = r / gcdiv(g.N, 256), n = (r % gcdiv(g.N, 256)) * 256 + tid;
            const long long NN = (long long)g.N * g.N;
            mirror_reduce(g.Dpart + d * NN, 1, g.N, g.D + d * NN, m, n);
        } else if constexpr (PHASE == GP_GAMMA) {
            const int bx = local % g.neg, by = local / g.neg;
            const long long NN = (long long)g.N * g.N;
            const TB* B = (const TB*)g.B;
            if (g.nd == 2) {
                DPtrs<2> dp{{g.D, g.D + NN}};
                j_gamma_kernel_body<TB, 2, 4, 2>(B, g.N, g.Q, dp, g.N, g.rowoff, g.nvec, 32, g.gpart, bx, by);
            } else {
                DPtrs<1> dp{{g.D}};
                j_gamma_kernel_body<TB, 2, 4, 1>(B, g.N, g.Q, dp, g.N, g.rowoff, g.nvec, 32, g.gpart, bx, by);
            }
        } else if constexpr (PHASE == GP_GRED) {
            const long long t = (long long)local * 256 + tid;
            if (t < (long long)g.nd * g.Q) {
                const int dd = (int)(t / g.Q), P = (int)(t % g.Q);
                double s = 0.0;
                for (int q = 0; q < g.neg; ++q) s += g.gpart[((long long)q * g.nd + dd) * g.Q + P];
                g.gamma[t] = s;
            }
        } else if constexpr (PHASE == GP_JACC) {
            const int bx = local % g.neg, by = local / g.neg;
            const TB* B = (const TB*)g.B;
            if (g.nd == 2)
                j_accum_kernel_body<TB, 2, 4, 2>(B, g.N, g.Q, g.gamma, g.rowoff, g.nvec, g.ppw2, g.jpart, bx, by);
            else
                j_accum_kernel_body<TB, 2, 4, 1>(B, g.N, g.Q, g.gamma, g.rowoff, g.nvec, g.ppw2, g.jpart, bx, by);
        } else if constexpr (PHASE == GP_JRED) {
            const int e = local * 256 + tid;
            if (e < g.nvec) {
                const int m = find_row(g.rowoff, g.N, e);
                const int col = (e - g.rowoff[m]) * 2;
                for (int dd = 0; dd < g.nd; ++dd) {
                    double s0 = 0.0, s1 = 0.0;
                    for (int c = 0; c < g.npc2; ++c) {
                        const double* src = g.jpart + (((long long)c * g.nd + dd) * g.nvec + e) * 2;
                        s0 += src[0];
                        s1 += src[1];
                    }
                    double* Jd = g.J[dd];
                    Jd[(long long)m * g.N + col] = s0;
                    Jd[(long long)col * g.N + m] = s0;
                    if (col + 1 <= m) {
                        Jd[(long long)m * g.N + col + 1] = s1;
                        Jd[(long long)(col + 1) * g.N + m] = s1;
                    }
                }
            }
        } else if constexpr (PHASE == GP_KX) {
            const int nb = phase_blocks<PREC>(g, GP_KX);
            const long long R = (long long)g.Q * g.N;
            if constexpr (PREC == F2JK_PREC_FP64) {
                kx_kernel_body<double, 2>((const double*)g.B, R, g.N, g.Cp, g.OPAD, (double*)g.X, g.ncolt, local, nb);
            } else if constexpr (PREC == F2JK_PREC_FP32) {
                kx_kernel_body<float, 2>((const float*)g.B, R, g.N, g.Cp32, g.OPAD, (float*)g.X, g.ncolt, local, nb);
            } else {
                u16* xh = (u16*)g.X;
                kx_bf16x3_kernel_body<NBW, VA>((const float*)g.B, R, g.N, g.KP, g.Cth, g.Ctl, g.OPAD, xh,
                                               xh + R * g.OPAD, g.ncolt, local, nb);
            }
        } else if constexpr (PHASE == GP_KS) {
            const int per = g.nlower * g.nsplit;
            const int d = local / per, b = local % per;
            double* kp = g.Kpart + (long long)d * g.nsplit * g.N * g.N;
            if constexpr (PREC == F2JK_PREC_BF16X3) {
                const u16* xh = (const u16*)g.X;
                ks_bf16x3_kernel_body<kTile, 2, 2>(xh, xh + (long long)g.Q * g.N * g.OPAD, g.N, g.OPAD, g.ioff[d], g.width[d],
                                      g.Q, g.nsplit, g.nlower, kp, 0, b, per);
            } else {
                ks_kernel_body<TB>((const TB*)g.X, g.N, g.OPAD, g.ioff[d], g.width[d], g.Q, g.nsplit, g.nlower,
                                   kp, 0, b, per);
            }
        } else if constexpr (PHASE == GP_KRED) {
            const int per = g.N * gcdiv(g.N, 256);
            const int d = local / per, r = local % per;
            const int m = r / gcdiv(g.N, 256), n = (r % gcdiv(g.N, 256)) * 256 + tid;
            mirror_reduce(g.Kpart + (long long)d * g.nsplit * g.N * g.N, g.nsplit, g.N, g.K[d], m, n);
        }
    }
}

}  // namespace f2jk
