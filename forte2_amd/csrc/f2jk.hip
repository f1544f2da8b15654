// C ABI of the MI355X density-fitted J/K builder.  Declarations and the
// reference functions each entry point replaces are in include/f2jk.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "f2jk.h"
#include "f2jk_kernels.h"
#include "f2jk_group.h"

using namespace f2jk;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define F2JK_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(F2JK_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define F2JK_LAUNCHED(what)                                                             \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return fail(F2JK_EHIP, std::string("launch ") + what + ": " +               \
                                       hipGetErrorString(e_));                         \
    } while (0)

constexpr int kMaxDens = 8;
constexpr int kMaxEvents = 512;
constexpr int kEPT = 4;                 // vector elements per thread in the J kernels
constexpr int kJ1PerBlock = 32;         // P values per block in J pass 1
constexpr int kTargetBlocks = 2048;     // blocks wanted per launch to fill 256 CUs
constexpr size_t kDefaultXChunkBytes = size_t(4) << 30;

size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }
int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

struct f2jk_ctx {
    int device = 0;
    const void* B[2] = {nullptr, nullptr};
    int64_t naux = 0, nbf = 0;
    int* rowoff[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // by VEC (1, 2, 4)
    int nvec[5] = {0, 0, 0, 0, 0};
    // J over the nonzero vector elements only ("j_sparse"): per B copy (dtype)
    // the list of lower-triangle vector elements (of that copy's J width)
    // where B is nonzero for some P
    bool j_sparse = false;
    uint32_t* jlist[2] = {nullptr, nullptr};
    int jlist_vec[2] = {0, 0};
    int jlist_n[2] = {0, 0};
    size_t xchunk_bytes = kDefaultXChunkBytes;
    bool j_nontemporal = true;
    int bf16_nbw_max = 10;          // widest BF16x3 half-transform tile: 32 * 10 columns
    // J (HBM-bound) runs on a side stream concurrently with K (MFMA-bound)
    bool overlap_jk = false;
    int kx_waves = 8;               // waves per BF16x3 half-transform workgroup (4 or 8)
    int ks_tile = 0;                // BF16x3 contraction tile edge: 128, 256, 0 = by nbf
    int kx_rows = 0;                // BF16x3 half-transform tile rows: 128, 256, 0 = by nbf
    int ks_dma = -1;                // BF16x3 contraction tiles by direct-to-LDS loads: 0, 1, -1 = with 256 tiles
    int kx_wide = -1;               // FP64/FP32 half-transform in 320-column tiles: 0, 1, -1 = above 128 columns
    hipStream_t side = nullptr;     // default priority, created with the context
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int target_blocks = kTargetBlocks;
    // timing
    hipEvent_t ev[kMaxEvents];      // (start, end) pairs of non-nested phases
    int ev_phase[kMaxEvents / 2];
    int nev = 0;
    hipEvent_t ev_total[2];         // the whole build
    bool total_recorded = false;
    bool timing = false;
};

namespace {

// ---------------------------------------------------------------------------
// Launch plan and workspace layout of one build.
// ---------------------------------------------------------------------------
struct Plan {
    int N = 0, Q = 0, nd = 0;
    bool want_j = false, want_k = false, need_c = false;
    int prec = 0;
    size_t tk = 8;             // bytes of a K-path element (8 fp64, 4 fp32)
    // C packing (K and the D of build_jk)
    int OPAD = 0;
    int width[kMaxDens] = {}, ioff[kMaxDens] = {};
    // J
    int jvec = 1, nvec = 0, neg = 0, npc1 = 0, ppw2 = 0, npc2 = 0;
    const uint32_t* jlist = nullptr;  // element list of the J passes (null: every element)
    size_t jb = 8;             // bytes of a B element read by J
    // K
    int Pc = 0, nchunks = 0, ntile = 0, nlower = 0, nsplit = 0, ncolt = 0, kva = 1, nbw = 0, kx_waves = 4;
    int ks_tile = 128;  // edge of the contraction's lower tiles (256: BF16x3 8-wave tiles)
    int kx_rows = 128;  // rows of a BF16x3 half-transform tile (256: the tall kernel)
    bool kx_wide = false;  // FP64/FP32 half-transform in 320-column tiles (kx_wide_kernel)
    int ctrows = 0;     // rows of the padded bf16 coefficient planes (>= ncolt * tile columns)
    // workspace offsets
    int KP = 0;                // nbf rounded up to the bf16 k step
    size_t off_cp32 = 0, off_cth = 0, off_ctl = 0;
    size_t off_cp = 0, off_d = 0, off_dpart = 0, off_gpart = 0, off_gamma = 0, off_jpart = 0,
           off_x = 0, off_kpart = 0, total = 0;
    // sizes of the regions that scale with the launch-shape knobs
    size_t jpart_bytes = 0, x_bytes = 0, kpart_bytes = 0;
    // the workspace is the library's scratch array: the build may not allocate
    bool no_alloc = false;
};

int j_vec_for(int dtype, int64_t N) {
    if (dtype == F2JK_F64) return (N % 2 == 0) ? 2 : 1;
    return (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 1);
}

int ensure_rowoff(f2jk_ctx* ctx, int vec) {
    if (ctx->rowoff[vec]) return F2JK_OK;
    const int64_t N = ctx->nbf;
    std::vector<int> h(N + 1);
    int64_t acc = 0;
    for (int64_t m = 0; m < N; ++m) {
        h[m] = (int)acc;
        acc += (m + 1 + vec - 1) / vec;
    }
    h[N] = (int)acc;
    if (acc > (int64_t)INT32_MAX) return fail(F2JK_EUNSUPPORTED, "nbf too large for J element tables");
    F2JK_HIP(hipSetDevice(ctx->device));
    F2JK_HIP(hipMalloc(&ctx->rowoff[vec], sizeof(int) * (N + 1)));
    F2JK_HIP(hipMemcpy(ctx->rowoff[vec], h.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    ctx->nvec[vec] = (int)acc;
    return F2JK_OK;
}

// The J element list of B copy `dtype` (see f2jk_ctx::jlist): the device
// flags the lower-triangle vector elements where B is nonzero for some P (one
// pass over the lower triangle), the host compacts them.  Set-up work: it
// synchronises the device.
template <typename TB, int VEC>
int scan_jlist(f2jk_ctx* ctx, const TB* B, unsigned char* flag) {
    const int nvec = ctx->nvec[VEC];
    hipLaunchKernelGGL((j_scan_kernel<TB, VEC>), dim3((unsigned)cdiv(nvec, 256)), dim3(256), 0, nullptr, B,
                       (int)ctx->nbf, (int)ctx->naux, ctx->rowoff[VEC], nvec, flag);
    F2JK_LAUNCHED("j_scan");
    return F2JK_OK;
}

void free_jlist(f2jk_ctx* ctx, int dtype) {
    if (ctx->jlist[dtype]) (void)hipFree(ctx->jlist[dtype]);
    ctx->jlist[dtype] = nullptr;
    ctx->jlist_vec[dtype] = ctx->jlist_n[dtype] = 0;
}

int build_jlist(f2jk_ctx* ctx, int dtype) {
    free_jlist(ctx, dtype);
    if (!ctx->B[dtype]) return F2JK_OK;
    const int vec = j_vec_for(dtype, ctx->nbf);
    int rc = ensure_rowoff(ctx, vec);
    if (rc) return rc;
    const int N = (int)ctx->nbf, nvec = ctx->nvec[vec];
    F2JK_HIP(hipSetDevice(ctx->device));
    F2JK_HIP(hipDeviceSynchronize());  // B may have been written on any stream
    unsigned char* flag = nullptr;
    F2JK_HIP(hipMalloc(&flag, nvec));
    if (dtype == F2JK_F64)
        rc = vec == 2 ? scan_jlist<double, 2>(ctx, (const double*)ctx->B[dtype], flag)
                      : scan_jlist<double, 1>(ctx, (const double*)ctx->B[dtype], flag);
    else
        rc = vec == 4 ? scan_jlist<float, 4>(ctx, (const float*)ctx->B[dtype], flag)
           : vec == 2 ? scan_jlist<float, 2>(ctx, (const float*)ctx->B[dtype], flag)
                      : scan_jlist<float, 1>(ctx, (const float*)ctx->B[dtype], flag);
    std::vector<unsigned char> h(nvec);
    hipError_t e = rc ? hipSuccess : hipMemcpy(h.data(), flag, nvec, hipMemcpyDeviceToHost);
    (void)hipFree(flag);
    if (rc) return rc;
    if (e != hipSuccess) return fail(F2JK_EHIP, std::string("j_scan flags: ") + hipGetErrorString(e));
    std::vector<uint32_t> list;
    int el = 0;
    for (int m = 0; m < N; ++m)
        for (int c = 0; c <= m; c += vec, ++el)
            if (h[el]) list.push_back(((uint32_t)m << 16) | (uint32_t)c);
    if (list.empty()) list.push_back(0u);  // an all-zero B: one element keeps the launches non-empty
    F2JK_HIP(hipMalloc(&ctx->jlist[dtype], list.size() * sizeof(uint32_t)));
    F2JK_HIP(hipMemcpy(ctx->jlist[dtype], list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    ctx->jlist_vec[dtype] = vec;
    ctx->jlist_n[dtype] = (int)list.size();
    return F2JK_OK;
}

// Limits of a plan that are otherwise the context's launch-shape knobs; the
// no-alloc entry point lowers them until the plan fits its scratch array.
struct PlanCaps {
    size_t xchunk_bytes;
    int target_blocks;
};

int make_plan(const f2jk_ctx* ctx, int ndens, const int64_t* nocc, int precision, int flags,
              bool j_from_c, Plan& p, int jvec_override = 0, bool grouped = false,
              const PlanCaps* caps = nullptr) {
    const size_t xchunk_bytes = caps ? caps->xchunk_bytes : ctx->xchunk_bytes;
    const int target_blocks = caps ? caps->target_blocks : ctx->target_blocks;
    if (ctx->nbf <= 0 || ctx->naux <= 0) return fail(F2JK_ESTATE, "B has not been set");
    if (ndens < 1 || ndens > kMaxDens)
        return fail(F2JK_EINVAL, "ndens must be in [1, " + std::to_string(kMaxDens) + "]");
    if (precision < F2JK_PREC_FP64 || precision > F2JK_PREC_BF16X3)
        return fail(F2JK_EINVAL, "unknown precision");
    p.N = (int)ctx->nbf;
    p.Q = (int)ctx->naux;
    p.nd = ndens;
    p.prec = precision;
    p.want_j = flags & F2JK_WANT_J;
    p.want_k = flags & F2JK_WANT_K;
    p.need_c = p.want_k || (p.want_j && j_from_c);
    const int jdtype = (precision == F2JK_PREC_FP64) ? F2JK_F64 : F2JK_F32;
    const int kdtype = jdtype;
    if (p.want_j && !ctx->B[jdtype])
        return fail(F2JK_ESTATE, jdtype == F2JK_F64 ? "FP64 B not set" : "FP32 B not set");
    if (p.want_k && !ctx->B[kdtype])
        return fail(F2JK_ESTATE, kdtype == F2JK_F64 ? "FP64 B not set" : "FP32 B not set");
    p.tk = (kdtype == F2JK_F64) ? 8 : 4;
    const int N = p.N, Q = p.Q;

    size_t off = 0;
    if (p.need_c) {
        if (!nocc) return fail(F2JK_EINVAL, "nocc is required");
        int o = 0;
        const int kalign = (precision == F2JK_PREC_BF16X3) ? kBK16 : 16;
        for (int d = 0; d < ndens; ++d) {
            if (nocc[d] < 0 || nocc[d] > N) return fail(F2JK_EINVAL, "nocc out of range");
            p.width[d] = (int)(cdiv(std::max<int64_t>(nocc[d], 1), kalign) * kalign);
            p.ioff[d] = o;
            o += p.width[d];
        }
        p.OPAD = o;
        // the C copy used for D (always fp64) and the K copy (fp64 or fp32)
        p.off_cp = off;
        off += align_up((size_t)N * p.OPAD * 8);
        p.off_cp32 = off;
        off += align_up((size_t)N * p.OPAD * 4);  // fp32 copy (FP32 K)
        p.KP = (int)(cdiv(N, kBK16) * kBK16);
        if (precision == F2JK_PREC_BF16X3) {
            // transposed hi/lo bf16 copies (BF16X3 K), with zero rows up to a
            // whole number of half-transform column tiles (at most 320 wide)
            p.ctrows = p.OPAD + 320;
            p.off_cth = off;
            off += align_up((size_t)p.ctrows * p.KP * 2);
            p.off_ctl = off;
            off += align_up((size_t)p.ctrows * p.KP * 2);
        }
    }
    if (p.want_j) {
        p.jb = (jdtype == F2JK_F64) ? 8 : 4;
        p.jvec = jvec_override ? jvec_override : j_vec_for(jdtype, N);
        int64_t nvec = 0;
        for (int64_t m = 0; m < N; ++m) nvec += (m + 1 + p.jvec - 1) / p.jvec;
        if (!grouped && ctx->jlist[jdtype] && ctx->jlist_vec[jdtype] == p.jvec) {
            p.jlist = ctx->jlist[jdtype];
            nvec = ctx->jlist_n[jdtype];
        }
        p.nvec = (int)nvec;
        p.neg = (int)cdiv(nvec, (int64_t)kJThreads * kEPT);
        p.npc1 = (int)cdiv(Q, kJ1PerBlock);
        p.npc2 = (int)std::max<int64_t>(1, std::min<int64_t>(Q, cdiv(target_blocks, p.neg)));
        p.ppw2 = (int)cdiv(Q, p.npc2);
        p.npc2 = (int)cdiv(Q, p.ppw2);
        if (j_from_c) {
            p.off_d = off;
            off += align_up((size_t)ndens * N * N * 8);
            p.off_dpart = off;
            off += align_up((size_t)N * N * 8);
        }
        p.off_gpart = off;
        off += align_up((size_t)p.neg * ndens * Q * 8);
        p.off_gamma = off;
        off += align_up((size_t)ndens * Q * 8);
        p.off_jpart = off;
        p.jpart_bytes = align_up((size_t)p.npc2 * ndens * (size_t)nvec * p.jvec * 8);
        off += p.jpart_bytes;
    }
    if (p.want_k) {
        const size_t per_p = (size_t)N * p.OPAD * p.tk;
        int64_t pc = (int64_t)(xchunk_bytes / std::max<size_t>(per_p, 1));
        pc = std::max<int64_t>(1, std::min<int64_t>(pc, Q));
        // as many chunks as the cap needs, of balanced size: no short last
        // chunk whose launches would be all set-up and Kpart traffic
        pc = cdiv(Q, cdiv(Q, pc));
        p.Pc = (int)pc;
        p.nchunks = (int)cdiv(Q, pc);
        // BF16x3 single-job builds of large nbf contract in 256-wide tiles
        // (one 8-wave workgroup per CU); everything else in 128-wide tiles
        p.ks_tile = 128;
        if (precision == F2JK_PREC_BF16X3 && !grouped)
            p.ks_tile = ctx->ks_tile ? ctx->ks_tile : (N >= 768 ? 256 : 128);
        p.ntile = (int)cdiv(N, p.ks_tile);
        p.nlower = p.ntile * (p.ntile + 1) / 2;
        // splits of the P range: at most wg_target workgroups per launch
        // (rounded down, so that a launch fills whole rounds of the CUs
        // instead of starting a nearly empty last one)
        const int64_t wg_target = p.ks_tile == 256 ? target_blocks / 4 : target_blocks / 2;
        p.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(pc, wg_target / p.nlower));
        if (precision == F2JK_PREC_BF16X3) {
            // single-job builds of large nbf run the tall (256-row) half-transform
            p.kx_rows = 128;
            if (!grouped) p.kx_rows = ctx->kx_rows ? ctx->kx_rows : (N >= 768 ? 256 : 128);
            if (N % 4 || N < 32) p.kx_rows = 128;  // the tall kernel loads A rows 16 bytes at a time
            // the tile width (in 32-column units) within the cap that computes
            // the fewest column blocks; ties go to the wider tile.  The tall
            // kernel's 2 x 4 waves take even widths only.
            const int need = (int)cdiv(p.OPAD, 32);
            int best = 0, cost = 0;
            for (int w : {10, 8, 6, 5, 4, 2}) {
                if (w > ctx->bf16_nbw_max || (p.kx_rows == 256 && w % 2)) continue;
                const int c = (int)cdiv(need, w) * w;
                if (best == 0 || c < cost) best = w, cost = c;
            }
            if (best == 0) {  // a cap of 5 leaves no even width: the 128-row kernel
                p.kx_rows = 128;
                for (int w : {10, 8, 6, 5, 4, 2}) {
                    if (w > ctx->bf16_nbw_max) continue;
                    const int c = (int)cdiv(need, w) * w;
                    if (best == 0 || c < cost) best = w, cost = c;
                }
            }
            p.nbw = best;
            p.kx_waves = ctx->kx_waves;
            p.ncolt = (int)cdiv(p.OPAD, 32 * p.nbw);
        } else {
            // one 320-column tile reads each row block of B once (the
            // 128-column kernel once per column tile)
            p.kx_wide = !grouped && (ctx->kx_wide < 0 ? p.OPAD > kTile : ctx->kx_wide == 1);
            p.ncolt = (int)cdiv(p.OPAD, p.kx_wide ? kWideCols : kTile);
        }
        if (p.tk == 8)
            p.kva = (N % 2 == 0) ? 2 : 1;
        else
            p.kva = (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 1);
        p.off_x = off;
        p.x_bytes = align_up((size_t)p.Pc * per_p);
        off += p.x_bytes;
        p.off_kpart = off;
        p.kpart_bytes = align_up((size_t)p.nsplit * ndens * N * N * 8);
        off += p.kpart_bytes;
    }
    p.total = std::max<size_t>(off, 256);
    return F2JK_OK;
}

// --------------------------------------------------------------------------- timing
void tim_begin(f2jk_ctx* ctx, bool on) {
    ctx->timing = on;
    ctx->nev = 0;
    ctx->total_recorded = false;
}
int tim_mark(f2jk_ctx* ctx, int phase, hipStream_t st, bool start) {
    if (!ctx->timing) return F2JK_OK;
    if (phase == F2JK_PH_TOTAL) {
        F2JK_HIP(hipEventRecord(ctx->ev_total[start ? 0 : 1], st));
        if (!start) ctx->total_recorded = true;
        return F2JK_OK;
    }
    if (start && ctx->nev + 2 > kMaxEvents) return F2JK_OK;  // table full: phase not timed
    if (!start && (ctx->nev % 2) == 0) return F2JK_OK;       // end of an untimed phase
    if (start) ctx->ev_phase[ctx->nev / 2] = phase;
    F2JK_HIP(hipEventRecord(ctx->ev[ctx->nev], st));
    ctx->nev++;
    return F2JK_OK;
}
#define TIM(phase, start)                                        \
    do {                                                         \
        int rc_ = tim_mark(ctx, phase, st, start);               \
        if (rc_) return rc_;                                     \
    } while (0)

// --------------------------------------------------------------------------- J
template <typename TB, int VEC, int ND, bool NT>
int launch_j(f2jk_ctx* ctx, const Plan& p, const TB* B, const double* const* D, int64_t ldd,
             double* const* J, char* ws, hipStream_t st) {
    DPtrs<ND> dp;
    JPtrs<ND> jp;
    for (int d = 0; d < ND; ++d) {
        dp.p[d] = D[d];
        jp.p[d] = J[d];
    }
    double* gpart = (double*)(ws + p.off_gpart);
    double* gamma = (double*)(ws + p.off_gamma);
    double* jpart = (double*)(ws + p.off_jpart);
    const int* rowoff = ctx->rowoff[VEC];
    const uint32_t* elist = p.jlist;
    TIM(F2JK_PH_JGAMMA, true);
    hipLaunchKernelGGL((j_gamma_kernel<TB, VEC, kEPT, ND, NT>), dim3(p.neg, p.npc1), dim3(kJThreads), 0,
                       st, B, p.N, p.Q, dp, ldd, rowoff, elist, p.nvec, kJ1PerBlock, gpart);
    F2JK_LAUNCHED("j_gamma");
    const int64_t ng = (int64_t)ND * p.Q;
    hipLaunchKernelGGL(j_gamma_reduce_kernel, dim3((unsigned)cdiv(ng, 256)), dim3(256), 0, st,
                       gpart, p.neg, ND, p.Q, gamma);
    F2JK_LAUNCHED("j_gamma_reduce");
    TIM(F2JK_PH_JGAMMA, false);
    TIM(F2JK_PH_JACC, true);
    hipLaunchKernelGGL((j_accum_kernel<TB, VEC, kEPT, ND, NT>), dim3(p.neg, p.npc2), dim3(kJThreads), 0,
                       st, B, p.N, p.Q, gamma, rowoff, elist, p.nvec, p.ppw2, jpart);
    F2JK_LAUNCHED("j_accum");
    if (elist) {
        // elements outside the list are exactly zero (B is zero there for every P)
        const int64_t NN = (int64_t)p.N * p.N;
        for (int d = 0; d < ND; ++d) {
            hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)cdiv(NN, 256)), dim3(256), 0, st, J[d], NN);
            F2JK_LAUNCHED("fill_zero");
        }
    }
    hipLaunchKernelGGL((j_reduce_kernel<VEC, ND>), dim3((unsigned)cdiv(p.nvec, 256)), dim3(256), 0,
                       st, jpart, p.npc2, p.N, rowoff, elist, p.nvec, jp);
    F2JK_LAUNCHED("j_reduce");
    TIM(F2JK_PH_JACC, false);
    return F2JK_OK;
}

template <typename TB, int VEC>
int launch_j_nd(f2jk_ctx* ctx, const Plan& p, const TB* B, int nd, const double* const* D,
                int64_t ldd, double* const* J, char* ws, hipStream_t st) {
    // densities are processed two at a time: one read of B serves both
    int d = 0;
    while (d < nd) {
        int rc;
        const bool nt = ctx->j_nontemporal;
        if (nd - d >= 2) {
            rc = nt ? launch_j<TB, VEC, 2, true>(ctx, p, B, D + d, ldd, J + d, ws, st)
                    : launch_j<TB, VEC, 2, false>(ctx, p, B, D + d, ldd, J + d, ws, st);
            d += 2;
        } else {
            rc = nt ? launch_j<TB, VEC, 1, true>(ctx, p, B, D + d, ldd, J + d, ws, st)
                    : launch_j<TB, VEC, 1, false>(ctx, p, B, D + d, ldd, J + d, ws, st);
            d += 1;
        }
        if (rc) return rc;
    }
    return F2JK_OK;
}

int run_j(f2jk_ctx* ctx, const Plan& p, const double* const* D, int64_t ldd, double* const* J,
          char* ws, hipStream_t st) {
    if (p.no_alloc && !ctx->rowoff[p.jvec])
        return fail(F2JK_ESTATE, "J element table missing (set B of this precision first)");
    int rc = ensure_rowoff(ctx, p.jvec);
    if (rc) return rc;
    if (p.prec == F2JK_PREC_FP64) {
        const double* B = (const double*)ctx->B[F2JK_F64];
        if (p.jvec == 2) return launch_j_nd<double, 2>(ctx, p, B, p.nd, D, ldd, J, ws, st);
        return launch_j_nd<double, 1>(ctx, p, B, p.nd, D, ldd, J, ws, st);
    }
    const float* B = (const float*)ctx->B[F2JK_F32];
    if (p.jvec == 4) return launch_j_nd<float, 4>(ctx, p, B, p.nd, D, ldd, J, ws, st);
    if (p.jvec == 2) return launch_j_nd<float, 2>(ctx, p, B, p.nd, D, ldd, J, ws, st);
    return launch_j_nd<float, 1>(ctx, p, B, p.nd, D, ldd, J, ws, st);
}

// --------------------------------------------------------------------------- C packing, D
int pack_c(f2jk_ctx* ctx, const Plan& p, const double* const* C, const int64_t* nocc,
           const int64_t* ldc, char* ws, hipStream_t st, bool want_f32) {
    double* cp = (double*)(ws + p.off_cp);
    float* cp32 = (float*)(ws + p.off_cp32);
    for (int d = 0; d < p.nd; ++d) {
        if (nocc[d] > 0 && (!C[d] || !ldc || ldc[d] < nocc[d]))
            return fail(F2JK_EINVAL, "C pointer / ldc invalid");
        const int64_t n = (int64_t)p.N * p.width[d];
        hipLaunchKernelGGL(pad_c_kernel<double>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st,
                           C[d], nocc[d] > 0 ? ldc[d] : 1, p.N, (int)nocc[d], p.width[d], p.ioff[d],
                           p.OPAD, cp);
        F2JK_LAUNCHED("pad_c");
        if (want_f32) {
            hipLaunchKernelGGL(pad_c_kernel<float>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st,
                               C[d], nocc[d] > 0 ? ldc[d] : 1, p.N, (int)nocc[d], p.width[d],
                               p.ioff[d], p.OPAD, cp32);
            F2JK_LAUNCHED("pad_c32");
        }
        if (p.want_k && p.prec == F2JK_PREC_BF16X3) {
            const int64_t nt = (int64_t)p.width[d] * p.KP;
            hipLaunchKernelGGL(pad_ct_bf16_kernel, dim3((unsigned)cdiv(nt, 256)), dim3(256), 0, st,
                               C[d], nocc[d] > 0 ? ldc[d] : 1, p.N, (int)nocc[d], p.width[d],
                               p.ioff[d], p.KP, (u16*)(ws + p.off_cth), (u16*)(ws + p.off_ctl));
            F2JK_LAUNCHED("pad_ct_bf16");
        }
    }
    if (p.want_k && p.prec == F2JK_PREC_BF16X3) {
        // zero rows past the occupied columns: the tall half-transform loads
        // whole column tiles without clamping
        const int zr = p.ctrows - p.OPAD;
        const int64_t nt = (int64_t)zr * p.KP;
        hipLaunchKernelGGL(pad_ct_bf16_kernel, dim3((unsigned)cdiv(nt, 256)), dim3(256), 0, st, C[0], 1, p.N, 0,
                           zr, p.OPAD, p.KP, (u16*)(ws + p.off_cth), (u16*)(ws + p.off_ctl));
        F2JK_LAUNCHED("pad_ct_bf16 zero rows");
    }
    return F2JK_OK;
}

// D_d = Cp_d Cp_d^T via the contraction kernel with a single "P".
int build_d(f2jk_ctx* ctx, const Plan& p, char* ws, hipStream_t st) {
    const double* cp = (const double*)(ws + p.off_cp);
    double* dbuf = (double*)(ws + p.off_d);
    double* dpart = (double*)(ws + p.off_dpart);
    const int nt = (int)cdiv(p.N, kTile);
    const int nlower = nt * (nt + 1) / 2;
    for (int d = 0; d < p.nd; ++d) {
        hipLaunchKernelGGL(ks_kernel<double>, dim3(nlower), dim3(kGThreads), 0, st, cp, p.N, p.OPAD,
                           p.ioff[d], p.width[d], 1, 1, nlower, dpart, 0);
        F2JK_LAUNCHED("ks(D)");
        hipLaunchKernelGGL(k_reduce_kernel, dim3((unsigned)cdiv(p.N, 256), p.N), dim3(256), 0, st,
                           dpart, 1, p.N, dbuf + (size_t)d * p.N * p.N);
        F2JK_LAUNCHED("k_reduce(D)");
    }
    return F2JK_OK;
}

// --------------------------------------------------------------------------- K
// X = A Cp over `grid` tiles: 128 x 320 tiles (kx_wide_kernel) or 128 x 128
// (kx_kernel), as the plan's column tiling says
template <typename T>
void launch_kx_typed(const Plan& p, unsigned grid, const T* A, int64_t R, const T* cp, T* X, hipStream_t st) {
    constexpr int V4 = sizeof(T) == 4 ? 4 : 2;
    if (p.kx_wide) {
        if (p.kva == 4)
            hipLaunchKernelGGL((kx_wide_kernel<T, V4>), dim3(grid), dim3(kWideThreads), 0, st, A, R, p.N, cp,
                               p.OPAD, X, p.ncolt);
        else if (p.kva == 2)
            hipLaunchKernelGGL((kx_wide_kernel<T, 2>), dim3(grid), dim3(kWideThreads), 0, st, A, R, p.N, cp,
                               p.OPAD, X, p.ncolt);
        else
            hipLaunchKernelGGL((kx_wide_kernel<T, 1>), dim3(grid), dim3(kWideThreads), 0, st, A, R, p.N, cp,
                               p.OPAD, X, p.ncolt);
    } else if (p.kva == 4) {
        hipLaunchKernelGGL((kx_kernel<T, V4>), dim3(grid), dim3(kGThreads), 0, st, A, R, p.N, cp, p.OPAD, X,
                           p.ncolt);
    } else if (p.kva == 2) {
        hipLaunchKernelGGL((kx_kernel<T, 2>), dim3(grid), dim3(kGThreads), 0, st, A, R, p.N, cp, p.OPAD, X,
                           p.ncolt);
    } else {
        hipLaunchKernelGGL((kx_kernel<T, 1>), dim3(grid), dim3(kGThreads), 0, st, A, R, p.N, cp, p.OPAD, X,
                           p.ncolt);
    }
}

template <typename T>
int run_k_typed(f2jk_ctx* ctx, const Plan& p, double* const* K, char* ws, hipStream_t st) {
    const T* B = (const T*)ctx->B[sizeof(T) == 8 ? F2JK_F64 : F2JK_F32];
    const T* cp = (sizeof(T) == 8) ? (const T*)(ws + p.off_cp) : (const T*)(ws + p.off_cp32);
    T* X = (T*)(ws + p.off_x);
    double* kpart = (double*)(ws + p.off_kpart);
    const int64_t NN = (int64_t)p.N * p.N;
    for (int c = 0; c < p.nchunks; ++c) {
        const int P0 = c * p.Pc;
        const int pc = std::min(p.Pc, p.Q - P0);
        const int64_t R = (int64_t)pc * p.N;
        const int64_t nrowt = cdiv(R, kTile);
        TIM(F2JK_PH_KX, true);
        const T* A = B + (int64_t)P0 * NN;
        launch_kx_typed<T>(p, (unsigned)(nrowt * p.ncolt), A, R, cp, X, st);
        F2JK_LAUNCHED("kx");
        TIM(F2JK_PH_KX, false);
        TIM(F2JK_PH_KS, true);
        const int nsplit = std::min(p.nsplit, pc);
        for (int d = 0; d < p.nd; ++d) {
            hipLaunchKernelGGL(ks_kernel<T>, dim3(p.nlower * nsplit), dim3(kGThreads), 0, st, X, p.N,
                               p.OPAD, p.ioff[d], p.width[d], pc, nsplit, p.nlower,
                               kpart + (size_t)d * p.nsplit * NN, c > 0 ? 1 : 0);
            F2JK_LAUNCHED("ks");
        }
        TIM(F2JK_PH_KS, false);
    }
    TIM(F2JK_PH_KRED, true);
    const int nsplit_used = std::min(p.nsplit, std::min(p.Pc, p.Q));
    for (int d = 0; d < p.nd; ++d) {
        hipLaunchKernelGGL(k_reduce_kernel, dim3((unsigned)cdiv(p.N, 256), p.N), dim3(256), 0, st,
                           kpart + (size_t)d * p.nsplit * NN, nsplit_used, p.N, K[d]);
        F2JK_LAUNCHED("k_reduce");
    }
    TIM(F2JK_PH_KRED, false);
    return F2JK_OK;
}

template <int NBW, int WAVES>
void launch_kx_bf16_w(const Plan& p, unsigned grid, const float* A, int64_t R, const u16* th,
                      const u16* tl, u16* xh, u16* xl, hipStream_t st) {
    if (p.kva == 4)
        hipLaunchKernelGGL((kx_bf16x3_kernel<NBW, 4, WAVES>), dim3(grid), dim3(64 * WAVES), 0, st, A, R, p.N,
                           p.KP, th, tl, p.OPAD, xh, xl, p.ncolt);
    else if (p.kva == 2)
        hipLaunchKernelGGL((kx_bf16x3_kernel<NBW, 2, WAVES>), dim3(grid), dim3(64 * WAVES), 0, st, A, R, p.N,
                           p.KP, th, tl, p.OPAD, xh, xl, p.ncolt);
    else
        hipLaunchKernelGGL((kx_bf16x3_kernel<NBW, 1, WAVES>), dim3(grid), dim3(64 * WAVES), 0, st, A, R, p.N,
                           p.KP, th, tl, p.OPAD, xh, xl, p.ncolt);
}

// 8 waves for tiles of an even number of 32-column units (each wave then
// holds NBW/2 blocks of 16 columns); 4 waves otherwise
template <int NBW>
void launch_kx_bf16(const Plan& p, unsigned grid, const float* A, int64_t R, const u16* th,
                    const u16* tl, u16* xh, u16* xl, hipStream_t st) {
    if constexpr (NBW % 2 == 0 && NBW >= 4) {
        if (p.kx_waves == 8) return launch_kx_bf16_w<NBW, 8>(p, grid, A, R, th, tl, xh, xl, st);
    }
    launch_kx_bf16_w<NBW, 4>(p, grid, A, R, th, tl, xh, xl, st);
}

// the tall (256-row, 2 x 4 waves) half-transform; even widths only
template <int NBW>
void launch_kx_bf16_tall(const Plan& p, unsigned grid, const float* A, int64_t R, const u16* th,
                         const u16* tl, u16* xh, u16* xl, hipStream_t st) {
    hipLaunchKernelGGL((kx_bf16x3_tall_kernel<NBW, 2, 4>), dim3(grid), dim3(512), 0, st, A, R, p.N, p.KP, th, tl,
                       p.OPAD, xh, xl, p.ncolt);
}

int run_k_bf16x3(f2jk_ctx* ctx, const Plan& p, double* const* K, char* ws, hipStream_t st) {
    const float* B = (const float*)ctx->B[F2JK_F32];
    const u16* th = (const u16*)(ws + p.off_cth);
    const u16* tl = (const u16*)(ws + p.off_ctl);
    // X chunk: hi half then lo half
    u16* xh = (u16*)(ws + p.off_x);
    u16* xl = xh + (size_t)p.Pc * p.N * p.OPAD;
    double* kpart = (double*)(ws + p.off_kpart);
    const int64_t NN = (int64_t)p.N * p.N;
    for (int c = 0; c < p.nchunks; ++c) {
        const int P0 = c * p.Pc;
        const int pc = std::min(p.Pc, p.Q - P0);
        const int64_t R = (int64_t)pc * p.N;
        const float* A = B + (int64_t)P0 * NN;
        TIM(F2JK_PH_KX, true);
        if (p.kx_rows == 256) {
            const unsigned grid = (unsigned)(cdiv(R, kTallRows) * p.ncolt);
            switch (p.nbw) {
                case 2: launch_kx_bf16_tall<2>(p, grid, A, R, th, tl, xh, xl, st); break;
                case 4: launch_kx_bf16_tall<4>(p, grid, A, R, th, tl, xh, xl, st); break;
                case 6: launch_kx_bf16_tall<6>(p, grid, A, R, th, tl, xh, xl, st); break;
                case 8: launch_kx_bf16_tall<8>(p, grid, A, R, th, tl, xh, xl, st); break;
                default: launch_kx_bf16_tall<10>(p, grid, A, R, th, tl, xh, xl, st); break;
            }
        } else {
        const unsigned grid = (unsigned)(cdiv(R, kTile) * p.ncolt);
        switch (p.nbw) {
            case 2: launch_kx_bf16<2>(p, grid, A, R, th, tl, xh, xl, st); break;
            case 4: launch_kx_bf16<4>(p, grid, A, R, th, tl, xh, xl, st); break;
            case 5: launch_kx_bf16<5>(p, grid, A, R, th, tl, xh, xl, st); break;
            case 6: launch_kx_bf16<6>(p, grid, A, R, th, tl, xh, xl, st); break;
            case 8: launch_kx_bf16<8>(p, grid, A, R, th, tl, xh, xl, st); break;
            default: launch_kx_bf16<10>(p, grid, A, R, th, tl, xh, xl, st); break;
        }
        }
        F2JK_LAUNCHED("kx_bf16x3");
        TIM(F2JK_PH_KX, false);
        TIM(F2JK_PH_KS, true);
        const int nsplit = std::min(p.nsplit, pc);
        const bool ks_dma = ctx->ks_dma < 0 ? p.ks_tile == 256 : ctx->ks_dma == 1;
        for (int d = 0; d < p.nd; ++d) {
            if (ks_dma && p.ks_tile == 256)
                hipLaunchKernelGGL((ks_bf16x3_dma_kernel<256, 4, 2>), dim3(p.nlower * nsplit), dim3(512), 0, st,
                                   xh, xl, p.N, p.OPAD, p.ioff[d], p.width[d], pc, nsplit, p.nlower,
                                   kpart + (size_t)d * p.nsplit * NN, c > 0 ? 1 : 0);
            else if (ks_dma)
                hipLaunchKernelGGL((ks_bf16x3_dma_kernel<kTile, 2, 2>), dim3(p.nlower * nsplit), dim3(kGThreads), 0,
                                   st, xh, xl, p.N, p.OPAD, p.ioff[d], p.width[d], pc, nsplit, p.nlower,
                                   kpart + (size_t)d * p.nsplit * NN, c > 0 ? 1 : 0);
            else if (p.ks_tile == 256)
                hipLaunchKernelGGL((ks_bf16x3_kernel<256, 4, 2>), dim3(p.nlower * nsplit), dim3(512), 0, st,
                                   xh, xl, p.N, p.OPAD, p.ioff[d], p.width[d], pc, nsplit, p.nlower,
                                   kpart + (size_t)d * p.nsplit * NN, c > 0 ? 1 : 0);
            else
                hipLaunchKernelGGL((ks_bf16x3_kernel<kTile, 2, 2>), dim3(p.nlower * nsplit), dim3(kGThreads), 0,
                                   st, xh, xl, p.N, p.OPAD, p.ioff[d], p.width[d], pc, nsplit, p.nlower,
                                   kpart + (size_t)d * p.nsplit * NN, c > 0 ? 1 : 0);
            F2JK_LAUNCHED("ks_bf16x3");
        }
        TIM(F2JK_PH_KS, false);
    }
    TIM(F2JK_PH_KRED, true);
    const int nsplit_used = std::min(p.nsplit, std::min(p.Pc, p.Q));
    for (int d = 0; d < p.nd; ++d) {
        hipLaunchKernelGGL(k_reduce_kernel, dim3((unsigned)cdiv(p.N, 256), p.N), dim3(256), 0, st,
                           kpart + (size_t)d * p.nsplit * NN, nsplit_used, p.N, K[d]);
        F2JK_LAUNCHED("k_reduce");
    }
    TIM(F2JK_PH_KRED, false);
    return F2JK_OK;
}

int check_ws(const Plan& p, void* ws, size_t ws_bytes) {
    if (!ws) return fail(F2JK_EINVAL, "workspace pointer is NULL");
    if (ws_bytes < p.total)
        return fail(F2JK_ENOMEM, "workspace too small: need " + std::to_string(p.total) + " bytes, got " +
                                     std::to_string(ws_bytes));
    if (((uintptr_t)ws) % 256) return fail(F2JK_EINVAL, "workspace must be 256-byte aligned");
    return F2JK_OK;
}

int build_common(f2jk_ctx* ctx, int ndens, const double* const* C, const int64_t* nocc,
                 const int64_t* ldc, const double* const* Din, int64_t ldd, double* const* J,
                 double* const* K, int precision, int flags, void* ws, size_t ws_bytes, void* stream,
                 bool j_from_c, const PlanCaps* caps = nullptr, bool no_alloc = false) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    Plan p;
    int rc = make_plan(ctx, ndens, nocc, precision, flags, j_from_c, p, 0, false, caps);
    if (rc) return rc;
    p.no_alloc = no_alloc;
    if ((rc = check_ws(p, ws, ws_bytes))) return rc;
    for (int d = 0; d < ndens; ++d) {
        if (p.want_j && !J[d]) return fail(F2JK_EINVAL, "J output pointer is NULL");
        if (p.want_k && !K[d]) return fail(F2JK_EINVAL, "K output pointer is NULL");
        if (p.want_j && !j_from_c && !Din[d]) return fail(F2JK_EINVAL, "D pointer is NULL");
    }
    F2JK_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)ws;
    tim_begin(ctx, flags & F2JK_TIMING);
    TIM(F2JK_PH_TOTAL, true);
    if (p.need_c) {
        TIM(F2JK_PH_PREP, true);
        if ((rc = pack_c(ctx, p, C, nocc, ldc, w, st, p.want_k && p.prec == F2JK_PREC_FP32))) return rc;
        if (p.want_j && j_from_c)
            if ((rc = build_d(ctx, p, w, st))) return rc;
        TIM(F2JK_PH_PREP, false);
    }
    // fork: the J passes go to the side stream once the prepared inputs exist
    const bool overlap = p.want_j && p.want_k && ctx->overlap_jk;
    hipStream_t sj = st;
    if (overlap) {
        F2JK_HIP(hipEventRecord(ctx->ev_fork, st));
        F2JK_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        sj = ctx->side;
    }
    if (p.want_j) {
        std::vector<const double*> dptr(ndens);
        int64_t lddu = ldd;
        for (int d = 0; d < ndens; ++d)
            dptr[d] = j_from_c ? (const double*)(w + p.off_d) + (size_t)d * p.N * p.N : Din[d];
        if (j_from_c) lddu = p.N;
        if ((rc = run_j(ctx, p, dptr.data(), lddu, J, w, sj))) return rc;
    }
    if (p.want_k) {
        if (p.prec == F2JK_PREC_BF16X3)
            rc = run_k_bf16x3(ctx, p, K, w, st);
        else
            rc = (p.tk == 8) ? run_k_typed<double>(ctx, p, K, w, st) : run_k_typed<float>(ctx, p, K, w, st);
        if (rc) return rc;
    }
    // join: later work on the caller's stream sees J
    if (overlap) {
        F2JK_HIP(hipEventRecord(ctx->ev_join, sj));
        F2JK_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
    }
    TIM(F2JK_PH_TOTAL, false);
    g_last_error.clear();
    return F2JK_OK;
}

// --------------------------------------------------------------------------- no-alloc builds
// Scratch of f2jk_build_jk_noalloc: one array per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory; the device address comes from hipGetSymbolAddress, once per
// device (f2jk_ctx_create), and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[F2JK_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// with `device` current
int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(F2JK_EUNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return F2JK_OK;
    void* ptr = nullptr;
    F2JK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr || ((uintptr_t)ptr) % 256) return fail(F2JK_EHIP, "scratch array has no aligned device address");
    g_scratch_ptr[device] = ptr;
    return F2JK_OK;
}

// The widest launch shapes whose build_jk plan fits `budget` bytes: starting
// from the context's knobs, the X chunk or the split count (the larger of the
// two regions first) is halved until the plan fits; ENOMEM when even one P
// per chunk and one split do not.
int fit_plan(const f2jk_ctx* ctx, int ndens, const int64_t* nocc, int precision, int flags, size_t budget,
             PlanCaps& caps) {
    caps = PlanCaps{ctx->xchunk_bytes, ctx->target_blocks};
    for (int it = 0; it < 256; ++it) {
        Plan p;
        int rc = make_plan(ctx, ndens, nocc, precision, flags, true, p, 0, false, &caps);
        if (rc) return rc;
        if (p.total <= budget) return F2JK_OK;
        const size_t per_p = p.want_k ? (size_t)p.N * p.OPAD * p.tk : 0;
        if (p.want_k && p.Pc > 1 && p.x_bytes >= p.jpart_bytes + p.kpart_bytes)
            caps.xchunk_bytes = per_p * (size_t)(p.Pc / 2);
        else if (caps.target_blocks > 1)
            caps.target_blocks /= 2;
        else if (p.want_k && p.Pc > 1)
            caps.xchunk_bytes = per_p * (size_t)(p.Pc / 2);
        else
            return fail(F2JK_ENOMEM, "build needs " + std::to_string(p.total) +
                                         " bytes of scratch at its smallest launch shapes; the no-alloc scratch "
                                         "array holds " + std::to_string(budget) + " (use f2jk_build_jk with a "
                                         "caller workspace)");
    }
    return fail(F2JK_ENOMEM, "no launch shape fits the no-alloc scratch array");
}

// --------------------------------------------------------------------------- grouped builds
int group_plan(f2jk_ctx* const* ctxs, int njobs, const int* ndens, const int64_t* nocc, int precision,
               int flags, std::vector<Plan>& plans) {
    if (njobs < 1) return fail(F2JK_EINVAL, "njobs must be >= 1");
    if (!ctxs || !ndens || !nocc) return fail(F2JK_EINVAL, "null argument");
    if (precision < F2JK_PREC_FP64 || precision > F2JK_PREC_BF16X3) return fail(F2JK_EINVAL, "unknown precision");
    const int jk = flags & (F2JK_WANT_J | F2JK_WANT_K);
    if (jk != (F2JK_WANT_J | F2JK_WANT_K) && jk != F2JK_WANT_K)
        return fail(F2JK_EUNSUPPORTED, "grouped builds compute J and K, or K only");
    plans.resize(njobs);
    int64_t k = 0;
    for (int j = 0; j < njobs; ++j) {
        if (!ctxs[j]) return fail(F2JK_EINVAL, "null context in group");
        if (ndens[j] < 1 || ndens[j] > 2) return fail(F2JK_EUNSUPPORTED, "grouped jobs have 1 or 2 densities");
        // one J vector width (2) for the whole group, so one kernel serves every job
        int rc = make_plan(ctxs[j], ndens[j], nocc + k, precision, flags, true, plans[j], 2, true);
        if (rc) return rc;
        const Plan& p = plans[j];
        if (p.N % 2) return fail(F2JK_EUNSUPPORTED, "grouped builds need an even nbf");
        if (p.nchunks != 1) return fail(F2JK_EUNSUPPORTED, "grouped builds need X in one chunk");
        if (ctxs[j]->device != ctxs[0]->device) return fail(F2JK_EINVAL, "group spans devices");
        k += ndens[j];
    }
    return F2JK_OK;
}

template <int PH, int PREC, int NBW = 2, int VA = 2>
int launch_group_phase(const GArgs& a, hipStream_t st) {
    long long total = 0;
    for (int j = 0; j < a.njobs; ++j) total += phase_blocks<PREC>(a.job[j], PH);
    if (total == 0) return F2JK_OK;
    const unsigned grid = (unsigned)std::min<long long>(total, 4096);
    hipLaunchKernelGGL((grouped_kernel<PH, PREC, NBW, VA>), dim3(grid), dim3(256), 0, st, a);
    F2JK_LAUNCHED("grouped phase " + std::to_string(PH));
    return F2JK_OK;
}

// the half-transform of a BF16x3 group: tile width NBW (the widest job's) and load width VA
template <int VA>
int launch_group_kx_bf16(const GArgs& a, int nbw, hipStream_t st) {
    switch (nbw) {
        case 2: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 2, VA>(a, st);
        case 4: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 4, VA>(a, st);
        case 5: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 5, VA>(a, st);
        case 6: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 6, VA>(a, st);
        case 8: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 8, VA>(a, st);
        default: return launch_group_phase<GP_KX, F2JK_PREC_BF16X3, 10, VA>(a, st);
    }
}

template <int PREC>
int run_group(f2jk_ctx* ctx, const GArgs& a, int nbw, int va, hipStream_t st) {
    int rc;
    TIM(F2JK_PH_PREP, true);
    if ((rc = launch_group_phase<GP_PAD, PREC>(a, st))) return rc;
    if ((rc = launch_group_phase<GP_DKS, PREC>(a, st))) return rc;
    if ((rc = launch_group_phase<GP_DRED, PREC>(a, st))) return rc;
    TIM(F2JK_PH_PREP, false);
    TIM(F2JK_PH_JGAMMA, true);
    if ((rc = launch_group_phase<GP_GAMMA, PREC>(a, st))) return rc;
    if ((rc = launch_group_phase<GP_GRED, PREC>(a, st))) return rc;
    TIM(F2JK_PH_JGAMMA, false);
    TIM(F2JK_PH_JACC, true);
    if ((rc = launch_group_phase<GP_JACC, PREC>(a, st))) return rc;
    if ((rc = launch_group_phase<GP_JRED, PREC>(a, st))) return rc;
    TIM(F2JK_PH_JACC, false);
    TIM(F2JK_PH_KX, true);
    if constexpr (PREC == F2JK_PREC_BF16X3) {
        rc = (va == 4) ? launch_group_kx_bf16<4>(a, nbw, st) : launch_group_kx_bf16<2>(a, nbw, st);
        if (rc) return rc;
    } else {
        if ((rc = launch_group_phase<GP_KX, PREC>(a, st))) return rc;
    }
    TIM(F2JK_PH_KX, false);
    TIM(F2JK_PH_KS, true);
    if ((rc = launch_group_phase<GP_KS, PREC>(a, st))) return rc;
    TIM(F2JK_PH_KS, false);
    TIM(F2JK_PH_KRED, true);
    if ((rc = launch_group_phase<GP_KRED, PREC>(a, st))) return rc;
    TIM(F2JK_PH_KRED, false);
    return F2JK_OK;
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

int f2jk_ctx_create(int device, f2jk_ctx** out) {
    if (!out) return fail(F2JK_EINVAL, "out is NULL");
    int ndev = 0;
    F2JK_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(F2JK_EINVAL, "device index out of range");
    F2JK_HIP(hipSetDevice(device));
    int rc = scratch_init(device);
    if (rc) return rc;
    f2jk_ctx* ctx = new f2jk_ctx();
    ctx->device = device;
    for (int i = 0; i < kMaxEvents + 2; ++i) {
        hipEvent_t* slot = (i < kMaxEvents) ? &ctx->ev[i] : &ctx->ev_total[i - kMaxEvents];
        hipError_t e = hipEventCreate(slot);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j)
                (void)hipEventDestroy(j < kMaxEvents ? ctx->ev[j] : ctx->ev_total[j - kMaxEvents]);
            delete ctx;
            return fail(F2JK_EHIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    hipError_t e1 = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    hipError_t e2 = (e1 == hipSuccess) ? hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) : e1;
    hipError_t e3 = (e2 == hipSuccess) ? hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) : e2;
    if (e3 != hipSuccess) {
        f2jk_ctx_destroy(ctx);
        return fail(F2JK_EHIP, std::string("side stream: ") + hipGetErrorString(e3));
    }
    *out = ctx;
    return F2JK_OK;
}

int f2jk_ctx_destroy(f2jk_ctx* ctx) {
    if (!ctx) return F2JK_OK;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < kMaxEvents; ++i) (void)hipEventDestroy(ctx->ev[i]);
    for (int i = 0; i < 2; ++i) (void)hipEventDestroy(ctx->ev_total[i]);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    for (int v = 0; v < 5; ++v)
        if (ctx->rowoff[v]) (void)hipFree(ctx->rowoff[v]);
    free_jlist(ctx, F2JK_F64);
    free_jlist(ctx, F2JK_F32);
    delete ctx;
    return F2JK_OK;
}

int f2jk_set_B(f2jk_ctx* ctx, int dtype, const void* B, int64_t naux, int64_t nbf) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    if (dtype != F2JK_F64 && dtype != F2JK_F32) return fail(F2JK_EINVAL, "unknown dtype");
    if (!B) return fail(F2JK_EINVAL, "B is NULL");
    if (naux <= 0 || nbf <= 0 || naux > INT32_MAX || nbf > 65535)
        return fail(F2JK_EINVAL, "naux/nbf out of range");
    if (((uintptr_t)B) % 16) return fail(F2JK_EINVAL, "B must be 16-byte aligned");
    const void* other = ctx->B[1 - dtype];
    if (other && (ctx->naux != naux || ctx->nbf != nbf))
        return fail(F2JK_EINVAL, "B copies must share naux and nbf");
    if (ctx->nbf != nbf) {
        (void)hipSetDevice(ctx->device);
        for (int v = 0; v < 5; ++v)
            if (ctx->rowoff[v]) {
                (void)hipFree(ctx->rowoff[v]);
                ctx->rowoff[v] = nullptr;
                ctx->nvec[v] = 0;
            }
    }
    ctx->B[dtype] = B;
    ctx->naux = naux;
    ctx->nbf = nbf;
    // build the J element tables now so that builds allocate nothing
    int rc = ensure_rowoff(ctx, j_vec_for(dtype, nbf));
    if (rc) return rc;
    if (ctx->j_sparse && (rc = build_jlist(ctx, dtype))) return rc;
    g_last_error.clear();
    return F2JK_OK;
}

int f2jk_get_shape(const f2jk_ctx* ctx, int64_t* naux, int64_t* nbf) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    if (naux) *naux = ctx->naux;
    if (nbf) *nbf = ctx->nbf;
    return F2JK_OK;
}

int f2jk_set_tuning(f2jk_ctx* ctx, const char* key, int64_t value) {
    if (!ctx || !key) return fail(F2JK_EINVAL, "null argument");
    if (!strcmp(key, "xchunk_bytes")) {
        if (value < (1 << 20)) return fail(F2JK_EINVAL, "xchunk_bytes must be >= 1 MiB");
        ctx->xchunk_bytes = (size_t)value;
    } else if (!strcmp(key, "j_nontemporal")) {
        ctx->j_nontemporal = value != 0;
    } else if (!strcmp(key, "j_sparse")) {
        ctx->j_sparse = value != 0;
        for (int dt = 0; dt < 2; ++dt) {
            if (ctx->j_sparse) {
                int rc = build_jlist(ctx, dt);
                if (rc) return rc;
            } else {
                free_jlist(ctx, dt);
            }
        }
    } else if (!strcmp(key, "ks_dma")) {
        if (value < -1 || value > 1) return fail(F2JK_EINVAL, "ks_dma must be -1, 0 or 1");
        ctx->ks_dma = (int)value;
    } else if (!strcmp(key, "kx_wide")) {
        if (value < -1 || value > 1) return fail(F2JK_EINVAL, "kx_wide must be -1, 0 or 1");
        ctx->kx_wide = (int)value;
    } else if (!strcmp(key, "kx_rows")) {
        if (value != 0 && value != 128 && value != 256) return fail(F2JK_EINVAL, "kx_rows must be 0, 128 or 256");
        ctx->kx_rows = (int)value;
    } else if (!strcmp(key, "bf16_nbw_max")) {
        if (value != 2 && value != 4 && value != 5 && value != 6 && value != 8 && value != 10)
            return fail(F2JK_EINVAL, "bf16_nbw_max must be 2, 4, 5, 6, 8 or 10");
        ctx->bf16_nbw_max = (int)value;
    } else if (!strcmp(key, "overlap_jk")) {
        ctx->overlap_jk = value != 0;
    } else if (!strcmp(key, "ks_tile")) {
        if (value != 0 && value != 128 && value != 256) return fail(F2JK_EINVAL, "ks_tile must be 0, 128 or 256");
        ctx->ks_tile = (int)value;
    } else if (!strcmp(key, "kx_waves")) {
        if (value != 4 && value != 8) return fail(F2JK_EINVAL, "kx_waves must be 4 or 8");
        ctx->kx_waves = (int)value;

    } else if (!strcmp(key, "target_blocks")) {
        if (value < 1 || value > (1 << 20)) return fail(F2JK_EINVAL, "target_blocks out of range");
        ctx->target_blocks = (int)value;
    } else {
        return fail(F2JK_EINVAL, std::string("unknown tuning key ") + key);
    }
    return F2JK_OK;
}

int f2jk_workspace_size(const f2jk_ctx* ctx, int ndens, const int64_t* nocc, int precision, int flags,
                        size_t* bytes) {
    if (!ctx || !bytes) return fail(F2JK_EINVAL, "null argument");
    Plan p;
    // a build_jk plan (J from C) is the largest; use it whenever C is known
    int rc = make_plan(ctx, ndens, nocc, precision, flags, nocc != nullptr, p);
    if (rc) return rc;
    *bytes = p.total;
    return F2JK_OK;
}

int f2jk_build_jk(f2jk_ctx* ctx, int ndens, const double* const* C, const int64_t* nocc,
                  const int64_t* ldc, double* const* J, double* const* K, int precision, int flags,
                  void* ws, size_t ws_bytes, void* stream) {
    if (!C || !nocc) return fail(F2JK_EINVAL, "C and nocc are required");
    return build_common(ctx, ndens, C, nocc, ldc, nullptr, 0, J, K, precision, flags, ws, ws_bytes,
                        stream, true);
}

int f2jk_build_jk_noalloc(f2jk_ctx* ctx, int ndens, const double* const* C, const int64_t* nocc,
                          const int64_t* ldc, double* const* J, double* const* K, int precision, int flags,
                          void* stream) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    if (!C || !nocc) return fail(F2JK_EINVAL, "C and nocc are required");
    void* ws = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    if (!ws) return fail(F2JK_ESTATE, "no scratch address for this device");
    PlanCaps caps;
    int rc = fit_plan(ctx, ndens, nocc, precision, flags, F2JK_NOALLOC_SCRATCH_BYTES, caps);
    if (rc) return rc;
    return build_common(ctx, ndens, C, nocc, ldc, nullptr, 0, J, K, precision, flags, ws,
                        F2JK_NOALLOC_SCRATCH_BYTES, stream, true, &caps, true);
}

int f2jk_build_j(f2jk_ctx* ctx, int ndens, const double* const* D, int64_t ldd, double* const* J,
                 int precision, int flags, void* ws, size_t ws_bytes, void* stream) {
    if (!D || !J) return fail(F2JK_EINVAL, "D and J are required");
    if (ctx && ldd < ctx->nbf) return fail(F2JK_EINVAL, "ldd < nbf");
    const int fl = (flags & F2JK_TIMING) | F2JK_WANT_J;
    return build_common(ctx, ndens, nullptr, nullptr, nullptr, D, ldd, J, nullptr, precision, fl, ws,
                        ws_bytes, stream, false);
}

int f2jk_build_k(f2jk_ctx* ctx, int ndens, const double* const* C, const int64_t* nocc,
                 const int64_t* ldc, double* const* K, int precision, int flags, void* ws,
                 size_t ws_bytes, void* stream) {
    if (!C || !nocc || !K) return fail(F2JK_EINVAL, "C, nocc and K are required");
    const int fl = (flags & F2JK_TIMING) | F2JK_WANT_K;
    return build_common(ctx, ndens, C, nocc, ldc, nullptr, 0, nullptr, K, precision, fl, ws, ws_bytes,
                        stream, false);
}

int f2jk_last_timings(f2jk_ctx* ctx, double* ms, int n) {
    if (!ctx || !ms) return fail(F2JK_EINVAL, "null argument");
    double acc[F2JK_NPHASE] = {0};
    for (int i = 0; i + 1 < ctx->nev; i += 2) {
        F2JK_HIP(hipEventSynchronize(ctx->ev[i + 1]));
        float t = 0.f;
        F2JK_HIP(hipEventElapsedTime(&t, ctx->ev[i], ctx->ev[i + 1]));
        acc[ctx->ev_phase[i / 2]] += t;
    }
    if (ctx->total_recorded) {
        F2JK_HIP(hipEventSynchronize(ctx->ev_total[1]));
        float t = 0.f;
        F2JK_HIP(hipEventElapsedTime(&t, ctx->ev_total[0], ctx->ev_total[1]));
        acc[F2JK_PH_TOTAL] = t;
    }
    for (int i = 0; i < n && i < F2JK_NPHASE; ++i) ms[i] = acc[i];
    return F2JK_OK;
}

int f2jk_last_launches(const f2jk_ctx* ctx, int* counts, int n) {
    if (!ctx || !counts) return fail(F2JK_EINVAL, "null argument");
    int c[F2JK_NPHASE] = {0};
    for (int i = 0; i + 1 < ctx->nev; i += 2) c[ctx->ev_phase[i / 2]]++;
    c[F2JK_PH_TOTAL] = ctx->total_recorded ? 1 : 0;
    for (int i = 0; i < n && i < F2JK_NPHASE; ++i) counts[i] = c[i];
    return F2JK_OK;
}

int f2jk_work(const f2jk_ctx* ctx, int ndens, const int64_t* nocc, int precision, int flags,
              double* j_bytes, double* k_flops) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    const double N = (double)ctx->nbf, Q = (double)ctx->naux;
    const double sb = (precision == F2JK_PREC_FP64) ? 8.0 : 4.0;
    if (j_bytes) {
        // two passes over the lower triangle of every B_P, plus D/J and gamma traffic
        *j_bytes = (flags & F2JK_WANT_J)
                       ? 2.0 * Q * (N * (N + 1) / 2) * sb + 8.0 * ndens * (2.0 * Q + 2.0 * N * N)
                       : 0.0;
    }
    if (k_flops) {
        double f = 0.0;
        if ((flags & F2JK_WANT_K) && nocc)
            for (int d = 0; d < ndens; ++d) {
                const double o = (double)nocc[d];
                f += 2.0 * Q * N * N * o + N * (N + 1) * Q * o;
            }
        *k_flops = f;
    }
    return F2JK_OK;
}

int f2jk_group_workspace_size(f2jk_ctx* const* ctxs, int njobs, const int* ndens, const int64_t* nocc,
                              int precision, int flags, size_t* job_bytes) {
    if (!job_bytes) return fail(F2JK_EINVAL, "null argument");
    std::vector<Plan> plans;
    int rc = group_plan(ctxs, njobs, ndens, nocc, precision, flags, plans);
    if (rc) return rc;
    for (int j = 0; j < njobs; ++j) job_bytes[j] = plans[j].total;
    return F2JK_OK;
}

int f2jk_group_build_jk(f2jk_ctx* const* ctxs, int njobs, const int* ndens, const double* const* C,
                        const int64_t* nocc, const int64_t* ldc, double* const* J, double* const* K,
                        int precision, int flags, void* const* ws, const size_t* ws_bytes, void* stream) {
    if (!C || !ldc || !K || !ws || !ws_bytes) return fail(F2JK_EINVAL, "null argument");
    std::vector<Plan> plans;
    int rc = group_plan(ctxs, njobs, ndens, nocc, precision, flags, plans);
    if (rc) return rc;
    const bool want_j = flags & F2JK_WANT_J;
    if (want_j && !J) return fail(F2JK_EINVAL, "null argument");
    F2JK_HIP(hipSetDevice(ctxs[0]->device));
    hipStream_t st = (hipStream_t)stream;
    // the group's BF16x3 tile width (widest job) and load width (4 when every nbf allows it)
    int nbw = 2, va = 4;
    for (int j = 0; j < njobs; ++j) {
        nbw = std::max(nbw, plans[j].nbw);
        if (plans[j].N % 4) va = 2;
    }
    std::vector<GJob> jobs(njobs);
    int64_t k = 0;
    for (int j = 0; j < njobs; ++j) {
        const Plan& p = plans[j];
        if ((rc = check_ws(p, ws[j], ws_bytes[j]))) return rc;
        if (want_j && (rc = ensure_rowoff(ctxs[j], p.jvec))) return rc;
        char* w = (char*)ws[j];
        GJob& g = jobs[j];
        std::memset(&g, 0, sizeof(g));
        g.B = ctxs[j]->B[precision == F2JK_PREC_FP64 ? F2JK_F64 : F2JK_F32];
        g.rowoff = want_j ? ctxs[j]->rowoff[p.jvec] : nullptr;
        g.Cp = (double*)(w + p.off_cp);
        g.Cp32 = (float*)(w + p.off_cp32);
        if (precision == F2JK_PREC_BF16X3) {
            g.Cth = (u16*)(w + p.off_cth);
            g.Ctl = (u16*)(w + p.off_ctl);
        }
        g.X = (void*)(w + p.off_x);
        g.Kpart = (double*)(w + p.off_kpart);
        if (want_j) {
            g.D = (double*)(w + p.off_d);
            g.Dpart = (double*)(w + p.off_dpart);
            g.gpart = (double*)(w + p.off_gpart);
            g.gamma = (double*)(w + p.off_gamma);
            g.jpart = (double*)(w + p.off_jpart);
        }
        g.want_j = want_j ? 1 : 0;
        g.N = p.N;
        g.Q = p.Q;
        g.nd = p.nd;
        g.OPAD = p.OPAD;
        g.KP = p.KP;
        for (int d = 0; d < p.nd; ++d, ++k) {
            if (!C[k] || !K[k] || (want_j && !J[k])) return fail(F2JK_EINVAL, "null C/J/K pointer in group");
            if (nocc[k] > 0 && ldc[k] < nocc[k]) return fail(F2JK_EINVAL, "ldc < nocc");
            g.C[d] = C[k];
            g.ldc[d] = nocc[k] > 0 ? ldc[k] : 1;
            g.nocc[d] = (int)nocc[k];
            g.width[d] = p.width[d];
            g.ioff[d] = p.ioff[d];
            g.J[d] = want_j ? J[k] : nullptr;
            g.K[d] = K[k];
        }
        g.nvec = p.nvec;
        g.neg = p.neg;
        g.npc1 = p.npc1;
        g.npc2 = p.npc2;
        g.ppw2 = p.ppw2;
        g.nlower = p.nlower;
        g.nsplit = std::min(p.nsplit, std::min(p.Pc, p.Q));
        g.ncolt = (precision == F2JK_PREC_BF16X3) ? (int)cdiv(p.OPAD, 32 * nbw) : p.ncolt;
        // the D-build tiles use the same lower-tile count as the K tiles
        if (p.nlower != (int)(cdiv(p.N, kTile) * (cdiv(p.N, kTile) + 1) / 2))
            return fail(F2JK_EINVAL, "inconsistent tile plan");
    }
    f2jk_ctx* ctx = ctxs[0];  // timing events of the group
    tim_begin(ctx, flags & F2JK_TIMING);
    TIM(F2JK_PH_TOTAL, true);
    for (int j0 = 0; j0 < njobs; j0 += kMaxGroup) {
        GArgs a;
        std::memset(&a, 0, sizeof(a));
        a.njobs = std::min(kMaxGroup, njobs - j0);
        for (int j = 0; j < a.njobs; ++j) a.job[j] = jobs[j0 + j];
        if (precision == F2JK_PREC_FP64)
            rc = run_group<F2JK_PREC_FP64>(ctx, a, nbw, va, st);
        else if (precision == F2JK_PREC_FP32)
            rc = run_group<F2JK_PREC_FP32>(ctx, a, nbw, va, st);
        else
            rc = run_group<F2JK_PREC_BF16X3>(ctx, a, nbw, va, st);
        if (rc) return rc;
    }
    TIM(F2JK_PH_TOTAL, false);
    g_last_error.clear();
    return F2JK_OK;
}

int f2jk_half_transform(f2jk_ctx* ctx, const double* C, int64_t ncol, int64_t ldc, int64_t p0, int64_t p1,
                        double* X, int64_t ldx, int precision, void* ws, size_t ws_bytes, void* stream) {
    if (!ctx) return fail(F2JK_EINVAL, "null context");
    if (!C || !X) return fail(F2JK_EINVAL, "C and X are required");
    if (precision != F2JK_PREC_FP64) return fail(F2JK_EUNSUPPORTED, "the half-transform is FP64");
    if (ncol < 1 || ldc < ncol || ldx < ncol) return fail(F2JK_EINVAL, "ncol, ldc or ldx invalid");
    Plan p;
    int rc = make_plan(ctx, 1, &ncol, precision, F2JK_WANT_K, false, p);
    if (rc) return rc;
    if (p0 < 0 || p1 > p.Q || p0 >= p1) return fail(F2JK_EINVAL, "aux range [p0, p1) invalid");
    if ((rc = check_ws(p, ws, ws_bytes))) return rc;
    F2JK_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)ws;
    const double* Cs[1] = {C};
    const int64_t lds[1] = {ldc};
    if ((rc = pack_c(ctx, p, Cs, &ncol, lds, w, st, false))) return rc;
    const double* B = (const double*)ctx->B[F2JK_F64];
    const double* cp = (const double*)(w + p.off_cp);
    double* Xc = (double*)(w + p.off_x);
    const int64_t NN = (int64_t)p.N * p.N;
    // the kernel writes rows of OPAD columns: straight into the caller's X
    // when its row stride is exactly that, through the workspace otherwise
    const bool direct = (ldx == p.OPAD);
    for (int64_t P0 = p0; P0 < p1; P0 += p.Pc) {
        const int pc = (int)std::min<int64_t>(p.Pc, p1 - P0);
        const int64_t R = (int64_t)pc * p.N;
        const unsigned grid = (unsigned)(cdiv(R, kTile) * p.ncolt);
        double* dst = X + (P0 - p0) * p.N * ldx;
        double* Xo = direct ? dst : Xc;
        launch_kx_typed<double>(p, grid, B + (int64_t)P0 * NN, R, cp, Xo, st);
        F2JK_LAUNCHED("kx");
        if (!direct) {
            hipLaunchKernelGGL(x_copy_kernel, dim3((unsigned)cdiv(R * ncol, 256)), dim3(256), 0, st, Xc, R,
                               p.OPAD, (int)ncol, dst, ldx);
            F2JK_LAUNCHED("x_copy");
        }
    }
    g_last_error.clear();
    return F2JK_OK;
}

const char* f2jk_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
