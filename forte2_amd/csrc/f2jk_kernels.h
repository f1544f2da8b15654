// Device kernels of the density-fitted J/K builder (gfx950 / CDNA4).
//
// Data layout in HBM
//   B[P][m][n]   naux x nbf x nbf, row-major, symmetric in (m, n) -- the
//                reference layout of FockBuilder.B_Pmn
//                (forte2/jkbuilder/jkbuilder.py:52-72).
//   Cp[n][i]     nbf x OPAD, occupied coefficients of all densities side by
//                side, each density's block padded with zero columns to a
//                multiple of 16 (the MFMA k step of the contraction).
//   X[P][m][i]   half-transformed tensor of one P chunk, Pc x nbf x OPAD.
//
// J (two passes over the lower triangle of B, HBM-bound):
//   pass 1  gamma_P = sum_{m>=n} w_mn B_Pmn D_mn,  w = 2 (m>n), 1 (m=n)
//   pass 2  J_mn    = sum_P gamma_P B_Pmn           (m >= n, mirrored)
//   The lower triangle is walked as "vector elements": row m is cut into
//   ceil((m+1)/VEC) aligned VEC-wide pieces (16-byte loads), numbered
//   row after row; rowoff[m] is the number of pieces in rows < m.
// K (MFMA, compute-bound):
//   KX  X = B_chunk Cp           one tall GEMM, (Pc*N) x N x OPAD
//   KS  Kpart[s] += sum_{P in split s} X_P X_P^T   lower 128x128 tiles only
//   KR  K = sum_s Kpart[s], lower triangle mirrored to the upper one
// All reductions run in a fixed order (no atomics): the result of a build is
// a function of its inputs and shapes only.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace f2jk {

constexpr int kWave = 64;
constexpr int kJThreads = 256;
constexpr int kJFlush = 32;  // P values reduced across waves per LDS flush

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
    T v[V];
};

template <typename T, int V>
__device__ __forceinline__ Vec<T, V> load_vec(const T* p) {
    return *reinterpret_cast<const Vec<T, V>*>(p);
}

template <typename T, int V>
struct NTVec {
    typedef T type __attribute__((ext_vector_type(V)));
};

// streaming load that bypasses cache retention (data read once per pass)
template <typename T, int V, bool NT>
__device__ __forceinline__ Vec<T, V> stream_vec(const T* p) {
    if constexpr (NT && V > 1) {
        typedef typename NTVec<T, V>::type vt;
        const vt x = __builtin_nontemporal_load(reinterpret_cast<const vt*>(p));
        Vec<T, V> r;
#pragma unroll
        for (int v = 0; v < V; ++v) r.v[v] = x[v];
        return r;
    } else if constexpr (NT) {
        Vec<T, V> r;
        r.v[0] = __builtin_nontemporal_load(p);
        return r;
    } else {
        return load_vec<T, V>(p);
    }
}

template <int ND>
struct DPtrs {
    const double* p[ND];
};
template <int ND>
struct JPtrs {
    double* p[ND];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// largest m in [0, n) with rowoff[m] <= e  (rowoff[0] = 0, rowoff[n] = total)
__device__ __forceinline__ int find_row(const int* __restrict__ rowoff, int n, int e) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (rowoff[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Give every XCD a contiguous range of logical tile ids so that tiles that
// share operands run on the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    constexpr int NX = 8;
    const int xcd = b % NX, local = b / NX;
    const int base = nblocks / NX, rem = nblocks % NX;
    return xcd * base + (xcd < rem ? xcd : rem) + local;
}

// ---------------------------------------------------------------------------
// J pass 1: partial gamma per (element group, density, P).
// grid = (n_element_groups, ceil(Q / ppw)), block = 256.
// ---------------------------------------------------------------------------
// Vector element e of the lower triangle: row m and first column col.  With
// an element list (only the elements where B is nonzero for some P, packed
// as m << 16 | col) e indexes the list; otherwise elements are numbered row
// after row through rowoff.
template <int VEC>
__device__ __forceinline__ void j_element(const int* __restrict__ rowoff, const uint32_t* __restrict__ elist, int N,
                                          int e, int& m, int& col) {
    if (elist) {
        const uint32_t x = elist[e];
        m = (int)(x >> 16);
        col = (int)(x & 0xffffu);
    } else {
        m = find_row(rowoff, N, e);
        col = (e - rowoff[m]) * VEC;
    }
}

template <typename TB, int VEC, int EPT, int ND, bool NT = false>
__device__ __forceinline__ void j_gamma_kernel_body(
    const TB* __restrict__ B, int N, int Q, DPtrs<ND> D, int64_t ldd,
    const int* __restrict__ rowoff, int nvec, int ppw, double* __restrict__ gpart, int bx, int by,
    const uint32_t* __restrict__ elist = nullptr) {
    __shared__ double red[kJThreads / kWave][kJFlush][ND];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int64_t NN = (int64_t)N * N;

    int64_t off[EPT];
    double d[EPT][ND][VEC];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = bx * (kJThreads * EPT) + k * kJThreads + tid;
        off[k] = 0;
#pragma unroll
        for (int dd = 0; dd < ND; ++dd)
#pragma unroll
            for (int v = 0; v < VEC; ++v) d[k][dd][v] = 0.0;
        if (e < nvec) {
            int m, col;
            j_element<VEC>(rowoff, elist, N, e, m, col);
            off[k] = (int64_t)m * N + col;
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const int c = col + v;
                const double w = (c < m) ? 2.0 : (c == m ? 1.0 : 0.0);
#pragma unroll
                for (int dd = 0; dd < ND; ++dd)
                    d[k][dd][v] = (w != 0.0) ? w * D.p[dd][(int64_t)m * ldd + c] : 0.0;
            }
        }
    }

    const int P0 = by * ppw;
    const int P1 = min(Q, P0 + ppw);
    for (int Pb = P0; Pb < P1; Pb += kJFlush) {
        const int np = min(kJFlush, P1 - Pb);
        int j = 0;
        for (; j + 1 < np; j += 2) {
            const TB* B0 = B + (int64_t)(Pb + j) * NN;
            const TB* B1 = B0 + NN;
            Vec<TB, VEC> b0[EPT], b1[EPT];
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                b0[k] = stream_vec<TB, VEC, NT>(B0 + off[k]);
                b1[k] = stream_vec<TB, VEC, NT>(B1 + off[k]);
            }
            double a0[ND], a1[ND];
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) a0[dd] = a1[dd] = 0.0;
#pragma unroll
            for (int k = 0; k < EPT; ++k)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
#pragma unroll
                    for (int dd = 0; dd < ND; ++dd) {
                        a0[dd] = fma((double)b0[k].v[v], d[k][dd][v], a0[dd]);
                        a1[dd] = fma((double)b1[k].v[v], d[k][dd][v], a1[dd]);
                    }
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) {
                a0[dd] = wave_sum(a0[dd]);
                a1[dd] = wave_sum(a1[dd]);
            }
            if (lane == 0) {
#pragma unroll
                for (int dd = 0; dd < ND; ++dd) {
                    red[wid][j][dd] = a0[dd];
                    red[wid][j + 1][dd] = a1[dd];
                }
            }
        }
        if (j < np) {
            const TB* B0 = B + (int64_t)(Pb + j) * NN;
            double a0[ND];
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) a0[dd] = 0.0;
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                Vec<TB, VEC> b0 = stream_vec<TB, VEC, NT>(B0 + off[k]);
#pragma unroll
                for (int v = 0; v < VEC; ++v)
#pragma unroll
                    for (int dd = 0; dd < ND; ++dd)
                        a0[dd] = fma((double)b0.v[v], d[k][dd][v], a0[dd]);
            }
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) a0[dd] = wave_sum(a0[dd]);
            if (lane == 0) {
#pragma unroll
                for (int dd = 0; dd < ND; ++dd) red[wid][j][dd] = a0[dd];
            }
        }
        __syncthreads();
        for (int t = tid; t < np * ND; t += kJThreads) {
            const int jj = t / ND, dd = t % ND;
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kJThreads / kWave; ++w) s += red[w][jj][dd];
            gpart[((int64_t)bx * ND + dd) * Q + Pb + jj] = s;
        }
        __syncthreads();
    }
}
template <typename TB, int VEC, int EPT, int ND, bool NT = false>
__global__ __launch_bounds__(kJThreads) void j_gamma_kernel(
    const TB* __restrict__ B, int N, int Q, DPtrs<ND> D, int64_t ldd,
    const int* __restrict__ rowoff, const uint32_t* __restrict__ elist, int nvec, int ppw,
    double* __restrict__ gpart) {
    j_gamma_kernel_body<TB, VEC, EPT, ND, NT>(B, N, Q, D, ldd, rowoff, nvec, ppw, gpart, blockIdx.x, blockIdx.y,
                                              elist);
}


// gamma[dd][P] = sum_g gpart[g][dd][P], g in a fixed order.
__global__ void j_gamma_reduce_kernel(const double* __restrict__ gpart, int ngroups,
                                      int ND, int Q, double* __restrict__ gamma) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ND * Q) return;
    const int dd = (int)(t / Q), P = (int)(t % Q);
    double s = 0.0;
    for (int g = 0; g < ngroups; ++g) s += gpart[((int64_t)g * ND + dd) * Q + P];
    gamma[t] = s;
}

// ---------------------------------------------------------------------------
// J pass 2: partial J per (P chunk, density, element).
// grid = (n_element_groups, n_p_chunks), block = 256.
// ---------------------------------------------------------------------------
template <typename TB, int VEC, int EPT, int ND, bool NT = false>
__device__ __forceinline__ void j_accum_kernel_body(
    const TB* __restrict__ B, int N, int Q, const double* __restrict__ gamma,
    const int* __restrict__ rowoff, int nvec, int ppw, double* __restrict__ jpart, int bx, int by,
    const uint32_t* __restrict__ elist = nullptr) {
    const int tid = threadIdx.x;
    const int64_t NN = (int64_t)N * N;
    int64_t off[EPT];
    int eidx[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = bx * (kJThreads * EPT) + k * kJThreads + tid;
        eidx[k] = e;
        off[k] = 0;
        if (e < nvec) {
            int m, col;
            j_element<VEC>(rowoff, elist, N, e, m, col);
            off[k] = (int64_t)m * N + col;
        }
    }
    double acc[EPT][ND][VEC];
#pragma unroll
    for (int k = 0; k < EPT; ++k)
#pragma unroll
        for (int dd = 0; dd < ND; ++dd)
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[k][dd][v] = 0.0;

    const int P0 = by * ppw;
    const int P1 = min(Q, P0 + ppw);
    int P = P0;
    for (; P + 1 < P1; P += 2) {
        const TB* B0 = B + (int64_t)P * NN;
        const TB* B1 = B0 + NN;
        Vec<TB, VEC> b0[EPT], b1[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            b0[k] = stream_vec<TB, VEC, NT>(B0 + off[k]);
            b1[k] = stream_vec<TB, VEC, NT>(B1 + off[k]);
        }
        double g0[ND], g1[ND];
#pragma unroll
        for (int dd = 0; dd < ND; ++dd) {
            g0[dd] = gamma[(int64_t)dd * Q + P];
            g1[dd] = gamma[(int64_t)dd * Q + P + 1];
        }
#pragma unroll
        for (int k = 0; k < EPT; ++k)
#pragma unroll
            for (int dd = 0; dd < ND; ++dd)
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    acc[k][dd][v] = fma(g0[dd], (double)b0[k].v[v], acc[k][dd][v]);
                    acc[k][dd][v] = fma(g1[dd], (double)b1[k].v[v], acc[k][dd][v]);
                }
    }
    if (P < P1) {
        const TB* B0 = B + (int64_t)P * NN;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            Vec<TB, VEC> b0 = stream_vec<TB, VEC, NT>(B0 + off[k]);
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) {
                const double g = gamma[(int64_t)dd * Q + P];
#pragma unroll
                for (int v = 0; v < VEC; ++v) acc[k][dd][v] = fma(g, (double)b0.v[v], acc[k][dd][v]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        if (eidx[k] >= nvec) continue;
#pragma unroll
        for (int dd = 0; dd < ND; ++dd) {
            double* dst = jpart + (((int64_t)by * ND + dd) * nvec + eidx[k]) * VEC;
#pragma unroll
            for (int v = 0; v < VEC; ++v) dst[v] = acc[k][dd][v];
        }
    }
}
template <typename TB, int VEC, int EPT, int ND, bool NT = false>
__global__ __launch_bounds__(kJThreads) void j_accum_kernel(
    const TB* __restrict__ B, int N, int Q, const double* __restrict__ gamma,
    const int* __restrict__ rowoff, const uint32_t* __restrict__ elist, int nvec, int ppw,
    double* __restrict__ jpart) {
    j_accum_kernel_body<TB, VEC, EPT, ND, NT>(B, N, Q, gamma, rowoff, nvec, ppw, jpart, blockIdx.x, blockIdx.y,
                                              elist);
}


// J[dd][m][c] = J[dd][c][m] = sum_chunk jpart[chunk][dd][e][v]  (c = col+v <= m)
template <int VEC, int ND>
__global__ void j_reduce_kernel(const double* __restrict__ jpart, int nchunks, int N,
                                const int* __restrict__ rowoff, const uint32_t* __restrict__ elist, int nvec,
                                JPtrs<ND> J) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvec) return;
    int m, col;
    j_element<VEC>(rowoff, elist, N, e, m, col);
#pragma unroll
    for (int dd = 0; dd < ND; ++dd) {
        double s[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) s[v] = 0.0;
        for (int c = 0; c < nchunks; ++c) {
            const double* src = jpart + (((int64_t)c * ND + dd) * nvec + e) * VEC;
#pragma unroll
            for (int v = 0; v < VEC; ++v) s[v] += src[v];
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            const int cc = col + v;
            if (cc <= m) {
                J.p[dd][(int64_t)m * N + cc] = s[v];
                J.p[dd][(int64_t)cc * N + m] = s[v];
            }
        }
    }
}

// flag[e] = 1 when B[P][m][c] != 0 for some P and some c <= m of vector
// element e (lower triangle, numbered row after row): the elements the J
// passes have to read.  One thread per element, all P.
template <typename TB, int VEC>
__global__ void j_scan_kernel(const TB* __restrict__ B, int N, int Q, const int* __restrict__ rowoff, int nvec,
                              unsigned char* __restrict__ flag) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvec) return;
    const int m = find_row(rowoff, N, e);
    const int col = (e - rowoff[m]) * VEC;
    const int64_t NN = (int64_t)N * N;
    const TB* p = B + (int64_t)m * N + col;
    int nz = 0;
    for (int P = 0; P < Q && !nz; ++P) {
        const Vec<TB, VEC> b = load_vec<TB, VEC>(p + P * NN);
#pragma unroll
        for (int v = 0; v < VEC; ++v) nz |= (col + v <= m) && (b.v[v] != TB(0));
    }
    flag[e] = (unsigned char)nz;
}

// x[i] = 0 for i < n
__global__ void fill_zero_kernel(double* __restrict__ x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// ---------------------------------------------------------------------------
// MFMA helpers.  v_mfma_{f64,f32}_16x16x4: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Accumulator element r of lane l holds D[i][l%16] with
//   f32:  i = 4*(l/16) + r
//   f64:  i = (l/16) + 4*r          (measured on gfx950)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int acc_row(int lane, int r) {
    if constexpr (sizeof(T) == 8)
        return (lane >> 4) + 4 * r;
    else
        return 4 * (lane >> 4) + r;
}

__device__ __forceinline__ dx4 mfma4(double a, double b, dx4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 mfma4(float a, float b, fx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
struct Acc;
template <>
struct Acc<double> {
    typedef dx4 type;
};
template <>
struct Acc<float> {
    typedef fx4 type;
};

constexpr int kTile = 128;   // output tile edge (rows and columns)
constexpr int kBK = 16;      // k extent of one LDS stage
constexpr int kGThreads = 256;

// ---------------------------------------------------------------------------
// KX:  X[r][c] = sum_k A[r][k] Cp[k][c]
//   A  = B of the chunk viewed as R x K row-major (R = Pc*N rows, K = N)
//   Cp = K x OPAD row-major (zero-padded columns), X = R x OPAD row-major.
// Tile 128 x 128, 4 waves as 2 x 2 blocks of 64 x 64 (4 x 4 MFMA blocks each).
// VA: elements per global load of A (16 B when the row length allows).
// ---------------------------------------------------------------------------
template <typename T, int VA>
__device__ __forceinline__ void kx_kernel_body(const T* __restrict__ A, int64_t R,
                                                       int K, const T* __restrict__ Cp,
                                                       int OPAD, T* __restrict__ X,
                                                       int ncolt, int b, int nblocks) {
    typedef typename Acc<T>::type acc_t;
    constexpr int VB = 16 / sizeof(T);       // B tile loads are always 16 B (OPAD % 16 == 0)
    __shared__ T As[2][kTile][kBK + 1];  // two stages: MFMAs read one, the next k-step fills the other
    __shared__ T Bs[2][kBK][kTile];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int t = xcd_remap(b, nblocks);
    const int64_t r0 = (int64_t)(t / ncolt) * kTile;
    const int c0 = (t % ncolt) * kTile;
    const int wr = wid >> 1, wc = wid & 1;
    const int wcol0 = c0 + wc * 64;
    int nbv = (OPAD - wcol0) / 16;
    nbv = nbv < 0 ? 0 : (nbv > 4 ? 4 : nbv);

    constexpr int A_LD = (kTile * kBK) / (kGThreads * VA);  // loads per thread
    constexpr int B_LD = (kTile * kBK) / (kGThreads * VB);
    Vec<T, VA> ra[A_LD];
    Vec<T, VB> rb[B_LD];

    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int s = 0; s < A_LD; ++s) {
            const int idx = tid + kGThreads * s;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
            const int64_t gr = r0 + row;
            const int gk = k0 + kc;
            if (gr < R && gk < K) {
                ra[s] = load_vec<T, VA>(A + gr * K + gk);
            } else {
#pragma unroll
                for (int v = 0; v < VA; ++v) ra[s].v[v] = T(0);
            }
        }
#pragma unroll
        for (int s = 0; s < B_LD; ++s) {
            const int idx = tid + kGThreads * s;
            const int kr = idx / (kTile / VB), cc = (idx % (kTile / VB)) * VB;
            const int gk = k0 + kr, gc = c0 + cc;
            if (gk < K && gc < OPAD) {
                rb[s] = load_vec<T, VB>(Cp + (int64_t)gk * OPAD + gc);
            } else {
#pragma unroll
                for (int v = 0; v < VB; ++v) rb[s].v[v] = T(0);
            }
        }
    };
    auto store_tiles = [&](int stg) {
#pragma unroll
        for (int s = 0; s < A_LD; ++s) {
            const int idx = tid + kGThreads * s;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
#pragma unroll
            for (int v = 0; v < VA; ++v) As[stg][row][kc + v] = ra[s].v[v];
        }
#pragma unroll
        for (int s = 0; s < B_LD; ++s) {
            const int idx = tid + kGThreads * s;
            const int kr = idx / (kTile / VB), cc = (idx % (kTile / VB)) * VB;
#pragma unroll
            for (int v = 0; v < VB; ++v) Bs[stg][kr][cc + v] = rb[s].v[v];
        }
    };

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nk = (K + kBK - 1) / kBK;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();
    if (nk > 1) load_tiles(kBK);
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[cur][wr * 64 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[cur][kk + (lane >> 4)][wc * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (j < nbv) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[i][j] = mfma4(a[i], b[j], acc[i][j]);
                }
            }
        }
        if (kt + 1 < nk) store_tiles(cur ^ 1);  // the other stage was last read before the previous barrier
        if (kt + 2 < nk) load_tiles((kt + 2) * kBK);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= nbv) continue;
            const int col = wcol0 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + wr * 64 + i * 16 + acc_row<T>(lane, r);
                if (row < R) X[row * OPAD + col] = acc[i][j][r];
            }
        }
}
template <typename T, int VA>
__global__ __launch_bounds__(kGThreads) void kx_kernel(const T* __restrict__ A, int64_t R,
                                                       int K, const T* __restrict__ Cp,
                                                       int OPAD, T* __restrict__ X,
                                                       int ncolt) {
    kx_kernel_body<T, VA>(A, R, K, Cp, OPAD, X, ncolt, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// KX, wide tiles (FP64 / FP32): the product of kx_kernel with 128 x 320
// output tiles, so that a 128-row block of B is read once for up to 320
// occupied columns (the 128-column kernel reads it once per column tile: 3x
// at 320 columns).  8 waves as 2 x 4, each 64 x 80 (4 x 5 MFMA blocks, 160
// accumulator registers in FP64), one workgroup per CU.  The k loop, the
// MFMA and each output element's accumulation order are those of kx_kernel,
// so X is bitwise the same.  The k-major coefficient tile is padded by 16
// elements per row: its fragment reads (4 k rows x 16 columns per wave
// instruction) then fall in distinct LDS banks.
// ---------------------------------------------------------------------------
constexpr int kWideCols = 320;
constexpr int kWideThreads = 512;

template <typename T, int VA>
__device__ __forceinline__ void kx_wide_body(const T* __restrict__ A, int64_t R, int K,
                                             const T* __restrict__ Cp, int OPAD, T* __restrict__ X,
                                             int ncolt, int b, int nblocks) {
    typedef typename Acc<T>::type acc_t;
    constexpr int TN = kWideCols, BS = TN + 16;
    constexpr int WC = 4, NB = TN / WC / 16;       // 5 column blocks per wave
    constexpr int VB = 2;                          // coefficient loads: 2 elements each
    constexpr int A_LD = (kTile * kBK) / (kWideThreads * VA);
    constexpr int B_LD = (kBK * TN) / (kWideThreads * VB);
    static_assert(A_LD * kWideThreads * VA == kTile * kBK && B_LD * kWideThreads * VB == kBK * TN,
                  "tiles must split evenly over the threads");
    __shared__ T As[2][kTile][kBK + 1];
    __shared__ T Bs[2][kBK][BS];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = xcd_remap(b, nblocks);
    const int64_t r0 = (int64_t)(t / ncolt) * kTile;
    const int c0 = (t % ncolt) * TN;
    const int wr = wid / WC, wc = wid % WC;
    const int wcol0 = c0 + wc * 16 * NB;
    int nbv = (OPAD - wcol0) / 16;
    nbv = nbv < 0 ? 0 : (nbv > NB ? NB : nbv);

    Vec<T, VA> ra[A_LD];
    Vec<T, VB> rb[B_LD];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int s = 0; s < A_LD; ++s) {
            const int idx = tid + kWideThreads * s;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
            const int64_t gr = r0 + row;
            const int gk = k0 + kc;
            if (gr < R && gk < K) {
                ra[s] = load_vec<T, VA>(A + gr * K + gk);
            } else {
#pragma unroll
                for (int v = 0; v < VA; ++v) ra[s].v[v] = T(0);
            }
        }
#pragma unroll
        for (int s = 0; s < B_LD; ++s) {
            const int idx = tid + kWideThreads * s;
            const int kr = idx / (TN / VB), cc = (idx % (TN / VB)) * VB;
            const int gk = k0 + kr, gc = c0 + cc;
            if (gk < K && gc < OPAD) {
                rb[s] = load_vec<T, VB>(Cp + (int64_t)gk * OPAD + gc);
            } else {
#pragma unroll
                for (int v = 0; v < VB; ++v) rb[s].v[v] = T(0);
            }
        }
    };
    auto store_tiles = [&](int stg) {
#pragma unroll
        for (int s = 0; s < A_LD; ++s) {
            const int idx = tid + kWideThreads * s;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
#pragma unroll
            for (int v = 0; v < VA; ++v) As[stg][row][kc + v] = ra[s].v[v];
        }
#pragma unroll
        for (int s = 0; s < B_LD; ++s) {
            const int idx = tid + kWideThreads * s;
            const int kr = idx / (TN / VB), cc = (idx % (TN / VB)) * VB;
#pragma unroll
            for (int v = 0; v < VB; ++v) Bs[stg][kr][cc + v] = rb[s].v[v];
        }
    };

    acc_t acc[4][NB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nk = (K + kBK - 1) / kBK;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();
    if (nk > 1) load_tiles(kBK);
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            T a[4], bv[NB];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[cur][wr * 64 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < NB; ++j) bv[j] = Bs[cur][kk + (lane >> 4)][wc * 16 * NB + j * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < nbv) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[i][j] = mfma4(a[i], bv[j], acc[i][j]);
                }
            }
        }
        if (kt + 1 < nk) store_tiles(cur ^ 1);  // the other stage was last read before the previous barrier
        if (kt + 2 < nk) load_tiles((kt + 2) * kBK);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= nbv) continue;
            const int col = wcol0 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + wr * 64 + i * 16 + acc_row<T>(lane, r);
                if (row < R) X[row * OPAD + col] = acc[i][j][r];
            }
        }
}
template <typename T, int VA>
__global__ __launch_bounds__(kWideThreads) void kx_wide_kernel(const T* __restrict__ A, int64_t R, int K,
                                                               const T* __restrict__ Cp, int OPAD,
                                                               T* __restrict__ X, int ncolt) {
    kx_wide_body<T, VA>(A, R, K, Cp, OPAD, X, ncolt, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// KS:  Kpart[s][m][n] (+)= sum_{P in split s} sum_{i < width} X[P][m][ioff+i] X[P][n][ioff+i]
// Only lower tiles (mt >= nt) are computed.  grid = nlower * nsplit.
// FP32 inputs accumulate each P in FP32 and fold it into FP64 registers.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ks_kernel_body(const T* __restrict__ X, int N, int OPAD,
                                                       int ioff, int width, int Pc, int nsplit,
                                                       int nlower, double* __restrict__ Kpart,
                                                       int accumulate, int b, int nblocks) {
    typedef typename Acc<T>::type acc_t;
    constexpr int VA = 16 / sizeof(T);
    __shared__ T As[2][kTile][kBK + 1];  // two LDS stages
    __shared__ T Bs[2][kTile][kBK + 1];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int lt = xcd_remap(b, nblocks);
    const int s = lt / nlower;
    int tl = lt % nlower;
    int mt = 0;
    while ((mt + 1) * (mt + 2) / 2 <= tl) ++mt;
    const int nt = tl - mt * (mt + 1) / 2;
    const int m0 = mt * kTile, n0 = nt * kTile;
    const int wr = wid >> 1, wc = wid & 1;
    const int Pa = (int)(((int64_t)Pc * s) / nsplit);
    const int Pb = (int)(((int64_t)Pc * (s + 1)) / nsplit);

    constexpr int LD = (kTile * kBK) / (kGThreads * VA);
    Vec<T, VA> ra[LD], rb[LD];

    auto load_tiles = [&](int P, int i0) {
        const T* XP = X + (int64_t)P * N * OPAD + ioff + i0;
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + kGThreads * q;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
            if (m0 + row < N) {
                ra[q] = load_vec<T, VA>(XP + (int64_t)(m0 + row) * OPAD + kc);
            } else {
#pragma unroll
                for (int v = 0; v < VA; ++v) ra[q].v[v] = T(0);
            }
            if (n0 + row < N) {
                rb[q] = load_vec<T, VA>(XP + (int64_t)(n0 + row) * OPAD + kc);
            } else {
#pragma unroll
                for (int v = 0; v < VA; ++v) rb[q].v[v] = T(0);
            }
        }
    };
    auto store_tiles = [&](int stg) {
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + kGThreads * q;
            const int row = idx / (kBK / VA), kc = (idx % (kBK / VA)) * VA;
#pragma unroll
            for (int v = 0; v < VA; ++v) {
                As[stg][row][kc + v] = ra[q].v[v];
                Bs[stg][row][kc + v] = rb[q].v[v];
            }
        }
    };

    dx4 acc64[4][4];
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc64[i][j] = dx4{0, 0, 0, 0};
            acc[i][j] = acc_t{0, 0, 0, 0};
        }

    const int nkp = width / kBK;  // k steps per P (width % 16 == 0)
    const int nsteps = (Pb - Pa) * nkp;
    auto load_step = [&](int st) { load_tiles(Pa + st / nkp, (st % nkp) * kBK); };
    if (nsteps > 0) {
        load_step(0);
        store_tiles(0);
    }
    __syncthreads();
    if (nsteps > 1) load_step(1);
    for (int st = 0; st < nsteps; ++st) {
        const int cur = st & 1;
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[cur][wr * 64 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[cur][wc * 64 + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma4(a[i], b[j], acc[i][j]);
        }
        if (st + 1 < nsteps) store_tiles(cur ^ 1);
        if (st + 2 < nsteps) load_step(st + 2);
        __syncthreads();
        if constexpr (sizeof(T) == 4) {
            if ((st + 1) % nkp == 0) {  // end of one P: fold FP32 partials into FP64
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc64[i][j][r] += (double)acc[i][j][r];
                        acc[i][j] = acc_t{0, 0, 0, 0};
                    }
            }
        }
    }
    auto outval = [&](int i, int j, int r) -> double {
        if constexpr (sizeof(T) == 8)
            return acc[i][j][r];
        else
            return acc64[i][j][r];
    };
    double* Ks = Kpart + (int64_t)s * N * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wc * 64 + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 64 + i * 16 + acc_row<T>(lane, r);
                if (row >= N) continue;
                double* dst = Ks + (int64_t)row * N + col;
                const double v = outval(i, j, r);
                *dst = accumulate ? (*dst + v) : v;
            }
        }
}
// two workgroups per CU: 4 waves of one workgroup leave each SIMD idle at the
// k-step barrier and the LDS-read latency
template <typename T>
__global__ __launch_bounds__(kGThreads, 2) void ks_kernel(const T* __restrict__ X, int N, int OPAD,
                                                       int ioff, int width, int Pc, int nsplit,
                                                       int nlower, double* __restrict__ Kpart,
                                                       int accumulate) {
    ks_kernel_body<T>(X, N, OPAD, ioff, width, Pc, nsplit, nlower, Kpart, accumulate, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// BF16x3 precision form.  An FP32 value x is carried as the bf16 pair
//   hi = rne_bf16(x),  lo = rne_bf16(x - hi)      (x - hi is exact in FP32)
// which keeps ~16 significant bits; a product is formed from three bf16 MFMAs
//   a*b ~ hi_a*hi_b + hi_a*lo_b + lo_a*hi_b     (lo_a*lo_b ~ 2^-32 relative, dropped)
// with FP32 accumulation.  v_mfma_f32_16x16x32_bf16: lane l supplies
// A[l%16][8*(l/16) + 0..7] and B[8*(l/16) + 0..7][l%16]; accumulator element
// r holds D[4*(l/16) + r][l%16].
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;
// 16-byte staging registers: a native vector (plain loads/stores, no struct copies)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBK16 = 32;              // k extent of one LDS stage = one MFMA k step

// LDS rows of the bf16 tiles are 32 elements (64 B) with the 16-byte chunks
// swizzled: chunk c of row r sits at element (c ^ ((r >> 2) & 2)) * 8.  gfx950
// serves a ds_read_b128 in passes of 32 lanes over 128 banks (512 B); a
// fragment read (lanes 0..31 = rows 0..15, chunks 0..1) would hit rows r and
// r + 8 in the same banks with plain 64 B (or 80 B) rows.  With the swizzle
// both those reads and row-ordered 16-byte stores are conflict-free.
__device__ __forceinline__ int swz16(int row, int chunk) { return (chunk ^ ((row >> 2) & 2)) << 3; }

// f(integral_constant<int, I>) for I = B..E-1: loop bodies whose register
// arrays are indexed by compile-time constants before any loop unrolling, so
// that the arrays are split into registers and never placed in memory.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// A register prefetched in one iteration and consumed in the next is a phi of
// two loads (prologue and loop body); the compiler folds such a phi into one
// load at the loop head, right before the use, which removes the prefetch
// distance.  Passing the prologue's values through an empty asm that claims
// to modify them keeps the phi from being made of loads only (the cost is one
// wait for the prologue loads).
__device__ __forceinline__ void pin(u32x4& v) { asm volatile("" : "+v"(v)); }
template <int VA>
__device__ __forceinline__ void pin(Vec<float, VA>& v) {
#pragma unroll
    for (int i = 0; i < VA; ++i) asm volatile("" : "+v"(v.v[i]));
}

__device__ __forceinline__ fx4 mfma_bf16(bf16x8 a, bf16x8 b, fx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// A double rounded to float and then split.  The asm keeps the compiler from
// merging the two roundings (double -> float -> bf16) into one, which can give
// another hi for a few values: every pad of the coefficients rounds the same way.
__device__ __forceinline__ void split_bf16_from_double(double x, __bf16& hi, __bf16& lo) {
    float f = (float)x;
    asm volatile("" : "+v"(f));
    split_bf16(f, hi, lo);
}

// CpT[c][k] (hi/lo) = Cp[k][c] for c < OPAD, k < N; zero for N <= k < KP.
__global__ void pad_ct_bf16_kernel(const double* __restrict__ C, int64_t ldc, int N, int nocc,
                                   int width, int ioff, int KP, u16* __restrict__ Th,
                                   u16* __restrict__ Tl) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)width * KP) return;
    const int i = (int)(t / KP), k = (int)(t % KP);
    const double x = (i < nocc && k < N) ? C[(int64_t)k * ldc + i] : 0.0;
    __bf16 h, l;
    split_bf16_from_double(x, h, l);
    const int64_t o = (int64_t)(ioff + i) * KP + k;
    Th[o] = __builtin_bit_cast(u16, h);
    Tl[o] = __builtin_bit_cast(u16, l);
}

// KX (bf16x3): X[r][c] = sum_k A[r][k] Ct[c][k]
//   A  = FP32 B chunk, R x K row-major, split into hi/lo while staged to LDS
//   Ct = pre-split transposed coefficients, OPAD x KP (KP = K rounded up to 32)
//   X  = hi/lo bf16 pairs, R x OPAD.
// Tile 128 x (32*NBW): one block holds up to 320 occupied columns so that B is
// read from HBM once.  WAVES waves as 2 x (WAVES/2), each 64 rows x NB blocks
// of 16 columns (NB = 4*NBW/WAVES).  With 8 waves (two per SIMD, ~190
// registers each) the tile's staging is shared by twice the waves.
template <int NBW, int VA, int WAVES = 4>
__device__ __forceinline__ void kx_bf16x3_kernel_body(
    const float* __restrict__ A, int64_t R, int K, int KP, const u16* __restrict__ Cth,
    const u16* __restrict__ Ctl, int OPAD, u16* __restrict__ Xh, u16* __restrict__ Xl, int ncolt, int b, int nblocks) {
    constexpr int BN = 32 * NBW;
    constexpr int T = 64 * WAVES;       // threads
    constexpr int WC = WAVES / 2;       // waves across the columns
    constexpr int NB = 2 * NBW / WC;    // 16-column blocks per wave
    static_assert((2 * NBW) % WC == 0 && NB >= 1, "tile must split over the waves");
    // two LDS stages: the MFMAs read one while the next k-step is written to the other
    __shared__ __bf16 Ah[2][kTile][kBK16];
    __shared__ __bf16 Al[2][kTile][kBK16];
    __shared__ __bf16 Bh[2][BN][kBK16];
    __shared__ __bf16 Bl[2][BN][kBK16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int t = xcd_remap(b, nblocks);
    const int64_t r0 = (int64_t)(t / ncolt) * kTile;
    const int c0 = (t % ncolt) * BN;
    const int wr = wid / WC, wc = wid % WC;
    const int wcol0 = c0 + wc * 16 * NB;
    int nbv = (OPAD - wcol0) / 16;
    nbv = nbv < 0 ? 0 : (nbv > NB ? NB : nbv);

    constexpr int A_LD = (kTile * kBK16) / (T * VA);
    // the B tile (hi and lo arrays, BN rows of four 16-byte chunks each) is
    // staged as NB chunks per thread over both arrays (one per MFMA group); a
    // wave's 64 chunks lie in one array
    constexpr int BCH = BN * (kBK16 / 8);
    constexpr int B_LD = 2 * BCH / T;
    static_assert((2 * BCH) % T == 0 && BCH % 64 == 0 && B_LD == NB, "B tile must split evenly");
    Vec<float, VA> ra[A_LD];
    u32x4 rb[B_LD];

    // The loads carry no branches: rows past R (or columns past OPAD) are
    // clamped onto the last valid row and feed only output rows/columns that
    // are never stored; k past K is clamped onto the row's last VA elements,
    // finite values that meet the zero padding of Ct.
    // parts [S0, S1) of the A tile; the MFMA loop stores and reloads it in two halves
    auto load_a = [&](int k0, auto S0, auto S1) {
        static_for<decltype(S0)::value, decltype(S1)::value>([&](auto S) {
            constexpr int s = decltype(S)::value;
            const int idx = tid + T * s;
            const int row = idx / (kBK16 / VA), kc = (idx % (kBK16 / VA)) * VA;
            const int64_t gr = min(r0 + row, R - 1);
            const int gk = min(k0 + kc, K - VA);
            ra[s] = load_vec<float, VA>(A + gr * K + gk);
        });
    };
    auto load_b = [&](auto q, int k0) {
        const int idx = tid + T * q;
        const int lo = idx >= BCH, rem = idx - (lo ? BCH : 0);
        const int row = rem >> 2, kc = (rem & 3) * 8;
        const int64_t o = (int64_t)min(c0 + row, OPAD - 1) * KP + k0 + kc;
        rb[q] = *reinterpret_cast<const u32x4*>((lo ? Ctl : Cth) + o);
    };
    auto store_a = [&](int stg, auto S0, auto S1) {
        static_for<decltype(S0)::value, decltype(S1)::value>([&](auto S) {
            constexpr int s = decltype(S)::value;
            const int idx = tid + T * s;
            const int row = idx / (kBK16 / VA), kc = (idx % (kBK16 / VA)) * VA;
            const int e = swz16(row, kc >> 3) + (kc & 7);
#pragma unroll
            for (int v = 0; v < VA; ++v) split_bf16(ra[s].v[v], Ah[stg][row][e + v], Al[stg][row][e + v]);
        });
    };
    using A0 = std::integral_constant<int, 0>;
    using AH = std::integral_constant<int, (NB > 2 ? A_LD / 2 : A_LD)>;
    using AE = std::integral_constant<int, A_LD>;
    auto store_b = [&](auto q, int stg) {
        const int idx = tid + T * q;
        const int lo = idx >= BCH, rem = idx - (lo ? BCH : 0);
        const int row = rem >> 2, e = swz16(row, rem & 3);
        *reinterpret_cast<u32x4*>(lo ? &Bl[stg][row][e] : &Bh[stg][row][e]) = rb[q];
    };

    fx4 acc[4][NB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = fx4{0, 0, 0, 0};

    const int nk = KP / kBK16;
    load_a(0, A0{}, AE{});
#pragma unroll
    for (int q = 0; q < B_LD; ++q) load_b(q, 0);
    store_a(0, A0{}, AE{});
#pragma unroll
    for (int q = 0; q < B_LD; ++q) store_b(q, 0);
    __syncthreads();
    {
        const int k1 = (nk > 1 ? 1 : 0) * kBK16;
        load_a(k1, A0{}, AE{});
#pragma unroll
        for (int q = 0; q < B_LD; ++q) load_b(q, k1);
#pragma unroll
        for (int s = 0; s < A_LD; ++s) pin(ra[s]);
#pragma unroll
        for (int q = 0; q < B_LD; ++q) pin(rb[q]);
    }
    const int fr = lane & 15, fc = lane >> 4;  // fragment row / 16-byte chunk of this lane
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1, nxt = cur ^ 1;
        // registers hold step kt+1 (loaded one step ago); they are stored to the
        // other stage and refilled with step kt+2 (a clamped repeat near the end)
        const int kpre = min(kt + 2, nk - 1) * kBK16;
        // fragment reads in the order the first group consumes them (bh, al for
        // the first sweep, then bl, ah), so its first MFMA waits for two reads
        bf16x8 ah[4], al[4], bh[2], bl[2];
        const int rb0 = wc * 16 * NB + fr;
        bh[0] = *reinterpret_cast<const bf16x8*>(&Bh[cur][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = wr * 64 + i * 16 + fr;
            al[i] = *reinterpret_cast<const bf16x8*>(&Al[cur][r][swz16(r, fc)]);
        }
        bl[0] = *reinterpret_cast<const bf16x8*>(&Bl[cur][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = wr * 64 + i * 16 + fr;
            ah[i] = *reinterpret_cast<const bf16x8*>(&Ah[cur][r][swz16(r, fc)]);
        }
        // One group per 16-column block j: the fragments of block j+1 are read
        // while the 12 MFMAs of block j run (three sweeps, so consecutive MFMAs
        // never depend on each other), and the staging work of the next steps
        // sits in fixed groups so that it issues under the MFMAs.  Every block
        // is computed, also those past OPAD (their B rows are clamped copies).
        static_for<0, NB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if constexpr (j + 1 < NB) {
                const int r = wc * 16 * NB + (j + 1) * 16 + fr;
                bh[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Bh[cur][r][swz16(r, fc)]);
                bl[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Bl[cur][r][swz16(r, fc)]);
            }
            static_for<0, 4>([&](auto I) { acc[I][j] = mfma_bf16(al[I], bh[j & 1], acc[I][j]); });
            static_for<0, 4>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bl[j & 1], acc[I][j]); });
            static_for<0, 4>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bh[j & 1], acc[I][j]); });
            // the other stage was last read before the previous barrier
            if constexpr (j == 0) {
                store_a(nxt, A0{}, AH{});
                load_a(kpre, A0{}, AH{});
            }
            if constexpr (j == 2 && AH::value < A_LD) {
                store_a(nxt, AH{}, AE{});
                load_a(kpre, AH{}, AE{});
            }
            // one B chunk per group
            store_b(std::integral_constant<int, j>{}, nxt);
            load_b(std::integral_constant<int, j>{}, kpre);
            __builtin_amdgcn_sched_barrier(0);
        });
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= nbv) continue;
            const int col = wcol0 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + wr * 64 + i * 16 + 4 * (lane >> 4) + r;
                if (row < R) {
                    __bf16 h, l;
                    split_bf16(acc[i][j][r], h, l);
                    Xh[row * OPAD + col] = __builtin_bit_cast(u16, h);
                    Xl[row * OPAD + col] = __builtin_bit_cast(u16, l);
                }
            }
        }
}
// Waves of at most 5 column blocks need at most 256 registers: two waves per
// SIMD (two 4-wave workgroups per CU, or one 8-wave workgroup).
template <int NBW, int WAVES = 4>
constexpr int kx_bf16_waves_per_eu() { return (WAVES == 8 || NBW <= 5) ? 2 : 1; }

template <int NBW, int VA, int WAVES>
__global__ __launch_bounds__(64 * WAVES, (kx_bf16_waves_per_eu<NBW, WAVES>())) void kx_bf16x3_kernel(
    const float* __restrict__ A, int64_t R, int K, int KP, const u16* __restrict__ Cth,
    const u16* __restrict__ Ctl, int OPAD, u16* __restrict__ Xh, u16* __restrict__ Xl, int ncolt) {
    kx_bf16x3_kernel_body<NBW, VA, WAVES>(A, R, K, KP, Cth, Ctl, OPAD, Xh, Xl, ncolt, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// KX (bf16x3), tall tiles: the same product as kx_bf16x3_kernel with 256-row
// tiles, so that each staged coefficient tile feeds twice the rows (the
// coefficient tile, 2.5x the bytes of a 128-row A tile at 320 columns, is
// what the 128-row kernel spends most of its staging on).  Per k-step:
//   * A (FP32 rows of B) is loaded into registers one step ahead, split into
//     bf16 hi/lo and stored to the next LDS stage at the start of the step;
//   * the pre-split coefficient tile (hi and lo planes) goes from global
//     memory straight into the next LDS stage (global_load_lds_dwordx4, no
//     registers): a wave's instruction fills 1 KB = 16 swizzled 64-byte rows,
//     lane l writing slot l, so lane l fetches the logical chunk that the
//     swizzle puts in slot l (the XOR swizzle is its own inverse);
//   * the MFMAs of the current stage run under both.
// The direct-to-LDS loads are issued from inline asm: the compiler does not
// track them, so it never waits for them before unrelated LDS reads (its own
// tracking waits for every pending one before any LDS read).  The kernel
// waits for them itself, with vmcnt(0) at the end of each step (the register
// loads of A are waited for too: a direct-to-LDS load need not complete in
// order with register loads, so no partial count is relied on); the barrier
// then publishes them.  Both kinds of load are issued at the start of the
// step, a full step ahead of their use.
// WR x WC waves; each wave holds a (256/WR) x (32*NBW/WC) block of the tile.
// ---------------------------------------------------------------------------
constexpr int kTallRows = 256;

__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
// 16 bytes per lane from global memory (wave-uniform base + per-lane 32-bit
// byte offset) into LDS at m0 + 16 * lane
__device__ __forceinline__ void lds_load16(const void* base, uint32_t off, uint32_t lds_base) {
    asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(off), "s"(base), "s"(lds_base)
                 : "memory", "m0");
}
#pragma clang diagnostic pop
// Every global load (direct-to-LDS ones included) and LDS access of this wave
// has completed, then the workgroup barrier.  One asm statement with a memory
// clobber: the compiler moves no memory access across it.
__device__ __forceinline__ void drain_and_barrier() {
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
// every global load of this wave has completed
__device__ __forceinline__ void drain_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <int NBW, int WR, int WC>
__device__ __forceinline__ void kx_bf16x3_tall_body(
    const float* __restrict__ A, int64_t R, int K, int KP, const u16* __restrict__ Cth,
    const u16* __restrict__ Ctl, int OPAD, u16* __restrict__ Xh, u16* __restrict__ Xl, int ncolt, int b,
    int nblocks) {
    constexpr int TM = kTallRows;
    constexpr int BN = 32 * NBW;
    constexpr int NW = WR * WC;
    constexpr int T = 64 * NW;
    constexpr int IB = TM / WR / 16;     // 16-row blocks per wave
    constexpr int NB = BN / WC / 16;     // 16-column blocks per wave
    static_assert(TM % (16 * WR) == 0 && BN % (16 * WC) == 0 && NB >= 1, "tile must split over the waves");
    // A: TM x 32 FP32 = TM*8 16-byte chunks, VA = 4 elements per chunk
    constexpr int A_LD = TM * 8 / T;
    static_assert(A_LD * T == TM * 8 && A_LD <= 15, "A tile must split evenly");
    // coefficient tile: 2 planes x BN rows x 4 chunks, in wave instructions of 16 rows
    constexpr int B_INS = 2 * BN / 16;
    constexpr int B_LD = B_INS / NW;
    static_assert(B_LD * NW == B_INS, "coefficient tile must split over the waves");
    __shared__ __attribute__((aligned(16))) __bf16 Ah[2][TM][kBK16];
    __shared__ __attribute__((aligned(16))) __bf16 Al[2][TM][kBK16];
    __shared__ __attribute__((aligned(16))) __bf16 Bt[2][2][BN][kBK16];  // [stage][hi, lo][row][k]

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = xcd_remap(b, nblocks);
    const int64_t r0 = (int64_t)(t / ncolt) * TM;
    const int c0 = (t % ncolt) * BN;
    const int wr = wid / WC, wc = wid % WC;
    const int wcol0 = c0 + wc * 16 * NB;
    int nbv = (OPAD - wcol0) / 16;
    nbv = nbv < 0 ? 0 : (nbv > NB ? NB : nbv);

    Vec<float, 4> ra[A_LD];
    // Rows past R are clamped onto the last row (their outputs are never
    // stored); k past K is clamped onto the row's last 4 elements, finite
    // values that meet the zero padding of Ct.  The tile's base is
    // wave-uniform; each load adds a loop-invariant 32-bit row offset.
    const float* Atile = A + r0 * K;
    uint32_t aoff[A_LD];
    static_for<0, A_LD>([&](auto S) {
        constexpr int s = decltype(S)::value;
        const int row = (tid + T * s) >> 3;
        aoff[s] = (uint32_t)((min(r0 + row, R - 1) - r0) * K);
    });
    auto load_a = [&](int k0) {
        static_for<0, A_LD>([&](auto S) {
            constexpr int s = decltype(S)::value;
            const int kc = ((tid + T * s) & 7) * 4;
            ra[s] = load_vec<float, 4>(Atile + aoff[s] + min(k0 + kc, K - 4));
        });
    };
    auto store_a = [&](int stg) {
        static_for<0, A_LD>([&](auto S) {
            constexpr int s = decltype(S)::value;
            const int idx = tid + T * s;
            const int row = idx >> 3, kc = (idx & 7) * 4;
            const int e = swz16(row, kc >> 3) + (kc & 7);
#pragma unroll
            for (int v = 0; v < 4; ++v) split_bf16(ra[s].v[v], Ah[stg][row][e + v], Al[stg][row][e + v]);
        });
    };
    // This wave's direct-to-LDS loads of the coefficient tile of step k0.
    // Ct has ncolt * BN rows (zero past OPAD), so no row is clamped: a load's
    // address is a wave-uniform base (plane, 16-row group, k0) plus the
    // lane's offset, the same for every load (the swizzle of row 16g + lrow
    // depends on lrow only).
    const int lrow = lane >> 2, lpos = lane & 3;
    const uint32_t boff = (uint32_t)(lrow * KP + (lpos ^ ((lrow >> 2) & 2)) * 8) * 2;
    auto load_b = [&](int k0, int stg) {
        static_for<0, B_LD>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const int g = wid + NW * q;                // wave instruction index (wave-uniform)
            const int plane = g / (BN / 16), grp = g % (BN / 16);
            const u16* base = (plane ? Ctl : Cth) + (int64_t)(c0 + grp * 16) * KP + k0;
            lds_load16(base, boff, lds_addr(&Bt[stg][plane][grp * 16][0]));
        });
    };

    fx4 acc[IB][NB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = fx4{0, 0, 0, 0};

    const int nk = KP / kBK16;
    load_a(0);
    store_a(0);
    load_b(0, 0);
    load_a((nk > 1 ? 1 : 0) * kBK16);
    drain_and_barrier();

    const int fr = lane & 15, fc = lane >> 4;
    const int ra0 = wr * (TM / WR), rb0 = wc * 16 * NB;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1, nxt = cur ^ 1;
        const int k1 = min(kt + 1, nk - 1) * kBK16;  // clamped: the last steps repeat harmless loads
        const int k2 = min(kt + 2, nk - 1) * kBK16;
        bf16x8 bh[NB], bl[NB], ah[2], al[2];
        // row block 0 streams the column fragments in; later row blocks reuse them
        {
            const int r = ra0 + fr;
            al[0] = *reinterpret_cast<const bf16x8*>(&Al[cur][r][swz16(r, fc)]);
            ah[0] = *reinterpret_cast<const bf16x8*>(&Ah[cur][r][swz16(r, fc)]);
        }
        static_for<0, NB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const int r = rb0 + j * 16 + fr;
            bh[j] = *reinterpret_cast<const bf16x8*>(&Bt[cur][0][r][swz16(r, fc)]);
            bl[j] = *reinterpret_cast<const bf16x8*>(&Bt[cur][1][r][swz16(r, fc)]);
        });
        static_for<0, IB>([&](auto I) {
            constexpr int i = decltype(I)::value;
            if constexpr (i + 1 < IB) {
                const int r = ra0 + (i + 1) * 16 + fr;
                al[(i + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Al[cur][r][swz16(r, fc)]);
                ah[(i + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Ah[cur][r][swz16(r, fc)]);
            }
            // three sweeps over the column blocks: consecutive MFMAs never share an accumulator
            static_for<0, NB>([&](auto J) { acc[i][J] = mfma_bf16(al[i & 1], bh[J], acc[i][J]); });
            static_for<0, NB>([&](auto J) { acc[i][J] = mfma_bf16(ah[i & 1], bl[J], acc[i][J]); });
            static_for<0, NB>([&](auto J) { acc[i][J] = mfma_bf16(ah[i & 1], bh[J], acc[i][J]); });
            // staging of the next steps, under the MFMAs of row block 0: the
            // split of A(kt+1) (loaded in the previous step), the register
            // loads of A(kt+2), the LDS loads of Ct(kt+1)
            if constexpr (i == 0) {
                store_a(nxt);
                load_a(k2);
                load_b(k1, nxt);
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        // every load and LDS store of this wave has landed, then publish
        drain_and_barrier();
    }
    // no direct-to-LDS load may outlive the workgroup
    drain_vm();
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= nbv) continue;
            const int col = wcol0 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + ra0 + i * 16 + 4 * (lane >> 4) + r;
                if (row < R) {
                    __bf16 h, l;
                    split_bf16(acc[i][j][r], h, l);
                    Xh[row * OPAD + col] = __builtin_bit_cast(u16, h);
                    Xl[row * OPAD + col] = __builtin_bit_cast(u16, l);
                }
            }
        }
}

template <int NBW, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, (WR * WC == 8 ? 2 : 1)) void kx_bf16x3_tall_kernel(
    const float* __restrict__ A, int64_t R, int K, int KP, const u16* __restrict__ Cth,
    const u16* __restrict__ Ctl, int OPAD, u16* __restrict__ Xh, u16* __restrict__ Xl, int ncolt) {
    kx_bf16x3_tall_body<NBW, WR, WC>(A, R, K, KP, Cth, Ctl, OPAD, Xh, Xl, ncolt, blockIdx.x, gridDim.x);
}

// KS (bf16x3): Kpart[s][m][n] (+)= sum_{P in split s} sum_{i < width} X[P][m][ioff+i] X[P][n][ioff+i]
// with X given as hi/lo bf16 pairs; width % 32 == 0.  FP32 accumulation over
// the P range of a split (the MFMA rounds once per 32 products), FP64 across
// splits.  Lower TS x TS tiles; WR x WC waves, each holding a (TS/WR) x
// (TS/WC) block of the tile.  TS = 128 with 2 x 2 waves fits two workgroups
// per CU (no FP64 registers); TS = 256 with 4 x 2 waves is one 8-wave
// workgroup per CU that reads half the LDS bytes per MFMA (each fragment
// feeds more blocks), for large nbf.  The split length (target_blocks) sets
// the length of the FP32 sums.
template <int TS, int WR, int WC>
__device__ __forceinline__ void ks_bf16x3_kernel_body(
    const u16* __restrict__ Xh, const u16* __restrict__ Xl, int N, int OPAD, int ioff, int width,
    int Pc, int nsplit, int nlower, double* __restrict__ Kpart, int accumulate, int b, int nblocks) {
    constexpr int T = 64 * WR * WC;        // threads
    constexpr int IB = TS / WR / 16;       // 16-row blocks per wave
    constexpr int JB = TS / WC / 16;       // 16-column blocks per wave
    static_assert(JB >= 2, "staging is spread over the first two column blocks");
    __shared__ __bf16 Ah[2][TS][kBK16];
    __shared__ __bf16 Al[2][TS][kBK16];
    __shared__ __bf16 Bh[2][TS][kBK16];
    __shared__ __bf16 Bl[2][TS][kBK16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int lt = xcd_remap(b, nblocks);
    const int s = lt / nlower;
    const int tl = lt % nlower;
    int mt = 0;
    while ((mt + 1) * (mt + 2) / 2 <= tl) ++mt;
    const int nt = tl - mt * (mt + 1) / 2;
    const int m0 = mt * TS, n0 = nt * TS;
    const int wr = wid / WC, wc = wid % WC;
    const int Pa = (int)(((int64_t)Pc * s) / nsplit);
    const int Pb = (int)(((int64_t)Pc * (s + 1)) / nsplit);

    constexpr int LD = (TS * kBK16 / 8) / T;  // 16-byte loads per array (= 2)
    static_assert(LD * T * 8 == TS * kBK16, "tile rows must split evenly over the threads");
    u32x4 rah[LD], ral[LD], rbh[LD], rbl[LD];

    // Branch-free loads: rows past N are clamped onto row N-1 and feed only
    // rows/columns of the tile that are never stored.
    auto load_a = [&](int P, int i0) {
        const int64_t base = (int64_t)P * N * OPAD + ioff + i0;
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + T * q;
            const int row = idx >> 2, kc = (idx & 3) * 8;
            const int64_t o = base + (int64_t)min(m0 + row, N - 1) * OPAD + kc;
            rah[q] = *reinterpret_cast<const u32x4*>(Xh + o);
            ral[q] = *reinterpret_cast<const u32x4*>(Xl + o);
        }
    };
    auto load_b = [&](int P, int i0) {
        const int64_t base = (int64_t)P * N * OPAD + ioff + i0;
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + T * q;
            const int row = idx >> 2, kc = (idx & 3) * 8;
            const int64_t o = base + (int64_t)min(n0 + row, N - 1) * OPAD + kc;
            rbh[q] = *reinterpret_cast<const u32x4*>(Xh + o);
            rbl[q] = *reinterpret_cast<const u32x4*>(Xl + o);
        }
    };
    auto store_a = [&](int stg) {
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + T * q;
            const int row = idx >> 2, e = swz16(row, idx & 3);
            *reinterpret_cast<u32x4*>(&Ah[stg][row][e]) = rah[q];
            *reinterpret_cast<u32x4*>(&Al[stg][row][e]) = ral[q];
        }
    };
    auto store_b = [&](int stg) {
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            const int idx = tid + T * q;
            const int row = idx >> 2, e = swz16(row, idx & 3);
            *reinterpret_cast<u32x4*>(&Bh[stg][row][e]) = rbh[q];
            *reinterpret_cast<u32x4*>(&Bl[stg][row][e]) = rbl[q];
        }
    };

    fx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = fx4{0, 0, 0, 0};

    const int nkp = width / kBK16;
    // steps are numbered across P; step st+1's tiles are written to the other
    // LDS stage while the MFMAs of step st run, and step st+2 is loaded
    const int nsteps = (Pb - Pa) * nkp;
    if (nsteps > 0) {
        load_a(Pa, 0);
        load_b(Pa, 0);
        store_a(0);
        store_b(0);
        const int s1 = nsteps > 1 ? 1 : 0;
        load_a(Pa + s1 / nkp, (s1 % nkp) * kBK16);
        load_b(Pa + s1 / nkp, (s1 % nkp) * kBK16);
#pragma unroll
        for (int q = 0; q < LD; ++q) {
            pin(rah[q]);
            pin(ral[q]);
            pin(rbh[q]);
            pin(rbl[q]);
        }
    }
    __syncthreads();
    const int fr = lane & 15, fc = lane >> 4;  // fragment row / 16-byte chunk of this lane
    const int ra0 = wr * (TS / WR), rb0w = wc * (TS / WC);
    int st = 0;
    for (int P = Pa; P < Pb; ++P) {
    for (int kt = 0; kt < nkp; ++kt, ++st) {
        const int cur = st & 1, nxt = cur ^ 1;
        const int sp = min(st + 2, nsteps - 1);  // clamped prefetch: the last loads are harmless repeats
        const int Ppre = Pa + sp / nkp, ipre = (sp % nkp) * kBK16;
        // fragment reads in the order the first group consumes them
        bf16x8 ah[IB], al[IB], bh[2], bl[2];
        const int rb0 = rb0w + fr;
        bh[0] = *reinterpret_cast<const bf16x8*>(&Bh[cur][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int r = ra0 + i * 16 + fr;
            al[i] = *reinterpret_cast<const bf16x8*>(&Al[cur][r][swz16(r, fc)]);
        }
        bl[0] = *reinterpret_cast<const bf16x8*>(&Bl[cur][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int r = ra0 + i * 16 + fr;
            ah[i] = *reinterpret_cast<const bf16x8*>(&Ah[cur][r][swz16(r, fc)]);
        }
        // one group per 16-column block j (as in the half-transform): block
        // j+1's fragments are read and one staging task is issued under the
        // 3 * IB MFMAs of block j
        static_for<0, JB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if constexpr (j + 1 < JB) {
                const int r = rb0w + (j + 1) * 16 + fr;
                bh[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Bh[cur][r][swz16(r, fc)]);
                bl[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Bl[cur][r][swz16(r, fc)]);
            }
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(al[I], bh[j & 1], acc[I][j]); });
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bl[j & 1], acc[I][j]); });
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bh[j & 1], acc[I][j]); });
            if constexpr (j == 0) {
                store_a(nxt);
                load_a(Ppre, ipre);
            }
            if constexpr (j == 1) {
                store_b(nxt);
                load_b(Ppre, ipre);
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        __syncthreads();
    }
    }
    double* Ks = Kpart + (int64_t)s * N * N;
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int col = n0 + rb0w + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + ra0 + i * 16 + 4 * (lane >> 4) + r;
                if (row >= N) continue;
                double* dst = Ks + (int64_t)row * N + col;
                const double v = (double)acc[i][j][r];
                *dst = accumulate ? (*dst + v) : v;
            }
        }
}
template <int TS, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, (TS == 128 ? 2 : 1)) void ks_bf16x3_kernel(
    const u16* __restrict__ Xh, const u16* __restrict__ Xl, int N, int OPAD, int ioff, int width,
    int Pc, int nsplit, int nlower, double* __restrict__ Kpart, int accumulate) {
    ks_bf16x3_kernel_body<TS, WR, WC>(Xh, Xl, N, OPAD, ioff, width, Pc, nsplit, nlower, Kpart, accumulate,
                                      blockIdx.x, gridDim.x);
}


// KS (bf16x3) with the tiles of X brought into LDS by direct-to-LDS loads:
// the same product, tiles, split ranges and MFMA sequence as
// ks_bf16x3_kernel (so the same Kpart, bitwise), without staging registers
// or LDS stores.  Each step issues the loads of the next step into the other
// stage right after the barrier and waits for them (vmcnt 0: no other global
// loads are in flight in the loop) at the end of the step.  A wave
// instruction fills 16 swizzled 64-byte rows of one plane (A hi, A lo, B hi,
// B lo); rows past N are clamped onto row N-1 (per lane, loop-invariant).
template <int TS, int WR, int WC>
__device__ __forceinline__ void ks_bf16x3_dma_body(
    const u16* __restrict__ Xh, const u16* __restrict__ Xl, int N, int OPAD, int ioff, int width,
    int Pc, int nsplit, int nlower, double* __restrict__ Kpart, int accumulate, int b, int nblocks) {
    constexpr int NW = WR * WC;
    constexpr int IB = TS / WR / 16;
    constexpr int JB = TS / WC / 16;
    constexpr int INS = 4 * TS / 16;   // wave instructions per step: 4 planes x TS/16 row groups
    constexpr int LD = INS / NW;
    static_assert(LD * NW == INS && JB >= 2, "tile must split over the waves");
    __shared__ __attribute__((aligned(16))) __bf16 Xt[2][4][TS][kBK16];  // [stage][Ah, Al, Bh, Bl][row][k]

    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lt = xcd_remap(b, nblocks);
    const int s = lt / nlower;
    const int tl = lt % nlower;
    int mt = 0;
    while ((mt + 1) * (mt + 2) / 2 <= tl) ++mt;
    const int nt = tl - mt * (mt + 1) / 2;
    const int m0 = mt * TS, n0 = nt * TS;
    const int wr = wid / WC, wc = wid % WC;
    const int Pa = (int)(((int64_t)Pc * s) / nsplit);
    const int Pb = (int)(((int64_t)Pc * (s + 1)) / nsplit);

    // per load q: wave-uniform row of its 16-row group (clamped) and the
    // lane's 32-bit byte offset from it (clamped row, swizzled chunk)
    const int lrow = lane >> 2, lpos = lane & 3;
    const int lchunk = lpos ^ ((lrow >> 2) & 2);
    int qrow[LD];
    uint32_t qoff[LD];
    static_for<0, LD>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        const int g = wid + NW * q;
        const int plane = g / (TS / 16), grp = g % (TS / 16);
        const int r0 = (plane < 2 ? m0 : n0) + grp * 16;
        const int rb = min(r0, N - 1);
        qrow[q] = rb;
        qoff[q] = (uint32_t)(((min(r0 + lrow, N - 1) - rb) * OPAD + lchunk * 8) * 2);
    });
    auto load = [&](int P, int i0, int stg) {
        static_for<0, LD>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const int g = wid + NW * q;
            const int plane = g / (TS / 16), grp = g % (TS / 16);
            const u16* base = ((plane & 1) ? Xl : Xh) + ((int64_t)P * N + qrow[q]) * OPAD + ioff + i0;
            lds_load16(base, qoff[q], lds_addr(&Xt[stg][plane][grp * 16][0]));
        });
    };

    fx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = fx4{0, 0, 0, 0};

    const int nkp = width / kBK16;
    const int nsteps = (Pb - Pa) * nkp;
    if (nsteps > 0) load(Pa, 0, 0);
    drain_and_barrier();
    const int fr = lane & 15, fc = lane >> 4;
    const int ra0 = wr * (TS / WR), rb0w = wc * (TS / WC);
    int st = 0;
    for (int P = Pa; P < Pb; ++P) {
    for (int kt = 0; kt < nkp; ++kt, ++st) {
        const int cur = st & 1, nxt = cur ^ 1;
        const int sp = min(st + 1, nsteps - 1);  // clamped: the last loads are harmless repeats
        load(Pa + sp / nkp, (sp % nkp) * kBK16, nxt);
        bf16x8 ah[IB], al[IB], bh[2], bl[2];
        const int rb0 = rb0w + fr;
        bh[0] = *reinterpret_cast<const bf16x8*>(&Xt[cur][2][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int r = ra0 + i * 16 + fr;
            al[i] = *reinterpret_cast<const bf16x8*>(&Xt[cur][1][r][swz16(r, fc)]);
        }
        bl[0] = *reinterpret_cast<const bf16x8*>(&Xt[cur][3][rb0][swz16(rb0, fc)]);
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int r = ra0 + i * 16 + fr;
            ah[i] = *reinterpret_cast<const bf16x8*>(&Xt[cur][0][r][swz16(r, fc)]);
        }
        static_for<0, JB>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if constexpr (j + 1 < JB) {
                const int r = rb0w + (j + 1) * 16 + fr;
                bh[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Xt[cur][2][r][swz16(r, fc)]);
                bl[(j + 1) & 1] = *reinterpret_cast<const bf16x8*>(&Xt[cur][3][r][swz16(r, fc)]);
            }
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(al[I], bh[j & 1], acc[I][j]); });
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bl[j & 1], acc[I][j]); });
            static_for<0, IB>([&](auto I) { acc[I][j] = mfma_bf16(ah[I], bh[j & 1], acc[I][j]); });
            __builtin_amdgcn_sched_barrier(0);
        });
        drain_and_barrier();
    }
    }
    double* Ks = Kpart + (int64_t)s * N * N;
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int col = n0 + rb0w + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + ra0 + i * 16 + 4 * (lane >> 4) + r;
                if (row >= N) continue;
                double* dst = Ks + (int64_t)row * N + col;
                const double v = (double)acc[i][j][r];
                *dst = accumulate ? (*dst + v) : v;
            }
        }
}
template <int TS, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, (TS == 128 ? 2 : 1)) void ks_bf16x3_dma_kernel(
    const u16* __restrict__ Xh, const u16* __restrict__ Xl, int N, int OPAD, int ioff, int width,
    int Pc, int nsplit, int nlower, double* __restrict__ Kpart, int accumulate) {
    ks_bf16x3_dma_body<TS, WR, WC>(Xh, Xl, N, OPAD, ioff, width, Pc, nsplit, nlower, Kpart, accumulate,
                                   blockIdx.x, gridDim.x);
}

// out[r][c] = X[r][c] for r < R, c < ncol: a half-transformed chunk (row
// stride OPAD) copied to a caller buffer with row stride ldx.
__global__ void x_copy_kernel(const double* __restrict__ X, int64_t R, int OPAD, int ncol,
                              double* __restrict__ out, int64_t ldx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= R * ncol) return;
    const int64_t r = t / ncol;
    const int c = (int)(t % ncol);
    out[r * ldx + c] = X[r * OPAD + c];
}

// K[m][n] = K[n][m] = sum_s Kpart[s][m][n] for m >= n (fixed order in s).
__global__ void k_reduce_kernel(const double* __restrict__ Kpart, int nsplit, int N,
                                double* __restrict__ K) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (n > m || m >= N) return;
    const int64_t NN = (int64_t)N * N;
    double acc = 0.0;
    for (int s = 0; s < nsplit; ++s) acc += Kpart[s * NN + (int64_t)m * N + n];
    K[(int64_t)m * N + n] = acc;
    K[(int64_t)n * N + m] = acc;
}

// Cp[n][ioff + i] = C[n][i] (i < nocc), zero for nocc <= i < width.
template <typename T>
__global__ void pad_c_kernel(const double* __restrict__ C, int64_t ldc, int N, int nocc,
                             int width, int ioff, int OPAD, T* __restrict__ Cp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * width) return;
    const int n = (int)(t / width), i = (int)(t % width);
    Cp[(int64_t)n * OPAD + ioff + i] = (i < nocc) ? (T)C[(int64_t)n * ldc + i] : T(0);
}

}  // namespace f2jk
