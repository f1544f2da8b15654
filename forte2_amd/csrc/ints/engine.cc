// Host integral engine: McMurchie-Davidson Hermite expansions, solid-harmonic
// shells, OpenMP over shell pairs.  ABI in include/f2ints.h.
//
// Method (T. Helgaker, P. Jorgensen, J. Olsen, "Molecular Electronic-Structure
// Theory", ch. 9):
//   * a Cartesian Gaussian pair x_A^i x_B^j e^{-a x_A^2 - b x_B^2} is expanded in
//     Hermite Gaussians Lambda_t(p, P) with coefficients E^{ij}_t (eq. 9.5.6-7);
//   * Coulomb-type integrals over Hermite Gaussians reduce to the auxiliary
//     integrals R_{tuv}(alpha, R_PQ) built from the Boys function (9.9.18-20);
//   * a single-centre solid-harmonic Gaussian S_lm(r_C) e^{-g r_C^2} equals
//     (2g)^{-l} S_lm(d/dC) e^{-g r_C^2} exactly (Hobson's theorem, S_lm is
//     harmonic), so auxiliary shells contribute only order-l Hermite terms;
//   * the Boys function F_n(T) is tabulated on a grid (series, then downward
//     recursion), evaluated by a 7-term Taylor expansion at the highest order
//     and downward recursion below it; for T > 36 the asymptotic F_0 and
//     upward recursion are exact to double precision (9.8.2).
#include <omp.h>

#include <mutex>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "f2ints.h"

namespace {

constexpr int LMAX = 6;                  // up to i shells
constexpr int NBOYS = 4 * LMAX + 1;      // highest Boys order needed by (ab|cd)
constexpr double PI = 3.14159265358979323846;
constexpr double kPrimScreen = 1.0e-15;  // primitive pairs below this prefactor are skipped

thread_local std::string g_err;
int g_threads = 0;

int fail(const std::string& m) {
    g_err = m;
    return -1;
}

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int nsph(int l) { return 2 * l + 1; }
inline int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// ---------------------------------------------------------------- Boys function
struct BoysTable {
    static constexpr double h = 0.05;
    static constexpr double Tswitch = 36.0;
    static constexpr int NT = NBOYS + 8;  // orders kept per grid point
    int npts;
    std::vector<double> tab;
    BoysTable() {
        npts = (int)(Tswitch / h) + 3;
        tab.resize((size_t)npts * NT);
        for (int i = 0; i < npts; ++i) {
            const long double T = (long double)i * h;
            // F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
            const int n = NT - 1;
            long double term = 1.0L / (2 * n + 1), sum = term;
            for (int k = 1; k < 2000; ++k) {
                term *= 2.0L * T / (2 * n + 2 * k + 1);
                sum += term;
                if (term < 1e-22L * sum) break;
            }
            const long double eT = expl(-T);
            long double F = eT * sum;
            tab[(size_t)i * NT + n] = (double)F;
            for (int m = n - 1; m >= 0; --m) {
                F = (2.0L * T * F + eT) / (2 * m + 1);
                tab[(size_t)i * NT + m] = (double)F;
            }
        }
    }
    void eval(int nmax, double T, double* F) const {
        if (T < Tswitch) {
            const int i = (int)(T / h + 0.5);
            const double d = T - i * h;
            const double* t = &tab[(size_t)i * NT];
            // F_nmax(T) = sum_k F_{nmax+k}(T_i) (-d)^k / k!
            double f = 0.0, c = 1.0;
            for (int k = 0; k <= 6; ++k) {
                f += t[nmax + k] * c;
                c *= -d / (k + 1);
            }
            F[nmax] = f;
            const double eT = std::exp(-T);
            for (int m = nmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
        } else {
            const double eT = std::exp(-T);
            F[0] = 0.5 * std::sqrt(PI / T);
            for (int m = 0; m < nmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0 * T);
        }
    }
    // eval for K arguments at once, F[lane][n]: each lane the arithmetic of
    // eval in the same order (bit for bit its values); when every argument
    // is on the grid the Taylor sums and the downward recursions run across
    // the lanes
    template <int K, int NF>
    void eval_lanes(int nmax, const double* T, double (*F)[NF]) const {
        for (int l = 0; l < K; ++l)
            if (!(T[l] < Tswitch)) {
                for (int q = 0; q < K; ++q) eval(nmax, T[q], F[q]);
                return;
            }
        const double* t[K];
        double d[K], f[K], c[K], eT[K];
        for (int l = 0; l < K; ++l) {
            const int i = (int)(T[l] / h + 0.5);
            d[l] = T[l] - i * h;
            t[l] = &tab[(size_t)i * NT];
            f[l] = 0.0;
            c[l] = 1.0;
        }
        for (int k = 0; k <= 6; ++k)
            for (int l = 0; l < K; ++l) {
                f[l] += t[l][nmax + k] * c[l];
                c[l] *= -d[l] / (k + 1);
            }
        for (int l = 0; l < K; ++l) {
            F[l][nmax] = f[l];
            eT[l] = std::exp(-T[l]);
        }
        for (int m = nmax - 1; m >= 0; --m)
            for (int l = 0; l < K; ++l) F[l][m] = (2.0 * T[l] * F[l][m + 1] + eT[l]) / (2 * m + 1);
    }
};
const BoysTable& boys() {
    static BoysTable t;
    return t;
}

// ---------------------------------------------------------------- Cartesian / solid harmonics
struct CartComp {
    int x, y, z;
};
struct SphTerm {
    int cart;
    double c;
};
struct Harmonics {
    std::vector<std::vector<CartComp>> cart;                 // [l][ncart]
    std::vector<std::vector<double>> T;                      // [l][(2l+1) * ncart]
    std::vector<std::vector<std::vector<SphTerm>>> sparse;   // [l][m] -> terms
    Harmonics() {
        cart.resize(LMAX + 1);
        T.resize(LMAX + 1);
        sparse.resize(LMAX + 1);
        for (int l = 0; l <= LMAX; ++l) {
            for (int a = l; a >= 0; --a)
                for (int b = l - a; b >= 0; --b) cart[l].push_back({a, b, l - a - b});
            T[l].assign((size_t)nsph(l) * ncart(l), 0.0);
            for (int m = -l; m <= l; ++m) build(l, m);
            sparse[l].resize(nsph(l));
            for (int s = 0; s < nsph(l); ++s)
                for (int c = 0; c < ncart(l); ++c) {
                    const double v = T[l][(size_t)s * ncart(l) + c];
                    if (v != 0.0) sparse[l][s].push_back({c, v});
                }
        }
    }
    static double fact(int n) {
        double f = 1;
        for (int i = 2; i <= n; ++i) f *= i;
        return f;
    }
    static double binom(double n, double k) {
        if (k < 0 || k > n) return 0.0;
        // integer arguments only (k may be passed as an exact integer double)
        return fact((int)n) / (fact((int)k) * fact((int)(n - k)));
    }
    int cart_index(int l, int x, int y, int z) const {
        for (int c = 0; c < ncart(l); ++c)
            if (cart[l][c].x == x && cart[l][c].y == y && cart[l][c].z == z) return c;
        return -1;
    }
    // real solid harmonic S_lm (Helgaker eq. 6.4.47-6.4.50), phases and norms of
    // forte2/integrals/real_spherical_harmonics.hpp (same norm as z^l)
    void build(int l, int m) {
        const int am = std::abs(m);
        const double N = 1.0 / (std::pow(2.0, am) * fact(l)) *
                         std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
        const int half = (m < 0) ? 1 : 0;  // v = half/2 + k
        for (int t = 0; t <= (l - am) / 2; ++t)
            for (int u = 0; u <= t; ++u) {
                const int kmax = (m < 0) ? (am - 1) / 2 : am / 2;
                for (int k = 0; k <= kmax; ++k) {
                    const int twov = 2 * k + half;  // 2v
                    const double sign = ((t + k) % 2) ? -1.0 : 1.0;
                    const double C = sign * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                                     binom(t, u) * binom(am, twov);
                    if (C == 0.0) continue;
                    // x^{2t+|m|-2(u+v)} y^{2(u+v)} z^{l-2t-|m|}
                    const int py = 2 * u + twov;
                    const int px = 2 * t + am - py;
                    const int pz = l - 2 * t - am;
                    const int c = cart_index(l, px, py, pz);
                    T[l][(size_t)(m + l) * ncart(l) + c] += N * C;
                }
            }
    }
};
const Harmonics& harm() {
    static Harmonics h;
    return h;
}

// ---------------------------------------------------------------- Hermite index tables
struct HermIndex {
    // idx[L][t][u][v] dense (-1 where t+u+v > L) and lists of (t,u,v) per L
    std::vector<std::vector<std::array<int, 3>>> list;  // list[L]
    int LM;
    std::vector<int> dense;  // [(t*(LM+1)+u)*(LM+1)+v] -> position in list[LM] (order is nested)
    HermIndex() {
        LM = 4 * LMAX + 2;
        list.resize(LM + 1);
        for (int L = 0; L <= LM; ++L)
            for (int N = 0; N <= L; ++N)
                for (int t = N; t >= 0; --t)
                    for (int u = N - t; u >= 0; --u) list[L].push_back({t, u, N - t - u});
        // list[L] is a prefix of list[L+1]: enumeration by order N
        dense.assign((size_t)(LM + 1) * (LM + 1) * (LM + 1), -1);
        for (size_t i = 0; i < list[LM].size(); ++i) {
            const auto& q = list[LM][i];
            dense[((size_t)q[0] * (LM + 1) + q[1]) * (LM + 1) + q[2]] = (int)i;
        }
    }
    inline int at(int t, int u, int v) const {
        return dense[((size_t)t * (LM + 1) + u) * (LM + 1) + v];
    }
};
const HermIndex& hidx() {
    static HermIndex h;
    return h;
}

// One step of the Hermite recursion R^n_{tuv} = P_d R^{n+1}_{t-1..} + (k-1) R^{n+1}_{t-2..}
// along the first non-zero direction d of (t, u, v), with its source positions
// in hidx().list (b < 0: no second term).  The order is that of list[LM], so
// the steps of any L are a prefix.
struct RStep {
    int N, dir, a, b;
    double k1;  // the (k - 1) factor of the second term
};
const std::vector<RStep>& r_steps() {
    static const std::vector<RStep> steps = [] {
        const HermIndex& H = hidx();
        std::vector<RStep> out;
        const auto& lst = H.list[H.LM];
        for (size_t i = 0; i < lst.size(); ++i) {
            const int t = lst[i][0], u = lst[i][1], v = lst[i][2];
            RStep st{t + u + v, 0, -1, -1, 0.0};
            if (i == 0) {
                out.push_back(st);
                continue;
            }
            if (t > 0) {
                st.dir = 0;
                st.a = H.at(t - 1, u, v);
                if (t > 1) st.b = H.at(t - 2, u, v), st.k1 = t - 1;
            } else if (u > 0) {
                st.dir = 1;
                st.a = H.at(t, u - 1, v);
                if (u > 1) st.b = H.at(t, u - 2, v), st.k1 = u - 1;
            } else {
                st.dir = 2;
                st.a = H.at(t, u, v - 1);
                if (v > 1) st.b = H.at(t, u, v - 2), st.k1 = v - 1;
            }
            out.push_back(st);
        }
        return out;
    }();
    return steps;
}

// R^0_{tuv}(alpha, X, Y, Z) for t+u+v <= L, stored by position in hidx().list[L].
// work must hold (L+1) * nherm(L) doubles.
void compute_R(int L, double alpha, double X, double Y, double Z, double* R0, double* work) {
    const int nh = nherm(L);
    double F[NBOYS + 2];
    boys().eval(L, alpha * std::fma(Z, Z, std::fma(Y, Y, X * X)), F);
    // work[n * nh + idx]
    double m2a = 1.0;
    for (int n = 0; n <= L; ++n) {
        work[(size_t)n * nh + 0] = m2a * F[n];
        m2a *= -2.0 * alpha;
    }
    const RStep* steps = r_steps().data();
    const double P[3] = {X, Y, Z};
    for (int i = 1; i < nh; ++i) {
        const RStep& st = steps[i];
        const double Pd = P[st.dir];
        for (int n = 0; n <= L - st.N; ++n) {
            const double* Rn1 = work + (size_t)(n + 1) * nh;
            const double val = Pd * Rn1[st.a];
            work[(size_t)n * nh + i] = st.b >= 0 ? std::fma(st.k1, Rn1[st.b], val) : val;
        }
    }
    std::memcpy(R0, work, sizeof(double) * nh);
}

// compute_R for kLanes independent (alpha, X, Y, Z): R4[idx * kLanes + lane],
// work4 of (L + 1) * nherm(L) * kLanes.  Every lane does the arithmetic of
// compute_R in the same order, so each lane's values are those of a call of
// compute_R; the lane loops vectorise.
constexpr int kLanes = 4;

void compute_R4(int L, const double* alpha, const double* X, const double* Y, const double* Z, double* R4,
                double* work4) {
    const int nh = nherm(L);
    double F[kLanes][NBOYS + 2], T[kLanes];
    for (int l = 0; l < kLanes; ++l) T[l] = alpha[l] * std::fma(Z[l], Z[l], std::fma(Y[l], Y[l], X[l] * X[l]));
    boys().eval_lanes<kLanes, NBOYS + 2>(L, T, F);
    double m2a[kLanes];
    for (int l = 0; l < kLanes; ++l) m2a[l] = 1.0;
    for (int n = 0; n <= L; ++n)
        for (int l = 0; l < kLanes; ++l) {
            work4[((size_t)n * nh + 0) * kLanes + l] = m2a[l] * F[l][n];
            m2a[l] *= -2.0 * alpha[l];
        }
    const RStep* steps = r_steps().data();
    const double* P[3] = {X, Y, Z};
    for (int i = 1; i < nh; ++i) {
        const RStep& st = steps[i];
        const double* Pd = P[st.dir];
        for (int n = 0; n <= L - st.N; ++n) {
            const double* Rn1 = work4 + (size_t)(n + 1) * nh * kLanes;
            double* out = work4 + ((size_t)n * nh + i) * kLanes;
            const double* ra = Rn1 + (size_t)st.a * kLanes;
            if (st.b >= 0) {
                const double* rb = Rn1 + (size_t)st.b * kLanes;
                for (int l = 0; l < kLanes; ++l) out[l] = std::fma(st.k1, rb[l], Pd[l] * ra[l]);
            } else {
                for (int l = 0; l < kLanes; ++l) out[l] = Pd[l] * ra[l];
            }
        }
    }
    std::memcpy(R4, work4, sizeof(double) * nh * kLanes);
}

// ---------------------------------------------------------------- basis
struct Shell {
    int l;
    double A[3];
    std::vector<double> alpha, coef;
    int64_t off;
};

}  // namespace

struct ShellPairCacheEntry;

struct f2ints_basis {
    std::vector<Shell> shells;
    int64_t nbf = 0;
    int maxl = 0;
    // Significant shell pairs (a >= b) and the ket-form Hermite data of their
    // shell families (consecutive shells of one centre, l and exponent set:
    // the columns of a general contraction), built once on the first
    // 3-centre call and reused by every later block
    std::once_flag pairs_once;
    std::vector<std::pair<int, int>> pair_index;
    std::vector<void*> fam_pairs;   // FamPair*, owned
    std::vector<int> fam_of_pair;   // family pair of each significant shell pair
    std::vector<int> col_of_pair;   // its column pair within that family pair
    ~f2ints_basis();
};

namespace {

// 1-D Hermite coefficients E^{ij}_t, i <= imax, j <= jmax
struct E1D {
    static constexpr int D = LMAX + 3;
    double e[D][D][2 * D];
    void compute(int imax, int jmax, double a, double b, double Ax, double Bx) {
        const double p = a + b, mu = a * b / p, P = (a * Ax + b * Bx) / p;
        const double XPA = P - Ax, XPB = P - Bx, o2p = 0.5 / p;
        for (int i = 0; i <= imax; ++i)
            for (int j = 0; j <= jmax; ++j)
                for (int t = 0; t <= i + j + 1; ++t) e[i][j][t] = 0.0;
        e[0][0][0] = std::exp(-mu * (Ax - Bx) * (Ax - Bx));
        for (int i = 0; i < imax; ++i)
            for (int t = 0; t <= i + 1; ++t) {
                double v = XPA * e[i][0][t];
                if (t > 0) v += o2p * e[i][0][t - 1];
                if (t + 1 <= i) v += (t + 1) * e[i][0][t + 1];
                e[i + 1][0][t] = v;
            }
        for (int i = 0; i <= imax; ++i)
            for (int j = 0; j < jmax; ++j)
                for (int t = 0; t <= i + j + 1; ++t) {
                    double v = XPB * e[i][j][t];
                    if (t > 0) v += o2p * e[i][j][t - 1];
                    if (t + 1 <= i + j) v += (t + 1) * e[i][j][t + 1];
                    e[i][j + 1][t] = v;
                }
    }
};

// Hermite expansion of a primitive pair in the solid-harmonic product basis:
//   E[(ma*nsb + mb) * nh + h]  (h over hidx().list[la+lb]),
// with coefficient product and the Hermite sign (-1)^{t+u+v} when `ket`.
struct PrimPair {
    double p, P[3], scale;  // scale: ca*cb (exp factor is inside E)
    std::vector<double> E;
};

struct ShellPair {
    int a, b;
    std::vector<PrimPair> prims;
};

// Spherical Hermite expansion E[sa * nsb + sb][h] of the primitive pair
// (a on A, b on B) without contraction coefficients; ket form flips odd h.
struct PrimE {
    E1D ex, ey, ez;
    std::vector<double> ecart, etmp;
    void compute(const Shell& A, const Shell& B, double a, double b, bool ket, std::vector<double>& E);
};

void make_pair(const Shell& A, const Shell& B, bool ket, ShellPair& sp) {
    const double AB2 = (A.A[0] - B.A[0]) * (A.A[0] - B.A[0]) + (A.A[1] - B.A[1]) * (A.A[1] - B.A[1]) +
                       (A.A[2] - B.A[2]) * (A.A[2] - B.A[2]);
    PrimE pe;
    sp.prims.clear();
    for (size_t ia = 0; ia < A.alpha.size(); ++ia)
        for (size_t ib = 0; ib < B.alpha.size(); ++ib) {
            const double a = A.alpha[ia], b = B.alpha[ib], p = a + b;
            const double cc = A.coef[ia] * B.coef[ib];
            const double K = std::exp(-a * b / p * AB2);
            if (std::fabs(cc) * K * std::pow(PI / p, 1.5) < kPrimScreen) continue;
            PrimPair pp;
            pp.p = p;
            for (int d = 0; d < 3; ++d) pp.P[d] = (a * A.A[d] + b * B.A[d]) / p;
            pp.scale = cc;
            pe.compute(A, B, a, b, ket, pp.E);
            sp.prims.push_back(std::move(pp));
        }
}

void PrimE::compute(const Shell& A, const Shell& B, double a, double b, bool ket, std::vector<double>& E) {
    const Harmonics& Hm = harm();
    const HermIndex& H = hidx();
    const int la = A.l, lb = B.l, L = la + lb, nh = nherm(L);
    const int nca = ncart(la), ncb = ncart(lb), nsa = nsph(la), nsb = nsph(lb);
    ecart.resize((size_t)nca * ncb * nh);
    etmp.resize((size_t)nsa * ncb * nh);
    ex.compute(la, lb, a, b, A.A[0], B.A[0]);
    ey.compute(la, lb, a, b, A.A[1], B.A[1]);
    ez.compute(la, lb, a, b, A.A[2], B.A[2]);
    // Cartesian products
    for (int ca = 0; ca < nca; ++ca)
        for (int cb = 0; cb < ncb; ++cb) {
            const CartComp& qa = Hm.cart[la][ca];
            const CartComp& qb = Hm.cart[lb][cb];
            double* dst = &ecart[((size_t)ca * ncb + cb) * nh];
            for (int h = 0; h < nh; ++h) {
                const auto& tuv = H.list[L][h];
                if (tuv[0] > qa.x + qb.x || tuv[1] > qa.y + qb.y || tuv[2] > qa.z + qb.z) {
                    dst[h] = 0.0;
                    continue;
                }
                double v = ex.e[qa.x][qb.x][tuv[0]] * ey.e[qa.y][qb.y][tuv[1]] *
                           ez.e[qa.z][qb.z][tuv[2]];
                if (ket && ((tuv[0] + tuv[1] + tuv[2]) & 1)) v = -v;
                dst[h] = v;
            }
        }
    // transform a, then b
    std::fill(etmp.begin(), etmp.end(), 0.0);
    for (int sa = 0; sa < nsa; ++sa)
        for (const SphTerm& ta : Hm.sparse[la][sa])
            for (int cb = 0; cb < ncb; ++cb) {
                const double* src = &ecart[((size_t)ta.cart * ncb + cb) * nh];
                double* dst = &etmp[((size_t)sa * ncb + cb) * nh];
                for (int h = 0; h < nh; ++h) dst[h] += ta.c * src[h];
            }
    E.assign((size_t)nsa * nsb * nh, 0.0);
    for (int sa = 0; sa < nsa; ++sa)
        for (int sb = 0; sb < nsb; ++sb)
            for (const SphTerm& tb : Hm.sparse[lb][sb]) {
                const double* src = &etmp[((size_t)sa * ncb + tb.cart) * nh];
                double* dst = &E[((size_t)sa * nsb + sb) * nh];
                for (int h = 0; h < nh; ++h) dst[h] += tb.c * src[h];
            }
}

int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

// ---------------------------------------------------------------- one-electron
enum OneE { kOverlap, kKinetic };

int one_electron(const f2ints_basis* bs, OneE kind, double* out) {
    const Harmonics& Hm = harm();
    const int ns = (int)bs->shells.size();
    const int64_t nbf = bs->nbf;
    std::vector<std::pair<int, int>> pairs;
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b <= a; ++b) pairs.push_back({a, b});
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
    for (size_t q = 0; q < pairs.size(); ++q) {
        const Shell& A = bs->shells[pairs[q].first];
        const Shell& B = bs->shells[pairs[q].second];
        const int la = A.l, lb = B.l, nca = ncart(la), ncb = ncart(lb);
        std::vector<double> cartv((size_t)nca * ncb, 0.0);
        E1D ex, ey, ez;
        for (size_t ia = 0; ia < A.alpha.size(); ++ia)
            for (size_t ib = 0; ib < B.alpha.size(); ++ib) {
                const double a = A.alpha[ia], b = B.alpha[ib], p = a + b;
                const double cc = A.coef[ia] * B.coef[ib] * std::pow(PI / p, 1.5);
                const int jx = (kind == kKinetic) ? lb + 2 : lb;
                ex.compute(la, jx, a, b, A.A[0], B.A[0]);
                ey.compute(la, jx, a, b, A.A[1], B.A[1]);
                ez.compute(la, jx, a, b, A.A[2], B.A[2]);
                for (int ca = 0; ca < nca; ++ca)
                    for (int cb = 0; cb < ncb; ++cb) {
                        const CartComp& qa = Hm.cart[la][ca];
                        const CartComp& qb = Hm.cart[lb][cb];
                        const double sx = ex.e[qa.x][qb.x][0], sy = ey.e[qa.y][qb.y][0],
                                     sz = ez.e[qa.z][qb.z][0];
                        double v;
                        if (kind == kOverlap) {
                            v = sx * sy * sz;
                        } else {
                            // -1/2 d^2/dx^2 acting on x^j e^{-b x^2}
                            auto t1 = [&](const E1D& e, int i, int j) {
                                double r = -2.0 * b * b * e.e[i][j + 2][0] + b * (2 * j + 1) * e.e[i][j][0];
                                if (j >= 2) r -= 0.5 * j * (j - 1) * e.e[i][j - 2][0];
                                return r;
                            };
                            v = t1(ex, qa.x, qb.x) * sy * sz + sx * t1(ey, qa.y, qb.y) * sz +
                                sx * sy * t1(ez, qa.z, qb.z);
                        }
                        cartv[(size_t)ca * ncb + cb] += cc * v;
                    }
            }
        const int nsa = nsph(la), nsb = nsph(lb);
        for (int sa = 0; sa < nsa; ++sa)
            for (int sb = 0; sb < nsb; ++sb) {
                double v = 0.0;
                for (const SphTerm& ta : Hm.sparse[la][sa])
                    for (const SphTerm& tb : Hm.sparse[lb][sb]) v += ta.c * tb.c * cartv[(size_t)ta.cart * ncb + tb.cart];
                const int64_t m = A.off + sa, n = B.off + sb;
                out[m * nbf + n] = v;
                out[n * nbf + m] = v;
            }
    }
    return 0;
}

int nuclear(const f2ints_basis* bs, int nc, const double* Z, const double* xyz, double* out) {
    const HermIndex& H = hidx();
    const int ns = (int)bs->shells.size();
    const int64_t nbf = bs->nbf;
    std::vector<std::pair<int, int>> pairs;
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b <= a; ++b) pairs.push_back({a, b});
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
    for (size_t q = 0; q < pairs.size(); ++q) {
        const Shell& A = bs->shells[pairs[q].first];
        const Shell& B = bs->shells[pairs[q].second];
        ShellPair sp;
        make_pair(A, B, false, sp);
        const int L = A.l + B.l, nh = nherm(L), nsa = nsph(A.l), nsb = nsph(B.l);
        std::vector<double> R(nh), work((size_t)(L + 1) * nh), acc((size_t)nsa * nsb, 0.0);
        for (const PrimPair& pp : sp.prims) {
            for (int c = 0; c < nc; ++c) {
                const double X = pp.P[0] - xyz[3 * c], Y = pp.P[1] - xyz[3 * c + 1],
                             Zc = pp.P[2] - xyz[3 * c + 2];
                compute_R(L, pp.p, X, Y, Zc, R.data(), work.data());
                const double pref = -Z[c] * 2.0 * PI / pp.p * pp.scale;
                for (int s = 0; s < nsa * nsb; ++s) {
                    const double* e = &pp.E[(size_t)s * nh];
                    double v = 0.0;
                    for (int h = 0; h < nh; ++h) v += e[h] * R[h];
                    acc[s] += pref * v;
                }
            }
        }
        (void)H;
        for (int sa = 0; sa < nsa; ++sa)
            for (int sb = 0; sb < nsb; ++sb) {
                const int64_t m = A.off + sa, n = B.off + sb;
                out[m * nbf + n] = acc[(size_t)sa * nsb + sb];
                out[n * nbf + m] = acc[(size_t)sa * nsb + sb];
            }
    }
    return 0;
}

// ---------------------------------------------------------------- 2-centre Coulomb
int coulomb_2c(const f2ints_basis* a, const f2ints_basis* b, double* out) {
    const Harmonics& Hm = harm();
    const HermIndex& H = hidx();
    const bool same = (a == b);
    const int na = (int)a->shells.size(), nb = (int)b->shells.size();
    const int64_t nbfb = b->nbf;
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < (same ? i + 1 : nb); ++j) pairs.push_back({i, j});
    int lmax = 0;
    for (const Shell& s : a->shells) lmax = std::max(lmax, s.l);
    for (const Shell& s : b->shells) lmax = std::max(lmax, s.l);
#pragma omp parallel num_threads(nthreads())
    {
    // per-thread buffers, sized once for the largest pair
    std::vector<double> R(nherm(2 * lmax)), work((size_t)(2 * lmax + 1) * nherm(2 * lmax)),
        acc((size_t)nsph(lmax) * nsph(lmax));
#pragma omp for schedule(dynamic)
    for (size_t q = 0; q < pairs.size(); ++q) {
        const Shell& P = a->shells[pairs[q].first];
        const Shell& Q = b->shells[pairs[q].second];
        const int L = P.l + Q.l, nsP = nsph(P.l), nsQ = nsph(Q.l);
        std::fill(acc.begin(), acc.begin() + (size_t)nsP * nsQ, 0.0);
        const double X = P.A[0] - Q.A[0], Y = P.A[1] - Q.A[1], Z = P.A[2] - Q.A[2];
        const double sgnQ = (Q.l & 1) ? -1.0 : 1.0;
        for (size_t ip = 0; ip < P.alpha.size(); ++ip)
            for (size_t iq = 0; iq < Q.alpha.size(); ++iq) {
                const double g = P.alpha[ip], d = Q.alpha[iq];
                const double al = g * d / (g + d);
                compute_R(L, al, X, Y, Z, R.data(), work.data());
                const double pref = 2.0 * std::pow(PI, 2.5) / (g * d * std::sqrt(g + d)) * P.coef[ip] *
                                    Q.coef[iq] * std::pow(2.0 * g, -P.l) * std::pow(2.0 * d, -Q.l) * sgnQ;
                for (int sP = 0; sP < nsP; ++sP)
                    for (int sQ = 0; sQ < nsQ; ++sQ) {
                        double v = 0.0;
                        for (const SphTerm& tp : Hm.sparse[P.l][sP]) {
                            const CartComp& cp = Hm.cart[P.l][tp.cart];
                            for (const SphTerm& tq : Hm.sparse[Q.l][sQ]) {
                                const CartComp& cq = Hm.cart[Q.l][tq.cart];
                                v += tp.c * tq.c * R[H.at(cp.x + cq.x, cp.y + cq.y, cp.z + cq.z)];
                            }
                        }
                        acc[(size_t)sP * nsQ + sQ] += pref * v;
                    }
            }
        for (int sP = 0; sP < nsP; ++sP)
            for (int sQ = 0; sQ < nsQ; ++sQ) {
                const int64_t m = P.off + sP, n = Q.off + sQ;
                out[m * nbfb + n] = acc[(size_t)sP * nsQ + sQ];
                if (same) out[n * nbfb + m] = acc[(size_t)sP * nsQ + sQ];
            }
    }
    }
    return 0;
}

// ---------------------------------------------------------------- 3-centre Coulomb
// A shell family is a run of consecutive shells with one centre, l and
// exponent set: the coefficient columns of a general contraction, which the
// basis builder splits into one shell per column.  A family pair shares its
// primitive pairs (and their Hermite expansions, Boys values and aux
// contractions) between all its column pairs; each column pair only weights
// them with its own coefficient products.
struct FamPrim {
    double p, P[3];
    std::vector<double> E;  // [nh][nsa * nsb] ket form, no contraction coefficients
    std::vector<double> w;  // per column pair: ca_i * cb_j, 0 where that pair screens the primitive out
};

struct FamPair {
    int a0, na, b0, nb;                     // shell ranges of the two families
    std::vector<std::pair<int, int>> cols;  // column pairs (ca, cb): shell a0 + ca >= b0 + cb
    std::vector<int64_t> q;                 // significant shell-pair index of each column pair, -1 if none
    std::vector<FamPrim> prims;
};

// significant shell pairs (a >= b) of `bs`, ordered by a then b, with the
// ket-form Hermite data of their family pairs
const std::vector<std::pair<int, int>>& ket_pairs(const f2ints_basis* cbs) {
    f2ints_basis* bs = const_cast<f2ints_basis*>(cbs);
    std::call_once(bs->pairs_once, [bs]() {
        const int ns = (int)bs->shells.size();
        std::vector<int> f0;  // first shell of each family
        for (int s = 0; s < ns; ++s) {
            const Shell& A = bs->shells[s];
            bool same = s > 0;
            if (same) {
                const Shell& Z = bs->shells[s - 1];
                same = Z.l == A.l && Z.A[0] == A.A[0] && Z.A[1] == A.A[1] && Z.A[2] == A.A[2] && Z.alpha == A.alpha;
            }
            if (!same) f0.push_back(s);
        }
        const int nf = (int)f0.size();
        f0.push_back(ns);
        std::vector<std::pair<int, int>> fps;
        for (int fa = 0; fa < nf; ++fa)
            for (int fb = 0; fb <= fa; ++fb) fps.push_back({fa, fb});
        std::vector<FamPair*> built(fps.size(), nullptr);
#pragma omp parallel num_threads(nthreads())
        {
            PrimE pe;
            std::vector<double> Etmp;
#pragma omp for schedule(dynamic)
            for (size_t k = 0; k < fps.size(); ++k) {
                const int fa = fps[k].first, fb = fps[k].second;
                FamPair* f = new FamPair();
                f->a0 = f0[fa];
                f->na = f0[fa + 1] - f0[fa];
                f->b0 = f0[fb];
                f->nb = f0[fb + 1] - f0[fb];
                for (int ca = 0; ca < f->na; ++ca)
                    for (int cb = 0; cb < f->nb; ++cb)
                        if (f->a0 + ca >= f->b0 + cb) f->cols.push_back({ca, cb});
                const int ncp = (int)f->cols.size();
                const Shell& A = bs->shells[f->a0];
                const Shell& B = bs->shells[f->b0];
                const double AB2 = (A.A[0] - B.A[0]) * (A.A[0] - B.A[0]) + (A.A[1] - B.A[1]) * (A.A[1] - B.A[1]) +
                                   (A.A[2] - B.A[2]) * (A.A[2] - B.A[2]);
                std::vector<char> used(ncp, 0);
                for (size_t ia = 0; ia < A.alpha.size(); ++ia)
                    for (size_t ib = 0; ib < B.alpha.size(); ++ib) {
                        const double a = A.alpha[ia], b = B.alpha[ib], p = a + b;
                        const double K = std::exp(-a * b / p * AB2);
                        const double G = K * std::pow(PI / p, 1.5);
                        FamPrim pp;
                        pp.w.assign(ncp, 0.0);
                        bool any = false;
                        for (int c = 0; c < ncp; ++c) {
                            // the screening test of make_pair, per column pair
                            const double cc = bs->shells[f->a0 + f->cols[c].first].coef[ia] *
                                              bs->shells[f->b0 + f->cols[c].second].coef[ib];
                            if (std::fabs(cc) * G < kPrimScreen) continue;
                            pp.w[c] = cc;
                            used[c] = 1;
                            any = true;
                        }
                        if (!any) continue;
                        pp.p = p;
                        for (int d = 0; d < 3; ++d) pp.P[d] = (a * A.A[d] + b * B.A[d]) / p;
                        // stored h-major so that the aux contraction runs over the
                        // pair components in the inner loop
                        pe.compute(A, B, a, b, true, Etmp);
                        const int nab = nsph(A.l) * nsph(B.l), nh = nherm(A.l + B.l);
                        pp.E.resize(Etmp.size());
                        for (int s2 = 0; s2 < nab; ++s2)
                            for (int h = 0; h < nh; ++h) pp.E[(size_t)h * nab + s2] = Etmp[(size_t)s2 * nh + h];
                        f->prims.push_back(std::move(pp));
                    }
                f->q.assign(ncp, -1);
                for (int c = 0; c < ncp; ++c)
                    if (used[c]) f->q[c] = 0;  // significant; numbered below
                built[k] = f;
            }
        }
        // number the significant shell pairs in (a, b) order
        std::vector<std::tuple<int, int, int, int>> sig;  // (a, b, family pair, column pair)
        for (size_t k = 0; k < fps.size(); ++k) {
            FamPair* f = built[k];
            for (size_t c = 0; c < f->cols.size(); ++c)
                if (f->q[c] >= 0)
                    sig.emplace_back(f->a0 + f->cols[c].first, f->b0 + f->cols[c].second, (int)k, (int)c);
        }
        std::sort(sig.begin(), sig.end());
        for (size_t q = 0; q < sig.size(); ++q) {
            const auto& t = sig[q];
            bs->pair_index.push_back({std::get<0>(t), std::get<1>(t)});
            bs->fam_of_pair.push_back(std::get<2>(t));
            bs->col_of_pair.push_back(std::get<3>(t));
            built[std::get<2>(t)]->q[std::get<3>(t)] = (int64_t)q;
        }
        for (FamPair* f : built) bs->fam_pairs.push_back(f);
    });
    return bs->pair_index;
}

// scratch of one thread for eval_fam
struct Scratch3c {
    std::vector<double> R, work, Y, V, acc;
    std::vector<double> R4, work4, Y4, V4, acc4[kLanes];  // lane-batched (eval_fam4)
    Scratch3c(int maxlaux, int maxl) {
        const int Lmax = maxlaux + 2 * maxl;
        R.resize(nherm(Lmax));
        work.resize((size_t)(Lmax + 1) * nherm(Lmax));
        Y.resize((size_t)nsph(maxlaux) * nherm(2 * maxl));
        V.resize((size_t)nsph(maxlaux) * nsph(maxl) * nsph(maxl));
        R4.resize(R.size() * kLanes);
        work4.resize(work.size() * kLanes);
        Y4.resize(Y.size() * kLanes);
        V4.resize(V.size() * kLanes);
    }
};

// Positions in hidx().list of R_{c + h} for the Cartesian component c of an
// aux shell of angular momentum lP and the Hermite index h of a pair of total
// angular momentum lab: [c * nherm(lab) + h].
const std::vector<int>& y_index(int lP, int lab) {
    static const std::vector<std::vector<std::vector<int>>> tab = [] {
        const Harmonics& Hm = harm();
        const HermIndex& H = hidx();
        std::vector<std::vector<std::vector<int>>> t(LMAX + 1, std::vector<std::vector<int>>(2 * LMAX + 1));
        for (int l = 0; l <= LMAX; ++l)
            for (int L = 0; L <= 2 * LMAX; ++L) {
                const int nh = nherm(L);
                std::vector<int>& v = t[l][L];
                v.resize((size_t)ncart(l) * nh);
                for (int c = 0; c < ncart(l); ++c) {
                    const CartComp& cp = Hm.cart[l][c];
                    for (int h = 0; h < nh; ++h) {
                        const auto& tuv = H.list[L][h];
                        v[(size_t)c * nh + h] = H.at(cp.x + tuv[0], cp.y + tuv[1], cp.z + tuv[2]);
                    }
                }
            }
        return t;
    }();
    return tab[lP][lab];
}

// (P|ab) of aux shell P for every column pair c of family pair f, into
// w.acc[(c * nsP + sP) * nab + sa * nsb + sb].  Per column pair the terms are
// added in the order of a single-shell evaluation (aux primitive, then
// primitive pair), each as (prefactor * coefficient product) * integral.
void eval_fam(const f2ints_basis* bs, const FamPair& f, const Shell& P, Scratch3c& w) {
    const Harmonics& Hm = harm();
    const HermIndex& H = hidx();
    const Shell& A = bs->shells[f.a0];
    const Shell& B = bs->shells[f.b0];
    const int lab = A.l + B.l, nhab = nherm(lab), nab = nsph(A.l) * nsph(B.l);
    const int lP = P.l, nsP = nsph(lP), L = lP + lab;
    const int ncp = (int)f.cols.size();
    w.acc.assign((size_t)ncp * nsP * nab, 0.0);
    static const double two_pi52 = 2.0 * std::pow(PI, 2.5);
    const std::vector<int>& yix = y_index(lP, lab);
    for (size_t k = 0; k < P.alpha.size(); ++k) {
        const double g = P.alpha[k];
        const double gl = std::pow(2.0 * g, -lP);  // once per aux primitive
        for (const FamPrim& pp : f.prims) {
            const double p = pp.p, al = g * p / (g + p);
            compute_R(L, al, P.A[0] - pp.P[0], P.A[1] - pp.P[1], P.A[2] - pp.P[2], w.R.data(), w.work.data());
            const double base = two_pi52 / (g * p * std::sqrt(g + p)) * P.coef[k] * gl;
            // Y[sP][h] = sum_c T_{lP}(sP, c) R[c + h]
            for (int sP = 0; sP < nsP; ++sP) {
                double* y = &w.Y[(size_t)sP * nhab];
                std::fill(y, y + nhab, 0.0);
                for (const SphTerm& tp : Hm.sparse[lP][sP]) {
                    const int* ix = &yix[(size_t)tp.cart * nhab];
                    for (int h = 0; h < nhab; ++h) y[h] = std::fma(tp.c, w.R[ix[h]], y[h]);
                }
            }
            // V[sP][s2] = sum_h E[s2][h] Y[sP][h], shared by the column pairs;
            // each V[sP][s2] sums over h in ascending order, the inner loop
            // runs over independent s2
            for (int sP = 0; sP < nsP; ++sP) {
                const double* y = &w.Y[(size_t)sP * nhab];
                double* v = &w.V[(size_t)sP * nab];
                std::fill(v, v + nab, 0.0);
                for (int h = 0; h < nhab; ++h) {
                    const double yh = y[h];
                    const double* e = &pp.E[(size_t)h * nab];
                    for (int s2 = 0; s2 < nab; ++s2) v[s2] = std::fma(e[s2], yh, v[s2]);
                }
            }
            for (int c = 0; c < ncp; ++c) {
                if (pp.w[c] == 0.0) continue;
                const double pref = base * pp.w[c];
                double* a = &w.acc[(size_t)c * nsP * nab];
                for (int i = 0; i < nsP * nab; ++i) a[i] = std::fma(pref, w.V[i], a[i]);
            }
        }
    }
}

// eval_fam for kLanes aux shells of one angular momentum with one primitive
// each: w.acc4[lane] as w.acc of eval_fam(.., *Ps[lane], w), lane by lane the
// same arithmetic in the same order (bit for bit the same values); R, the
// aux-side transform and the E contraction run across the lanes.
void eval_fam4(const f2ints_basis* bs, const FamPair& f, const Shell* const* Ps, Scratch3c& w) {
    const Harmonics& Hm = harm();
    const Shell& A = bs->shells[f.a0];
    const Shell& B = bs->shells[f.b0];
    const int lab = A.l + B.l, nhab = nherm(lab), nab = nsph(A.l) * nsph(B.l);
    const int lP = Ps[0]->l, nsP = nsph(lP), L = lP + lab;
    const int ncp = (int)f.cols.size();
    for (int l = 0; l < kLanes; ++l) w.acc4[l].assign((size_t)ncp * nsP * nab, 0.0);
    static const double two_pi52 = 2.0 * std::pow(PI, 2.5);
    const std::vector<int>& yix = y_index(lP, lab);
    double g[kLanes], gl[kLanes], coef[kLanes];
    for (int l = 0; l < kLanes; ++l) {
        g[l] = Ps[l]->alpha[0];
        gl[l] = std::pow(2.0 * g[l], -lP);
        coef[l] = Ps[l]->coef[0];
    }
    double al[kLanes], X[kLanes], Y[kLanes], Z[kLanes], base[kLanes];
    for (const FamPrim& pp : f.prims) {
        const double p = pp.p;
        for (int l = 0; l < kLanes; ++l) {
            al[l] = g[l] * p / (g[l] + p);
            X[l] = Ps[l]->A[0] - pp.P[0];
            Y[l] = Ps[l]->A[1] - pp.P[1];
            Z[l] = Ps[l]->A[2] - pp.P[2];
            base[l] = two_pi52 / (g[l] * p * std::sqrt(g[l] + p)) * coef[l] * gl[l];
        }
        compute_R4(L, al, X, Y, Z, w.R4.data(), w.work4.data());
        // Y4[(sP * nhab + h) * kLanes + lane]
        for (int sP = 0; sP < nsP; ++sP) {
            double* y = &w.Y4[(size_t)sP * nhab * kLanes];
            std::fill(y, y + (size_t)nhab * kLanes, 0.0);
            for (const SphTerm& tp : Hm.sparse[lP][sP]) {
                const int* ix = &yix[(size_t)tp.cart * nhab];
                for (int h = 0; h < nhab; ++h) {
                    const double* r = &w.R4[(size_t)ix[h] * kLanes];
                    for (int l = 0; l < kLanes; ++l) y[h * kLanes + l] = std::fma(tp.c, r[l], y[h * kLanes + l]);
                }
            }
        }
        // V4[lane * nsP * nab + sP * nab + s2], h ascending as in eval_fam; the
        // kLanes sums of one s2 stay in registers over the h loop, and each
        // lane's values are contiguous for the column accumulation below
        const int nv = nsP * nab;
        for (int sP = 0; sP < nsP; ++sP) {
            const double* y = &w.Y4[(size_t)sP * nhab * kLanes];
            for (int s2 = 0; s2 < nab; ++s2) {
                double acc[kLanes] = {};
                const double* e = &pp.E[s2];
                for (int h = 0; h < nhab; ++h) {
                    const double eh = e[(size_t)h * nab];
                    const double* yh = y + (size_t)h * kLanes;
                    for (int l = 0; l < kLanes; ++l) acc[l] = std::fma(eh, yh[l], acc[l]);
                }
                for (int l = 0; l < kLanes; ++l) w.V4[(size_t)l * nv + sP * nab + s2] = acc[l];
            }
        }
        for (int c = 0; c < ncp; ++c) {
            if (pp.w[c] == 0.0) continue;
            for (int l = 0; l < kLanes; ++l) {
                const double pref = base[l] * pp.w[c];
                double* a = &w.acc4[l][(size_t)c * nv];
                const double* v = &w.V4[(size_t)l * nv];
                for (int i = 0; i < nv; ++i) a[i] = std::fma(pref, v[i], a[i]);
            }
        }
    }
}

// The aux shells [s0, s1) in batches: runs of kLanes shells of one angular
// momentum with one primitive each (taken in order of l, then of index), and
// the rest one by one.  Each batch is a list of shell indices.
std::vector<std::vector<int>> aux_batches(const f2ints_basis* aux, int s0, int s1) {
    std::vector<std::vector<int>> out;
    std::vector<std::vector<int>> by_l(LMAX + 1);
    for (int s = s0; s < s1; ++s) {
        const Shell& P = aux->shells[s];
        if (P.alpha.size() == 1) by_l[P.l].push_back(s);
        else out.push_back({s});
    }
    for (int l = 0; l <= LMAX; ++l) {
        const std::vector<int>& v = by_l[l];
        size_t i = 0;
        for (; i + kLanes <= v.size(); i += kLanes) out.emplace_back(v.begin() + i, v.begin() + i + kLanes);
        for (; i < v.size(); ++i) out.push_back({v[i]});
    }
    return out;
}

int max_aux_l(const f2ints_basis* aux, int sh0, int sh1) {
    int m = 0;
    for (int s = sh0; s < sh1; ++s) m = std::max(m, aux->shells[s].l);
    return m;
}

int coulomb_3c(const f2ints_basis* aux, const f2ints_basis* bs, int sh0, int sh1, double* out) {
    const int64_t nbf = bs->nbf;
    const int64_t first = aux->shells[sh0].off;
    ket_pairs(bs);
    const int64_t nP = (sh1 < (int)aux->shells.size() ? aux->shells[sh1].off : aux->nbf) - first;
    std::fill(out, out + (size_t)nP * nbf * nbf, 0.0);
    const int maxlaux = max_aux_l(aux, sh0, sh1);
#pragma omp parallel num_threads(nthreads())
    {
        Scratch3c w(maxlaux, bs->maxl);
#pragma omp for schedule(dynamic)
        for (size_t k = 0; k < bs->fam_pairs.size(); ++k) {
            const FamPair& f = *static_cast<const FamPair*>(bs->fam_pairs[k]);
            const int nsa = nsph(bs->shells[f.a0].l), nsbb = nsph(bs->shells[f.b0].l), nab = nsa * nsbb;
            for (int s = sh0; s < sh1; ++s) {
                const Shell& P = aux->shells[s];
                const int nsP = nsph(P.l);
                eval_fam(bs, f, P, w);
                for (size_t c = 0; c < f.cols.size(); ++c) {
                    if (f.q[c] < 0) continue;
                    const Shell& A = bs->shells[f.a0 + f.cols[c].first];
                    const Shell& B = bs->shells[f.b0 + f.cols[c].second];
                    const double* acc = &w.acc[c * nsP * nab];
                    for (int sP = 0; sP < nsP; ++sP) {
                        double* o = out + (size_t)(P.off + sP - first) * nbf * nbf;
                        for (int sa = 0; sa < nsa; ++sa)
                            for (int sb = 0; sb < nsbb; ++sb) {
                                const double v = acc[(size_t)sP * nab + sa * nsbb + sb];
                                const int64_t m = A.off + sa, n = B.off + sb;
                                o[m * nbf + n] = v;
                                o[n * nbf + m] = v;
                            }
                    }
                }
            }
        }
    }
    return 0;
}

// V[m][n] = sum over every aux function P of (P|mn), both triangles
int coulomb_3c_sum(const f2ints_basis* aux, const f2ints_basis* bs, double* out) {
    const int64_t nbf = bs->nbf;
    ket_pairs(bs);
    std::fill(out, out + (size_t)nbf * nbf, 0.0);
    const int maxlaux = max_aux_l(aux, 0, (int)aux->shells.size());
    const std::vector<std::vector<int>> batches = aux_batches(aux, 0, (int)aux->shells.size());
#pragma omp parallel num_threads(nthreads())
    {
        Scratch3c w(maxlaux, bs->maxl);
        std::vector<double> sum;
#pragma omp for schedule(dynamic)
        for (size_t k = 0; k < bs->fam_pairs.size(); ++k) {
            const FamPair& f = *static_cast<const FamPair*>(bs->fam_pairs[k]);
            const int nsa = nsph(bs->shells[f.a0].l), nsbb = nsph(bs->shells[f.b0].l), nab = nsa * nsbb;
            const size_t ncp = f.cols.size();
            sum.assign(ncp * nab, 0.0);
            auto add = [&](const Shell& P, const std::vector<double>& acc) {
                const int nsP = nsph(P.l);
                for (size_t c = 0; c < ncp; ++c)
                    for (int sP = 0; sP < nsP; ++sP)
                        for (int i = 0; i < nab; ++i) sum[c * nab + i] += acc[(c * nsP + sP) * nab + i];
            };
            // aux shells in batch order (by l, then index)
            for (const std::vector<int>& bt : batches) {
                if (bt.size() == (size_t)kLanes) {
                    const Shell* Ps[kLanes];
                    for (int l = 0; l < kLanes; ++l) Ps[l] = &aux->shells[bt[l]];
                    eval_fam4(bs, f, Ps, w);
                    for (int l = 0; l < kLanes; ++l) add(*Ps[l], w.acc4[l]);
                } else {
                    eval_fam(bs, f, aux->shells[bt[0]], w);
                    add(aux->shells[bt[0]], w.acc);
                }
            }
            // this family pair owns these entries: no other thread writes them
            for (size_t c = 0; c < ncp; ++c) {
                if (f.q[c] < 0) continue;
                const Shell& A = bs->shells[f.a0 + f.cols[c].first];
                const Shell& B = bs->shells[f.b0 + f.cols[c].second];
                for (int sa = 0; sa < nsa; ++sa)
                    for (int sb = 0; sb < nsbb; ++sb) {
                        const int64_t m = A.off + sa, n = B.off + sb;
                        out[m * nbf + n] = sum[c * nab + sa * nsbb + sb];
                        out[n * nbf + m] = sum[c * nab + sa * nsbb + sb];
                    }
            }
        }
    }
    return 0;
}

// packed columns of pair q: nsa * nsb, or nsa (nsa + 1) / 2 for a == b
int64_t pair_ncols(const f2ints_basis* bs, const std::pair<int, int>& ab) {
    const int na = nsph(bs->shells[ab.first].l), nb = nsph(bs->shells[ab.second].l);
    return ab.first == ab.second ? (int64_t)na * (na + 1) / 2 : (int64_t)na * nb;
}

// (P|ab) for every aux function and the significant pairs [q0, q1), packed:
// out[P * ldo + c], column c running over the pairs in order; a pair a > b
// has nsa * nsb columns (row-major in (sa, sb)), a pair a == b the
// nsa (nsa + 1) / 2 components sa >= sb (row-major), so every (m, n) with
// m >= n appears once.  The family pairs that hold these pairs are evaluated
// whole; their other column pairs are not written.
// Rows P of the aux shells [s0, s1) (row 0 = the first function of shell s0),
// packed columns of the pairs [q0, q1).  With `colmajor` the same values go
// to out[c * ldo + P]: a thread then fills whole rows of its columns instead
// of a narrow strip of every row.
int coulomb_3c_pairs(const f2ints_basis* aux, const f2ints_basis* bs, int s0, int s1, int64_t q0, int64_t q1,
                     double* out, int64_t ldo, bool colmajor = false) {
    const std::vector<std::pair<int, int>>& pairs = ket_pairs(bs);
    if (q0 < 0 || q1 > (int64_t)pairs.size() || q0 > q1) return fail("f2ints_coulomb_3c_pairs: bad pair range");
    if (s0 < 0 || s1 > (int)aux->shells.size() || s0 >= s1) return fail("f2ints_coulomb_3c_pairs: bad aux shell range");
    const int64_t prow0 = aux->shells[s0].off;
    std::vector<int64_t> col(q1 - q0 + 1, 0);
    for (int64_t q = q0; q < q1; ++q) col[q - q0 + 1] = col[q - q0] + pair_ncols(bs, pairs[q]);
    const int64_t nrows = aux->shells[s1 - 1].off + nsph(aux->shells[s1 - 1].l) - prow0;
    if (!colmajor && ldo < col.back()) return fail("f2ints_coulomb_3c_pairs: ldo smaller than the block's columns");
    if (colmajor && ldo < nrows) return fail("f2ints_coulomb_3c_pairs: ldo smaller than the aux functions");
    std::vector<int> fams;
    for (int64_t q = q0; q < q1; ++q) fams.push_back(bs->fam_of_pair[q]);
    std::sort(fams.begin(), fams.end());
    fams.erase(std::unique(fams.begin(), fams.end()), fams.end());
    const int maxlaux = max_aux_l(aux, s0, s1);
    const std::vector<std::vector<int>> batches = aux_batches(aux, s0, s1);
#pragma omp parallel num_threads(nthreads())
    {
        Scratch3c w(maxlaux, bs->maxl);
#pragma omp for schedule(dynamic)
        for (size_t k = 0; k < fams.size(); ++k) {
            const FamPair& f = *static_cast<const FamPair*>(bs->fam_pairs[fams[k]]);
            const int na = nsph(bs->shells[f.a0].l), nb = nsph(bs->shells[f.b0].l), nab = na * nb;
            auto store = [&](const Shell& P, const std::vector<double>& acc) {
                const int nsP = nsph(P.l);
                for (size_t c = 0; c < f.cols.size(); ++c) {
                    const int64_t q = f.q[c];
                    if (q < q0 || q >= q1) continue;
                    const bool diag = pairs[q].first == pairs[q].second;
                    if (colmajor) {
                        // column t of the pair: rows P.off - prow0 + sP, sP ascending
                        double* o = out + (size_t)col[q - q0] * ldo + (P.off - prow0);
                        int t = 0;
                        for (int sa = 0; sa < na; ++sa)
                            for (int sb = 0; sb < (diag ? sa + 1 : nb); ++sb, ++t) {
                                double* dst = o + (size_t)t * ldo;
                                for (int sP = 0; sP < nsP; ++sP) dst[sP] = acc[(c * nsP + sP) * nab + sa * nb + sb];
                            }
                        continue;
                    }
                    double* o = out + col[q - q0];
                    for (int sP = 0; sP < nsP; ++sP) {
                        double* row = o + (size_t)(P.off - prow0 + sP) * ldo;
                        const double* a = &acc[(c * nsP + sP) * nab];
                        if (!diag) {
                            std::copy(a, a + nab, row);
                        } else {
                            int t = 0;
                            for (int sa = 0; sa < na; ++sa)
                                for (int sb = 0; sb <= sa; ++sb) row[t++] = a[sa * nb + sb];
                        }
                    }
                }
            };
            for (const std::vector<int>& bt : batches) {
                if (bt.size() == (size_t)kLanes) {
                    const Shell* Ps[kLanes];
                    for (int l = 0; l < kLanes; ++l) Ps[l] = &aux->shells[bt[l]];
                    eval_fam4(bs, f, Ps, w);
                    for (int l = 0; l < kLanes; ++l) store(*Ps[l], w.acc4[l]);
                } else {
                    const Shell& P = aux->shells[bt[0]];
                    eval_fam(bs, f, P, w);
                    store(P, w.acc);
                }
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------- 4-centre (tests)
int coulomb_4c(const f2ints_basis* bs, double* out) {
    const HermIndex& H = hidx();
    const int ns = (int)bs->shells.size();
    const int64_t n = bs->nbf;
    std::vector<ShellPair> bra(ns * ns), ket(ns * ns);
    for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b) {
            make_pair(bs->shells[a], bs->shells[b], false, bra[a * ns + b]);
            make_pair(bs->shells[a], bs->shells[b], true, ket[a * ns + b]);
        }
#pragma omp parallel for schedule(dynamic) num_threads(nthreads())
    for (int ab = 0; ab < ns * ns; ++ab) {
        const Shell& A = bs->shells[ab / ns];
        const Shell& B = bs->shells[ab % ns];
        const int lab = A.l + B.l, nhab = nherm(lab), nsa = nsph(A.l), nsbb = nsph(B.l);
        for (int cd = 0; cd < ns * ns; ++cd) {
            const Shell& C = bs->shells[cd / ns];
            const Shell& D = bs->shells[cd % ns];
            const int lcd = C.l + D.l, nhcd = nherm(lcd), nsc = nsph(C.l), nsd = nsph(D.l);
            const int L = lab + lcd;
            std::vector<double> R(nherm(L)), work((size_t)(L + 1) * nherm(L));
            std::vector<double> acc((size_t)nsa * nsbb * nsc * nsd, 0.0);
            for (const PrimPair& p1 : bra[ab].prims)
                for (const PrimPair& p2 : ket[cd].prims) {
                    const double p = p1.p, qq = p2.p, al = p * qq / (p + qq);
                    compute_R(L, al, p1.P[0] - p2.P[0], p1.P[1] - p2.P[1], p1.P[2] - p2.P[2], R.data(),
                              work.data());
                    const double pref = 2.0 * std::pow(PI, 2.5) / (p * qq * std::sqrt(p + qq)) * p1.scale * p2.scale;
                    for (int s1 = 0; s1 < nsa * nsbb; ++s1)
                        for (int s2 = 0; s2 < nsc * nsd; ++s2) {
                            double v = 0.0;
                            for (int h1 = 0; h1 < nhab; ++h1) {
                                const double e1 = p1.E[(size_t)s1 * nhab + h1];
                                if (e1 == 0.0) continue;
                                const auto& t1 = H.list[lab][h1];
                                double w = 0.0;
                                for (int h2 = 0; h2 < nhcd; ++h2) {
                                    const auto& t2 = H.list[lcd][h2];
                                    w += p2.E[(size_t)s2 * nhcd + h2] * R[H.at(t1[0] + t2[0], t1[1] + t2[1], t1[2] + t2[2])];
                                }
                                v += e1 * w;
                            }
                            acc[(size_t)s1 * nsc * nsd + s2] += pref * v;
                        }
                }
            for (int sa = 0; sa < nsa; ++sa)
                for (int sb = 0; sb < nsbb; ++sb)
                    for (int sc = 0; sc < nsc; ++sc)
                        for (int sd = 0; sd < nsd; ++sd) {
                            const int64_t i = A.off + sa, j = B.off + sb, k = C.off + sc, l = D.off + sd;
                            out[((i * n + j) * n + k) * n + l] =
                                acc[((size_t)(sa * nsbb + sb)) * nsc * nsd + sc * nsd + sd];
                        }
        }
    }
    return 0;
}

}  // namespace

f2ints_basis::~f2ints_basis() {
    for (void* p : fam_pairs) delete static_cast<FamPair*>(p);
}

// =========================================================================== C ABI
extern "C" {

int f2ints_basis_create(int nshell, const int* l, const int* nprim, const double* xyz, const double* exps,
                        const double* coefs, f2ints_basis** out) {
    if (!out || nshell < 0 || (nshell > 0 && (!l || !nprim || !xyz || !exps || !coefs)))
        return fail("f2ints_basis_create: invalid arguments");
    f2ints_basis* b = new f2ints_basis();
    int64_t off = 0, k = 0;
    for (int s = 0; s < nshell; ++s) {
        if (l[s] < 0 || l[s] > LMAX || nprim[s] < 1) {
            delete b;
            return fail("f2ints_basis_create: shell " + std::to_string(s) + " has l or nprim out of range");
        }
        Shell sh;
        sh.l = l[s];
        for (int d = 0; d < 3; ++d) sh.A[d] = xyz[3 * s + d];
        for (int p = 0; p < nprim[s]; ++p, ++k) {
            if (coefs[k] == 0.0) continue;  // zero columns of general contractions
            sh.alpha.push_back(exps[k]);
            sh.coef.push_back(coefs[k]);
        }
        if (sh.alpha.empty()) {  // keep one zero primitive so the shell exists
            sh.alpha.push_back(exps[k - 1]);
            sh.coef.push_back(0.0);
        }
        sh.off = off;
        off += nsph(sh.l);
        b->maxl = std::max(b->maxl, sh.l);
        b->shells.push_back(std::move(sh));
    }
    b->nbf = off;
    *out = b;
    return 0;
}

int f2ints_basis_destroy(f2ints_basis* b) {
    delete b;
    return 0;
}

int64_t f2ints_basis_nbf(const f2ints_basis* b) { return b ? b->nbf : -1; }

int f2ints_max_l(void) { return LMAX; }

int f2ints_overlap(const f2ints_basis* b, double* S) {
    if (!b || !S) return fail("f2ints_overlap: null argument");
    return one_electron(b, kOverlap, S);
}

int f2ints_kinetic(const f2ints_basis* b, double* T) {
    if (!b || !T) return fail("f2ints_kinetic: null argument");
    return one_electron(b, kKinetic, T);
}

int f2ints_nuclear(const f2ints_basis* b, int nc, const double* Z, const double* xyz, double* V) {
    if (!b || !V || nc < 0 || (nc > 0 && (!Z || !xyz))) return fail("f2ints_nuclear: invalid arguments");
    return nuclear(b, nc, Z, xyz, V);
}

double f2ints_nuclear_repulsion(int natom, const double* Z, const double* xyz) {
    double e = 0.0;
    for (int i = 0; i < natom; ++i)
        for (int j = 0; j < i; ++j) {
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                         dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            e += Z[i] * Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

int f2ints_coulomb_2c(const f2ints_basis* a, const f2ints_basis* b, double* M) {
    if (!a || !b || !M) return fail("f2ints_coulomb_2c: null argument");
    return coulomb_2c(a, b, M);
}

int f2ints_coulomb_3c(const f2ints_basis* aux, const f2ints_basis* bs, int sb, int se, double* out) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c: null argument");
    const int ns = (int)aux->shells.size();
    if (se < 0 || se > ns) se = ns;
    if (sb < 0 || sb > se) return fail("f2ints_coulomb_3c: bad shell range");
    if (sb == se) return 0;
    return coulomb_3c(aux, bs, sb, se, out);
}

int64_t f2ints_ket_pairs(const f2ints_basis* bs, int* a, int* b, int64_t cap) {
    if (!bs) {
        fail("f2ints_ket_pairs: null basis");
        return -1;
    }
    const std::vector<std::pair<int, int>>& pairs = ket_pairs(bs);
    const int64_t n = (int64_t)pairs.size();
    if (a && b)
        for (int64_t q = 0; q < std::min(n, cap); ++q) {
            a[q] = pairs[q].first;
            b[q] = pairs[q].second;
        }
    return n;
}

int f2ints_coulomb_3c_sum(const f2ints_basis* aux, const f2ints_basis* bs, double* out) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c_sum: null argument");
    return coulomb_3c_sum(aux, bs, out);
}

int f2ints_coulomb_3c_pairs(const f2ints_basis* aux, const f2ints_basis* bs, int64_t q0, int64_t q1,
                            double* out, int64_t ldo) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c_pairs: null argument");
    if (q0 == q1 || aux->shells.empty()) return 0;
    return coulomb_3c_pairs(aux, bs, 0, (int)aux->shells.size(), q0, q1, out, ldo);
}

int f2ints_coulomb_3c_pairs_t(const f2ints_basis* aux, const f2ints_basis* bs, int64_t q0, int64_t q1, double* out,
                              int64_t ldo) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c_pairs_t: null argument");
    if (q0 == q1 || aux->shells.empty()) return 0;
    return coulomb_3c_pairs(aux, bs, 0, (int)aux->shells.size(), q0, q1, out, ldo, true);
}

int f2ints_coulomb_3c_pairs_aux(const f2ints_basis* aux, const f2ints_basis* bs, int shell_begin, int shell_end,
                                int64_t q0, int64_t q1, double* out, int64_t ldo) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c_pairs_aux: null argument");
    if (q0 == q1 || shell_begin == shell_end) return 0;
    return coulomb_3c_pairs(aux, bs, shell_begin, shell_end, q0, q1, out, ldo);
}

int f2ints_coulomb_3c_pairs_aux_t(const f2ints_basis* aux, const f2ints_basis* bs, int shell_begin, int shell_end,
                                  int64_t q0, int64_t q1, double* out, int64_t ldo) {
    if (!aux || !bs || !out) return fail("f2ints_coulomb_3c_pairs_aux_t: null argument");
    if (q0 == q1 || shell_begin == shell_end) return 0;
    return coulomb_3c_pairs(aux, bs, shell_begin, shell_end, q0, q1, out, ldo, true);
}

int f2ints_coulomb_4c(const f2ints_basis* b, double* out) {
    if (!b || !out) return fail("f2ints_coulomb_4c: null argument");
    if (b->nbf > 200) return fail("f2ints_coulomb_4c: basis too large for the test-only 4-centre routine");
    return coulomb_4c(b, out);
}

int f2ints_boys(int nmax, double T, double* out) {
    if (nmax < 0 || nmax > NBOYS || !out || T < 0) return fail("f2ints_boys: invalid arguments");
    boys().eval(nmax, T, out);
    return 0;
}

int f2ints_cart2sph(int l, double* out) {
    if (l < 0 || l > LMAX || !out) return fail("f2ints_cart2sph: invalid arguments");
    const auto& T = harm().T[l];
    std::copy(T.begin(), T.end(), out);
    return 0;
}

int f2ints_set_threads(int n) {
    g_threads = n;
    return 0;
}

const char* f2ints_last_error(void) { return g_err.c_str(); }

}  // extern "C"
