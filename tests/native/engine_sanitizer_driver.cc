// Host integral engine under AddressSanitizer / UndefinedBehaviorSanitizer:
// every entry point of include/f2ints.h on a small basis whose fitting set
// has runs of four and more single-primitive shells per l (the lane-batched
// path), a contracted shell and a leftover shell per l (the one-shell path).
// Exit code 0 when every call succeeds and the packed, ranged and square
// three-centre results agree bit for bit.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "f2ints.h"

namespace {

struct Spec {
    std::vector<int> l, nprim;
    std::vector<double> xyz, exps, coefs;
    void add(int L, const double* c, std::vector<double> e, std::vector<double> k) {
        l.push_back(L);
        nprim.push_back((int)e.size());
        xyz.insert(xyz.end(), c, c + 3);
        exps.insert(exps.end(), e.begin(), e.end());
        coefs.insert(coefs.end(), k.begin(), k.end());
    }
    f2ints_basis* make() const {
        f2ints_basis* b = nullptr;
        if (f2ints_basis_create((int)l.size(), l.data(), nprim.data(), xyz.data(), exps.data(), coefs.data(), &b) != 0)
            return nullptr;
        return b;
    }
};

int fail(const char* what) {
    std::fprintf(stderr, "FAIL %s: %s\n", what, f2ints_last_error());
    return 1;
}

}  // namespace

int main() {
    const double A[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 1.43, 1.11}, C[3] = {2.9, -0.3, 0.4};
    Spec orb, aux;
    for (const double* c : {A, B, C}) {
        orb.add(0, c, {130.7, 23.8, 6.44}, {0.154, 0.535, 0.445});
        orb.add(0, c, {5.03, 1.17, 0.38}, {-0.1, 0.4, 0.7});
        orb.add(1, c, {5.03, 1.17, 0.38}, {0.156, 0.607, 0.392});
        orb.add(2, c, {0.8}, {1.0});
    }
    for (const double* c : {A, B, C}) {
        for (int L = 0; L <= 4; ++L) {
            const int n = L < 2 ? 6 : (L < 4 ? 5 : 1);  // runs of 4 plus leftovers
            for (int i = 0; i < n; ++i) aux.add(L, c, {0.3 * std::pow(2.2, i) + 0.1 * L}, {1.0});
        }
        aux.add(0, c, {9.0, 2.0}, {0.6, 0.5});  // contracted: one-shell path
    }
    f2ints_basis* bs = orb.make();
    f2ints_basis* ax = aux.make();
    if (!bs || !ax) return fail("basis_create");
    const int64_t N = f2ints_basis_nbf(bs), Q = f2ints_basis_nbf(ax);
    std::vector<double> S(N * N), T(N * N), V(N * N), M(Q * Q);
    const double Z[3] = {8.0, 1.0, 1.0};
    const double centres[9] = {A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]};
    if (f2ints_overlap(bs, S.data()) || f2ints_kinetic(bs, T.data()) || f2ints_nuclear(bs, 3, Z, centres, V.data()))
        return fail("one-electron");
    if (f2ints_coulomb_2c(ax, ax, M.data())) return fail("coulomb_2c");
    const int nsh_aux = (int)aux.l.size();
    std::vector<double> sq(Q * N * N);
    if (f2ints_coulomb_3c(ax, bs, 0, nsh_aux, sq.data())) return fail("coulomb_3c");
    const int64_t npairs = f2ints_ket_pairs(bs, nullptr, nullptr, 0);
    std::vector<int> pa(npairs), pb(npairs);
    f2ints_ket_pairs(bs, pa.data(), pb.data(), npairs);
    // packed column count, as pair_columns in forte2_amd/df.py lays it out
    std::vector<int> off(orb.l.size() + 1, 0);
    for (size_t s = 0; s < orb.l.size(); ++s) off[s + 1] = off[s] + 2 * orb.l[s] + 1;
    std::vector<int64_t> colm, coln;
    for (int64_t q = 0; q < npairs; ++q) {
        const int a = pa[q], b = pb[q], na = off[a + 1] - off[a], nb = off[b + 1] - off[b];
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j)
                if (a != b || j <= i) {
                    colm.push_back(off[a] + i);
                    coln.push_back(off[b] + j);
                }
    }
    const int64_t nc = (int64_t)colm.size();
    std::vector<double> packed(Q * nc);
    if (f2ints_coulomb_3c_pairs(ax, bs, 0, npairs, packed.data(), nc)) return fail("coulomb_3c_pairs");
    for (int64_t P = 0; P < Q; ++P)
        for (int64_t c = 0; c < nc; ++c)
            if (packed[P * nc + c] != sq[(P * N + colm[c]) * N + coln[c]]) return fail("packed != square");
    // the column-major layout holds the same values, one row per column
    // (row stride Q + 3: the padding must stay untouched)
    std::vector<double> packed_t(nc * (Q + 3), -7.0);
    if (f2ints_coulomb_3c_pairs_t(ax, bs, 0, npairs, packed_t.data(), Q + 3)) return fail("coulomb_3c_pairs_t");
    for (int64_t c = 0; c < nc; ++c) {
        for (int64_t P = 0; P < Q; ++P)
            if (packed_t[c * (Q + 3) + P] != packed[P * nc + c]) return fail("column-major != row-major");
        for (int64_t P = Q; P < Q + 3; ++P)
            if (packed_t[c * (Q + 3) + P] != -7.0) return fail("column-major wrote past naux");
    }
    if (f2ints_coulomb_3c_pairs_t(ax, bs, 0, npairs, packed_t.data(), Q - 1) == 0) return fail("short ldo accepted");
    // aux-shell ranges that cut the runs of one l in different places
    std::vector<int> aoff(nsh_aux + 1, 0);
    for (int s = 0; s < nsh_aux; ++s) aoff[s + 1] = aoff[s] + 2 * aux.l[s] + 1;
    const int cuts[][2] = {{0, 3}, {3, 17}, {17, nsh_aux}, {5, 6}};
    for (const auto& r : cuts) {
        const int64_t rows = aoff[r[1]] - aoff[r[0]];
        std::vector<double> part(rows * nc);
        if (f2ints_coulomb_3c_pairs_aux(ax, bs, r[0], r[1], 0, npairs, part.data(), nc)) return fail("pairs_aux");
        std::vector<double> part_t(nc * rows);
        if (f2ints_coulomb_3c_pairs_aux_t(ax, bs, r[0], r[1], 0, npairs, part_t.data(), rows)) return fail("pairs_aux_t");
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t P = 0; P < rows; ++P)
                if (part_t[c * rows + P] != part[P * nc + c]) return fail("pairs_aux_t != pairs_aux");
        if (std::memcmp(part.data(), packed.data() + aoff[r[0]] * nc, sizeof(double) * rows * nc) != 0)
            return fail("pairs_aux != pairs");
    }
    std::vector<double> sum(N * N);
    if (f2ints_coulomb_3c_sum(ax, bs, sum.data())) return fail("coulomb_3c_sum");
    double maxdiff = 0.0, maxabs = 0.0;
    for (int64_t i = 0; i < N * N; ++i) {
        double ref = 0.0;
        for (int64_t P = 0; P < Q; ++P) ref += sq[P * N * N + i];
        maxdiff = std::fmax(maxdiff, std::fabs(ref - sum[i]));
        maxabs = std::fmax(maxabs, std::fabs(ref));
    }
    if (maxdiff > 1e-12 * maxabs) return fail("coulomb_3c_sum != sum of blocks");
    // the remaining entry points
    std::vector<double> eri(N * N * N * N);
    if (f2ints_coulomb_4c(bs, eri.data())) return fail("coulomb_4c");
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j < N; ++j)
            if (std::fabs(eri[((i * N + j) * N + j) * N + i] - eri[((j * N + i) * N + i) * N + j]) > 1e-13)
                return fail("4c symmetry");
    // E_nn = sum_{a<b} Z_a Z_b / R_ab, summed here by hand over the three centres
    double enn = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < a; ++b) {
            const double dx = centres[3 * a] - centres[3 * b], dy = centres[3 * a + 1] - centres[3 * b + 1],
                         dz = centres[3 * a + 2] - centres[3 * b + 2];
            enn += Z[a] * Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    if (std::fabs(f2ints_nuclear_repulsion(3, Z, centres) - enn) > 1e-13 * enn) return fail("nuclear_repulsion");
    double F[32];
    if (f2ints_boys(10, 3.7, F) || !(F[0] > F[10] && F[10] > 0.0)) return fail("boys");
    const int lmax = f2ints_max_l();
    std::vector<double> T3((size_t)(lmax + 1) * (lmax + 2) / 2 * (2 * lmax + 1));
    if (lmax < 4 || f2ints_cart2sph(lmax, T3.data())) return fail("cart2sph");
    if (f2ints_set_threads(2)) return fail("set_threads");
    std::vector<double> sq2(Q * N * N);
    if (f2ints_coulomb_3c(ax, bs, 0, nsh_aux, sq2.data()) || std::memcmp(sq2.data(), sq.data(), sizeof(double) * sq.size()))
        return fail("coulomb_3c with 2 threads");
    // error paths
    if (f2ints_coulomb_3c_pairs_aux(ax, bs, 4, 2, 0, npairs, packed.data(), nc) == 0) return fail("bad range accepted");
    if (f2ints_boys(-1, 1.0, F) == 0) return fail("bad Boys order accepted");
    f2ints_basis_destroy(bs);
    f2ints_basis_destroy(ax);
    std::printf("ok N=%lld Q=%lld packed=%lld\n", (long long)N, (long long)Q, (long long)nc);
    return 0;
}
