/*
 * f2ints.h -- C ABI of the host integral engine (libf2ints.so).
 *
 * forte2 evaluates its one-time integrals with libint2 / libcint behind
 * nanobind (forte2/integrals/integrals.py, two_electron_compute.h); those
 * libraries are not part of this build, so this engine provides the same
 * integrals with the same AO conventions:
 *   - real solid-harmonic (pure) shells, components ordered m = -l..l with
 *     the phases of forte2/integrals/real_spherical_harmonics.hpp:33-60
 *     (p: y, z, x; d: xy, yz, z^2-(x^2+y^2)/2, xz, (x^2-y^2)*sqrt(3)/2, ...);
 *   - every component normalised like the axial Cartesian z^l;
 *   - contraction coefficients given with the primitive normalisation
 *     already embedded (the Python basis builder does that), or raw when
 *     the caller asks for unnormalised primitives (SAP density basis).
 * Results are row-major FP64 arrays owned by the caller.
 *
 * Replaces:
 *   integrals.overlap   (forte2/integrals/integrals.py:198)      -> f2ints_overlap
 *   integrals.kinetic   (forte2/integrals/integrals.py:224)      -> f2ints_kinetic
 *   integrals.nuclear   (forte2/integrals/integrals.py:250)      -> f2ints_nuclear
 *   integrals.nuclear_repulsion (integrals.py:156)               -> f2ints_nuclear_repulsion
 *   integrals.coulomb_2c (integrals.py:903-934, two_electron_compute.h:434-502) -> f2ints_coulomb_2c
 *   integrals.coulomb_3c (integrals.py:811-862, two_electron_compute.h:108-253) -> f2ints_coulomb_3c
 *   ints.coulomb_4c (used by tests/integrals/test_two_electron.py) -> f2ints_coulomb_4c
 */
#ifndef F2INTS_H
#define F2INTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct f2ints_basis f2ints_basis;

/* Build a basis of nshell pure shells.  Shell s has angular momentum l[s],
 * nprim[s] primitives (exponents/coefs taken in order from the flat arrays)
 * and centre xyz[3s..3s+2] (bohr).  Coefficients are used as given. */
int f2ints_basis_create(int nshell, const int* l, const int* nprim, const double* xyz,
                        const double* exps, const double* coefs, f2ints_basis** out);
int f2ints_basis_destroy(f2ints_basis* b);
int64_t f2ints_basis_nbf(const f2ints_basis* b);
int f2ints_max_l(void);

/* One-electron integrals, nbf x nbf. */
int f2ints_overlap(const f2ints_basis* b, double* S);
int f2ints_kinetic(const f2ints_basis* b, double* T);
/* sum over point charges Z_c at xyz_c of -Z_c <m|1/|r - xyz_c||n> */
int f2ints_nuclear(const f2ints_basis* b, int ncharge, const double* Z, const double* xyz, double* V);
/* sum_{a<b} Z_a Z_b / R_ab */
double f2ints_nuclear_repulsion(int natom, const double* Z, const double* xyz);

/* (P|Q) for P in a, Q in b: na x nb. */
int f2ints_coulomb_2c(const f2ints_basis* a, const f2ints_basis* b, double* M);

/* (P|mn) for aux shells [shell_begin, shell_end) of `aux` and all pairs of
 * `bs`: written as out[P - first][m][n] (first = first function of
 * shell_begin), both (m,n) and (n,m) filled.  shell_end < 0 means all. */
int f2ints_coulomb_3c(const f2ints_basis* aux, const f2ints_basis* bs, int shell_begin,
                      int shell_end, double* out);

/* The significant shell pairs (a >= b) of `bs`, in the order the 3-centre
 * routines use: returns their number and, when a and b are not NULL, writes
 * the first min(count, cap) of them.  -1 on error.  Used to stream B by
 * pair blocks (the mn-blocks of forte2/jkbuilder/jkbuilder.py:135-159). */
int64_t f2ints_ket_pairs(const f2ints_basis* bs, int* a, int* b, int64_t cap);

/* (P|ab) for every aux function and the significant pairs [pair_begin,
 * pair_end), packed: out[P * ldo + c], where c runs over the pairs in order.
 * A pair a > b has nsph(la) * nsph(lb) columns (row-major in the two shells'
 * components); a pair a == b has the n (n + 1) / 2 components sa >= sb
 * (row-major), n = nsph(la).  So every (m, n) with m >= n appears once. */
int f2ints_coulomb_3c_pairs(const f2ints_basis* aux, const f2ints_basis* bs, int64_t pair_begin,
                            int64_t pair_end, double* out, int64_t ldo);

/* The same packed (P|ab) stored column-major, out[c * ldo + P] (ldo >= naux):
 * each packed column (one (m, n), m >= n) is a contiguous row of naux values.
 * The device B formation multiplies these rows by L^-1 as they are. */
int f2ints_coulomb_3c_pairs_t(const f2ints_basis* aux, const f2ints_basis* bs, int64_t pair_begin,
                              int64_t pair_end, double* out, int64_t ldo);

/* The same packed (P|ab) for the aux shells [shell_begin, shell_end) only:
 * row 0 of out is the first function of shell_begin.  Used by the
 * integral-direct builder for its aux-shell blocks of (P|mn)
 * (forte2/jkbuilder/jkbuilder.py:676-740, 813-875). */
int f2ints_coulomb_3c_pairs_aux(const f2ints_basis* aux, const f2ints_basis* bs, int shell_begin, int shell_end,
                                int64_t pair_begin, int64_t pair_end, double* out, int64_t ldo);

/* f2ints_coulomb_3c_pairs_aux stored column-major: out[c * ldo + P - first]
 * (ldo >= the rows of those shells), one row per packed column. */
int f2ints_coulomb_3c_pairs_aux_t(const f2ints_basis* aux, const f2ints_basis* bs, int shell_begin,
                                  int shell_end, int64_t pair_begin, int64_t pair_end, double* out, int64_t ldo);

/* V[m][n] = sum over all functions P of `aux` of (P|mn), nbf x nbf: the
 * superposition-of-atomic-potentials guess potential of
 * forte2/scf/scf_utils.py:10-51 (aux = the SAP charge-density basis), without
 * forming (P|mn) per P. */
int f2ints_coulomb_3c_sum(const f2ints_basis* aux, const f2ints_basis* bs, double* out);

/* (mn|rs) chemist's notation, nbf^4 -- small bases only (tests). */
int f2ints_coulomb_4c(const f2ints_basis* b, double* out);

/* Boys function F_n(T), n = 0..nmax, written to out[0..nmax]. */
int f2ints_boys(int nmax, double T, double* out);

/* Cartesian -> solid-harmonic coefficients for angular momentum l:
 * (2l+1) x (l+1)(l+2)/2, Cartesian components ordered x^a y^b z^c with a
 * descending, then b descending. */
int f2ints_cart2sph(int l, double* out);

/* Threads used by the integral loops (default: all cores). */
int f2ints_set_threads(int n);

const char* f2ints_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* F2INTS_H */
