/* Test infrastructure: FP64 CPU restatement of TBTK's Chebyshev path.
 *
 * This is the oracle the HIP path is checked against. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker; the product (tbtk_amd) never calls it.
 *
 * Complex numbers are interleaved (re, im) doubles. Arrays are row-major:
 * moments out[k*M + n], lookup table L[n*E + e], Green's function G[e].
 * The arithmetic order of every routine follows the reference line by line so
 * that FP64 results agree bit for bit (build with -O2 -ffp-contract=off and
 * no -march, as the reference build in oracle/Makefile.ref).
 */
#ifndef KPM_ORACLE_H
#define KPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Green's function types, numbered as Solver::ChebyshevExpander::Type
 * (Lib/include/TBTK/Solver/ChebyshevExpander.h:206-211). */
enum { KPM_ADVANCED = 0, KPM_RETARDED = 1, KPM_PRINCIPAL = 2, KPM_NONPRINCIPAL = 3 };

/* Raw Chebyshev moments c[k][n] = <to_k| T_n(H/s) |from>, n < M.
 * Follows calculateCoefficientsCPU, Lib/src/Solver/ChebyshevExpander.cpp:103-188
 * (without the Lorentz step, see kpm_lorentz). */
void kpm_moments(int64_t N, const uint32_t *rowptr, const uint32_t *col,
                 const double *val, double scale, int64_t from, int K,
                 const int64_t *to, int M, double *out);

/* Lorentz damping as applied by the reference when broadening != 0,
 * Lib/src/Solver/ChebyshevExpander.cpp:190-196. In place on c[K][M]. */
void kpm_lorentz(int K, int M, double broadening, double *c);

/* Jackson kernel g_n, n < M (Weisse, Wellein, Alvermann, Fehske,
 * Rev. Mod. Phys. 78, 275 (2006), eq. 71). Not in the reference. */
void kpm_jackson(int M, double *g);

/* Lookup table L[n][e], Lib/src/Solver/ChebyshevExpander.cpp:256-268, with the
 * energy grid of Range(lo, hi, E) (Lib/src/Utilities/Range.cpp:30-57). */
void kpm_lookup_table(int M, int E, double lo, double hi, double scale,
                      double *L);

/* G[e] from coefficients c[M] through the table,
 * Lib/src/Solver/ChebyshevExpander.cpp:324-370. */
void kpm_gf_lookup(int type, int M, int E, const double *L, const double *c,
                   double *G);

/* G[e] evaluated directly, Lib/src/Solver/ChebyshevExpander.cpp:371-432. */
void kpm_gf_direct(int type, int M, int E, double lo, double hi,
                   double scale, const double *c, double *G);

/* Stochastic-trace moments mu_n = sum_r <r| T_n(H/s) |r>, n < M, for R
 * vectors given as v[r*N + i] (complex). Same recurrence as kpm_moments;
 * the dot product is accumulated sequentially over i in FP64. */
void kpm_moments_trace(int64_t N, const uint32_t *rowptr, const uint32_t *col,
                       const double *val, double scale, int R,
                       const double *v, int M, double *mu);

/* The random-phase start vectors used by the stochastic trace:
 * v[i] = exp(i*2*pi*u), u = kpm_uniform(seed, vec*N + i). */
double kpm_uniform(uint64_t seed, uint64_t counter);

#ifdef __cplusplus
}
#endif
#endif
