/* Test infrastructure: FP64 CPU restatement of TBTK's Chebyshev path.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, as the checker. The product path never calls it.
 *
 * Each routine cites the reference lines it restates. Complex arithmetic is
 * spelled out in real arithmetic in the order libstdc++'s std::complex<double>
 * performs it (a*b = (ar*br - ai*bi, ar*bi + ai*br); complex*real and
 * complex/real act componentwise), so FP64 results match the reference build
 * bit for bit when compiled with -O2 -ffp-contract=off and no -march.
 *
 * Pinned by: tests/test_oracle.py (golden vectors from the compiled reference,
 * closed forms of Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-565,
 * dense-eigensolver moments, sum rules).
 */
#include "kpm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b){
	cplx r = { a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re };
	return r;
}
static inline cplx cadd(cplx a, cplx b){ cplx r = { a.re + b.re, a.im + b.im }; return r; }
static inline cplx csub(cplx a, cplx b){ cplx r = { a.re - b.re, a.im - b.im }; return r; }
static inline cplx cscale(cplx a, double d){ cplx r = { a.re*d, a.im*d }; return r; }
static inline cplx cdivr(cplx a, double d){ cplx r = { a.re/d, a.im/d }; return r; }
static inline cplx cconj(cplx a){ cplx r = { a.re, -a.im }; return r; }
static inline cplx mk(double re, double im){ cplx r = { re, im }; return r; }

/* exp(-i*n*theta) exactly as std::exp(-i*((double)n)*acos(E)) evaluates it:
 * -i = (-0, -1); times n; times theta; then cexp. */
static inline cplx exp_minus_i_n_theta(int n, double theta){
	double xr = -0.0*(double)n, xi = -1.0*(double)n;
	xr = xr*theta; xi = xi*theta;
	double complex z = cexp(CMPLX(xr, xi));
	return mk(creal(z), cimag(z));
}

/* Lib/src/Solver/ChebyshevExpander.cpp:65-82 (addHamiltonianProduct). */
static void add_product(int64_t N, const uint32_t *rowptr, const uint32_t *col,
                        const cplx *v, const cplx *x, cplx *y){
	for(int64_t row = 0; row < N; row++)
		for(uint32_t k = rowptr[row]; k < rowptr[row + 1]; k++)
			y[row] = cadd(y[row], cmul(v[k], x[col[k]]));
}

/* SparseMatrix::operator*= with a complex scalar (d, 0),
 * Lib/include/Utilities/TBTK/SparseMatrix.h:1106-1125. */
static void scale_values(int64_t Z, cplx *v, double d){
	cplx s = mk(d, 0.0);
	for(int64_t k = 0; k < Z; k++)
		v[k] = cmul(v[k], s);
}

static cplx *scaled_copy(int64_t Z, const double *val, double scale){
	cplx *v = (cplx*)malloc((size_t)(Z > 0 ? Z : 1)*sizeof(cplx));
	memcpy(v, val, (size_t)Z*sizeof(cplx));
	scale_values(Z, v, 1/scale);   /* ChebyshevExpander.cpp:147 */
	return v;
}

/* Lib/src/Solver/ChebyshevExpander.cpp:103-188. The coefficient map keeps
 * the reference's semantics: if a to-index repeats, only its last position
 * receives moments (coefficientMap is overwritten, :125-126). */
void kpm_moments(int64_t N, const uint32_t *rowptr, const uint32_t *col,
                 const double *val, double scale, int64_t from, int K,
                 const int64_t *to, int M, double *out){
	int64_t Z = rowptr[N];
	cplx *v = scaled_copy(Z, val, scale);
	cplx *jIn1 = (cplx*)calloc((size_t)N, sizeof(cplx));
	cplx *jIn2 = (cplx*)calloc((size_t)N, sizeof(cplx));
	cplx *jRes = (cplx*)calloc((size_t)N, sizeof(cplx));
	int *map = (int*)malloc((size_t)N*sizeof(int));
	cplx *c = (cplx*)out;
	for(int64_t n = 0; n < N; n++) map[n] = -1;
	for(int k = 0; k < K; k++) map[to[k]] = k;
	for(int64_t k = 0; k < (int64_t)K*M; k++) c[k] = mk(0, 0);

	jIn1[from] = mk(1, 0);
	for(int64_t n = 0; n < N; n++)
		if(map[n] != -1) c[(int64_t)map[n]*M + 0] = jIn1[n];

	if(M > 1){
		add_product(N, rowptr, col, v, jIn1, jRes);
		cplx *t = jIn2; jIn2 = jIn1; jIn1 = jRes; jRes = t;
		for(int64_t n = 0; n < N; n++)
			if(map[n] != -1) c[(int64_t)map[n]*M + 1] = jIn1[n];
	}
	scale_values(Z, v, 2);          /* ChebyshevExpander.cpp:168 */
	for(int m = 2; m < M; m++){
		for(int64_t n = 0; n < N; n++) jRes[n] = mk(-jIn2[n].re, -jIn2[n].im);
		add_product(N, rowptr, col, v, jIn1, jRes);
		cplx *t = jIn2; jIn2 = jIn1; jIn1 = jRes; jRes = t;
		for(int64_t n = 0; n < N; n++)
			if(map[n] != -1) c[(int64_t)map[n]*M + m] = jIn1[n];
	}
	free(v); free(jIn1); free(jIn2); free(jRes); free(map);
}

/* Lib/src/Solver/ChebyshevExpander.cpp:190-196. */
void kpm_lorentz(int K, int M, double broadening, double *cc){
	if(broadening == 0) return;
	cplx *c = (cplx*)cc;
	double lambda = broadening*M;
	for(int n = 0; n < M; n++)
		for(int k = 0; k < K; k++)
			c[(int64_t)k*M + n] = cdivr(
				cscale(c[(int64_t)k*M + n], sinh(lambda*(1 - n/(double)M))),
				sinh(lambda));
}

/* Jackson kernel, Rev. Mod. Phys. 78, 275 (2006), eq. 71:
 * g_n = [(M-n+1) cos(pi n/(M+1)) + sin(pi n/(M+1)) cot(pi/(M+1))]/(M+1). */
void kpm_jackson(int M, double *g){
	double q = M_PI/(M + 1);
	for(int n = 0; n < M; n++)
		g[n] = ((M - n + 1)*cos(q*n) + sin(q*n)*cos(q)/sin(q))/(M + 1);
}

/* Range::operator[] with dx = (hi-lo)/(res-1) (Range.cpp:44, Range.h:130). */
static inline double range_at(double lo, double hi, int res, int e){
	double dx = (hi - lo)/(res - 1);
	return lo + e*dx;
}

/* Lib/src/Solver/ChebyshevExpander.cpp:256-268. */
void kpm_lookup_table(int M, int E, double lo, double hi, double scale,
                      double *LL){
	cplx *L = (cplx*)LL;
	for(int n = 0; n < M; n++){
		double denominator = 1.;
		if(n == 0) denominator = 2.;
		for(int e = 0; e < E; e++){
			double En = range_at(lo, hi, E, e)/scale;
			/* (1/s)*(-2.*i/sqrt(1 - E*E)) */
			cplx a = cdivr(mk(-2.*0.0, -2.*1.0), sqrt(1 - En*En));
			a = cscale(a, 1/scale);
			cplx x = exp_minus_i_n_theta(n, acos(En));
			L[(int64_t)n*E + e] = cdivr(cmul(a, x), denominator);
		}
	}
}

/* Lib/src/Solver/ChebyshevExpander.cpp:324-370. */
void kpm_gf_lookup(int type, int M, int E, const double *LL, const double *cc,
                   double *GG){
	const cplx *L = (const cplx*)LL, *c = (const cplx*)cc;
	cplx *G = (cplx*)GG;
	for(int e = 0; e < E; e++) G[e] = mk(0, 0);
	for(int n = 0; n < M; n++){
		for(int e = 0; e < E; e++){
			cplx l = L[(int64_t)n*E + e];
			switch(type){
			case KPM_RETARDED:
				G[e] = cadd(G[e], cmul(l, c[n]));
				break;
			case KPM_ADVANCED:
				G[e] = cadd(G[e], cmul(c[n], cconj(l)));
				break;
			case KPM_PRINCIPAL:
				G[e] = cadd(G[e], cscale(mk(-c[n].re, -c[n].im), l.re));
				break;
			case KPM_NONPRINCIPAL:
				G[e] = csub(G[e], cscale(cmul(c[n], mk(0, 1)), l.im));
				break;
			}
		}
	}
}

/* Lib/src/Solver/ChebyshevExpander.cpp:371-432. */
void kpm_gf_direct(int type, int M, int E, double lo, double hi,
                   double scale, const double *cc, double *GG){
	const cplx *c = (const cplx*)cc;
	cplx *G = (cplx*)GG;
	for(int e = 0; e < E; e++){
		cplx g = type == KPM_PRINCIPAL ? mk(0, 0) : cdivr(c[0], 2.);
		double En = range_at(lo, hi, E, e)/scale;
		double acosE = acos(En);
		switch(type){
		case KPM_RETARDED:
			for(int n = 1; n < M; n++)
				g = cadd(g, cmul(c[n], exp_minus_i_n_theta(n, acosE)));
			g = cmul(g, cdivr(mk(-2.*0.0, -2.*1.0), scale));
			break;
		case KPM_ADVANCED:
			for(int n = 1; n < M; n++)
				g = cadd(g, cconj(cmul(c[n], exp_minus_i_n_theta(n, acosE))));
			g = cmul(g, cdivr(mk(2.*0.0, 2.*1.0), scale));
			break;
		case KPM_PRINCIPAL:
			for(int n = 1; n < M; n++)
				g = cadd(g, cscale(c[n], sin(((double)n)*acosE)));
			g = cscale(g, 2./scale);
			break;
		case KPM_NONPRINCIPAL:
			for(int n = 1; n < M; n++)
				g = cadd(g, cscale(c[n], cos(((double)n)*acosE)));
			g = cmul(g, cdivr(mk(2.*0.0, 2.*1.0), scale));
			break;
		}
		G[e] = cdivr(g, sqrt(1 - En*En));
	}
}

void kpm_moments_trace(int64_t N, const uint32_t *rowptr, const uint32_t *col,
                       const double *val, double scale, int R,
                       const double *vv, int M, double *mu){
	int64_t Z = rowptr[N];
	cplx *v = scaled_copy(Z, val, scale);
	cplx *v2 = (cplx*)malloc((size_t)(Z > 0 ? Z : 1)*sizeof(cplx));
	memcpy(v2, v, (size_t)Z*sizeof(cplx));
	scale_values(Z, v2, 2);
	cplx *a = (cplx*)malloc((size_t)N*sizeof(cplx));
	cplx *b = (cplx*)malloc((size_t)N*sizeof(cplx));
	cplx *y = (cplx*)malloc((size_t)N*sizeof(cplx));
	const cplx *start = (const cplx*)vv;
	cplx *out = (cplx*)mu;
	for(int m = 0; m < M; m++) out[m] = mk(0, 0);
	for(int r = 0; r < R; r++){
		const cplx *s = start + (int64_t)r*N;
		/* a = T_{m-2} v, b = T_{m-1} v */
		for(int m = 0; m < M; m++){
			if(m == 0){
				memcpy(y, s, (size_t)N*sizeof(cplx));
			}else if(m == 1){
				for(int64_t n = 0; n < N; n++) y[n] = mk(0, 0);
				add_product(N, rowptr, col, v, b, y);
			}else{
				for(int64_t n = 0; n < N; n++) y[n] = mk(-a[n].re, -a[n].im);
				add_product(N, rowptr, col, v2, b, y);
			}
			cplx acc = mk(0, 0);
			for(int64_t n = 0; n < N; n++)
				acc = cadd(acc, cmul(cconj(s[n]), y[n]));
			out[m] = cadd(out[m], acc);
			cplx *t = a; a = b; b = y; y = t;
		}
	}
	free(v); free(v2); free(a); free(b); free(y);
}

/* splitmix64 finaliser over seed and counter; 53-bit uniform in [0, 1). */
double kpm_uniform(uint64_t seed, uint64_t counter){
	uint64_t z = seed*0x9E3779B97F4A7C15ull + counter*0xD1B54A32D192ED03ull
	             + 0x632BE59BD9B4E019ull;
	z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27))*0x94D049BB133111EBull;
	z = z ^ (z >> 31);
	return (double)(z >> 11)*(1.0/9007199254740992.0);
}
