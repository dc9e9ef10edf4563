// Lookup table and FP64 Green's-function contraction.
//
// Table (Lib/src/Solver/ChebyshevExpander.cpp:256-268), E~ = E_e/s:
//   L[n][e] = (1/s) * (-2i/sqrt(1 - E~^2)) * exp(-i n acos E~) / (n == 0 ? 2 : 1)
// Contraction (Lib/src/Solver/ChebyshevExpander.cpp:324-370), n ascending:
//   Retarded      G += L*c            Advanced     G += c*conj(L)
//   Principal     G += -c*Re(L)       NonPrincipal G -= (c*i)*Im(L)
// The FP64 kernel keeps that order and operation sequence (no fused
// multiply-add), so with an uploaded host table it reproduces the reference
// bit for bit; with the device-generated table the difference is that of
// the device acos/sqrt/sincos against the host libm (a few ulp).
//
// Traffic: the table is read once per coefficient vector by a thread per
// energy (16 B per (n, e)); it is an FP64 GEMV/GEMM of 8 K M E flops.
#include "common.hpp"
#include "host_math.hpp"

namespace tbtkx {
namespace {

__global__ __launch_bounds__(kThreads) void lut_generate(
    double2 *__restrict__ L, int M, int E, double lo, double dx, double s) {
#pragma clang fp contract(off)
	const int e = blockIdx.x * kThreads + threadIdx.x;
	const int n = blockIdx.y;
	if (e >= E) return;
	const double En = (lo + e * dx) / s;
	const double r = sqrt(1 - En * En);
	// (1/s) * ((-0, -2) / r)
	const double ar = (-2.0 * 0.0 / r) * (1 / s);
	const double ai = (-2.0 / r) * (1 / s);
	double sn, cs;
	sincos((-1.0 * (double)n) * acos(En), &sn, &cs);
	const double denom = n == 0 ? 2.0 : 1.0;
	L[(int64_t)n * E + e] = make_double2((ar * cs - ai * sn) / denom,
	                                     (ar * sn + ai * cs) / denom);
}

// Damping of K coefficient vectors C[k*M + n] in FP64, in place; one thread
// per moment n computes its factor once and walks the vectors.
//   Lorentz (Lib/src/Solver/ChebyshevExpander.cpp:190-196; the reference's
//   GPU path damps on the device too, Lib/src/cuda/ChebyshevSolver.cu:632-645):
//     c *= sinh(lambda (1 - n/M)); c /= sinh(lambda), lambda = broadening*M,
//     real and imaginary part each, as std::complex<double> * and / double do.
//   Jackson (Weisse et al., Rev. Mod. Phys. 78, 275 (2006), eq. 71):
//     g_n = ((M - n + 1) cos(q n) + sin(q n) cos(q) / sin(q)) / (M + 1),
//     q = pi/(M + 1); c *= g_n.
__global__ __launch_bounds__(kThreads) void damp_coeffs(double2 *__restrict__ C, int K, int M,
                                                        int kind, double broadening) {
#pragma clang fp contract(off)
	const int n = blockIdx.x * kThreads + threadIdx.x;
	if (n >= M) return;
	double num, den;
	if (kind == TBTKX_DAMP_LORENTZ) {
		const double lambda = broadening * M;
		num = sinh(lambda * (1 - n / (double)M));
		den = sinh(lambda);
	} else {
		const double q = M_PI / (M + 1);
		num = (double)(M - n + 1) * cos(q * n) + sin(q * n) * cos(q) / sin(q);
		den = (double)(M + 1);
	}
	const double g = num / den;
	for (int k = blockIdx.y; k < K; k += gridDim.y) {
		double2 c = C[(int64_t)k * M + n];
		if (kind == TBTKX_DAMP_LORENTZ) {
			c.x = (c.x * num) / den;
			c.y = (c.y * num) / den;
		} else {
			c.x = c.x * g;
			c.y = c.y * g;
		}
		C[(int64_t)k * M + n] = c;
	}
}

// One term of the contraction, in the reference's operation order.
template <int TYPE>
__device__ __forceinline__ void contract_term(double2 l, double2 cn, double &gr, double &gi) {
#pragma clang fp contract(off)
	if (TYPE == TBTKX_RETARDED) {
		gr = gr + (l.x * cn.x - l.y * cn.y);
		gi = gi + (l.x * cn.y + l.y * cn.x);
	} else if (TYPE == TBTKX_ADVANCED) {
		gr = gr + (cn.x * l.x - cn.y * (-l.y));
		gi = gi + (cn.x * (-l.y) + cn.y * l.x);
	} else if (TYPE == TBTKX_PRINCIPAL) {
		gr = gr + (-cn.x) * l.x;
		gi = gi + (-cn.y) * l.x;
	} else {
		// (c * i) = (cr*0 - ci*1, cr*1 + ci*0), then times Im(L)
		const double pr = cn.x * 0.0 - cn.y * 1.0;
		const double pi = cn.x * 1.0 + cn.y * 0.0;
		gr = gr - pr * l.y;
		gi = gi - pi * l.y;
	}
}

// A thread per (energy, coefficient vector). A few vectors of few energies
// leave little parallelism, so each thread issues kAhead table loads before
// it adds them, in order.
constexpr int kAhead = 16;

template <int TYPE>
__global__ __launch_bounds__(kThreads) void contract_fp64(
    const double2 *__restrict__ L, const double2 *__restrict__ C,
    double2 *__restrict__ G, int M, int E) {
	const int e = blockIdx.x * kThreads + threadIdx.x;
	const int k = blockIdx.y;
	if (e >= E) return;
	const double2 *c = C + (int64_t)k * M;
	const double2 *l = L + e;
	double gr = 0.0, gi = 0.0;
	int n = 0;
	for (; n + kAhead <= M; n += kAhead) {
		double2 lv[kAhead];
#pragma unroll
		for (int i = 0; i < kAhead; i++) lv[i] = l[(int64_t)(n + i) * E];
#pragma unroll
		for (int i = 0; i < kAhead; i++) contract_term<TYPE>(lv[i], c[n + i], gr, gi);
	}
	for (; n < M; n++) contract_term<TYPE>(l[(int64_t)n * E], c[n], gr, gi);
	G[(int64_t)k * E + e] = make_double2(gr, gi);
}

// Direct evaluation without a table, Lib/src/Solver/ChebyshevExpander.cpp:
// 371-432: per energy E~ = E/s, theta = acos(E~),
//   Retarded     (c0/2 + sum_n c_n e^{-in theta}) * (-2i/s)
//   Advanced     (c0/2 + sum_n conj(c_n e^{-in theta})) * (2i/s)
//   Principal    (sum_n c_n sin(n theta)) * (2/s)
//   NonPrincipal (c0/2 + sum_n c_n cos(n theta)) * (2i/s)
// all divided by sqrt(1 - E~^2); energies e0 + e*de as Range::operator[].
// The trigonometric factors come from the host's libm (direct_factors), the
// functions the reference calls, so the sums below reproduce its results bit
// for bit: X holds e^{-in theta} as (re, im) pairs (Retarded, Advanced),
// sin(n theta) (Principal) or cos(n theta) (NonPrincipal) at X[(n-1)*E + e],
// and den[e] = sqrt(1 - E~^2).
template <int TYPE>
__device__ __forceinline__ void direct_term(double2 cn, double2 x, double &gr, double &gi) {
#pragma clang fp contract(off)
	if (TYPE == TBTKX_RETARDED || TYPE == TBTKX_ADVANCED) {
		const double pr = cn.x * x.x - cn.y * x.y;
		const double pi = cn.x * x.y + cn.y * x.x;
		gr = gr + pr;
		gi = TYPE == TBTKX_RETARDED ? gi + pi : gi + (-pi);
	} else {
		gr = gr + cn.x * x.x;
		gi = gi + cn.y * x.x;
	}
}

template <int TYPE>
__global__ __launch_bounds__(kThreads) void contract_direct(
    const double2 *__restrict__ C, const double *__restrict__ X,
    const double *__restrict__ den, double2 *__restrict__ G, int M, int E, double s) {
#pragma clang fp contract(off)
	constexpr bool kPair = TYPE == TBTKX_RETARDED || TYPE == TBTKX_ADVANCED;
	const int e = blockIdx.x * kThreads + threadIdx.x;
	const int k = blockIdx.y;
	if (e >= E) return;
	const double2 *c = C + (int64_t)k * M;
	auto factor = [&](int n) {
		const int64_t at = (int64_t)(n - 1) * E + e;
		return kPair ? reinterpret_cast<const double2 *>(X)[at] : make_double2(X[at], 0.0);
	};
	double gr, gi;
	if (TYPE == TBTKX_PRINCIPAL) {
		gr = 0.0;
		gi = 0.0;
	} else {
		gr = c[0].x / 2.;
		gi = c[0].y / 2.;
	}
	int n = 1;
	for (; n + kAhead <= M; n += kAhead) {
		double2 xv[kAhead];
#pragma unroll
		for (int i = 0; i < kAhead; i++) xv[i] = factor(n + i);
#pragma unroll
		for (int i = 0; i < kAhead; i++) direct_term<TYPE>(c[n + i], xv[i], gr, gi);
	}
	for (; n < M; n++) direct_term<TYPE>(c[n], factor(n), gr, gi);
	double fr, fi;   // the type's prefactor
	if (TYPE == TBTKX_RETARDED) {
		fr = -2. * 0.0 / s;
		fi = -2. / s;
	} else if (TYPE == TBTKX_PRINCIPAL) {
		fr = 2. / s;
		fi = 0.0;
	} else {
		fr = 2. * 0.0 / s;
		fi = 2. / s;
	}
	double hr, hi;
	if (TYPE == TBTKX_PRINCIPAL) {
		hr = gr * fr;
		hi = gi * fr;
	} else {
		hr = gr * fr - gi * fi;
		hi = gr * fi + gi * fr;
	}
	const double r = den[e];
	G[(int64_t)k * E + e] = make_double2(hr / r, hi / r);
}

}  // namespace

// Applies the handle's damping (tbtkx_set_damping) to K x M coefficients
// already in device memory, on the handle's stream; no-op without one.
int apply_damping(tbtkx_context *h, double2 *C, int K, int M) {
	if (h->damp_kind == TBTKX_DAMP_NONE || K == 0) return TBTKX_OK;
	if (h->damp_kind == TBTKX_DAMP_LORENTZ && h->damp_broadening == 0) return TBTKX_OK;
	dim3 grid((M + kThreads - 1) / kThreads, K < 1024 ? K : 1024);
	damp_coeffs<<<grid, kThreads, 0, h->stream>>>(C, K, M, h->damp_kind, h->damp_broadening);
	TBTKX_CHECK_HIP(hipGetLastError());
	return TBTKX_OK;
}

// The factors of the direct contraction for one Green's function kind (0:
// e^{-in theta} for Retarded and Advanced, 1: sin(n theta), 2: cos(n theta))
// over n = 1 .. M-1 and the energies e0 + e*de, computed on the host with the
// expressions of generateGreensFunctionCPU (host_math.cpp), so the device
// sums use the reference's values. Kept on the device until another kind,
// order or grid is asked for.
static int direct_factors(tbtkx_context *h, int kind, int M, int E, double e0, double de,
                          double s) {
	if (h->dfac_kind == kind && h->dfac_M == M && h->dfac_E == E && h->dfac_e0 == e0 &&
	    h->dfac_de == de && h->dfac_s == s)
		return TBTKX_OK;
	h->dfac_kind = -1;
	const size_t per = kind == 0 ? 2 : 1;
	const size_t rows = M > 1 ? (size_t)(M - 1) : 0;
	std::vector<double> fac(rows * E * per + per), den(E);
	direct_factor_table(kind, M, E, e0, de, s, fac.data(), den.data());
	int rc;
	if ((rc = h->dfac.reserve(fac.size() * 8))) return rc;
	if ((rc = h->dden.reserve((size_t)E * 8))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->dfac.ptr, fac.data(), fac.size() * 8,
	                               hipMemcpyHostToDevice, h->stream));
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->dden.ptr, den.data(), (size_t)E * 8,
	                               hipMemcpyHostToDevice, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->dfac_kind = kind;
	h->dfac_M = M;
	h->dfac_E = E;
	h->dfac_e0 = e0;
	h->dfac_de = de;
	h->dfac_s = s;
	return TBTKX_OK;
}

}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

int tbtkx_lut_upload(tbtkx_handle h, int M, int E, const double *table) {
	TBTKX_REQUIRE(h && table, TBTKX_ERR_ARG, "tbtkx_lut_upload: null pointer");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(M > 0 && E > 0, TBTKX_ERR_ARG,
	              "tbtkx_lut_upload: table dimensions must be positive");
	int rc = set_device(h);
	if (rc) return rc;
	if ((rc = h->lut.reserve((size_t)M * E * 16))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->lut.ptr, table, (size_t)M * E * 16,
	                               hipMemcpyHostToDevice, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->lut_M = M;
	h->lplanes_valid = 0;
	h->lut_E = E;
	return TBTKX_OK;
}

int tbtkx_lut_generate(tbtkx_handle h, int M, int E, double lo, double hi,
                       double s) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_lut_generate: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(M > 0, TBTKX_ERR_ARG,
	              "tbtkx_lut_generate: numCoefficients has to be larger than 0");
	TBTKX_REQUIRE(E > 1, TBTKX_ERR_ARG,
	              "tbtkx_lut_generate: the energy resolution has to be larger than 1");
	TBTKX_REQUIRE(lo < hi, TBTKX_ERR_ARG,
	              "tbtkx_lut_generate: the lower bound has to be smaller than the upper bound");
	TBTKX_REQUIRE(s > 0 && lo > -s && hi < s, TBTKX_ERR_ARG,
	              "tbtkx_lut_generate: the energy window has to lie inside (-s, s)");
	int rc = set_device(h);
	if (rc) return rc;
	if ((rc = h->lut.reserve((size_t)M * E * 16))) return rc;
	const double dx = (hi - lo) / (E - 1);
	dim3 grid((E + kThreads - 1) / kThreads, M);
	lut_generate<<<grid, kThreads, 0, h->stream>>>(h->lut.as<double2>(), M, E,
	                                               lo, dx, s);
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->lut_M = M;
	h->lplanes_valid = 0;
	h->lut_E = E;
	return TBTKX_OK;
}

int tbtkx_lut_destroy(tbtkx_handle h) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_lut_destroy: null handle");
	TBTKX_LOCK(h);
	int rc = set_device(h);
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->lut.release();
	h->lut_M = h->lut_E = 0;
	h->lplanes_valid = 0;
	return TBTKX_OK;
}

int tbtkx_set_damping(tbtkx_handle h, int kind, double broadening) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_set_damping: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(kind == TBTKX_DAMP_NONE || kind == TBTKX_DAMP_LORENTZ ||
	                  kind == TBTKX_DAMP_JACKSON,
	              TBTKX_ERR_ARG, "tbtkx_set_damping: unknown damping kind");
	TBTKX_REQUIRE(std::isfinite(broadening), TBTKX_ERR_ARG,
	              "tbtkx_set_damping: the broadening has to be finite");
	h->damp_kind = kind;
	h->damp_broadening = kind == TBTKX_DAMP_LORENTZ ? broadening : 0.0;
	return TBTKX_OK;
}

int tbtkx_contract(tbtkx_handle h, int K, const double *coeffs, int type,
                   double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_contract: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->lut_M > 0, TBTKX_ERR_STATE,
	              "tbtkx_contract: no lookup table loaded");
	TBTKX_REQUIRE(K >= 0 && (K == 0 || (coeffs && out)), TBTKX_ERR_ARG,
	              "tbtkx_contract: bad coefficient arrays");
	TBTKX_REQUIRE(type >= TBTKX_ADVANCED && type <= TBTKX_NONPRINCIPAL,
	              TBTKX_ERR_ARG, "tbtkx_contract: unknown Green's function type");
	TBTKX_REQUIRE(K <= 65535, TBTKX_ERR_ARG,
	              "tbtkx_contract: at most 65535 vectors per call");
	if (K == 0) return TBTKX_OK;
	const int M = h->lut_M, E = h->lut_E;
	int rc = set_device(h);
	if (rc) return rc;
	if ((rc = h->gf_in.reserve((size_t)K * M * 16))) return rc;
	if ((rc = h->gf_out.reserve((size_t)K * E * 16))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->gf_in.ptr, coeffs, (size_t)K * M * 16,
	                               hipMemcpyHostToDevice, h->stream));
	if ((rc = apply_damping(h, h->gf_in.as<double2>(), K, M))) return rc;
	dim3 grid((E + kThreads - 1) / kThreads, K);
	const double2 *L = h->lut.as<double2>(), *C = h->gf_in.as<double2>();
	double2 *G = h->gf_out.as<double2>();
	switch (type) {
	case TBTKX_RETARDED:
		contract_fp64<TBTKX_RETARDED><<<grid, kThreads, 0, h->stream>>>(L, C, G, M, E);
		break;
	case TBTKX_ADVANCED:
		contract_fp64<TBTKX_ADVANCED><<<grid, kThreads, 0, h->stream>>>(L, C, G, M, E);
		break;
	case TBTKX_PRINCIPAL:
		contract_fp64<TBTKX_PRINCIPAL><<<grid, kThreads, 0, h->stream>>>(L, C, G, M, E);
		break;
	default:
		contract_fp64<TBTKX_NONPRINCIPAL><<<grid, kThreads, 0, h->stream>>>(L, C, G, M, E);
		break;
	}
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipMemcpyAsync(out, h->gf_out.ptr, (size_t)K * E * 16,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	return TBTKX_OK;
}

int tbtkx_contract_direct(tbtkx_handle h, int K, int M, const double *coeffs, int type,
                          int E, double e0, double de, double s, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_contract_direct: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(K >= 0 && K <= 65535 && M > 0 && E > 0 && (K == 0 || (coeffs && out)),
	              TBTKX_ERR_ARG, "tbtkx_contract_direct: bad sizes or arrays");
	TBTKX_REQUIRE(type >= TBTKX_ADVANCED && type <= TBTKX_NONPRINCIPAL, TBTKX_ERR_ARG,
	              "tbtkx_contract_direct: unknown Green's function type");
	TBTKX_REQUIRE(s > 0 && e0 > -s && e0 + (E - 1) * de < s, TBTKX_ERR_ARG,
	              "tbtkx_contract_direct: the energy window has to lie inside (-s, s)");
	if (K == 0) return TBTKX_OK;
	int rc = set_device(h);
	if (rc) return rc;
	if ((rc = h->gf_in.reserve((size_t)K * M * 16))) return rc;
	if ((rc = h->gf_out.reserve((size_t)K * E * 16))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->gf_in.ptr, coeffs, (size_t)K * M * 16,
	                               hipMemcpyHostToDevice, h->stream));
	if ((rc = apply_damping(h, h->gf_in.as<double2>(), K, M))) return rc;
	const int kind = type == TBTKX_PRINCIPAL ? 1 : type == TBTKX_NONPRINCIPAL ? 2 : 0;
	if ((rc = direct_factors(h, kind, M, E, e0, de, s))) return rc;
	dim3 grid((E + kThreads - 1) / kThreads, K);
	const double2 *C = h->gf_in.as<double2>();
	const double *X = h->dfac.as<double>(), *D = h->dden.as<double>();
	double2 *G = h->gf_out.as<double2>();
	switch (type) {
	case TBTKX_RETARDED:
		contract_direct<TBTKX_RETARDED><<<grid, kThreads, 0, h->stream>>>(C, X, D, G, M, E, s);
		break;
	case TBTKX_ADVANCED:
		contract_direct<TBTKX_ADVANCED><<<grid, kThreads, 0, h->stream>>>(C, X, D, G, M, E, s);
		break;
	case TBTKX_PRINCIPAL:
		contract_direct<TBTKX_PRINCIPAL><<<grid, kThreads, 0, h->stream>>>(C, X, D, G, M, E, s);
		break;
	default:
		contract_direct<TBTKX_NONPRINCIPAL><<<grid, kThreads, 0, h->stream>>>(C, X, D, G, M, E, s);
		break;
	}
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipMemcpyAsync(out, h->gf_out.ptr, (size_t)K * E * 16,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	return TBTKX_OK;
}

}  // extern "C"
