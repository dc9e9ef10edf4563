// Batched path: blocks of B start vectors through one SpMM recurrence
//   Phi_1 = (H/s) Phi_0,   Phi_{n+1} = (2H/s) Phi_n - Phi_{n-1}
// with Phi stored row-major (N rows x B columns, each row contiguous) in FP32
// complex (or FP64 with TBTKX_VEC_FP64), H in FP32 (or FP64 with
// TBTKX_H_FP64), and every product and row sum formed in FP64 and rounded once
// on store.
//
// Moments (diagonal, per column b):
//  - gather mode: mu_n[b] = Phi_n[site_b][b], one recurrence step per moment;
//  - doubling mode: per step the dot products D0_n = <Phi_n|Phi_n> and
//    D1_n = <Phi_{n+1}|Phi_n> are fused into the SpMM pass (FP64 sums), and
//    mu_0 = D0_0, mu_1 = D1_0, mu_2n = 2 D0_n - mu_0, mu_2n+1 = 2 D1_n - mu_1
//    (Weisse et al., Rev. Mod. Phys. 78, 275 (2006), eq. 31), so M moments
//    take ceil(M/2) steps.
//
// Work mapping: one 64-lane wavefront per row; lane l owns B/64 consecutive
// columns (B = 64, 128, 256), so each non-zero streams one contiguous B-wide
// row of Phi_n (coalesced); column index and value are wave-uniform scalar
// loads. Each wavefront walks a contiguous range of rows; the workgroup id is
// remapped so every XCD owns one contiguous slice of the rows.
//
// Algorithmic HBM bytes per step (FP32 vectors, FP32 H):
//   12 Z (value + column) + 4 N (row pointer) + 24 B N (read Phi_n, read
//   Phi_{n-1}, write Phi_{n+1}, 8 B each per element).
#include <cmath>
#include <map>
#include <cstring>
#include <type_traits>

#include "common.hpp"

namespace tbtkx {
namespace {

constexpr int kWaves = kThreads / 64;
constexpr int kDotChunk = 128;     // steps whose dot partials are kept

// A real H's value in the 8 bytes of an FP32 complex entry: one FP64 value
// (FP32 rounding of H moves its eigenvalues by ~1e-7 of the band, which
// thousands of moments resolve). Its imaginary part is zero.
struct RealD {
	double x;
	static constexpr double y = 0.0;
};

// 8-byte H entries (FP32 complex, or FP64 real): the copies v1f/v2f and
// ell_v1f/ell_v2f.
template <typename HT>
constexpr bool kHt8 = std::is_same<HT, float2>::value || std::is_same<HT, RealD>::value;

template <typename T> struct C2;
template <> struct C2<float2> {
	__device__ static double2 load(const float2 &a) {
		return make_double2((double)a.x, (double)a.y);
	}
	__device__ static float2 store(double2 a) {
		return make_float2((float)a.x, (float)a.y);
	}
};
template <> struct C2<double2> {
	__device__ static double2 load(const double2 &a) { return a; }
	__device__ static double2 store(double2 a) { return a; }
};

// Packed real vectors: with a real H and real start vectors the real and
// imaginary parts of a complex column evolve independently, so column c
// carries two real vectors, vector 2c in the real part and 2c + 1 in the
// imaginary part. The step kernels are the same; only the start block, the
// moment gather and the doubling dots see the packing.

// Moment of live vector g at its site.
template <typename VT>
__device__ __forceinline__ double2 site_moment(const VT *__restrict__ X,
                                               const int64_t *__restrict__ sites, int g,
                                               int B, int packed) {
	if (!packed) return C2<VT>::load(X[sites[g] * B + g]);
	const double2 v = C2<VT>::load(X[sites[g] * B + (g >> 1)]);
	return make_double2((g & 1) ? v.y : v.x, 0.0);
}

// Doubling dots of one element, FP64: d0 += |x|^2 and d1 += conj(y) x for a
// complex column; per component for a packed column, d0 += (x.re^2, x.im^2)
// and d1 += (y.re x.re, y.im x.im).
__device__ __forceinline__ void dot_update(double2 x, double2 y, int packed, double2 &d0,
                                           double2 &d1) {
	if (packed) {
		d0.x = fma(x.x, x.x, d0.x);
		d0.y = fma(x.y, x.y, d0.y);
		d1.x = fma(y.x, x.x, d1.x);
		d1.y = fma(y.y, x.y, d1.y);
	} else {
		d0.x = fma(x.x, x.x, fma(x.y, x.y, d0.x));
		d1.x = fma(y.x, x.x, fma(y.y, x.y, d1.x));
		d1.y = fma(y.x, x.y, fma(-y.y, x.x, d1.y));
	}
}

// A lane's CPL consecutive complex values of one block row, moved with the
// widest aligned accesses (16 B): FP32 pairs as float4, FP64 as double2.
template <typename VT, int CPL> struct Row {
	__device__ static void load(const VT *p, VT (&o)[CPL]) {
#pragma unroll
		for (int j = 0; j < CPL; j++) o[j] = p[j];
	}
	__device__ static void store(VT *p, const VT (&o)[CPL]) {
#pragma unroll
		for (int j = 0; j < CPL; j++) p[j] = o[j];
	}
};
template <int CPL> struct Row<float2, CPL> {
	static_assert(CPL == 1 || CPL % 2 == 0, "FP32 rows move in pairs");
	__device__ static void load(const float2 *p, float2 (&o)[CPL]) {
		if (CPL == 1) {
			o[0] = p[0];
		} else {
#pragma unroll
			for (int j = 0; j < CPL; j += 2) {
				const float4 q = *reinterpret_cast<const float4 *>(p + j);
				o[j] = make_float2(q.x, q.y);
				o[j + 1] = make_float2(q.z, q.w);
			}
		}
	}
	__device__ static void store(float2 *p, const float2 (&o)[CPL]) {
		if (CPL == 1) {
			p[0] = o[0];
		} else {
#pragma unroll
			for (int j = 0; j < CPL; j += 2)
				*reinterpret_cast<float4 *>(p + j) =
				    make_float4(o[j].x, o[j].y, o[j + 1].x, o[j + 1].y);
		}
	}
};

// The same moves with the non-temporal hint, for the rows that are read or
// written exactly once per step (Phi_{n-1} in, Phi_{n+1} out): they stream
// past L2 instead of evicting the Phi_n rows that neighbouring rows gather.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename VT, int CPL> struct RowNT;
template <int CPL> struct RowNT<float2, CPL> {
	__device__ static void load(const float2 *p, float2 (&o)[CPL]) {
		if (CPL == 1) {
			const f32x2 q = __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(p));
			o[0] = make_float2(q.x, q.y);
		} else {
#pragma unroll
			for (int j = 0; j < CPL; j += 2) {
				const f32x4 q =
				    __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p + j));
				o[j] = make_float2(q.x, q.y);
				o[j + 1] = make_float2(q.z, q.w);
			}
		}
	}
	__device__ static void store(float2 *p, const float2 (&o)[CPL]) {
		if (CPL == 1) {
			f32x2 q = {o[0].x, o[0].y};
			__builtin_nontemporal_store(q, reinterpret_cast<f32x2 *>(p));
		} else {
#pragma unroll
			for (int j = 0; j < CPL; j += 2) {
				f32x4 q = {o[j].x, o[j].y, o[j + 1].x, o[j + 1].y};
				__builtin_nontemporal_store(q, reinterpret_cast<f32x4 *>(p + j));
			}
		}
	}
};
template <int CPL> struct RowNT<double2, CPL> {
	__device__ static void load(const double2 *p, double2 (&o)[CPL]) {
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			const f64x2 q = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p + j));
			o[j] = make_double2(q.x, q.y);
		}
	}
	__device__ static void store(double2 *p, const double2 (&o)[CPL]) {
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			f64x2 q = {o[j].x, o[j].y};
			__builtin_nontemporal_store(q, reinterpret_cast<f64x2 *>(p + j));
		}
	}
};

// Initial block: unit vectors at sites[b] (mode 0; packed: two per column)
// or random phases exp(2 pi i u(seed, (col0 + b) N + row)) (mode 1). Columns
// past the live vectors are zero.
template <typename VT>
__global__ __launch_bounds__(kThreads) void block_init(
    VT *__restrict__ X, int64_t N, int B, int live, int mode,
    const int64_t *__restrict__ sites, uint64_t seed, int64_t col0, int packed) {
	const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (i >= N * B) return;
	const int64_t row = i / B;
	const int b = (int)(i - row * B);
	double2 v = make_double2(0.0, 0.0);
	if (mode == 0 && packed) {
		// live counts real vectors: 2b in the real part, 2b + 1 in the
		// imaginary part of column b
		if (2 * b < live && sites[2 * b] == row) v.x = 1.0;
		if (2 * b + 1 < live && sites[2 * b + 1] == row) v.y = 1.0;
	} else if (b < live) {
		if (mode == 0) {
			if (sites[b] == row) v.x = 1.0;
		} else {
			uint64_t c = (uint64_t)(col0 + b) * (uint64_t)N + (uint64_t)row;
			uint64_t z = seed * 0x9E3779B97F4A7C15ull + c * 0xD1B54A32D192ED03ull +
			             0x632BE59BD9B4E019ull;
			z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
			z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
			z = z ^ (z >> 31);
			const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
			double sn, cs;
			sincos(2.0 * M_PI * u, &sn, &cs);
			v = make_double2(cs, sn);
		}
	}
	X[i] = C2<VT>::store(v);
}

// Complex value in the accumulation precision AT (float or double).
template <typename AT> struct CA { AT x, y; };
template <typename AT, typename T>
__device__ __forceinline__ CA<AT> to_acc(const T &v) {
	return CA<AT>{(AT)v.x, (AT)v.y};
}

// Broadcast lane `idx` (uniform) of a per-lane value to the whole wavefront.
__device__ __forceinline__ uint32_t bcast(uint32_t v, int idx) {
	return (uint32_t)__builtin_amdgcn_readlane((int)v, idx);
}
__device__ __forceinline__ float bcast(float v, int idx) {
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), idx));
}
__device__ __forceinline__ double bcast(double v, int idx) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), idx);
	const int hi = __builtin_amdgcn_readlane((int)(b >> 32), idx);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Column indices, weights and the lane's slice of the Phi rows for the
// non-zeros k0 .. k0+U-1 of a row ending at ke (weights past ke are zero).
// The tile's CSR entries kb0 .. kb0+63 are held one per lane in wcol/wval
// (loaded with one coalesced vector load); entries inside that window are
// broadcast with readlane, others (rows longer than the window) come from
// scalar loads. The CSR arrays carry kCsrPad >= 64 zero entries of padding,
// so every read stays in bounds and entries past ke get weight zero.
template <typename VT, typename HT, typename AT, int CPL, int U>
__device__ __forceinline__ void gather_chunk(
    const uint32_t *__restrict__ col, const HT *__restrict__ v,
    const VT *__restrict__ X, uint32_t kb0, uint32_t wcol, HT wval,
    uint32_t k0, uint32_t ke, int lane, CA<AT> (&aa)[U], VT (&xv)[U][CPL]) {
	constexpr int B = CPL * 64;
	uint32_t cc[U];
	if (k0 - kb0 + U <= 64) {
#pragma unroll
		for (int u = 0; u < U; u++) {
			const int idx = (int)(k0 - kb0) + u;
			cc[u] = bcast(wcol, idx);
			const AT live = k0 + u < ke ? (AT)1 : (AT)0;
			aa[u] = CA<AT>{(AT)bcast(wval.x, idx) * live, (AT)bcast(wval.y, idx) * live};
		}
	} else {
#pragma unroll
		for (int u = 0; u < U; u++) {
			cc[u] = col[k0 + u];
			const CA<AT> w = to_acc<AT>(v[k0 + u]);
			const AT live = k0 + u < ke ? (AT)1 : (AT)0;
			aa[u] = CA<AT>{w.x * live, w.y * live};
		}
	}
#pragma unroll
	for (int u = 0; u < U; u++)
		Row<VT, CPL>::load(X + (int64_t)cc[u] * B + lane * CPL, xv[u]);
}

// EXACT: the reference's operation sequence instead of fused multiply-adds,
// acc + (a.re x.re - a.im x.im, a.re x.im + a.im x.re) with every product and
// sum rounded (for a real H: acc + a.re x, per component), non-zeros in CSR
// order; with FP64 H and vectors each column then reproduces the plain
// recurrence (and the reference's) bit for bit. Padding entries past a row's
// end add (+-0).
template <typename VT, typename AT, int CPL, int U, bool REALH, bool EXACT = false>
__device__ __forceinline__ void accumulate_chunk(const CA<AT> (&aa)[U],
                                                 const VT (&xv)[U][CPL],
                                                 CA<AT> (&acc)[CPL]) {
#pragma clang fp contract(off)
#pragma unroll
	for (int u = 0; u < U; u++) {
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			const CA<AT> x = to_acc<AT>(xv[u][j]);
			const CA<AT> a = aa[u];
			if (EXACT && REALH) {
				acc[j].x = acc[j].x + a.x * x.x;
				acc[j].y = acc[j].y + a.x * x.y;
			} else if (EXACT) {
				const AT pr = a.x * x.x - a.y * x.y;
				const AT pi = a.x * x.y + a.y * x.x;
				acc[j].x = acc[j].x + pr;
				acc[j].y = acc[j].y + pi;
			} else if (REALH) {
				acc[j].x = fma(a.x, x.x, acc[j].x);
				acc[j].y = fma(a.x, x.y, acc[j].y);
			} else {
				acc[j].x = fma(a.x, x.x, fma(-a.y, x.y, acc[j].x));
				acc[j].y = fma(a.x, x.y, fma(a.y, x.x, acc[j].y));
			}
		}
	}
}

// One recurrence step on a block. Y holds Phi_{n-1} on entry and Phi_{n+1}
// on exit (in place, same lane reads and writes a row's columns). Row sums
// are formed in AT and rounded once to the storage type VT. With DOTS, the
// workgroup's FP64 sums of conj(X)X and conj(Y)X per column (from the stored
// values) go to part[wg][0..1][b]; in gather mode column b's moment of X goes
// to mu[b*M + nprev]. REALH: every value of H is real (2 FMAs per complex
// product instead of 4).
template <typename VT, typename HT, typename AT, int B, bool FIRST, bool DOTS,
          bool REALH, bool EXACT = false>
__global__ __launch_bounds__(kThreads) void spmm_step(
    const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
    const HT *__restrict__ v, const VT *__restrict__ X, VT *Y, int64_t N,
    int tile, int nwg, int remap, int nt, double2 *__restrict__ part,
    const int64_t *__restrict__ sites, int live, double2 *__restrict__ mu,
    int M, int nprev, int packed) {
	constexpr int CPL = B / 64;   // complex values per lane
	// Rows in flight per wavefront: CPL * RPI = 4 complex columns per lane.
	constexpr int RPI = CPL >= 4 ? 1 : 4 / CPL;
	// Non-zeros gathered together (bounded by the CSR padding kCsrPad).
	constexpr int kUnroll = sizeof(VT) == 8 ? 8 : 4;
	static_assert(kUnroll <= kCsrPad, "CSR padding too short");
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;

	if (!DOTS) {
		const int g = blockIdx.x * kThreads + threadIdx.x;
		if (g < live)
			mu[(int64_t)g * M + nprev] = site_moment<VT>(X, sites, g, B, packed);
	}

	double2 d0[CPL], d1[CPL];
#pragma unroll
	for (int j = 0; j < CPL; j++) {
		d0[j] = make_double2(0.0, 0.0);
		d1[j] = make_double2(0.0, 0.0);
	}

	// Persistent grid (nwg, a multiple of 8, is what fits on the chip at
	// once). The workgroups of one XCD (hardware id mod 8, the round-robin
	// dispatch order) own one contiguous slice of the rows, and their
	// wavefronts take tiles of `tile` rows round-robin. At any time an XCD's
	// resident wavefronts therefore sweep one window of about
	// (wavefronts x tile) rows, so the Phi_n rows of a row's near neighbours
	// are still in that XCD's L2 when they are gathered again.
	const int per = nwg >> 3;
	int xcd, local;
	if (remap) {
		xcd = blockIdx.x & 7;
		local = blockIdx.x >> 3;
	} else {
		xcd = blockIdx.x / per;
		local = blockIdx.x - xcd * per;
	}
	const int64_t s_lo = N * xcd / 8, s_hi = N * (xcd + 1) / 8;
	const int64_t stride = (int64_t)per * kWaves * tile;
	for (int64_t t0 = s_lo + ((int64_t)local * kWaves + wave) * tile; t0 < s_hi;
	     t0 += stride) {
		// Row pointers of the tile (tile < 64): lane i holds rowptr[t0 + i];
		// rows read theirs back with readlane (uniform).
		const uint32_t rp_lane = rowptr[t0 + lane < N ? t0 + lane : N];
		const int nrows = s_hi - t0 < tile ? (int)(s_hi - t0) : tile;
		// The tile's CSR window, one entry per lane (coalesced).
		const uint32_t kb0 = bcast(rp_lane, 0);
		const uint32_t wcol = col[kb0 + lane];
		const HT wval = v[kb0 + lane];
		for (int i0 = 0; i0 < nrows; i0 += RPI) {
			// RPI rows at once. For each row, kUnroll non-zeros: their
			// column indices and values are uniform (scalar) loads, and the
			// gathers of all RPI rows, their Phi_{n-1} rows (and, for the dot
			// products, their own Phi_n rows) are issued before the first
			// multiply, so a wavefront keeps RPI*(kUnroll+2) row loads in
			// flight. The CSR arrays carry kCsrPad zero entries of padding,
			// so reads past a row's end stay in bounds; those entries get
			// weight zero. Rows past the tile end are clamped to a valid
			// row for their loads and not stored.
			uint32_t kb[RPI], ke[RPI];
			int64_t base[RPI];
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				const int i = i0 + r < nrows ? i0 + r : nrows - 1;
				kb[r] = (uint32_t)__builtin_amdgcn_readlane((int)rp_lane, i);
				ke[r] = (uint32_t)__builtin_amdgcn_readlane((int)rp_lane, i + 1);
				base[r] = (t0 + i) * B + lane * CPL;
			}
			CA<AT> aa[RPI][kUnroll];
			VT xv[RPI][kUnroll][CPL];
			VT yo[RPI][CPL], xs[RPI][CPL];
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				if (!FIRST) {
					if (nt) RowNT<VT, CPL>::load(Y + base[r], yo[r]);
					else Row<VT, CPL>::load(Y + base[r], yo[r]);
				}
				if (DOTS) Row<VT, CPL>::load(X + base[r], xs[r]);
				gather_chunk<VT, HT, AT, CPL, kUnroll>(col, v, X, kb0, wcol, wval, kb[r],
				                                       ke[r], lane, aa[r], xv[r]);
			}
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				CA<AT> acc[CPL];
#pragma unroll
				for (int j = 0; j < CPL; j++) {
					if (FIRST) {
						acc[j] = CA<AT>{(AT)0, (AT)0};
					} else {
						const CA<AT> p = to_acc<AT>(yo[r][j]);
						acc[j] = CA<AT>{-p.x, -p.y};
					}
				}
				accumulate_chunk<VT, AT, CPL, kUnroll, REALH, EXACT>(aa[r], xv[r], acc);
				// Rows longer than kUnroll (none in nearest-neighbour
				// lattices) continue chunk by chunk.
				for (uint32_t k0 = kb[r] + kUnroll; k0 < ke[r]; k0 += kUnroll) {
					gather_chunk<VT, HT, AT, CPL, kUnroll>(col, v, X, kb0, wcol, wval,
					                                       k0, ke[r], lane, aa[r], xv[r]);
					accumulate_chunk<VT, AT, CPL, kUnroll, REALH, EXACT>(aa[r], xv[r], acc);
				}
				if (i0 + r >= nrows) continue;
				VT st[CPL];
#pragma unroll
				for (int j = 0; j < CPL; j++)
					st[j] = C2<VT>::store(
					    make_double2((double)acc[j].x, (double)acc[j].y));
				if (nt) RowNT<VT, CPL>::store(Y + base[r], st);
				else Row<VT, CPL>::store(Y + base[r], st);
				if (DOTS) {
#pragma unroll
					for (int j = 0; j < CPL; j++) {
						const double2 y = C2<VT>::load(st[j]);
						const double2 x = C2<VT>::load(xs[r][j]);
						dot_update(x, y, packed, d0[j], d1[j]);
					}
				}
			}
		}
	}

	if (DOTS) {
		// Sum the four wavefronts' partials in LDS, in wave order.
		__shared__ double2 red[kWaves][2][B];
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			red[wave][0][lane * CPL + j] = d0[j];
			red[wave][1][lane * CPL + j] = d1[j];
		}
		__syncthreads();
		for (int i = threadIdx.x; i < 2 * B; i += kThreads) {
			const int w = i / B, b = i - w * B;
			double2 s = red[0][w][b];
#pragma unroll
			for (int q = 1; q < kWaves; q++) {
				s.x += red[q][w][b].x;
				s.y += red[q][w][b].y;
			}
			part[(int64_t)blockIdx.x * 2 * B + i] = s;
		}
	}
}

// Tile-major ELL form of H for the batched step: tiles of T = 64/W rows,
// W slots per row (W >= the longest row), stored tile by tile as
// ecol[tile][row][slot] and the values likewise, so one coalesced load of
// T*W <= 64 lanes brings a whole tile's column indices (and one its
// values) without reading row pointers first. Unused slots hold the row's
// own index with value zero.
//
// The wavefront prefetches the next tile's window while it gathers and
// multiplies the current tile, so a tile costs one exposed round trip to
// memory (the gathers), hidden by the other resident wavefronts.
template <typename VT, typename HT, typename AT, int B, bool FIRST, bool DOTS,
          bool REALH, int W>
__global__ __launch_bounds__(kThreads) void spmm_step_ell(
    const uint32_t *__restrict__ ecol, const HT *__restrict__ eval,
    const VT *__restrict__ X, VT *Y, int64_t N, int nwg, int remap, int nt,
    double2 *__restrict__ part, const int64_t *__restrict__ sites, int live,
    double2 *__restrict__ mu, int M, int nprev, int packed) {
	constexpr int CPL = B / 64;
	constexpr int T = 64 / W;          // rows per tile
	constexpr int TW = T * W;          // entries per tile (<= 64)
	constexpr int RPI = CPL >= 4 ? 1 : 4 / CPL;
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;

	if (!DOTS) {
		const int g = blockIdx.x * kThreads + threadIdx.x;
		if (g < live)
			mu[(int64_t)g * M + nprev] = site_moment<VT>(X, sites, g, B, packed);
	}

	double2 d0[CPL], d1[CPL];
#pragma unroll
	for (int j = 0; j < CPL; j++) {
		d0[j] = make_double2(0.0, 0.0);
		d1[j] = make_double2(0.0, 0.0);
	}

	// Persistent grid; tile slices per XCD (remap) or per block of
	// consecutive workgroups, wavefronts round-robin over the slice's tiles.
	const int per = nwg >> 3;
	int xcd, local;
	if (remap) {
		xcd = blockIdx.x & 7;
		local = blockIdx.x >> 3;
	} else {
		xcd = blockIdx.x / per;
		local = blockIdx.x - xcd * per;
	}
	const int64_t ntiles = (N + T - 1) / T;
	const int64_t s_lo = ntiles * xcd / 8, s_hi = ntiles * (xcd + 1) / 8;
	const int64_t stride = (int64_t)per * kWaves;
	int64_t t = s_lo + (int64_t)local * kWaves + wave;

	uint32_t wcol = 0;
	HT wval = {};
	if (t < s_hi && lane < TW) {
		wcol = ecol[t * TW + lane];
		wval = eval[t * TW + lane];
	}
	for (; t < s_hi; t += stride) {
		// Prefetch the next tile's window.
		const int64_t tn = t + stride;
		uint32_t ncol = 0;
		HT nval = {};
		if (tn < s_hi && lane < TW) {
			ncol = ecol[tn * TW + lane];
			nval = eval[tn * TW + lane];
		}
		const int64_t r0 = t * T;
		const int nrows = N - r0 < T ? (int)(N - r0) : T;
		for (int i0 = 0; i0 < nrows; i0 += RPI) {
			int64_t base[RPI];
			CA<AT> aa[RPI][W];
			VT xv[RPI][W][CPL];
			VT yo[RPI][CPL], xs[RPI][CPL];
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				const int i = i0 + r < nrows ? i0 + r : nrows - 1;
				base[r] = (r0 + i) * B + lane * CPL;
				if (!FIRST) {
					if (nt) RowNT<VT, CPL>::load(Y + base[r], yo[r]);
					else Row<VT, CPL>::load(Y + base[r], yo[r]);
				}
#pragma unroll
				for (int u = 0; u < W; u++) {
					const uint32_t c = bcast(wcol, i * W + u);
					aa[r][u] = CA<AT>{(AT)bcast(wval.x, i * W + u), (AT)bcast(wval.y, i * W + u)};
					Row<VT, CPL>::load(X + (int64_t)c * B + lane * CPL, xv[r][u]);
				}
				// Slot 0 of an ELL row is the row itself (diagonal first).
				if (DOTS) {
#pragma unroll
					for (int j = 0; j < CPL; j++) xs[r][j] = xv[r][0][j];
				}
			}
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				CA<AT> acc[CPL];
#pragma unroll
				for (int j = 0; j < CPL; j++) {
					if (FIRST) {
						acc[j] = CA<AT>{(AT)0, (AT)0};
					} else {
						const CA<AT> p = to_acc<AT>(yo[r][j]);
						acc[j] = CA<AT>{-p.x, -p.y};
					}
				}
				accumulate_chunk<VT, AT, CPL, W, REALH>(aa[r], xv[r], acc);
				if (i0 + r >= nrows) continue;
				VT st[CPL];
#pragma unroll
				for (int j = 0; j < CPL; j++)
					st[j] = C2<VT>::store(make_double2((double)acc[j].x, (double)acc[j].y));
				if (nt) RowNT<VT, CPL>::store(Y + base[r], st);
				else Row<VT, CPL>::store(Y + base[r], st);
				if (DOTS) {
#pragma unroll
					for (int j = 0; j < CPL; j++) {
						const double2 y = C2<VT>::load(st[j]);
						const double2 x = C2<VT>::load(xs[r][j]);
						dot_update(x, y, packed, d0[j], d1[j]);
					}
				}
			}
		}
		wcol = ncol;
		wval = nval;
	}

	if (DOTS) {
		__shared__ double2 red[kWaves][2][B];
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			red[wave][0][lane * CPL + j] = d0[j];
			red[wave][1][lane * CPL + j] = d1[j];
		}
		__syncthreads();
		for (int i = threadIdx.x; i < 2 * B; i += kThreads) {
			const int w = i / B, b = i - w * B;
			double2 s = red[0][w][b];
#pragma unroll
			for (int q = 1; q < kWaves; q++) {
				s.x += red[q][w][b].x;
				s.y += red[q][w][b].y;
			}
			part[(int64_t)blockIdx.x * 2 * B + i] = s;
		}
	}
}

// Traversal of the rows in L2-sized strips. Rows are viewed as layers of D
// (the band: max |col - row|, the slowest lattice stride) and every layer
// is cut into strips of S rows; the traversal visits strip k of layers 0, 1,
// 2, ... before strip k+1, so a row's neighbours one stride away (the
// previous and next layer) are visited S rows before and after it instead of
// D rows. Segments (strip k, layer a) are cut into tiles of at most T rows.
// XCD x (hardware workgroup id mod 8) owns the x-th eighth of that tile
// list, and its wavefronts take consecutive tiles round-robin, so the rows in
// flight on one XCD (and the Phi_n rows they gather) form a narrow front
// whose footprint (front + 2 S rows) fits its L2.
struct StripMap {
	int64_t N, D;     // rows, layer length (band)
	int S, A, tps;    // strip rows, layers, tiles per segment
	int64_t ntiles;   // strips * A * tps
	int ch;           // consecutive tiles a workgroup's 4 wavefronts share
};

// Tile t of the strip traversal: first row and row count (0 for the empty
// tail of a partial strip or layer). The tile list has fewer than 2^31
// entries (checked on the host), so the divisions are 32-bit.
__device__ __forceinline__ int strip_tile(const StripMap &m, int T, int64_t t, int64_t *row0) {
	const uint32_t tt = (uint32_t)t;
	const uint32_t seg = tt / (uint32_t)m.tps;
	const int j = (int)(tt - seg * (uint32_t)m.tps);
	const uint32_t k = seg / (uint32_t)m.A;
	const int64_t a = seg - k * (uint32_t)m.A;
	const int64_t c0 = (int64_t)k * m.S + (int64_t)j * T;
	int64_t n = m.S - (int64_t)j * T < T ? m.S - (int64_t)j * T : T;
	if (c0 + n > m.D) n = m.D - c0;
	*row0 = a * m.D + c0;
	if (*row0 + n > m.N) n = m.N - *row0;
	if (n <= 0) {
		// Empty tail tile: row 0 keeps the (unused) loads of the tile in
		// bounds; it stores nothing.
		*row0 = 0;
		return 0;
	}
	return (int)n;
}

template <typename VT, typename HT, typename AT, int B, bool FIRST, bool DOTS,
          bool REALH, int W, int RMUL>
__global__ __launch_bounds__(kThreads) void spmm_step_strip(
    const uint32_t *__restrict__ ecol, const HT *__restrict__ eval,
    const VT *__restrict__ X, VT *Y, StripMap sm, int nwg, int nt,
    double2 *__restrict__ part, const int64_t *__restrict__ sites, int live,
    double2 *__restrict__ mu, int M, int nprev, int packed) {
	constexpr int CPL = B / 64;
	// Rows a wavefront keeps in flight: 4 / CPL (at least 1), times RMUL
	// (2 doubles the bytes in flight per wavefront for the price of fewer
	// resident wavefronts).
	constexpr int RPI = (CPL >= 4 ? 1 : 4 / CPL) * RMUL;
	// A tile is the rows a wavefront keeps in flight, so the XCD's front is
	// (wavefronts x RPI) rows wide.
	constexpr int T = RPI;
	const int wave = threadIdx.x >> 6;
	const int lane = threadIdx.x & 63;

	if (!DOTS) {
		const int g = blockIdx.x * kThreads + threadIdx.x;
		if (g < live)
			mu[(int64_t)g * M + nprev] = site_moment<VT>(X, sites, g, B, packed);
	}

	double2 d0[CPL], d1[CPL];
#pragma unroll
	for (int j = 0; j < CPL; j++) {
		d0[j] = make_double2(0.0, 0.0);
		d1[j] = make_double2(0.0, 0.0);
	}

	// The XCD's part of the tile list is cut into chunks of sm.ch consecutive
	// tiles dealt round-robin to its workgroups; the 4 wavefronts of a
	// workgroup interleave over the chunk's tiles, so the rows one CU works
	// on at a time are adjacent and their nearby neighbours hit its L1.
	const int xcd = blockIdx.x & 7;
	const int local = blockIdx.x >> 3;
	const int per = nwg >> 3;
	const int64_t s_lo = sm.ntiles * xcd / 8, s_hi = sm.ntiles * (xcd + 1) / 8;
	const int spc = sm.ch / kWaves;          // tiles per wavefront per chunk
	// k-th tile of this wavefront (s_hi when past the end).
	auto tile_of = [&](int64_t k) -> int64_t {
		const int64_t chunk = local + (k / spc) * per;
		const int64_t tt = s_lo + chunk * sm.ch + wave + (k % spc) * kWaves;
		return tt < s_hi ? tt : s_hi;
	};
	int64_t kk = 0;
	int64_t t = tile_of(0);

	// Per tile, in issue order: broadcast the tile's window (loaded one tile
	// ahead), issue its gathers and Phi_{n-1} / Phi_n row loads, issue the
	// stores of the previous tile's rows (kept in registers until now),
	// prefetch the next window, then wait for this tile's loads and compute.
	// The window registers are consumed before any of the tile's loads are
	// issued, so the wait on them never covers this tile's own round trip,
	// and the stores overlap the next tile's loads.
	uint32_t wcol = 0;
	HT wval = {};
	int64_t r0 = 0;
	int nrows = 0;
	if (t < s_hi) {
		nrows = strip_tile(sm, T, t, &r0);
		if (lane < nrows * W) {
			wcol = ecol[r0 * W + lane];
			wval = eval[r0 * W + lane];
		}
	}
	VT pst[RPI][CPL];
	int64_t pbase[RPI];
	int pn = 0;
	while (t < s_hi) {
		uint32_t cc[RPI][W];
		CA<AT> aa[RPI][W];
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			const int i = r < nrows ? r : (nrows > 0 ? nrows - 1 : 0);
#pragma unroll
			for (int u = 0; u < W; u++) {
				cc[r][u] = bcast(wcol, i * W + u);
				aa[r][u] = CA<AT>{(AT)bcast(wval.x, i * W + u), (AT)bcast(wval.y, i * W + u)};
			}
		}
		int64_t base[RPI];
		VT xv[RPI][W][CPL];
		VT yo[RPI][CPL], xs[RPI][CPL];
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			const int i = r < nrows ? r : (nrows > 0 ? nrows - 1 : 0);
			base[r] = (r0 + i) * B + lane * CPL;
#pragma unroll
			for (int u = 0; u < W; u++)
				Row<VT, CPL>::load(X + (int64_t)cc[r][u] * B + lane * CPL, xv[r][u]);
			if (!FIRST) {
				if (nt) RowNT<VT, CPL>::load(Y + base[r], yo[r]);
				else Row<VT, CPL>::load(Y + base[r], yo[r]);
			}
			// Slot 0 of an ELL row is the row itself (diagonal first).
			if (DOTS) {
#pragma unroll
				for (int j = 0; j < CPL; j++) xs[r][j] = xv[r][0][j];
			}
		}
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			if (r >= pn) break;
			if (nt) RowNT<VT, CPL>::store(Y + pbase[r], pst[r]);
			else Row<VT, CPL>::store(Y + pbase[r], pst[r]);
		}
		// Tiles of a chunk beyond the part's end are skipped; a later chunk
		// of this wavefront can only start further on, so the first tile
		// index at or past s_hi ends the walk.
		const int64_t tn = tile_of(++kk);
		uint32_t ncol = 0;
		HT nval = {};
		int64_t rn = 0;
		int nn = 0;
		if (tn < s_hi) {
			nn = strip_tile(sm, T, tn, &rn);
			if (lane < nn * W) {
				ncol = ecol[rn * W + lane];
				nval = eval[rn * W + lane];
			}
		}
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			CA<AT> acc[CPL];
#pragma unroll
			for (int j = 0; j < CPL; j++) {
				if (FIRST) {
					acc[j] = CA<AT>{(AT)0, (AT)0};
				} else {
					const CA<AT> p = to_acc<AT>(yo[r][j]);
					acc[j] = CA<AT>{-p.x, -p.y};
				}
			}
			accumulate_chunk<VT, AT, CPL, W, REALH>(aa[r], xv[r], acc);
#pragma unroll
			for (int j = 0; j < CPL; j++)
				pst[r][j] = C2<VT>::store(make_double2((double)acc[j].x, (double)acc[j].y));
			pbase[r] = base[r];
			if (DOTS && r < nrows) {
#pragma unroll
				for (int j = 0; j < CPL; j++) {
					const double2 y = C2<VT>::load(pst[r][j]);
					const double2 x = C2<VT>::load(xs[r][j]);
					dot_update(x, y, packed, d0[j], d1[j]);
				}
			}
		}
		pn = nrows;
		wcol = ncol;
		wval = nval;
		r0 = rn;
		nrows = nn;
		t = tn;
	}
#pragma unroll
	for (int r = 0; r < RPI; r++) {
		if (r >= pn) break;
		if (nt) RowNT<VT, CPL>::store(Y + pbase[r], pst[r]);
		else Row<VT, CPL>::store(Y + pbase[r], pst[r]);
	}

	if (DOTS) {
		__shared__ double2 red[kWaves][2][B];
#pragma unroll
		for (int j = 0; j < CPL; j++) {
			red[wave][0][lane * CPL + j] = d0[j];
			red[wave][1][lane * CPL + j] = d1[j];
		}
		__syncthreads();
		for (int i = threadIdx.x; i < 2 * B; i += kThreads) {
			const int w = i / B, b = i - w * B;
			double2 s = red[0][w][b];
#pragma unroll
			for (int q = 1; q < kWaves; q++) {
				s.x += red[q][w][b].x;
				s.y += red[q][w][b].y;
			}
			part[(int64_t)blockIdx.x * 2 * B + i] = s;
		}
	}
}

// D[step][w][b] = sum over workgroups of part[step][wg][w][b], in workgroup
// order (deterministic).
__global__ __launch_bounds__(kThreads) void reduce_parts(
    const double2 *__restrict__ part, int nsteps, int nwg, int B,
    double2 *__restrict__ D) {
	const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	const int64_t per = 2 * (int64_t)B;
	if (i >= nsteps * per) return;
	const int64_t step = i / per, c = i - step * per;
	const double2 *p = part + step * nwg * per + c;
	double sx = 0, sy = 0;
	for (int w = 0; w < nwg; w++) {
		sx += p[(int64_t)w * per].x;
		sy += p[(int64_t)w * per].y;
	}
	D[i] = make_double2(sx, sy);
}

template <typename VT>
__global__ __launch_bounds__(kThreads) void gather_block(
    const VT *__restrict__ X, const int64_t *__restrict__ sites, int live,
    int B, double2 *__restrict__ mu, int M, int n, int packed) {
	const int g = blockIdx.x * kThreads + threadIdx.x;
	if (g < live) mu[(int64_t)g * M + n] = site_moment<VT>(X, sites, g, B, packed);
}

// General moment gather: mu[l*M + n] = component of Phi_n at (row[l],
// vector vec[l]) for every entry l < L (vector v in column v, or with packed
// real vectors in the real (v even) or imaginary (v odd) part of column v/2).
template <typename VT>
__global__ __launch_bounds__(kThreads) void gather_list(
    const VT *__restrict__ X, const int64_t *__restrict__ row, const int *__restrict__ vec,
    int L, int B, double2 *__restrict__ mu, int M, int n, int packed) {
	const int l = blockIdx.x * kThreads + threadIdx.x;
	if (l >= L) return;
	const int v = vec[l];
	double2 x;
	if (packed) {
		const double2 c = C2<VT>::load(X[row[l] * B + (v >> 1)]);
		x = make_double2((v & 1) ? c.y : c.x, 0.0);
	} else {
		x = C2<VT>::load(X[row[l] * B + v]);
	}
	mu[(int64_t)l * M + n] = x;
}

inline unsigned grid_for(int64_t n) {
	return (unsigned)((n + kThreads - 1) / kThreads);
}

struct BlockRun {
	tbtkx_context *h;
	int B;             // block width
	int flags;
	int nwg;
	int tile;          // rows per wavefront tile
	int64_t launches = 0;
	int S = -1;        // strip rows of the strip traversal (-1: natural order)
	int share = 1;     // grid = occupancy x CUs / share
	int rmul = 1;      // rows in flight multiplier of the strip step (1 or 2)
	StripMap sm{};
	bool exact = false;   // CSR step in the reference's operation order
};

// The strip step with doubled rows in flight is built for FP32 vectors and
// 8-byte H entries only (the batched default).
template <typename VT, typename HT>
constexpr bool kFp32Pair = std::is_same<VT, float2>::value && kHt8<HT>;
// The exact step exists for FP64 H and vectors only.
template <typename VT, typename HT>
constexpr bool kFp64Pair = std::is_same<VT, double2>::value && std::is_same<HT, double2>::value;

// Strip map for S rows per strip on the resident model (ELL layout), tiles
// of T rows (the step kernel's rows in flight).
StripMap make_strip(const tbtkx_context *h, int S, int T) {
	StripMap m{};
	m.N = h->N;
	m.D = h->band < h->N ? h->band : h->N;
	if (S > m.D) S = (int)m.D;
	m.S = S;
	m.A = (int)((m.N + m.D - 1) / m.D);
	m.tps = (S + T - 1) / T;
	m.ntiles = (int64_t)((m.D + S - 1) / S) * m.A * m.tps;
	m.ch = h->strip_chunk > 0 ? h->strip_chunk : 4;
	return m;
}

template <typename VT, typename HT, int B>
int launch_step(BlockRun &r, bool first, bool dots, const VT *X, VT *Y,
                double2 *part, const int64_t *sites, int live, double2 *mu,
                int M, int nprev, int packed) {
	tbtkx_context *h = r.h;
	const HT *v = first ? h->v1.as<HT>() : h->v2.as<HT>();
	if (kHt8<HT>)
		v = first ? (const HT *)h->v1f.ptr : (const HT *)h->v2f.ptr;
	const uint32_t *rp = h->rowptr.as<uint32_t>(), *cl = h->col.as<uint32_t>();
	// FP32 H and vectors accumulate rows in FP32; anything with FP64 in it
	// (a real H's FP64 entries included) accumulates in FP64.
	using AT = typename std::conditional<std::is_same<HT, float2>::value &&
	                                         std::is_same<VT, float2>::value,
	                                     float, double>::type;
	if (h->ell_W > 0 && h->use_ell && !r.exact) {
		const uint32_t *ec = h->ell_col.as<uint32_t>();
		const HT *ev;
		if (kHt8<HT>)
			ev = first ? (const HT *)h->ell_v1f.ptr : (const HT *)h->ell_v2f.ptr;
		else
			ev = first ? (const HT *)h->ell_v1d.ptr : (const HT *)h->ell_v2d.ptr;
#define TBTKX_ELL(F, D, R, WW)                                                 \
	do {                                                                   \
		if (r.S > 0 && r.rmul == 2 && kFp32Pair<VT, HT>)               \
			spmm_step_strip<VT, HT, AT, B, F, D, R, WW, 2>         \
			    <<<r.nwg, kThreads, 0, h->stream>>>(               \
			        ec, ev, X, Y, r.sm, r.nwg, h->nt_stream, part, \
			        sites, live, mu, M, nprev, packed);            \
		else if (r.S > 0)                                              \
			spmm_step_strip<VT, HT, AT, B, F, D, R, WW, 1>         \
			    <<<r.nwg, kThreads, 0, h->stream>>>(               \
			        ec, ev, X, Y, r.sm, r.nwg, h->nt_stream, part, \
			        sites, live, mu, M, nprev, packed);            \
		else                                                           \
			spmm_step_ell<VT, HT, AT, B, F, D, R, WW>              \
			    <<<r.nwg, kThreads, 0, h->stream>>>(               \
			        ec, ev, X, Y, h->N, r.nwg, h->xcd_remap,       \
			        h->nt_stream, part, sites, live, mu, M, nprev, packed); \
	} while (0)
#define TBTKX_ELL_W(F, D, R)                                                   \
	do {                                                                   \
		if (h->ell_W == 4) TBTKX_ELL(F, D, R, 4);                      \
		else if (h->ell_W == 6) TBTKX_ELL(F, D, R, 6);                 \
		else if (h->ell_W == 7) TBTKX_ELL(F, D, R, 7);                 \
		else TBTKX_ELL(F, D, R, 8);                                    \
	} while (0)
#define TBTKX_ELL_R(F, D)                                                      \
	do {                                                                   \
		if (h->h_real) TBTKX_ELL_W(F, D, true);                        \
		else TBTKX_ELL_W(F, D, false);                                 \
	} while (0)
		if (first && dots) TBTKX_ELL_R(true, true);
		else if (first) TBTKX_ELL_R(true, false);
		else if (dots) TBTKX_ELL_R(false, true);
		else TBTKX_ELL_R(false, false);
#undef TBTKX_ELL_R
#undef TBTKX_ELL_W
#undef TBTKX_ELL
		r.launches++;
		return TBTKX_OK;
	}
#define TBTKX_STEP(F, D, R)                                                    \
	do {                                                                   \
		if constexpr (kFp64Pair<VT, HT>) {                             \
			if (r.exact) {                                         \
				spmm_step<VT, HT, AT, B, F, D, R, true>        \
				    <<<r.nwg, kThreads, 0, h->stream>>>(       \
				        rp, cl, v, X, Y, h->N, r.tile, r.nwg,  \
				        h->xcd_remap, h->nt_stream, part, sites, \
				        live, mu, M, nprev, packed);           \
				break;                                         \
			}                                                      \
		}                                                              \
		spmm_step<VT, HT, AT, B, F, D, R><<<r.nwg, kThreads, 0, h->stream>>>( \
		    rp, cl, v, X, Y, h->N, r.tile, r.nwg, h->xcd_remap, h->nt_stream, \
		    part, sites, live, mu, M, nprev, packed);                  \
	} while (0)
#define TBTKX_STEP_R(F, D)                                                     \
	do {                                                                   \
		if (h->h_real) TBTKX_STEP(F, D, true);                         \
		else TBTKX_STEP(F, D, false);                                  \
	} while (0)
	if (first && dots) TBTKX_STEP_R(true, true);
	else if (first) TBTKX_STEP_R(true, false);
	else if (dots) TBTKX_STEP_R(false, true);
	else TBTKX_STEP_R(false, false);
#undef TBTKX_STEP_R
#undef TBTKX_STEP
	r.launches++;
	return TBTKX_OK;
}

// Run one block: S steps, moments into mu_dev (gather: live x M; doubling:
// D buffer of S x 2 x B). Returns the number of steps.
template <typename VT, typename HT, int B>
int run_block(BlockRun &r, int live, const int64_t *d_sites, int mode,
              uint64_t seed, int64_t col0, int M, double2 *mu_dev,
              double2 *D_dev, int packed) {
	tbtkx_context *h = r.h;
	const bool dots = (r.flags & TBTKX_DOUBLING) != 0;
	VT *X = h->blk_x.as<VT>(), *Y = h->blk_y.as<VT>();
	double2 *part = h->blk_part.as<double2>();
	block_init<VT><<<grid_for(h->N * B), kThreads, 0, h->stream>>>(
	    X, h->N, B, live, mode, d_sites, seed, col0, packed);
	const int S = dots ? (M + 1) / 2 : M - 1;
	int chunk0 = 0;
	for (int n = 0; n < S; n++) {
		// Step n computes Phi_{n+1} from Phi_n (X) and Phi_{n-1} (Y).
		double2 *p = part + (int64_t)(n - chunk0) * r.nwg * 2 * B;
		int rc = launch_step<VT, HT, B>(r, n == 0, dots, X, Y, p, d_sites,
		                                live, mu_dev, M, n, packed);
		if (rc) return rc;
		VT *t = X;
		X = Y;
		Y = t;
		if (dots && (n - chunk0 + 1 == kDotChunk || n == S - 1)) {
			const int ns = n - chunk0 + 1;
			reduce_parts<<<grid_for((int64_t)ns * 2 * B), kThreads, 0,
			               h->stream>>>(part, ns, r.nwg, B,
			                            D_dev + (int64_t)chunk0 * 2 * B);
			chunk0 = n + 1;
		}
	}
	if (!dots)
		gather_block<VT><<<grid_for(live), kThreads, 0, h->stream>>>(
		    X, d_sites, live, B, mu_dev, M, M - 1, packed);
	TBTKX_CHECK_HIP(hipGetLastError());
	return TBTKX_OK;
}

// Workgroups of the persistent step grid: as many as fit on the device at
// once for the steady-state kernel variant, a multiple of 8 (one share per
// XCD), and no more than the row tiles need.
template <typename VT, typename HT, typename AT, int B, bool D, bool R>
const void *ell_fn(int W, bool strip, int rmul) {
	if (strip && rmul == 2 && kFp32Pair<VT, HT>) {
		if (W == 4) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 4, 2>;
		if (W == 6) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 6, 2>;
		if (W == 7) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 7, 2>;
		return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 8, 2>;
	}
	if (strip) {
		if (W == 4) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 4, 1>;
		if (W == 6) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 6, 1>;
		if (W == 7) return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 7, 1>;
		return (const void *)spmm_step_strip<VT, HT, AT, B, false, D, R, 8, 1>;
	}
	if (W == 4) return (const void *)spmm_step_ell<VT, HT, AT, B, false, D, R, 4>;
	if (W == 6) return (const void *)spmm_step_ell<VT, HT, AT, B, false, D, R, 6>;
	if (W == 7) return (const void *)spmm_step_ell<VT, HT, AT, B, false, D, R, 7>;
	return (const void *)spmm_step_ell<VT, HT, AT, B, false, D, R, 8>;
}

// Workgroups of the persistent step grid: the occupancy of the steady-state
// kernel variant times the CUs, divided by `share`, a multiple of 8 (one
// share per XCD), and no more than the row tiles need. With strips (r.S > 0)
// r.sm is set up as well.
template <typename VT, typename HT, int B>
int persistent_grid(tbtkx_context *h, bool dots, BlockRun &r) {
	int tile = r.tile;
	int64_t tiles_total = 0;
	using AT = typename std::conditional<std::is_same<HT, float2>::value &&
	                                         std::is_same<VT, float2>::value,
	                                     float, double>::type;
	const void *fn;
	const bool ell = h->ell_W > 0 && h->use_ell && !r.exact;
	if (!ell) r.S = -1;
	if (ell) {
		// The occupancy of the steady-state ELL variant.
		tile = 64 / h->ell_W;
		const bool strip = r.S > 0;
		if (!kFp32Pair<VT, HT>) r.rmul = 1;
		if (strip) {
			// Tile rows = rows in flight of the step kernel.
			const int kStripT = (B >= 256 ? 1 : 256 / B) * r.rmul;
			r.sm = make_strip(h, r.S, kStripT);
			TBTKX_REQUIRE(r.sm.ntiles < ((int64_t)1 << 31), TBTKX_ERR_ARG,
			              "batched step: too many strip tiles; use a larger strip");
			tiles_total = r.sm.ntiles;
		}
		if (dots)
			fn = h->h_real ? ell_fn<VT, HT, AT, B, true, true>(h->ell_W, strip, r.rmul)
			               : ell_fn<VT, HT, AT, B, true, false>(h->ell_W, strip, r.rmul);
		else
			fn = h->h_real ? ell_fn<VT, HT, AT, B, false, true>(h->ell_W, strip, r.rmul)
			               : ell_fn<VT, HT, AT, B, false, false>(h->ell_W, strip, r.rmul);
	} else if (r.exact) {
		if constexpr (kFp64Pair<VT, HT>)
			fn = h->h_real ? (const void *)spmm_step<VT, HT, AT, B, false, false, true, true>
			               : (const void *)spmm_step<VT, HT, AT, B, false, false, false, true>;
		else
			return TBTKX_ERR_STATE;
	} else if (dots) {
		fn = h->h_real ? (const void *)spmm_step<VT, HT, AT, B, false, true, true>
		               : (const void *)spmm_step<VT, HT, AT, B, false, true, false>;
	} else {
		fn = h->h_real ? (const void *)spmm_step<VT, HT, AT, B, false, false, true>
		               : (const void *)spmm_step<VT, HT, AT, B, false, false, false>;
	}
	int per_cu = 0, cus = 0;
	TBTKX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
	TBTKX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
	                                      h->device));
	int64_t n = (int64_t)(per_cu > 0 ? per_cu : 1) * cus / (r.share > 0 ? r.share : 1);
	const int64_t tiles = tiles_total > 0 ? tiles_total : (h->N + tile - 1) / tile;
	const int64_t need = (tiles + kWaves - 1) / kWaves;
	if (n > need) n = need;
	n = (n + 7) / 8 * 8;
	if (n < 8) n = 8;
	r.nwg = (int)n;
	return TBTKX_OK;
}

template <typename VT, typename HT, int B>
int run_all(tbtkx_context *h, int total, const int64_t *sites, int mode,
            uint64_t seed, int M, int flags, int packed, std::vector<double2> &host_mu) {
	// Tile: rows a wavefront takes per round-robin turn, a multiple of the
	// rows it keeps in flight (RPI of spmm_step), at most 32.
	constexpr int kRpi = B >= 256 ? 1 : 256 / B;
	int tile = h->tile_rows > 0 ? h->tile_rows : 4;
	tile = (tile + kRpi - 1) / kRpi * kRpi;
	if (tile > 32) tile = 32;
	BlockRun r{h, B, flags, 0, tile, 0};
	r.S = h->run_strip;
	r.share = h->run_share;
	r.rmul = h->run_rmul;
	int rc = persistent_grid<VT, HT, B>(h, (flags & TBTKX_DOUBLING) != 0, r);
	if (rc) return rc;
	const bool dots = (flags & TBTKX_DOUBLING) != 0;
	const int S = dots ? (M + 1) / 2 : M - 1;
	// Vectors per block: B columns, two real vectors per column when packed.
	const int vpb = packed ? 2 * B : B;
	const int nblocks = (total + vpb - 1) / vpb;
	// Every block writes its own moment region, so the blocks run back to
	// back with one copy to the host at the end.
	const size_t per_block = dots ? (size_t)(S > 0 ? S : 1) * 2 * B : (size_t)vpb * M;
	if ((rc = h->blk_x.reserve((size_t)h->N * B * sizeof(VT)))) return rc;
	if ((rc = h->blk_y.reserve((size_t)h->N * B * sizeof(VT)))) return rc;
	if (dots &&
	    (rc = h->blk_part.reserve((size_t)kDotChunk * r.nwg * 2 * B * 16)))
		return rc;
	if ((rc = h->blk_mu.reserve(per_block * nblocks * 16))) return rc;
	if ((rc = h->blk_sites.reserve((size_t)(mode == 0 ? total : 1) * 8))) return rc;
	if (mode == 0)
		TBTKX_CHECK_HIP(hipMemcpyAsync(h->blk_sites.ptr, sites, (size_t)total * 8,
		                               hipMemcpyHostToDevice, h->stream));

	auto enqueue = [&]() -> int {
		for (int blk = 0; blk < nblocks; blk++) {
			const int live = total - blk * vpb < vpb ? total - blk * vpb : vpb;
			double2 *mu = h->blk_mu.as<double2>() + per_block * blk;
			int erc = run_block<VT, HT, B>(r, live, h->blk_sites.as<int64_t>() + (int64_t)blk * vpb,
			                               mode, seed, (int64_t)blk * B, M, mu, mu, packed);
			if (erc) return erc;
		}
		return TBTKX_OK;
	};
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	if (h->use_graph) {
		GraphCache &gc = h->blk_graph;
		int64_t key[GraphCache::kKey];
		const bool key_ok = GraphCache::make_key(
		    key,
		    {total, M, flags, mode, (int64_t)seed, B, (int64_t)r.nwg * 64 + r.tile,
		     h->use_ell * 4 + h->nt_stream * 2 + h->xcd_remap, h->N, h->ell_W, r.S,
		     h->strip_chunk, r.rmul, packed},
		    {h->blk_x.ptr, h->blk_y.ptr, h->blk_part.ptr, h->blk_mu.ptr, h->blk_sites.ptr,
		     h->rowptr.ptr, h->col.ptr, h->v1.ptr, h->v2.ptr, h->v1f.ptr, h->v2f.ptr,
		     h->ell_col.ptr, h->ell_v1f.ptr, h->ell_v2f.ptr});
		TBTKX_REQUIRE(key_ok, TBTKX_ERR_STATE, "batched step: graph key too long");
		if (!gc.matches(key)) {
			gc.reset();
			TBTKX_CHECK_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
			int erc = enqueue();
			hipError_t ec = hipStreamEndCapture(h->stream, &gc.graph);
			if (erc) return erc;
			TBTKX_CHECK_HIP(ec);
			TBTKX_CHECK_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
			memcpy(gc.key, key, sizeof key);
		}
		TBTKX_CHECK_HIP(hipGraphLaunch(gc.exec, h->stream));
	} else {
		r.launches = 0;
		if ((rc = enqueue())) return rc;
	}
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	std::vector<double2> buf(per_block * nblocks);
	TBTKX_CHECK_HIP(hipMemcpyAsync(buf.data(), h->blk_mu.ptr, buf.size() * 16,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_launches = (int64_t)nblocks * S;
	h->last_block_width = B;
	h->last_strip = r.S;
	h->last_share = r.share;
	h->last_rmul = r.rmul;
	h->last_real = packed;

	host_mu.assign((size_t)total * M, make_double2(0, 0));
	for (int blk = 0; blk < nblocks; blk++) {
		const int live = total - blk * vpb < vpb ? total - blk * vpb : vpb;
		const double2 *bb = buf.data() + per_block * blk;
		for (int b = 0; b < live; b++) {
			double2 *o = host_mu.data() + (size_t)(blk * vpb + b) * M;
			if (!dots) {
				memcpy(o, bb + (size_t)b * M, (size_t)M * 16);
				continue;
			}
			// D[n][0][c] = <Phi_n|Phi_n>, D[n][1][c] = <Phi_n+1|Phi_n> of
			// column c; packed: per component, vector b in column b / 2.
			const int c = packed ? b >> 1 : b;
			auto dval = [&](size_t i) {
				const double2 d = bb[i + c];
				if (!packed) return d;
				return make_double2((b & 1) ? d.y : d.x, 0.0);
			};
			const double2 mu0 = dval(0);
			const double2 mu1 = M > 1 ? dval(B) : make_double2(0, 0);
			o[0] = mu0;
			if (M > 1) o[1] = mu1;
			for (int n = 1; n < S; n++) {
				const double2 a = dval((size_t)n * 2 * B);
				const double2 e = dval((size_t)n * 2 * B + B);
				if (2 * n < M) o[2 * n] = make_double2(2 * a.x - mu0.x, 2 * a.y - mu0.y);
				if (2 * n + 1 < M)
					o[2 * n + 1] = make_double2(2 * e.x - mu1.x, 2 * e.y - mu1.y);
			}
		}
	}
	return TBTKX_OK;
}

// Time a few steps of one block of width B on the resident model (random
// start vectors, the same kernel variant a real run uses) for the width
// autotuner; returns milliseconds per step.
template <typename VT, typename HT, int B>
int trial_width(tbtkx_context *h, int flags, int S, int share, int rmul, float *ms_per_step) {
	constexpr int kRpi = B >= 256 ? 1 : 256 / B;
	int tile = h->tile_rows > 0 ? h->tile_rows : 4;
	tile = (tile + kRpi - 1) / kRpi * kRpi;
	if (tile > 32) tile = 32;
	BlockRun r{h, B, flags, 0, tile, 0};
	r.S = S;
	r.share = share;
	r.rmul = rmul;
	const bool dots = (flags & TBTKX_DOUBLING) != 0;
	int rc = persistent_grid<VT, HT, B>(h, dots, r);
	if (rc) return rc;
	if ((rc = h->blk_x.reserve((size_t)h->N * B * sizeof(VT)))) return rc;
	if ((rc = h->blk_y.reserve((size_t)h->N * B * sizeof(VT)))) return rc;
	if (dots && (rc = h->blk_part.reserve((size_t)kDotChunk * r.nwg * 2 * B * 16)))
		return rc;
	VT *X = h->blk_x.as<VT>(), *Y = h->blk_y.as<VT>();
	double2 *part = h->blk_part.as<double2>();
	constexpr int kTrial = 3;
	block_init<VT><<<grid_for(h->N * B), kThreads, 0, h->stream>>>(X, h->N, B, B, 1,
	                                                              nullptr, 1, 0, 0);
	// live = 0: no moment gathers (there are no sites), full vector traffic.
	if ((rc = launch_step<VT, HT, B>(r, true, dots, X, Y, part, nullptr, 0, nullptr, 1, 0, 0)))
		return rc;
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	for (int i = 0; i < kTrial; i++) {
		if ((rc = launch_step<VT, HT, B>(r, false, dots, Y, X, part + (int64_t)(i + 1) * r.nwg * 2 * B,
		                                 nullptr, 0, nullptr, 1, 0, 0)))
			return rc;
		VT *t = X;
		X = Y;
		Y = t;
	}
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	TBTKX_CHECK_HIP(hipEventSynchronize(h->ev_stop));
	TBTKX_CHECK_HIP(hipGetLastError());
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	*ms_per_step = ms / kTrial;
	return TBTKX_OK;
}

// Launch-shape autotuner: every free parameter among the block width B, the
// strip rows S (natural order or two strip widths sized from the L2: rows
// of about 1 MB and 256 KB of Phi) and the grid share is tried for three
// steps on the resident model; the shape with the least time for all blocks
// of the call wins.
// `total` counts complex columns (packed real vectors: two per column).
template <typename VT, typename HT>
int tune_shape(tbtkx_context *h, int total, int flags, int *best_B, int *best_S,
               int *best_share, int *best_rmul) {
	std::vector<int> widths;
	if (h->block_width == 64 || h->block_width == 128 || h->block_width == 256)
		widths.push_back(h->block_width);
	else
		widths = {64, 128, 256};
	const bool ell = h->ell_W > 0 && h->use_ell;
	double best = 0;
	*best_B = widths[0];
	*best_S = -1;
	*best_share = 1;
	*best_rmul = 1;
	const int cap = (total + 63) / 64 * 64;
	for (int B : widths) {
		if (B > cap && B != widths[0]) continue;
		std::vector<int> strips;
		if (h->strip_rows != 0 || !ell) {
			strips.push_back(ell ? h->strip_rows : -1);
		} else {
			strips.push_back(-1);
			const int64_t rowbytes = (int64_t)B * sizeof(VT);
			for (int64_t target : {(int64_t)1 << 20, (int64_t)1 << 18}) {
				int S = 32;
				while ((int64_t)S * 2 * rowbytes <= target) S *= 2;
				if (h->band >= 2 * S) strips.push_back(S);
			}
		}
		std::vector<int> shares;
		if (h->grid_share > 0) shares.push_back(h->grid_share);
		else shares = {1, 2, 4};
		std::vector<int> rmuls;
		if (h->rows_in_flight > 0) rmuls.push_back(h->rows_in_flight);
		else if (kFp32Pair<VT, HT>) rmuls = {1, 2};
		else rmuls.push_back(1);
		for (int S : strips)
			for (int share : shares)
				for (int rmul : rmuls) {
					if (S < 0 && (share > 2 || rmul > 1)) continue;
					float ms = 0;
					int rc = B == 64 ? trial_width<VT, HT, 64>(h, flags, S, share, rmul, &ms)
					         : B == 128
					             ? trial_width<VT, HT, 128>(h, flags, S, share, rmul, &ms)
					             : trial_width<VT, HT, 256>(h, flags, S, share, rmul, &ms);
					if (rc) return rc;
					const double cost = (double)ms * ((total + B - 1) / B);
					if (best == 0 || cost < best) {
						best = cost;
						*best_B = B;
						*best_S = S;
						*best_share = share;
						*best_rmul = rmul;
					}
				}
	}
	return TBTKX_OK;
}

template <typename VT, typename HT>
int dispatch_width(tbtkx_context *h, int B, int total, const int64_t *sites,
                   int mode, uint64_t seed, int M, int flags, int packed,
                   std::vector<double2> &mu) {
	switch (B) {
	case 64: return run_all<VT, HT, 64>(h, total, sites, mode, seed, M, flags, packed, mu);
	case 128: return run_all<VT, HT, 128>(h, total, sites, mode, seed, M, flags, packed, mu);
	default: return run_all<VT, HT, 256>(h, total, sites, mode, seed, M, flags, packed, mu);
	}
}

// Exact FP64 blocks for many unit start vectors (tbtkx_moments_multi): each
// column (or packed real half-column) runs the plain recurrence of its start
// vector bit for bit (spmm_step<..., EXACT>: CSR order, no fused
// multiply-add), and after every step a gather list copies the requested
// components of Phi_n into the moments. Vectors are taken vpb at a time
// (B columns, two real vectors per column when packed); entries of the
// gather list are ordered by vector, so every block gathers one contiguous
// range of them.
template <int B>
int run_exact(tbtkx_context *h, int nvec, const int64_t *from, const std::vector<int64_t> &grow,
              const std::vector<int> &gvec, int M, int packed, std::vector<double2> &mu) {
	constexpr int kRpi = B >= 256 ? 1 : 256 / B;
	int tile = h->tile_rows > 0 ? h->tile_rows : 4;
	tile = (tile + kRpi - 1) / kRpi * kRpi;
	if (tile > 32) tile = 32;
	BlockRun r{h, B, TBTKX_VEC_FP64 | TBTKX_H_FP64, 0, tile, 0};
	r.exact = true;
	int rc = persistent_grid<double2, double2, B>(h, false, r);
	if (rc) return rc;
	const int vpb = packed ? 2 * B : B;
	const int nblocks = (nvec + vpb - 1) / vpb;
	const int L = (int)grow.size();
	// first gather entry of every block (entries are ordered by vector)
	std::vector<int> first(nblocks + 1, L);
	for (int l = L - 1; l >= 0; l--) first[gvec[l] / vpb] = l;
	for (int b = nblocks - 1; b >= 0; b--)
		if (first[b] > first[b + 1]) first[b] = first[b + 1];
	std::vector<int> lvec(L);
	for (int l = 0; l < L; l++) lvec[l] = gvec[l] % vpb;
	if ((rc = h->blk_x.reserve((size_t)h->N * B * 16))) return rc;
	if ((rc = h->blk_y.reserve((size_t)h->N * B * 16))) return rc;
	if ((rc = h->blk_mu.reserve((size_t)(L > 0 ? L : 1) * M * 16))) return rc;
	if ((rc = h->blk_sites.reserve((size_t)nvec * 8))) return rc;
	if ((rc = h->blk_grow.reserve((size_t)(L > 0 ? L : 1) * 8))) return rc;
	if ((rc = h->blk_gvec.reserve((size_t)(L > 0 ? L : 1) * 4))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->blk_sites.ptr, from, (size_t)nvec * 8,
	                               hipMemcpyHostToDevice, h->stream));
	if (L > 0) {
		TBTKX_CHECK_HIP(hipMemcpyAsync(h->blk_grow.ptr, grow.data(), (size_t)L * 8,
		                               hipMemcpyHostToDevice, h->stream));
		TBTKX_CHECK_HIP(hipMemcpyAsync(h->blk_gvec.ptr, lvec.data(), (size_t)L * 4,
		                               hipMemcpyHostToDevice, h->stream));
	}
	double2 *X0 = h->blk_x.as<double2>(), *Y0 = h->blk_y.as<double2>();
	double2 *mu_dev = h->blk_mu.as<double2>();
	const int64_t *grow_dev = h->blk_grow.as<int64_t>();
	const int *gvec_dev = h->blk_gvec.as<int>();
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	for (int blk = 0; blk < nblocks; blk++) {
		const int live = nvec - blk * vpb < vpb ? nvec - blk * vpb : vpb;
		const int g0 = first[blk], gl = first[blk + 1] - first[blk];
		double2 *X = X0, *Y = Y0;
		block_init<double2><<<grid_for(h->N * B), kThreads, 0, h->stream>>>(
		    X, h->N, B, live, 0, h->blk_sites.as<int64_t>() + (int64_t)blk * vpb, 0, 0, packed);
		for (int n = 0; n < M; n++) {
			if (gl > 0)
				gather_list<double2><<<grid_for(gl), kThreads, 0, h->stream>>>(
				    X, grow_dev + g0, gvec_dev + g0, gl, B, mu_dev + (int64_t)g0 * M, M, n,
				    packed);
			if (n == M - 1) break;
			// Phi_{n+1} from Phi_n (X) and Phi_{n-1} (Y), into Y.
			if ((rc = launch_step<double2, double2, B>(r, n == 0, false, X, Y, nullptr,
			                                           nullptr, 0, nullptr, M, n, packed)))
				return rc;
			double2 *t = X;
			X = Y;
			Y = t;
		}
		TBTKX_CHECK_HIP(hipGetLastError());
	}
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	mu.assign((size_t)L * M, make_double2(0, 0));
	if (L > 0)
		TBTKX_CHECK_HIP(hipMemcpyAsync(mu.data(), mu_dev, (size_t)L * M * 16,
		                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_launches = (int64_t)nblocks * (M - 1);
	h->last_block_width = B;
	h->last_real = packed;
	return TBTKX_OK;
}

// Algorithmic HBM bytes of the whole call: plain recurrences (one per start
// vector, in the plain kernel's current form) or exact FP64 blocks of width
// B (20 Z + 4 N for FP64 H and row pointers, 48 B N for the three block
// passes, per step and block).
double plain_call_bytes(const tbtkx_context *h, int nvec) {
	return plain_form(h).bytes * nvec;
}
double block_call_bytes(const tbtkx_context *h, int nvec, int B, int packed) {
	const double N = (double)h->N, Z = (double)h->Z;
	const int cols = packed ? (nvec + 1) / 2 : nvec;
	const int blocks = (cols + B - 1) / B;
	return blocks * (20 * Z + 4 * N + 48.0 * B * N);
}

int block_entry(tbtkx_context *h, int total, const int64_t *sites, int mode,
                uint64_t seed, int M, int flags, std::vector<double2> &mu) {
	int rc = set_device(h);
	if (rc) return rc;
	if (flags & TBTKX_VEC_FP64) flags |= TBTKX_H_FP64;
	int B = h->block_width;
	if (B == 32) B = 0;
	// Unit start vectors on a real H: two real vectors per complex column.
	const int packed = mode == 0 && h->real_path && h->h_real;
	const int cols = packed ? (total + 1) / 2 : total;
	const bool ell = h->ell_W > 0 && h->use_ell;
	h->run_strip = ell ? (h->strip_rows != 0 ? h->strip_rows : -1) : -1;
	h->run_share = h->grid_share > 0 ? h->grid_share : 1;
	h->run_rmul = h->rows_in_flight > 0 ? h->rows_in_flight : 1;
	const bool free_shape = B == 0 || (ell && h->strip_rows == 0) || h->grid_share == 0 ||
	                        (ell && h->rows_in_flight == 0);
	if (free_shape && !h->autotune) {
		if (B == 0) B = cols <= 64 ? 64 : 128;
	} else if (free_shape) {
		// Launch-shape autotuner: time the candidates on the resident
		// model once per (model, precision, mode, size class, options) and
		// keep the fastest for the whole call.
		const int64_t key[8] = {h->csr_gen, flags + (packed << 16),
		                        cols <= 64 ? 64 : cols <= 128 ? 128 : 256,
		                        h->use_ell * 4 + h->nt_stream * 2 + h->xcd_remap,
		                        h->tile_rows * 65536 + h->strip_chunk, h->block_width,
		                        h->strip_rows, h->grid_share * 16 + h->rows_in_flight};
		if (memcmp(key, h->blk_tune_key, sizeof key) != 0) {
			int tb = 64, ts = -1, tsh = 1, trm = 1;
			if (flags & TBTKX_VEC_FP64)
				rc = tune_shape<double2, double2>(h, cols, flags, &tb, &ts, &tsh, &trm);
			else if (flags & TBTKX_H_FP64)
				rc = tune_shape<float2, double2>(h, cols, flags, &tb, &ts, &tsh, &trm);
			else if (h->h_real)
				rc = tune_shape<float2, RealD>(h, cols, flags, &tb, &ts, &tsh, &trm);
			else
				rc = tune_shape<float2, float2>(h, cols, flags, &tb, &ts, &tsh, &trm);
			if (rc) return rc;
			memcpy(h->blk_tune_key, key, sizeof key);
			h->blk_tune_B = tb;
			h->blk_tune_S = ts;
			h->blk_tune_share = tsh;
			h->blk_tune_rmul = trm;
		}
		B = h->blk_tune_B;
		h->run_strip = h->blk_tune_S;
		h->run_share = h->blk_tune_share;
		h->run_rmul = h->blk_tune_rmul;
	}
	if (flags & TBTKX_VEC_FP64)
		return dispatch_width<double2, double2>(h, B, total, sites, mode, seed, M,
		                                        flags, packed, mu);
	if (flags & TBTKX_H_FP64)
		return dispatch_width<float2, double2>(h, B, total, sites, mode, seed, M,
		                                       flags, packed, mu);
	// A real H's 8-byte copies hold FP64 values (tbtkx_set_csr, tbtkx_set_lattice).
	if (h->h_real)
		return dispatch_width<float2, RealD>(h, B, total, sites, mode, seed, M, flags,
		                                     packed, mu);
	return dispatch_width<float2, float2>(h, B, total, sites, mode, seed, M, flags,
	                                      packed, mu);
}

}  // namespace
}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

int tbtkx_moments_block(tbtkx_handle h, int num_sites, const int64_t *sites,
                        int M, int flags, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_moments_block: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_moments_block: no model set");
	TBTKX_REQUIRE(M > 0, TBTKX_ERR_ARG,
	              "tbtkx_moments_block: numCoefficients has to be larger than 0");
	TBTKX_REQUIRE(num_sites >= 0 && (num_sites == 0 || (sites && out)),
	              TBTKX_ERR_ARG, "tbtkx_moments_block: bad site list");
	TBTKX_REQUIRE((flags & ~(TBTKX_H_FP64 | TBTKX_DOUBLING | TBTKX_VEC_FP64)) == 0,
	              TBTKX_ERR_ARG, "tbtkx_moments_block: unknown flags");
	for (int b = 0; b < num_sites; b++)
		TBTKX_REQUIRE(sites[b] >= 0 && sites[b] < h->N, TBTKX_ERR_ARG,
		              "tbtkx_moments_block: site outside the basis");
	if (num_sites == 0) return TBTKX_OK;
	std::vector<double2> mu;
	int rc = block_entry(h, num_sites, sites, 0, 0, M, flags, mu);
	if (rc) return rc;
	memcpy(out, mu.data(), mu.size() * 16);
	return TBTKX_OK;
}

int tbtkx_moments_multi(tbtkx_handle h, int num_from, const int64_t *from,
                        const int *to_offsets, const int64_t *to, int M, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_moments_multi: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_moments_multi: no model set");
	TBTKX_REQUIRE(M > 0, TBTKX_ERR_ARG,
	              "tbtkx_moments_multi: numCoefficients has to be larger than 0");
	TBTKX_REQUIRE(num_from >= 0 && (num_from == 0 || (from && to_offsets)), TBTKX_ERR_ARG,
	              "tbtkx_moments_multi: bad from-index list");
	if (num_from == 0) return TBTKX_OK;
	TBTKX_REQUIRE(to_offsets[0] == 0, TBTKX_ERR_ARG,
	              "tbtkx_moments_multi: to_offsets[0] must be 0");
	for (int j = 0; j < num_from; j++) {
		TBTKX_REQUIRE(from[j] >= 0 && from[j] < h->N, TBTKX_ERR_ARG,
		              "tbtkx_moments_multi: from-index outside the basis");
		TBTKX_REQUIRE(to_offsets[j + 1] >= to_offsets[j], TBTKX_ERR_ARG,
		              "tbtkx_moments_multi: to_offsets not ascending");
	}
	const int T = to_offsets[num_from];
	TBTKX_REQUIRE(T == 0 || (to && out), TBTKX_ERR_ARG, "tbtkx_moments_multi: null array");
	for (int k = 0; k < T; k++)
		TBTKX_REQUIRE(to[k] >= 0 && to[k] < h->N, TBTKX_ERR_ARG,
		              "tbtkx_moments_multi: to-index outside the basis");
	if (T == 0) return TBTKX_OK;
	int rc = set_device(h);
	if (!rc) rc = build_plain_forms(h);
	if (rc) return rc;
	const int packed = h->real_path && h->h_real;
	int B = 64;
	for (int b : {128, 256})
		if (block_call_bytes(h, num_from, b, packed) < block_call_bytes(h, num_from, B, packed))
			B = b;
	bool blocks = h->multi_mode == 2 ||
	              (h->multi_mode == 0 &&
	               block_call_bytes(h, num_from, B, packed) < plain_call_bytes(h, num_from));
	if (!blocks) {
		for (int j = 0; j < num_from; j++) {
			const int k0 = to_offsets[j], k1 = to_offsets[j + 1];
			if (k1 == k0) continue;
			if ((rc = tbtkx_moments(h, from[j], k1 - k0, to + k0, M, out + (size_t)k0 * M * 2)))
				return rc;
		}
		h->last_multi_blocks = 0;
		return TBTKX_OK;
	}
	// Gather entries: per from-index its distinct to-indices (a repeated
	// to-index keeps only its last position, as tbtkx_moments), ordered by
	// from-index.
	std::vector<int64_t> grow;
	std::vector<int> gvec, gslot;
	for (int j = 0; j < num_from; j++) {
		std::map<int64_t, int> last;
		for (int k = to_offsets[j]; k < to_offsets[j + 1]; k++) last[to[k]] = k;
		for (auto &kv : last) {
			grow.push_back(kv.first);
			gvec.push_back(j);
			gslot.push_back(kv.second);
		}
	}
	std::vector<double2> mu;
	if (B == 64) rc = run_exact<64>(h, num_from, from, grow, gvec, M, packed, mu);
	else if (B == 128) rc = run_exact<128>(h, num_from, from, grow, gvec, M, packed, mu);
	else rc = run_exact<256>(h, num_from, from, grow, gvec, M, packed, mu);
	if (rc) return rc;
	memset(out, 0, (size_t)T * M * 16);
	for (size_t l = 0; l < grow.size(); l++)
		memcpy(out + (size_t)gslot[l] * M * 2, mu.data() + l * M, (size_t)M * 16);
	h->last_multi_blocks = 1;
	return TBTKX_OK;
}

int tbtkx_moments_trace(tbtkx_handle h, int R, uint64_t seed, int M, int flags,
                        double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_moments_trace: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_moments_trace: no model set");
	TBTKX_REQUIRE(M > 0 && R > 0 && out, TBTKX_ERR_ARG,
	              "tbtkx_moments_trace: need M > 0, R > 0 and an output array");
	TBTKX_REQUIRE((flags & ~(TBTKX_H_FP64 | TBTKX_DOUBLING | TBTKX_VEC_FP64)) == 0,
	              TBTKX_ERR_ARG, "tbtkx_moments_trace: unknown flags");
	// The random start vectors are never stored, so the moments come from the
	// doubling identities.
	std::vector<double2> mu;
	int rc = block_entry(h, R, nullptr, 1, seed, M, flags | TBTKX_DOUBLING, mu);
	if (rc) return rc;
	for (int n = 0; n < M; n++) {
		double sx = 0, sy = 0;
		for (int r = 0; r < R; r++) {
			sx += mu[(size_t)r * M + n].x;
			sy += mu[(size_t)r * M + n].y;
		}
		out[2 * n] = sx;
		out[2 * n + 1] = sy;
	}
	return TBTKX_OK;
}

}  // extern "C"
