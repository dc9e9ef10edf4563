// Plain FP64 path: one start vector, the three-term Chebyshev recurrence
//   j_1 = (H/s) j_0,   j_n = (2H/s) j_{n-1} - j_{n-2},
// with the moments <to_k|j_n> gathered as the recurrence runs.
//
// Arithmetic order is that of the reference CPU recurrence
// (Lib/src/Solver/ChebyshevExpander.cpp:65-82, 157-186): every row starts
// from -j_{n-2}[row] (0 for j_1) and adds v_k * x[col_k] for k ascending, each
// product formed as (vr*xr - vi*xi, vr*xi + vi*xr) with no fused multiply-add.
// The moments are therefore bitwise equal to the reference's.
//
// HBM traffic per step (algorithmic): 20 B per non-zero (16 B value + 4 B
// column), 4 B row pointer, 16 B read of x, 16 B read and 16 B write of the
// in-place j_{n-2} -> j_n row: 20 Z + 52 N bytes. For a real H (and the real
// unit start vector) the recurrence runs on real FP64 values and vectors:
// 12 Z + 28 N bytes. With 16-bit columns (offsets from the first row of
// the row block, used when every offset fits) a non-zero costs 2 bytes less:
// 18 Z + 52 N complex, 10 Z + 28 N real. With the value palette as well
// (a 1-byte code per non-zero: an index into at most 255 distinct
// off-diagonal values, or 255 for the row's diagonal, read from a per-row
// array): 3 Z + 68 N complex, 3 Z + 36 N real.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "common.hpp"

namespace tbtkx {
namespace {

// FP64 arithmetic of one recurrence element: complex (double2) or, for a
// real H and a real start vector, real (double). A complex product is formed
// as (vr*xr - vi*xi, vr*xi + vi*xr) with no fused multiply-add. With real
// operands the complex recurrence keeps every imaginary part exactly zero,
// and its real parts come out of the same operations in the same order as
// the real recurrence's values, so both give the same moments (a zero may
// differ in sign).
template <typename S> struct F64;
template <> struct F64<double2> {
	__device__ static double2 mul(double2 a, double2 c) {
#pragma clang fp contract(off)
		return make_double2(a.x * c.x - a.y * c.y, a.x * c.y + a.y * c.x);
	}
	__device__ static double2 add(double2 a, double2 b) {
#pragma clang fp contract(off)
		return make_double2(a.x + b.x, a.y + b.y);
	}
	__device__ static double2 neg(double2 p) { return make_double2(-p.x, -p.y); }
	__device__ static double2 zero() { return make_double2(0.0, 0.0); }
	__device__ static double2 unit(bool one) { return make_double2(one ? 1.0 : 0.0, 0.0); }
	__device__ static double2 moment(double2 x) { return x; }
};
template <> struct F64<double> {
	__device__ static double mul(double a, double c) {
#pragma clang fp contract(off)
		return a * c;
	}
	__device__ static double add(double a, double b) {
#pragma clang fp contract(off)
		return a + b;
	}
	__device__ static double neg(double p) { return -p; }
	__device__ static double zero() { return 0.0; }
	__device__ static double unit(bool one) { return one ? 1.0 : 0.0; }
	__device__ static double2 moment(double x) { return make_double2(x, 0.0); }
};

// Column of non-zero k in a row block starting at row r0: the 32-bit
// column itself, or a 16-bit offset from r0.
__device__ __forceinline__ uint32_t col_at(const uint32_t *col, uint32_t k, uint32_t) {
	return col[k];
}
__device__ __forceinline__ uint32_t col_at(const int16_t *col, uint32_t k, uint32_t r0) {
	return (uint32_t)((int32_t)r0 + (int32_t)col[k]);
}

// 16-bit column offsets: workgroup b writes col[k] - r0 for the non-zeros of
// row block b and raises the flag if one does not fit in int16.
__global__ __launch_bounds__(kThreads) void make_col16(
    const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ rblk, int16_t *__restrict__ col16, int *overflow) {
	const uint32_t r0 = rblk[blockIdx.x], r1 = rblk[blockIdx.x + 1];
	const uint32_t k0 = rowptr[r0], k1 = rowptr[r1];
	for (uint32_t k = k0 + threadIdx.x; k < k1; k += kThreads) {
		const int64_t d = (int64_t)col[k] - (int64_t)r0;
		const bool fits = d >= -32768 && d <= 32767;
		col16[k] = fits ? (int16_t)d : (int16_t)0;
		if (!fits) atomicOr(overflow, 1);
	}
}

// Slot-major ELL form of the plain kernel (TBTKX_OPT_PLAIN_KERNEL = 2):
// W slots per row (W >= the longest row), slot s of row r at s*N + r, so a
// wavefront's 64 rows read each slot with one coalesced load. A slot holds
// the column as a 16-bit offset from the row itself (kSlotPad marks the
// padding after a row's last non-zero) and the value as a palette code
// (PAL) or as H/s and 2H/s. The slots of a row keep the CSR order (ascending
// columns), so the sums are the reference's, bit for bit.
constexpr int16_t kSlotPad = -32768;

template <typename C, typename R>
__global__ __launch_bounds__(kThreads) void make_slots(
    const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
    const uint8_t *__restrict__ pcode, const C *__restrict__ v1, const C *__restrict__ v2,
    const R *__restrict__ v1r, const R *__restrict__ v2r, int64_t N, int W,
    int16_t *__restrict__ scol, uint8_t *__restrict__ scode, C *__restrict__ sv1,
    C *__restrict__ sv2, R *__restrict__ sv1r, R *__restrict__ sv2r, int *overflow) {
	const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (r >= N) return;
	const uint32_t k0 = rowptr[r];
	const int len = (int)(rowptr[r + 1] - k0);
	for (int s = 0; s < W; s++) {
		const int64_t i = (int64_t)s * N + r;
		if (s < len) {
			const uint32_t k = k0 + s;
			const int64_t d = (int64_t)col[k] - r;
			const bool fits = d >= -32767 && d <= 32767;
			scol[i] = fits ? (int16_t)d : (int16_t)0;
			if (!fits) atomicOr(overflow, 1);
			if (scode) scode[i] = pcode[k];
			if (sv1) {
				sv1[i] = v1[k];
				sv2[i] = v2[k];
			}
			if (sv1r) {
				sv1r[i] = v1r[k];
				sv2r[i] = v2r[k];
			}
		} else {
			scol[i] = kSlotPad;
			if (scode) scode[i] = 0;
			if (sv1) {
				sv1[i] = C{};
				sv2[i] = C{};
			}
			if (sv1r) {
				sv1r[i] = R{};
				sv2r[i] = R{};
			}
		}
	}
}

// One thread per row on the slot-major form: the row's W column offsets and
// codes (or values), its j_{n-2} entry and diagonal value are loaded at
// once, then its W gathers of x, then the products are added in slot order
// from -j_{n-2} (padding slots skipped). Two dependent memory trips per
// step and no LDS or barrier.
// PAT (with PAL): the row's slots come from its row pattern instead, the
// pattern id pid[r] indexing a table of 8 offsets (scol) and 8 codes (scode)
// per pattern. R rows per thread (rows b*R*256 + j*256 + tid, j < R): all
// their loads are issued before the first sum, R times the loads in flight.
// With pchunk (one byte per 64-row chunk: the pattern all its rows share, or
// 255), a wavefront whose R chunks share one pattern reads its offsets,
// codes and palette values through scalar loads at wave-uniform addresses
// and no per-row pattern ids: the gathers wait for a cache hit instead of a
// memory round trip. Same values, same order, same results.
template <typename S, bool FIRST, bool PAL, int W, bool PAT = false, int R = 1>
__global__ __launch_bounds__(kThreads) void plain_step_slot(
    const int16_t *__restrict__ scol, const uint8_t *__restrict__ scode,
    const S *__restrict__ sval, const S *__restrict__ pal, const S *__restrict__ diag,
    const S *__restrict__ x, S *y, int64_t N, const int64_t *__restrict__ to, int K,
    double2 *__restrict__ coef, int M, int nprev, const uint8_t *__restrict__ pid,
    int remap, const uint8_t *__restrict__ pchunk) {
#pragma clang fp contract(off)
	typedef F64<S> A;
	for (int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x; g < K;
	     g += (int64_t)gridDim.x * kThreads)
		coef[g * M + nprev] = A::moment(x[to[g]]);
	// remap: each XCD takes one contiguous eighth of the rows, so the x
	// entries of a row's far neighbours (a lattice stride away) are read
	// through the same L2 as the row's own (measured on the 1024x1024 spin
	// lattice: HBM reads equal to the algorithmic bytes, 1.34x without).
	const int64_t b = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
	int64_t row[R];
#pragma unroll
	for (int j = 0; j < R; j++) {
		const int64_t rj = (b * R + j) * kThreads + threadIdx.x;
		row[j] = rj < N ? rj : N - 1;   // rows past the end load row N-1, store nothing
	}
	if ((b * R) * kThreads + threadIdx.x >= N) return;
	// j_{n-2} and the diagonal value, which do not depend on the row's slots.
	// Vector loads complete in issue order, so they are issued after the
	// loads the gathers' addresses wait for.
	S yo[R], dg[R];
	auto load_row_values = [&]() {
#pragma unroll
		for (int j = 0; j < R; j++) {
			yo[j] = FIRST ? A::zero() : y[row[j]];
			dg[j] = PAL ? diag[row[j]] : A::zero();
		}
	};
	if (PAT && pchunk) {
		// The wavefront's R chunks of 64 rows are chunks (b R + j) 4 + w:
		// byte w of 32-bit word b R + j.
		const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
		const uint32_t *pc = reinterpret_cast<const uint32_t *>(pchunk);
		int up = (int)((pc[b * R] >> (8 * w)) & 255u);
#pragma unroll
		for (int j = 1; j < R; j++)
			if ((int)((pc[b * R + j] >> (8 * w)) & 255u) != up) up = 255;
		up = __builtin_amdgcn_readfirstlane(up);
		if (up != 255) {
			// Every row of the wavefront is below N (a chunk past the end is
			// never uniform), so no store is guarded.
			load_row_values();
			// Whole words through readfirstlane, so that the compiler keeps
			// them as scalar loads instead of narrowing one half-word to a
			// vector load (which would wait on the row loads above).
			const int4 o = *reinterpret_cast<const int4 *>(scol + up * 8);
			const int2 e = *reinterpret_cast<const int2 *>(scode + up * 8);
			const uint32_t ow[4] = {(uint32_t)__builtin_amdgcn_readfirstlane(o.x),
			                        (uint32_t)__builtin_amdgcn_readfirstlane(o.y),
			                        (uint32_t)__builtin_amdgcn_readfirstlane(o.z),
			                        (uint32_t)__builtin_amdgcn_readfirstlane(o.w)};
			const uint32_t ew[2] = {(uint32_t)__builtin_amdgcn_readfirstlane(e.x),
			                        (uint32_t)__builtin_amdgcn_readfirstlane(e.y)};
			int16_t oo[8];
			uint8_t ee[8];
#pragma unroll
			for (int s = 0; s < 8; s++) {
				oo[s] = (int16_t)(ow[s >> 1] >> (16 * (s & 1)));
				ee[s] = (uint8_t)(ew[s >> 2] >> (8 * (s & 3)));
			}
			S pv[W], xu[R][W];
			// Unconditional loads (entry 0 stands in for the diagonal code;
			// each palette section has 256 entries), so they issue together.
#pragma unroll
			for (int s = 0; s < W; s++) pv[s] = pal[ee[s] == 255 ? 0 : ee[s]];
#pragma unroll
			for (int j = 0; j < R; j++)
#pragma unroll
				for (int s = 0; s < W; s++)
					xu[j][s] = x[row[j] + (oo[s] == kSlotPad ? 0 : (int64_t)oo[s])];
			// Every row's gathers are in flight before the first product (the
			// scheduler would otherwise issue row j + 1's gathers only after
			// row j's sum, one memory round trip per row).
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int j = 0; j < R; j++) {
				S acc = FIRST ? A::zero() : A::neg(yo[j]);
#pragma unroll
				for (int s = 0; s < W; s++) {
					const S p = A::mul(ee[s] == 255 ? dg[j] : pv[s], xu[j][s]);
					if (oo[s] != kSlotPad) acc = A::add(acc, p);
				}
				y[row[j]] = acc;
			}
			return;
		}
	}
	int16_t c[R][W];
	uint8_t q[R][W];
	S v[R][W], xx[R][W];
#pragma unroll
	for (int j = 0; j < R; j++) {
		const int64_t r = row[j];
		if (PAT) {
			const int p = pid[r];
			const int4 o = *reinterpret_cast<const int4 *>(scol + p * 8);
			const int2 e = *reinterpret_cast<const int2 *>(scode + p * 8);
			int16_t oo[8];
			uint8_t ee[8];
			memcpy(oo, &o, 16);
			memcpy(ee, &e, 8);
#pragma unroll
			for (int s = 0; s < W; s++) {
				c[j][s] = oo[s];
				q[j][s] = ee[s];
			}
		} else {
#pragma unroll
			for (int s = 0; s < W; s++) {
				c[j][s] = scol[(int64_t)s * N + r];
				if (PAL) q[j][s] = scode[(int64_t)s * N + r];
				else v[j][s] = sval[(int64_t)s * N + r];
			}
		}
	}
	load_row_values();
#pragma unroll
	for (int j = 0; j < R; j++)
#pragma unroll
		for (int s = 0; s < W; s++)
			xx[j][s] = x[row[j] + (c[j][s] == kSlotPad ? 0 : (int64_t)c[j][s])];
	if (PAL) {
#pragma unroll
		for (int j = 0; j < R; j++)
#pragma unroll
			for (int s = 0; s < W; s++) v[j][s] = q[j][s] == 255 ? dg[j] : pal[q[j][s]];
	}
#pragma unroll
	for (int j = 0; j < R; j++) {
		S acc = FIRST ? A::zero() : A::neg(yo[j]);
#pragma unroll
		for (int s = 0; s < W; s++) {
			const S p = A::mul(v[j][s], xx[j][s]);
			if (c[j][s] != kSlotPad) acc = A::add(acc, p);
		}
		if ((b * R + j) * kThreads + threadIdx.x < N) y[row[j]] = acc;
	}
}

// Row-block kernel: the workgroup streams its block's non-zeros with
// consecutive threads on consecutive non-zeros (coalesced value/column loads
// and gathers of x), forms the products in registers, stages them in LDS and
// then lets one thread per row add them up in column order.
// PAL: values from the palette (code < 255) or the per-row diagonal array
// (code 255, where the column is the row itself); else from v.
template <typename S, bool FIRST, typename CI, bool PAL>
__global__ __launch_bounds__(kThreads) void plain_step_lds(
    const uint32_t *__restrict__ rowptr, const CI *__restrict__ col,
    const S *__restrict__ v, const S *__restrict__ x,
    S *y, const uint32_t *__restrict__ rblk, int nblk, int remap,
    const int64_t *__restrict__ to, int K, double2 *__restrict__ coef,
    int M, int nprev, const uint8_t *__restrict__ code, const S *__restrict__ pal,
    const S *__restrict__ diag) {
#pragma clang fp contract(off)
	typedef F64<S> A;
	__shared__ S prod[kPlainCap];
	const int tid = threadIdx.x;
	// The palette (at most 255 values) is read from global memory, where it
	// stays cache-resident; staging it in LDS would put a load and a barrier
	// in front of every workgroup's first product.
	auto value = [&](uint32_t k, uint32_t c) -> S {
		if (!PAL) return v[k];
		const uint32_t q = code[k];
		return q == 255 ? diag[c] : pal[q];
	};

	// Moments of the previous vector (x = j_{nprev}); x is not written in
	// this launch, so any workgroup may read it.
	const int64_t g = (int64_t)blockIdx.x * kThreads + tid;
	if (g < K) coef[g * M + nprev] = A::moment(x[to[g]]);

	const int b = remap ? xcd_remap(blockIdx.x, nblk) : (int)blockIdx.x;
	const uint32_t r0 = rblk[b], r1 = rblk[b + 1];
	const uint32_t k0 = rowptr[r0], k1 = rowptr[r1];

	if (k1 - k0 <= (uint32_t)kPlainCap) {
		// Loads are issued for several non-zeros at once (measured fastest
		// per form): with the value palette four per iteration; with full
		// values all of a thread's non-zeros (at most kPlainCap / kThreads)
		// in one batch, indices past the block's last non-zero clamped to it
		// so every load is issued without a branch, and only the products
		// of real non-zeros stored.
		constexpr int U = (kPlainCap + kThreads - 1) / kThreads;
		if (PAL) {
			uint32_t k = k0 + tid;
			for (; k + 3 * kThreads < k1; k += 4 * kThreads) {
				uint32_t c[4];
				S vv[4], xx[4];
#pragma unroll
				for (int i = 0; i < 4; i++) c[i] = col_at(col, k + i * kThreads, r0);
#pragma unroll
				for (int i = 0; i < 4; i++) {
					vv[i] = value(k + i * kThreads, c[i]);
					xx[i] = x[c[i]];
				}
#pragma unroll
				for (int i = 0; i < 4; i++) prod[k + i * kThreads - k0] = A::mul(vv[i], xx[i]);
			}
			for (; k < k1; k += kThreads) {
				const uint32_t c = col_at(col, k, r0);
				prod[k - k0] = A::mul(value(k, c), x[c]);
			}
		} else if (k1 > k0) {
			uint32_t c[U];
			S vv[U], xx[U];
#pragma unroll
			for (int i = 0; i < U; i++) {
				const uint32_t k = k0 + tid + i * kThreads;
				c[i] = col_at(col, k < k1 ? k : k1 - 1, r0);
			}
#pragma unroll
			for (int i = 0; i < U; i++) {
				const uint32_t k = k0 + tid + i * kThreads;
				vv[i] = value(k < k1 ? k : k1 - 1, c[i]);
				xx[i] = x[c[i]];
			}
#pragma unroll
			for (int i = 0; i < U; i++) {
				const uint32_t k = k0 + tid + i * kThreads;
				if (k < k1) prod[k - k0] = A::mul(vv[i], xx[i]);
			}
		}
		__syncthreads();
		const uint32_t row = r0 + tid;
		if (row < r1) {
			S acc = FIRST ? A::zero() : A::neg(y[row]);
			const uint32_t e = rowptr[row + 1] - k0;
			for (uint32_t k = rowptr[row] - k0; k < e; k++) acc = A::add(acc, prod[k]);
			y[row] = acc;
		}
		return;
	}

	// A single row with more than kPlainCap non-zeros: stage it chunk by
	// chunk and let thread 0 keep the sequential sum.
	S acc = A::zero();
	if (tid == 0 && !FIRST) acc = A::neg(y[r0]);
	for (uint32_t base = k0; base < k1; base += kPlainCap) {
		const uint32_t end = base + kPlainCap < k1 ? base + kPlainCap : k1;
		for (uint32_t k = base + tid; k < end; k += kThreads) {
			const uint32_t c = col_at(col, k, r0);
			prod[k - base] = A::mul(value(k, c), x[c]);
		}
		__syncthreads();
		if (tid == 0)
			for (uint32_t k = 0; k < end - base; k++) acc = A::add(acc, prod[k]);
		__syncthreads();
	}
	if (tid == 0) y[r0] = acc;
}

// One thread per row, loads straight from memory (no LDS).
template <typename S, bool FIRST>
__global__ __launch_bounds__(kThreads) void plain_step_row(
    const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
    const S *__restrict__ v, const S *__restrict__ x,
    S *y, int64_t N, int nwg, int remap, const int64_t *__restrict__ to,
    int K, double2 *__restrict__ coef, int M, int nprev) {
#pragma clang fp contract(off)
	typedef F64<S> A;
	const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (g < K) coef[g * M + nprev] = A::moment(x[to[g]]);
	const int b = remap ? xcd_remap(blockIdx.x, nwg) : (int)blockIdx.x;
	const int64_t row = (int64_t)b * kThreads + threadIdx.x;
	if (row >= N) return;
	S acc = FIRST ? A::zero() : A::neg(y[row]);
	const uint32_t e = rowptr[row + 1];
	for (uint32_t k = rowptr[row]; k < e; k++) acc = A::add(acc, A::mul(v[k], x[col[k]]));
	y[row] = acc;
}

template <typename S>
__global__ __launch_bounds__(kThreads) void init_unit(S *a, int64_t N, int64_t from) {
	const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (i < N) a[i] = F64<S>::unit(i == from);
}

template <typename S>
__global__ __launch_bounds__(kThreads) void gather_moments(
    const S *__restrict__ x, const int64_t *__restrict__ to, int K,
    double2 *__restrict__ coef, int M, int n) {
	const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (g < K) coef[g * M + n] = F64<S>::moment(x[to[g]]);
}

inline unsigned grid_for(int64_t n) {
	return (unsigned)((n + kThreads - 1) / kThreads);
}

template <typename S, typename CI, bool PAL>
void launch_lds(tbtkx_context *h, bool first, const CI *col, const S *v, const S *x, S *y,
                const int64_t *to, int K, double2 *coef, int M, int nprev,
                const S *pal, const S *diag) {
	const unsigned grid = (unsigned)h->nblk;
	const uint8_t *code = PAL ? h->pcode.as<uint8_t>() : nullptr;
	if (first)
		plain_step_lds<S, true, CI, PAL><<<grid, kThreads, 0, h->stream>>>(
		    h->rowptr.as<uint32_t>(), col, v, x, y, h->rblk.as<uint32_t>(), h->nblk,
		    h->xcd_remap, to, K, coef, M, nprev, code, pal, diag);
	else
		plain_step_lds<S, false, CI, PAL><<<grid, kThreads, 0, h->stream>>>(
		    h->rowptr.as<uint32_t>(), col, v, x, y, h->rblk.as<uint32_t>(), h->nblk,
		    h->xcd_remap, to, K, coef, M, nprev, code, pal, diag);
}

// Rows per thread of the slot step: the option, or by default 1, and 2
// when every row reads its pattern byte (TBTKX_OPT_ROW_PATTERNS = 2: one
// byte per row leaves the loads of one row too few to cover the gathers'
// latency). Measured on the 1024x1024 spin lattice: wave-uniform patterns
// 12.05 us with 1 row per thread, 13.3 with 2, 13.6 with 4; pattern bytes
// for every row 15.9 with 2, 16.1 with 1; per-slot data 16.4 with 1,
// 16.7 with 2.
inline int slot_rows(const tbtkx_context *h, bool pat) {
	return h->slot_rows > 0 ? h->slot_rows : pat && h->use_patterns == 2 ? 2 : 1;
}

template <typename S, bool FIRST, bool PAL, bool PAT>
void launch_slot_w(tbtkx_context *h, int W, const int16_t *scol, const uint8_t *scode,
                   const S *sv, const S *pal, const S *diag, const S *x, S *y,
                   const int64_t *to, int K, double2 *coef, int M, int nprev,
                   const uint8_t *pid) {
	const int R = slot_rows(h, PAT);
	const unsigned grid = (unsigned)((h->N + (int64_t)kThreads * R - 1) / ((int64_t)kThreads * R));
	const uint8_t *pchunk = PAT && h->use_patterns == 1 ? h->pchunk.as<uint8_t>() : nullptr;
#define TBTKX_SLOT_R(WW, RR)                                                            \
	plain_step_slot<S, FIRST, PAL, WW, PAT, RR><<<grid, kThreads, 0, h->stream>>>(    \
	    scol, scode, sv, pal, diag, x, y, h->N, to, K, coef, M, nprev, pid, 1, pchunk)
#define TBTKX_SLOT(WW)                                                                  \
	do {                                                                            \
		if (R == 4) TBTKX_SLOT_R(WW, 4);                                        \
		else if (R == 2) TBTKX_SLOT_R(WW, 2);                                   \
		else TBTKX_SLOT_R(WW, 1);                                               \
	} while (0)
	switch (W) {
	case 4: TBTKX_SLOT(4); break;
	case 5: TBTKX_SLOT(5); break;
	case 6: TBTKX_SLOT(6); break;
	case 7: TBTKX_SLOT(7); break;
	default: TBTKX_SLOT(8); break;
	}
#undef TBTKX_SLOT
#undef TBTKX_SLOT_R
}

template <typename S>
void launch_slot(tbtkx_context *h, bool first, bool pal, bool pat, int W, const int16_t *scol,
                 const uint8_t *scode, const S *sv, const S *pp, const S *dp, const S *x, S *y,
                 const int64_t *to, int K, double2 *coef, int M, int nprev,
                 const uint8_t *pid) {
#define TBTKX_SLOTL(F, P, T)                                                            \
	launch_slot_w<S, F, P, T>(h, W, scol, scode, sv, pp, dp, x, y, to, K, coef, M, nprev, pid)
	if (pat && first) TBTKX_SLOTL(true, true, true);
	else if (pat) TBTKX_SLOTL(false, true, true);
	else if (first && pal) TBTKX_SLOTL(true, true, false);
	else if (first) TBTKX_SLOTL(true, false, false);
	else if (pal) TBTKX_SLOTL(false, true, false);
	else TBTKX_SLOTL(false, false, false);
#undef TBTKX_SLOTL
}

// Enqueue the step launches n = 1 .. M-1 and the final gather; S = double2
// (complex) or double (real H and start vector: v1r/v2r, vectors of N
// doubles in the same buffers).
template <typename S>
int enqueue_steps(tbtkx_context *h, int K, int M, S *a, S *b, double2 *coef,
                  const int64_t *to) {
	const bool real = std::is_same<S, double>::value;
	const S *v1 = real ? h->v1r.as<S>() : h->v1.as<S>();
	const S *v2 = real ? h->v2r.as<S>() : h->v2.as<S>();
	const uint32_t *rp = h->rowptr.as<uint32_t>(), *cl = h->col.as<uint32_t>();
	const int remap = h->xcd_remap;
	const bool lds = h->plain_kernel != 1;
	const bool c16 = lds && h->use_col16 && h->col16_ok;
	const bool usepal = c16 && h->use_palette && h->palette_ok;
	const S *palp = real ? h->palr.as<S>() : h->pal.as<S>();
	const S *diagp = real ? h->pdiagr.as<S>() : h->pdiag.as<S>();
	const int nwg_row = (int)grid_for(h->N);
	const unsigned grid = lds ? (unsigned)h->nblk : (unsigned)nwg_row;
	const PlainForm pf = plain_form(h);
	// The moment gather rides on the step launches; their grid covers
	// K <= N threads (to-indices are unique basis indices).
	for (int n = 1; n < M; n++) {
		const S *x = (n & 1) ? a : b;
		S *y = (n & 1) ? b : a;
		const S *v = n == 1 ? v1 : v2;
		if (pf.slot) {
			const bool first = n == 1;
			const S *sv = nullptr;
			if (!pf.pal)
				sv = real ? (first ? h->sv1r.as<S>() : h->sv2r.as<S>())
				          : (first ? h->sv1.as<S>() : h->sv2.as<S>());
			const S *pp = pf.pal ? palp + (first ? 0 : 256) : nullptr;
			const S *dp = pf.pal ? diagp + (first ? 0 : h->N) : nullptr;
			if (pf.pat)
				launch_slot<S>(h, first, true, true, h->slot_W, h->ptoff.as<int16_t>(),
				               h->ptcode.as<uint8_t>(), sv, pp, dp, x, y, to, K, coef, M, n - 1,
				               h->ppid.as<uint8_t>());
			else
				launch_slot<S>(h, first, pf.pal, false, h->slot_W, h->scol.as<int16_t>(),
				               h->scode.as<uint8_t>(), sv, pp, dp, x, y, to, K, coef, M, n - 1,
				               nullptr);
			continue;
		}
		if (lds) {
			if (usepal)
				launch_lds<S, int16_t, true>(h, n == 1, h->col16.as<int16_t>(), v, x, y, to,
				                             K, coef, M, n - 1, palp + (n == 1 ? 0 : 256),
				                             diagp + (n == 1 ? 0 : h->N));
			else if (c16)
				launch_lds<S, int16_t, false>(h, n == 1, h->col16.as<int16_t>(), v, x, y, to,
				                              K, coef, M, n - 1, nullptr, nullptr);
			else
				launch_lds<S, uint32_t, false>(h, n == 1, cl, v, x, y, to, K, coef, M, n - 1,
				                               nullptr, nullptr);
		} else {
			if (n == 1)
				plain_step_row<S, true><<<grid, kThreads, 0, h->stream>>>(
				    rp, cl, v, x, y, h->N, nwg_row, remap, to, K, coef,
				    M, n - 1);
			else
				plain_step_row<S, false><<<grid, kThreads, 0, h->stream>>>(
				    rp, cl, v, x, y, h->N, nwg_row, remap, to, K, coef,
				    M, n - 1);
		}
	}
	const S *last = ((M - 1) & 1) ? b : a;
	gather_moments<S><<<grid_for(K), kThreads, 0, h->stream>>>(last, to, K, coef, M, M - 1);
	TBTKX_CHECK_HIP(hipGetLastError());
	return TBTKX_OK;
}

}  // namespace

// Slot columns, and palette codes when the palette exists (values are added
// by build_slot_values when a call needs them).
static int make_slot_arrays(tbtkx_context *h, bool codes, bool values) {
	const int W = h->slot_W;
	const int64_t WN = (int64_t)W * h->N;
	int rc;
	if ((rc = h->scol.reserve((size_t)WN * 2))) return rc;
	if (codes && (rc = h->scode.reserve((size_t)WN))) return rc;
	const bool realv = values && h->h_real;
	if (values) {
		if ((rc = h->sv1.reserve((size_t)WN * 16))) return rc;
		if ((rc = h->sv2.reserve((size_t)WN * 16))) return rc;
		if (realv) {
			if ((rc = h->sv1r.reserve((size_t)WN * 8))) return rc;
			if ((rc = h->sv2r.reserve((size_t)WN * 8))) return rc;
		}
	}
	if ((rc = h->col16_flag.reserve(4))) return rc;
	TBTKX_CHECK_HIP(hipMemsetAsync(h->col16_flag.ptr, 0, 4, h->stream));
	make_slots<double2, double><<<grid_for(h->N), kThreads, 0, h->stream>>>(
	    h->rowptr.as<uint32_t>(), h->col.as<uint32_t>(), codes ? h->pcode.as<uint8_t>() : nullptr,
	    h->v1.as<double2>(), h->v2.as<double2>(), h->v1r.as<double>(), h->v2r.as<double>(),
	    h->N, W, h->scol.as<int16_t>(), codes ? h->scode.as<uint8_t>() : nullptr,
	    values ? h->sv1.as<double2>() : nullptr, values ? h->sv2.as<double2>() : nullptr,
	    realv ? h->sv1r.as<double>() : nullptr, realv ? h->sv2r.as<double>() : nullptr,
	    h->col16_flag.as<int>());
	TBTKX_CHECK_HIP(hipGetLastError());
	int overflow = 1;
	TBTKX_CHECK_HIP(hipMemcpyAsync(&overflow, h->col16_flag.ptr, 4, hipMemcpyDeviceToHost,
	                               h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	if (overflow) h->slot_ok = 0;
	return TBTKX_OK;
}

int build_slots(tbtkx_context *h, const uint32_t *rowptr) {
	h->slot_ok = 0;
	h->slot_pal = 0;
	h->slot_vals = 0;
	h->slot_W = 0;
	int64_t maxrow = 0;
	for (int64_t r = 0; r < h->N; r++)
		maxrow = std::max<int64_t>(maxrow, rowptr[r + 1] - rowptr[r]);
	if (h->N <= 0 || maxrow > 8) return TBTKX_OK;
	h->slot_W = maxrow < 4 ? 4 : (int)maxrow;
	h->slot_ok = 1;
	const bool pal = h->palette_ok != 0;
	int rc = make_slot_arrays(h, pal, !pal);
	if (rc || !h->slot_ok) return rc;
	h->slot_pal = pal;
	h->slot_vals = !pal;
	return TBTKX_OK;
}

int build_plain_forms(tbtkx_context *h) {
	if (!h->forms_pending) return TBTKX_OK;
	const int64_t N = h->N, Z = h->Z;
	std::vector<uint32_t> hrp(N + 1), hcol(Z > 0 ? Z : 1);
	std::vector<double> hv1(2 * (Z > 0 ? Z : 1)), hv2(2 * (Z > 0 ? Z : 1));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	TBTKX_CHECK_HIP(hipMemcpy(hrp.data(), h->rowptr.ptr, (N + 1) * 4, hipMemcpyDeviceToHost));
	if (Z > 0) {
		TBTKX_CHECK_HIP(hipMemcpy(hcol.data(), h->col.ptr, Z * 4, hipMemcpyDeviceToHost));
		TBTKX_CHECK_HIP(hipMemcpy(hv1.data(), h->v1.ptr, Z * 16, hipMemcpyDeviceToHost));
		TBTKX_CHECK_HIP(hipMemcpy(hv2.data(), h->v2.ptr, Z * 16, hipMemcpyDeviceToHost));
	}
	int rc = build_palette(h, N, Z, hrp.data(), hcol.data(), hv1.data(), hv2.data());
	if (rc) return rc;
	if ((rc = build_slots(h, hrp.data()))) return rc;
	h->forms_pending = 0;
	return TBTKX_OK;
}

int build_slot_values(tbtkx_context *h) {
	if (!h->slot_ok || h->slot_vals) return TBTKX_OK;
	int rc = make_slot_arrays(h, h->slot_pal, true);
	if (rc) return rc;
	h->slot_vals = h->slot_ok;
	return TBTKX_OK;
}

int build_col16(tbtkx_context *h) {
	h->col16_ok = 0;
	if (h->nblk <= 0) return TBTKX_OK;
	int rc;
	if ((rc = h->col16.reserve((size_t)(h->Z + 64) * 2))) return rc;
	if ((rc = h->col16_flag.reserve(4))) return rc;
	TBTKX_CHECK_HIP(hipMemsetAsync(h->col16_flag.ptr, 0, 4, h->stream));
	make_col16<<<(unsigned)h->nblk, kThreads, 0, h->stream>>>(
	    h->rowptr.as<uint32_t>(), h->col.as<uint32_t>(), h->rblk.as<uint32_t>(),
	    h->col16.as<int16_t>(), h->col16_flag.as<int>());
	TBTKX_CHECK_HIP(hipGetLastError());
	int overflow = 1;
	TBTKX_CHECK_HIP(hipMemcpyAsync(&overflow, h->col16_flag.ptr, 4, hipMemcpyDeviceToHost,
	                               h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->col16_ok = overflow == 0;
	return TBTKX_OK;
}

// Row patterns: the slots of a row (column offsets from the row, palette
// codes) as one of at most 255 distinct patterns of the model, so the step
// reads one byte per row instead of W column offsets and W codes. A lattice
// with uniform hoppings and an on-site term has a few dozen (its boundary
// classes); rows whose off-diagonal values vary (bond disorder, a magnetic
// flux) make more and keep the per-slot form.
static int build_row_patterns(tbtkx_context *h, int64_t N, const uint32_t *rowptr,
                              const uint32_t *col, const uint8_t *code) {
	h->pat_ok = 0;
	struct Pat {
		int16_t off[8];
		uint8_t code[8];
		bool operator==(const Pat &o) const { return memcmp(this, &o, sizeof(Pat)) == 0; }
	};
	struct PatHash {
		size_t operator()(const Pat &p) const {
			uint64_t w[3];
			memcpy(w, &p, sizeof(Pat));
			uint64_t x = w[0] * 0x9E3779B97F4A7C15ull;
			x = (x ^ (x >> 29)) * 0xBF58476D1CE4E5B9ull + w[1];
			x = (x ^ (x >> 32)) * 0x94D049BB133111EBull + w[2];
			return (size_t)(x ^ (x >> 31));
		}
	};
	static_assert(sizeof(Pat) == 24, "pattern key layout");
	std::unordered_map<Pat, int, PatHash> index;
	std::vector<Pat> table;
	std::vector<uint8_t> pid((size_t)N);
	for (int64_t r = 0; r < N; r++) {
		const uint32_t k0 = rowptr[r], len = rowptr[r + 1] - k0;
		if (len > 8) return TBTKX_OK;
		Pat p;
		for (int s = 0; s < 8; s++) {
			p.off[s] = kSlotPad;
			p.code[s] = 0;
		}
		for (uint32_t s = 0; s < len; s++) {
			const int64_t d = (int64_t)col[k0 + s] - r;
			if (d < -32767 || d > 32767) return TBTKX_OK;
			p.off[s] = (int16_t)d;
			p.code[s] = code[k0 + s];
		}
		auto it = index.find(p);
		if (it == index.end()) {
			if (table.size() >= 255) return TBTKX_OK;   // too many patterns
			it = index.emplace(p, (int)table.size()).first;
			table.push_back(p);
		}
		pid[r] = (uint8_t)it->second;
	}
	const size_t P = table.size();
	std::vector<int16_t> toff(P * 8);
	std::vector<uint8_t> tcode(P * 8);
	for (size_t i = 0; i < P; i++) {
		memcpy(&toff[i * 8], table[i].off, 16);
		memcpy(&tcode[i * 8], table[i].code, 8);
	}
	// 64-row chunks: the pattern of all the chunk's rows, or 255 (mixed, or
	// a chunk running past row N - 1). Sized for the chunk indices of every
	// rows-per-thread setting (1, 2, 4): ceil(N / 1024) x 16.
	const int64_t nchunk = (N + 1023) / 1024 * 16;
	std::vector<uint8_t> chunk((size_t)nchunk, 255);
	int64_t uniform_rows = 0;
	for (int64_t c = 0; c < nchunk && (c + 1) * 64 <= N; c++) {
		const uint8_t p0 = pid[(size_t)(c * 64)];
		bool same = true;
		for (int64_t r = c * 64 + 1; r < (c + 1) * 64 && same; r++) same = pid[(size_t)r] == p0;
		if (same) {
			chunk[(size_t)c] = p0;
			uniform_rows += 64;
		}
	}
	int rc;
	if ((rc = h->ppid.reserve((size_t)N))) return rc;
	if ((rc = h->pchunk.reserve((size_t)nchunk))) return rc;
	if ((rc = h->ptoff.reserve(P * 16 > 0 ? P * 16 : 16))) return rc;
	if ((rc = h->ptcode.reserve(P * 8 > 0 ? P * 8 : 8))) return rc;
	TBTKX_CHECK_HIP(hipMemcpy(h->ppid.ptr, pid.data(), (size_t)N, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->pchunk.ptr, chunk.data(), (size_t)nchunk,
	                          hipMemcpyHostToDevice));
	if (P) {
		TBTKX_CHECK_HIP(hipMemcpy(h->ptoff.ptr, toff.data(), P * 16, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->ptcode.ptr, tcode.data(), P * 8, hipMemcpyHostToDevice));
	}
	h->pat_ok = 1;
	h->pat_count = (int)P;
	h->pat_byte_rows = N - uniform_rows;
	return TBTKX_OK;
}

int build_palette(tbtkx_context *h, int64_t N, int64_t Z, const uint32_t *rowptr,
                  const uint32_t *col, const double *v1, const double *v2) {
	h->pat_ok = 0;
	h->palette_ok = 0;
	struct Key {
		uint64_t w[4];
		bool operator==(const Key &o) const {
			return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
		}
	};
	struct KeyHash {
		size_t operator()(const Key &k) const {
			uint64_t x = k.w[0] * 0x9E3779B97F4A7C15ull;
			for (int i = 1; i < 4; i++) x = (x ^ (x >> 29)) * 0xBF58476D1CE4E5B9ull + k.w[i];
			return (size_t)(x ^ (x >> 32));
		}
	};
	std::unordered_map<Key, int, KeyHash> index;
	std::vector<uint8_t> code((size_t)Z + kCsrPad, 0);
	// pal: [H/s][256] then [2H/s][256] complex; diag: [H/s][N] then [2H/s][N]
	std::vector<double> pal(2 * 256 * 2, 0.0), diag((size_t)2 * N * 2, 0.0);
	for (int64_t r = 0; r < N; r++) {
		bool have_diag = false;
		for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
			if ((int64_t)col[k] == r) {
				// One diagonal value per row: a row that lists its diagonal
				// twice (a CSR with unmerged entries) keeps full values.
				if (have_diag) return TBTKX_OK;
				have_diag = true;
				code[k] = 255;
				diag[2 * r] = v1[2 * k];
				diag[2 * r + 1] = v1[2 * k + 1];
				diag[2 * (N + r)] = v2[2 * k];
				diag[2 * (N + r) + 1] = v2[2 * k + 1];
				continue;
			}
			Key key;
			memcpy(&key.w[0], &v1[2 * k], 16);
			memcpy(&key.w[2], &v2[2 * k], 16);
			auto it = index.find(key);
			int q;
			if (it != index.end()) {
				q = it->second;
			} else {
				if (index.size() >= 255) return TBTKX_OK;   // too many values
				q = (int)index.size();
				index.emplace(key, q);
				pal[2 * q] = v1[2 * k];
				pal[2 * q + 1] = v1[2 * k + 1];
				pal[2 * (256 + q)] = v2[2 * k];
				pal[2 * (256 + q) + 1] = v2[2 * k + 1];
			}
			code[k] = (uint8_t)q;
		}
	}
	std::vector<double> palr(2 * 256), diagr((size_t)2 * N);
	for (size_t i = 0; i < palr.size(); i++) palr[i] = pal[2 * i];
	for (size_t i = 0; i < diagr.size(); i++) diagr[i] = diag[2 * i];
	int rc;
	if ((rc = h->pcode.reserve(code.size()))) return rc;
	if ((rc = h->pal.reserve(pal.size() * 8))) return rc;
	if ((rc = h->palr.reserve(palr.size() * 8))) return rc;
	if ((rc = h->pdiag.reserve(diag.size() * 8))) return rc;
	if ((rc = h->pdiagr.reserve(diagr.size() * 8))) return rc;
	TBTKX_CHECK_HIP(hipMemcpy(h->pcode.ptr, code.data(), code.size(), hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->pal.ptr, pal.data(), pal.size() * 8, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->palr.ptr, palr.data(), palr.size() * 8, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->pdiag.ptr, diag.data(), diag.size() * 8, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->pdiagr.ptr, diagr.data(), diagr.size() * 8,
	                          hipMemcpyHostToDevice));
	h->palette_ok = 1;
	return build_row_patterns(h, N, rowptr, col, code.data());
}

// Scratch of tbtkx_moments_noalloc: the two recurrence vectors, the moment
// rows and the to-list of one call, in a fixed device array of
// TBTKX_NOALLOC_BYTES per GPU (allocated with the code object when the
// library is loaded on that GPU, held until the process ends). Host code
// reaches it only through the address hipGetSymbolAddress gives on each
// device (noalloc_scratch); the array's name is not a device address there.
static __device__ __align__(256) unsigned char scratch[TBTKX_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices];
// Held for a whole tbtkx_moments_noalloc call: calls on one GPU share its
// scratch, so they run one after another.
std::mutex g_scratch_busy[kMaxDevices];
}  // namespace

// The device address of `scratch` on the handle's device, looked up once per
// device (with that device current) and kept.
static int noalloc_scratch(tbtkx_context *h, unsigned char **ptr) {
	TBTKX_REQUIRE(h->device >= 0 && h->device < kMaxDevices, TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: device id beyond the scratch table");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_ptr[h->device]) {
		int rc = set_device(h);
		if (rc) return rc;
		void *p = nullptr;
		TBTKX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		TBTKX_REQUIRE(p != nullptr, TBTKX_ERR_HIP,
		              "tbtkx_moments_noalloc: no device address for the scratch array");
		g_scratch_ptr[h->device] = p;
	}
	*ptr = static_cast<unsigned char *>(g_scratch_ptr[h->device]);
	return TBTKX_OK;
}

}  // namespace tbtkx

using namespace tbtkx;

extern "C" int tbtkx_moments_noalloc(tbtkx_handle h, int64_t from, int K,
                                     const int64_t *to, int M, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_moments_noalloc: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_moments_noalloc: no model set");
	TBTKX_REQUIRE(M > 0, TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: numCoefficients has to be larger than 0");
	TBTKX_REQUIRE(K >= 0 && (K == 0 || (to && out)), TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: bad to-index list");
	TBTKX_REQUIRE(from >= 0 && from < h->N, TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: from-index outside the basis");
	TBTKX_REQUIRE(K <= h->N && K <= TBTKX_NOALLOC_MAX_TO, TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: at most min(N, " +
	                  std::to_string(TBTKX_NOALLOC_MAX_TO) + ") to-indices");
	for (int k = 0; k < K; k++)
		TBTKX_REQUIRE(to[k] >= 0 && to[k] < h->N, TBTKX_ERR_ARG,
		              "tbtkx_moments_noalloc: to-index outside the basis");
	if (K == 0) return TBTKX_OK;
	const bool real = h->real_path && h->h_real;
	const size_t elem = real ? 8 : 16;
	// Layout: vector a, vector b, moments (K x M), to-list; each 256-aligned.
	auto up = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t vbytes = up((size_t)h->N * elem), cbytes = up((size_t)K * M * 16);
	const size_t need = 2 * vbytes + cbytes + up((size_t)K * 8);
	TBTKX_REQUIRE(need <= (size_t)TBTKX_NOALLOC_BYTES, TBTKX_ERR_ARG,
	              "tbtkx_moments_noalloc: the call needs " + std::to_string(need) +
	                  " bytes of scratch (2 vectors of N, K x M moments, K to-indices), more "
	                  "than the " + std::to_string((size_t)TBTKX_NOALLOC_BYTES) +
	                  " available; use tbtkx_moments");
	unsigned char *base = nullptr;
	int rc = noalloc_scratch(h, &base);
	if (rc) return rc;
	std::lock_guard<std::mutex> busy(g_scratch_busy[h->device]);
	if ((rc = set_device(h))) return rc;
	void *va = base, *vb = base + vbytes;
	double2 *coef = reinterpret_cast<double2 *>(base + 2 * vbytes);
	int64_t *tod = reinterpret_cast<int64_t *>(base + 2 * vbytes + cbytes);
	TBTKX_CHECK_HIP(hipMemcpyAsync(tod, to, (size_t)K * 8, hipMemcpyHostToDevice, h->stream));
	// Every scratch byte the call reads is written first: the start vector
	// by init_unit, the other vector by the first step (which does not read
	// it), every moment by a gather.
	const PlainForm pform = plain_form(h);
	const bool c16 = pform.c16, usepal = pform.pal;
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	if (real) {
		init_unit<double><<<grid_for(h->N), kThreads, 0, h->stream>>>(
		    static_cast<double *>(va), h->N, from);
		rc = enqueue_steps<double>(h, K, M, static_cast<double *>(va), static_cast<double *>(vb),
		                           coef, tod);
	} else {
		init_unit<double2><<<grid_for(h->N), kThreads, 0, h->stream>>>(
		    static_cast<double2 *>(va), h->N, from);
		rc = enqueue_steps<double2>(h, K, M, static_cast<double2 *>(va),
		                            static_cast<double2 *>(vb), coef, tod);
	}
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	TBTKX_CHECK_HIP(hipMemcpyAsync(out, coef, (size_t)K * M * 16, hipMemcpyDeviceToHost,
	                               h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_launches = M > 1 ? M - 1 : 0;
	h->last_real = real;
	h->last_col16 = c16;
	h->last_palette = usepal;
	h->last_plain_kernel = pform.slot ? 2 : h->plain_kernel == 1 ? 1 : 0;
	h->last_patterns = pform.pat;
	// A repeated to-index keeps only its last position (as tbtkx_moments).
	for (int k = 0; k < K; k++)
		for (int l = k + 1; l < K; l++)
			if (to[l] == to[k]) {
				memset(out + (size_t)k * M * 2, 0, (size_t)M * 16);
				break;
			}
	return TBTKX_OK;
}

extern "C" int tbtkx_moments(tbtkx_handle h, int64_t from, int K,
                             const int64_t *to, int M, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_moments: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_moments: no model set");
	TBTKX_REQUIRE(M > 0, TBTKX_ERR_ARG,
	              "tbtkx_moments: numCoefficients has to be larger than 0");
	TBTKX_REQUIRE(K >= 0 && (K == 0 || (to && out)), TBTKX_ERR_ARG,
	              "tbtkx_moments: bad to-index list");
	TBTKX_REQUIRE(from >= 0 && from < h->N, TBTKX_ERR_ARG,
	              "tbtkx_moments: from-index outside the basis");
	if (K == 0) return TBTKX_OK;

	// Unique to-indices; a repeated index keeps only its last position, as
	// the reference's coefficientMap (ChebyshevExpander.cpp:121-126).
	std::vector<int64_t> uniq;
	std::vector<int> slot(K, -1);
	{
		std::map<int64_t, int> last;
		for (int k = 0; k < K; k++) {
			TBTKX_REQUIRE(to[k] >= 0 && to[k] < h->N, TBTKX_ERR_ARG,
			              "tbtkx_moments: to-index outside the basis");
			last[to[k]] = k;
		}
		for (auto &kv : last) {
			slot[kv.second] = (int)uniq.size();
			uniq.push_back(kv.first);
		}
	}
	const int U = (int)uniq.size();
	int rc = set_device(h);
	if (rc) return rc;
	// A device-built model gets its palette, patterns and slots now.
	if ((rc = build_plain_forms(h))) return rc;
	// The slot form without the palette needs the values in slots.
	if (h->plain_kernel == 2 && !(h->use_palette && h->slot_pal) &&
	    (rc = build_slot_values(h)))
		return rc;
	if ((rc = h->to.reserve((size_t)U * 8))) return rc;
	if ((rc = h->coef.reserve((size_t)U * M * 16))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->to.ptr, uniq.data(), (size_t)U * 8,
	                               hipMemcpyHostToDevice, h->stream));
	// Real arithmetic when every value of H is real (the start vector is a
	// unit vector).
	const bool real = h->real_path && h->h_real;
	const PlainForm pform = plain_form(h);
	const bool c16 = pform.c16, usepal = pform.pal;
	if (real)
		init_unit<double><<<grid_for(h->N), kThreads, 0, h->stream>>>(
		    h->vec_a.as<double>(), h->N, from);
	else
		init_unit<double2><<<grid_for(h->N), kThreads, 0, h->stream>>>(
		    h->vec_a.as<double2>(), h->N, from);
	TBTKX_CHECK_HIP(hipGetLastError());
	auto enqueue = [&]() {
		double2 *coef = h->coef.as<double2>();
		const int64_t *tod = h->to.as<int64_t>();
		return real ? enqueue_steps<double>(h, U, M, h->vec_a.as<double>(),
		                                    h->vec_b.as<double>(), coef, tod)
		            : enqueue_steps<double2>(h, U, M, h->vec_a.as<double2>(),
		                                     h->vec_b.as<double2>(), coef, tod);
	};

	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	if (h->use_graph) {
		GraphCache &gc = h->plain_graph;
		int64_t key[GraphCache::kKey];
		const bool key_ok = GraphCache::make_key(
		    key, {U, M, h->plain_kernel * 4 + pform.slot * 2 + pform.pat + 16 * h->slot_rows +
		                    256 * h->use_patterns,
		          h->xcd_remap, h->N,
		          real, c16, usepal},
		    {h->vec_a.ptr, h->vec_b.ptr, h->coef.ptr, h->to.ptr, h->rowptr.ptr, h->col.ptr,
		     h->v1.ptr, h->v2.ptr, h->v1r.ptr, h->v2r.ptr, h->rblk.ptr, h->col16.ptr,
		     h->pcode.ptr, h->pal.ptr, h->palr.ptr, h->pdiag.ptr, h->pdiagr.ptr, h->scol.ptr,
		     h->scode.ptr, h->sv1.ptr, h->sv2.ptr, h->sv1r.ptr, h->sv2r.ptr, h->ppid.ptr,
		     h->pchunk.ptr,
		     h->ptoff.ptr, h->ptcode.ptr});
		TBTKX_REQUIRE(key_ok, TBTKX_ERR_STATE, "tbtkx_moments: graph key too long");
		if (!gc.matches(key)) {
			gc.reset();
			TBTKX_CHECK_HIP(hipStreamBeginCapture(h->stream,
			                                      hipStreamCaptureModeThreadLocal));
			int erc = enqueue();
			hipError_t ec = hipStreamEndCapture(h->stream, &gc.graph);
			if (erc) return erc;
			TBTKX_CHECK_HIP(ec);
			TBTKX_CHECK_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr,
			                                    nullptr, 0));
			memcpy(gc.key, key, sizeof key);
		}
		TBTKX_CHECK_HIP(hipGraphLaunch(gc.exec, h->stream));
	} else {
		if ((rc = enqueue())) return rc;
	}
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));

	std::vector<double> dense((size_t)U * M * 2);
	TBTKX_CHECK_HIP(hipMemcpyAsync(dense.data(), h->coef.ptr, dense.size() * 8,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_launches = M > 1 ? M - 1 : 0;
	h->last_real = real;
	h->last_col16 = c16;
	h->last_palette = usepal;
	h->last_plain_kernel = pform.slot ? 2 : h->plain_kernel == 1 ? 1 : 0;
	h->last_patterns = pform.pat;

	for (int k = 0; k < K; k++) {
		double *o = out + (size_t)k * M * 2;
		if (slot[k] < 0)
			memset(o, 0, (size_t)M * 16);
		else
			memcpy(o, dense.data() + (size_t)slot[k] * M * 2, (size_t)M * 16);
	}
	return TBTKX_OK;
}
