// Green's-function contraction on MFMA: G[k][e] = sum_n op_type(L[n][e], c[k][n])
// as a complex GEMM (K x M) x (M x E) in split-BF16 with FP32 accumulation.
//
// Each FP32 value x is split as x = hi + lo, hi = bf16(x), lo = bf16(x - hi)
// (about 16 significant bits), and a product is taken as three BF16 terms,
//   a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi,
// so a complex product (ar + i ai)(br + i bi) costs 12 real BF16 products
// (6 for the Principal / NonPrincipal types, whose table operand is purely
// real / imaginary). The Green's function type acts on the table operand
// L' (Lib/src/Solver/ChebyshevExpander.cpp:324-370):
//   Retarded L' = L, Advanced L' = conj(L), Principal L' = -Re L,
//   NonPrincipal L' = -i Im L.
//
// Data: coefficient planes Cp[4][K][Mp] and table planes Lp[4][E][Mp]
// (plane = re_hi, re_lo, im_hi, im_lo; the table is stored transposed so
// the reduction index is contiguous for both operands; Mp = M rounded up to
// 32, zero padded). Output G[K][E] complex FP32.
//
// Kernels. contract_mfma_pipe8 (the default): double-buffered direct
// global->LDS staging, 8 wavefronts per workgroup (2 per SIMD); 128 x 128
// tiles for Retarded/Advanced, 128 x 256 tiles over the two used table
// planes for Principal/NonPrincipal. contract_mfma (TBTKX_OPT_MFMA_PIPE = 0):
// single-stage, 4 wavefronts, 128 x 128 tiles, operands staged through
// VGPRs; it adds the same products in the same order per accumulator, so
// both give identical outputs. Both use v_mfma_f32_32x32x16_bf16 with FP32
// accumulators for the real and imaginary parts, the reduction in chunks of
// 32 staged in LDS.
// Executed MFMA flops per launch: 2 * K * E * M * 12 (Retarded/Advanced)
// or * 6 (Principal/NonPrincipal).
#include <cstring>

#include "common.hpp"

namespace tbtkx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTileK = 128;   // output rows (coefficient vectors) per workgroup
constexpr int kTileE = 128;   // output energies per workgroup
constexpr int kChunk = 32;    // reduction chunk staged in LDS
constexpr int kRowPad = 40;   // LDS row stride in bf16 (32 + 8 against bank conflicts)

__device__ __forceinline__ uint16_t bf16_bits(float x) {
	// round to nearest even
	uint32_t u = __float_as_uint(x);
	u += 0x7fff + ((u >> 16) & 1);
	return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t b) {
	return __uint_as_float((uint32_t)b << 16);
}

// planes[p][row][m] (Mp columns) from complex double src[row][m] (M columns),
// zero past M; with `transpose` src is read as src[m][row] (the table).
__global__ __launch_bounds__(kThreads) void split_planes(
    const double2 *__restrict__ src, uint16_t *__restrict__ planes, int64_t rows,
    int M, int Mp, int transpose) {
	const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (i >= rows * Mp) return;
	const int64_t row = i / Mp;
	const int m = (int)(i - row * Mp);
	double2 v = make_double2(0.0, 0.0);
	if (m < M) v = transpose ? src[(int64_t)m * rows + row] : src[row * M + m];
	const float re = (float)v.x, im = (float)v.y;
	const uint16_t rh = bf16_bits(re), ih = bf16_bits(im);
	const uint16_t rl = bf16_bits(re - bf16_value(rh)), il = bf16_bits(im - bf16_value(ih));
	const int64_t plane = rows * Mp;
	planes[0 * plane + i] = rh;
	planes[1 * plane + i] = rl;
	planes[2 * plane + i] = ih;
	planes[3 * plane + i] = il;
}

__device__ __forceinline__ bf16x8 neg(bf16x8 a) {
	typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
	u32x4 u = __builtin_bit_cast(u32x4, a);
	u ^= (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
	return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c) {
	return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

template <int TYPE>
__global__ __launch_bounds__(kThreads) void contract_mfma(
    const uint16_t *__restrict__ Cp, const uint16_t *__restrict__ Lp,
    float2 *__restrict__ G, int K, int E, int Mp, int ntile_e) {
	__shared__ uint16_t sA[4][kTileK][kRowPad];
	__shared__ uint16_t sB[4][kTileE][kRowPad];
	const int tid = threadIdx.x;
	const int wave = tid >> 6, lane = tid & 63;
	const int tk = blockIdx.x / ntile_e, te = blockIdx.x - tk * ntile_e;
	const int k0 = tk * kTileK, e0 = te * kTileE;
	const int wr = wave >> 1, wc = wave & 1;   // 2 x 2 wavefronts
	const int64_t planeC = (int64_t)K * Mp, planeL = (int64_t)E * Mp;

	f32x16 accR[2][2], accI[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++) {
			accR[a][b] = (f32x16){};
			accI[a][b] = (f32x16){};
		}

	for (int m0 = 0; m0 < Mp; m0 += kChunk) {
		// Stage 4 planes x 128 rows x 32 bf16 of each operand: 64 B per
		// row and plane, as 4 loads of 16 B; rows past K or E are zero.
#pragma unroll
		for (int it = 0; it < 8; it++) {
			const int q = tid + it * kThreads;      // 0 .. 2047
			const int p = q >> 9, r = (q >> 2) & 127, c4 = q & 3;
			uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
			if (k0 + r < K)
				va = *reinterpret_cast<const uint4 *>(Cp + p * planeC + (int64_t)(k0 + r) * Mp + m0 + c4 * 8);
			if (e0 + r < E)
				vb = *reinterpret_cast<const uint4 *>(Lp + p * planeL + (int64_t)(e0 + r) * Mp + m0 + c4 * 8);
			*reinterpret_cast<uint4 *>(&sA[p][r][c4 * 8]) = va;
			*reinterpret_cast<uint4 *>(&sB[p][r][c4 * 8]) = vb;
		}
		__syncthreads();
#pragma unroll
		for (int ks = 0; ks < kChunk; ks += 16) {
			const int kk = ks + (lane >> 5) * 8;
			bf16x8 A[4][2], Bm[4][2];
#pragma unroll
			for (int p = 0; p < 4; p++)
#pragma unroll
				for (int t = 0; t < 2; t++) {
					A[p][t] = *reinterpret_cast<const bf16x8 *>(
					    &sA[p][wr * 64 + t * 32 + (lane & 31)][kk]);
					Bm[p][t] = *reinterpret_cast<const bf16x8 *>(
					    &sB[p][wc * 64 + t * 32 + (lane & 31)][kk]);
				}
#pragma unroll
			for (int a = 0; a < 2; a++)
#pragma unroll
				for (int b = 0; b < 2; b++) {
					const bf16x8 arh = A[0][a], arl = A[1][a], aih = A[2][a], ail = A[3][a];
					const bf16x8 brh = Bm[0][b], brl = Bm[1][b], bih = Bm[2][b], bil = Bm[3][b];
					if (TYPE == TBTKX_RETARDED || TYPE == TBTKX_ADVANCED) {
						// L' = (br, +-bi); re += ar br - ai bi', im += ar bi' + ai br
						const bf16x8 sih = TYPE == TBTKX_RETARDED ? bih : neg(bih);
						const bf16x8 sil = TYPE == TBTKX_RETARDED ? bil : neg(bil);
						const bf16x8 naih = neg(aih), nail = neg(ail);
						f32x16 r = accR[a][b], i = accI[a][b];
						r = mma(arh, brh, r); r = mma(arh, brl, r); r = mma(arl, brh, r);
						r = mma(naih, sih, r); r = mma(naih, sil, r); r = mma(nail, sih, r);
						i = mma(arh, sih, i); i = mma(arh, sil, i); i = mma(arl, sih, i);
						i = mma(aih, brh, i); i = mma(aih, brl, i); i = mma(ail, brh, i);
						accR[a][b] = r;
						accI[a][b] = i;
					} else if (TYPE == TBTKX_PRINCIPAL) {
						// L' = -Re L: re += ar (-br), im += ai (-br)
						const bf16x8 nbh = neg(brh), nbl = neg(brl);
						f32x16 r = accR[a][b], i = accI[a][b];
						r = mma(arh, nbh, r); r = mma(arh, nbl, r); r = mma(arl, nbh, r);
						i = mma(aih, nbh, i); i = mma(aih, nbl, i); i = mma(ail, nbh, i);
						accR[a][b] = r;
						accI[a][b] = i;
					} else {
						// L' = -i Im L = (0, -bi): re += ai bi, im += -ar bi
						const bf16x8 nbh = neg(bih), nbl = neg(bil);
						f32x16 r = accR[a][b], i = accI[a][b];
						r = mma(aih, bih, r); r = mma(aih, bil, r); r = mma(ail, bih, r);
						i = mma(arh, nbh, i); i = mma(arh, nbl, i); i = mma(arl, nbh, i);
						accR[a][b] = r;
						accI[a][b] = i;
					}
				}
		}
		__syncthreads();
	}
	// D layout of v_mfma_f32_32x32x16 (checked on the device): element r of
	// a lane is row 8*(r/4) + 4*(lane/32) + r%4, column lane%32.
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
#pragma unroll
			for (int r = 0; r < 16; r++) {
				const int row = k0 + wr * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
				const int col = e0 + wc * 64 + b * 32 + (lane & 31);
				if (row < K && col < E)
					G[(int64_t)row * E + col] = make_float2(accR[a][b][r], accI[a][b][r]);
			}
}

// Staging of the pipelined kernel. The operand tiles of chunk c+1 stream
// into the second of two LDS stages by direct global->LDS loads
// (global_load_lds_dwordx4: lane i of a wavefront writes 16 B at the
// wave-uniform LDS address + 16 i) while the MFMAs of chunk c run, so no
// VGPRs hold staged data and the load latency hides behind a whole chunk of
// MFMAs. LDS rows are 64 B (32 bf16, no padding); the four 16 B quads of
// row r are stored XOR-swizzled by (r >> 2) & 3, which makes the 16 lanes of
// each ds_read_b128 phase hit 16 distinct bank groups. 2 stages x 64 KB =
// 128 KB of LDS.
constexpr int kRowB = 32;

__device__ __forceinline__ int swz(int row, int quad) { return quad ^ ((row >> 2) & 3); }

typedef __attribute__((address_space(3))) void lds_void;

// Pipelined kernel with two wavefronts per SIMD. A workgroup of 8
// wavefronts (2 x 4) shares each double-buffered 64 KB stage, so one
// wavefront's fragment reads overlap the other's MFMAs on the same SIMD.
//   Retarded / Advanced: 128 x 128 tile, all four planes of both operands
//   staged; each wavefront computes 64 x 32 (2 x 1 MFMA tiles for the real
//   and imaginary parts).
//   Principal / NonPrincipal: the table operand is one real plane pair
//   (Re L or Im L), so a stage holds the 4 coefficient planes of 128 rows and
//   the 2 used table planes of 256 rows; 128 x 256 tile, each wavefront
//   64 x 64 (2 x 2 MFMA tiles).
// Either way a k-step reads 12 fragments per 24 MFMAs per wavefront. Each
// accumulator receives the same products in the same order as in the
// single-stage kernel.
constexpr int kThreads8 = 512;

template <int TYPE>
struct PipeShape {
	static constexpr bool kHalf = TYPE == TBTKX_PRINCIPAL || TYPE == TBTKX_NONPRINCIPAL;
	static constexpr int kTE = kHalf ? 256 : 128;   // energies per workgroup
	static constexpr int kNBP = kHalf ? 2 : 4;      // table planes staged
	static constexpr int kNB = kHalf ? 2 : 1;       // MFMA tiles along E per wavefront
};

template <int TYPE>
__global__ __launch_bounds__(kThreads8) void contract_mfma_pipe8(
    const uint16_t *__restrict__ Cp, const uint16_t *__restrict__ Lp,
    float2 *__restrict__ G, int K, int E, int Mp, int ntile_e) {
	typedef PipeShape<TYPE> S;
	constexpr int kNB = S::kNB, kNBP = S::kNBP, kTE = S::kTE;
	constexpr int kPlaneA = kTileK * kRowB, kPlaneB = kTE * kRowB;
	constexpr int kGroupsB = kTE / 16;   // 16-row load groups per table plane
	static_assert(4 * kPlaneA == kNBP * kPlaneB, "stage halves differ");
	__shared__ __attribute__((aligned(16))) uint16_t sm[2][2][4 * kPlaneA];
	const int tid = threadIdx.x;
	const int wave = tid >> 6, lane = tid & 63;
	const int tk = blockIdx.x / ntile_e, te = blockIdx.x - tk * ntile_e;
	const int k0 = tk * kTileK, e0 = te * kTE;
	const int wr = wave >> 2, wc = wave & 3;   // 2 x 4 wavefronts
	const int64_t planeC = (int64_t)K * Mp, planeL = (int64_t)E * Mp;
	const int nc = Mp / kChunk;
	// first table plane staged: Re L (Principal) or Im L (NonPrincipal)
	const int lp0 = TYPE == TBTKX_NONPRINCIPAL ? 2 : 0;

	// 64 loads of 1 KB per stage: j < 32 are (coefficient plane j / 8, rows
	// 16 (j % 8) ..), j >= 32 are (staged table plane (j - 32) / kGroupsB,
	// rows 16 ((j - 32) % kGroupsB) ..). Wavefront w issues j = w, w + 8,
	// ..., w + 56; lane l fetches row (group start) + l / 4, the quad that
	// belongs in LDS slot l % 4. Rows past K or E re-read the last valid row
	// (their results are never stored).
	const int lrow = lane >> 2;
	auto issue = [&](int c, int buf) {
		const int m0 = c * kChunk;
#pragma unroll
		for (int i = 0; i < 8; i++) {
			const int j = wave + 8 * i;
			const uint16_t *src;
			int off;
			if (i < 4) {
				const int p = j >> 3, rb = 16 * (j & 7);
				const int r = rb + lrow;
				const int row = k0 + r < K ? k0 + r : K - 1;
				src = Cp + p * planeC + (int64_t)row * Mp + m0 + swz(r, lane & 3) * 8;
				off = p * kPlaneA + rb * kRowB;
			} else {
				const int jj = j - 32;
				const int p = jj / kGroupsB, rb = 16 * (jj % kGroupsB);
				const int r = rb + lrow;
				const int row = e0 + r < E ? e0 + r : E - 1;
				src = Lp + (lp0 + p) * planeL + (int64_t)row * Mp + m0 + swz(r, lane & 3) * 8;
				off = p * kPlaneB + rb * kRowB;
			}
			__builtin_amdgcn_global_load_lds((const void *)src,
			                                 (lds_void *)&sm[buf][i >= 4][off], 16, 0, 0);
		}
	};

	f32x16 accR[2][kNB], accI[2][kNB];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < kNB; b++) {
			accR[a][b] = (f32x16){};
			accI[a][b] = (f32x16){};
		}

	issue(0, 0);
	for (int c = 0; c < nc; c++) {
		// This wavefront's loads of chunk c are done; the barrier makes every
		// wavefront's visible and tells that all finished reading the stage
		// the next loads overwrite (chunk c - 1's).
		__builtin_amdgcn_s_waitcnt(0);
		__syncthreads();
		if (c + 1 < nc) issue(c + 1, (c + 1) & 1);
		const int buf = c & 1;
#pragma unroll
		for (int ks = 0; ks < kChunk; ks += 16) {
			const int kq = ks / 8 + (lane >> 5);
			bf16x8 A[4][2], Bm[kNBP][kNB];
#pragma unroll
			for (int p = 0; p < 4; p++)
#pragma unroll
				for (int t = 0; t < 2; t++) {
					const int ra = wr * 64 + t * 32 + (lane & 31);
					A[p][t] = *reinterpret_cast<const bf16x8 *>(
					    &sm[buf][0][p * kPlaneA + ra * kRowB + swz(ra, kq) * 8]);
				}
#pragma unroll
			for (int p = 0; p < kNBP; p++)
#pragma unroll
				for (int t = 0; t < kNB; t++) {
					const int rb = wc * (32 * kNB) + t * 32 + (lane & 31);
					Bm[p][t] = *reinterpret_cast<const bf16x8 *>(
					    &sm[buf][1][p * kPlaneB + rb * kRowB + swz(rb, kq) * 8]);
				}
#define TBTKX_MM(ACC, X, Y)                                                     \
	_Pragma("unroll") for (int a = 0; a < 2; a++)                          \
	    _Pragma("unroll") for (int b = 0; b < kNB; b++) ACC[a][b] = mma(X, Y, ACC[a][b]);
			if (TYPE == TBTKX_RETARDED || TYPE == TBTKX_ADVANCED) {
				// L' = (br, +-bi): re += ar br - ai bi', im += ar bi' + ai br
				bf16x8 sih[kNB], sil[kNB], naih[2], nail[2];
#pragma unroll
				for (int t = 0; t < kNB; t++) {
					sih[t] = TYPE == TBTKX_RETARDED ? Bm[2 % kNBP][t] : neg(Bm[2 % kNBP][t]);
					sil[t] = TYPE == TBTKX_RETARDED ? Bm[3 % kNBP][t] : neg(Bm[3 % kNBP][t]);
				}
#pragma unroll
				for (int t = 0; t < 2; t++) {
					naih[t] = neg(A[2][t]);
					nail[t] = neg(A[3][t]);
				}
				TBTKX_MM(accR, A[0][a], Bm[0][b])
				TBTKX_MM(accI, A[0][a], sih[b])
				TBTKX_MM(accR, A[0][a], Bm[1][b])
				TBTKX_MM(accI, A[0][a], sil[b])
				TBTKX_MM(accR, A[1][a], Bm[0][b])
				TBTKX_MM(accI, A[1][a], sih[b])
				TBTKX_MM(accR, naih[a], sih[b])
				TBTKX_MM(accI, A[2][a], Bm[0][b])
				TBTKX_MM(accR, naih[a], sil[b])
				TBTKX_MM(accI, A[2][a], Bm[1][b])
				TBTKX_MM(accR, nail[a], sih[b])
				TBTKX_MM(accI, A[3][a], Bm[0][b])
			} else {
				bf16x8 nB[2][kNB];
#pragma unroll
				for (int p = 0; p < 2; p++)
#pragma unroll
					for (int t = 0; t < kNB; t++) nB[p][t] = neg(Bm[p][t]);
				if (TYPE == TBTKX_PRINCIPAL) {
					// L' = -Re L: re += ar (-br), im += ai (-br)
					TBTKX_MM(accR, A[0][a], nB[0][b])
					TBTKX_MM(accI, A[2][a], nB[0][b])
					TBTKX_MM(accR, A[0][a], nB[1][b])
					TBTKX_MM(accI, A[2][a], nB[1][b])
					TBTKX_MM(accR, A[1][a], nB[0][b])
					TBTKX_MM(accI, A[3][a], nB[0][b])
				} else {
					// L' = -i Im L = (0, -bi): re += ai bi, im += -ar bi
					TBTKX_MM(accR, A[2][a], Bm[0][b])
					TBTKX_MM(accI, A[0][a], nB[0][b])
					TBTKX_MM(accR, A[2][a], Bm[1][b])
					TBTKX_MM(accI, A[0][a], nB[1][b])
					TBTKX_MM(accR, A[3][a], Bm[0][b])
					TBTKX_MM(accI, A[1][a], nB[0][b])
				}
			}
#undef TBTKX_MM
		}
	}
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < kNB; b++)
#pragma unroll
			for (int r = 0; r < 16; r++) {
				const int row = k0 + wr * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
				const int col = e0 + wc * (32 * kNB) + b * 32 + (lane & 31);
				if (row < K && col < E)
					G[(int64_t)row * E + col] = make_float2(accR[a][b][r], accI[a][b][r]);
			}
}

}  // namespace
}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

int tbtkx_coeffs_upload(tbtkx_handle h, int K, int M, const double *coeffs) {
	TBTKX_REQUIRE(h && coeffs, TBTKX_ERR_ARG, "tbtkx_coeffs_upload: null pointer");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(K > 0 && M > 0, TBTKX_ERR_ARG, "tbtkx_coeffs_upload: K and M must be positive");
	int rc = set_device(h);
	if (rc) return rc;
	const int Mp = (M + kChunk - 1) / kChunk * kChunk;
	if ((rc = h->gf_in.reserve((size_t)K * M * 16))) return rc;
	if ((rc = h->cplanes.reserve((size_t)4 * K * Mp * 2))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->gf_in.ptr, coeffs, (size_t)K * M * 16,
	                               hipMemcpyHostToDevice, h->stream));
	if ((rc = apply_damping(h, h->gf_in.as<double2>(), K, M))) return rc;
	const int64_t n = (int64_t)K * Mp;
	split_planes<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, h->stream>>>(
	    h->gf_in.as<double2>(), h->cplanes.as<uint16_t>(), K, M, Mp, 0);
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->cplanes_K = K;
	h->cplanes_M = M;
	return TBTKX_OK;
}

int tbtkx_contract_mfma(tbtkx_handle h, int type, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_contract_mfma: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->lut_M > 0, TBTKX_ERR_STATE, "tbtkx_contract_mfma: no lookup table loaded");
	TBTKX_REQUIRE(h->cplanes_K > 0, TBTKX_ERR_STATE,
	              "tbtkx_contract_mfma: no coefficients uploaded (tbtkx_coeffs_upload)");
	TBTKX_REQUIRE(h->cplanes_M == h->lut_M, TBTKX_ERR_ARG,
	              "tbtkx_contract_mfma: coefficient count differs from the table's");
	TBTKX_REQUIRE(type >= TBTKX_ADVANCED && type <= TBTKX_NONPRINCIPAL, TBTKX_ERR_ARG,
	              "tbtkx_contract_mfma: unknown Green's function type");
	const int K = h->cplanes_K, M = h->lut_M, E = h->lut_E;
	const int Mp = (M + kChunk - 1) / kChunk * kChunk;
	int rc = set_device(h);
	if (rc) return rc;
	if (h->lplanes_valid != 1) {
		if ((rc = h->lplanes.reserve((size_t)4 * E * Mp * 2))) return rc;
		const int64_t n = (int64_t)E * Mp;
		split_planes<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, h->stream>>>(
		    h->lut.as<double2>(), h->lplanes.as<uint16_t>(), E, M, Mp, 1);
		TBTKX_CHECK_HIP(hipGetLastError());
		h->lplanes_valid = 1;
	}
	if ((rc = h->gf_mfma_out.reserve((size_t)K * E * 8))) return rc;
	const int ntk = (K + kTileK - 1) / kTileK, nte = (E + kTileE - 1) / kTileE;
	const unsigned grid = (unsigned)ntk * (unsigned)nte;
	const uint16_t *Cp = h->cplanes.as<uint16_t>(), *Lp = h->lplanes.as<uint16_t>();
	float2 *G = h->gf_mfma_out.as<float2>();
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
#define TBTKX_CM(T)                                                              \
	do {                                                                     \
		if (h->mfma_pipe) {                                              \
			const int n8 = (E + PipeShape<T>::kTE - 1) / PipeShape<T>::kTE; \
			contract_mfma_pipe8<T><<<(unsigned)ntk * (unsigned)n8, kThreads8, 0, h->stream>>>( \
			    Cp, Lp, G, K, E, Mp, n8);                            \
		} else {                                                         \
			contract_mfma<T><<<grid, kThreads, 0, h->stream>>>(Cp, Lp, G, K, E, Mp, nte); \
		}                                                                \
	} while (0)
	switch (type) {
	case TBTKX_RETARDED: TBTKX_CM(TBTKX_RETARDED); break;
	case TBTKX_ADVANCED: TBTKX_CM(TBTKX_ADVANCED); break;
	case TBTKX_PRINCIPAL: TBTKX_CM(TBTKX_PRINCIPAL); break;
	default: TBTKX_CM(TBTKX_NONPRINCIPAL); break;
	}
#undef TBTKX_CM
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	TBTKX_CHECK_HIP(hipEventSynchronize(h->ev_stop));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_launches = 1;
	if (out) {
		std::vector<float2> g((size_t)K * E);
		TBTKX_CHECK_HIP(hipMemcpy(g.data(), G, g.size() * 8, hipMemcpyDeviceToHost));
		for (size_t i = 0; i < g.size(); i++) {
			out[2 * i] = g[i].x;
			out[2 * i + 1] = g[i].y;
		}
	}
	return TBTKX_OK;
}

}  // extern "C"
