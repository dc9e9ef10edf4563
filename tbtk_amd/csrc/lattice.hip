// Lattice models built on the device: tbtkx_set_lattice (one model, every
// layout the recurrence kernels read) and tbtkx_multi_set_lattices (the
// uneven batch's row-major ELL).
//
// The host path builds H through HoppingAmplitudeSet -> getSparseMatrix ->
// SparseMatrix LIL -> CSR (Lib/include/Core/TBTK/HoppingAmplitudeSet.h:397-424,
// Lib/include/Utilities/TBTK/SparseMatrix.h:1309-1530): rows = basis(to),
// columns = basis(from) ascending in a row, entries of one (row, column)
// summed. For a lattice the same CSR follows from the unit cell: every row
// is one thread, which lists the row's entries (the on-site amplitude, then
// every term whose to-orbital is the row's orbital, at the column of the
// partner cell), keeps them sorted by column and sums equal columns in that
// order. Basis numbering: depth-first over the subindex values of
// {x, y, z, orbital}, i.e. row = ((x Ly + y) Lz + z) O + o.
//
// Derived copies are formed exactly as tbtkx_set_csr forms them on the host
// (scaling by 1/s and 2 as complex products without contraction, FP32 by
// rounding, ELL slots diagonal first), so both give identical device data.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace tbtkx {
namespace {

constexpr int kMaxRow = 16;   // distinct columns a row may hold

struct LatDev {
	int64_t L[3];
	int per[3];
	int O, nt, has_onsite;
};

// The sorted, merged entries of row r. Returns the number of entries.
__device__ int lattice_row(const LatDev &g, const tbtkx_lattice_term *__restrict__ terms,
                           const double2 *__restrict__ onsite, int64_t r, uint32_t (&cols)[kMaxRow],
                           double2 (&vals)[kMaxRow]) {
	const int o = (int)(r % g.O);
	int64_t c = r / g.O;
	const int64_t z = c % g.L[2];
	c /= g.L[2];
	const int64_t y = c % g.L[1];
	const int64_t x = c / g.L[1];
	int n = 0;
	auto add = [&](uint32_t col, double2 a) {
		int p = n;
		while (p > 0 && cols[p - 1] > col) p--;
		if (p > 0 && cols[p - 1] == col) {
			vals[p - 1].x = vals[p - 1].x + a.x;
			vals[p - 1].y = vals[p - 1].y + a.y;
			return;
		}
		for (int q = n; q > p; q--) {
			cols[q] = cols[q - 1];
			vals[q] = vals[q - 1];
		}
		cols[p] = col;
		vals[p] = a;
		n++;
	};
	if (g.has_onsite) add((uint32_t)r, onsite[r]);
	for (int t = 0; t < g.nt; t++) {
		const tbtkx_lattice_term &tm = terms[t];
		if (tm.to != o) continue;
		// from-cell = to-cell - d
		const int64_t p[3] = {x - tm.d[0], y - tm.d[1], z - tm.d[2]};
		int64_t q[3];
		bool inside = true;
#pragma unroll
		for (int a = 0; a < 3; a++) {
			int64_t v = p[a];
			if (v < 0 || v >= g.L[a]) {
				if (!g.per[a]) {
					inside = false;
					break;
				}
				v %= g.L[a];
				if (v < 0) v += g.L[a];
			}
			q[a] = v;
		}
		if (!inside) continue;
		const int64_t col = ((q[0] * g.L[1] + q[1]) * g.L[2] + q[2]) * g.O + tm.from;
		add((uint32_t)col, make_double2(tm.re, tm.im));
	}
	return n;
}

// Per row: the number of entries; per workgroup: the largest ELL row length
// (entries, plus one when the diagonal is absent), the band max |col - r|
// and the largest row sum of |H_rc| (the Gershgorin bound), as
// part[3 b .. 3 b + 2] (the bound's bits in the third).
__global__ __launch_bounds__(kThreads) void lattice_count(
    LatDev g, const tbtkx_lattice_term *__restrict__ terms, const double2 *__restrict__ onsite,
    int64_t N, uint32_t *__restrict__ count, int64_t *__restrict__ part) {
	__shared__ int64_t s_len[kThreads], s_band[kThreads];
	__shared__ double s_sum[kThreads];
	const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	int64_t len = 0, band = 0;
	double rsum = 0.0;
	if (r < N) {
		uint32_t cols[kMaxRow];
		double2 vals[kMaxRow];
		const int n = lattice_row(g, terms, onsite, r, cols, vals);
		bool diag = false;
		for (int k = 0; k < n; k++) {
			const int64_t d = (int64_t)cols[k] - r;
			band = max(band, d < 0 ? -d : d);
			if (d == 0) diag = true;
			rsum += hypot(vals[k].x, vals[k].y);
		}
		len = n + (diag ? 0 : 1);
		if (count) count[r] = (uint32_t)n;
	}
	s_len[threadIdx.x] = len;
	s_band[threadIdx.x] = band;
	s_sum[threadIdx.x] = rsum;
	__syncthreads();
	for (int s = kThreads / 2; s > 0; s >>= 1) {
		if ((int)threadIdx.x < s) {
			s_len[threadIdx.x] = max(s_len[threadIdx.x], s_len[threadIdx.x + s]);
			s_band[threadIdx.x] = max(s_band[threadIdx.x], s_band[threadIdx.x + s]);
			s_sum[threadIdx.x] = fmax(s_sum[threadIdx.x], s_sum[threadIdx.x + s]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		part[3 * blockIdx.x] = s_len[0];
		part[3 * blockIdx.x + 1] = s_band[0];
		part[3 * blockIdx.x + 2] = __double_as_longlong(s_sum[0]);
	}
}

// (a + ib)(d + 0i) without contraction: the scalings by 1/s and by 2.
__device__ __forceinline__ double2 cscale(double2 v, double d) {
#pragma clang fp contract(off)
	return make_double2(v.x * d - v.y * 0.0, v.x * 0.0 + v.y * d);
}

// CSR entries (col, H/s, 2H/s in FP64 and FP32, real parts when real) at
// rowptr[r], and the tile-major ELL row (W slots, diagonal first, padding
// slots = the row itself with value 0) for r < ell_rows.
__global__ __launch_bounds__(kThreads) void lattice_fill(
    LatDev g, const tbtkx_lattice_term *__restrict__ terms, const double2 *__restrict__ onsite,
    int64_t N, double inv_s, const uint32_t *__restrict__ rowptr, uint32_t *__restrict__ col,
    double2 *__restrict__ v1, double2 *__restrict__ v2, float2 *__restrict__ v1f,
    float2 *__restrict__ v2f, double *__restrict__ v1r, double *__restrict__ v2r, int W,
    int64_t ell_rows, uint32_t *__restrict__ ecol, float2 *__restrict__ e1f,
    float2 *__restrict__ e2f, double2 *__restrict__ e1d, double2 *__restrict__ e2d) {
	const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (r >= N) {
		if (W > 0 && r < ell_rows)
			for (int u = 0; u < W; u++) {
				const int64_t e = r * W + u;
				ecol[e] = 0;
				e1f[e] = e2f[e] = make_float2(0.0f, 0.0f);
				e1d[e] = e2d[e] = make_double2(0.0, 0.0);
			}
		return;
	}
	uint32_t cols[kMaxRow];
	double2 vals[kMaxRow];
	const int n = lattice_row(g, terms, onsite, r, cols, vals);
	double2 a1[kMaxRow], a2[kMaxRow];
	const int64_t k0 = rowptr[r];
	int diag = -1;
	for (int k = 0; k < n; k++) {
		a1[k] = cscale(vals[k], inv_s);
		a2[k] = cscale(a1[k], 2.0);
		col[k0 + k] = cols[k];
		v1[k0 + k] = a1[k];
		v2[k0 + k] = a2[k];
		// The 8-byte copies: FP32 complex, or for a real H (v1r set) one
		// FP64 value each.
		if (v1r) {
			v1r[k0 + k] = a1[k].x;
			v2r[k0 + k] = a2[k].x;
			reinterpret_cast<double *>(v1f)[k0 + k] = a1[k].x;
			reinterpret_cast<double *>(v2f)[k0 + k] = a2[k].x;
		} else {
			v1f[k0 + k] = make_float2((float)a1[k].x, (float)a1[k].y);
			v2f[k0 + k] = make_float2((float)a2[k].x, (float)a2[k].y);
		}
		if ((int64_t)cols[k] == r && diag < 0) diag = k;
	}
	if (W <= 0) return;
	// slot 0: the diagonal (or a zero at the row itself), then the others
	int u = 0;
	auto put = [&](uint32_t c, double2 x1, double2 x2) {
		const int64_t e = r * W + u++;
		ecol[e] = c;
		e1d[e] = x1;
		e2d[e] = x2;
		if (v1r) {
			reinterpret_cast<double *>(e1f)[e] = x1.x;
			reinterpret_cast<double *>(e2f)[e] = x2.x;
		} else {
			e1f[e] = make_float2((float)x1.x, (float)x1.y);
			e2f[e] = make_float2((float)x2.x, (float)x2.y);
		}
	};
	const double2 zero = make_double2(0.0, 0.0);
	if (diag >= 0) put(cols[diag], a1[diag], a2[diag]);
	else put((uint32_t)r, zero, zero);
	for (int k = 0; k < n; k++)
		if (k != diag) put(cols[k], a1[k], a2[k]);
	while (u < W) put((uint32_t)r, zero, zero);
}

// The uneven batch's row-major ELL rows of one model (FP32 H/s and 2H/s,
// diagonal first, padding slots = the row itself with value 0).
__global__ __launch_bounds__(kThreads) void lattice_fill_multi(
    LatDev g, const tbtkx_lattice_term *__restrict__ terms, const double2 *__restrict__ onsite,
    int64_t N, double inv_s, int W, uint32_t *__restrict__ ecol, float2 *__restrict__ e1,
    float2 *__restrict__ e2, int realh) {
	const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
	if (r >= N) return;
	uint32_t cols[kMaxRow];
	double2 vals[kMaxRow];
	const int n = lattice_row(g, terms, onsite, r, cols, vals);
	int diag = -1;
	for (int k = 0; k < n; k++)
		if ((int64_t)cols[k] == r && diag < 0) diag = k;
	int u = 0;
	// realh: every model of the batch is real; an entry holds one FP64
	// value in the 8 bytes of an FP32 complex value.
	auto put = [&](uint32_t c, double2 x) {
		const double2 x2 = cscale(x, 2.0);
		const int64_t e = r * W + u++;
		ecol[e] = c;
		if (realh) {
			reinterpret_cast<double *>(e1)[e] = x.x;
			reinterpret_cast<double *>(e2)[e] = x2.x;
		} else {
			e1[e] = make_float2((float)x.x, (float)x.y);
			e2[e] = make_float2((float)x2.x, (float)x2.y);
		}
	};
	if (diag >= 0) put(cols[diag], cscale(vals[diag], inv_s));
	else {
		ecol[r * W] = (uint32_t)r;
		e1[r * W] = e2[r * W] = make_float2(0.0f, 0.0f);
		u = 1;
	}
	for (int k = 0; k < n; k++)
		if (k != diag) put(cols[k], cscale(vals[k], inv_s));
	for (; u < W; u++) {
		ecol[r * W + u] = (uint32_t)r;
		e1[r * W + u] = e2[r * W + u] = make_float2(0.0f, 0.0f);
	}
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Host-side checks of one lattice; fills the device description. `max_row`
// is an upper bound of a row's entries.
int check_lattice(const tbtkx_lattice *lat, const double *onsite, const std::string &tag,
                  LatDev *g, int64_t *N, int *max_row, bool *real) {
	TBTKX_REQUIRE(lat, TBTKX_ERR_ARG, tag + "null lattice");
	TBTKX_REQUIRE(lat->orbitals >= 1 && lat->num_terms >= 0 &&
	                  (lat->num_terms == 0 || lat->terms),
	              TBTKX_ERR_ARG, tag + "need orbitals >= 1 and a term array");
	int64_t n = lat->orbitals;
	for (int a = 0; a < 3; a++) {
		TBTKX_REQUIRE(lat->cells[a] >= 1, TBTKX_ERR_ARG, tag + "cells must be >= 1");
		n *= lat->cells[a];
		TBTKX_REQUIRE(n < ((int64_t)1 << 31), TBTKX_ERR_ARG,
		              tag + "more than 2^31 - 1 states");
		g->L[a] = lat->cells[a];
		g->per[a] = lat->periodic[a] != 0;
	}
	std::vector<int> per_orbital(lat->orbitals, onsite ? 1 : 0);
	*real = true;
	for (int t = 0; t < lat->num_terms; t++) {
		const tbtkx_lattice_term &tm = lat->terms[t];
		TBTKX_REQUIRE(tm.to >= 0 && tm.to < lat->orbitals && tm.from >= 0 &&
		                  tm.from < lat->orbitals,
		              TBTKX_ERR_ARG, tag + "term " + std::to_string(t) + ": orbital out of range");
		per_orbital[tm.to]++;
		if (tm.im != 0.0) *real = false;
	}
	if (onsite)
		for (int64_t r = 0; r < n && *real; r++)
			if (onsite[2 * r + 1] != 0.0) *real = false;
	*max_row = *std::max_element(per_orbital.begin(), per_orbital.end());
	TBTKX_REQUIRE(*max_row <= kMaxRow, TBTKX_ERR_ARG,
	              tag + "a row may collect at most 16 entries (on-site and terms of one "
	                    "to-orbital)");
	g->O = lat->orbitals;
	g->nt = lat->num_terms;
	g->has_onsite = onsite != nullptr;
	*N = n;
	return TBTKX_OK;
}

// Upload a lattice's terms and on-site amplitudes to the handle's staging
// buffers.
int stage_lattice(tbtkx_context *h, const tbtkx_lattice *lat, const double *onsite, int64_t N) {
	int rc;
	const size_t tb = (size_t)std::max(lat->num_terms, 1) * sizeof(tbtkx_lattice_term);
	if ((rc = h->lat_terms.reserve(tb))) return rc;
	if (lat->num_terms)
		TBTKX_CHECK_HIP(hipMemcpy(h->lat_terms.ptr, lat->terms,
		                          (size_t)lat->num_terms * sizeof(tbtkx_lattice_term),
		                          hipMemcpyHostToDevice));
	if (onsite) {
		if ((rc = h->lat_onsite.reserve((size_t)N * 16))) return rc;
		TBTKX_CHECK_HIP(hipMemcpy(h->lat_onsite.ptr, onsite, (size_t)N * 16,
		                          hipMemcpyHostToDevice));
	}
	return TBTKX_OK;
}

// Largest ELL row length, band and Gershgorin bound of a lattice (and
// per-row counts into h->lat_count when `counts`).
int lattice_extent(tbtkx_context *h, const LatDev &g, int64_t N, bool counts, int64_t *maxlen,
                   int64_t *band, double *bound = nullptr) {
	int rc;
	const unsigned nb = grid_for(N);
	if ((rc = h->lat_part.reserve((size_t)nb * 24))) return rc;
	if (counts && (rc = h->lat_count.reserve((size_t)N * 4))) return rc;
	lattice_count<<<nb, kThreads, 0, h->stream>>>(
	    g, h->lat_terms.as<tbtkx_lattice_term>(),
	    g.has_onsite ? h->lat_onsite.as<double2>() : nullptr, N,
	    counts ? h->lat_count.as<uint32_t>() : nullptr, h->lat_part.as<int64_t>());
	TBTKX_CHECK_HIP(hipGetLastError());
	std::vector<int64_t> part((size_t)nb * 3);
	TBTKX_CHECK_HIP(hipMemcpyAsync(part.data(), h->lat_part.ptr, part.size() * 8,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	*maxlen = 0;
	*band = 0;
	double gb = 0.0;
	for (unsigned b = 0; b < nb; b++) {
		*maxlen = std::max(*maxlen, part[3 * b]);
		*band = std::max(*band, part[3 * b + 1]);
		double x;
		memcpy(&x, &part[3 * b + 2], 8);
		gb = std::max(gb, x);
	}
	if (bound) *bound = gb;
	return TBTKX_OK;
}

}  // namespace
}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

int tbtkx_set_lattice(tbtkx_handle h, const tbtkx_lattice *lat, const double *onsite,
                      double scale) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_set_lattice: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(scale > 0, TBTKX_ERR_ARG,
	              "tbtkx_set_lattice: the scale factor has to be larger than 0");
	LatDev g{};
	int64_t N = 0;
	int max_row = 0;
	bool real = true;
	int rc = check_lattice(lat, onsite, "tbtkx_set_lattice: ", &g, &N, &max_row, &real);
	if (rc) return rc;
	TBTKX_REQUIRE(N * max_row + kCsrPad < ((int64_t)1 << 32), TBTKX_ERR_ARG,
	              "tbtkx_set_lattice: more than 2^32 non-zeros");
	if ((rc = set_device(h))) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->plain_graph.reset();
	h->blk_graph.reset();
	if ((rc = stage_lattice(h, lat, onsite, N))) return rc;
	int64_t maxlen = 0, band = 0;
	if ((rc = lattice_extent(h, g, N, true, &maxlen, &band))) return rc;

	// rowptr = 0, inclusive sum of the counts
	if ((rc = h->rowptr.reserve((size_t)(N + 1) * 4))) return rc;
	uint32_t *rp = h->rowptr.as<uint32_t>();
	TBTKX_CHECK_HIP(hipMemsetAsync(rp, 0, 4, h->stream));
	size_t tmp = 0;
	TBTKX_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, h->lat_count.as<uint32_t>(),
	                                                 rp + 1, (int)N, h->stream));
	if ((rc = h->lat_tmp.reserve(tmp > 0 ? tmp : 1))) return rc;
	TBTKX_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(h->lat_tmp.ptr, tmp,
	                                                 h->lat_count.as<uint32_t>(), rp + 1, (int)N,
	                                                 h->stream));
	// The row pointer comes back for the plain kernel's row blocks.
	std::vector<uint32_t> hrp((size_t)N + 1);
	TBTKX_CHECK_HIP(hipMemcpyAsync(hrp.data(), rp, hrp.size() * 4, hipMemcpyDeviceToHost,
	                               h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	const int64_t Z = hrp[N];
	const int64_t Zp = Z + kCsrPad;
	std::vector<uint32_t> blocks = make_row_blocks(N, hrp.data());

	if ((rc = h->col.reserve(Zp * 4))) return rc;
	if ((rc = h->v1.reserve(Zp * 16))) return rc;
	if ((rc = h->v2.reserve(Zp * 16))) return rc;
	if ((rc = h->v1f.reserve(Zp * 8))) return rc;
	if ((rc = h->v2f.reserve(Zp * 8))) return rc;
	if (real) {
		if ((rc = h->v1r.reserve(Zp * 8))) return rc;
		if ((rc = h->v2r.reserve(Zp * 8))) return rc;
	}
	if ((rc = h->rblk.reserve(blocks.size() * 4))) return rc;
	if ((rc = h->vec_a.reserve(N * 16))) return rc;
	if ((rc = h->vec_b.reserve(N * 16))) return rc;
	const int W = ell_slots_block(maxlen);
	int64_t ell_rows = 0;
	if (W) {
		const int T = 64 / W;
		ell_rows = (N + T - 1) / T * T;
		const int64_t E = ell_rows * W;
		if ((rc = h->ell_col.reserve(E * 4))) return rc;
		if ((rc = h->ell_v1f.reserve(E * 8))) return rc;
		if ((rc = h->ell_v2f.reserve(E * 8))) return rc;
		if ((rc = h->ell_v1d.reserve(E * 16))) return rc;
		if ((rc = h->ell_v2d.reserve(E * 16))) return rc;
	}
	// Padding after the last non-zero: column 0, value 0.
	TBTKX_CHECK_HIP(hipMemsetAsync(h->col.as<uint32_t>() + Z, 0, kCsrPad * 4, h->stream));
	TBTKX_CHECK_HIP(hipMemsetAsync(h->v1.as<double2>() + Z, 0, kCsrPad * 16, h->stream));
	TBTKX_CHECK_HIP(hipMemsetAsync(h->v2.as<double2>() + Z, 0, kCsrPad * 16, h->stream));
	TBTKX_CHECK_HIP(hipMemsetAsync(h->v1f.as<float2>() + Z, 0, kCsrPad * 8, h->stream));
	TBTKX_CHECK_HIP(hipMemsetAsync(h->v2f.as<float2>() + Z, 0, kCsrPad * 8, h->stream));
	if (real) {
		TBTKX_CHECK_HIP(hipMemsetAsync(h->v1r.as<double>() + Z, 0, kCsrPad * 8, h->stream));
		TBTKX_CHECK_HIP(hipMemsetAsync(h->v2r.as<double>() + Z, 0, kCsrPad * 8, h->stream));
	}
	lattice_fill<<<grid_for(std::max(N, ell_rows)), kThreads, 0, h->stream>>>(
	    g, h->lat_terms.as<tbtkx_lattice_term>(), onsite ? h->lat_onsite.as<double2>() : nullptr,
	    N, 1 / scale, rp, h->col.as<uint32_t>(), h->v1.as<double2>(), h->v2.as<double2>(),
	    h->v1f.as<float2>(), h->v2f.as<float2>(), real ? h->v1r.as<double>() : nullptr,
	    real ? h->v2r.as<double>() : nullptr, W, ell_rows, W ? h->ell_col.as<uint32_t>() : nullptr,
	    W ? h->ell_v1f.as<float2>() : nullptr, W ? h->ell_v2f.as<float2>() : nullptr,
	    W ? h->ell_v1d.as<double2>() : nullptr, W ? h->ell_v2d.as<double2>() : nullptr);
	TBTKX_CHECK_HIP(hipGetLastError());
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->rblk.ptr, blocks.data(), blocks.size() * 4,
	                               hipMemcpyHostToDevice, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->ell_W = W;
	h->band = band > 0 ? band : N;
	h->N = N;
	h->Z = Z;
	h->scale = scale;
	h->csr_gen++;
	h->h_real = real;
	h->nblk = (int)blocks.size() - 1;
	if ((rc = build_col16(h))) return rc;
	// The value palette, the row patterns and the slot form of the plain
	// step come from a host copy of the device CSR; they are built on the
	// first plain recurrence (build_plain_forms), so a model used only by
	// the batched paths never copies its CSR back.
	h->palette_ok = 0;
	h->pat_ok = 0;
	h->slot_ok = 0;
	h->slot_pal = 0;
	h->slot_vals = 0;
	h->forms_pending = 1;
	return TBTKX_OK;
}

int tbtkx_lattice_gershgorin(tbtkx_handle h, const tbtkx_lattice *lat, const double *onsite,
                             double *bound) {
	TBTKX_REQUIRE(h && bound, TBTKX_ERR_ARG, "tbtkx_lattice_gershgorin: null pointer");
	TBTKX_LOCK(h);
	LatDev g{};
	int64_t N = 0;
	int max_row = 0;
	bool real = true;
	int rc = check_lattice(lat, onsite, "tbtkx_lattice_gershgorin: ", &g, &N, &max_row, &real);
	if (rc) return rc;
	if ((rc = set_device(h))) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	if ((rc = stage_lattice(h, lat, onsite, N))) return rc;
	int64_t maxlen = 0, band = 0;
	return lattice_extent(h, g, N, false, &maxlen, &band, bound);
}

int tbtkx_csr_size(tbtkx_handle h, int64_t *num_rows, int64_t *nnz) {
	TBTKX_REQUIRE(h && num_rows && nnz, TBTKX_ERR_ARG, "tbtkx_csr_size: null pointer");
	TBTKX_LOCK(h);
	*num_rows = h->N;
	*nnz = h->Z;
	return TBTKX_OK;
}

int tbtkx_csr_copy(tbtkx_handle h, uint32_t *rowptr, uint32_t *col, double *values) {
	TBTKX_REQUIRE(h && rowptr && (h->Z == 0 || (col && values)), TBTKX_ERR_ARG,
	              "tbtkx_csr_copy: null pointer");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->N > 0, TBTKX_ERR_STATE, "tbtkx_csr_copy: no model set");
	int rc = set_device(h);
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	TBTKX_CHECK_HIP(hipMemcpy(rowptr, h->rowptr.ptr, (size_t)(h->N + 1) * 4,
	                          hipMemcpyDeviceToHost));
	if (h->Z) {
		TBTKX_CHECK_HIP(hipMemcpy(col, h->col.ptr, (size_t)h->Z * 4, hipMemcpyDeviceToHost));
		TBTKX_CHECK_HIP(hipMemcpy(values, h->v1.ptr, (size_t)h->Z * 16, hipMemcpyDeviceToHost));
	}
	return TBTKX_OK;
}

int tbtkx_multi_set_lattices(tbtkx_handle h, int K, const tbtkx_lattice *lats,
                             const double *const *onsite, const double *scales) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_multi_set_lattices: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(K > 0 && K <= 65536 && lats && scales, TBTKX_ERR_ARG,
	              "tbtkx_multi_set_lattices: need 1..65536 models and arrays");
	std::vector<LatDev> g(K);
	std::vector<int64_t> N(K);
	std::vector<int> realm(K, 1);   // per model: every value of its H is real
	bool real = true;
	int64_t rows = 0;
	for (int m = 0; m < K; m++) {
		const std::string tag = "tbtkx_multi_set_lattices: model " + std::to_string(m) + ": ";
		TBTKX_REQUIRE(scales[m] > 0, TBTKX_ERR_ARG,
		              tag + "the scale factor has to be larger than 0");
		int max_row = 0;
		bool rm = true;
		int rc = check_lattice(&lats[m], onsite ? onsite[m] : nullptr, tag, &g[m], &N[m],
		                       &max_row, &rm);
		if (rc) return rc;
		real = real && rm;
		realm[m] = rm ? 1 : 0;
		rows += N[m];
	}
	TBTKX_REQUIRE(rows * 8 < ((int64_t)1 << 40), TBTKX_ERR_ARG,
	              "tbtkx_multi_set_lattices: batch too large");
	int rc = set_device(h);
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	// Pass 1: row lengths and bands, for each model's slot count.
	std::vector<int64_t> band(K);
	std::vector<int> Wm(K);
	int64_t maxlen = 0, entries = 0;
	for (int m = 0; m < K; m++) {
		if ((rc = stage_lattice(h, &lats[m], onsite ? onsite[m] : nullptr, N[m]))) return rc;
		int64_t ml = 0;
		if ((rc = lattice_extent(h, g[m], N[m], false, &ml, &band[m]))) return rc;
		maxlen = std::max(maxlen, ml);
		Wm[m] = ell_slots_multi(ml);
		entries += N[m] * Wm[m];
	}
	TBTKX_REQUIRE(maxlen <= 8, TBTKX_ERR_ARG,
	              "tbtkx_multi_set_lattices: rows with more than 8 non-zeros (7 without a "
	              "diagonal entry) are not supported");
	const int W = ell_slots_multi(maxlen);
	h->reset_mm_graphs();
	h->mm_K = 0;
	if ((rc = h->mm_ecol.reserve((size_t)entries * 4))) return rc;
	if ((rc = h->mm_ev1.reserve((size_t)entries * 8))) return rc;
	if ((rc = h->mm_ev2.reserve((size_t)entries * 8))) return rc;
	h->mm_N.assign(K, 0);
	h->mm_e0.assign(K, 0);
	h->mm_Wm = Wm;
	h->mm_band.assign(K, 0);
	// Pass 2: each model's ELL rows.
	int64_t e0 = 0;
	for (int m = 0; m < K; m++) {
		if ((rc = stage_lattice(h, &lats[m], onsite ? onsite[m] : nullptr, N[m]))) return rc;
		lattice_fill_multi<<<grid_for(N[m]), kThreads, 0, h->stream>>>(
		    g[m], h->lat_terms.as<tbtkx_lattice_term>(),
		    g[m].has_onsite ? h->lat_onsite.as<double2>() : nullptr, N[m], 1 / scales[m], Wm[m],
		    h->mm_ecol.as<uint32_t>() + e0, h->mm_ev1.as<float2>() + e0,
		    h->mm_ev2.as<float2>() + e0, realm[m]);
		TBTKX_CHECK_HIP(hipGetLastError());
		// The staging buffers are reused by the next model.
		TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
		h->mm_band[m] = band[m] > 0 ? band[m] : N[m];
		h->mm_N[m] = N[m];
		h->mm_e0[m] = e0;
		e0 += N[m] * Wm[m];
	}
	h->mm_K = K;
	h->mm_W = W;
	h->mm_real = real;
	h->mm_realm = realm;
	h->mm_gen++;
	for (auto &k : h->mm_tune_key) k = -1;
	return TBTKX_OK;
}

}  // extern "C"
