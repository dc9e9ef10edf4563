// Shared definitions of the tbtkx engine (HIP, gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tbtkx.h"

namespace tbtkx {

// ---------------------------------------------------------------- errors --

void set_error(const std::string &msg);

#define TBTKX_CHECK_HIP(expr)                                                  \
	do {                                                                   \
		hipError_t _e = (expr);                                        \
		if (_e != hipSuccess) {                                        \
			::tbtkx::set_error(std::string(#expr) + ": " +         \
			                   hipGetErrorString(_e) + " at " +    \
			                   __FILE__ + ":" +                    \
			                   std::to_string(__LINE__));          \
			return _e == hipErrorOutOfMemory ? TBTKX_ERR_OOM       \
			                                 : TBTKX_ERR_HIP;      \
		}                                                              \
	} while (0)

#define TBTKX_REQUIRE(cond, code, msg)                                         \
	do {                                                                   \
		if (!(cond)) {                                                 \
			::tbtkx::set_error(msg);                               \
			return (code);                                         \
		}                                                              \
	} while (0)

// Entry points that take a handle hold its lock for the whole call, so one
// handle may be shared by several host threads (their calls run one after
// another). Recursive: an entry point may call another on the same handle.
#define TBTKX_LOCK(h) std::lock_guard<std::recursive_mutex> tbtkx_call_lock_((h)->call_mutex)

// ------------------------------------------------------------ device mem --

// Owning device buffer; grows on demand, never shrinks until destroyed.
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	int reserve(size_t want);
	void release();
	template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// ----------------------------------------------------------------- state --

constexpr int kThreads = 256;       // workgroup size of every step kernel
constexpr int kCsrPad = 64;         // zero entries after the last non-zero
constexpr int kPlainCap = 1536;     // products staged in LDS per row block
constexpr int kPlainRows = 256;     // rows per row block (one per thread)

// One instantiated graph and the key it was captured for. Keys hold the
// launch parameters and the device addresses of every buffer the graph
// touches, so a buffer that was reallocated since the capture never gets
// replayed.
struct GraphCache {
	static constexpr int kKey = 40;
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	int64_t key[kKey];
	GraphCache() { reset_key(); }
	void reset_key() {
		for (auto &k : key) k = -1;
	}
	void reset();
	// Fill `out` with the values, then the addresses, zero after. Returns
	// false (and writes nothing past the array) if they do not fit.
	static bool make_key(int64_t (&out)[kKey], std::initializer_list<int64_t> vals,
	                     std::initializer_list<const void *> ptrs) {
		if (vals.size() + ptrs.size() > (size_t)kKey) return false;
		int i = 0;
		for (int64_t v : vals) out[i++] = v;
		for (const void *q : ptrs) out[i++] = (int64_t)(intptr_t)q;
		while (i < kKey) out[i++] = 0;
		return true;
	}
	bool matches(const int64_t (&k)[kKey]) const {
		return exec && memcmp(k, key, sizeof key) == 0;
	}
};

}  // namespace tbtkx

struct tbtkx_context {
	std::recursive_mutex call_mutex;   // held by every entry point (TBTKX_LOCK)
	int device = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev_start = nullptr, ev_stop = nullptr;
	double last_ms = 0;
	int64_t last_launches = 0;

	// Hamiltonian (CSR), scaled by 1/s (v1) and 2/s (v2).
	int64_t N = 0, Z = 0;
	double scale = 0;
	bool h_real = false;   // every value of H has a zero imaginary part
	tbtkx::DevBuf rowptr, col, v1, v2, v1f, v2f;
	// Real parts of v1 and v2 (FP64), present when h_real.
	tbtkx::DevBuf v1r, v2r;
	// Staging of tbtkx_set_lattice / tbtkx_multi_set_lattices: terms,
	// on-site amplitudes, row counts, per-workgroup extents, scan storage.
	tbtkx::DevBuf lat_terms, lat_onsite, lat_count, lat_part, lat_tmp;
	// Tile-major ELL copy of H for the batched step (ell_W = slots per row,
	// ell_slots_block; 0 when a row is longer than 8): tiles of 64/ell_W
	// rows.
	tbtkx::DevBuf ell_col, ell_v1f, ell_v2f, ell_v1d, ell_v2d;
	int ell_W = 0;
	int64_t band = 0;      // max |col - row| over the non-zeros (N if diagonal)
	// Row blocks for the LDS-staged plain kernel.
	tbtkx::DevBuf rblk;
	int nblk = 0;
	// Columns of the plain kernel as 16-bit offsets from the first row of
	// their row block (col16_ok when every offset fits in int16).
	tbtkx::DevBuf col16, col16_flag;
	int col16_ok = 0;
	// Value palette of the plain kernel (palette_ok): one code per non-zero,
	// 255 = the row's diagonal (pdiag: H/s then 2H/s, N each), else an index
	// into pal (256 values of H/s, then 256 of 2H/s); palr/pdiagr: real parts.
	tbtkx::DevBuf pcode, pal, palr, pdiag, pdiagr;
	int palette_ok = 0;
	// Slot-major ELL form of the plain kernel (slot_ok: every row has at
	// most 8 non-zeros and every column lies within int16 of its row):
	// slot_W slots per row, 16-bit column offsets from the row, palette codes
	// (slot_pal) or H/s, 2H/s values (complex, and real for a real H).
	tbtkx::DevBuf scol, scode, sv1, sv2, sv1r, sv2r;
	int slot_ok = 0, slot_pal = 0, slot_vals = 0, slot_W = 0;
	// Row patterns of the slot form (pat_ok): one byte per row (ppid) into
	// pat_count patterns of 8 offsets (ptoff) and 8 palette codes (ptcode).
	tbtkx::DevBuf ppid, ptoff, ptcode;
	tbtkx::DevBuf pchunk;   // per 64-row chunk: its rows' common pattern, or 255
	int pat_ok = 0, pat_count = 0;
	// Device-built models: palette, patterns and slots not built yet (they
	// are built from a host copy of the CSR on the first plain recurrence).
	int forms_pending = 0;
	int64_t pat_byte_rows = 0;   // rows outside uniform 64-row chunks
	int use_patterns = 1;
	int slot_rows = 0;       // rows per thread of the slot step (1, 2, 4; 0 = by form)
	int last_patterns = 0;   // the last plain call used the row patterns
	int last_plain_kernel = 0;   // kernel of the last plain call (0, 1, 2)

	// Plain path vectors and moment output.
	tbtkx::DevBuf vec_a, vec_b, coef, to;

	// Lookup table L[n*E + e], complex double.
	tbtkx::DevBuf lut;
	int lut_M = 0, lut_E = 0;
	tbtkx::DevBuf gf_in, gf_out;
	// Factors of the direct contraction (host libm values for one Green's
	// function kind, M moments and one energy grid) and sqrt(1 - E~^2).
	tbtkx::DevBuf dfac, dden;
	int dfac_kind = -1, dfac_M = 0, dfac_E = 0;
	double dfac_e0 = 0, dfac_de = 0, dfac_s = 0;
	// Split-BF16 planes for the MFMA contraction.
	tbtkx::DevBuf cplanes, lplanes, gf_mfma_out;
	int cplanes_K = 0, cplanes_M = 0;
	int lplanes_valid = 0;   // the table planes match the current table

	// Batched path buffers.
	tbtkx::DevBuf blk_x, blk_y, blk_mu, blk_sites, blk_part;
	// Gather list of the exact blocks (tbtkx_moments_multi): rows and
	// vectors within the block.
	tbtkx::DevBuf blk_grow, blk_gvec;
	int64_t csr_gen = 0;                          // bumped by every tbtkx_set_csr
	int64_t blk_tune_key[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
	int blk_tune_B = 0, blk_tune_S = -1, blk_tune_share = 1, blk_tune_rmul = 1;
	int run_strip = -1, run_share = 1, run_rmul = 1;   // shape of the current call
	int last_block_width = 0;                     // shape of the last batched call
	int last_strip = -1, last_share = 1, last_rmul = 1;
	int last_tile_groups = 0;   // row groups per tile of the last uneven-batch call
	tbtkx::GraphCache blk_graph;

	// Uneven multi-model batch (tbtkx_multi_*): every model's H as row-major
	// ELL with mm_Wm[m] slots per row (mm_W: the widest; FP32, scaled by
	// 1/s_m and 2/s_m), rows of
	// all models concatenated in upload order.
	int mm_K = 0, mm_W = 0;
	bool mm_real = false;
	int mm_packed = 0;                        // current call packs real vectors
	int64_t mm_gen = 0;                       // bumped by every tbtkx_multi_set
	std::vector<int64_t> mm_N, mm_e0;         // rows, first ELL entry per model
	std::vector<int> mm_Wm;                   // ELL slots per row, per model
	std::vector<int> mm_realm;                // per model: H real (FP64 ELL values)
	std::vector<int64_t> mm_band;             // max |col - row| per model
	tbtkx::DevBuf mm_ecol, mm_ev1, mm_ev2;
	tbtkx::DevBuf mm_desc, mm_tab, mm_x, mm_y, mm_part, mm_seg, mm_D, mm_cols;
	int64_t mm_tune_key[3] = {-1, -1, -1};    // (generation, flags, widths hash)
	int mm_tune_g = 0, mm_tune_occ = 0;       // tuned tile groups, grid share
	std::vector<tbtkx::GraphCache> mm_graphs;   // the run in bounded chunks
	void reset_mm_graphs() {
		for (tbtkx::GraphCache &g : mm_graphs) g.reset();
		mm_graphs.clear();
	}

	// Options.
	// Damping of the contraction's coefficients (TBTKX_DAMP_*).
	int damp_kind = 0;
	double damp_broadening = 0;

	int plain_kernel = 2;
	int use_col16 = 1;
	int last_col16 = 0;   // the last plain call used 16-bit columns
	int use_palette = 1;
	int last_palette = 0; // the last plain call used the value palette
	int use_graph = 0;
	int block_width = 0;   // 0 = autotune
	int xcd_remap = 0;     // measured faster without for both step kernels
	int tile_rows = 0;     // 0 = default
	int nt_stream = 1;     // non-temporal hints on once-used SpMM rows
	int use_ell = 1;       // batched step on the ELL copy when it exists
	int autotune = 1;      // time candidate launch shapes on first use
	int mfma_pipe = 1;     // double-buffered direct-to-LDS MFMA contraction
	int strip_rows = 0;    // L2 strip width S of the batched step (0 = tuned,
	                       // -1 = natural row order, the tile-major ELL step)
	int grid_share = 0;    // batched grid = occupancy x CUs / share (0 = tuned)
	int strip_chunk = 0;   // consecutive strip tiles per workgroup (multiple
	                       // of 4; 0 = 4)
	int rows_in_flight = 0;   // strip step rows-in-flight multiplier (0 = tuned)
	int multi_mode = 0;    // tbtkx_moments_multi: 0 by cost, 1 plain, 2 exact blocks
	int last_multi_blocks = 0;   // the last tbtkx_moments_multi ran exact blocks
	int real_path = 1;     // real arithmetic for real H and real start vectors
	int last_real = 0;     // the last recurrence call ran in real arithmetic

	tbtkx::GraphCache plain_graph;
};

namespace tbtkx {

// Map the hardware workgroup id to a logical one so that each of the 8 XCDs
// (which receive workgroups round-robin) works on one contiguous range of
// rows and keeps its neighbours' vector entries in its own L2.
__device__ __forceinline__ int xcd_remap(int hw, int n) {
	const int xcd = hw & 7;
	const int local = hw >> 3;
	const int per = n >> 3;
	const int rem = n & 7;
	return xcd * per + (xcd < rem ? xcd : rem) + local;
}

int set_device(tbtkx_context *h);

// Damping set by tbtkx_set_damping, applied in FP64 to K x M coefficients
// in device memory on the handle's stream (gf.hip).
int apply_damping(tbtkx_context *h, double2 *C, int K, int M);

// Builds the 16-bit column offsets of the resident CSR (plain.hip); sets
// col16_ok. Needs rowptr, col, rblk and nblk in place.
int build_col16(tbtkx_context *h);

// Builds the slot-major ELL form of the plain kernel from the resident CSR
// and palette (plain.hip); `rowptr` is a host copy of the row pointer.
int build_slots(tbtkx_context *h, const uint32_t *rowptr);
// Adds H/s and 2H/s values to the slots (for calls without the palette).
int build_slot_values(tbtkx_context *h);
// Builds the palette, the row patterns and the slot form of a device-built
// model from a host copy of its CSR, once (forms_pending).
int build_plain_forms(tbtkx_context *h);

// The form the plain recurrence step runs in: the slot-major ELL kernel
// (TBTKX_OPT_PLAIN_KERNEL 2, needs 16-bit columns), 16-bit column offsets,
// the value palette; and the algorithmic HBM bytes of one step in that form.
struct PlainForm {
	bool slot, c16, pal, pat;
	double bytes;
};
inline PlainForm plain_form(const tbtkx_context *h) {
	const bool real = h->real_path && h->h_real;
	PlainForm f{};
	f.slot = h->plain_kernel == 2 && h->slot_ok && h->use_col16 &&
	         ((h->use_palette && h->slot_pal) || h->slot_vals);
	f.c16 = f.slot || (h->plain_kernel == 0 && h->use_col16 && h->col16_ok);
	f.pal = f.c16 && h->use_palette && (f.slot ? h->slot_pal : h->palette_ok);
	f.pat = f.slot && f.pal && h->use_patterns && h->pat_ok;
	const double N = (double)h->N, Z = (double)h->Z, e = real ? 8 : 16;
	if (f.pat)
		// per row: the pattern id, x, j read and written, the diagonal
		f.bytes = N * (1 + 4 * e);
	else if (f.slot)
		// per slot: 2 B column + 1 B code or a value; per row: x, j read and
		// written, and the diagonal with the palette
		f.bytes = (double)h->slot_W * N * (2 + (f.pal ? 1 : e)) + N * (3 * e + (f.pal ? e : 0));
	else if (f.pal)
		f.bytes = 3 * Z + N * (4 + 4 * e);
	else
		f.bytes = (e + (f.c16 ? 2 : 4)) * Z + N * (4 + 3 * e);
	return f;
}

// Builds the plain kernel's value palette from host copies of the CSR
// (v1 = H/s, v2 = 2H/s, interleaved complex); sets palette_ok when the
// off-diagonal values take at most 255 distinct (H/s, 2H/s) pairs.
int build_palette(tbtkx_context *h, int64_t N, int64_t Z, const uint32_t *rowptr,
                  const uint32_t *col, const double *v1, const double *v2);

// Slot order of a row in the ELL copies of H: the diagonal entry first (k =
// -1 when the row has none: a zero value in column r), then the other
// entries in ascending column order. With the row itself in slot 0, the
// kernels that need <Phi_n| for the moment dot products reuse that gather.
// Returns the number of slots; `slots` must hold the row length + 1.
inline int ell_row_slots(const uint32_t *rowptr, const uint32_t *col, int64_t r,
                         int64_t *slots) {
	int64_t diag = -1;
	for (int64_t k = rowptr[r]; k < (int64_t)rowptr[r + 1]; k++)
		if ((int64_t)col[k] == r && diag < 0) diag = k;
	int n = 0;
	slots[n++] = diag;
	for (int64_t k = rowptr[r]; k < (int64_t)rowptr[r + 1]; k++)
		if (k != diag) slots[n++] = k;
	return n;
}

// out = in * (d, 0) as complex products (contraction off): the scaling of the
// CSR values by 1/s and by 2.
void scale_complex(const double *in, double *out, int64_t Z, double d);
// Row blocks of the LDS-staged plain kernel from a host row pointer.
std::vector<uint32_t> make_row_blocks(int64_t N, const uint32_t *rowptr);

// ELL slots per row for a longest row of maxlen entries (diagonal
// included): the single-model batched step has kernels for 4, 6, 7 and 8
// slots, the uneven batch for 4, 5, 6 and 8; 0: longer than 8.
inline int ell_slots_block(int64_t maxlen) {
	return maxlen <= 4 ? 4 : maxlen <= 6 ? 6 : maxlen <= 8 ? (int)maxlen : 0;
}
inline int ell_slots_multi(int64_t maxlen) {
	return maxlen <= 4 ? 4 : maxlen <= 6 ? (int)maxlen : maxlen <= 8 ? 8 : 0;
}

}  // namespace tbtkx
