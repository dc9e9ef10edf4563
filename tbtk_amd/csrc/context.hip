// Handle lifetime, device pool, Hamiltonian upload and options.
#include <condition_variable>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace tbtkx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int DevBuf::reserve(size_t want) {
	if (want <= bytes && ptr) return TBTKX_OK;
	release();
	if (want == 0) want = 256;
	hipError_t e = hipMalloc(&ptr, want);
	if (e != hipSuccess) {
		ptr = nullptr;
		set_error("hipMalloc(" + std::to_string(want) + " B): " +
		          hipGetErrorString(e));
		return e == hipErrorOutOfMemory ? TBTKX_ERR_OOM : TBTKX_ERR_HIP;
	}
	bytes = want;
	return TBTKX_OK;
}

void DevBuf::release() {
	if (ptr) (void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}

void GraphCache::reset() {
	if (exec) (void)hipGraphExecDestroy(exec);
	if (graph) (void)hipGraphDestroy(graph);
	exec = nullptr;
	graph = nullptr;
	reset_key();
}

int set_device(tbtkx_context *h) {
	TBTKX_CHECK_HIP(hipSetDevice(h->device));
	return TBTKX_OK;
}

// Device pool: one flag per device, guarded by a mutex. acquire() waits on a
// condition variable instead of the reference's spin on an omp_lock_t.
namespace {
std::mutex g_pool_mutex;
std::condition_variable g_pool_cv;
std::vector<char> g_pool_busy;
std::vector<int> g_pool_handles;   // live handles per device
bool g_pool_ready = false;
}  // namespace

static int pool_init() {
	if (g_pool_ready) return TBTKX_OK;
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n <= 0) {
		set_error("no HIP device available");
		return TBTKX_ERR_NODEVICE;
	}
	g_pool_busy.assign(n, 0);
	g_pool_handles.assign(n, 0);
	g_pool_ready = true;
	return TBTKX_OK;
}

// The reference scales the CSR values in place, first by the complex scalar
// (1/s, 0) and later by (2, 0) (Lib/src/Solver/ChebyshevExpander.cpp:147,168,
// SparseMatrix::operator*= at Lib/include/Utilities/TBTK/SparseMatrix.h:
// 1106-1125). The same two complex products are formed here on the host.
void scale_complex(const double *in, double *out, int64_t Z, double d) {
#pragma clang fp contract(off)
	for (int64_t k = 0; k < Z; k++) {
		const double a = in[2 * k], b = in[2 * k + 1];
		out[2 * k] = a * d - b * 0.0;
		out[2 * k + 1] = a * 0.0 + b * d;
	}
}

// Row blocks for the LDS-staged kernel: at most kPlainRows rows and
// kPlainCap non-zeros each; a row longer than kPlainCap is a block of its own.
std::vector<uint32_t> make_row_blocks(int64_t N, const uint32_t *rowptr) {
	std::vector<uint32_t> b;
	b.push_back(0);
	int64_t r = 0;
	while (r < N) {
		const int64_t start = r;
		int64_t nnz = 0;
		while (r < N && r - start < kPlainRows &&
		       nnz + (rowptr[r + 1] - rowptr[r]) <= kPlainCap) {
			nnz += rowptr[r + 1] - rowptr[r];
			r++;
		}
		if (r == start) r++;
		b.push_back((uint32_t)r);
	}
	return b;
}

}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

const char *tbtkx_last_error(void) { return g_last_error.c_str(); }

int tbtkx_device_count(int *count) {
	TBTKX_REQUIRE(count, TBTKX_ERR_ARG, "tbtkx_device_count: null pointer");
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) n = 0;
	*count = n;
	return TBTKX_OK;
}

int tbtkx_device_acquire(int *device) {
	TBTKX_REQUIRE(device, TBTKX_ERR_ARG, "tbtkx_device_acquire: null pointer");
	std::unique_lock<std::mutex> lock(g_pool_mutex);
	int rc = pool_init();
	if (rc) return rc;
	for (;;) {
		for (size_t d = 0; d < g_pool_busy.size(); d++) {
			if (!g_pool_busy[d]) {
				g_pool_busy[d] = 1;
				*device = (int)d;
				return TBTKX_OK;
			}
		}
		g_pool_cv.wait(lock);
	}
}

int tbtkx_device_release(int device) {
	std::lock_guard<std::mutex> lock(g_pool_mutex);
	TBTKX_REQUIRE(g_pool_ready && device >= 0 &&
	                  device < (int)g_pool_busy.size() && g_pool_busy[device],
	              TBTKX_ERR_ARG, "tbtkx_device_release: device not held");
	g_pool_busy[device] = 0;
	g_pool_cv.notify_one();
	return TBTKX_OK;
}

int tbtkx_create(int device, tbtkx_handle *handle) {
	TBTKX_REQUIRE(handle, TBTKX_ERR_ARG, "tbtkx_create: null handle pointer");
	*handle = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
		set_error("tbtkx_create: no HIP device available");
		return TBTKX_ERR_NODEVICE;
	}
	TBTKX_REQUIRE(device == TBTKX_ANY_DEVICE || (device >= 0 && device < n), TBTKX_ERR_ARG,
	              "tbtkx_create: device " + std::to_string(device) +
	                  " out of range (" + std::to_string(n) + " devices)");
	{
		// Count live handles per device; TBTKX_ANY_DEVICE takes the device
		// with the fewest (the lowest id on a tie), so solvers created by
		// different threads spread over the devices.
		std::lock_guard<std::mutex> lock(g_pool_mutex);
		int rc = pool_init();
		if (rc) return rc;
		if (device == TBTKX_ANY_DEVICE) {
			device = 0;
			for (int d = 1; d < (int)g_pool_handles.size(); d++)
				if (g_pool_handles[d] < g_pool_handles[device]) device = d;
		}
		g_pool_handles[device]++;
	}
	auto *h = new tbtkx_context();
	h->device = device;
	hipError_t e = hipSetDevice(device);
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
	if (e == hipSuccess) e = hipEventCreate(&h->ev_start);
	if (e == hipSuccess) e = hipEventCreate(&h->ev_stop);
	if (e != hipSuccess) {
		set_error(std::string("tbtkx_create: ") + hipGetErrorString(e));
		tbtkx_destroy(h);
		return TBTKX_ERR_HIP;
	}
	*handle = h;
	return TBTKX_OK;
}

int tbtkx_destroy(tbtkx_handle h) {
	if (!h) return TBTKX_OK;
	(void)hipSetDevice(h->device);
	if (h->stream) (void)hipStreamSynchronize(h->stream);
	h->plain_graph.reset();
	h->reset_mm_graphs();
	h->blk_graph.reset();
	for (DevBuf *b : {&h->rowptr, &h->col, &h->v1, &h->v2, &h->v1f, &h->v2f,
	                  &h->v1r, &h->v2r, &h->lat_terms, &h->lat_onsite, &h->lat_count,
	                  &h->lat_part, &h->lat_tmp, &h->rblk, &h->col16, &h->col16_flag,
	                  &h->pcode, &h->pal, &h->palr, &h->pdiag, &h->pdiagr, &h->scol, &h->scode, &h->sv1, &h->sv2, &h->sv1r, &h->sv2r, &h->ppid, &h->ptoff, &h->ptcode, &h->pchunk, &h->vec_a, &h->vec_b, &h->coef, &h->to,
	                  &h->lut, &h->gf_in, &h->gf_out, &h->dfac, &h->dden, &h->blk_x, &h->blk_y,
	                  &h->blk_mu, &h->blk_sites, &h->blk_part, &h->blk_grow, &h->blk_gvec, &h->cplanes,
	                  &h->lplanes, &h->gf_mfma_out, &h->ell_col, &h->ell_v1f,
	                  &h->ell_v2f, &h->ell_v1d, &h->ell_v2d,
	                  &h->mm_ecol, &h->mm_ev1, &h->mm_ev2, &h->mm_desc, &h->mm_tab,
	                  &h->mm_x, &h->mm_y, &h->mm_part, &h->mm_seg, &h->mm_D,
	                  &h->mm_cols})
		b->release();
	if (h->ev_start) (void)hipEventDestroy(h->ev_start);
	if (h->ev_stop) (void)hipEventDestroy(h->ev_stop);
	if (h->stream) (void)hipStreamDestroy(h->stream);
	{
		std::lock_guard<std::mutex> lock(g_pool_mutex);
		if (h->device < (int)g_pool_handles.size() && g_pool_handles[h->device] > 0)
			g_pool_handles[h->device]--;
	}
	delete h;
	return TBTKX_OK;
}

int tbtkx_set_csr(tbtkx_handle h, int64_t N, int64_t Z, const uint32_t *rowptr,
                  const uint32_t *col, const double *val, double scale) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_set_csr: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(N > 0 && N < (int64_t)1 << 31, TBTKX_ERR_ARG,
	              "tbtkx_set_csr: num_rows must be in [1, 2^31)");
	TBTKX_REQUIRE(Z >= 0 && Z < ((int64_t)1 << 32), TBTKX_ERR_ARG,
	              "tbtkx_set_csr: nnz must be in [0, 2^32)");
	TBTKX_REQUIRE(rowptr && (Z == 0 || (col && val)), TBTKX_ERR_ARG,
	              "tbtkx_set_csr: null array");
	TBTKX_REQUIRE(scale > 0, TBTKX_ERR_ARG,
	              "tbtkx_set_csr: the scale factor has to be larger than 0");
	TBTKX_REQUIRE(rowptr[0] == 0 && (int64_t)rowptr[N] == Z, TBTKX_ERR_ARG,
	              "tbtkx_set_csr: rowptr[0] must be 0 and rowptr[N] == nnz");
	for (int64_t r = 0; r < N; r++)
		TBTKX_REQUIRE(rowptr[r] <= rowptr[r + 1], TBTKX_ERR_ARG,
		              "tbtkx_set_csr: rowptr not monotone");
	for (int64_t k = 0; k < Z; k++)
		TBTKX_REQUIRE(col[k] < (uint64_t)N, TBTKX_ERR_ARG,
		              "tbtkx_set_csr: column index out of range at nnz " +
		                  std::to_string(k));
	int rc = set_device(h);
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	h->plain_graph.reset();

	// Device copies carry kCsrPad zero entries (column 0, value 0) after the
	// last non-zero, so the batched kernels may read a fixed number of
	// entries past a row's end without leaving the arrays.
	const int64_t Zp = Z + kCsrPad;
	std::vector<uint32_t> colp(Zp, 0);
	if (Z) memcpy(colp.data(), col, Z * 4);
	std::vector<double> v1(2 * Zp, 0.0), v2(2 * Zp, 0.0);
	scale_complex(val, v1.data(), Z, 1 / scale);
	scale_complex(v1.data(), v2.data(), Z, 2);
	bool real = true;
	for (int64_t k = 0; k < Z && real; k++)
		if (val[2 * k + 1] != 0.0) real = false;
	// The 8-byte copies of the batched paths: FP32 complex values, or for a
	// real H one FP64 value each.
	std::vector<float> v1f(2 * Zp, 0.0f), v2f(2 * Zp, 0.0f);
	for (int64_t k = 0; k < Z; k++) {
		if (real) {
			memcpy(&v1f[2 * k], &v1[2 * k], 8);
			memcpy(&v2f[2 * k], &v2[2 * k], 8);
		} else {
			for (int c = 0; c < 2; c++) {
				v1f[2 * k + c] = (float)v1[2 * k + c];
				v2f[2 * k + c] = (float)v2[2 * k + c];
			}
		}
	}
	std::vector<uint32_t> blocks = make_row_blocks(N, rowptr);

	if ((rc = h->rowptr.reserve((N + 1) * 4))) return rc;
	if ((rc = h->col.reserve(Zp * 4))) return rc;
	if ((rc = h->v1.reserve(Zp * 16))) return rc;
	if ((rc = h->v2.reserve(Zp * 16))) return rc;
	if ((rc = h->v1f.reserve(Zp * 8))) return rc;
	if ((rc = h->v2f.reserve(Zp * 8))) return rc;
	if ((rc = h->rblk.reserve(blocks.size() * 4))) return rc;
	if ((rc = h->vec_a.reserve(N * 16))) return rc;
	if ((rc = h->vec_b.reserve(N * 16))) return rc;
	TBTKX_CHECK_HIP(hipMemcpy(h->rowptr.ptr, rowptr, (N + 1) * 4, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->col.ptr, colp.data(), Zp * 4, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->v1.ptr, v1.data(), Zp * 16, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->v2.ptr, v2.data(), Zp * 16, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->v1f.ptr, v1f.data(), Zp * 8, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->v2f.ptr, v2f.data(), Zp * 8, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->rblk.ptr, blocks.data(), blocks.size() * 4,
	                          hipMemcpyHostToDevice));
	if (real) {
		// Real parts of H/s and 2H/s for the real plain recurrence.
		std::vector<double> r1(Zp, 0.0), r2(Zp, 0.0);
		for (int64_t k = 0; k < Z; k++) {
			r1[k] = v1[2 * k];
			r2[k] = v2[2 * k];
		}
		if ((rc = h->v1r.reserve(Zp * 8))) return rc;
		if ((rc = h->v2r.reserve(Zp * 8))) return rc;
		TBTKX_CHECK_HIP(hipMemcpy(h->v1r.ptr, r1.data(), Zp * 8, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->v2r.ptr, r2.data(), Zp * 8, hipMemcpyHostToDevice));
	}
	// Tile-major ELL copy (see spmm_step_ell): W slots per row in the order
	// of ell_row_slots (diagonal first), tiles of 64/W rows; padding slots
	// point at the row itself with value zero.
	uint32_t maxlen = 0;
	int64_t band = 0;   // max |col - row|: the slowest lattice stride
	std::vector<int64_t> slots;
	for (int64_t r = 0; r < N; r++) {
		uint32_t len = rowptr[r + 1] - rowptr[r];
		bool diag = false;
		for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
			const int64_t d = (int64_t)col[k] - r;
			if (d > band) band = d;
			if (-d > band) band = -d;
			if (d == 0) diag = true;
		}
		if (!diag) len++;
		if (len > maxlen) maxlen = len;
	}
	h->band = band > 0 ? band : N;
	h->ell_W = ell_slots_block(maxlen);
	if (h->ell_W) {
		const int W = h->ell_W, T = 64 / W;
		const int64_t ntiles = (N + T - 1) / T, E = ntiles * T * W;
		std::vector<uint32_t> ec(E);
		std::vector<float> e1f(2 * E, 0.0f), e2f(2 * E, 0.0f);
		std::vector<double> e1d(2 * E, 0.0), e2d(2 * E, 0.0);
		slots.resize(W + 1);
		for (int64_t r = 0; r < ntiles * T; r++) {
			const int n = r < N ? ell_row_slots(rowptr, col, r, slots.data()) : 0;
			for (int u = 0; u < W; u++) {
				const int64_t e = r * W + u;
				ec[e] = (uint32_t)(r < N ? r : 0);
				if (u >= n || slots[u] < 0) continue;
				const int64_t k = slots[u];
				ec[e] = col[k];
				for (int c = 0; c < 2; c++) {
					e1d[2 * e + c] = v1[2 * k + c];
					e2d[2 * e + c] = v2[2 * k + c];
					e1f[2 * e + c] = v1f[2 * k + c];
					e2f[2 * e + c] = v2f[2 * k + c];
				}
			}
		}
		if ((rc = h->ell_col.reserve(E * 4))) return rc;
		if ((rc = h->ell_v1f.reserve(E * 8))) return rc;
		if ((rc = h->ell_v2f.reserve(E * 8))) return rc;
		if ((rc = h->ell_v1d.reserve(E * 16))) return rc;
		if ((rc = h->ell_v2d.reserve(E * 16))) return rc;
		TBTKX_CHECK_HIP(hipMemcpy(h->ell_col.ptr, ec.data(), E * 4, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->ell_v1f.ptr, e1f.data(), E * 8, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->ell_v2f.ptr, e2f.data(), E * 8, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->ell_v1d.ptr, e1d.data(), E * 16, hipMemcpyHostToDevice));
		TBTKX_CHECK_HIP(hipMemcpy(h->ell_v2d.ptr, e2d.data(), E * 16, hipMemcpyHostToDevice));
	}
	h->N = N;
	h->Z = Z;
	h->scale = scale;
	h->csr_gen++;
	h->blk_graph.reset();
	h->h_real = real;
	h->nblk = (int)blocks.size() - 1;
	if ((rc = build_col16(h))) return rc;
	h->forms_pending = 0;
	if ((rc = build_palette(h, N, Z, rowptr, col, v1.data(), v2.data()))) return rc;
	return build_slots(h, rowptr);
}

int tbtkx_last_timing(tbtkx_handle h, double *ms, int64_t *launches) {
	TBTKX_REQUIRE(h && ms && launches, TBTKX_ERR_ARG,
	              "tbtkx_last_timing: null pointer");
	TBTKX_LOCK(h);
	*ms = h->last_ms;
	*launches = h->last_launches;
	return TBTKX_OK;
}

int tbtkx_set_option(tbtkx_handle h, int option, int64_t value) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_set_option: null handle");
	TBTKX_LOCK(h);
	switch (option) {
	case TBTKX_OPT_PLAIN_KERNEL:
		TBTKX_REQUIRE(value >= 0 && value <= 2, TBTKX_ERR_ARG,
		              "tbtkx_set_option: plain kernel must be 0, 1 or 2");
		h->plain_kernel = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_GRAPH:
		h->use_graph = value != 0;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_BLOCK_WIDTH:
		TBTKX_REQUIRE(value == 0 || value == 32 || value == 64 ||
		                  value == 128 || value == 256,
		              TBTKX_ERR_ARG,
		              "tbtkx_set_option: block width must be 0, 32, 64, 128 or 256");
		h->block_width = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_XCD_REMAP:
		h->xcd_remap = value != 0;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_TILE_ROWS:
		TBTKX_REQUIRE(value >= 0 && value <= 32, TBTKX_ERR_ARG,
		              "tbtkx_set_option: tile rows must be in [1, 32] (0 = default)");
		h->tile_rows = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_NT_STREAM:
		h->nt_stream = value != 0;
		return TBTKX_OK;
	case TBTKX_OPT_ELL:
		h->use_ell = value != 0;
		return TBTKX_OK;
	case TBTKX_OPT_AUTOTUNE:
		h->autotune = value != 0;
		for (auto &k : h->mm_tune_key) k = -1;
		return TBTKX_OK;
	case TBTKX_OPT_MFMA_PIPE:
		h->mfma_pipe = value != 0;
		return TBTKX_OK;
	case TBTKX_OPT_STRIP_ROWS:
		TBTKX_REQUIRE(value >= -1 && value <= (1 << 24), TBTKX_ERR_ARG,
		              "tbtkx_set_option: strip rows must be -1, 0 or in [1, 2^24]");
		h->strip_rows = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_GRID_SHARE:
		TBTKX_REQUIRE(value >= 0 && value <= 64, TBTKX_ERR_ARG,
		              "tbtkx_set_option: grid share must be in [0, 64]");
		h->grid_share = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_STRIP_CHUNK:
		TBTKX_REQUIRE(value >= 0 && value <= 4096 && value % 4 == 0, TBTKX_ERR_ARG,
		              "tbtkx_set_option: strip chunk must be a multiple of 4 in [0, 4096]");
		h->strip_chunk = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_ROWS_IN_FLIGHT:
		TBTKX_REQUIRE(value >= 0 && value <= 2, TBTKX_ERR_ARG,
		              "tbtkx_set_option: rows in flight multiplier must be 0, 1 or 2");
		h->rows_in_flight = (int)value;
		return TBTKX_OK;
	case TBTKX_OPT_REAL:
		TBTKX_REQUIRE(value == 0 || value == 1, TBTKX_ERR_ARG,
		              "tbtkx_set_option: real arithmetic option must be 0 or 1");
		h->real_path = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_COL16:
		TBTKX_REQUIRE(value == 0 || value == 1, TBTKX_ERR_ARG,
		              "tbtkx_set_option: 16-bit column option must be 0 or 1");
		h->use_col16 = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_PALETTE:
		TBTKX_REQUIRE(value == 0 || value == 1, TBTKX_ERR_ARG,
		              "tbtkx_set_option: palette option must be 0 or 1");
		h->use_palette = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_PATTERN_BYTE_ROWS:
		TBTKX_REQUIRE(false, TBTKX_ERR_ARG,
		              "tbtkx_set_option: TBTKX_OPT_PATTERN_BYTE_ROWS is read-only");
	case TBTKX_OPT_SLOT_ROWS:
		TBTKX_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4, TBTKX_ERR_ARG,
		              "tbtkx_set_option: slot rows per thread must be 1, 2 or 4 (0 = by form)");
		h->slot_rows = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_ROW_PATTERNS:
		TBTKX_REQUIRE(value >= 0 && value <= 2, TBTKX_ERR_ARG,
		              "tbtkx_set_option: row pattern option must be 0, 1 or 2");
		h->use_patterns = (int)value;
		h->plain_graph.reset();
		return TBTKX_OK;
	case TBTKX_OPT_MULTI:
		TBTKX_REQUIRE(value >= 0 && value <= 2, TBTKX_ERR_ARG,
		              "tbtkx_set_option: multi-vector mode must be 0, 1 or 2");
		h->multi_mode = (int)value;
		return TBTKX_OK;
	default:
		set_error("tbtkx_set_option: unknown option " + std::to_string(option));
		return TBTKX_ERR_ARG;
	}
}

int tbtkx_get_option(tbtkx_handle h, int option, int64_t *value) {
	TBTKX_REQUIRE(h && value, TBTKX_ERR_ARG, "tbtkx_get_option: null pointer");
	TBTKX_LOCK(h);
	switch (option) {
	case TBTKX_OPT_PLAIN_KERNEL: *value = h->last_plain_kernel; return TBTKX_OK;
	case TBTKX_OPT_GRAPH: *value = h->use_graph; return TBTKX_OK;
	case TBTKX_OPT_BLOCK_WIDTH:
		*value = h->block_width ? h->block_width : h->last_block_width;
		return TBTKX_OK;
	case TBTKX_OPT_XCD_REMAP: *value = h->xcd_remap; return TBTKX_OK;
	case TBTKX_OPT_TILE_ROWS:
		*value = h->tile_rows ? h->tile_rows : h->last_tile_groups;
		return TBTKX_OK;
	case TBTKX_OPT_NT_STREAM: *value = h->nt_stream; return TBTKX_OK;
	case TBTKX_OPT_ELL: *value = h->use_ell && h->ell_W > 0; return TBTKX_OK;
	case TBTKX_OPT_AUTOTUNE: *value = h->autotune; return TBTKX_OK;
	case TBTKX_OPT_MFMA_PIPE: *value = h->mfma_pipe; return TBTKX_OK;
	case TBTKX_OPT_MULTI: *value = h->last_multi_blocks ? 2 : 1; return TBTKX_OK;
	case TBTKX_OPT_ROW_PATTERNS: *value = h->last_patterns; return TBTKX_OK;
	case TBTKX_OPT_SLOT_ROWS: *value = h->slot_rows; return TBTKX_OK;
	case TBTKX_OPT_PATTERN_BYTE_ROWS: *value = h->pat_ok ? h->pat_byte_rows : -1; return TBTKX_OK;
	case TBTKX_OPT_STRIP_ROWS:
		*value = h->strip_rows ? h->strip_rows : h->last_strip;
		return TBTKX_OK;
	case TBTKX_OPT_GRID_SHARE:
		*value = h->grid_share ? h->grid_share : h->last_share;
		return TBTKX_OK;
	case TBTKX_OPT_STRIP_CHUNK:
		*value = h->strip_chunk ? h->strip_chunk : 4;
		return TBTKX_OK;
	case TBTKX_OPT_ROWS_IN_FLIGHT:
		*value = h->rows_in_flight ? h->rows_in_flight : h->last_rmul;
		return TBTKX_OK;
	case TBTKX_OPT_REAL:
		*value = h->last_real;
		return TBTKX_OK;
	case TBTKX_OPT_COL16:
		*value = h->last_col16;
		return TBTKX_OK;
	case TBTKX_OPT_PALETTE:
		*value = h->last_palette;
		return TBTKX_OK;
	default:
		set_error("tbtkx_get_option: unknown option " + std::to_string(option));
		return TBTKX_ERR_ARG;
	}
}

}  // extern "C"
