// Uneven batch: many independent models in one launch per step.
//
// K models, each with its own Hamiltonian H_m, scale factor s_m, B_m <= 64
// unit start vectors (LDOS sites) and M_m moments, advance through the
// Chebyshev recurrence together: step n is ONE kernel launch over the row
// tiles of every model that still needs step n. Models are ordered by their
// step count, longest first, so the models active at step n are a prefix of
// that order and the launch covers a prefix of the tile list; the launch
// shape changes only when a model finishes (at most K "phases").
//
// Storage: H in row-major ELL (W_m slots per row: the model's longest row
// rounded up to 4, 5, 6 or 8; scaled by 1/s_m and 2/s_m exactly as the
// single-model path scales its copies), 8 bytes per entry: FP32 complex
// values, or, for a model whose H is real, one FP64 value (FP32 rounding
// of H moves its eigenvalues by ~1e-7 of the band, which thousands of
// moments resolve; a batch mixes both forms, one per model). FP32 complex vectors, row-major per model (N_m rows x
// B_m columns). Rows are summed with the same fused multiply-add order as
// the single-model step, in FP32 (complex H) or FP64 (real H, rounded once
// to FP32 on store), so the recurrence vectors are the ones
// tbtkx_moments_block produces for the same model and sites (it keeps a
// complex H in FP32 and a real H in FP64 the same way). Moments are FP64:
// gather mode reads mu_n[b] = Phi_n[site_b][b] each step; doubling mode fuses
// D0_n = <Phi_n|Phi_n> and D1_n = <Phi_n+1|Phi_n> into the step (FP64 per
// lane, then ordered reductions) and uses mu_2n = 2 D0_n - mu_0,
// mu_2n+1 = 2 D1_n - mu_1 (Weisse et al., Rev. Mod. Phys. 78, 275, eq. 31).
//
// Packed real vectors: when every model's H is real, the real and imaginary
// parts of a complex column evolve independently, so a model's B_m real
// unit vectors take w = ceil(B_m / 2) complex columns (vector 2c in the real
// part of column c, 2c + 1 in the imaginary part); the step is unchanged,
// the gathers and dots read per component. Otherwise w = B_m.
//
// Lane mapping for a model of width w: a lane moves one column (8 B) or,
// for wide models, two adjacent columns (16 B; an odd w is then stored with
// one zero column more, wp = w + 1, which stays zero through the
// recurrence). With lpr lanes per row (w, or wp / 2), a wavefront covers
// Q = floor(64 / lpr) consecutive rows at once, lane l -> row slot
// q = l / lpr, first column (l % lpr) times the columns per lane; lanes
// l >= Q lpr are idle (they mirror a valid row of the same group). A tile is
// G row groups (G Q rows). The phase's tile list is cut into 8 XCD parts of
// equal algorithmic bytes; each XCD's wavefronts take the part's tiles
// round-robin, either in a persistent grid (several tiles each) or one tile
// each in a grid as large as the longest part.
//
// Dot partials: a wavefront sums its lanes' FP64 partials per column when it
// leaves a model (lanes of one column, ascending row slot) and writes them to
// its own slot; slots of one model are contiguous. Two ordered passes reduce
// them (groups of <= 64 slots, then the groups of a model), so the result is
// deterministic for a given launch shape.
//
// Algorithmic HBM bytes per step: sum over active models of
// 12 W_m N_m (ELL column + value) + 24 w N_m (read Phi_n, read Phi_{n-1},
// write Phi_{n+1}); a padding column adds 24 N_m that are not counted.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "common.hpp"

namespace tbtkx {
namespace {

constexpr int kWaves = kThreads / 64;
constexpr int kRpi = 2;          // row groups in flight per lane
constexpr int kSegSlots = 64;    // dot slots summed per group in pass 1
constexpr int kMaxWidth = 64;    // start vectors per model
constexpr int64_t kMaxSlotBytes = (int64_t)1 << 30;   // dot slots of a one-tile-per-wavefront plan

struct MDesc {
	int64_t N;       // rows
	int64_t e0;      // first ELL entry of the model (its rows have W slots)
	int64_t vec0;    // first element of the model's vectors in X / Y
	int64_t dof;     // first element of the model's D block (doubling)
	int64_t D;       // layer length of the strip traversal (the model's band)
	int w;           // columns (start vectors)
	int Q;           // rows per row group: 64 / lanes per row
	int S;           // strip rows
	int A;           // layers, ceil(N / D)
	int tps;         // tiles per strip segment
	int wp;          // stored columns: w, or w rounded up to even for pairs
	int pairs;       // 1: a lane moves two columns (16 B), else one
	int W;           // ELL slots per row of this model (4, 5, 6 or 8)
	int realh;       // the model's H is real: FP64 ELL values, FP64 row sums
};

struct GCol {
	int64_t xoff;    // X element of the vector's site in its column
	int64_t out;     // first output moment of the vector
	int M;           // moments of the vector's model
	int comp;        // 0: complex column; packed real vectors: 1 = the
	                 // column's real part, 2 = its imaginary part
};

typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 load_nt(const float2 *p) {
	const f32x2 q = __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(p));
	return make_float2(q.x, q.y);
}
__device__ __forceinline__ void store_nt(float2 *p, float x, float y) {
	f32x2 q = {x, y};
	__builtin_nontemporal_store(q, reinterpret_cast<f32x2 *>(p));
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// CPL consecutive complex columns of one row at byte address p (8 CPL B,
// aligned to that): plain and non-temporal loads, non-temporal stores.
template <int CPL>
__device__ __forceinline__ void lane_load(const char *p, float2 (&o)[CPL]) {
	if constexpr (CPL == 1) {
		o[0] = *reinterpret_cast<const float2 *>(p);
	} else {
		const float4 v = *reinterpret_cast<const float4 *>(p);
		o[0] = make_float2(v.x, v.y);
		o[1] = make_float2(v.z, v.w);
	}
}
template <int CPL>
__device__ __forceinline__ void lane_load_nt(const char *p, float2 (&o)[CPL]) {
	if constexpr (CPL == 1) {
		o[0] = load_nt(reinterpret_cast<const float2 *>(p));
	} else {
		const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
		o[0] = make_float2(v.x, v.y);
		o[1] = make_float2(v.z, v.w);
	}
}
template <int CPL>
__device__ __forceinline__ void lane_store_nt(char *p, const float2 (&v)[CPL]) {
	if constexpr (CPL == 1) {
		store_nt(reinterpret_cast<float2 *>(p), v[0].x, v[0].y);
	} else {
		f32x4 q = {v[0].x, v[0].y, v[1].x, v[1].y};
		__builtin_nontemporal_store(q, reinterpret_cast<f32x4 *>(p));
	}
}

inline unsigned grid_for(int64_t n) {
	return (unsigned)((n + kThreads - 1) / kThreads);
}

__global__ __launch_bounds__(kThreads) void multi_fill(float2 *__restrict__ X,
                                                         int64_t n) {
	for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * kThreads)
		X[i] = make_float2(0.0f, 0.0f);
}

__global__ __launch_bounds__(kThreads) void multi_units(float2 *__restrict__ X,
                                                          const GCol *__restrict__ cols,
                                                          int ncols) {
	const int g = blockIdx.x * kThreads + threadIdx.x;
	if (g >= ncols) return;
	const GCol c = cols[g];
	// Two packed vectors of one column may start at the same site: each
	// writes only its own component.
	float *e = reinterpret_cast<float *>(X + c.xoff);
	e[c.comp == 2 ? 1 : 0] = 1.0f;
}

// Gather mode: mu_n of every column whose model has more than n moments.
__global__ __launch_bounds__(kThreads) void multi_gather(
    const GCol *__restrict__ cols, int ncols, const float2 *__restrict__ X,
    double2 *__restrict__ mu, int n) {
	const int g = blockIdx.x * kThreads + threadIdx.x;
	if (g >= ncols) return;
	const GCol c = cols[g];
	if (n < c.M) {
		const float2 x = X[c.xoff];
		if (c.comp == 0) mu[c.out + n] = make_double2((double)x.x, (double)x.y);
		else mu[c.out + n] = make_double2((double)(c.comp == 2 ? x.y : x.x), 0.0);
	}
}

// Doubling dots of one element, FP64: d0 += |x|^2 and d1 += conj(y) x for a
// complex column; per component for a column of two packed real vectors,
// d0 += (x.re^2, x.im^2) and d1 += (y.re x.re, y.im x.im).
__device__ __forceinline__ void dot_acc(double xx, double xy, double yx, double yy, int packed,
                                        double &d0x, double &d0y, double &d1x, double &d1y) {
	if (packed) {
		d0x = fma(xx, xx, d0x);
		d0y = fma(xy, xy, d0y);
		d1x = fma(yx, xx, d1x);
		d1y = fma(yy, xy, d1y);
	} else {
		d0x = fma(xx, xx, fma(xy, xy, d0x));
		d1x = fma(yx, xx, fma(yy, xy, d1x));
		d1y = fma(yx, xy, fma(-yy, xx, d1y));
	}
}

template <bool REALH>
__device__ __forceinline__ void mac(float2 a, float2 x, float &ax, float &ay) {
	if (REALH) {
		ax = fmaf(a.x, x.x, ax);
		ay = fmaf(a.x, x.y, ay);
	} else {
		ax = fmaf(a.x, x.x, fmaf(-a.y, x.y, ax));
		ay = fmaf(a.x, x.y, fmaf(a.y, x.x, ay));
	}
}
// Real H: one FP64 value per ELL entry (the 8 bytes of an FP32 complex
// value), the row summed in FP64 and rounded once to FP32 on store.
template <bool REALH>
__device__ __forceinline__ void mac(double a, float2 x, double &ax, double &ay) {
	ax = fma(a, (double)x.x, ax);
	ay = fma(a, (double)x.y, ay);
}

// ELL value and row accumulator types: FP32 complex values and FP32 sums,
// or (real H) FP64 values and FP64 sums.
template <bool REALH> struct EllT {
	typedef float2 V;
	typedef float A;
};
template <> struct EllT<true> {
	typedef double V;
	typedef double A;
};

// Tile lt of a model (strip k, layer a, j-th tile of the segment): its first
// row, and in rows its row count (0 for a tile past the end of the model).
__device__ __forceinline__ int64_t tile_first_row(const MDesc &d, uint32_t lt, int64_t tr,
                                                  int64_t &rows) {
	const uint32_t seg = lt / (uint32_t)d.tps;
	const uint32_t j = lt - seg * (uint32_t)d.tps;
	const uint32_t kk = seg / (uint32_t)d.A;
	const int64_t a = seg - kk * (uint32_t)d.A;
	const int64_t c0 = (int64_t)kk * d.S + (int64_t)j * tr;
	rows = d.S - (int64_t)j * tr < tr ? d.S - (int64_t)j * tr : tr;
	if (c0 + rows > d.D) rows = d.D - c0;
	const int64_t row0 = a * d.D + c0;
	if (row0 + rows > d.N) rows = d.N - row0;
	if (rows <= 0) {
		rows = 0;
		return 0;
	}
	return row0;
}

// Row groups with the ELL entries of the next iteration loaded one iteration
// ahead. A lane moves CPL consecutive columns of its row (CPL = 2: 16-byte
// loads and stores). An iteration is RPI row groups of one tile. Their
// entries come either from NW = RPI / GPW shared windows (Q W <= 64;
// window k: the GPW Q W consecutive entries of groups k GPW .. k GPW + GPW
// - 1, lane l holding entry l, read by the other lanes through shuffles) or,
// with GPW = 0 (narrow models, Q W > 64), from each lane's own loads of its
// rows' W entries. Per iteration, in issue order: the Phi_{n-1} rows and
// the gathers (columns taken from the window registers), the stores of the
// previous iteration's rows (kept in registers until now), the next
// iteration's windows, then the wait for this iteration's loads and the
// multiply-adds. Vector memory operations retire in issue order on one
// counter, so that wait never covers the stores or the prefetch and an
// iteration costs one memory round trip instead of two.
//
// The stores and the prefetch are issued by every lane on every iteration
// (no branch around them, so the compiler's wait counts are exact): the
// last iteration loads its own window again, iterations and tiles without
// rows are skipped, and a lane whose row slot lies past the tile's last row
// (or an idle lane, q >= Q) takes that last row, which a live lane of the
// same iteration computes from the same entries and the same Phi_{n-1}
// value: it stores the same value to the same address.
template <int W, int GPW, int RPI, int CPL, bool FIRST, bool DOTS, bool REALH>
__device__ __forceinline__ void multi_rows_pipe(
    const MDesc &d, int &t, int tend, int nwx, int toff_m, int G, int Q, int qc, bool act,
    int b, int64_t tr, int lane, const uint32_t *__restrict__ ecol,
    const float2 *__restrict__ ev, const float2 *__restrict__ X, float2 *__restrict__ Y,
    int packed, double (&dd)[2][4]) {
	// GPW = 0: each lane loads its own rows' W entries (RPI W registers).
	constexpr bool PERLANE = GPW == 0;
	constexpr int NW = PERLANE ? RPI * W : RPI / (PERLANE ? 1 : GPW);
	typedef typename EllT<REALH>::V EV;
	typedef typename EllT<REALH>::A AC;
	const EV *__restrict__ evv = reinterpret_cast<const EV *>(ev);
	const uint32_t wp = (uint32_t)d.wp;
	// The model's blocks of X and Y (wave-uniform bases) and 32-bit byte
	// offsets into them (N_m wp 8 B < 4 GiB, checked on the host).
	const char *Xm = reinterpret_cast<const char *>(X + d.vec0);
	char *Ym = reinterpret_cast<char *>(Y + d.vec0);
	const int wrow_l = lane / W, wslot = lane - wrow_l * W;
	// The wave's first tile with rows.
	int64_t rows = 0, row0 = 0;
	for (; t < tend; t += nwx) {
		row0 = tile_first_row(d, (uint32_t)(t - toff_m), tr, rows);
		if (rows > 0) break;
	}
	if (t >= tend) return;
	int gg = 0;
	uint32_t wc[NW];
	EV wv[NW];
	// Windows of the iteration at groups g0.. of the tile at r0 (nr rows).
	auto load_windows = [&](int64_t r0, int64_t nr, int g0) {
		const int64_t glast = r0 + nr - 1;
		if constexpr (PERLANE) {
#pragma unroll
			for (int r = 0; r < RPI; r++) {
				const int64_t o = (int64_t)(g0 + r) * Q + qc;
				const int64_t e = d.e0 + (r0 + (o < nr ? o : nr - 1)) * W;
#pragma unroll
				for (int u = 0; u < W; u++) {
					wc[r * W + u] = ecol[e + u];
					wv[r * W + u] = evv[e + u];
				}
			}
		} else {
#pragma unroll
			for (int k = 0; k < NW; k++) {
				int64_t wrow = r0 + (int64_t)(g0 + k * GPW) * Q + wrow_l;
				if (wrow > glast) wrow = glast;
				const int64_t we = d.e0 + wrow * W + wslot;
				wc[k] = ecol[we];
				wv[k] = evv[we];
			}
		}
	};
	load_windows(row0, rows, 0);
	float2 ps[RPI][CPL];   // the previous iteration's results, stored late
	uint32_t po[RPI];
	// One iteration; STORE: the previous iteration's rows are stored (every
	// iteration but the first). Returns whether another iteration follows.
	auto iteration = [&](auto store_prev) -> bool {
		constexpr bool STORE = decltype(store_prev)::value;
		uint32_t off[RPI];
		bool live[RPI];
		float2 yo[RPI][CPL], xv[RPI][W][CPL];
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			const int64_t o = (int64_t)(gg + r) * Q + qc;
			live[r] = act && o < rows;
			const int64_t row = row0 + (o < rows ? o : rows - 1);
			off[r] = ((uint32_t)row * wp + (uint32_t)b) * 8u;
			if (!FIRST) lane_load_nt<CPL>(Ym + off[r], yo[r]);
		}
#pragma unroll
		for (int r = 0; r < RPI; r++) {
#pragma unroll
			for (int u = 0; u < W; u++) {
				uint32_t c;
				if constexpr (PERLANE)
					c = wc[r * W + u];
				else
					c = (uint32_t)__shfl((int)wc[r / GPW], ((r % GPW) * Q + qc) * W + u);
				lane_load<CPL>(Xm + (c * wp + (uint32_t)b) * 8u, xv[r][u]);
			}
		}
		// This iteration's values; the window registers take the next one.
		// (The empty asm keeps the copy, so the prefetch below can load into
		// the window registers themselves and the loop carries no copies of
		// pending loads.)
		EV vc[NW];
#pragma unroll
		for (int k = 0; k < NW; k++) {
			vc[k] = wv[k];
			if constexpr (REALH) asm volatile("" : "+v"(vc[k]));
			else asm volatile("" : "+v"(vc[k].x), "+v"(vc[k].y));
		}
		if constexpr (STORE) {
#pragma unroll
			for (int r = 0; r < RPI; r++) lane_store_nt<CPL>(Ym + po[r], ps[r]);
		}
		// Next iteration: the next RPI groups of this tile while they hold
		// rows, else the wave's next tile with rows.
		int tn = t, gn = gg + RPI;
		int64_t row0n = row0, rowsn = rows;
		if (gn >= G || (int64_t)gn * Q >= rows) {
			gn = 0;
			for (;;) {
				tn += nwx;
				if (tn >= tend) break;
				row0n = tile_first_row(d, (uint32_t)(tn - toff_m), tr, rowsn);
				if (rowsn > 0) break;
			}
		}
		const bool more = tn < tend;
		if (more) load_windows(row0n, rowsn, gn);
		else load_windows(row0, rows, gg);
#pragma unroll
		for (int r = 0; r < RPI; r++) {
			AC ax[CPL], ay[CPL];
#pragma unroll
			for (int j = 0; j < CPL; j++) {
				ax[j] = FIRST ? (AC)0 : -(AC)yo[r][j].x;
				ay[j] = FIRST ? (AC)0 : -(AC)yo[r][j].y;
			}
#pragma unroll
			for (int u = 0; u < W; u++) {
				EV av;
				if constexpr (PERLANE) {
					av = vc[r * W + u];
				} else {
					const int src = ((r % GPW) * Q + qc) * W + u;
					if constexpr (REALH)
						av = __shfl(vc[r / GPW], src);
					else
						av = make_float2(__shfl(vc[r / GPW].x, src),
						                 __shfl(vc[r / GPW].y, src));
				}
#pragma unroll
				for (int j = 0; j < CPL; j++) mac<REALH>(av, xv[r][u][j], ax[j], ay[j]);
			}
#pragma unroll
			for (int j = 0; j < CPL; j++) ps[r][j] = make_float2((float)ax[j], (float)ay[j]);
			po[r] = off[r];
			// Slot 0 is the row itself (diagonal first); the dots use the
			// stored (FP32) values, as a gather of the vectors would.
			if (DOTS && live[r]) {
#pragma unroll
				for (int j = 0; j < CPL; j++)
					dot_acc(xv[r][0][j].x, xv[r][0][j].y, ps[r][j].x, ps[r][j].y, packed,
					        dd[j][0], dd[j][1], dd[j][2], dd[j][3]);
			}
		}
		t = tn;
		gg = gn;
		row0 = row0n;
		rows = rowsn;
		return more;
	};
	if (iteration(std::false_type{}))
		while (iteration(std::true_type{})) {
		}
#pragma unroll
	for (int r = 0; r < RPI; r++) lane_store_nt<CPL>(Ym + po[r], ps[r]);
}

// One model's tiles t, t + nwx, ... below tend with its ELL width W: the
// path for its lane mode and row-group size.
template <int W, bool FIRST, bool DOTS, bool REALH>
__device__ __forceinline__ void model_rows_w(
    const MDesc &d, int &t, int tend, int nwx, int toff_m, int G, int Q, int qc, bool act,
    int b, int64_t tr, int lane, const uint32_t *__restrict__ ecol,
    const float2 *__restrict__ ev, const float2 *__restrict__ X, float2 *__restrict__ Y,
    int packed, double (&dd)[2][4]) {
	// A row group's entries fit one window (Q W <= 64, guaranteed on the
	// host for pair models): the pipelined path, with as many groups per
	// window as fit; narrow models load their rows' entries per lane.
#define TBTKX_PIPE(GPW, RPI, CPL)                                                        \
	multi_rows_pipe<W, GPW, RPI, CPL, FIRST, DOTS, REALH>(d, t, tend, nwx, toff_m, G, Q, qc, \
	                                                      act, b, tr, lane, ecol, ev, X, Y,  \
	                                                      packed, dd)
	if (d.pairs) {
		if (Q * W <= 32) TBTKX_PIPE(2, 2, 2);
		else TBTKX_PIPE(1, 2, 2);
	} else if (Q * W <= 16) {
		TBTKX_PIPE(4, 4, 1);
	} else if (Q * W <= 32) {
		TBTKX_PIPE(2, 4, 1);
	} else if (Q * W <= 64) {
		TBTKX_PIPE(1, 4, 1);
	} else {
		TBTKX_PIPE(0, kRpi, 1);
	}
#undef TBTKX_PIPE
}

// Dispatch on the model's ELL width (4, 5, 6 or 8, at most the batch's
// widest, WMAX, which fixes the kernel's registers).
template <int WMAX, bool FIRST, bool DOTS, bool REALH>
__device__ __forceinline__ void model_rows(
    const MDesc &d, int &t, int tend, int nwx, int toff_m, int G, int Q, int qc, bool act,
    int b, int64_t tr, int lane, const uint32_t *__restrict__ ecol,
    const float2 *__restrict__ ev, const float2 *__restrict__ X, float2 *__restrict__ Y,
    int packed, double (&dd)[2][4]) {
	// REALH: every model of the batch is real. Otherwise each model takes
	// the path of its own H (a real model's entries hold FP64 values).
	if constexpr (!REALH) {
		if (d.realh) {
			model_rows<WMAX, FIRST, DOTS, true>(d, t, tend, nwx, toff_m, G, Q, qc, act, b, tr,
			                                    lane, ecol, ev, X, Y, packed, dd);
			return;
		}
	}
#define TBTKX_MR(WW)                                                                 \
	model_rows_w<WW, FIRST, DOTS, REALH>(d, t, tend, nwx, toff_m, G, Q, qc, act, b, tr, lane, \
	                                     ecol, ev, X, Y, packed, dd)
	if (d.W == WMAX) {
		TBTKX_MR(WMAX);
	} else if constexpr (WMAX > 4) {
		if (d.W == 4) {
			TBTKX_MR(4);
		} else if constexpr (WMAX > 5) {
			if (d.W == 5) {
				TBTKX_MR(5);
			} else if constexpr (WMAX > 6) {
				TBTKX_MR(6);
			}
		}
	}
#undef TBTKX_MR
}

// One recurrence step of every active model. Each model is traversed in
// L2 strips (layers of D rows, strips of S rows; a tile is G row groups of
// Q rows inside one strip segment), the models' tile lists are concatenated
// (toff[k]: first tile of sorted model k) and the phase's list is cut into
// 8 byte-balanced parts, one per XCD (xpart). Inside a part the XCD's
// wavefronts take consecutive tiles round-robin, so the rows in flight on an
// XCD form a narrow front. The grid is persistent (wavefront wx takes tiles
// wx, wx + nwx, ...) or has a wavefront per tile (nwx >= the longest part):
// the hardware dispatches workgroups x, x + 8, ... to XCD x in order as
// earlier ones finish, so the front cannot drift apart. Dot partials: a
// wavefront flushes its sums when
// it leaves a model, into slot pbase[k*8+x] + its rank among the
// wavefronts that touch the model's piece of part x (pstart: the piece's
// first tile).
// Batches of at most 5 ELL slots per row are held to 4 wavefronts per SIMD
// (128 VGPRs): the doubling variants then keep 4-6 values in scratch and
// still finish the 64-model batch 1.3% sooner than at 3 wavefronts.
template <int WMAX, bool FIRST, bool DOTS, bool REALH>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(WMAX <= 5 ? 4 : 1))) void multi_step(
    const MDesc *__restrict__ md, const int *__restrict__ toff, int kact,
    const int *__restrict__ xpart, const int *__restrict__ pbase,
    const int *__restrict__ pstart, int G, const uint32_t *__restrict__ ecol,
    const float2 *__restrict__ ev, const float2 *__restrict__ X,
    float2 *__restrict__ Y, double2 *__restrict__ part, int packed) {
	const int lane = threadIdx.x & 63;
	const int x = blockIdx.x & 7;
	const int nwx = (gridDim.x >> 3) * kWaves;                  // wavefronts per XCD
	// Wave-uniform (readfirstlane tells the compiler): the tile, the model
	// and its descriptor then live in scalar registers.
	const int wx = __builtin_amdgcn_readfirstlane((blockIdx.x >> 3) * kWaves + (threadIdx.x >> 6));
	const int plo = xpart[x], phi = xpart[x + 1];
	int t = plo + wx;
	if (t >= phi) return;
	// Model of tile t: the last k with toff[k] <= t (every model has a tile).
	int lo = 0, hi = kact;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (toff[mid] <= t) lo = mid;
		else hi = mid;
	}
	int m = lo;
	for (;;) {
		const MDesc d = md[m];
		// Lane l: row slot q = l / lpr of the group, columns b .. b + cpl - 1.
		const int Q = d.Q, cpl = d.pairs ? 2 : 1;
		const int lpr = d.wp / cpl;
		const int q = lane / lpr;
		const int j = lane - q * lpr;
		const int b = j * cpl;
		const bool act = q < Q;
		const int qc = act ? q : Q - 1;
		const int64_t tr = (int64_t)G * Q;   // rows per tile
		const int tend = phi < toff[m + 1] ? phi : toff[m + 1];
		double dd[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
		model_rows<WMAX, FIRST, DOTS, REALH>(d, t, tend, nwx, toff[m], G, Q, qc, act, b, tr, lane,
		                                     ecol, ev, X, Y, packed, dd);
		if (DOTS) {
			// Column sums over the row slots q = 0 .. Q-1, ascending; lane
			// j < lpr writes columns j cpl .. j cpl + cpl - 1.
			const int pi = m * 8 + x;
			const int s0w = (pstart[pi] - plo) % nwx;
			const int slot = pbase[pi] + (wx - s0w + nwx) % nwx;
			for (int c = 0; c < cpl; c++) {
				double s[4] = {0.0, 0.0, 0.0, 0.0};
				for (int k = 0; k < Q; k++) {
					const int src = j + k * lpr;
#pragma unroll
					for (int i = 0; i < 4; i++) s[i] += __shfl(c ? dd[1][i] : dd[0][i], src);
				}
				if (lane < lpr) {
					part[((int64_t)slot * 2 + 0) * kMaxWidth + b + c] = make_double2(s[0], s[1]);
					part[((int64_t)slot * 2 + 1) * kMaxWidth + b + c] = make_double2(s[2], s[3]);
				}
			}
		}
		if (t >= phi) break;
		while (toff[m + 1] <= t) m++;
	}
}

// Pass 1: segment s = (model, slot_lo, slot_hi) of <= 64 slots; one
// wavefront per (segment, which), lane = column (the model's w columns).
__global__ __launch_bounds__(kThreads) void multi_reduce_slots(
    const MDesc *__restrict__ md, const int *__restrict__ seg, int nseg,
    const double2 *__restrict__ part, double2 *__restrict__ segsum) {
	const int lane = threadIdx.x & 63;
	const int gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
	if (gw >= 2 * nseg) return;
	const int s = gw >> 1, which = gw & 1;
	if (lane >= md[seg[3 * s]].w) return;
	const int lo = seg[3 * s + 1], hi = seg[3 * s + 2];
	// Slots in ascending order; eight loads are issued before their adds.
	double2 acc = make_double2(0.0, 0.0);
	for (int k0 = lo; k0 < hi; k0 += 8) {
		double2 v[8];
#pragma unroll
		for (int i = 0; i < 8; i++)
			if (k0 + i < hi) v[i] = part[((int64_t)(k0 + i) * 2 + which) * kMaxWidth + lane];
#pragma unroll
		for (int i = 0; i < 8; i++) {
			if (k0 + i < hi) {
				acc.x += v[i].x;
				acc.y += v[i].y;
			}
		}
	}
	segsum[((int64_t)s * 2 + which) * kMaxWidth + lane] = acc;
}

// Pass 2: one wavefront per (active model, which): the model's segments in
// order, into D[dof + (n * 2 + which) * w + b].
__global__ __launch_bounds__(kThreads) void multi_reduce_models(
    const MDesc *__restrict__ md, const int *__restrict__ mseg, int kact,
    const double2 *__restrict__ segsum, double2 *__restrict__ D, int n) {
	const int lane = threadIdx.x & 63;
	const int gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
	if (gw >= 2 * kact) return;
	const int k = gw >> 1, which = gw & 1;
	const MDesc d = md[k];
	if (lane >= d.w) return;
	// Segments in ascending order; eight loads are issued before their adds.
	double2 acc = make_double2(0.0, 0.0);
	const int lo = mseg[k], hi = mseg[k + 1];
	for (int s0 = lo; s0 < hi; s0 += 8) {
		double2 v[8];
#pragma unroll
		for (int i = 0; i < 8; i++)
			if (s0 + i < hi) v[i] = segsum[((int64_t)(s0 + i) * 2 + which) * kMaxWidth + lane];
#pragma unroll
		for (int i = 0; i < 8; i++) {
			if (s0 + i < hi) {
				acc.x += v[i].x;
				acc.y += v[i].y;
			}
		}
	}
	D[d.dof + ((int64_t)n * 2 + which) * d.w + lane] = acc;
}

// ------------------------------------------------------------------ host --

struct Phase {
	int kact;          // active models (a prefix of the sorted order)
	int nwg;           // workgroups of the phase's step launches
	int n_lo, n_hi;    // steps of the phase
	int64_t toff, xpart, pbase, pstart, mseg, seg;   // offsets into the table buffer
	int nseg, nslots;
};

struct Plan {
	std::vector<int> order;      // sorted position -> model
	std::vector<MDesc> desc;     // in sorted order
	std::vector<int> S;          // steps, sorted order
	std::vector<int64_t> tiles;  // tiles per model, sorted order
	std::vector<Phase> phases;
	std::vector<int> tab;        // all phase tables
	int nwg = 0, G = 8, Smax = 0, max_slots = 0, max_seg = 0;
	int64_t vec_elems = 0, d_elems = 0;
};

// Two columns per lane (16-byte accesses, half the memory instructions per
// byte) for models whose rows are wide enough that the padding column of an
// odd width costs little (w even, or w >= 13) and whose row group of
// Q = 64 / (lanes per row) rows still fits one ELL window (Q W <= 64).
// Models wider than 32 columns always qualify.
bool use_pairs(int w, int W) {
	const int wp = w + (w & 1);
	const int Q = kMaxWidth / (wp / 2);
	if (w < 2 || Q * W > 64) return false;
	return w > 32 || w % 2 == 0 || w >= 13;
}

// Tables of one phase: tile offsets, the byte-balanced cut of the tile list
// into 8 XCD parts, the pieces (model x part) with the base of their dot
// slots (model-major, so a model's slots are contiguous) and the slot groups
// of each model for the ordered reduction. nwx: wavefronts per XCD of the
// persistent grid (p.nwg workgroups), or 0 for one tile per wavefront: the
// grid then has as many wavefronts per XCD as the longest part has tiles,
// and the hardware's in-order dispatch walks each part front to back.
void plan_phase(const Plan &p, int kact, int nwx, Phase &ph, std::vector<int> &tab) {
	std::vector<int> toff(kact + 1, 0);
	std::vector<double> cost(kact), cum(kact + 1, 0.0);
	for (int k = 0; k < kact; k++) {
		toff[k + 1] = toff[k] + (int)p.tiles[k];
		cost[k] = (double)p.desc[k].N * (12.0 * p.desc[k].W + 24.0 * p.desc[k].wp);
		cum[k + 1] = cum[k] + cost[k];
	}
	const int T = toff[kact];
	// XCD parts: cut where the cumulative cost (uniform per tile inside a
	// model) crosses i/8 of the total.
	std::vector<int> xpart(9, T);
	xpart[0] = 0;
	int k = 0;
	for (int i = 1; i < 8; i++) {
		const double target = cum[kact] * i / 8;
		while (k < kact - 1 && cum[k + 1] <= target) k++;
		const double per_tile = cost[k] / (double)(toff[k + 1] - toff[k]);
		int tcut = toff[k] + (int)((target - cum[k]) / per_tile);
		tcut = std::min(std::max(tcut, toff[k]), toff[k + 1]);
		xpart[i] = std::max(tcut, xpart[i - 1]);
	}
	if (nwx > 0) {
		ph.nwg = p.nwg;
	} else {
		int maxpart = 1;
		for (int x = 0; x < 8; x++) maxpart = std::max(maxpart, xpart[x + 1] - xpart[x]);
		nwx = (maxpart + kWaves - 1) / kWaves * kWaves;
		ph.nwg = 8 * (nwx / kWaves);
	}
	// Pieces and slots.
	std::vector<int> pbase(kact * 8, -1), pstart(kact * 8, 0), mslot(kact + 1, 0);
	int slot = 0;
	for (int j = 0; j < kact; j++) {
		mslot[j] = slot;
		for (int x = 0; x < 8; x++) {
			const int a = std::max(toff[j], xpart[x]), e = std::min(toff[j + 1], xpart[x + 1]);
			if (a >= e) continue;
			pbase[j * 8 + x] = slot;
			pstart[j * 8 + x] = a;
			slot += std::min(e - a, nwx);
		}
	}
	mslot[kact] = slot;
	std::vector<int> mseg(kact + 1, 0), seg;
	for (int j = 0; j < kact; j++) {
		mseg[j] = (int)(seg.size() / 3);
		for (int s = mslot[j]; s < mslot[j + 1]; s += kSegSlots) {
			seg.push_back(j);
			seg.push_back(s);
			seg.push_back(std::min(s + kSegSlots, mslot[j + 1]));
		}
	}
	mseg[kact] = (int)(seg.size() / 3);
	ph.kact = kact;
	ph.nslots = slot;
	ph.nseg = (int)(seg.size() / 3);
	auto put = [&tab](const std::vector<int> &v) {
		const int64_t off = (int64_t)tab.size();
		tab.insert(tab.end(), v.begin(), v.end());
		return off;
	};
	ph.toff = put(toff);
	ph.xpart = put(xpart);
	ph.pbase = put(pbase);
	ph.pstart = put(pstart);
	ph.mseg = put(mseg);
	ph.seg = put(seg.empty() ? std::vector<int>(3, 0) : seg);
}

template <int W, bool DOTS, bool REALH>
const void *step_fn() {
	return (const void *)multi_step<W, false, DOTS, REALH>;
}

int grid_size(tbtkx_context *h, bool dots, int share, int *nwg) {
	const void *fn;
	const bool R = h->mm_real;
#define TBTKX_FN(WW) \
	fn = dots ? (R ? step_fn<WW, true, true>() : step_fn<WW, true, false>()) \
	          : (R ? step_fn<WW, false, true>() : step_fn<WW, false, false>())
	if (h->mm_W == 4) TBTKX_FN(4);
	else if (h->mm_W == 5) TBTKX_FN(5);
	else if (h->mm_W == 6) TBTKX_FN(6);
	else TBTKX_FN(8);
#undef TBTKX_FN
	int per_cu = 0, cus = 0;
	TBTKX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
	TBTKX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
	                                      h->device));
	int64_t n = (int64_t)(per_cu > 0 ? per_cu : 1) * cus / (share > 0 ? share : 1);
	n = (n + 7) / 8 * 8;
	if (n < 8) n = 8;
	*nwg = (int)n;
	return TBTKX_OK;
}

// Launch shape and tables for the given widths and step counts.
int make_plan(tbtkx_context *h, const std::vector<int> &width, const std::vector<int> &steps,
              bool dots, int Gmax, int share, Plan &p) {
	const int K = h->mm_K;
	p.order.resize(K);
	std::iota(p.order.begin(), p.order.end(), 0);
	std::stable_sort(p.order.begin(), p.order.end(),
	                 [&steps](int x, int y) { return steps[x] > steps[y]; });
	p.desc.resize(K);
	p.S.resize(K);
	int64_t vec = 0, dof = 0;
	p.G = Gmax;
	p.tiles.resize(K);
	int64_t tiles_all = 0;
	for (int k = 0; k < K; k++) {
		const int m = p.order[k];
		MDesc &d = p.desc[k];
		d.N = h->mm_N[m];
		d.e0 = h->mm_e0[m];
		d.W = h->mm_Wm[m];
		d.realh = h->mm_realm[m];
		d.w = width[m];
		d.pairs = use_pairs(d.w, d.W) ? 1 : 0;
		d.wp = d.pairs ? d.w + (d.w & 1) : d.w;
		d.Q = kMaxWidth / (d.pairs ? d.wp / 2 : d.wp);
		d.vec0 = vec;
		d.dof = dof;
		// Strips: the rows whose Phi rows are reused one layer later span
		// 2 S rows; about 512 KB of them per strip, at least one tile.
		const int64_t tr = (int64_t)p.G * d.Q;
		d.D = h->mm_band[m];
		int64_t S = (int64_t)32768 / d.wp;
		if (S < tr) S = tr;
		if (2 * S > d.D) S = d.D;
		d.S = (int)S;
		d.A = (int)((d.N + d.D - 1) / d.D);
		d.tps = (int)((S + tr - 1) / tr);
		p.tiles[k] = ((d.D + S - 1) / S) * d.A * d.tps;
		tiles_all += p.tiles[k];
		// Every model's block starts on 16 bytes (two-column lanes).
		vec += d.N * d.wp;
		vec += vec & 1;
		dof += (int64_t)steps[m] * 2 * d.w;
		p.S[k] = steps[m];
	}
	TBTKX_REQUIRE(tiles_all < ((int64_t)1 << 31), TBTKX_ERR_ARG,
	              "tbtkx_multi_moments: batch too large for one launch");
	p.vec_elems = vec;
	p.d_elems = dof;
	p.Smax = p.S[0];
	// share > 0: persistent grid of occupancy x CUs / share workgroups;
	// share == 0: one tile per wavefront (per-phase grids).
	p.nwg = 0;
	if (share > 0) {
		int rc = grid_size(h, dots, share, &p.nwg);
		if (rc) return rc;
	}
	const int nwx = p.nwg / 8 * kWaves;
	p.phases.clear();
	p.tab.clear();
	for (int n = 0; n < p.Smax;) {
		int kact = 0;
		while (kact < K && p.S[kact] > n) kact++;
		const int n_hi = p.S[kact - 1];
		Phase ph;
		ph.n_lo = n;
		ph.n_hi = n_hi;
		plan_phase(p, kact, nwx, ph, p.tab);
		p.max_slots = std::max(p.max_slots, ph.nslots);
		p.max_seg = std::max(p.max_seg, ph.nseg);
		p.phases.push_back(ph);
		n = n_hi;
	}
	return TBTKX_OK;
}

template <int W, bool REALH>
void launch_step_w(tbtkx_context *h, const Plan &p, const Phase &ph, bool first,
                   bool dots, const float2 *X, float2 *Y) {
	const int *tab = h->mm_tab.as<int>();
	const MDesc *md = h->mm_desc.as<MDesc>();
	const float2 *ev = first ? h->mm_ev1.as<float2>() : h->mm_ev2.as<float2>();
	double2 *part = h->mm_part.as<double2>();
#define TBTKX_MS(F, D)                                                          \
	multi_step<W, F, D, REALH><<<ph.nwg, kThreads, 0, h->stream>>>(           \
	    md, tab + ph.toff, ph.kact, tab + ph.xpart, tab + ph.pbase,            \
	    tab + ph.pstart, p.G, h->mm_ecol.as<uint32_t>(), ev, X, Y, part, h->mm_packed)
	if (first && dots) TBTKX_MS(true, true);
	else if (first) TBTKX_MS(true, false);
	else if (dots) TBTKX_MS(false, true);
	else TBTKX_MS(false, false);
#undef TBTKX_MS
}

void launch_step(tbtkx_context *h, const Plan &p, const Phase &ph, bool first,
                 bool dots, const float2 *X, float2 *Y) {
	const bool R = h->mm_real;
#define TBTKX_LS(WW)                                                  \
	do {                                                          \
		if (R) launch_step_w<WW, true>(h, p, ph, first, dots, X, Y);  \
		else launch_step_w<WW, false>(h, p, ph, first, dots, X, Y);   \
	} while (0)
	if (h->mm_W == 4) TBTKX_LS(4);
	else if (h->mm_W == 5) TBTKX_LS(5);
	else if (h->mm_W == 6) TBTKX_LS(6);
	else TBTKX_LS(8);
#undef TBTKX_LS
}

int reserve_plan(tbtkx_context *h, const Plan &p, int ncols, int64_t mu_elems) {
	int rc;
	if ((rc = h->mm_desc.reserve(p.desc.size() * sizeof(MDesc)))) return rc;
	if ((rc = h->mm_tab.reserve(p.tab.size() * 4))) return rc;
	if ((rc = h->mm_x.reserve((size_t)p.vec_elems * 8))) return rc;
	if ((rc = h->mm_y.reserve((size_t)p.vec_elems * 8))) return rc;
	if ((rc = h->mm_part.reserve((size_t)(p.max_slots > 0 ? p.max_slots : 1) * 2 *
	                             kMaxWidth * 16)))
		return rc;
	if ((rc = h->mm_seg.reserve((size_t)(p.max_seg > 0 ? p.max_seg : 1) * 2 *
	                            kMaxWidth * 16)))
		return rc;
	if ((rc = h->mm_D.reserve((size_t)(mu_elems > 0 ? mu_elems : 1) * 16))) return rc;
	if ((rc = h->mm_cols.reserve((size_t)ncols * sizeof(GCol)))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->mm_desc.ptr, p.desc.data(),
	                               p.desc.size() * sizeof(MDesc),
	                               hipMemcpyHostToDevice, h->stream));
	if (!p.tab.empty())
		TBTKX_CHECK_HIP(hipMemcpyAsync(h->mm_tab.ptr, p.tab.data(), p.tab.size() * 4,
		                               hipMemcpyHostToDevice, h->stream));
	return TBTKX_OK;
}

// The two ordered reduction passes of the doubling dots of step n.
void launch_reduce(tbtkx_context *h, const Phase &ph, double2 *D, int n) {
	const int *tab = h->mm_tab.as<int>();
	multi_reduce_slots<<<(2 * ph.nseg + kWaves - 1) / kWaves, kThreads, 0, h->stream>>>(
	    h->mm_desc.as<MDesc>(), tab + ph.seg, ph.nseg, h->mm_part.as<double2>(),
	    h->mm_seg.as<double2>());
	multi_reduce_models<<<(2 * ph.kact + kWaves - 1) / kWaves, kThreads, 0, h->stream>>>(
	    h->mm_desc.as<MDesc>(), tab + ph.mseg, ph.kact, h->mm_seg.as<double2>(), D, n);
}

// Everything from the initial vectors to the last step, on h->stream.
// Enqueue steps s_lo .. s_hi - 1 of the run (all phases' steps counted in
// order; the vectors' roles alternate with the step's parity), with the
// start vectors before step 0 and, without the doubling dots, the final
// gather after the last step.
int enqueue_run(tbtkx_context *h, const Plan &p, bool dots, int ncols, int64_t s_lo,
                int64_t s_hi) {
	const GCol *cols = h->mm_cols.as<GCol>();
	double2 *D = h->mm_D.as<double2>();
	const int *tab = h->mm_tab.as<int>();
	float2 *X = h->mm_x.as<float2>(), *Y = h->mm_y.as<float2>();
	if (s_lo & 1) std::swap(X, Y);
	if (s_lo == 0) {
		multi_fill<<<2048, kThreads, 0, h->stream>>>(X, p.vec_elems);
		multi_units<<<grid_for(ncols), kThreads, 0, h->stream>>>(X, cols, ncols);
	}
	int64_t launches = 0, step = 0;
	for (const Phase &ph : p.phases) {
		for (int n = ph.n_lo; n < ph.n_hi; n++, step++) {
			if (step < s_lo || step >= s_hi) continue;
			if (!dots)
				multi_gather<<<grid_for(ncols), kThreads, 0, h->stream>>>(
				    cols, ncols, X, D, n);
			launch_step(h, p, ph, n == 0, dots, X, Y);
			launches++;
			if (dots) launch_reduce(h, ph, D, n);
			float2 *tmp = X;
			X = Y;
			Y = tmp;
		}
	}
	if (!dots && s_hi >= step)
		multi_gather<<<grid_for(ncols), kThreads, 0, h->stream>>>(cols, ncols, X, D,
		                                                          p.Smax);
	TBTKX_CHECK_HIP(hipGetLastError());
	h->last_launches = launches;
	return TBTKX_OK;
}

// Steps per captured graph of the run (3 kernel nodes per step with the
// doubling dots).
constexpr int64_t kGraphSteps = 256;

// Steps of the run (over all phases).
int64_t run_steps(const Plan &p) {
	int64_t n = 0;
	for (const Phase &ph : p.phases) n += ph.n_hi - ph.n_lo;
	return n;
}

// Autotuner: time a few steps of the first phase for each (tile groups, grid
// share) candidate on the actual batch and keep the fastest: persistent grids
// (share 1 or 2) with tiles of 2-16 row groups, and one tile per wavefront
// (share 0) with tiles of 8-64 row groups.
int autotune(tbtkx_context *h, const std::vector<int> &width, const std::vector<int> &S,
             bool dots, int ncols, int64_t mu_elems) {
	static const int kCand[][2] = {{1, 2}, {1, 4}, {1, 8}, {1, 16}, {2, 2},  {2, 4},
	                               {2, 8}, {2, 16}, {0, 8}, {0, 16}, {0, 32}, {0, 64}};
	constexpr int kTrial = 4;
	float best = 0;
	int bestG = 8, bestShare = 1;
	for (const auto &cand : kCand) {
		const int share = cand[0], G = cand[1];
		{
			Plan p;
			int rc = make_plan(h, width, S, dots, G, share, p);
			if (rc) return rc;
			// One dot slot per tile: skip shapes whose slots would take
			// more than kMaxSlotBytes (very large batches of small tiles).
			if (share == 0 && dots &&
			    (int64_t)p.max_slots * 2 * kMaxWidth * 16 > kMaxSlotBytes)
				continue;
			if ((rc = reserve_plan(h, p, ncols, mu_elems))) return rc;
			const Phase &ph = p.phases[0];
			float2 *X = h->mm_x.as<float2>(), *Y = h->mm_y.as<float2>();
			multi_fill<<<2048, kThreads, 0, h->stream>>>(X, p.vec_elems);
			launch_step(h, p, ph, true, dots, X, Y);
			TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
			for (int i = 0; i < kTrial; i++) {
				launch_step(h, p, ph, false, dots, Y, X);
				if (dots) launch_reduce(h, ph, h->mm_D.as<double2>(), 0);
				float2 *tmp = X;
				X = Y;
				Y = tmp;
			}
			TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
			TBTKX_CHECK_HIP(hipEventSynchronize(h->ev_stop));
			TBTKX_CHECK_HIP(hipGetLastError());
			float ms = 0;
			TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
			if (best == 0 || ms < best) {
				best = ms;
				bestG = G;
				bestShare = share;
			}
		}
	}
	h->mm_tune_g = bestG;
	h->mm_tune_occ = bestShare;
	return TBTKX_OK;
}

}  // namespace
}  // namespace tbtkx

using namespace tbtkx;

extern "C" {

int tbtkx_multi_set(tbtkx_handle h, int K, const int64_t *num_rows,
                    const int64_t *nnz, const uint32_t *rowptr, const uint32_t *col,
                    const double *val, const double *scales) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_multi_set: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(K > 0 && K <= 65536 && num_rows && nnz && rowptr && scales,
	              TBTKX_ERR_ARG, "tbtkx_multi_set: need 1..65536 models and arrays");
	int64_t rp0 = 0, z0 = 0, rows = 0, entries = 0;
	uint32_t maxlen = 0;
	bool real = true;
	std::vector<int> Wm(K, 4);   // ELL slots per row, per model
	std::vector<int> realm(K, 1);   // per model: every value of its H is real
	for (int m = 0; m < K; m++) {
		const int64_t N = num_rows[m], Z = nnz[m];
		const std::string tag = "tbtkx_multi_set: model " + std::to_string(m) + ": ";
		TBTKX_REQUIRE(N > 0 && N < ((int64_t)1 << 31) && Z >= 0 && Z < ((int64_t)1 << 32),
		              TBTKX_ERR_ARG, tag + "bad size");
		TBTKX_REQUIRE(scales[m] > 0, TBTKX_ERR_ARG,
		              tag + "the scale factor has to be larger than 0");
		TBTKX_REQUIRE(Z == 0 || (col && val), TBTKX_ERR_ARG, tag + "null array");
		const uint32_t *rp = rowptr + rp0;
		TBTKX_REQUIRE(rp[0] == 0 && (int64_t)rp[N] == Z, TBTKX_ERR_ARG,
		              tag + "rowptr[0] must be 0 and rowptr[N] == nnz");
		for (int64_t r = 0; r < N; r++)
			TBTKX_REQUIRE(rp[r] <= rp[r + 1], TBTKX_ERR_ARG, tag + "rowptr not monotone");
		for (int64_t k = 0; k < Z; k++) {
			TBTKX_REQUIRE(col[z0 + k] < (uint64_t)N, TBTKX_ERR_ARG,
			              tag + "column index out of range");
			if (val[2 * (z0 + k) + 1] != 0.0) {
				real = false;
				realm[m] = 0;
			}
		}
		uint32_t ml = 0;
		for (int64_t r = 0; r < N; r++) {
			uint32_t len = rp[r + 1] - rp[r];
			bool diag = false;
			for (uint32_t k = rp[r]; k < rp[r + 1]; k++)
				if ((int64_t)col[z0 + k] == r) diag = true;
			ml = std::max(ml, len + (diag ? 0u : 1u));
		}
		maxlen = std::max(maxlen, ml);
		Wm[m] = ell_slots_multi(ml);
		entries += N * Wm[m];
		rp0 += N + 1;
		z0 += Z;
		rows += N;
	}
	TBTKX_REQUIRE(maxlen <= 8, TBTKX_ERR_ARG,
	              "tbtkx_multi_set: rows with more than 8 non-zeros (7 without a "
	              "diagonal entry) are not supported");
	TBTKX_REQUIRE(rows * 8 < ((int64_t)1 << 40), TBTKX_ERR_ARG,
	              "tbtkx_multi_set: batch too large");
	int rc = set_device(h);
	if (rc) return rc;
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	const int W = ell_slots_multi(maxlen);
	std::vector<uint32_t> ec((size_t)entries);
	std::vector<float> e1((size_t)entries * 2, 0.0f), e2((size_t)entries * 2, 0.0f);
	h->mm_N.assign(K, 0);
	h->mm_e0.assign(K, 0);
	h->mm_Wm = Wm;
	h->mm_band.assign(K, 0);
	rp0 = 0;
	z0 = 0;
	int64_t e0 = 0;
	std::vector<double> v1, v2;
	std::vector<int64_t> slots(W + 1);
	for (int m = 0; m < K; m++) {
		const int64_t N = num_rows[m], Z = nnz[m];
		const int Wmod = Wm[m];
		const uint32_t *rp = rowptr + rp0;
		v1.assign(2 * Z, 0.0);
		v2.assign(2 * Z, 0.0);
		scale_complex(val + 2 * z0, v1.data(), Z, 1 / scales[m]);
		scale_complex(v1.data(), v2.data(), Z, 2);
		// Row slots as the single-model ELL copy (ell_row_slots: diagonal
		// first), on the model's local column indices.
		for (int64_t r = 0; r < N; r++) {
			const int64_t base = e0 + r * Wmod;
			const int n = ell_row_slots(rp, col + z0, r, slots.data());
			for (int u = 0; u < Wmod; u++) {
				const int64_t k = u < n ? slots[u] : -1;
				if (k >= 0) {
					ec[base + u] = col[z0 + k];
					if (realm[m]) {
						// One FP64 value in the entry's 8 bytes.
						memcpy(&e1[2 * (base + u)], &v1[2 * k], 8);
						memcpy(&e2[2 * (base + u)], &v2[2 * k], 8);
					} else {
						e1[2 * (base + u)] = (float)v1[2 * k];
						e1[2 * (base + u) + 1] = (float)v1[2 * k + 1];
						e2[2 * (base + u)] = (float)v2[2 * k];
						e2[2 * (base + u) + 1] = (float)v2[2 * k + 1];
					}
				} else {
					ec[base + u] = (uint32_t)r;
				}
			}
		}
		int64_t band = 0;   // max |col - row|: the model's slowest stride
		for (int64_t r = 0; r < N; r++)
			for (int64_t k = rp[r]; k < (int64_t)rp[r + 1]; k++) {
				const int64_t dd = (int64_t)col[z0 + k] - r;
				band = std::max(band, dd < 0 ? -dd : dd);
			}
		h->mm_band[m] = band > 0 ? band : N;
		h->mm_N[m] = N;
		h->mm_e0[m] = e0;
		e0 += N * Wmod;
		rp0 += N + 1;
		z0 += Z;
	}
	h->reset_mm_graphs();
	h->mm_K = 0;
	if ((rc = h->mm_ecol.reserve(ec.size() * 4))) return rc;
	if ((rc = h->mm_ev1.reserve(e1.size() * 4))) return rc;
	if ((rc = h->mm_ev2.reserve(e2.size() * 4))) return rc;
	TBTKX_CHECK_HIP(hipMemcpy(h->mm_ecol.ptr, ec.data(), ec.size() * 4, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->mm_ev1.ptr, e1.data(), e1.size() * 4, hipMemcpyHostToDevice));
	TBTKX_CHECK_HIP(hipMemcpy(h->mm_ev2.ptr, e2.data(), e2.size() * 4, hipMemcpyHostToDevice));
	h->mm_K = K;
	h->mm_W = W;
	h->mm_real = real;
	h->mm_realm = realm;
	h->mm_gen++;
	for (auto &k : h->mm_tune_key) k = -1;
	return TBTKX_OK;
}

int tbtkx_multi_moments(tbtkx_handle h, const int *num_sites, const int64_t *sites,
                        const int *num_moments, int flags, double *out) {
	TBTKX_REQUIRE(h, TBTKX_ERR_ARG, "tbtkx_multi_moments: null handle");
	TBTKX_LOCK(h);
	TBTKX_REQUIRE(h->mm_K > 0, TBTKX_ERR_STATE, "tbtkx_multi_moments: no models set");
	TBTKX_REQUIRE(num_sites && sites && num_moments && out, TBTKX_ERR_ARG,
	              "tbtkx_multi_moments: null array");
	TBTKX_REQUIRE((flags & ~TBTKX_DOUBLING) == 0, TBTKX_ERR_ARG,
	              "tbtkx_multi_moments: only TBTKX_DOUBLING is supported (FP32 storage)");
	const int K = h->mm_K;
	const bool dots = (flags & TBTKX_DOUBLING) != 0;
	// Real H in every model (the start vectors are unit vectors): two real
	// vectors per complex column.
	const int packed = h->real_path && h->mm_real;
	h->mm_packed = packed;
	std::vector<int> width(K), S(K);
	std::vector<int64_t> out_off(K + 1, 0);
	int ncols = 0;
	for (int m = 0; m < K; m++) {
		const std::string tag = "tbtkx_multi_moments: model " + std::to_string(m) + ": ";
		TBTKX_REQUIRE(num_sites[m] >= 1 && num_sites[m] <= kMaxWidth, TBTKX_ERR_ARG,
		              tag + "needs 1..64 sites");
		TBTKX_REQUIRE(num_moments[m] >= 1, TBTKX_ERR_ARG,
		              tag + "numCoefficients has to be larger than 0");
		for (int b = 0; b < num_sites[m]; b++)
			TBTKX_REQUIRE(sites[ncols + b] >= 0 && sites[ncols + b] < h->mm_N[m],
			              TBTKX_ERR_ARG, tag + "site outside the basis");
		width[m] = packed ? (num_sites[m] + 1) / 2 : num_sites[m];
		// The step addresses a model's vectors with 32-bit byte offsets
		// (stored columns: at most one past the model's).
		TBTKX_REQUIRE(h->mm_N[m] * (width[m] + 1) * 8 < ((int64_t)1 << 32), TBTKX_ERR_ARG,
		              tag + "rows x (columns + 1) x 8 B must stay below 4 GiB");
		S[m] = dots ? (num_moments[m] + 1) / 2 : num_moments[m] - 1;
		out_off[m + 1] = out_off[m] + (int64_t)num_sites[m] * num_moments[m];
		ncols += num_sites[m];
	}
	int rc = set_device(h);
	if (rc) return rc;
	const int Smax = *std::max_element(S.begin(), S.end());

	// Launch shape: explicit (TBTKX_OPT_TILE_ROWS = row groups per tile, with
	// TBTKX_OPT_GRID_SHARE: 0 = one tile per wavefront, s > 0 = persistent
	// grid of occupancy x CUs / s workgroups) or autotuned once per (batch,
	// mode, widths).
	int64_t hw = 1469598103934665603ll;
	for (int m = 0; m < K; m++) hw = (hw ^ (width[m] * 131 + S[m])) * 1099511628211ll;
	const int64_t key[3] = {h->mm_gen, flags + (packed << 16), hw};
	int G = 8, share = 1;
	if (h->tile_rows > 0) {
		G = std::max(kRpi, (h->tile_rows + kRpi - 1) / kRpi * kRpi);
		share = h->grid_share;
	} else if (h->autotune && Smax > 0) {
		if (memcmp(key, h->mm_tune_key, sizeof key) != 0) {
			// Output sizes are needed to reserve buffers for the trial runs.
			int64_t mu_elems = 0;
			for (int m = 0; m < K; m++)
				mu_elems += dots ? (int64_t)S[m] * 2 * width[m] : out_off[m + 1] - out_off[m];
			if ((rc = autotune(h, width, S, dots, ncols, mu_elems))) return rc;
			memcpy(h->mm_tune_key, key, sizeof key);
		}
		G = h->mm_tune_g;
		share = h->mm_tune_occ;
	}
	h->last_share = share;
	h->last_tile_groups = G;
	Plan p;
	if (Smax > 0) {
		if ((rc = make_plan(h, width, S, dots, G, share, p))) return rc;
	} else {
		// One moment per model: only the initial gather.
		p.order.resize(K);
		std::iota(p.order.begin(), p.order.end(), 0);
		p.desc.resize(K);
		int64_t vec = 0;
		for (int m = 0; m < K; m++) {
			MDesc &d = p.desc[m];
			d = MDesc{};
			d.N = h->mm_N[m];
			d.e0 = h->mm_e0[m];
			d.W = h->mm_Wm[m];
		d.realh = h->mm_realm[m];
			d.vec0 = vec;
			d.w = width[m];
			d.wp = width[m];
			d.Q = kMaxWidth / width[m];
			vec += h->mm_N[m] * width[m];
		}
		p.vec_elems = vec;
	}
	// Columns: X offset of each unit start vector and its output row.
	std::vector<int> pos(K);
	for (int k = 0; k < K; k++) pos[p.order[k]] = k;
	std::vector<GCol> cols(ncols);
	for (int m = 0, c = 0; m < K; m++) {
		const MDesc &d = p.desc[pos[m]];
		for (int b = 0; b < num_sites[m]; b++, c++) {
			cols[c].xoff = d.vec0 + sites[c] * d.wp + (packed ? b >> 1 : b);
			cols[c].out = out_off[m] + (int64_t)b * num_moments[m];
			cols[c].M = num_moments[m];
			cols[c].comp = packed ? 1 + (b & 1) : 0;
		}
	}
	const int64_t mu_elems = dots ? p.d_elems : out_off[K];
	if ((rc = reserve_plan(h, p, ncols, mu_elems))) return rc;
	TBTKX_CHECK_HIP(hipMemcpyAsync(h->mm_cols.ptr, cols.data(), cols.size() * sizeof(GCol),
	                               hipMemcpyHostToDevice, h->stream));

	TBTKX_CHECK_HIP(hipEventRecord(h->ev_start, h->stream));
	const int64_t nsteps = run_steps(p);
	if (h->use_graph) {
		// The run is captured as graphs of at most kGraphSteps steps each (a
		// few hundred kernel nodes per graph rather than one graph of three
		// launches per step), all cached under one key and replayed in order.
		const int64_t nchunk = nsteps > 0 ? (nsteps + kGraphSteps - 1) / kGraphSteps : 1;
		int64_t mh = hw;
		for (int m = 0; m < K; m++) mh = (mh ^ num_moments[m]) * 1099511628211ll;
		int64_t gkey[GraphCache::kKey];
		const bool key_ok = GraphCache::make_key(
		    gkey, {h->mm_gen, flags, mh, G * 16 + share, p.nwg, packed, nchunk},
		    {h->mm_desc.ptr, h->mm_tab.ptr, h->mm_x.ptr, h->mm_y.ptr, h->mm_part.ptr,
		     h->mm_seg.ptr, h->mm_D.ptr, h->mm_cols.ptr, h->mm_ecol.ptr, h->mm_ev1.ptr,
		     h->mm_ev2.ptr});
		TBTKX_REQUIRE(key_ok, TBTKX_ERR_STATE, "tbtkx_multi_moments: graph key too long");
		bool cached = (int64_t)h->mm_graphs.size() == nchunk;
		for (int64_t c = 0; cached && c < nchunk; c++)
			cached = h->mm_graphs[c].matches(gkey);
		if (!cached) {
			for (GraphCache &g : h->mm_graphs) g.reset();
			h->mm_graphs.clear();
			h->mm_graphs.resize(nchunk);
			for (int64_t c = 0; c < nchunk; c++) {
				GraphCache &gc = h->mm_graphs[c];
				TBTKX_CHECK_HIP(hipStreamBeginCapture(h->stream,
				                                      hipStreamCaptureModeThreadLocal));
				int erc = enqueue_run(h, p, dots, ncols, c * kGraphSteps,
				                      std::min(nsteps, (c + 1) * kGraphSteps));
				hipError_t ec = hipStreamEndCapture(h->stream, &gc.graph);
				if (erc) return erc;
				TBTKX_CHECK_HIP(ec);
				TBTKX_CHECK_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
				memcpy(gc.key, gkey, sizeof gkey);
			}
		}
		for (GraphCache &gc : h->mm_graphs) TBTKX_CHECK_HIP(hipGraphLaunch(gc.exec, h->stream));
		h->last_launches = Smax;
	} else {
		if ((rc = enqueue_run(h, p, dots, ncols, 0, nsteps))) return rc;
	}
	TBTKX_CHECK_HIP(hipEventRecord(h->ev_stop, h->stream));
	std::vector<double2> buf(mu_elems > 0 ? mu_elems : 1);
	TBTKX_CHECK_HIP(hipMemcpyAsync(buf.data(), h->mm_D.ptr, (size_t)mu_elems * 16,
	                               hipMemcpyDeviceToHost, h->stream));
	TBTKX_CHECK_HIP(hipStreamSynchronize(h->stream));
	float ms = 0;
	TBTKX_CHECK_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
	h->last_ms = ms;
	h->last_real = packed;

	double2 *o2 = reinterpret_cast<double2 *>(out);
	if (!dots) {
		memcpy(o2, buf.data(), (size_t)out_off[K] * 16);
		return TBTKX_OK;
	}
	for (int m = 0; m < K; m++) {
		const MDesc &d = p.desc[pos[m]];
		const int M = num_moments[m], w = width[m];
		for (int b = 0; b < num_sites[m]; b++) {
			double2 *o = o2 + out_off[m] + (int64_t)b * M;
			// Column of vector b; packed: its component.
			const int c = packed ? b >> 1 : b;
			auto Dv = [&](int n, int which) {
				const double2 v = buf[d.dof + ((int64_t)n * 2 + which) * w + c];
				if (!packed) return v;
				return make_double2((b & 1) ? v.y : v.x, 0.0);
			};
			const double2 mu0 = Dv(0, 0), mu1 = Dv(0, 1);
			o[0] = mu0;
			if (M > 1) o[1] = mu1;
			for (int n = 1; n < S[m]; n++) {
				const double2 a = Dv(n, 0), c = Dv(n, 1);
				if (2 * n < M) o[2 * n] = make_double2(2 * a.x - mu0.x, 2 * a.y - mu0.y);
				if (2 * n + 1 < M)
					o[2 * n + 1] = make_double2(2 * c.x - mu1.x, 2 * c.y - mu1.y);
			}
		}
	}
	return TBTKX_OK;
}

}  // extern "C"
