/* tbtkx: C ABI of the MI355X (gfx950) Chebyshev-expansion engine.
 *
 * This is the device boundary of the drop-in Solver::ChebyshevExpander. In the
 * reference the boundary is a build-time object swap between
 * Lib/src/cuda/ChebyshevSolver.cu and Lib/src/nocuda/ChebyshevSolver.cpp,
 * both defining the private GPU members of Solver::ChebyshevExpander
 * (Lib/include/TBTK/Solver/ChebyshevExpander.h:337-389). Here those members
 * call the functions below instead.
 *
 * Conventions
 *  - Complex numbers are interleaved (re, im) doubles on the host side.
 *  - Every function returns 0 on success and a nonzero TBTKX_* status
 *    otherwise; tbtkx_last_error() gives the message. Nothing throws across
 *    the ABI.
 *  - A handle owns one HIP stream on one device and all device state (the
 *    Hamiltonian, recurrence vectors, the lookup table). Each call holds the
 *    handle's lock, so threads sharing a handle run their calls one after
 *    another; different handles are independent and run concurrently.
 *  - Host buffers are owned by the caller.
 */
#ifndef TBTKX_H
#define TBTKX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbtkx_context *tbtkx_handle;

enum {
	TBTKX_OK = 0,
	TBTKX_ERR_ARG = 1,        /* invalid argument */
	TBTKX_ERR_HIP = 2,        /* a HIP runtime call failed */
	TBTKX_ERR_STATE = 3,      /* call out of order (e.g. no model set) */
	TBTKX_ERR_NODEVICE = 4,   /* no usable GPU */
	TBTKX_ERR_OOM = 5         /* device allocation failed */
};

/* Green's function types, numbered as Solver::ChebyshevExpander::Type
 * (Lib/include/TBTK/Solver/ChebyshevExpander.h:206-211). */
enum {
	TBTKX_ADVANCED = 0,
	TBTKX_RETARDED = 1,
	TBTKX_PRINCIPAL = 2,
	TBTKX_NONPRINCIPAL = 3
};

/* Flags for the batched recurrences. */
enum {
	TBTKX_H_FP64 = 1 << 0,     /* H as FP64 complex values (default: FP32
	                              complex values, or one FP64 value per
	                              entry when every value of H is real) */
	TBTKX_DOUBLING = 1 << 1,   /* diagonal moments from the doubling
	                              identities mu_2n = 2<n|n> - mu_0,
	                              mu_2n+1 = 2<n+1|n> - mu_1 (M/2 steps) */
	TBTKX_VEC_FP64 = 1 << 2    /* FP64 recurrence vectors (reference block) */
};

/* Message of the last failing call on this thread. */
const char *tbtkx_last_error(void);

/* Number of visible GPUs (replaces GPUResourceManager::createDeviceTable,
 * Lib/src/cuda/GPUResourceManager.cu:28-46). */
int tbtkx_device_count(int *count);

/* Device pool: hand out a free device, or wait for one (replaces
 * GPUResourceManager::allocateDevice / freeDevice,
 * Lib/src/Utilities/GPUResourceManager.cpp:44-90). */
int tbtkx_device_acquire(int *device);
int tbtkx_device_release(int device);

/* A handle on `device`, or with TBTKX_ANY_DEVICE on the device that holds
 * the fewest live handles (the lowest id on a tie), so solvers created by
 * different threads spread over the GPUs the way the reference hands each
 * thread its own device (GPUResourceManager::allocateDevice,
 * Lib/src/Utilities/GPUResourceManager.cpp:44-72). */
enum { TBTKX_ANY_DEVICE = -1 };
int tbtkx_create(int device, tbtkx_handle *handle);
int tbtkx_destroy(tbtkx_handle handle);

/* Upload the Hamiltonian as CSR (rows = basis(to), columns = basis(from),
 * columns ascending in a row) and the scale factor s. The device keeps
 * H*(1/s) and H*(1/s)*2 in FP64 (the two scalings of
 * Lib/src/Solver/ChebyshevExpander.cpp:147,168) and FP32 copies for the
 * batched path. Replaces the per-call H upload of calculateCoefficientsGPU
 * (Lib/src/cuda/ChebyshevSolver.cu:153-317). */
int tbtkx_set_csr(tbtkx_handle handle, int64_t num_rows, int64_t nnz,
                  const uint32_t *rowptr, const uint32_t *col,
                  const double *val, double scale);

/* Plain FP64 path, one start vector: raw moments
 * out[k*M + n] = <to_k| T_n(H/s) |from>, 0 <= n < M, bitwise equal to the
 * reference CPU recurrence. If a to-index repeats, only its last position
 * receives moments and the others stay zero, as the reference's
 * coefficientMap does. Replaces calculateCoefficientsGPU(to, from)
 * (Lib/src/cuda/ChebyshevSolver.cu:74-648; damping stays on the host). */
int tbtkx_moments(tbtkx_handle handle, int64_t from, int num_to,
                  const int64_t *to, int num_moments, double *out);

/* Allocation-free variant of tbtkx_moments, with no counterpart in the
 * reference: for callers that cannot allocate device memory during the call
 * (tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results (bitwise equal to tbtkx_moments) and errors; the recurrence
 * vectors, the moments and the to-list live in a fixed scratch array of
 * TBTKX_NOALLOC_BYTES per GPU, held from the library's load until the
 * process ends, and the call makes no hipMalloc, hipFree or other
 * allocating call (the model itself is set up beforehand with
 * tbtkx_set_csr or tbtkx_set_lattice). A call needing more scratch than
 * that (2 vectors of N values, 8 B each for a real H and 16 B for a complex
 * one, plus K x M x 16 B of moments and K x 8 B, each rounded up to 256 B),
 * or more than min(N, TBTKX_NOALLOC_MAX_TO) to-indices, is rejected with
 * TBTKX_ERR_ARG. All handles on one GPU share its scratch, which keeps the
 * last call's contents; their calls take a per-GPU lock and run one after
 * another (calls never overlap on one GPU). */
#define TBTKX_NOALLOC_BYTES (64u << 20)
#define TBTKX_NOALLOC_MAX_TO 4096
int tbtkx_moments_noalloc(tbtkx_handle handle, int64_t from, int num_to,
                          const int64_t *to, int num_moments, double *out);

/* Load a host-generated lookup table L[n*E + e] (M x E complex) for the
 * contraction. Replaces loadLookupTableGPU
 * (Lib/src/cuda/ChebyshevSolver.cu:672-760). */
int tbtkx_lut_upload(tbtkx_handle handle, int num_moments, int resolution,
                     const double *table);

/* Generate the lookup table on the device in FP64: the formula of
 * generateLookupTable (Lib/src/Solver/ChebyshevExpander.cpp:256-268) over
 * Range(lo, hi, resolution). */
int tbtkx_lut_generate(tbtkx_handle handle, int num_moments, int resolution,
                       double lower, double upper, double scale);

/* Free the device lookup table (destroyLookupTableGPU,
 * Lib/src/cuda/ChebyshevSolver.cu:762-783). */
int tbtkx_lut_destroy(tbtkx_handle handle);

/* FP64 contraction of K coefficient vectors coeffs[k*M + n] with the loaded
 * table into out[k*E + e], for any of the four types, summing n in ascending
 * order as the reference (Lib/src/Solver/ChebyshevExpander.cpp:324-370).
 * Replaces generateGreensFunctionGPU (Lib/src/cuda/ChebyshevSolver.cu:786-921),
 * which supports Retarded only. */
int tbtkx_contract(tbtkx_handle handle, int num_vectors, const double *coeffs,
                   int type, double *out);

/* FP64 contraction without a lookup table: the direct evaluation of
 * generateGreensFunctionCPU (Lib/src/Solver/ChebyshevExpander.cpp:371-432)
 * over the energies e0 + e*de, e < E, for K vectors of M coefficients. The
 * trigonometric factors are computed on the host with the C library's acos,
 * complex exp, sin and cos, as the reference's CPU path computes them, and
 * kept on the device for the next call with the same type class, M and
 * grid; the sums run on the device in the reference's order, so the result
 * is bitwise the reference's. The reference has no GPU counterpart (its GPU
 * path requires the table, Lib/include/TBTK/Solver/ChebyshevExpander.h:
 * 555-565). */
int tbtkx_contract_direct(tbtkx_handle handle, int num_vectors, int num_moments,
                          const double *coeffs, int type, int resolution,
                          double e0, double de, double scale, double *out);

/* Coefficient vectors coeffs[k*M + n] (K x M complex) for the MFMA
 * contraction; kept on the device as split-BF16 planes until replaced. */
int tbtkx_coeffs_upload(tbtkx_handle handle, int num_vectors, int num_moments,
                        const double *coeffs);

/* Contraction of the uploaded coefficients with the loaded table on MFMA:
 * split-BF16 three-term products (about 16 significant bits per operand)
 * with FP32 accumulation, any of the four types. out[k*E + e] (complex, as
 * doubles) receives the result; with out == NULL it stays on the device
 * (for timing). The device time is reported by tbtkx_last_timing. No
 * reference counterpart (the reference contracts one vector at a time,
 * Lib/src/cuda/ChebyshevSolver.cu:650-670). */
int tbtkx_contract_mfma(tbtkx_handle handle, int type, double *out);

/* Damping of the contraction's coefficients, done on the device in FP64:
 * once set, tbtkx_contract, tbtkx_contract_direct and tbtkx_coeffs_upload
 * multiply the coefficients they are given by g_n before contracting.
 * TBTKX_DAMP_LORENTZ: c_n * sinh(lambda (1 - n/M)) / sinh(lambda), lambda =
 * broadening * M, per real and imaginary part as the reference (its CPU path,
 * Lib/src/Solver/ChebyshevExpander.cpp:190-196, and its GPU path on the
 * device, Lib/src/cuda/ChebyshevSolver.cu:632-645); broadening 0 leaves the
 * coefficients unchanged, as there. TBTKX_DAMP_JACKSON: the Jackson kernel
 * g_n = ((M - n + 1) cos(q n) + sin(q n) cot q) / (M + 1), q = pi/(M + 1)
 * (no reference counterpart). TBTKX_DAMP_NONE (the default): coefficients
 * are contracted as given. The factors use the device sinh/sin/cos, so they
 * may differ from the host libm's in the last bit. */
enum {
	TBTKX_DAMP_NONE = 0,
	TBTKX_DAMP_LORENTZ = 1,
	TBTKX_DAMP_JACKSON = 2
};
int tbtkx_set_damping(tbtkx_handle handle, int kind, double broadening);

/* Batched path: diagonal moments mu[b*M + n] = <s_b| T_n(H/s) |s_b> for B
 * unit start vectors at basis indices sites[b] (LDOS sites), run as blocks
 * of start vectors through one SpMM recurrence. FP32 complex vectors (unless
 * TBTKX_VEC_FP64), FP64 accumulation of products and moments. No reference
 * counterpart: the reference loops calculateCoefficients once per site
 * (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:661-684). */
int tbtkx_moments_block(tbtkx_handle handle, int num_sites,
                        const int64_t *sites, int num_moments, int flags,
                        double *out);

/* Moments of many start vectors with many to-indices each:
 * out[k*M + n] = <to_k| T_n(H/s) |from_j> for to_offsets[j] <= k <
 * to_offsets[j+1] (to_offsets[0] = 0, num_from + 1 entries), bitwise equal
 * to tbtkx_moments(from_j, its to-list) for every j, including its rule for
 * a repeated to-index. The engine runs either one plain recurrence per
 * from-index or exact FP64 blocks of 64-256 start vectors through one SpMM
 * per step (CSR order, no fused multiply-add: every column reproduces the
 * plain recurrence bit for bit; two real vectors per column for a real H),
 * whichever moves fewer bytes (TBTKX_OPT_MULTI). This batches the
 * reference's sequential per-index recurrences of
 * PropertyExtractor::ChebyshevExpander::calculateGreensFunction (pattern
 * form) and calculateLDOS (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:
 * 107-232, 475-505), each of which calls calculateCoefficients once per
 * from-index (Lib/src/Solver/ChebyshevExpander.cpp:103-199). */
int tbtkx_moments_multi(tbtkx_handle handle, int num_from, const int64_t *from,
                        const int *to_offsets, const int64_t *to, int num_moments,
                        double *out);

/* Stochastic trace: mu[n] = sum_r <r| T_n(H/s) |r> over R random-phase
 * vectors r_i = exp(2*pi*i*u), u = the 53-bit splitmix64 uniform of
 * (seed, r*N + i). No reference counterpart (the reference's sampleDOS,
 * Lib/src/PropertyExtractor/PropertyExtractor.cpp:234-290, sums LDOS of
 * sampled sites). */
int tbtkx_moments_trace(tbtkx_handle handle, int num_vectors, uint64_t seed,
                        int num_moments, int flags, double *out);

/* Lattice models built on the device. Replaces, for models given as a
 * lattice, the host pipeline HoppingAmplitudeSet -> getSparseMatrix ->
 * SparseMatrix LIL -> CSR (Lib/include/Core/TBTK/HoppingAmplitudeSet.h:397-424,
 * Lib/include/Utilities/TBTK/SparseMatrix.h:1309-1530) with the same basis
 * numbering: the model's Index is {x, y, z, orbital} (or a prefix with the
 * unused extents 1), numbered depth-first over subindex values,
 * row = ((x * cells[1] + y) * cells[2] + z) * orbitals + orbital.
 *
 * Term t adds amplitude (re, im) to H[(c + d, to), (c, from)] for every cell
 * c whose partner c + d lies inside the lattice, or wraps around along a
 * periodic axis (so a hermitian model lists both directions, as
 * "model << HoppingAmplitude(...) + HC" does). onsite (optional, 2 N
 * doubles) is every state's diagonal amplitude; with onsite, every row has
 * a diagonal entry. Entries of one (row, column) are summed in the order
 * onsite, then terms in list order. Rows hold at most 15 distinct columns
 * besides the diagonal. */
typedef struct {
	int32_t d[3];      /* cell offset of the to-cell from the from-cell */
	int32_t to, from;  /* orbitals */
	double re, im;     /* amplitude */
} tbtkx_lattice_term;

typedef struct {
	int32_t cells[3];     /* cells along x, y, z (>= 1) */
	int32_t periodic[3];  /* 1 = periodic boundary along the axis */
	int32_t orbitals;     /* states per cell (>= 1) */
	int32_t num_terms;
	const tbtkx_lattice_term *terms;
} tbtkx_lattice;

/* Build the Hamiltonian of a lattice on the device and make it the handle's
 * model, as tbtkx_set_csr does for a host CSR (same device layouts, so every
 * recurrence entry point works on it). */
int tbtkx_set_lattice(tbtkx_handle handle, const tbtkx_lattice *lattice,
                      const double *onsite, double scale);

/* Gershgorin bound of a lattice model, max over rows of sum_c |H_rc|
 * (the merged entries of the row), computed on the device without building
 * the model: an upper bound of the spectral radius from which to choose
 * the scale factor. */
int tbtkx_lattice_gershgorin(tbtkx_handle handle, const tbtkx_lattice *lattice,
                             const double *onsite, double *bound);

/* Size and copy of the handle's CSR (from tbtkx_set_csr or
 * tbtkx_set_lattice): rowptr (num_rows + 1), col (nnz) and the values
 * H * (1/s) as interleaved complex doubles, exactly as the device holds
 * them. */
int tbtkx_csr_size(tbtkx_handle handle, int64_t *num_rows, int64_t *nnz);
int tbtkx_csr_copy(tbtkx_handle handle, uint32_t *rowptr, uint32_t *col,
                   double *values_over_s);

/* Uneven batch: upload K independent models at once. Model m has
 * num_rows[m] states and nnz[m] non-zeros; its CSR (same conventions as
 * tbtkx_set_csr, indices local to the model) follows the previous model's in
 * rowptr (num_rows[m] + 1 entries per model), col and val, and scales[m] is
 * its scale factor. Rows may hold at most 8 non-zeros. No reference
 * counterpart: the reference handles one model per solver and one start
 * vector per recurrence (Lib/src/Solver/ChebyshevExpander.cpp:103-199). */
int tbtkx_multi_set(tbtkx_handle handle, int num_models, const int64_t *num_rows,
                    const int64_t *nnz, const uint32_t *rowptr, const uint32_t *col,
                    const double *val, const double *scales);

/* Uneven batch from lattices built on the device: model m is lattices[m]
 * with on-site amplitudes onsite[m] (2 N_m doubles, or NULL) and scale
 * factor scales[m]; otherwise as tbtkx_multi_set (rows of at most 8
 * non-zeros). Replaces the same host pipeline as tbtkx_set_lattice. */
int tbtkx_multi_set_lattices(tbtkx_handle handle, int num_models,
                             const tbtkx_lattice *lattices, const double *const *onsite,
                             const double *scales);

/* Diagonal moments of every uploaded model in one launch per
 * recurrence step: model m gets num_sites[m] (1..64) unit start vectors at
 * its local basis indices (consecutive in sites) and num_moments[m] moments.
 * out holds, model after model, num_sites[m] x num_moments[m] complex values
 * (site-major), as tbtkx_moments_block would give per model. FP32 storage,
 * FP64 moments; flags may be TBTKX_DOUBLING. The launch shape (tile size;
 * a persistent grid or one tile per wavefront) is autotuned on first use
 * unless TBTKX_OPT_AUTOTUNE is 0 or TBTKX_OPT_TILE_ROWS fixes it (the grid
 * then follows TBTKX_OPT_GRID_SHARE: 0 = one tile per wavefront, s > 0 = a
 * persistent grid of occupancy x CUs / s workgroups); TBTKX_OPT_GRAPH
 * replays the whole run as a hipGraph. A model's vectors are addressed with 32-bit byte offsets: a
 * model with num_rows x (columns + 1) x 8 B >= 4 GiB is refused with
 * TBTKX_ERR_ARG (columns: num_sites, or ceil(num_sites / 2) when every
 * model's H is real). */
int tbtkx_multi_moments(tbtkx_handle handle, const int *num_sites,
                        const int64_t *sites, const int *num_moments, int flags,
                        double *out);

/* Device time of the last recurrence call: total milliseconds between the
 * first and last step launch (HIP events on the handle's stream) and the
 * number of step launches in it. */
int tbtkx_last_timing(tbtkx_handle handle, double *milliseconds,
                      int64_t *launches);

/* Engine options (0 = default unless stated). TBTKX_OPT_PLAIN_KERNEL:
 * 2 (default) = slot-major ELL (one thread per row, 16-bit offsets from the
 * row, used when every row has at most 8 non-zeros and every offset fits;
 * else 0), 0 = LDS-staged row blocks, 1 = one thread per row on the CSR;
 * reading it gives the kernel the last plain call ran. TBTKX_OPT_GRAPH: 1 = replay the step
 * loop as a hipGraph. TBTKX_OPT_BLOCK_WIDTH: start vectors per SpMM block
 * (64, 128, 256; 0 = chosen by the width autotuner, which times each width on
 * the resident model once per model, precision and size class).
 * TBTKX_OPT_XCD_REMAP: 1 = give each XCD a contiguous row slice (the
 * batched steps; the slot-major plain step always does). TBTKX_OPT_TILE_ROWS: rows per wavefront tile of the SpMM (1-32).
 * TBTKX_OPT_NT_STREAM: 1 = non-temporal loads/stores for the SpMM rows that
 * are used once per step (default 1). */
enum {
	TBTKX_OPT_PLAIN_KERNEL = 1,
	TBTKX_OPT_GRAPH = 2,
	TBTKX_OPT_BLOCK_WIDTH = 3,
	TBTKX_OPT_XCD_REMAP = 4,
	TBTKX_OPT_TILE_ROWS = 5,
	TBTKX_OPT_NT_STREAM = 6,
	TBTKX_OPT_ELL = 7,      /* 1 = batched step on the tile-major ELL copy of H
	                           when no row has more than 8 entries (default) */
	TBTKX_OPT_AUTOTUNE = 8, /* 1 = time candidate launch shapes on first use
	                           (default); 0 = fixed defaults */
	TBTKX_OPT_MFMA_PIPE = 9, /* 1 = MFMA contraction with double-buffered
	                           direct-to-LDS operand loads (default); 0 = one
	                           LDS stage loaded through registers */
	TBTKX_OPT_STRIP_ROWS = 10, /* batched step traversal: S > 0 = rows in L2
	                           strips of S rows per layer of the band (the
	                           largest |col - row|), each XCD owning one part;
	                           -1 = natural row order; 0 = autotuned (default) */
	TBTKX_OPT_GRID_SHARE = 11, /* batched step grid = occupancy x CUs / share
	                           (0 = autotuned, default) */
	TBTKX_OPT_STRIP_CHUNK = 12, /* strip traversal: consecutive tiles dealt to
	                           one workgroup at a time (multiple of 4, 0 = 4) */
	TBTKX_OPT_ROWS_IN_FLIGHT = 13, /* strip traversal, FP32: 1 or 2 times the
	                           default rows a wavefront keeps in flight
	                           (0 = autotuned, default) */
	TBTKX_OPT_REAL = 14,    /* 1 = when every value of H is real and the start
	                           vectors are unit vectors, run the recurrence in
	                           real arithmetic (default): the plain path on
	                           real FP64 vectors, the batched paths with two
	                           start vectors per complex column (one in the
	                           real, one in the imaginary part), so half the
	                           vector bytes. The moments equal those of the
	                           complex recurrence (whose imaginary parts stay
	                           exactly zero). 0 = always complex */
	TBTKX_OPT_COL16 = 15    /* 1 = the plain FP64 step reads its column indices
	                           as 16-bit offsets from the first row of their row
	                           block when every offset fits (default): 2 bytes
	                           instead of 4 per non-zero, same arithmetic, same
	                           moments; 0 = 32-bit columns always. Reading it
	                           gives 1 when the last plain call used them */,
	TBTKX_OPT_PALETTE = 16  /* 1 = with 16-bit columns, the plain FP64 step
	                           reads each off-diagonal value as a 1-byte code
	                           into a palette of the model's distinct values
	                           and each diagonal value from a per-row array,
	                           when the off-diagonal values take at most 255
	                           distinct values (default): same values, same
	                           arithmetic, same moments, a third less HBM
	                           traffic on a square lattice; 0 = full values.
	                           Reading it gives 1 when the last plain call
	                           used the palette */,
	TBTKX_OPT_MULTI = 17    /* tbtkx_moments_multi: 0 = plain recurrences or
	                           exact FP64 blocks, whichever moves fewer bytes
	                           (default); 1 = plain recurrences; 2 = exact
	                           blocks. Reading it gives what the last call
	                           ran (1 plain, 2 blocks) */,
	TBTKX_OPT_ROW_PATTERNS = 18 /* 1 = the slot-major plain step with the
	                           palette reads one byte per row, an index into
	                           the model's at most 255 distinct rows (column
	                           offsets and palette codes), when the model has
	                           that few (default), and a wavefront whose rows
	                           all share one pattern reads none: it takes the
	                           pattern from a byte per 64-row chunk through
	                           scalar loads; 2 = pattern bytes for every row;
	                           all: same values, same order, same moments;
	                           0 = per-slot offsets and codes. Reading it
	                           gives 1 when the last plain call used patterns */,
	TBTKX_OPT_SLOT_ROWS = 19 /* rows per thread of the slot-major plain step:
	                           1, 2 or 4; 0 (default) = 2 when every row reads
	                           its pattern byte (TBTKX_OPT_ROW_PATTERNS = 2),
	                           else 1 */,
	TBTKX_OPT_PATTERN_BYTE_ROWS = 20 /* read-only: rows of the current model
	                           outside 64-row chunks of one shared pattern
	                           (the rows whose pattern byte the step reads
	                           with TBTKX_OPT_ROW_PATTERNS = 1); -1 when the
	                           model has no row patterns */
};
int tbtkx_set_option(tbtkx_handle handle, int option, int64_t value);

/* Effective value of an option: TBTKX_OPT_BLOCK_WIDTH, TBTKX_OPT_STRIP_ROWS
 * and TBTKX_OPT_GRID_SHARE give what the last batched call used (the
 * autotuner's choice when they are 0; the uneven batch sets the grid share
 * too: 0 there means one tile per wavefront, a grid as large as the tile
 * list, else a persistent grid of occupancy x CUs / share workgroups);
 * TBTKX_OPT_TILE_ROWS, when 0, gives the row groups per tile of the
 * last uneven-batch call; TBTKX_OPT_ELL gives 1 only when the
 * ELL copy exists and is enabled; TBTKX_OPT_REAL gives 1 when the last
 * recurrence call (plain, batched or uneven batch) ran in real arithmetic. */
int tbtkx_get_option(tbtkx_handle handle, int option, int64_t *value);

#ifdef __cplusplus
}
#endif
#endif
