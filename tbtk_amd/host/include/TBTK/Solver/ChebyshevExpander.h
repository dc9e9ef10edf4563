/* Drop-in Solver::ChebyshevExpander for MI355X.
 *
 * Same public interface, defaults, setters and asserts as the reference
 * (Lib/include/TBTK/Solver/ChebyshevExpander.h:72-566): scale factor 1.1,
 * 1000 coefficients, broadening 1e-6, energy window Range(-1, 1, 1000), GPU
 * flags and lookup table off. The GPU members that the reference builds from
 * Lib/src/cuda/ChebyshevSolver.cu call the tbtkx engine (include/tbtkx.h).
 *
 * Both settings of setCalculateCoefficientsOnGPU and
 * setGenerateGreensFunctionsOnGPU run on the GPU, and both give the
 * reference CPU path's results bit for bit: the FP64 recurrence of
 * calculateCoefficientsCPU and, for the Green's functions, the lookup-table
 * or direct evaluation of generateGreensFunctionCPU (Lib/src/Solver/
 * ChebyshevExpander.cpp:103-199, 277-435). So the reference's callers, which
 * set both flags false (Templates/BasicChebyshev/src/main.cpp:100-101,
 * doc/examples/Solver/ChebyshevExpander.cpp:28-29), run unchanged. As in the
 * reference, setGenerateGreensFunctionsOnGPU(true) requires the lookup table;
 * unlike its GPU path it supports all four Green's function types. There is
 * no CPU compute path: without a GPU the calls exit with code 1.
 *
 * Threads: each solver owns one engine handle, created on the device with
 * the fewest live handles; calls on one solver from several threads run one
 * after another (the handle is locked per call), calls on different solvers
 * run concurrently.
 *
 * Additions (no reference counterpart):
 *  - setUseBlockRecurrence: diagonal queries of the property extractor
 *    (LDOS, DOS) run as blocks of start vectors through one SpMM recurrence
 *    (FP32 vectors, a complex H in FP32 and a real H in FP64, FP64
 *    accumulation, doubling identities by default);
 *  - calculateCoefficientsBlock, calculateTraceCoefficients,
 *    generateGreensFunctions (many coefficient vectors in one contraction);
 *  - setDamping(Lorentz | Jackson).
 */
#ifndef TBTK_AMD_SOLVER_CHEBYSHEV_EXPANDER
#define TBTK_AMD_SOLVER_CHEBYSHEV_EXPANDER

#include "TBTK/Model.h"
#include "TBTK/Range.h"
#include "TBTK/Streams.h"

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct tbtkx_context;

namespace TBTK {
namespace Solver {

class Solver {
public:
	virtual ~Solver() {}
	virtual void setModel(Model &model) { this->model = &model; }
	Model &getModel();
	const Model &getModel() const;

private:
	Model *model = nullptr;
};

class ChebyshevExpander : public Solver, public Communicator {
public:
	ChebyshevExpander();
	virtual ~ChebyshevExpander();
	ChebyshevExpander(const ChebyshevExpander &) = delete;
	ChebyshevExpander &operator=(const ChebyshevExpander &) = delete;

	void setScaleFactor(double scaleFactor);
	double getScaleFactor() const { return scaleFactor; }
	void setNumCoefficients(int numCoefficients);
	int getNumCoefficients() const { return numCoefficients; }
	void setBroadening(double broadening) { this->broadening = broadening; }
	double getBroadening() const { return broadening; }
	void setEnergyWindow(const Range &energyWindow);
	const Range &getEnergyWindow() const { return energyWindow; }
	void setCalculateCoefficientsOnGPU(bool v) { calculateCoefficientsOnGPU = v; }
	bool getCalculateCoefficientsOnGPU() const { return calculateCoefficientsOnGPU; }
	void setGenerateGreensFunctionsOnGPU(bool v) { generateGreensFunctionsOnGPU = v; }
	bool getGenerateGreensFunctionsOnGPU() const { return generateGreensFunctionsOnGPU; }
	void setUseLookupTable(bool useLookupTable);
	bool getUseLookupTable() const { return useLookupTable; }

	std::vector<std::vector<std::complex<double>>> calculateCoefficients(
	    std::vector<Index> &to, Index from);
	std::vector<std::complex<double>> calculateCoefficients(Index to, Index from);

	enum class Type { Advanced, Retarded, Principal, NonPrincipal };

	std::vector<std::complex<double>> generateGreensFunction(
	    const std::vector<std::complex<double>> &coefficients,
	    Type type = Type::Retarded);

	// ---------------------------------------------------- additions --
	enum class Damping { Lorentz, Jackson };
	/* Lorentz (default; uses the broadening, as the reference) or Jackson
	 * (Rev. Mod. Phys. 78, 275 (2006), eq. 71). */
	void setDamping(Damping d) { damping = d; }
	Damping getDamping() const { return damping; }

	/* Diagonal queries of the property extractor (LDOS, DOS, Density) as
	 * FP32 blocks with the doubling identities (approximate, ~1e-6) instead
	 * of the exact path (off by default). */
	void setUseBlockRecurrence(bool v) { useBlockRecurrence = v; }
	bool getUseBlockRecurrence() const { return useBlockRecurrence; }
	/* tbtkx flags of the block recurrence (TBTKX_DOUBLING by default). */
	void setBlockFlags(int flags) { blockFlags = flags; }
	int getBlockFlags() const { return blockFlags; }

	/* Damped coefficients for many from-indices, each with its own
	 * to-list: result[j][k] equals calculateCoefficients(to[j], from[j])[k]
	 * bit for bit. One engine call: exact FP64 blocks of start vectors
	 * through one SpMM per step when that moves fewer bytes than one
	 * recurrence per from-index (tbtkx_moments_multi). */
	std::vector<std::vector<std::vector<std::complex<double>>>> calculateCoefficientsMulti(
	    const std::vector<Index> &from, const std::vector<std::vector<Index>> &to);
	/* Damped diagonal coefficients <i|T_n|i> for every index, computed as
	 * blocks of unit start vectors. */
	std::vector<std::vector<std::complex<double>>> calculateCoefficientsBlock(
	    const std::vector<Index> &indices);
	/* Damped trace coefficients sum_r <r|T_n|r> over numVectors random-phase
	 * vectors (seeded), divided by numVectors. */
	std::vector<std::complex<double>> calculateTraceCoefficients(int numVectors,
	                                                            uint64_t seed);
	/* Green's functions of many coefficient vectors in one contraction. */
	std::vector<std::vector<std::complex<double>>> generateGreensFunctions(
	    const std::vector<std::vector<std::complex<double>>> &coefficients,
	    Type type = Type::Retarded);

	/* Device time of the last recurrence (ms, step launches). */
	void getLastTiming(double &milliseconds, int64_t &launches) const;

private:
	double scaleFactor;
	int numCoefficients;
	double broadening;
	Range energyWindow;
	bool calculateCoefficientsOnGPU;
	bool generateGreensFunctionsOnGPU;
	bool useLookupTable;
	Damping damping = Damping::Lorentz;
	bool useBlockRecurrence = false;
	int blockFlags;

	/* Host lookup table L[n][e] (generateLookupTable,
	 * Lib/src/Solver/ChebyshevExpander.cpp:201-270) and its device copy. */
	std::vector<std::complex<double>> lookupTable;
	bool lookupTableValid = false;
	bool lookupTableOnDevice = false;
	int lookupTableNumCoefficients = 0;
	int lookupTableResolution = 0;

	tbtkx_context *handle = nullptr;
	/* Held by every computing call: threads sharing the solver run their
	 * calls one after another, and the lookup table is built once. */
	std::recursive_mutex callMutex;
	const Model *uploadedModel = nullptr;
	uint64_t uploadedGeneration = 0;
	double uploadedScale = 0;

	void ensureHandle();
	void ensureModelOnDevice();
	void ensureLookupTableIsReady();
	void generateLookupTable();
	void destroyLookupTable();
	void check(int status, const char *function) const;
	void applyDamping(std::vector<std::complex<double>> &c) const;
	int toTbtkxType(Type type) const;
};

}  // namespace Solver
}  // namespace TBTK
#endif
