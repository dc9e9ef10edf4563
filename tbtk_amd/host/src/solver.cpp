// Drop-in Solver::ChebyshevExpander on the tbtkx engine.
#include "TBTK/Solver/ChebyshevExpander.h"
#include "TBTK/TBTKMacros.h"

#include "tbtkx.h"

#include <cmath>
#include <limits>

namespace TBTK {
namespace Solver {

Model &Solver::getModel() {
	TBTKAssert(model != nullptr, "Solver::getModel()", "Model not set.",
	           "Use Solver::setModel() to set the model.");
	return *model;
}

const Model &Solver::getModel() const {
	TBTKAssert(model != nullptr, "Solver::getModel()", "Model not set.",
	           "Use Solver::setModel() to set the model.");
	return *model;
}

// Defaults of Lib/src/Solver/ChebyshevExpander.cpp:44-58.
ChebyshevExpander::ChebyshevExpander()
    : Communicator(false), scaleFactor(1.1), numCoefficients(1000), broadening(1e-6),
      energyWindow(-1, 1, 1000), calculateCoefficientsOnGPU(false),
      generateGreensFunctionsOnGPU(false), useLookupTable(false),
      blockFlags(TBTKX_DOUBLING) {}

ChebyshevExpander::~ChebyshevExpander() {
	if (handle) tbtkx_destroy(handle);
}

void ChebyshevExpander::setScaleFactor(double scaleFactor) {
	TBTKAssert(scaleFactor > 0, "Solver::ChebyshevExapnder::setScaleFactor()",
	           "The 'scaleFactor=" << scaleFactor << "' has to be larger than '0'.", "");
	destroyLookupTable();
	this->scaleFactor = scaleFactor;
}

void ChebyshevExpander::setNumCoefficients(int numCoefficients) {
	destroyLookupTable();
	this->numCoefficients = numCoefficients;
}

// Lib/include/TBTK/Solver/ChebyshevExpander.h:434-451.
void ChebyshevExpander::setEnergyWindow(const Range &energyWindow) {
	const double epsilon = std::numeric_limits<double>::epsilon();
	TBTKAssert(energyWindow[0] > -scaleFactor * (1 - 16 * epsilon) &&
	               energyWindow.getLast() < scaleFactor * (1 - 16 * epsilon),
	           "Solver::ChebyshevExpander::setEnergyWindow()",
	           "Invalid energy window. The 'energyWindow=[" << energyWindow[0] << ", "
	               << energyWindow.getLast() << "]' is not contained in the interval "
	               << "(-scaleFactor, scaleFactor)=(" << -scaleFactor << ", "
	               << scaleFactor << ").", "");
	destroyLookupTable();
	this->energyWindow = energyWindow;
}

void ChebyshevExpander::setUseLookupTable(bool useLookupTable) {
	if (!useLookupTable) destroyLookupTable();
	this->useLookupTable = useLookupTable;
}

void ChebyshevExpander::check(int status, const char *function) const {
	if (status != TBTKX_OK)
		TBTKExit(function, tbtkx_last_error(),
		         (status == TBTKX_ERR_NODEVICE ? "No GPU available." : ""));
}

void ChebyshevExpander::ensureHandle() {
	if (handle) return;
	int count = 0;
	tbtkx_device_count(&count);
	TBTKAssert(count > 0, "Solver::ChebyshevExpander", "No GPU available.",
	           "This build runs the Chebyshev path on an MI355X.");
	// The handle lives on the device with the fewest live handles, so
	// solvers of different threads spread over the GPUs; its own stream and
	// state stay there for the solver's lifetime.
	check(tbtkx_create(TBTKX_ANY_DEVICE, &handle), "Solver::ChebyshevExpander");
}

void ChebyshevExpander::ensureModelOnDevice() {
	ensureHandle();
	const Model &model = getModel();
	const HoppingAmplitudeSet &has = model.getHoppingAmplitudeSet();
	TBTKAssert(has.getIsConstructed(), "Solver::ChebyshevExpander",
	           "The model has not been constructed.", "Call Model::construct().");
	if (uploadedModel == &model && uploadedGeneration == has.getGeneration() &&
	    uploadedScale == scaleFactor)
		return;
	SparseMatrixCSR m = has.getSparseMatrix();
	check(tbtkx_set_csr(handle, m.numRows, (int64_t)m.values.size(), m.rowPointers.data(),
	                    m.columns.data(), reinterpret_cast<const double *>(m.values.data()),
	                    scaleFactor),
	      "Solver::ChebyshevExpander::calculateCoefficients()");
	uploadedModel = &model;
	uploadedGeneration = has.getGeneration();
	uploadedScale = scaleFactor;
}

// Lorentz: Lib/src/Solver/ChebyshevExpander.cpp:190-196 (same operation
// order). Jackson: Rev. Mod. Phys. 78, 275 (2006), eq. 71.
void ChebyshevExpander::applyDamping(std::vector<std::complex<double>> &c) const {
	const int M = (int)c.size();
	if (damping == Damping::Jackson) {
		const double q = M_PI / (M + 1);
		for (int n = 0; n < M; n++)
			c[n] *= ((M - n + 1) * std::cos(q * n) + std::sin(q * n) * std::cos(q) / std::sin(q)) /
			        (M + 1);
		return;
	}
	if (broadening == 0) return;
	const double lambda = broadening * numCoefficients;
	for (int n = 0; n < M; n++)
		c[n] = c[n] * std::sinh(lambda * (1 - n / (double)numCoefficients)) / std::sinh(lambda);
}

// Both settings of setCalculateCoefficientsOnGPU run the engine's FP64
// recurrence, whose moments are bitwise those of calculateCoefficientsCPU
// (Lib/src/Solver/ChebyshevExpander.cpp:103-199): the reference's dispatch
// (Lib/include/TBTK/Solver/ChebyshevExpander.h:491-527) picks its CPU or GPU
// member, here both are the same device path.
std::vector<std::vector<std::complex<double>>> ChebyshevExpander::calculateCoefficients(
    std::vector<Index> &to, Index from) {
	std::lock_guard<std::recursive_mutex> lock(callMutex);
	TBTKAssert(scaleFactor > 0, "ChebyshevExpander::calculateCoefficientsGPU()",
	           "Scale factor must be larger than zero.",
	           "Use ChebyshevExpander::setScaleFactor() to set scale factor.");
	TBTKAssert(numCoefficients > 0, "ChebyshevExpander::calculateCoefficients()",
	           "numCoefficients has to be larger than zero.", "");
	ensureModelOnDevice();
	const HoppingAmplitudeSet &has = getModel().getHoppingAmplitudeSet();
	const int fromBasis = has.getBasisIndex(from);
	TBTKAssert(fromBasis >= 0, "ChebyshevExpander::calculateCoefficients()",
	           "The from-Index " << from.toString() << " is not in the model.", "");
	std::vector<int64_t> toBasis(to.size());
	for (size_t n = 0; n < to.size(); n++) {
		toBasis[n] = has.getBasisIndex(to[n]);
		TBTKAssert(toBasis[n] >= 0, "ChebyshevExpander::calculateCoefficients()",
		           "The to-Index " << to[n].toString() << " is not in the model.", "");
	}
	std::vector<std::complex<double>> raw(to.size() * (size_t)numCoefficients);
	check(tbtkx_moments(handle, fromBasis, (int)to.size(), toBasis.data(), numCoefficients,
	                    reinterpret_cast<double *>(raw.data())),
	      "ChebyshevExpander::calculateCoefficients()");
	std::vector<std::vector<std::complex<double>>> out(to.size());
	for (size_t n = 0; n < to.size(); n++) {
		out[n].assign(raw.begin() + n * numCoefficients, raw.begin() + (n + 1) * numCoefficients);
		applyDamping(out[n]);
	}
	return out;
}

std::vector<std::complex<double>> ChebyshevExpander::calculateCoefficients(Index to, Index from) {
	std::vector<Index> tos = {to};
	return calculateCoefficients(tos, from)[0];
}

std::vector<std::vector<std::vector<std::complex<double>>>>
ChebyshevExpander::calculateCoefficientsMulti(const std::vector<Index> &from,
                                              const std::vector<std::vector<Index>> &to) {
	std::lock_guard<std::recursive_mutex> lock(callMutex);
	TBTKAssert(from.size() == to.size(), "ChebyshevExpander::calculateCoefficientsMulti()",
	           "Expected one to-list per from-index.", "");
	TBTKAssert(scaleFactor > 0, "ChebyshevExpander::calculateCoefficientsGPU()",
	           "Scale factor must be larger than zero.",
	           "Use ChebyshevExpander::setScaleFactor() to set scale factor.");
	TBTKAssert(numCoefficients > 0, "ChebyshevExpander::calculateCoefficients()",
	           "numCoefficients has to be larger than zero.", "");
	std::vector<std::vector<std::vector<std::complex<double>>>> out(from.size());
	if (from.empty()) return out;
	ensureModelOnDevice();
	const HoppingAmplitudeSet &has = getModel().getHoppingAmplitudeSet();
	std::vector<int64_t> fromBasis(from.size()), toBasis;
	std::vector<int> offsets(1, 0);
	for (size_t j = 0; j < from.size(); j++) {
		fromBasis[j] = has.getBasisIndex(from[j]);
		TBTKAssert(fromBasis[j] >= 0, "ChebyshevExpander::calculateCoefficients()",
		           "The from-Index " << from[j].toString() << " is not in the model.", "");
		for (const Index &t : to[j]) {
			const int b = has.getBasisIndex(t);
			TBTKAssert(b >= 0, "ChebyshevExpander::calculateCoefficients()",
			           "The to-Index " << t.toString() << " is not in the model.", "");
			toBasis.push_back(b);
		}
		TBTKAssert(toBasis.size() < (size_t)1 << 31, "ChebyshevExpander::calculateCoefficientsMulti()",
		           "Too many to-indices in one call.", "");
		offsets.push_back((int)toBasis.size());
	}
	std::vector<std::complex<double>> raw(toBasis.size() * (size_t)numCoefficients);
	check(tbtkx_moments_multi(handle, (int)from.size(), fromBasis.data(), offsets.data(),
	                          toBasis.data(), numCoefficients,
	                          reinterpret_cast<double *>(raw.data())),
	      "ChebyshevExpander::calculateCoefficients()");
	for (size_t j = 0; j < from.size(); j++) {
		out[j].resize(to[j].size());
		for (size_t k = 0; k < to[j].size(); k++) {
			const size_t at = ((size_t)offsets[j] + k) * numCoefficients;
			out[j][k].assign(raw.begin() + at, raw.begin() + at + numCoefficients);
			applyDamping(out[j][k]);
		}
	}
	return out;
}

std::vector<std::vector<std::complex<double>>> ChebyshevExpander::calculateCoefficientsBlock(
    const std::vector<Index> &indices) {
	std::lock_guard<std::recursive_mutex> lock(callMutex);
	TBTKAssert(numCoefficients > 0, "ChebyshevExpander::calculateCoefficientsBlock()",
	           "numCoefficients has to be larger than zero.", "");
	ensureModelOnDevice();
	const HoppingAmplitudeSet &has = getModel().getHoppingAmplitudeSet();
	std::vector<int64_t> sites(indices.size());
	for (size_t n = 0; n < indices.size(); n++) {
		sites[n] = has.getBasisIndex(indices[n]);
		TBTKAssert(sites[n] >= 0, "ChebyshevExpander::calculateCoefficientsBlock()",
		           "The Index " << indices[n].toString() << " is not in the model.", "");
	}
	std::vector<std::complex<double>> raw(sites.size() * (size_t)numCoefficients);
	check(tbtkx_moments_block(handle, (int)sites.size(), sites.data(), numCoefficients,
	                          blockFlags, reinterpret_cast<double *>(raw.data())),
	      "ChebyshevExpander::calculateCoefficientsBlock()");
	std::vector<std::vector<std::complex<double>>> out(sites.size());
	for (size_t n = 0; n < sites.size(); n++) {
		out[n].assign(raw.begin() + n * numCoefficients, raw.begin() + (n + 1) * numCoefficients);
		applyDamping(out[n]);
	}
	return out;
}

std::vector<std::complex<double>> ChebyshevExpander::calculateTraceCoefficients(int numVectors,
                                                                               uint64_t seed) {
	std::lock_guard<std::recursive_mutex> lock(callMutex);
	TBTKAssert(numVectors > 0 && numCoefficients > 0,
	           "ChebyshevExpander::calculateTraceCoefficients()",
	           "numVectors and numCoefficients have to be larger than zero.", "");
	ensureModelOnDevice();
	std::vector<std::complex<double>> mu(numCoefficients);
	check(tbtkx_moments_trace(handle, numVectors, seed, numCoefficients, blockFlags,
	                          reinterpret_cast<double *>(mu.data())),
	      "ChebyshevExpander::calculateTraceCoefficients()");
	for (auto &m : mu) m /= (double)numVectors;
	applyDamping(mu);
	return mu;
}

// generateLookupTable, Lib/src/Solver/ChebyshevExpander.cpp:201-270, with
// the same asserts and the same arithmetic (bitwise equal table).
void ChebyshevExpander::generateLookupTable() {
	TBTKAssert(numCoefficients > 0, "ChebyshevExpander::generateLookupTable()",
	           "numCoefficients has to be larger than 0.", "");
	TBTKAssert(energyWindow.getResolution() > 0, "ChebyshevExpander::generateLookupTable()",
	           "The energy windows resolution has to be larger than 0.", "");
	TBTKAssert(energyWindow[0] < energyWindow.getLast(),
	           "ChebyshevExpander::generateLookupTable()",
	           "The energy windows lower bound has to be smaller than its upper bound.", "");
	TBTKAssert(energyWindow[0] > -scaleFactor, "ChebyshevExpander::generateLookupTable()",
	           "The energy windows lower bound has to be larger than -scaleFactor.",
	           "Use ChebyshevExpander::setScaleFactor to set a larger scale factor.");
	TBTKAssert(energyWindow.getLast() < scaleFactor, "ChebyshevExpander::generateLookupTable()",
	           "The energy windows upper bound has to be smaller than scaleFactor.",
	           "Use ChebyshevExpander::setScaleFactor to set a larger scale factor.");
	const int M = numCoefficients, E = energyWindow.getResolution();
	lookupTable.assign((size_t)M * E, 0);
	const std::complex<double> i(0, 1);
#pragma omp parallel for
	for (int n = 0; n < M; n++) {
		double denominator = 1.;
		if (n == 0) denominator = 2.;
		for (int e = 0; e < E; e++) {
			double En = energyWindow[e] / scaleFactor;
			lookupTable[(size_t)n * E + e] = (1 / scaleFactor) * (-2. * i / sqrt(1 - En * En)) *
			                                 exp(-i * ((double)n) * acos(En)) / denominator;
		}
	}
	lookupTableNumCoefficients = M;
	lookupTableResolution = E;
	lookupTableValid = true;
	lookupTableOnDevice = false;
}

void ChebyshevExpander::destroyLookupTable() {
	lookupTable.clear();
	lookupTableValid = false;
	if (lookupTableOnDevice && handle) tbtkx_lut_destroy(handle);
	lookupTableOnDevice = false;
}

// Built before any contraction, never inside a parallel region (the
// reference builds it under an OpenMP parallel for, which races).
void ChebyshevExpander::ensureLookupTableIsReady() {
	if (!lookupTableValid) generateLookupTable();
	if (!lookupTableOnDevice) {
		ensureHandle();
		check(tbtkx_lut_upload(handle, lookupTableNumCoefficients, lookupTableResolution,
		                       reinterpret_cast<const double *>(lookupTable.data())),
		      "ChebyshevExpander::loadLookupTableGPU()");
		lookupTableOnDevice = true;
	}
}

int ChebyshevExpander::toTbtkxType(Type type) const {
	switch (type) {
	case Type::Advanced: return TBTKX_ADVANCED;
	case Type::Retarded: return TBTKX_RETARDED;
	case Type::Principal: return TBTKX_PRINCIPAL;
	case Type::NonPrincipal: return TBTKX_NONPRINCIPAL;
	}
	TBTKExit("Solver::ChebyshevExpander::generateGreensFunction()",
	         "Unkown Green's function type.", "This should never happen.");
}

// generateGreensFunction (Lib/include/TBTK/Solver/ChebyshevExpander.h:
// 536-545) for many coefficient vectors. Either setting of
// setGenerateGreensFunctionsOnGPU contracts on the device, bitwise as
// generateGreensFunctionCPU (Lib/src/Solver/ChebyshevExpander.cpp:277-435):
// with the lookup table through the host-built table, without it through
// the direct evaluation (factors from the host's libm). As in the
// reference, the GPU setting requires the lookup table
// (ensureLookupTableIsReady, ChebyshevExpander.h:548-565), and it supports
// all four types (the reference's GPU path only Retarded).
std::vector<std::vector<std::complex<double>>> ChebyshevExpander::generateGreensFunctions(
    const std::vector<std::vector<std::complex<double>>> &coefficients, Type type) {
	std::lock_guard<std::recursive_mutex> lock(callMutex);
	if (generateGreensFunctionsOnGPU && !useLookupTable)
		TBTKExit("Solver::ChebyshevSolver::ensureLookupTableIsReady()",
		         "Green's functions can only be generated on GPU using lookup tables.",
		         "Use Solver::ChebyshevExpander::setGenerateGreensFunctionOnGPU() and "
		         "Solver::ChebyshevExpander::setUseLookupTable() to configure the "
		         "Solver::ChebyshevExpander.");
	// Asserts of generateGreensFunctionCPU, ChebyshevExpander.cpp:281-317.
	TBTKAssert(numCoefficients > 0, "ChebyshevExpander::generateGreensFunction()",
	           "numCoefficients has to be larger than 0.", "");
	TBTKAssert(energyWindow.getResolution() > 0, "ChebyshevExpander::generateGreensFunction()",
	           "The energy resolution has to be larger than 0.", "");
	TBTKAssert(energyWindow[0] < energyWindow.getLast(),
	           "ChebyshevExpander::generateGreensFunction()",
	           "The energy windows lower bound has to be smaller than upper bound.", "");
	TBTKAssert(energyWindow[0] > -scaleFactor, "ChebyshevExpander::generateGreensFunction()",
	           "The energy windows lower bound has to be larger than -scaleFactor.",
	           "Use ChebyshevExpander::setScaleFactor to set a larger scale factor.");
	TBTKAssert(energyWindow.getLast() < scaleFactor,
	           "ChebyshevExpander::generateGreensFunction()",
	           "The energy windows upper bound has to be smaller than scaleFactor.",
	           "Use ChebyshevExpander::setScaleFactor to set a larger scale factor.");
	const int K = (int)coefficients.size();
	const int E = energyWindow.getResolution();
	std::vector<std::vector<std::complex<double>>> out(K);
	if (K == 0) return out;
	ensureHandle();
	const int M = useLookupTable ? (lookupTableValid ? lookupTableNumCoefficients : numCoefficients)
	                             : numCoefficients;
	std::vector<std::complex<double>> in((size_t)K * M);
	for (int k = 0; k < K; k++) {
		TBTKAssert((int)coefficients[k].size() >= M, "ChebyshevExpander::generateGreensFunction()",
		           "Expected " << M << " coefficients, got " << coefficients[k].size() << ".", "");
		std::copy(coefficients[k].begin(), coefficients[k].begin() + M, in.begin() + (size_t)k * M);
	}
	std::vector<std::complex<double>> g((size_t)K * E);
	if (useLookupTable) {
		ensureLookupTableIsReady();
		check(tbtkx_contract(handle, K, reinterpret_cast<const double *>(in.data()),
		                     toTbtkxType(type), reinterpret_cast<double *>(g.data())),
		      "ChebyshevExpander::generateGreensFunction()");
	} else {
		const double dx = energyWindow.getDx();
		check(tbtkx_contract_direct(handle, K, M, reinterpret_cast<const double *>(in.data()),
		                            toTbtkxType(type), E, energyWindow[0], dx, scaleFactor,
		                            reinterpret_cast<double *>(g.data())),
		      "ChebyshevExpander::generateGreensFunction()");
	}
	for (int k = 0; k < K; k++) out[k].assign(g.begin() + (size_t)k * E, g.begin() + (size_t)(k + 1) * E);
	return out;
}

std::vector<std::complex<double>> ChebyshevExpander::generateGreensFunction(
    const std::vector<std::complex<double>> &coefficients, Type type) {
	return generateGreensFunctions({coefficients}, type)[0];
}

void ChebyshevExpander::getLastTiming(double &milliseconds, int64_t &launches) const {
	milliseconds = 0;
	launches = 0;
	if (handle) tbtkx_last_timing(handle, &milliseconds, &launches);
}

}  // namespace Solver
}  // namespace TBTK
