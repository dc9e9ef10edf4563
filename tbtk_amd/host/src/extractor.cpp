// Drop-in PropertyExtractor::ChebyshevExpander and the PropertyExtractor
// base behaviour it uses.
#include "TBTK/PropertyExtractor/ChebyshevExpander.h"
#include "TBTK/Functions.h"
#include "TBTK/TBTKMacros.h"

#include <cmath>
#include <cstdlib>
#include <set>

namespace TBTK {
namespace PropertyExtractor {

// ------------------------------------------------------ PropertyExtractor --

void PropertyExtractor::setEnergyWindow(double lowerBound, double upperBound,
                                        int energyResolution) {
	energyWindow = Range(lowerBound, upperBound, energyResolution);
}

Property::DOS PropertyExtractor::calculateDOS() {
	TBTKExit("PropertyExtractor::calculateDOS()",
	         "The chosen property extractor does not support this function call.",
	         "See the API for list of supported calls.");
}

Property::LDOS PropertyExtractor::calculateLDOS(std::vector<Index>) {
	TBTKExit("PropertyExtractor::calculateLDOS()",
	         "The chosen property extractor does not support this function call.",
	         "See the API for list of supported calls.");
}

static bool isWild(int v) { return v == IDX_ALL || v == IDX_SUM_ALL || v == IDX_SPIN; }

// Lib/src/PropertyExtractor/PropertyExtractor.cpp:323-433.
IndexTree PropertyExtractor::generateIndexTree(const std::vector<Index> &patterns,
                                               const HoppingAmplitudeSet &has,
                                               bool keepSummationWildcards,
                                               bool keepSpinWildcards) {
	IndexTree tree;
	for (size_t n = 0; n < patterns.size(); n++) {
		std::vector<Index> components = patterns[n].split();
		if (components.size() == 1) {
			Index pattern = components[0];
			for (unsigned int c = 0; c < pattern.getSize(); c++) {
				TBTKAssert(isWild(pattern.at(c)) || pattern.at(c) >= 0,
				           "PropertyExtractor::generateIndexTree()",
				           "Subindex " << c << " of pattern " << n << " is invalid.",
				           "Must be non-negative, IDX_ALL, IDX_SUM_ALL, or IDX_SPIN.");
				if (isWild(pattern.at(c))) pattern.at(c) = IDX_ALL;
			}
			std::vector<Index> indices = has.getIndexList({pattern});
			const Index &p = components[0];
			for (Index &idx : indices) {
				for (unsigned int m = 0; m < p.getSize(); m++) {
					if (keepSummationWildcards && p.at(m).isSummationIndex())
						idx.at(m) = IDX_SUM_ALL;
					if (keepSpinWildcards && p.at(m).isSpinIndex()) idx.at(m) = IDX_SPIN;
				}
				tree.add(idx);
			}
		} else if (components.size() == 2) {
			IndexTree first = generateIndexTree({components[0]}, has, keepSummationWildcards,
			                                    keepSpinWildcards);
			IndexTree second = generateIndexTree({components[1]}, has, keepSummationWildcards,
			                                     keepSpinWildcards);
			for (const Index &a : first)
				for (const Index &b : second) tree.add(Index({a, b}));
		} else {
			TBTKExit("PropertyExtractor::generateIndexTree()",
			         "Only patterns with one and two component Indices are supported so far, "
			         "but the pattern '" << patterns[n].toString() << "' has '"
			                             << components.size() << "' components.", "");
		}
	}
	tree.generateLinearMap();
	return tree;
}

// Lib/src/PropertyExtractor/PropertyExtractor.cpp:234-290.
Property::DOS PropertyExtractor::sampleDOS(unsigned int numSamples,
                                           const std::vector<Index> &patterns,
                                           unsigned int seed) {
	const bool selective = !patterns.empty();
	const Model &model = getModelForSampling();
	const HoppingAmplitudeSet &has = model.getHoppingAmplitudeSet();
	const unsigned int space =
	    selective ? (unsigned int)has.getIndexList(patterns).size() : model.getBasisSize();
	TBTKAssert(space > 0, "PropertyExtractor::sampleDOS()",
	           "The provided 'patterns' does not match any Index in the Model.", "");
	Property::DOS dos(energyWindow);
	srand(seed);
	for (int n = 0; n < (int)numSamples; n++) {
		const int state = rand() % model.getBasisSize();
		const Index index = has.getPhysicalIndex(state);
		if (selective) {
			bool accepted = false;
			for (const Index &p : patterns)
				if (p.equals(index, true)) {
					accepted = true;
					break;
				}
			if (!accepted) {
				n--;
				continue;
			}
		}
		Property::LDOS ldos = calculateLDOS({index});
		for (unsigned int e = 0; e < dos.getResolution(); e++) dos(e) += ldos(index, e);
	}
	for (unsigned int e = 0; e < dos.getResolution(); e++) dos(e) *= space / (double)numSamples;
	return dos;
}

// ------------------------------------------------------ ChebyshevExpander --

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:35-74.
ChebyshevExpander::ChebyshevExpander(Solver::ChebyshevExpander &solver) : solver(&solver) {
	const Range &w = getEnergyWindow();
	TBTKAssert(w.getResolution() > 0, "PropertyExtractor::ChebyshevExpander::ChebyshevExpander()",
	           "The energy resolution has to be a positive number.", "");
	TBTKAssert(w[0] < w.getLast(), "PropertyExtractor::ChebyshevExpander::ChebyshevExpander()",
	           "The energy window must be accending.", "");
	TBTKAssert(w[0] >= -solver.getScaleFactor(),
	           "PropertyExtractor::ChebyshevExpander::ChebyshevExpander()",
	           "The lower bound for the energy window has to be larger than "
	           "-solver->getScaleFactor().",
	           "Use Solver::ChebyshevExpander::setScaleFactor() to set a larger scale factor.");
	TBTKAssert(w.getLast() <= solver.getScaleFactor(),
	           "PropertyExtractor::ChebyshevExapnder::ChebysheExpander()",
	           "The upper bound for the energy window has to be smaller than "
	           "solver->getScaleFactor().",
	           "Use Solver::ChebyshevExpnader::setScaleFactor() to set a larger scale factor.");
	setEnergyWindow(w[0], w.getLast(), w.getResolution());
}

void ChebyshevExpander::setEnergyWindow(double lowerBound, double upperBound,
                                        int energyResolution) {
	PropertyExtractor::setEnergyWindow(lowerBound, upperBound, energyResolution);
	solver->setEnergyWindow(Range(lowerBound, upperBound, energyResolution));
}

Solver::ChebyshevExpander::Type ChebyshevExpander::toSolverType(
    Property::GreensFunction::Type t) const {
	switch (t) {
	case Property::GreensFunction::Type::Advanced:
		return Solver::ChebyshevExpander::Type::Advanced;
	case Property::GreensFunction::Type::Retarded:
		return Solver::ChebyshevExpander::Type::Retarded;
	case Property::GreensFunction::Type::Principal:
		return Solver::ChebyshevExpander::Type::Principal;
	case Property::GreensFunction::Type::NonPrincipal:
		return Solver::ChebyshevExpander::Type::NonPrincipal;
	}
	TBTKExit("PropertyExtractor::ChebyshevExpander::calculateGreensFunctions()",
	         "Unknown GreensFunction type.", "This should never happen.");
}

Property::GreensFunction ChebyshevExpander::calculateGreensFunction(
    Index to, Index from, Property::GreensFunction::Type type) {
	std::vector<Index> tos = {to};
	return calculateGreensFunctions(tos, from, type);
}

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:234-295, with all
// to-indices contracted in one GPU call.
Property::GreensFunction ChebyshevExpander::calculateGreensFunctions(
    std::vector<Index> &to, Index from, Property::GreensFunction::Type type) {
	std::vector<std::vector<std::complex<double>>> coefficients =
	    solver->calculateCoefficients(to, from);
	IndexTree layout;
	for (const Index &t : to) layout.add(Index({t, from}));
	layout.generateLinearMap();
	Property::GreensFunction gf(layout, type, getEnergyWindow());
	std::vector<std::vector<std::complex<double>>> g =
	    solver->generateGreensFunctions(coefficients, toSolverType(type));
	std::vector<std::complex<double>> &data = gf.getDataRW();
	const unsigned int E = getEnergyWindow().getResolution();
	for (size_t n = 0; n < to.size(); n++) {
		const int offset = gf.getOffset(Index({to[n], from}));
		for (unsigned int c = 0; c < E; c++) data[offset + c] = g[n][c];
	}
	return gf;
}

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:107-232.
Property::GreensFunction ChebyshevExpander::calculateGreensFunction(
    std::vector<std::vector<Index>> patterns, Property::GreensFunction::Type type) {
	IndexTree layout, fromIndices;
	std::set<unsigned int> toSizes;
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	for (size_t n = 0; n < patterns.size(); n++) {
		TBTKAssert(patterns[n].size() == 2, "ChebyshevExpander::calculateGreensFunction()",
		           "Invalid pattern. Each pattern entry needs to contain exactly two Indices, "
		           "but '" << patterns[n].size() << "' found in entry '" << n << "'.", "");
		const Index &toPattern = patterns[n][0], &fromPattern = patterns[n][1];
		toSizes.insert(toPattern.getSize());
		IndexTree toTree = generateIndexTree({toPattern}, has, true, true);
		IndexTree fromTree = generateIndexTree({fromPattern}, has, true, true);
		for (const Index &t : toTree)
			for (const Index &f : fromTree) {
				layout.add(Index({t, f}));
				fromIndices.add(f);
			}
	}
	layout.generateLinearMap();
	fromIndices.generateLinearMap();
	Property::GreensFunction gf(layout, type, getEnergyWindow());
	const unsigned int E = getEnergyWindow().getResolution();
	// Every from-index with its to-indices (in the layout's order), all
	// recurrences in one engine call, all contractions in one more.
	std::vector<Index> froms;
	std::vector<std::vector<Index>> tos;
	for (const Index &from : fromIndices) {
		std::vector<Index> toIndices;
		for (unsigned int size : toSizes) {
			Index toPattern;
			for (unsigned int c = 0; c < size; c++) toPattern.pushBack(IDX_ALL);
			for (const Index &m : layout.getIndexList(Index({toPattern, from}))) {
				Index t;
				for (unsigned int c = 0; c < size; c++) t.pushBack(m.at(c));
				toIndices.push_back(t);
			}
		}
		froms.push_back(from);
		tos.push_back(toIndices);
	}
	std::vector<std::vector<std::vector<std::complex<double>>>> c =
	    solver->calculateCoefficientsMulti(froms, tos);
	std::vector<std::vector<std::complex<double>>> flat;
	for (auto &perFrom : c)
		for (auto &v : perFrom) flat.push_back(std::move(v));
	std::vector<std::vector<std::complex<double>>> g =
	    solver->generateGreensFunctions(flat, toSolverType(type));
	size_t n = 0;
	for (size_t j = 0; j < froms.size(); j++)
		for (const Index &t : tos[j]) {
			const int o = gf.getOffset(Index({t, froms[j]}));
			for (unsigned int e = 0; e < E; e++) gf.getDataRW()[o + e] = g[n][e];
			n++;
		}
	return gf;
}

// Damped diagonal coefficients <i|T_n|i> of the given indices: exact (the
// plain recurrence's bits) through calculateCoefficientsMulti, or the
// approximate FP32 blocks with setUseBlockRecurrence(true).
std::vector<std::vector<std::complex<double>>> ChebyshevExpander::diagonalCoefficients(
    const std::vector<Index> &indices) {
	if (solver->getUseBlockRecurrence()) return solver->calculateCoefficientsBlock(indices);
	std::vector<std::vector<Index>> tos;
	tos.reserve(indices.size());
	for (const Index &idx : indices) tos.push_back({idx});
	std::vector<std::vector<std::vector<std::complex<double>>>> c =
	    solver->calculateCoefficientsMulti(indices, tos);
	std::vector<std::vector<std::complex<double>>> out;
	out.reserve(indices.size());
	for (auto &perFrom : c) out.push_back(std::move(perFrom[0]));
	return out;
}

// The LDOS callback, Lib/src/PropertyExtractor/ChebyshevExpander.cpp:661-684:
// data[offset + n] += Im(G_NonPrincipal(index, index)[n])/pi, in the order of
// the indices. All diagonal moments come from one batched call.
void ChebyshevExpander::accumulateLDOS(const std::vector<Index> &indices,
                                       const std::vector<int> &offsets,
                                       Property::LDOS &ldos) {
	const unsigned int E = getEnergyWindow().getResolution();
	std::vector<std::vector<std::complex<double>>> g = solver->generateGreensFunctions(
	    diagonalCoefficients(indices), Solver::ChebyshevExpander::Type::NonPrincipal);
	std::vector<double> &data = ldos.getDataRW();
	for (size_t i = 0; i < indices.size(); i++)
		for (unsigned int n = 0; n < E; n++) data[offsets[i] + n] += std::imag(g[i][n]) / M_PI;
}

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:475-505.
Property::LDOS ChebyshevExpander::calculateLDOS(std::vector<Index> patterns) {
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	IndexTree allIndices = generateIndexTree(patterns, has, false, true);
	IndexTree layout = generateIndexTree(patterns, has, true, true);
	Property::LDOS ldos(layout, getEnergyWindow());
	std::vector<Index> indices;
	std::vector<int> offsets;
	for (const Index &idx : allIndices) {
		// The spin wildcard of the layout is only a marker here (the LDOS
		// callback does not use it); evaluate the concrete states.
		std::vector<Index> concrete = has.getIndexList({idx});
		for (const Index &c : concrete) {
			indices.push_back(c);
			offsets.push_back(ldos.getOffset(c));
		}
	}
	accumulateLDOS(indices, offsets, ldos);
	return ldos;
}

// Ranges format: the index walk of PropertyExtractor::calculate,
// Lib/include/TBTK/PropertyExtractor/PropertyExtractor.h:780-847, after
// ensureCompliantRanges (PropertyExtractor.cpp:292-303): every concrete index
// with its block offset, in the reference's order (last subindex first,
// IDX_SUM_ALL summed into one block).
void ChebyshevExpander::rangesIndices(Index pattern, Index ranges, int blockSize,
                                      std::vector<Index> &indices,
                                      std::vector<int> &offsets) {
	for (unsigned int n = 0; n < pattern.getSize(); n++)
		if (pattern.at(n) >= 0) ranges.at(n) = 1;
	struct Walker {
		const Index &ranges;
		std::vector<Index> &indices;
		std::vector<int> &offsets;
		void walk(Index pattern, int currentOffset, int offsetMultiplier) {
			int cs = (int)pattern.getSize() - 1;
			for (; cs >= 0; cs--)
				if (pattern.at(cs) < 0) break;
			if (cs == -1) {
				indices.push_back(pattern);
				offsets.push_back(currentOffset);
				return;
			}
			TBTKAssert(pattern.at(cs).isSummationIndex() || pattern.at(cs).isRangeIndex(),
			           "PropertyExtractor::calculate()",
			           "Invalid specifier found at subindex " << cs << ".",
			           "Did you mean IDX_SUM_ALL, IDX_X, IDX_Y, or IDX_Z?");
			const bool isSum = pattern.at(cs).isSummationIndex();
			const int next = isSum ? offsetMultiplier : offsetMultiplier * ranges.at(cs);
			for (int n = 0; n < ranges.at(cs); n++) {
				pattern.at(cs) = n;
				walk(pattern, currentOffset, next);
				if (!isSum) currentOffset += offsetMultiplier;
			}
		}
	} walker{ranges, indices, offsets};
	walker.walk(pattern, 0, blockSize);
}

// getLoopRanges, Lib/src/PropertyExtractor/PropertyExtractor.cpp:305-321.
std::vector<int> ChebyshevExpander::loopRanges(Index pattern, Index ranges) {
	for (unsigned int n = 0; n < pattern.getSize(); n++)
		if (pattern.at(n) >= 0) ranges.at(n) = 1;
	std::vector<int> r;
	for (unsigned int n = 0; n < ranges.getSize(); n++)
		if (pattern.at(n).isRangeIndex()) r.push_back(ranges.at(n));
	return r;
}

Property::LDOS ChebyshevExpander::calculateLDOS(Index pattern, Index ranges) {
	Property::LDOS ldos(loopRanges(pattern, ranges), getEnergyWindow());
	std::vector<Index> indices;
	std::vector<int> offsets;
	rangesIndices(pattern, ranges, (int)getEnergyWindow().getResolution(), indices, offsets);
	accumulateLDOS(indices, offsets, ldos);
	return ldos;
}

Property::DOS ChebyshevExpander::calculateDOS() {
	Property::DOS dos(getEnergyWindow());
	const unsigned int E = getEnergyWindow().getResolution();
	std::vector<double> &data = dos.getDataRW();
	if (stochasticVectors > 0) {
		// Tr delta(E - H) from sum_r <r|T_n|r>/R: the NonPrincipal
		// contraction is linear in the moments, so contracting the trace
		// moments gives sum_i Im G_ii(E).
		// The vectors have |r_i| = 1, so the averaged moments estimate
		// Tr T_n (mu_0 = N exactly).
		std::vector<std::complex<double>> mu =
		    solver->calculateTraceCoefficients(stochasticVectors, stochasticSeed);
		std::vector<std::complex<double>> g = solver->generateGreensFunction(
		    mu, Solver::ChebyshevExpander::Type::NonPrincipal);
		for (unsigned int n = 0; n < E; n++) data[n] = std::imag(g[n]) / M_PI;
		return dos;
	}
	// Exact trace: the LDOS of every basis state.
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	const int N = has.getBasisSize();
	const int chunk = 4096;
	for (int b0 = 0; b0 < N; b0 += chunk) {
		std::vector<Index> idx;
		for (int b = b0; b < N && b < b0 + chunk; b++) idx.push_back(has.getPhysicalIndex(b));
		std::vector<std::vector<std::complex<double>>> g = solver->generateGreensFunctions(
		    diagonalCoefficients(idx), Solver::ChebyshevExpander::Type::NonPrincipal);
		for (const auto &gi : g)
			for (unsigned int n = 0; n < E; n++) data[n] += std::imag(gi[n]) / M_PI;
	}
	return dos;
}

// ----------------------------------------- thermodynamic-equilibrium weights

// PropertyExtractor::getThermodynamicEquilibriumOccupation,
// Lib/include/TBTK/PropertyExtractor/PropertyExtractor.h:910-934.
double ChebyshevExpander::occupation(double energy) const {
	const Model &model = solver->getModel();
	switch (model.getStatistics()) {
	case Statistics::FermiDirac:
		return Functions::fermiDiracDistribution(energy, model.getChemicalPotential(),
		                                         model.getTemperature());
	case Statistics::BoseEinstein:
		return Functions::boseEinsteinDistribution(energy, model.getChemicalPotential(),
		                                           model.getTemperature());
	}
	TBTKExit("PropertyExtractor::getThermodynamicEquilibriumOccupation()",
	         "Unknown statistics.", "This should never happen, contact the developer.");
}

// Diagonal NonPrincipal Green's functions of the given indices.
std::vector<std::vector<std::complex<double>>> ChebyshevExpander::diagonalGreensFunctions(
    const std::vector<Index> &indices) {
	return solver->generateGreensFunctions(diagonalCoefficients(indices),
	                                       Solver::ChebyshevExpander::Type::NonPrincipal);
}

// The four NonPrincipal Green's functions G(to spin n/2, from spin n%2) of
// every index, n = 0..3 (up-up, up-down, down-up, down-down): one start
// vector per (index, from-spin) gathering both to-spins, all in one batched
// call; the moments are the ones separate calls give.
std::vector<std::array<std::vector<std::complex<double>>, 4>>
ChebyshevExpander::spinGreensFunctions(const std::vector<Index> &indices,
                                       const std::vector<int> &spinIndices) {
	std::vector<Index> froms;
	std::vector<std::vector<Index>> tos;
	for (size_t k = 0; k < indices.size(); k++)
		for (int c = 0; c < 2; c++) {
			const Index &index = indices[k];
			const int spinIndex = spinIndices[k];
			Index from(index);
			from.at(spinIndex) = c;
			std::vector<Index> to = {Index(index), Index(index)};
			to[0].at(spinIndex) = 0;
			to[1].at(spinIndex) = 1;
			froms.push_back(from);
			tos.push_back(to);
		}
	std::vector<std::vector<std::vector<std::complex<double>>>> c =
	    solver->calculateCoefficientsMulti(froms, tos);
	std::vector<std::vector<std::complex<double>>> flat;
	for (auto &perFrom : c)
		for (auto &v : perFrom) flat.push_back(std::move(v));
	std::vector<std::vector<std::complex<double>>> gf = solver->generateGreensFunctions(
	    flat, Solver::ChebyshevExpander::Type::NonPrincipal);
	std::vector<std::array<std::vector<std::complex<double>>, 4>> g(indices.size());
	for (size_t k = 0; k < indices.size(); k++)
		for (int c = 0; c < 2; c++) {
			g[k][0 * 2 + c] = gf[4 * k + 2 * c];
			g[k][1 * 2 + c] = gf[4 * k + 2 * c + 1];
		}
	return g;
}

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:297-327.
std::complex<double> ChebyshevExpander::calculateExpectationValue(Index to, Index from) {
	const std::complex<double> i(0, 1);
	std::complex<double> expectationValue = 0.;
	Property::GreensFunction greensFunction =
	    calculateGreensFunction(to, from, Property::GreensFunction::Type::NonPrincipal);
	const std::vector<std::complex<double>> &greensFunctionData = greensFunction.getData();
	const Range &energyWindow = getEnergyWindow();
	const double dE = (energyWindow[1] - energyWindow[0]);
	for (unsigned int e = 0; e < energyWindow.getResolution(); e++) {
		const double weight = occupation(energyWindow[e]);
		expectationValue -= weight * std::conj(i * greensFunctionData[e]) * dE / M_PI;
	}
	return expectationValue;
}

// The density callback, Lib/src/PropertyExtractor/ChebyshevExpander.cpp:
// 575-607: data[offset] += f(E) Im G_NonPrincipal(E)/pi dE, e ascending, in
// the order of the indices.
void ChebyshevExpander::accumulateDensity(const std::vector<Index> &indices,
                                          const std::vector<int> &offsets,
                                          Property::Density &density) {
	std::vector<std::vector<std::complex<double>>> g = diagonalGreensFunctions(indices);
	std::vector<double> &data = density.getDataRW();
	const Range &energyWindow = getEnergyWindow();
	const double dE = energyWindow[1] - energyWindow[0];
	for (size_t k = 0; k < indices.size(); k++)
		for (unsigned int e = 0; e < energyWindow.getResolution(); e++) {
			const double weight = occupation(energyWindow[e]);
			data[offsets[k]] += weight * std::imag(g[k][e]) / M_PI * dE;
		}
}

Property::Density ChebyshevExpander::calculateDensity(Index pattern, Index ranges) {
	Property::Density density(loopRanges(pattern, ranges));
	std::vector<Index> indices;
	std::vector<int> offsets;
	rangesIndices(pattern, ranges, 1, indices, offsets);
	accumulateDensity(indices, offsets, density);
	return density;
}

// Lib/src/PropertyExtractor/ChebyshevExpander.cpp:352-379.
Property::Density ChebyshevExpander::calculateDensity(std::vector<Index> patterns) {
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	IndexTree allIndices = generateIndexTree(patterns, has, false, false);
	IndexTree layout = generateIndexTree(patterns, has, true, true);
	Property::Density density(layout);
	std::vector<Index> indices;
	std::vector<int> offsets;
	for (const Index &idx : allIndices) {
		indices.push_back(idx);
		offsets.push_back(density.getOffset(idx));
	}
	accumulateDensity(indices, offsets, density);
	return density;
}

// The magnetization callback, Lib/src/PropertyExtractor/ChebyshevExpander.
// cpp:609-647: data[offset].at(n/2, n%2) += f(E) (-i) G_n(E)/pi dE.
void ChebyshevExpander::accumulateMagnetization(const std::vector<Index> &indices,
                                                const std::vector<int> &offsets,
                                                const std::vector<int> &spinIndices,
                                                Property::Magnetization &magnetization) {
	const std::complex<double> i(0, 1);
	std::vector<SpinMatrix> &data = magnetization.getDataRW();
	const Range &energyWindow = getEnergyWindow();
	const double dE = energyWindow[1] - energyWindow[0];
	const auto gs = spinGreensFunctions(indices, spinIndices);
	for (size_t k = 0; k < indices.size(); k++) {
		const auto &g = gs[k];
		for (int n = 0; n < 4; n++)
			for (unsigned int e = 0; e < energyWindow.getResolution(); e++) {
				const double weight = occupation(energyWindow[e]);
				data[offsets[k]].at(n / 2, n % 2) += weight * (-i) * g[n][e] / M_PI * dE;
			}
	}
}

// The spin-polarized LDOS callback, Lib/src/PropertyExtractor/
// ChebyshevExpander.cpp:686-717: data[offset + e].at(n/2, n%2) += -i G_n(E)/pi.
void ChebyshevExpander::accumulateSpinPolarizedLDOS(const std::vector<Index> &indices,
                                                    const std::vector<int> &offsets,
                                                    const std::vector<int> &spinIndices,
                                                    Property::SpinPolarizedLDOS &splodos) {
	const std::complex<double> i(0, 1);
	std::vector<SpinMatrix> &data = splodos.getDataRW();
	const unsigned int E = getEnergyWindow().getResolution();
	const auto gs = spinGreensFunctions(indices, spinIndices);
	for (size_t k = 0; k < indices.size(); k++) {
		const auto &g = gs[k];
		for (int n = 0; n < 4; n++)
			for (unsigned int e = 0; e < E; e++)
				data[offsets[k] + e].at(n / 2, n % 2) += -i * g[n][e] / M_PI;
	}
}

// Ranges format with IDX_SPIN: the spin subindex is fixed to 0 in the walk
// and set by the callback (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:
// 381-412, 507-546).
static int takeSpinIndex(Index &pattern, Index &ranges, const char *function) {
	for (unsigned int n = 0; n < pattern.getSize(); n++)
		if (pattern.at(n).isSpinIndex()) {
			pattern.at(n) = 0;
			ranges.at(n) = 1;
			return (int)n;
		}
	TBTKExit(function, "No spin index indicated.",
	         "Use IDX_SPIN to indicate the position of the spin index.");
}

// Custom format: the position of IDX_SPIN in an index of allIndices (kept as
// a wildcard), as PropertyExtractor::calculate finds it in the memory layout
// (PropertyExtractor.h:849-894).
static int spinPosition(const Index &index, const char *function) {
	int pos = -1;
	for (unsigned int n = 0; n < index.getSize(); n++)
		if (index.at(n).isSpinIndex()) {
			TBTKAssert(pos == -1, "PropertyExtractor::calculate()", "Several spin indeces found.",
			           "Use IDX_SPIN at most once per pattern to indicate spin index.");
			pos = (int)n;
		}
	if (pos == -1)
		TBTKExit(function, "No spin index indicated.",
		         "Use IDX_SPIN to indicate the position of the spin index.");
	return pos;
}

Property::Magnetization ChebyshevExpander::calculateMagnetization(Index pattern, Index ranges) {
	const int spinIndex = takeSpinIndex(
	    pattern, ranges, "PropertyExtractor::ChebyshevExpander::calculateMagnetization()");
	Property::Magnetization magnetization(loopRanges(pattern, ranges));
	std::vector<Index> indices;
	std::vector<int> offsets;
	rangesIndices(pattern, ranges, 1, indices, offsets);
	accumulateMagnetization(indices, offsets, std::vector<int>(indices.size(), spinIndex),
	                        magnetization);
	return magnetization;
}

Property::Magnetization ChebyshevExpander::calculateMagnetization(std::vector<Index> patterns) {
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	IndexTree allIndices = generateIndexTree(patterns, has, false, true);
	IndexTree layout = generateIndexTree(patterns, has, true, true);
	Property::Magnetization magnetization(layout);
	std::vector<Index> indices;
	std::vector<int> offsets, spins;
	for (const Index &idx : allIndices) {
		spins.push_back(
		    spinPosition(idx, "PropertyExtractor::ChebyshevExpander::calculateMagnetization()"));
		indices.push_back(idx);
		offsets.push_back(magnetization.getOffset(idx));
	}
	accumulateMagnetization(indices, offsets, spins, magnetization);
	return magnetization;
}

Property::SpinPolarizedLDOS ChebyshevExpander::calculateSpinPolarizedLDOS(Index pattern,
                                                                           Index ranges) {
	const int spinIndex = takeSpinIndex(
	    pattern, ranges, "PropertyExtractor::ChebsyhevExpander::calculateSpinPolarizedLDOS()");
	Property::SpinPolarizedLDOS splodos(loopRanges(pattern, ranges), getEnergyWindow());
	std::vector<Index> indices;
	std::vector<int> offsets;
	rangesIndices(pattern, ranges, (int)getEnergyWindow().getResolution(), indices, offsets);
	accumulateSpinPolarizedLDOS(indices, offsets, std::vector<int>(indices.size(), spinIndex),
	                            splodos);
	return splodos;
}

Property::SpinPolarizedLDOS ChebyshevExpander::calculateSpinPolarizedLDOS(
    std::vector<Index> patterns) {
	const HoppingAmplitudeSet &has = solver->getModel().getHoppingAmplitudeSet();
	IndexTree allIndices = generateIndexTree(patterns, has, false, true);
	IndexTree layout = generateIndexTree(patterns, has, true, true);
	Property::SpinPolarizedLDOS splodos(layout, getEnergyWindow());
	std::vector<Index> indices;
	std::vector<int> offsets, spins;
	for (const Index &idx : allIndices) {
		spins.push_back(spinPosition(
		    idx, "PropertyExtractor::ChebsyhevExpander::calculateSpinPolarizedLDOS()"));
		indices.push_back(idx);
		offsets.push_back(splodos.getOffset(idx));
	}
	accumulateSpinPolarizedLDOS(indices, offsets, spins, splodos);
	return splodos;
}

}  // namespace PropertyExtractor
}  // namespace TBTK
