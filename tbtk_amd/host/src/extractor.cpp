// Drop-in PropertyExtractor::ChebyshevExpander and the PropertyExtractor
// base behaviour it uses.
#include "TBTK/PropertyExtractor/ChebyshevExpander.h"
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
		Property::GreensFunction part = calculateGreensFunctions(toIndices, from, type);
		for (const Index &t : toIndices) {
			const Index compound({t, from});
			const int o = gf.getOffset(compound), op = part.getOffset(compound);
			for (unsigned int c = 0; c < E; c++) gf.getDataRW()[o + c] = part.getData()[op + c];
		}
	}
	return gf;
}

// The LDOS callback, Lib/src/PropertyExtractor/ChebyshevExpander.cpp:661-684:
// data[offset + n] += Im(G_NonPrincipal(index, index)[n])/pi, in the order of
// the indices. With the block recurrence all diagonal moments come from one
// batched run.
void ChebyshevExpander::accumulateLDOS(const std::vector<Index> &indices,
                                       const std::vector<int> &offsets,
                                       Property::LDOS &ldos) {
	const unsigned int E = getEnergyWindow().getResolution();
	std::vector<std::vector<std::complex<double>>> coefficients;
	if (solver->getUseBlockRecurrence()) {
		coefficients = solver->calculateCoefficientsBlock(indices);
	} else {
		coefficients.reserve(indices.size());
		for (const Index &idx : indices)
			coefficients.push_back(solver->calculateCoefficients(idx, idx));
	}
	std::vector<std::vector<std::complex<double>>> g = solver->generateGreensFunctions(
	    coefficients, Solver::ChebyshevExpander::Type::NonPrincipal);
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

// Ranges format: PropertyExtractor::calculate, Lib/include/TBTK/
// PropertyExtractor/PropertyExtractor.h:780-847, with ensureCompliantRanges
// and getLoopRanges (PropertyExtractor.cpp:292-321).
Property::LDOS ChebyshevExpander::calculateLDOS(Index pattern, Index ranges) {
	for (unsigned int n = 0; n < pattern.getSize(); n++)
		if (pattern.at(n) >= 0) ranges.at(n) = 1;
	std::vector<int> loopRanges;
	for (unsigned int n = 0; n < ranges.getSize(); n++)
		if (pattern.at(n).isRangeIndex()) loopRanges.push_back(ranges.at(n));
	Property::LDOS ldos(loopRanges, getEnergyWindow());

	std::vector<Index> indices;
	std::vector<int> offsets;
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
	walker.walk(pattern, 0, (int)getEnergyWindow().getResolution());
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
		std::vector<std::vector<std::complex<double>>> c =
		    solver->getUseBlockRecurrence() ? solver->calculateCoefficientsBlock(idx)
		                                    : std::vector<std::vector<std::complex<double>>>();
		if (!solver->getUseBlockRecurrence())
			for (const Index &i : idx) c.push_back(solver->calculateCoefficients(i, i));
		std::vector<std::vector<std::complex<double>>> g = solver->generateGreensFunctions(
		    c, Solver::ChebyshevExpander::Type::NonPrincipal);
		for (const auto &gi : g)
			for (unsigned int n = 0; n < E; n++) data[n] += std::imag(gi[n]) / M_PI;
	}
	return dos;
}

}  // namespace PropertyExtractor
}  // namespace TBTK
