/* The part of TBTK's model layer the Chebyshev path reads: HoppingAmplitude
 * (with + HC and amplitude callbacks), HoppingAmplitudeSet (basis numbering
 * and the CSR of H) and Model. Behaviour follows
 * Lib/include/Core/TBTK/HoppingAmplitude.h, HoppingAmplitudeSet.h:397-424,
 * Lib/src/Core/HoppingAmplitudeTree.cpp:240-616 and Model.h. */
#ifndef TBTK_AMD_MODEL
#define TBTK_AMD_MODEL

#include "TBTK/Index.h"
#include "TBTK/Streams.h"

#include <complex>
#include <cstdint>
#include <map>
#include <vector>

namespace TBTK {

class HoppingAmplitude {
public:
	/* Amplitude given by a function of (to, from), as
	 * HoppingAmplitude::AmplitudeCallback in the reference. */
	class AmplitudeCallback {
	public:
		virtual ~AmplitudeCallback() {}
		virtual std::complex<double> getHoppingAmplitude(const Index &to,
		                                                 const Index &from) const = 0;
	};

	HoppingAmplitude(std::complex<double> amplitude, Index to, Index from);
	HoppingAmplitude(const AmplitudeCallback &callback, Index to, Index from);

	std::complex<double> getAmplitude() const;
	const Index &getToIndex() const { return to; }
	const Index &getFromIndex() const { return from; }
	HoppingAmplitude getHermitianConjugate() const;

private:
	std::complex<double> amplitude;
	const AmplitudeCallback *callback;
	Index to, from;
};

enum HermitianConjugate { HC };

class HoppingAmplitudeList {
public:
	void pushBack(const HoppingAmplitude &ha) { list.push_back(ha); }
	const std::vector<HoppingAmplitude> &get() const { return list; }

private:
	std::vector<HoppingAmplitude> list;
};

/* ha + HC: the amplitude and its Hermitian conjugate. */
HoppingAmplitudeList operator+(const HoppingAmplitude &ha, HermitianConjugate);

/* H in compressed sparse row form: row = basis(to), column = basis(from),
 * columns ascending within a row, duplicate (to, from) pairs summed. */
struct SparseMatrixCSR {
	int64_t numRows = 0;
	std::vector<uint32_t> rowPointers;
	std::vector<uint32_t> columns;
	std::vector<std::complex<double>> values;
};

class HoppingAmplitudeSet {
public:
	void add(const HoppingAmplitude &ha);
	void construct();
	bool getIsConstructed() const { return constructed; }

	int getBasisSize() const { return (int)basis.size(); }
	/* Basis number of a physical index: depth-first order over subindex
	 * values among the indices that appear as from-index
	 * (HoppingAmplitudeTree::generateBasisIndices). -1 if absent. */
	int getBasisIndex(const Index &index) const;
	const Index &getPhysicalIndex(int basisIndex) const;
	/* All basis indices matching any pattern (IDX_ALL etc. as wildcards),
	 * in basis order. */
	std::vector<Index> getIndexList(const std::vector<Index> &patterns) const;
	const std::vector<HoppingAmplitude> &getHoppingAmplitudes() const { return amps; }

	/* CSR of H. Duplicates are summed right to left in insertion order,
	 * a + (b + c), which is the order SparseMatrix::mergeLIL
	 * (Lib/include/Utilities/TBTK/SparseMatrix.h:1400-1420) produces when
	 * its sorts keep insertion order. */
	SparseMatrixCSR getSparseMatrix() const;

	/* Changes every time construct() runs (used to detect a new model). */
	uint64_t getGeneration() const { return generation; }

private:
	std::vector<HoppingAmplitude> amps;
	std::map<Index, int> basisOf;
	std::vector<Index> basis;
	bool constructed = false;
	uint64_t generation = 0;
};

enum class Statistics { FermiDirac, BoseEinstein };

class Model : public Communicator {
public:
	Model() : Communicator(true) {}
	Model &operator<<(const HoppingAmplitude &ha);
	Model &operator<<(const HoppingAmplitudeList &list);
	void construct();
	int getBasisSize() const { return hoppingAmplitudeSet.getBasisSize(); }
	const HoppingAmplitudeSet &getHoppingAmplitudeSet() const {
		return hoppingAmplitudeSet;
	}
	void setChemicalPotential(double mu) { chemicalPotential = mu; }
	double getChemicalPotential() const { return chemicalPotential; }
	void setTemperature(double t) { temperature = t; }
	double getTemperature() const { return temperature; }
	void setStatistics(Statistics s) { statistics = s; }
	Statistics getStatistics() const { return statistics; }

private:
	HoppingAmplitudeSet hoppingAmplitudeSet;
	double chemicalPotential = 0;
	double temperature = 0;
	Statistics statistics = Statistics::FermiDirac;
};

}  // namespace TBTK
#endif
