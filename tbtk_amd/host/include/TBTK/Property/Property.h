/* Property containers the Chebyshev path returns: GreensFunction, LDOS,
 * DOS, Density, Magnetization and SpinPolarizedLDOS. Memory layout = (linear index of the
 * IndexTree entry, or the Ranges-format offset) x energy resolution, as
 * Lib/include/TBTK/Property/AbstractProperty.h:107-147 and
 * EnergyResolvedProperty.h:914-932 (getDeltaE). */
#ifndef TBTK_AMD_PROPERTY
#define TBTK_AMD_PROPERTY

#include "TBTK/IndexTree.h"
#include "TBTK/Range.h"
#include "TBTK/SpinMatrix.h"
#include "TBTK/TBTKMacros.h"

#include <complex>
#include <vector>

namespace TBTK {
namespace Property {

class Property {
public:
	virtual ~Property() {}
};

template <typename T> class AbstractProperty : public Property {
public:
	enum class IndexDescriptorFormat { None, Ranges, Custom };

	/* One block, no index (e.g. DOS). */
	explicit AbstractProperty(unsigned int blockSize)
	    : format(IndexDescriptorFormat::None), blockSize(blockSize),
	      data(blockSize, T(0)) {}
	/* Ranges format: a dense box of the given extents. */
	AbstractProperty(const std::vector<int> &ranges, unsigned int blockSize)
	    : format(IndexDescriptorFormat::Ranges), ranges(ranges), blockSize(blockSize) {
		size_t n = 1;
		for (int r : ranges) n *= (size_t)r;
		data.assign(n * blockSize, T(0));
	}
	/* Custom format: one block per IndexTree entry. */
	AbstractProperty(const IndexTree &tree, unsigned int blockSize)
	    : format(IndexDescriptorFormat::Custom), tree(tree), blockSize(blockSize),
	      data((size_t)tree.getSize() * blockSize, T(0)) {}

	std::vector<T> &getDataRW() { return data; }
	const std::vector<T> &getData() const { return data; }
	unsigned int getBlockSize() const { return blockSize; }
	unsigned int getSize() const { return (unsigned int)data.size(); }
	IndexDescriptorFormat getIndexDescriptorFormat() const { return format; }
	const IndexTree &getIndexTree() const { return tree; }

	/* Offset of the block of `index` (Custom format; wildcards in the
	 * stored layout match any value). */
	int getOffset(const Index &index) const {
		TBTKAssert(format == IndexDescriptorFormat::Custom,
		           "AbstractProperty::getOffset()",
		           "The index descriptor is not of the Custom format.", "");
		const int li = tree.getLinearIndex(index, IndexTree::SearchMode::MatchWildcards, true);
		TBTKAssert(li >= 0, "AbstractProperty::getOffset()",
		           "Index " << index.toString() << " is not in the property.", "");
		return li * (int)blockSize;
	}
	T &operator()(const Index &index, unsigned int n) { return data[getOffset(index) + n]; }
	const T &operator()(const Index &index, unsigned int n) const {
		return data[getOffset(index) + n];
	}
	T &operator()(unsigned int n) { return data[n]; }
	const T &operator()(unsigned int n) const { return data[n]; }

protected:
	IndexDescriptorFormat format;
	IndexTree tree;
	std::vector<int> ranges;
	unsigned int blockSize;
	std::vector<T> data;
};

template <typename T> class EnergyResolvedProperty : public AbstractProperty<T> {
public:
	EnergyResolvedProperty(const Range &window)
	    : AbstractProperty<T>(window.getResolution()), window(window) {}
	EnergyResolvedProperty(const std::vector<int> &ranges, const Range &window)
	    : AbstractProperty<T>(ranges, window.getResolution()), window(window) {}
	EnergyResolvedProperty(const IndexTree &tree, const Range &window)
	    : AbstractProperty<T>(tree, window.getResolution()), window(window) {}

	double getLowerBound() const { return window[0]; }
	double getUpperBound() const { return window.getLast(); }
	unsigned int getResolution() const { return window.getResolution(); }
	double getEnergy(unsigned int n) const { return window[n]; }
	double getDeltaE() const {
		if (window.getResolution() == 1) return 0;
		return (window.getLast() - window[0]) / (window.getResolution() - 1);
	}

private:
	Range window;
};

class GreensFunction : public EnergyResolvedProperty<std::complex<double>> {
public:
	enum class Type { Advanced, Retarded, Principal, NonPrincipal };
	GreensFunction(const IndexTree &tree, Type type, const Range &window)
	    : EnergyResolvedProperty<std::complex<double>>(tree, window), type(type) {}
	Type getType() const { return type; }

private:
	Type type;
};

class LDOS : public EnergyResolvedProperty<double> {
public:
	LDOS(const std::vector<int> &ranges, const Range &window)
	    : EnergyResolvedProperty<double>(ranges, window) {}
	LDOS(const IndexTree &tree, const Range &window)
	    : EnergyResolvedProperty<double>(tree, window) {}
};

/* Density of states, normalised so that sum_n DOS(n) dE = basis size
 * (Lib/include/TBTK/Property/DOS.h:40). */
class DOS : public EnergyResolvedProperty<double> {
public:
	explicit DOS(const Range &window) : EnergyResolvedProperty<double>(window) {}
};

/* Occupation per index (one double per block), as
 * Lib/include/TBTK/Property/Density.h. */
class Density : public AbstractProperty<double> {
public:
	explicit Density(const std::vector<int> &ranges) : AbstractProperty<double>(ranges, 1) {}
	explicit Density(const IndexTree &tree) : AbstractProperty<double>(tree, 1) {}
};

/* Spin density matrix per index, as Lib/include/TBTK/Property/Magnetization.h. */
class Magnetization : public AbstractProperty<SpinMatrix> {
public:
	explicit Magnetization(const std::vector<int> &ranges)
	    : AbstractProperty<SpinMatrix>(ranges, 1) {}
	explicit Magnetization(const IndexTree &tree) : AbstractProperty<SpinMatrix>(tree, 1) {}
};

/* Energy-resolved spin density matrix per index, as
 * Lib/include/TBTK/Property/SpinPolarizedLDOS.h. */
class SpinPolarizedLDOS : public EnergyResolvedProperty<SpinMatrix> {
public:
	SpinPolarizedLDOS(const std::vector<int> &ranges, const Range &window)
	    : EnergyResolvedProperty<SpinMatrix>(ranges, window) {}
	SpinPolarizedLDOS(const IndexTree &tree, const Range &window)
	    : EnergyResolvedProperty<SpinMatrix>(tree, window) {}
};

}  // namespace Property
}  // namespace TBTK
#endif
