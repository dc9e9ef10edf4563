/* Property containers the Chebyshev path returns: GreensFunction, LDOS,
 * DOS, Density, Magnetization and SpinPolarizedLDOS. Memory layout = (linear index of the
 * IndexTree entry, or the Ranges-format offset) x energy resolution, as
 * Lib/include/TBTK/Property/AbstractProperty.h:107-147 and
 * EnergyResolvedProperty.h:914-932 (getDeltaE). */
#ifndef TBTK_AMD_PROPERTY
#define TBTK_AMD_PROPERTY

#include "TBTK/IndexTree.h"
#include "TBTK/Json.h"
#include "TBTK/Range.h"
#include "TBTK/Serializable.h"
#include "TBTK/SpinMatrix.h"
#include "TBTK/TBTKMacros.h"

#include <complex>
#include <string>
#include <vector>

namespace TBTK {
namespace Property {

/* Pieces of the JSON serializations (src/serialize.cpp). Elements are
 * written as the reference writes them, through a std::stringstream with
 * its default precision (Lib/include/Utilities/TBTK/Serializable.h:464-524;
 * SpinMatrix elements are not supported there or here). */
namespace detail {
std::string serializeElement(double d);
std::string serializeElement(const std::complex<double> &c);
std::string serializeElement(const SpinMatrix &m);
void deserializeElement(const std::string &s, double *d);
void deserializeElement(const std::string &s, std::complex<double> *c);
void deserializeElement(const std::string &s, SpinMatrix *m);
JsonValue indexDescriptorJson(int format, const std::vector<int> &ranges, const IndexTree &tree);
int readIndexDescriptor(const JsonValue &j, std::vector<int> &ranges, IndexTree &tree);
JsonValue parseProperty(const std::string &text, const char *id, const char *function);
}  // namespace detail

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

	/* From the "abstractProperty" object of a serialization
	 * (AbstractProperty.h:1121-1171). */
	explicit AbstractProperty(const JsonValue &j) {
		TBTKAssert(j.at("id").getString() == "AbstractProperty",
		           "AbstractProperty::AbstractProperty()",
		           "Unable to parse string as AbstractProperty.", "");
		format = (IndexDescriptorFormat)detail::readIndexDescriptor(j.at("indexDescriptor"),
		                                                            ranges, tree);
		blockSize = (unsigned int)j.at("blockSize").getInt();
		if (j.has("data"))
			for (const JsonValue &e : j.at("data").elements()) {
				T v = T(0);
				detail::deserializeElement(e.getString(), &v);
				data.push_back(v);
			}
	}
	/* The "abstractProperty" object (AbstractProperty.h:922-955). The
	 * default value is 0 (the reference writes its defaultValue member,
	 * which some of its constructors leave uninitialized). */
	JsonValue toJson() const {
		JsonValue j;
		j["id"] = "AbstractProperty";
		j["indexDescriptor"] = detail::indexDescriptorJson((int)format, ranges, tree);
		j["blockSize"] = blockSize;
		for (const T &x : data) j["data"].push_back(detail::serializeElement(x));
		j["allowIndexOutOfBoundsAccess"] = false;
		j["defaultValue"] = detail::serializeElement(T(0));
		return j;
	}

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
	/* From the "energyResolvedProperty" object (real energies only, the
	 * only kind the Chebyshev path produces). */
	explicit EnergyResolvedProperty(const JsonValue &j)
	    : AbstractProperty<T>(j.at("abstractProperty")),
	      window(j.at("energyWindow").dump(), Serializable::Mode::JSON) {
		TBTKAssert(j.at("energyType").getString() == "Real",
		           "EnergyResolvedProperty::EnergyResolvedProperty()",
		           "Only real energies are supported.", "");
	}
	/* EnergyResolvedProperty.h:1156-1210, EnergyType::Real. */
	JsonValue toJson() const {
		JsonValue j;
		j["id"] = "EnergyResolvedProperty";
		j["energyType"] = "Real";
		j["energyWindow"] = JsonValue::parse(window.serialize(Serializable::Mode::JSON));
		j["abstractProperty"] = AbstractProperty<T>::toJson();
		return j;
	}
	std::string serialize(Serializable::Mode mode) const {
		TBTKAssert(mode == Serializable::Mode::JSON, "EnergyResolvedProperty::serialize()",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		return toJson().dump();
	}

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
	/* Lib/src/Property/LDOS.cpp:76-94, 121-140. */
	LDOS(const std::string &serialization, Serializable::Mode mode)
	    : EnergyResolvedProperty<double>(
	          detail::parseProperty(checkMode(serialization, mode), "LDOS", "LDOS::LDOS()")
	              .at("energyResolvedProperty")) {}
	std::string serialize(Serializable::Mode mode) const {
		checkMode("", mode);
		JsonValue j;
		j["id"] = "LDOS";
		j["energyResolvedProperty"] = toJson();
		return j.dump();
	}

private:
	static const std::string &checkMode(const std::string &s, Serializable::Mode mode) {
		TBTKAssert(mode == Serializable::Mode::JSON, "LDOS",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		return s;
	}
};

/* Density of states, normalised so that sum_n DOS(n) dE = basis size
 * (Lib/include/TBTK/Property/DOS.h:40). */
class DOS : public EnergyResolvedProperty<double> {
public:
	explicit DOS(const Range &window) : EnergyResolvedProperty<double>(window) {}
	/* Lib/src/Property/DOS.cpp:54-72, 99-118. */
	DOS(const std::string &serialization, Serializable::Mode mode)
	    : EnergyResolvedProperty<double>(
	          detail::parseProperty(checkMode(serialization, mode), "DOS", "DOS::DOS()")
	              .at("energyResolvedProperty")) {}
	std::string serialize(Serializable::Mode mode) const {
		checkMode("", mode);
		JsonValue j;
		j["id"] = "DOS";
		j["energyResolvedProperty"] = toJson();
		return j.dump();
	}

private:
	static const std::string &checkMode(const std::string &s, Serializable::Mode mode) {
		TBTKAssert(mode == Serializable::Mode::JSON, "DOS",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		return s;
	}
};

/* Occupation per index (one double per block), as
 * Lib/include/TBTK/Property/Density.h. */
class Density : public AbstractProperty<double> {
public:
	explicit Density(const std::vector<int> &ranges) : AbstractProperty<double>(ranges, 1) {}
	explicit Density(const IndexTree &tree) : AbstractProperty<double>(tree, 1) {}
	/* Lib/src/Property/Density.cpp:66-90, 112-131. */
	Density(const std::string &serialization, Serializable::Mode mode)
	    : AbstractProperty<double>(
	          detail::parseProperty(checkMode(serialization, mode), "Density",
	                                "Density::Density()")
	              .at("abstractProperty")) {}
	std::string serialize(Serializable::Mode mode) const {
		checkMode("", mode);
		JsonValue j;
		j["id"] = "Density";
		j["abstractProperty"] = toJson();
		return j.dump();
	}

private:
	static const std::string &checkMode(const std::string &s, Serializable::Mode mode) {
		TBTKAssert(mode == Serializable::Mode::JSON, "Density",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		return s;
	}
};

/* Spin density matrix per index, as Lib/include/TBTK/Property/Magnetization.h. */
class Magnetization : public AbstractProperty<SpinMatrix> {
public:
	explicit Magnetization(const std::vector<int> &ranges)
	    : AbstractProperty<SpinMatrix>(ranges, 1) {}
	explicit Magnetization(const IndexTree &tree) : AbstractProperty<SpinMatrix>(tree, 1) {}
	/* Lib/src/Property/Magnetization.cpp:81-100: exits on the SpinMatrix
	 * elements, as the reference does. */
	std::string serialize(Serializable::Mode mode) const {
		TBTKAssert(mode == Serializable::Mode::JSON, "Magnetization::serialize()",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		JsonValue j;
		j["id"] = "Magnetization";
		j["abstractProperty"] = toJson();
		return j.dump();
	}
};

/* Energy-resolved spin density matrix per index, as
 * Lib/include/TBTK/Property/SpinPolarizedLDOS.h. */
class SpinPolarizedLDOS : public EnergyResolvedProperty<SpinMatrix> {
public:
	SpinPolarizedLDOS(const std::vector<int> &ranges, const Range &window)
	    : EnergyResolvedProperty<SpinMatrix>(ranges, window) {}
	SpinPolarizedLDOS(const IndexTree &tree, const Range &window)
	    : EnergyResolvedProperty<SpinMatrix>(tree, window) {}
	/* Lib/src/Property/SpinPolarizedLDOS.cpp:123-142: exits on the
	 * SpinMatrix elements, as the reference does. */
	std::string serialize(Serializable::Mode mode) const {
		TBTKAssert(mode == Serializable::Mode::JSON, "SpinPolarizedLDOS::serialize()",
		           "Only Serializable::Mode::JSON is supported yet.", "");
		JsonValue j;
		j["id"] = "SpinPolarizedLDOS";
		j["energyResolvedProperty"] = toJson();
		return j.dump();
	}
};

}  // namespace Property
}  // namespace TBTK
#endif
