/* JSON serialization of Range, IndexTree and the property containers, in
 * the form of the reference's serialize(Serializable::Mode::JSON) output
 * (the object layouts of Lib/src/Utilities/Range.cpp:94-115,
 * Lib/src/Core/IndexTree.cpp:606-650, Lib/src/Property/IndexDescriptor.cpp:
 * 482-538, Lib/include/TBTK/Property/AbstractProperty.h:922-955 and
 * EnergyResolvedProperty.h:1156-1210), and the matching deserializing
 * constructors. */
#include "TBTK/Json.h"
#include "TBTK/Property/Property.h"
#include "TBTK/Range.h"

#include <sstream>

namespace TBTK {

namespace {

void requireJSON(Serializable::Mode mode, const char *function) {
	TBTKAssert(mode == Serializable::Mode::JSON, function,
	           "Only Serializable::Mode::JSON is supported yet.", "");
}

// Wildcard subindices as the reference encodes them
// (Lib/include/Core/TBTK/Subindex.h:395-398).
constexpr int kRefAll = (int)(0xBFFFFFFFu & ~0x20000000u);
constexpr int kRefSumAll = (int)(0xBFFFFFFFu & ~0x20000001u);
constexpr int kRefSpin = (int)(0xBFFFFFFFu & ~0x20000002u);

int toReferenceWildcard(int s) {
	if (s == (int)IDX_ALL) return kRefAll;
	if (s == (int)IDX_SUM_ALL) return kRefSumAll;
	if (s == (int)IDX_SPIN) return kRefSpin;
	TBTKExit("IndexTree::serialize()", "Unsupported wildcard " << s << ".", "");
}

int fromReferenceWildcard(int s) {
	if (s == kRefAll) return (int)IDX_ALL;
	if (s == kRefSumAll) return (int)IDX_SUM_ALL;
	if (s == kRefSpin) return (int)IDX_SPIN;
	TBTKExit("IndexTree::IndexTree()", "Unsupported wildcard type " << s << ".", "");
}

// One node of the reference's tree: children per subindex value (a
// wildcard occupies child 0), a separator marks the node where the next
// component of a compound index starts.
struct Node {
	std::vector<Node> children;
	bool indexIncluded = false, wildcardIndex = false, indexSeparator = false;
	int wildcardType = 0, linearIndex = -1, size = -1;

	void add(const Index &index, unsigned int s) {
		if (s == index.getSize()) {
			indexIncluded = true;
			return;
		}
		int c = index.at(s);
		if (index.at(s).isIndexSeparator()) {
			indexSeparator = true;
			add(index, s + 1);
			return;
		}
		if (c < 0) {
			wildcardIndex = true;
			wildcardType = toReferenceWildcard(c);
			c = 0;
		}
		if (c >= (int)children.size()) children.resize(c + 1);
		children[c].add(index, s + 1);
	}
	int linearMap(int i) {
		if (children.empty()) {
			if (!indexIncluded) return i;
			linearIndex = i;
			return i + 1;
		}
		for (Node &n : children) i = n.linearMap(i);
		return i;
	}
	JsonValue json() const {
		JsonValue j;
		j["id"] = "IndexTree";
		for (const Node &n : children) j["children"].push_back(n.json());
		j["indexIncluded"] = indexIncluded;
		j["wildcardIndex"] = wildcardIndex;
		j["wildcardType"] = wildcardType;
		j["indexSeparator"] = indexSeparator;
		j["linearIndex"] = linearIndex;
		j["size"] = size;
		return j;
	}
};

// The included indices of a serialized tree, depth first.
void collect(const JsonValue &node, std::vector<Subindex> &prefix, std::vector<Index> &out) {
	const bool hasChildren = node.has("children") && !node.at("children").elements().empty();
	if (!hasChildren) {
		if (node.at("indexIncluded").getBool()) out.push_back(Index(prefix));
		return;
	}
	const size_t base = prefix.size();
	if (node.at("indexSeparator").getBool()) prefix.push_back(IDX_SEPARATOR);
	const bool wildcard = node.at("wildcardIndex").getBool();
	const std::vector<JsonValue> &ch = node.at("children").elements();
	for (size_t k = 0; k < ch.size(); k++) {
		prefix.push_back(wildcard ? fromReferenceWildcard((int)node.at("wildcardType").getInt())
		                          : (int)k);
		collect(ch[k], prefix, out);
		prefix.pop_back();
	}
	prefix.resize(base);
}

JsonValue parseWithId(const std::string &text, const char *id, const char *function) {
	JsonValue j = JsonValue::parse(text);
	TBTKAssert(j.has("id") && j.at("id").getString() == id, function,
	           "Unable to parse string as " << id << ".", "");
	return j;
}

}  // namespace

Range::Range(const std::string &serialization, Serializable::Mode mode) {
	requireJSON(mode, "Range::Range()");
	const JsonValue j = parseWithId(serialization, "Range", "Range::Range()");
	start = j.at("start").getDouble();
	dx = j.at("dx").getDouble();
	resolution = (unsigned int)j.at("resolution").getInt();
}

std::string Range::serialize(Serializable::Mode mode) const {
	requireJSON(mode, "Range::serialize()");
	JsonValue j;
	j["id"] = "Range";
	j["start"] = start;
	j["dx"] = dx;
	j["resolution"] = resolution;
	return j.dump();
}

IndexTree::IndexTree(const std::string &serialization, Serializable::Mode mode) {
	requireJSON(mode, "IndexTree::IndexTree()");
	const JsonValue j = parseWithId(serialization, "IndexTree", "IndexTree::IndexTree()");
	std::vector<Subindex> prefix;
	std::vector<Index> included;
	collect(j, prefix, included);
	for (const Index &i : included) add(i);
	generateLinearMap();
}

std::string IndexTree::serialize(Serializable::Mode mode) const {
	requireJSON(mode, "IndexTree::serialize()");
	Node root;
	for (const Index &i : ordered) root.add(i, 0);
	root.size = root.linearMap(0);
	return root.json().dump();
}

namespace Property {
namespace detail {

std::string serializeElement(double d) {
	std::stringstream ss;
	ss << d;
	return ss.str();
}

std::string serializeElement(const std::complex<double> &c) {
	std::stringstream ss;
	ss << c;
	return ss.str();
}

std::string serializeElement(const SpinMatrix &) {
	// The reference has no serialization of SpinMatrix elements either
	// (Lib/include/Utilities/TBTK/Serializable.h:252-254).
	TBTKExit("Serializable::serialize<SpinMatrix>()", "Not yet implemented.", "");
}

void deserializeElement(const std::string &s, double *d) {
	std::stringstream ss(s);
	ss >> *d;
}

void deserializeElement(const std::string &s, std::complex<double> *c) {
	std::stringstream ss(s);
	ss >> *c;
}

void deserializeElement(const std::string &, SpinMatrix *) {
	TBTKExit("Serializable::deserialize<SpinMatrix>()", "Not yet implemented.", "");
}

JsonValue indexDescriptorJson(int format, const std::vector<int> &ranges, const IndexTree &tree) {
	JsonValue j;
	j["id"] = "IndexDescriptor";
	switch (format) {
	case 0:
		j["format"] = "None";
		break;
	case 1:
		j["format"] = "Ranges";
		j["dimensions"] = (unsigned int)ranges.size();
		for (int r : ranges) j["ranges"].push_back(r);
		break;
	default:
		j["format"] = "Custom";
		j["indexTree"] = JsonValue::parse(tree.serialize(Serializable::Mode::JSON));
	}
	return j;
}

int readIndexDescriptor(const JsonValue &j, std::vector<int> &ranges, IndexTree &tree) {
	TBTKAssert(j.at("id").getString() == "IndexDescriptor", "IndexDescriptor::IndexDescriptor()",
	           "Unable to parse string as IndexDescriptor.", "");
	const std::string &f = j.at("format").getString();
	if (f == "None") return 0;
	if (f == "Ranges") {
		ranges.clear();
		if (j.has("ranges"))
			for (const JsonValue &r : j.at("ranges").elements()) ranges.push_back((int)r.getInt());
		TBTKAssert((int64_t)ranges.size() == j.at("dimensions").getInt(),
		           "IndexDescriptor::IndexDescriptor()", "Inconsistent Ranges format.", "");
		return 1;
	}
	TBTKAssert(f == "Custom", "IndexDescriptor::IndexDescriptor()",
	           "Unsupported format '" << f << "'.", "");
	tree = IndexTree(j.at("indexTree").dump(), Serializable::Mode::JSON);
	return 2;
}

JsonValue parseProperty(const std::string &text, const char *id, const char *function) {
	return parseWithId(text, id, function);
}

}  // namespace detail
}  // namespace Property

}  // namespace TBTK
