/* Ordered set of (possibly compound, possibly wildcarded) Indices with a
 * linear numbering: the memory layout of properties. Order and matching
 * follow Lib/src/Core/IndexTree.cpp:166-454: indices are ordered
 * depth-first by subindex value, a wildcard subindex (IDX_SUM_ALL,
 * IDX_SPIN, IDX_ALL) occupies its level alone, and with MatchWildcards a
 * query index matches a stored index whose wildcard positions hold any
 * value. */
#ifndef TBTK_AMD_INDEX_TREE
#define TBTK_AMD_INDEX_TREE

#include "TBTK/Index.h"
#include "TBTK/Serializable.h"

#include <map>
#include <string>
#include <set>
#include <vector>

namespace TBTK {

class IndexTree {
public:
	enum class SearchMode { StrictMatch, MatchWildcards };

	IndexTree() {}
	/* From IndexTree::serialize output: the included indices of the tree,
	 * in its depth-first order (Lib/src/Core/IndexTree.cpp:41-131). */
	IndexTree(const std::string &serialization, Serializable::Mode mode);
	/* The reference's node tree of these indices (children per subindex
	 * value, flags, linear indices; Lib/src/Core/IndexTree.cpp:166-313,
	 * 606-650), as JSON. Wildcard types are written with the reference's
	 * values of IDX_ALL, IDX_SUM_ALL and IDX_SPIN
	 * (Lib/include/Core/TBTK/Subindex.h:395-398). */
	std::string serialize(Serializable::Mode mode) const;

	void add(const Index &index);
	void generateLinearMap();
	/* Linear index of `index`, or -1 if absent (returnNegativeForMissing),
	 * else exit(1). */
	int getLinearIndex(const Index &index,
	                   SearchMode mode = SearchMode::StrictMatch,
	                   bool returnNegativeForMissing = false) const;
	int getSize() const { return (int)ordered.size(); }
	const Index &getPhysicalIndex(int linearIndex) const {
		return ordered.at(linearIndex);
	}
	/* Stored indices matching `pattern` (IDX_ALL in the pattern matches any
	 * value), in order. */
	std::vector<Index> getIndexList(const Index &pattern) const;
	/* Subindex positions at which the stored index matching `index` has
	 * the wildcard `wildcard` (e.g. IDX_SPIN). */
	std::vector<unsigned int> getSubindicesMatching(Subindex wildcard,
	                                                const Index &index) const;

	typedef std::vector<Index>::const_iterator ConstIterator;
	ConstIterator cbegin() const { return ordered.cbegin(); }
	ConstIterator cend() const { return ordered.cend(); }
	ConstIterator begin() const { return ordered.cbegin(); }
	ConstIterator end() const { return ordered.cend(); }

private:
	struct TreeLess {
		bool operator()(const Index &a, const Index &b) const;
	};
	std::set<Index, TreeLess> items;
	std::vector<Index> ordered;
	std::map<Index, int> linear;
	std::set<std::vector<char>> wildcardMasks;
};

}  // namespace TBTK
#endif
