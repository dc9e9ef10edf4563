/* Index and Subindex: the physical index of a state ({x, y, spin}, ...),
 * with the wildcard values the pattern-based property calls use. Same names
 * and meaning as Lib/include/Core/TBTK/Index.h and Subindex.h:394-401; the
 * wildcard values are this build's own (negative integers). */
#ifndef TBTK_AMD_INDEX
#define TBTK_AMD_INDEX

#include <initializer_list>
#include <string>
#include <vector>

namespace TBTK {

class Subindex {
public:
	constexpr Subindex() : value(0) {}
	constexpr Subindex(int v) : value(v) {}
	constexpr operator int() const { return value; }
	bool isWildcard() const;          // IDX_ALL
	bool isSummationIndex() const;    // IDX_SUM_ALL
	bool isSpinIndex() const;         // IDX_SPIN
	bool isRangeIndex() const;        // IDX_X, IDX_Y, IDX_Z
	bool isIndexSeparator() const;    // IDX_SEPARATOR

private:
	int value;
};

constexpr Subindex IDX_ALL = -1;
constexpr Subindex _a_ = IDX_ALL;
constexpr Subindex IDX_SUM_ALL = -2;
constexpr Subindex IDX_SPIN = -3;
constexpr Subindex IDX_SEPARATOR = -4;
constexpr Subindex IDX_X = -5;
constexpr Subindex IDX_Y = -6;
constexpr Subindex IDX_Z = -7;

class Index {
public:
	Index() {}
	Index(std::initializer_list<Subindex> i) : indices(i) {}
	Index(std::vector<Subindex> i) : indices(std::move(i)) {}
	Index(const std::vector<int> &i) : indices(i.begin(), i.end()) {}
	/* Compound index {head, tail}: head, IDX_SEPARATOR, tail. */
	Index(const Index &head, const Index &tail);
	Index(std::initializer_list<Index> list);
	Index(const std::vector<Index> &list);

	unsigned int getSize() const { return (unsigned int)indices.size(); }
	Subindex &at(unsigned int n) { return indices.at(n); }
	const Subindex &at(unsigned int n) const { return indices.at(n); }
	Subindex &operator[](unsigned int n) { return indices[n]; }
	const Subindex &operator[](unsigned int n) const { return indices[n]; }
	void pushBack(Subindex s) { indices.push_back(s); }

	/* Components separated by IDX_SEPARATOR. */
	std::vector<Index> split() const;

	/* Equal sizes and subindices; with allowWildcards an IDX_ALL,
	 * IDX_SUM_ALL or IDX_SPIN on either side matches anything. */
	bool equals(const Index &other, bool allowWildcards = false) const;

	bool operator==(const Index &o) const { return indices == o.indices; }
	bool operator!=(const Index &o) const { return indices != o.indices; }
	bool operator<(const Index &o) const;

	std::string toString() const;
	const std::vector<Subindex> &getSubindices() const { return indices; }

private:
	std::vector<Subindex> indices;
};

}  // namespace TBTK
#endif
