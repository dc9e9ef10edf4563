// Index, Range, HoppingAmplitude, HoppingAmplitudeSet, Model, IndexTree,
// Streams of the drop-in model layer.
#include "TBTK/IndexTree.h"
#include "TBTK/Model.h"
#include "TBTK/Range.h"
#include "TBTK/Streams.h"
#include "TBTK/TBTKMacros.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <sstream>

namespace TBTK {

// ------------------------------------------------------------- Streams --

namespace {
bool g_globalVerbose = true;

// Forwards to a C stream until muted.
class MutableBuffer : public std::streambuf {
public:
	explicit MutableBuffer(FILE *target) : target(target) {}
	bool muted = false;

protected:
	int overflow(int c) override {
		if (c != EOF && !muted) fputc(c, target);
		return c == EOF ? 0 : c;
	}
	std::streamsize xsputn(const char *s, std::streamsize n) override {
		if (!muted) fwrite(s, 1, (size_t)n, target);
		return n;
	}
	int sync() override { return fflush(target) == 0 ? 0 : -1; }

private:
	FILE *target;
};

MutableBuffer g_outBuffer(stdout), g_errBuffer(stderr);
}  // namespace

std::ostream Streams::out(&g_outBuffer);
std::ostream Streams::err(&g_errBuffer);

void Streams::setStdMuteOut() { g_outBuffer.muted = true; }
void Streams::setStdMuteErr() { g_errBuffer.muted = true; }
void Communicator::setGlobalVerbose(bool v) { g_globalVerbose = v; }
bool Communicator::getGlobalVerbose() { return g_globalVerbose; }

// --------------------------------------------------------------- Index --

bool Subindex::isWildcard() const { return value == IDX_ALL; }
bool Subindex::isSummationIndex() const { return value == IDX_SUM_ALL; }
bool Subindex::isSpinIndex() const { return value == IDX_SPIN; }
bool Subindex::isRangeIndex() const {
	return value == IDX_X || value == IDX_Y || value == IDX_Z;
}
bool Subindex::isIndexSeparator() const { return value == IDX_SEPARATOR; }

Index::Index(const Index &head, const Index &tail) : indices(head.indices) {
	indices.push_back(IDX_SEPARATOR);
	indices.insert(indices.end(), tail.indices.begin(), tail.indices.end());
}

Index::Index(std::initializer_list<Index> list) : Index(std::vector<Index>(list)) {}

Index::Index(const std::vector<Index> &list) {
	for (size_t n = 0; n < list.size(); n++) {
		if (n) indices.push_back(IDX_SEPARATOR);
		indices.insert(indices.end(), list[n].indices.begin(), list[n].indices.end());
	}
}

std::vector<Index> Index::split() const {
	std::vector<Index> out(1);
	for (const Subindex &s : indices) {
		if (s.isIndexSeparator())
			out.push_back(Index());
		else
			out.back().pushBack(s);
	}
	return out;
}

static bool isPatternWildcard(int v) {
	return v == IDX_ALL || v == IDX_SUM_ALL || v == IDX_SPIN;
}

bool Index::equals(const Index &o, bool allowWildcards) const {
	if (indices.size() != o.indices.size()) return false;
	for (size_t n = 0; n < indices.size(); n++) {
		if (indices[n] == o.indices[n]) continue;
		if (allowWildcards && (isPatternWildcard(indices[n]) || isPatternWildcard(o.indices[n])))
			continue;
		return false;
	}
	return true;
}

bool Index::operator<(const Index &o) const {
	return std::lexicographical_compare(
	    indices.begin(), indices.end(), o.indices.begin(), o.indices.end(),
	    [](const Subindex &a, const Subindex &b) { return (int)a < (int)b; });
}

std::string Index::toString() const {
	std::stringstream ss;
	ss << "{";
	for (size_t n = 0; n < indices.size(); n++) {
		if (indices[n].isIndexSeparator()) {
			ss << "}, {";
			continue;
		}
		if (n && !indices[n - 1].isIndexSeparator()) ss << ", ";
		switch ((int)indices[n]) {
		case -1: ss << "IDX_ALL"; break;
		case -2: ss << "IDX_SUM_ALL"; break;
		case -3: ss << "IDX_SPIN"; break;
		case -5: ss << "IDX_X"; break;
		case -6: ss << "IDX_Y"; break;
		case -7: ss << "IDX_Z"; break;
		default: ss << (int)indices[n];
		}
	}
	ss << "}";
	return ss.str();
}

// --------------------------------------------------------------- Range --

Range::Range(double lowerBound, double upperBound, unsigned int resolution)
    : start(lowerBound), resolution(resolution) {
	TBTKAssert(resolution > 1, "Range::Range()", "'resolution' must be larger than 1.", "");
	dx = (upperBound - lowerBound) / (resolution - 1);
}

// ---------------------------------------------------- HoppingAmplitude --

HoppingAmplitude::HoppingAmplitude(std::complex<double> amplitude, Index to, Index from)
    : amplitude(amplitude), callback(nullptr), to(std::move(to)), from(std::move(from)) {}

HoppingAmplitude::HoppingAmplitude(const AmplitudeCallback &cb, Index to, Index from)
    : amplitude(0), callback(&cb), to(std::move(to)), from(std::move(from)) {}

std::complex<double> HoppingAmplitude::getAmplitude() const {
	return callback ? callback->getHoppingAmplitude(to, from) : amplitude;
}

HoppingAmplitude HoppingAmplitude::getHermitianConjugate() const {
	if (callback) return HoppingAmplitude(*callback, from, to);
	return HoppingAmplitude(std::conj(amplitude), from, to);
}

HoppingAmplitudeList operator+(const HoppingAmplitude &ha, HermitianConjugate) {
	HoppingAmplitudeList l;
	l.pushBack(ha);
	l.pushBack(ha.getHermitianConjugate());
	return l;
}

// ------------------------------------------------- HoppingAmplitudeSet --

void HoppingAmplitudeSet::add(const HoppingAmplitude &ha) {
	const Index &f = ha.getFromIndex();
	const Index &t = ha.getToIndex();
	for (unsigned int n = 0; n < f.getSize(); n++)
		TBTKAssert(f.at(n) >= 0, "HoppingAmplitudeSet::add()",
		           "Invalid Index. The from-Index " << f.toString()
		               << " has a negative subindex.", "");
	for (unsigned int n = 0; n < t.getSize(); n++)
		TBTKAssert(t.at(n) >= 0, "HoppingAmplitudeSet::add()",
		           "Invalid Index. The to-Index " << t.toString()
		               << " has a negative subindex.", "");
	amps.push_back(ha);
	constructed = false;
}

void HoppingAmplitudeSet::construct() {
	basisOf.clear();
	basis.clear();
	for (const HoppingAmplitude &ha : amps) basisOf[ha.getFromIndex()] = 0;
	// Depth-first order over subindex values equals lexicographic order as
	// long as no index is a proper prefix of another (the reference rejects
	// such sets too, HoppingAmplitudeTree::_add).
	const Index *prev = nullptr;
	for (auto &kv : basisOf) {
		if (prev && prev->getSize() < kv.first.getSize()) {
			bool prefix = true;
			for (unsigned int n = 0; n < prev->getSize(); n++)
				if (prev->at(n) != kv.first.at(n)) prefix = false;
			TBTKAssert(!prefix, "HoppingAmplitudeSet::construct()",
			           "Incompatible indices " << prev->toString() << " and "
			                                   << kv.first.toString() << ".", "");
		}
		kv.second = (int)basis.size();
		basis.push_back(kv.first);
		prev = &kv.first;
	}
	for (const HoppingAmplitude &ha : amps)
		TBTKAssert(basisOf.count(ha.getToIndex()), "HoppingAmplitudeSet::construct()",
		           "The to-Index " << ha.getToIndex().toString()
		                           << " is never a from-Index (H is not Hermitian).", "");
	constructed = true;
	generation++;
}

int HoppingAmplitudeSet::getBasisIndex(const Index &index) const {
	auto it = basisOf.find(index);
	return it == basisOf.end() ? -1 : it->second;
}

const Index &HoppingAmplitudeSet::getPhysicalIndex(int b) const {
	TBTKAssert(b >= 0 && b < (int)basis.size(), "HoppingAmplitudeSet::getPhysicalIndex()",
	           "Basis index out of range.", "");
	return basis[b];
}

std::vector<Index> HoppingAmplitudeSet::getIndexList(const std::vector<Index> &patterns) const {
	std::vector<Index> out;
	for (const Index &idx : basis) {
		for (const Index &p : patterns) {
			if (p.equals(idx, true)) {
				out.push_back(idx);
				break;
			}
		}
	}
	return out;
}

SparseMatrixCSR HoppingAmplitudeSet::getSparseMatrix() const {
	TBTKAssert(constructed, "HoppingAmplitudeSet::getSparseMatrix()",
	           "HoppingAmplitudeSet has to be constructed first.", "");
	const int64_t N = (int64_t)basis.size();
	struct Entry {
		uint32_t row, col;
		size_t order;
		std::complex<double> value;
	};
	std::vector<Entry> e;
	e.reserve(amps.size());
	for (size_t n = 0; n < amps.size(); n++)
		e.push_back({(uint32_t)getBasisIndex(amps[n].getToIndex()),
		             (uint32_t)getBasisIndex(amps[n].getFromIndex()), n,
		             amps[n].getAmplitude()});
	std::stable_sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) {
		return a.row != b.row ? a.row < b.row : a.col < b.col;
	});
	SparseMatrixCSR m;
	m.numRows = N;
	m.rowPointers.assign(N + 1, 0);
	for (size_t i = 0; i < e.size();) {
		size_t j = i + 1;
		while (j < e.size() && e[j].row == e[i].row && e[j].col == e[i].col) j++;
		// a + (b + (c + ...)): merge from the last duplicate backwards.
		std::complex<double> v = e[j - 1].value;
		for (size_t k = j - 1; k-- > i;) v = e[k].value + v;
		m.columns.push_back(e[i].col);
		m.values.push_back(v);
		m.rowPointers[e[i].row + 1]++;
		i = j;
	}
	for (int64_t r = 0; r < N; r++) m.rowPointers[r + 1] += m.rowPointers[r];
	return m;
}

// --------------------------------------------------------------- Model --

Model &Model::operator<<(const HoppingAmplitude &ha) {
	hoppingAmplitudeSet.add(ha);
	return *this;
}

Model &Model::operator<<(const HoppingAmplitudeList &list) {
	for (const HoppingAmplitude &ha : list.get()) hoppingAmplitudeSet.add(ha);
	return *this;
}

void Model::construct() { hoppingAmplitudeSet.construct(); }

// ----------------------------------------------------------- IndexTree --

// Order of the tree: subindex by subindex, a wildcard counting as value 0
// (it is the only child of its node), a separator as a level of its own.
bool IndexTree::TreeLess::operator()(const Index &a, const Index &b) const {
	const unsigned int n = std::min(a.getSize(), b.getSize());
	for (unsigned int i = 0; i < n; i++) {
		const int x = a.at(i) < 0 && !a.at(i).isIndexSeparator() ? 0 : (int)a.at(i);
		const int y = b.at(i) < 0 && !b.at(i).isIndexSeparator() ? 0 : (int)b.at(i);
		if (x != y) return x < y;
	}
	return a.getSize() < b.getSize();
}

void IndexTree::add(const Index &index) { items.insert(index); }

void IndexTree::generateLinearMap() {
	ordered.assign(items.begin(), items.end());
	linear.clear();
	wildcardMasks.clear();
	for (size_t n = 0; n < ordered.size(); n++) {
		linear[ordered[n]] = (int)n;
		std::vector<char> mask(ordered[n].getSize());
		for (unsigned int i = 0; i < ordered[n].getSize(); i++)
			mask[i] = isPatternWildcard(ordered[n].at(i));
		wildcardMasks.insert(mask);
	}
}

int IndexTree::getLinearIndex(const Index &index, SearchMode mode,
                              bool returnNegativeForMissing) const {
	auto it = linear.find(index);
	if (it != linear.end()) return it->second;
	if (mode == SearchMode::MatchWildcards) {
		for (const std::vector<char> &mask : wildcardMasks) {
			if (mask.size() != index.getSize()) continue;
			// Try the stored wildcard kinds at the masked positions.
			for (int kind : {(int)IDX_SUM_ALL, (int)IDX_SPIN, (int)IDX_ALL}) {
				Index q = index;
				for (unsigned int i = 0; i < q.getSize(); i++)
					if (mask[i]) q.at(i) = kind;
				auto jt = linear.find(q);
				if (jt != linear.end()) return jt->second;
			}
			// Mixed wildcard kinds: scan.
			for (size_t n = 0; n < ordered.size(); n++)
				if (ordered[n].getSize() == index.getSize() &&
				    ordered[n].equals(index, true)) {
					bool ok = true;
					for (unsigned int i = 0; i < index.getSize() && ok; i++)
						if (!isPatternWildcard(ordered[n].at(i)) &&
						    ordered[n].at(i) != index.at(i))
							ok = false;
					if (ok) return (int)n;
				}
		}
	}
	if (returnNegativeForMissing) return -1;
	TBTKExit("IndexTree::getLinearIndex()",
	         "Index not included in the IndexTree '" << index.toString() << "'.", "");
}

std::vector<Index> IndexTree::getIndexList(const Index &pattern) const {
	std::vector<Index> out;
	for (const Index &idx : ordered) {
		if (idx.getSize() != pattern.getSize()) continue;
		bool ok = true;
		for (unsigned int i = 0; i < idx.getSize() && ok; i++)
			if (pattern.at(i) != IDX_ALL && pattern.at(i) != idx.at(i)) ok = false;
		if (ok) out.push_back(idx);
	}
	return out;
}

std::vector<unsigned int> IndexTree::getSubindicesMatching(Subindex wildcard,
                                                          const Index &index) const {
	std::vector<unsigned int> out;
	const int li = getLinearIndex(index, SearchMode::MatchWildcards, true);
	if (li < 0) return out;
	const Index &stored = ordered[li];
	for (unsigned int i = 0; i < stored.getSize(); i++)
		if (stored.at(i) == wildcard) out.push_back(i);
	return out;
}

}  // namespace TBTK
