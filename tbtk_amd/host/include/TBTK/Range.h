/* Linear grid of `resolution` points with both bounds included, as
 * Lib/src/Utilities/Range.cpp:30-57 (dx = (upper - lower)/(resolution - 1))
 * and Range.h:130 (operator[] = start + n*dx). */
#ifndef TBTK_AMD_RANGE
#define TBTK_AMD_RANGE

#include "TBTK/Serializable.h"

#include <string>

namespace TBTK {

class Range {
public:
	Range(double lowerBound, double upperBound, unsigned int resolution);
	/* From Range::serialize output (Lib/src/Utilities/Range.cpp:59-92). */
	Range(const std::string &serialization, Serializable::Mode mode);
	/* {"dx", "id": "Range", "resolution", "start"} (Range.cpp:94-115). */
	std::string serialize(Serializable::Mode mode) const;
	double operator[](unsigned int n) const { return start + n * dx; }
	unsigned int getResolution() const { return resolution; }
	double getLast() const { return (*this)[resolution - 1]; }
	/* Grid spacing (not in the reference's Range interface). */
	double getDx() const { return dx; }
	bool operator==(const Range &o) const {
		return start == o.start && dx == o.dx && resolution == o.resolution;
	}

private:
	double start, dx;
	unsigned int resolution;
};

}  // namespace TBTK
#endif
