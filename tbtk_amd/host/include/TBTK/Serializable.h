/* Serialization modes of the property containers, as
 * Lib/include/Utilities/TBTK/Serializable.h (Serializable::Mode). Only
 * Mode::JSON is implemented for the Chebyshev properties, as in the
 * reference. */
#ifndef TBTK_AMD_SERIALIZABLE
#define TBTK_AMD_SERIALIZABLE

namespace TBTK {

class Serializable {
public:
	enum class Mode { Debug, JSON };
};

}  // namespace TBTK
#endif
