/* TBTK::Initialize (Lib/include/Core/TBTK/TBTK.h:28, Lib/src/Core/TBTK.cpp:
 * 28-33), which the reference's programs call first. The reference sets up
 * its unit system there; the Chebyshev path of this drop-in uses no units
 * (energies are in the Hamiltonian's own units, temperatures in kelvin with
 * the Boltzmann constant in eV/K, as the reference's defaults), so there is
 * nothing to initialize. */
#ifndef TBTK_AMD_TBTK
#define TBTK_AMD_TBTK

namespace TBTK {

inline void Initialize() {}

}  // namespace TBTK
#endif
