/* Occupation functions used for the thermodynamic-equilibrium weights of
 * Density, Magnetization and expectation values, as
 * Lib/include/Utilities/TBTK/Functions.h:52-105 with the default unit
 * scales (natural = base units: eV and K), so that the Boltzmann constant is
 * 1.380649e-23 J/K / 1.602176634e-19 J/eV in eV/K
 * (Lib/src/Quantity/Constants.cpp:38). */
#ifndef TBTK_AMD_FUNCTIONS
#define TBTK_AMD_FUNCTIONS

#include "TBTK/TBTKMacros.h"

#include <cmath>

namespace TBTK {

class Functions {
public:
	static double boltzmannConstant() { return 1.380649e-23 / 1.602176634e-19; }

	static double fermiDiracDistribution(double energy, double mu, double temperature) {
		const double e = energy - mu;
		const double t = temperature;
		if (t != 0.) return 1. / (std::exp(e / (boltzmannConstant() * t)) + 1.);
		if (e < 0.) return 1.;
		if (e == 0.) return 0.5;
		return 0.;
	}

	static double boseEinsteinDistribution(double energy, double mu, double temperature) {
		const double e = energy - mu;
		const double t = temperature;
		if (t != 0.) return 1. / (std::exp(e / (boltzmannConstant() * t)) - 1.);
		TBTKExit("Functions::boseEinsteinDistribution()",
		         "Bose-Einstein distribution not well behaved at T=0.",
		         "Use Model::setTemperature() to set a non-zero temperature.");
	}
};

}  // namespace TBTK
#endif
