// C++ tests of the drop-in Solver/PropertyExtractor::ChebyshevExpander,
// following the reference's own tests:
//   Api.*  setters, defaults, asserts, model layer (no GPU needed)
//          (Test/include/TBTK/Test/Solver/ChebyshevExpander.h:35-182)
//   Gpu.*  known answers and closed forms on the GPU
//          (Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-565,
//           Test/include/TBTK/Test/PropertyExtractor/ChebyshevExpander.h:19-45)
#include "minitest.h"

#include "TBTK/Functions.h"
#include "TBTK/Model.h"
#include "TBTK/PropertyExtractor/ChebyshevExpander.h"
#include "TBTK/Solver/ChebyshevExpander.h"

#include <complex>
#include <limits>
#include <thread>

using namespace TBTK;
using std::complex;

static const double EPSILON_100 = 100 * std::numeric_limits<double>::epsilon();

// ------------------------------------------------------------------ Api --

TEST(Api, setScaleFactor) {
	Solver::ChebyshevExpander solver;
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		solver.setScaleFactor(0);
	}, 1);
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		solver.setScaleFactor(-1);
	}, 1);
}

TEST(Api, getScaleFactor) {
	Solver::ChebyshevExpander solver;
	EXPECT_DOUBLE_EQ(solver.getScaleFactor(), 1.1);
	solver.setScaleFactor(10);
	EXPECT_DOUBLE_EQ(solver.getScaleFactor(), 10);
}

TEST(Api, getNumCoefficients) {
	Solver::ChebyshevExpander solver;
	EXPECT_EQ(solver.getNumCoefficients(), 1000);
	solver.setNumCoefficients(2000);
	EXPECT_EQ(solver.getNumCoefficients(), 2000);
}

TEST(Api, getBroadening) {
	Solver::ChebyshevExpander solver;
	EXPECT_DOUBLE_EQ(solver.getBroadening(), 1e-6);
	solver.setBroadening(0.01);
	EXPECT_DOUBLE_EQ(solver.getBroadening(), 0.01);
}

TEST(Api, setEnergyWindow) {
	Solver::ChebyshevExpander solver;
	solver.setScaleFactor(10);
	solver.setEnergyWindow(Range(-9.999, 9.999, 10));
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		solver.setEnergyWindow(Range(-10, 9.999, 10));
	}, 1);
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		solver.setEnergyWindow(Range(-9.999, 10, 10));
	}, 1);
}

TEST(Api, getEnergyWindow) {
	Solver::ChebyshevExpander solver;
	EXPECT_TRUE(solver.getEnergyWindow() == Range(-1, 1, 1000));
	solver.setEnergyWindow(Range(-0.1, 0.2, 10));
	EXPECT_TRUE(solver.getEnergyWindow() == Range(-0.1, 0.2, 10));
}

TEST(Api, gpuFlagsAndLookupTable) {
	Solver::ChebyshevExpander solver;
	EXPECT_FALSE(solver.getCalculateCoefficientsOnGPU());
	EXPECT_FALSE(solver.getGenerateGreensFunctionsOnGPU());
	EXPECT_FALSE(solver.getUseLookupTable());
	solver.setCalculateCoefficientsOnGPU(true);
	solver.setGenerateGreensFunctionsOnGPU(true);
	solver.setUseLookupTable(true);
	EXPECT_TRUE(solver.getCalculateCoefficientsOnGPU());
	EXPECT_TRUE(solver.getGenerateGreensFunctionsOnGPU());
	EXPECT_TRUE(solver.getUseLookupTable());
}

static Model periodicSquare(int SIZE, double mu, double t) {
	Model model;
	model.setVerbose(false);
	for (int x = 0; x < SIZE; x++) {
		for (int y = 0; y < SIZE; y++) {
			model << HoppingAmplitude(-mu, {x, y}, {x, y});
			model << HoppingAmplitude(-t, {(x + 1) % SIZE, y}, {x, y}) + HC;
			model << HoppingAmplitude(-t, {x, (y + 1) % SIZE}, {x, y}) + HC;
		}
	}
	model.construct();
	return model;
}

// As in the reference (ensureLookupTableIsReady,
// Lib/include/TBTK/Solver/ChebyshevExpander.h:548-565), Green's functions on
// the GPU need the lookup table. The check comes before any device work.
TEST(Api, gpuGreensFunctionNeedsLookupTable) {
	Solver::ChebyshevExpander solver;
	solver.setScaleFactor(10);
	solver.setNumCoefficients(3);
	solver.setGenerateGreensFunctionsOnGPU(true);
	std::vector<complex<double>> c = {1, 2, 3};
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		solver.generateGreensFunction(c);
	}, 1);
}

// Templates/BasicChebyshev/src/main.cpp:91-110 as written: the window
// [-5, 5] at scale factor 5 is not strictly inside (-s, s), so the
// reference exits there (Lib/include/TBTK/Solver/ChebyshevExpander.h:
// 434-451), before any computation; so does the drop-in.
TEST(Api, basicChebyshevTemplateWindow) {
	Model model;
	model.setVerbose(false);
	for (int x = 0; x < 4; x++)
		for (int s = 0; s < 2; s++) {
			model << HoppingAmplitude(1.0, {x, 0, s}, {x, 0, s});
			if (x + 1 < 4) model << HoppingAmplitude(-1.0, {x + 1, 0, s}, {x, 0, s}) + HC;
		}
	model.construct();
	const double SCALE_FACTOR = 5.;
	Solver::ChebyshevExpander solver;
	solver.setModel(model);
	solver.setScaleFactor(SCALE_FACTOR);
	solver.setCalculateCoefficientsOnGPU(false);
	solver.setGenerateGreensFunctionsOnGPU(false);
	solver.setUseLookupTable(true);
	solver.setNumCoefficients(500);
	PropertyExtractor::ChebyshevExpander propertyExtractor(solver);
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		propertyExtractor.setEnergyWindow(-SCALE_FACTOR, SCALE_FACTOR, 1000);
	}, 1);
}

TEST(Api, basisOrderAndSparseMatrix) {
	Model model = periodicSquare(5, -2, 1);
	const HoppingAmplitudeSet &has = model.getHoppingAmplitudeSet();
	EXPECT_EQ(has.getBasisSize(), 25);
	EXPECT_EQ(has.getBasisIndex({0, 0}), 0);
	EXPECT_EQ(has.getBasisIndex({0, 4}), 4);
	EXPECT_EQ(has.getBasisIndex({1, 0}), 5);
	EXPECT_TRUE(has.getPhysicalIndex(7) == Index({1, 2}));
	SparseMatrixCSR m = has.getSparseMatrix();
	EXPECT_EQ(m.rowPointers.back(), 125u);
	for (int r = 0; r < 25; r++) {
		EXPECT_EQ(m.rowPointers[r + 1] - m.rowPointers[r], 5u);
		for (uint32_t k = m.rowPointers[r] + 1; k < m.rowPointers[r + 1]; k++)
			EXPECT_TRUE(m.columns[k - 1] < m.columns[k]);
	}
	// 2x2 periodic: (x+1)%2 == (x-1)%2, so each neighbour pair appears twice
	// and is summed.
	Model m2 = periodicSquare(2, -2, 1);
	SparseMatrixCSR s2 = m2.getHoppingAmplitudeSet().getSparseMatrix();
	EXPECT_EQ(s2.rowPointers.back(), 12u);
	EXPECT_NEAR(std::real(s2.values[1]), -2.0, 0);
}

TEST(Api, indexTreeOrderAndWildcards) {
	IndexTree tree;
	tree.add({3, 1, IDX_SUM_ALL});
	tree.add({1, 2, IDX_SUM_ALL});
	tree.add({1, 0, IDX_SUM_ALL});
	tree.generateLinearMap();
	EXPECT_EQ(tree.getSize(), 3);
	EXPECT_TRUE(tree.getPhysicalIndex(0) == Index({1, 0, IDX_SUM_ALL}));
	EXPECT_EQ(tree.getLinearIndex({1, 2, 1}, IndexTree::SearchMode::MatchWildcards), 1);
	EXPECT_EQ(tree.getLinearIndex({3, 1, 0}, IndexTree::SearchMode::MatchWildcards), 2);
	EXPECT_EQ(tree.getLinearIndex({2, 1, 0}, IndexTree::SearchMode::MatchWildcards, true), -1);
	IndexTree c;
	c.add(Index({Index({1, 0}), Index({0, 0})}));
	c.add(Index({Index({0, 1}), Index({0, 0})}));
	c.generateLinearMap();
	EXPECT_TRUE(c.getPhysicalIndex(0) == Index({Index({0, 1}), Index({0, 0})}));
}

TEST(Api, propertyExtractorWindow) {
	Model model = periodicSquare(5, -2, 1);
	Solver::ChebyshevExpander solver;
	solver.setModel(model);
	solver.setScaleFactor(10);
	PropertyExtractor::ChebyshevExpander pe(solver);
	pe.setEnergyWindow(-9, 9, 100);
	EXPECT_TRUE(solver.getEnergyWindow() == Range(-9, 9, 100));
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		pe.setEnergyWindow(-10, 10, 100);
	}, 1);
}

TEST(Api, spinPropertiesNeedSpinIndex) {
	Model model = periodicSquare(5, -2, 1);
	Solver::ChebyshevExpander solver;
	solver.setModel(model);
	solver.setScaleFactor(10);
	PropertyExtractor::ChebyshevExpander pe(solver);
	pe.setEnergyWindow(-9, 9, 100);
	// No IDX_SPIN: the reference exits before any computation
	// (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:393-399, 520-526).
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		pe.calculateMagnetization({IDX_X, 0}, {5, 1});
	}, 1);
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		pe.calculateSpinPolarizedLDOS({IDX_X, 0}, {5, 1});
	}, 1);
}

TEST(Api, occupationFunctions) {
	// k_B in eV/K as the reference derives it (Constants.cpp:38).
	EXPECT_TRUE(Functions::boltzmannConstant() == 1.380649e-23 / 1.602176634e-19);
	EXPECT_TRUE(Functions::fermiDiracDistribution(0.2, 0.3, 0) == 1.);
	EXPECT_TRUE(Functions::fermiDiracDistribution(0.3, 0.3, 0) == 0.5);
	EXPECT_TRUE(Functions::fermiDiracDistribution(0.4, 0.3, 0) == 0.);
	EXPECT_TRUE(Functions::fermiDiracDistribution(0.3, 0.3, 300) == 0.5);
	// f(mu + x) + f(mu - x) = 1 and the Bose-Einstein function's pole
	const double x = 0.0123, T = 250;
	const double s = Functions::fermiDiracDistribution(0.3 + x, 0.3, T) +
	                 Functions::fermiDiracDistribution(0.3 - x, 0.3, T);
	EXPECT_TRUE(std::abs(s - 1) < 1e-15);
	EXPECT_TRUE(Functions::boseEinsteinDistribution(0.5, 0.3, 300) > 0);
	EXPECT_EXIT_CODE({
		Streams::setStdMuteErr();
		Functions::boseEinsteinDistribution(0.5, 0.3, 0);
	}, 1);
}

// ------------------------------------------------------------------ Gpu --

// The reference's own tests run with its default flags (CPU path); here
// the coefficient flag is set, the Green's-function flag left false so that
// both the table and the direct evaluation are reachable.
static void gpuSolver(Solver::ChebyshevExpander &solver) {
	solver.setVerbose(false);
	solver.setCalculateCoefficientsOnGPU(true);
	solver.setGenerateGreensFunctionsOnGPU(false);
}

TEST(Gpu, calculateCoefficients) {
	const double mu = -2, t = 1, SCALE_FACTOR = 10;
	Model model = periodicSquare(5, mu, t);
	Solver::ChebyshevExpander solver;
	gpuSolver(solver);
	solver.setModel(model);
	solver.setScaleFactor(SCALE_FACTOR);
	solver.setUseLookupTable(true);
	solver.setNumCoefficients(100);
	solver.setBroadening(0);
	std::vector<complex<double>> c0 = solver.calculateCoefficients({0, 0}, {0, 0});
	ASSERT_EQ(c0.size(), 100u);
	EXPECT_NEAR(real(c0[0]), 1, EPSILON_100);
	EXPECT_NEAR(imag(c0[0]), 0, EPSILON_100);
	EXPECT_NEAR(real(c0[1]), -mu / SCALE_FACTOR, EPSILON_100);
	EXPECT_NEAR(real(c0[2]), 2 * (mu * mu + 4 * t * t) / (SCALE_FACTOR * SCALE_FACTOR) - 1,
	            EPSILON_100);
	std::vector<complex<double>> c1 = solver.calculateCoefficients({1, 0}, {0, 0});
	EXPECT_NEAR(real(c1[0]), 0, EPSILON_100);
	EXPECT_NEAR(real(c1[1]), -t / SCALE_FACTOR, EPSILON_100);
	EXPECT_NEAR(real(c1[2]), 4 * mu * t / (SCALE_FACTOR * SCALE_FACTOR), EPSILON_100);
	std::vector<Index> toIndices = {{0, 0}, {1, 0}};
	std::vector<std::vector<complex<double>>> c2 = solver.calculateCoefficients(toIndices, {0, 0});
	for (int n = 0; n < 100; n++) {
		EXPECT_TRUE(c2[0][n] == c0[n]);
		EXPECT_TRUE(c2[1][n] == c1[n]);
	}
}

TEST(Gpu, generateGreensFunctionClosedForms) {
	const double S = 10;
	Range energyWindow(-5, 5, 10);
	const complex<double> i(0, 1);
	std::vector<complex<double>> c = {1, 2, 3};
	for (int lookup = 0; lookup < 2; lookup++) {
		Solver::ChebyshevExpander solver;
		gpuSolver(solver);
		solver.setScaleFactor(S);
		solver.setUseLookupTable(lookup);
		solver.setNumCoefficients(3);
		solver.setEnergyWindow(energyWindow);
		typedef Solver::ChebyshevExpander::Type T;
		std::vector<complex<double>> g[4] = {
		    solver.generateGreensFunction(c, T::Retarded),
		    solver.generateGreensFunction(c, T::Advanced),
		    solver.generateGreensFunction(c, T::Principal),
		    solver.generateGreensFunction(c, T::NonPrincipal)};
		for (unsigned int n = 0; n < energyWindow.getResolution(); n++) {
			const double th = acos(energyWindow[n] / S);
			const double den = sqrt(S * S - energyWindow[n] * energyWindow[n]);
			complex<double> ret = -c[0] / 2., adv = c[0] / 2., pri = 0, npr = c[0] / 2.;
			for (int k = 1; k < 3; k++) {
				ret -= c[k] * exp(-i * ((double)k) * th);
				adv += c[k] * exp(i * ((double)k) * th);
				pri += c[k] * sin(((double)k) * th);
				npr += c[k] * cos(((double)k) * th);
			}
			ret *= 2. * i / den;
			adv *= 2. * i / den;
			pri *= 2. / den;
			npr *= 2. * i / den;
			const complex<double> ref[4] = {ret, adv, pri, npr};
			for (int t = 0; t < 4; t++) {
				EXPECT_NEAR(real(g[t][n]), real(ref[t]), EPSILON_100);
				EXPECT_NEAR(imag(g[t][n]), imag(ref[t]), EPSILON_100);
			}
		}
	}
}

TEST(Gpu, calculateLDOSSumRule) {
	const unsigned int SIZE = 10;
	Model model;
	model.setVerbose(false);
	for (unsigned int n = 0; n < SIZE; n++)
		model << HoppingAmplitude(-1, {(int)n}, {(int)((n + 1) % SIZE)}) + HC;
	model.construct();
	Solver::ChebyshevExpander solver;
	gpuSolver(solver);
	solver.setModel(model);
	solver.setScaleFactor(10);
	PropertyExtractor::ChebyshevExpander pe(solver);
	const double fractions[2] = {0.9, 0.7};
	for (double f : fractions) {
		pe.setEnergyWindow(-10 * f, 10 * f, 1000);
		Property::LDOS ldos = pe.calculateLDOS({{0}});
		double integrated = 0;
		for (unsigned int n = 0; n < ldos.getResolution(); n++)
			integrated += ldos({0}, n) * ldos.getDeltaE();
		EXPECT_NEAR(integrated, 1, 0.001);
	}
}

TEST(Gpu, blockRecurrenceMatchesPlain) {
	// 12x12 square lattice with spin; LDOS of four sites both ways.
	Model model;
	model.setVerbose(false);
	for (int x = 0; x < 12; x++)
		for (int y = 0; y < 12; y++)
			for (int s = 0; s < 2; s++) {
				if (x + 1 < 12) model << HoppingAmplitude(-1, {x + 1, y, s}, {x, y, s}) + HC;
				if (y + 1 < 12) model << HoppingAmplitude(-1, {x, y + 1, s}, {x, y, s}) + HC;
				model << HoppingAmplitude(-0.5 * 2 * (0.5 - s), {x, y, s}, {x, y, s});
			}
	model.construct();
	std::vector<Index> patterns = {{3, 4, IDX_SUM_ALL}, {6, 6, IDX_SUM_ALL}, {0, 11, 1}};
	Property::LDOS ref(std::vector<int>{1}, Range(-1, 1, 2)), blk = ref;
	for (int pass = 0; pass < 2; pass++) {
		Solver::ChebyshevExpander solver;
		gpuSolver(solver);
		solver.setModel(model);
		solver.setScaleFactor(5);
		solver.setNumCoefficients(400);
		solver.setUseLookupTable(true);
		solver.setUseBlockRecurrence(pass == 1);
		solver.setBlockFlags(0);   // FP32 block, gather moments
		PropertyExtractor::ChebyshevExpander pe(solver);
		pe.setEnergyWindow(-4.9, 4.9, 300);
		(pass ? blk : ref) = pe.calculateLDOS(patterns);
	}
	double maxRef = 0, maxDiff = 0;
	for (size_t n = 0; n < ref.getData().size(); n++) {
		maxRef = std::max(maxRef, std::fabs(ref.getData()[n]));
		maxDiff = std::max(maxDiff, std::fabs(ref.getData()[n] - blk.getData()[n]));
	}
	EXPECT_TRUE(maxDiff / maxRef < 1e-5);
}

TEST(Gpu, calculateDOS) {
	Model model = periodicSquare(6, 0.3, 1);
	Solver::ChebyshevExpander solver;
	gpuSolver(solver);
	solver.setModel(model);
	solver.setScaleFactor(6);
	solver.setNumCoefficients(200);
	solver.setUseLookupTable(true);
	solver.setUseBlockRecurrence(true);
	solver.setBlockFlags(4 | 1 | 2);   // FP64 vectors and H, doubling
	PropertyExtractor::ChebyshevExpander pe(solver);
	pe.setEnergyWindow(-5.9, 5.9, 500);
	Property::DOS exact = pe.calculateDOS();
	double integral = 0;
	for (unsigned int n = 0; n < exact.getResolution(); n++)
		integral += exact(n) * exact.getDeltaE();
	EXPECT_NEAR(integral, 36, 0.05);
	pe.setStochasticVectors(400, 7);
	Property::DOS est = pe.calculateDOS();
	double diff = 0, peak = 0;
	for (unsigned int n = 0; n < est.getResolution(); n++) {
		diff = std::max(diff, std::fabs(est(n) - exact(n)));
		peak = std::max(peak, exact(n));
	}
	EXPECT_TRUE(diff < 0.1 * peak);
}

// The reference's callers leave both GPU flags false
// (Templates/BasicChebyshev/src/main.cpp:100-101,
// doc/examples/Solver/ChebyshevExpander.cpp:28-29). Their results equal
// those of the GPU flags bit for bit.
TEST(Gpu, referenceDefaultFlags) {
	Model model = periodicSquare(7, 0.4, 1);
	typedef Solver::ChebyshevExpander::Type T;
	const T types[4] = {T::Advanced, T::Retarded, T::Principal, T::NonPrincipal};
	std::vector<std::vector<complex<double>>> c[2];
	std::vector<complex<double>> g[2][4], direct[4];
	for (int flags = 0; flags < 2; flags++) {
		Solver::ChebyshevExpander solver;
		solver.setVerbose(false);
		solver.setModel(model);
		solver.setScaleFactor(6);
		solver.setNumCoefficients(150);
		solver.setEnergyWindow(Range(-5.5, 5.5, 120));
		if (flags) {
			solver.setCalculateCoefficientsOnGPU(true);
			solver.setGenerateGreensFunctionsOnGPU(true);
		} else {
			EXPECT_FALSE(solver.getCalculateCoefficientsOnGPU());
			EXPECT_FALSE(solver.getGenerateGreensFunctionsOnGPU());
			for (int t = 0; t < 4; t++)   // default: no lookup table
				direct[t] = solver.generateGreensFunction(
				    solver.calculateCoefficients({3, 3}, {3, 3}), types[t]);
		}
		solver.setUseLookupTable(true);
		std::vector<Index> to = {{3, 3}, {4, 3}, {0, 6}};
		c[flags] = solver.calculateCoefficients(to, {3, 3});
		for (int t = 0; t < 4; t++) g[flags][t] = solver.generateGreensFunction(c[flags][0], types[t]);
	}
	for (int k = 0; k < 3; k++) EXPECT_TRUE(c[0][k] == c[1][k]);
	for (int t = 0; t < 4; t++) {
		EXPECT_TRUE(g[0][t] == g[1][t]);
		// table and direct evaluation agree to rounding
		double diff = 0, peak = 0;
		for (size_t e = 0; e < direct[t].size(); e++) {
			diff = std::max(diff, std::abs(direct[t][e] - g[0][t][e]));
			peak = std::max(peak, std::abs(g[0][t][e]));
		}
		EXPECT_TRUE(diff <= 1e-12 * peak);
	}
}

// The reference hands each thread its own device per GPU call
// (Lib/src/Utilities/GPUResourceManager.cpp:44-90) and its property extractor
// calls the solver from an OpenMP parallel loop
// (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:278). Here: solvers in
// separate threads at once, and one solver shared by several threads, give
// the single-threaded results bit for bit.
static std::vector<double> threadWork(Model &model, int site) {
	Solver::ChebyshevExpander solver;
	gpuSolver(solver);
	solver.setModel(model);
	solver.setScaleFactor(6);
	solver.setNumCoefficients(300);
	solver.setUseLookupTable(site % 2 == 0);
	PropertyExtractor::ChebyshevExpander pe(solver);
	pe.setEnergyWindow(-5.5, 5.5, 200);
	std::vector<double> out;
	for (int rep = 0; rep < 3; rep++) {
		Property::LDOS ldos = pe.calculateLDOS({{site, site, IDX_SUM_ALL}, {1, site, 0}});
		out.insert(out.end(), ldos.getData().begin(), ldos.getData().end());
	}
	return out;
}

TEST(Gpu, threadsAndSolvers) {
	Model model;
	model.setVerbose(false);
	for (int x = 0; x < 9; x++)
		for (int y = 0; y < 9; y++)
			for (int s = 0; s < 2; s++) {
				if (x + 1 < 9) model << HoppingAmplitude(-1, {x + 1, y, s}, {x, y, s}) + HC;
				if (y + 1 < 9) model << HoppingAmplitude(-1, {x, y + 1, s}, {x, y, s}) + HC;
				model << HoppingAmplitude(0.3 * (s ? 1 : -1) + 0.01 * x, {x, y, s}, {x, y, s});
			}
	model.construct();
	const int T = 4;
	std::vector<double> single[T], multi[T];
	for (int t = 0; t < T; t++) single[t] = threadWork(model, t + 2);
	std::vector<std::thread> threads;
	for (int t = 0; t < T; t++)
		threads.emplace_back([&, t] { multi[t] = threadWork(model, t + 2); });
	for (auto &th : threads) th.join();
	for (int t = 0; t < T; t++) EXPECT_TRUE(single[t] == multi[t]);

	// One solver, Green's functions and coefficients from several threads.
	Solver::ChebyshevExpander solver;
	gpuSolver(solver);
	solver.setModel(model);
	solver.setScaleFactor(6);
	solver.setNumCoefficients(200);
	solver.setUseLookupTable(true);
	solver.setEnergyWindow(Range(-5.5, 5.5, 150));
	typedef Solver::ChebyshevExpander::Type Ty;
	std::vector<complex<double>> ref[T], got[T];
	for (int t = 0; t < T; t++)
		ref[t] = solver.generateGreensFunction(solver.calculateCoefficients({t, 4, 0}, {t, 4, 0}),
		                                       Ty::NonPrincipal);
	threads.clear();
	for (int t = 0; t < T; t++)
		threads.emplace_back([&, t] {
			for (int rep = 0; rep < 3; rep++)
				got[t] = solver.generateGreensFunction(
				    solver.calculateCoefficients({t, 4, 0}, {t, 4, 0}), Ty::NonPrincipal);
		});
	for (auto &th : threads) th.join();
	for (int t = 0; t < T; t++) EXPECT_TRUE(ref[t] == got[t]);
}

MINITEST_MAIN
