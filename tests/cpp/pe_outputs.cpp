// Runs the golden-fixture queries through the drop-in C++ API, building each
// model with the same loops a TBTK user writes (Model << HoppingAmplitude
// ... + HC), and writes the raw outputs to <outdir>/<name>.bin for
// comparison with tests/golden/*.npz (written by the compiled reference).
//
// Usage: pe_outputs <outdir> [<crand.bin>] [gpu]
//
// By default both GPU flags stay false, the reference's default and the
// setting of its own callers; with "gpu" the coefficients run with
// setCalculateCoefficientsOnGPU(true) and, where a lookup table is used,
// setGenerateGreensFunctionsOnGPU(true). Both must give the reference's
// outputs.
#include "TBTK/Model.h"
#include "TBTK/PropertyExtractor/ChebyshevExpander.h"
#include "TBTK/Solver/ChebyshevExpander.h"

#include <complex>
#include <cstdio>
#include <string>
#include <vector>

using namespace TBTK;
using std::complex;

static std::string outdir;
static bool gpuFlags = false;

template <typename T> static void write(const std::string &name, const std::vector<T> &v) {
	FILE *f = fopen((outdir + "/" + name + ".bin").c_str(), "wb");
	fwrite(v.data(), sizeof(T), v.size(), f);
	fclose(f);
}

static void gpu(Solver::ChebyshevExpander &s) {
	s.setVerbose(false);
	s.setCalculateCoefficientsOnGPU(gpuFlags);
	s.setGenerateGreensFunctionsOnGPU(false);
}

// The lookup-table setting, and with it the Green's-function flag (the
// reference evaluates Green's functions on the GPU only through the table).
static void lookup(Solver::ChebyshevExpander &s, bool useLookupTable) {
	s.setUseLookupTable(useLookupTable);
	s.setGenerateGreensFunctionsOnGPU(gpuFlags && useLookupTable);
}

// 30x30 with spin, README idiom: -t hopping, Zeeman -J*2*(1/2 - s).
static void squareSpin(Model &model, int L, double t, double J) {
	model.setVerbose(false);
	for (int x = 0; x < L; x++)
		for (int y = 0; y < L; y++)
			for (int s = 0; s < 2; s++) {
				if (x + 1 < L) model << HoppingAmplitude(-t, {x + 1, y, s}, {x, y, s}) + HC;
				if (y + 1 < L) model << HoppingAmplitude(-t, {x, y + 1, s}, {x, y, s}) + HC;
				model << HoppingAmplitude(-J * 2 * (1 / 2. - s), {x, y, s}, {x, y, s});
			}
	model.construct();
}

static void periodicSquare(Model &model, int SIZE, double mu, double t) {
	model.setVerbose(false);
	for (int x = 0; x < SIZE; x++)
		for (int y = 0; y < SIZE; y++) {
			model << HoppingAmplitude(-mu, {x, y}, {x, y});
			model << HoppingAmplitude(-t, {(x + 1) % SIZE, y}, {x, y}) + HC;
			model << HoppingAmplitude(-t, {x, (y + 1) % SIZE}, {x, y}) + HC;
		}
	model.construct();
}

template <typename C> static std::vector<complex<double>> flatten(const C &rows) {
	std::vector<complex<double>> out;
	for (const auto &r : rows) out.insert(out.end(), r.begin(), r.end());
	return out;
}

// A property's JSON serialization, as json_<name>.json.
static void writeJSON(const std::string &name, const std::string &text) {
	FILE *f = fopen((outdir + "/json_" + name + ".json").c_str(), "wb");
	if (!f) {
		fprintf(stderr, "cannot write %s\n", name.c_str());
		exit(2);
	}
	fwrite(text.data(), 1, text.size(), f);
	fclose(f);
}

int main(int argc, char **argv) {
	if (argc < 2) {
		fprintf(stderr, "usage: pe_outputs <outdir> [crand.bin]\n");
		return 2;
	}
	outdir = argv[1];
	for (int a = 2; a < argc; a++)
		if (std::string(argv[a]) == "gpu") gpuFlags = true;

	{	// square30_spin: coefficients, broadening 0 and 1e-6
		Model model;
		squareSpin(model, 30, 1, 0.5);
		Solver::ChebyshevExpander solver;
		gpu(solver);
		solver.setModel(model);
		solver.setScaleFactor(10);
		solver.setNumCoefficients(1000);
		std::vector<Index> to = {{15, 15, 0}, {16, 15, 0}, {15, 15, 1}, {0, 0, 0}, {15, 15, 0}};
		solver.setBroadening(0);
		write("square30_spin_coef_b0", flatten(solver.calculateCoefficients(to, {15, 15, 0})));
		solver.setBroadening(1e-6);
		write("square30_spin_coef_b1e-06", flatten(solver.calculateCoefficients(to, {15, 15, 0})));

		// LDOS at (15,15,IDX_SUM_ALL) on [-9.99, 9.99] x 1000
		for (int table = 0; table < 2; table++) {
			lookup(solver, table);
			PropertyExtractor::ChebyshevExpander pe(solver);
			pe.setEnergyWindow(-9.99, 9.99, 1000);
			Property::LDOS ldos = pe.calculateLDOS({{15, 15, IDX_SUM_ALL}});
			write("square30_spin_ldos_lookup" + std::to_string(table), ldos.getData());
		}
		// Green's functions, 4 types, M = 400, (-9.5, 9.5) x 300, lookup
		lookup(solver, true);
		solver.setNumCoefficients(400);
		PropertyExtractor::ChebyshevExpander pe(solver);
		pe.setEnergyWindow(-9.5, 9.5, 300);
		const char *names[4] = {"Advanced", "Retarded", "Principal", "NonPrincipal"};
		const Property::GreensFunction::Type types[4] = {
		    Property::GreensFunction::Type::Advanced, Property::GreensFunction::Type::Retarded,
		    Property::GreensFunction::Type::Principal,
		    Property::GreensFunction::Type::NonPrincipal};
		for (int t = 0; t < 4; t++) {
			std::vector<Index> tos = {{15, 15, 0}, {16, 15, 0}};
			Property::GreensFunction g = pe.calculateGreensFunctions(tos, {15, 15, 0}, types[t]);
			std::vector<complex<double>> out;
			for (const Index &ti : tos)
				for (int e = 0; e < 300; e++) out.push_back(g({ti, {15, 15, 0}}, e));
			write(std::string("square30_spin_gf_") + names[t], out);
			if (t == 3) writeJSON("gf_NonPrincipal", g.serialize(Serializable::Mode::JSON));
		}
		// JSON serializations of LDOS in the Custom and Ranges formats.
		writeJSON("ldos_custom",
		          pe.calculateLDOS({{15, 15, IDX_SUM_ALL}, {3, 4, IDX_ALL}})
		              .serialize(Serializable::Mode::JSON));
		writeJSON("ldos_ranges", pe.calculateLDOS({IDX_X, 15, IDX_SUM_ALL}, {2, 1, 2})
		                             .serialize(Serializable::Mode::JSON));
		// The pattern form of calculateGreensFunction must give the same.
		Property::GreensFunction gp = pe.calculateGreensFunction(
		    {{{15, 15, 0}, {15, 15, 0}}, {{16, 15, 0}, {15, 15, 0}}},
		    Property::GreensFunction::Type::NonPrincipal);
		std::vector<complex<double>> out;
		for (const Index &ti : std::vector<Index>{{15, 15, 0}, {16, 15, 0}})
			for (int e = 0; e < 300; e++) out.push_back(gp({ti, {15, 15, 0}}, e));
		write("square30_spin_gfpattern_NonPrincipal", out);
		// Pattern form with 900 from-indices: G((15,15,0), (x,y,0)), Retarded.
		{
			const Index toIdx = {15, 15, 0}, fromPattern = {IDX_ALL, IDX_ALL, 0};
			Property::GreensFunction g900 = pe.calculateGreensFunction(
			    {{toIdx, fromPattern}}, Property::GreensFunction::Type::Retarded);
			std::vector<complex<double>> flat;
			for (const Index &f : model.getHoppingAmplitudeSet().getIndexList({fromPattern}))
				for (int e = 0; e < 300; e++) flat.push_back(g900({toIdx, f}, e));
			write("square30_spin_gfpattern900", flat);
		}

		// Thermodynamic-equilibrium properties at mu = 0.3, T = 300 K and 0.
		auto spins = [](const std::vector<SpinMatrix> &d) {
			std::vector<complex<double>> flat;
			for (const SpinMatrix &m : d)
				for (int n = 0; n < 4; n++) flat.push_back(m.at(n / 2, n % 2));
			return flat;
		};
		model.setChemicalPotential(0.3);
		for (double T : {300.0, 0.0}) {
			model.setTemperature(T);
			const std::string tag = T == 0 ? "T0" : "T300";
			const Property::Density dc =
			    pe.calculateDensity({{15, 15, IDX_SUM_ALL}, {3, 4, IDX_ALL}, {0, 0, 1}});
			const Property::Density dr = pe.calculateDensity({IDX_X, 15, IDX_SUM_ALL}, {4, 1, 2});
			write("thermo_density_custom_" + tag, dc.getData());
			write("thermo_density_ranges_" + tag, dr.getData());
			if (T == 300) {
				writeJSON("density_custom_T300", dc.serialize(Serializable::Mode::JSON));
				writeJSON("density_ranges_T300", dr.serialize(Serializable::Mode::JSON));
			}
			write("thermo_mag_custom_" + tag,
			      spins(pe.calculateMagnetization({{15, 15, IDX_SPIN}, {2, 3, IDX_SPIN}})
			                .getData()));
			write("thermo_mag_ranges_" + tag,
			      spins(pe.calculateMagnetization({IDX_X, 15, IDX_SPIN}, {3, 1, 2}).getData()));
			write("thermo_expval_offdiag_" + tag,
			      std::vector<complex<double>>{
			          pe.calculateExpectationValue({15, 15, 0}, {16, 15, 0})});
			write("thermo_expval_diag_" + tag,
			      std::vector<complex<double>>{
			          pe.calculateExpectationValue({15, 15, 1}, {15, 15, 1})});
		}
		model.setTemperature(300);
		write("thermo_splodos_custom",
		      spins(pe.calculateSpinPolarizedLDOS({{15, 15, IDX_SPIN}}).getData()));
		write("thermo_splodos_ranges",
		      spins(pe.calculateSpinPolarizedLDOS({IDX_X, 14, IDX_SPIN}, {2, 1, 2}).getData()));
	}

	{	// periodic5: known-answer model, M = 100, broadening 0
		Model model;
		periodicSquare(model, 5, -2, 1);
		Solver::ChebyshevExpander solver;
		gpu(solver);
		solver.setModel(model);
		solver.setScaleFactor(10);
		solver.setNumCoefficients(100);
		solver.setBroadening(0);
		std::vector<Index> to = {{0, 0}, {1, 0}, {2, 3}};
		write("periodic5_coef_b0", flatten(solver.calculateCoefficients(to, {0, 0})));

		// sampleDOS: 7 samples, seed 3, M = 200, (-9, 9) x 200
		solver.setNumCoefficients(200);
		solver.setBroadening(1e-6);
		lookup(solver, true);
		PropertyExtractor::ChebyshevExpander pe(solver);
		pe.setEnergyWindow(-9, 9, 200);
		const Property::DOS dos = pe.sampleDOS(7, {}, 3);
		write("periodic5_sampledos", dos.getData());
		writeJSON("sampledos", dos.serialize(Serializable::Mode::JSON));
	}

	{	// periodic6 (mu = 0.3): calculateDOS, exact (plain and FP64 block
		// recurrences) and stochastic (FP64 and FP32 blocks), M = 200,
		// (-5.9, 5.9) x 500, lookup table
		Model model;
		periodicSquare(model, 6, 0.3, 1);
		for (int block = 0; block < 2; block++) {
			Solver::ChebyshevExpander solver;
			gpu(solver);
			solver.setModel(model);
			solver.setScaleFactor(6);
			solver.setNumCoefficients(200);
			lookup(solver, true);
			solver.setUseBlockRecurrence(block == 1);
			solver.setBlockFlags(4 | 1 | 2);   // FP64 vectors and H, doubling
			PropertyExtractor::ChebyshevExpander pe(solver);
			pe.setEnergyWindow(-5.9, 5.9, 500);
			write(std::string("periodic6_dos_exact_") + (block ? "block" : "plain"),
			      pe.calculateDOS().getData());
			if (block) {
				pe.setStochasticVectors(50, 7);
				write("periodic6_dos_stochastic_fp64", pe.calculateDOS().getData());
				solver.setBlockFlags(2);   // FP32 vectors and H, doubling
				write("periodic6_dos_stochastic_fp32", pe.calculateDOS().getData());
			}
		}
	}

	{	// ring10 LDOS, windows +-0.9s and +-0.7s, direct contraction
		Model model;
		model.setVerbose(false);
		for (int n = 0; n < 10; n++) model << HoppingAmplitude(-1, {n}, {(n + 1) % 10}) + HC;
		model.construct();
		Solver::ChebyshevExpander solver;
		gpu(solver);
		solver.setModel(model);
		solver.setScaleFactor(10);
		PropertyExtractor::ChebyshevExpander pe(solver);
		for (double f : {0.9, 0.7}) {
			pe.setEnergyWindow(-10 * f, 10 * f, 1000);
			write(std::string("ring10_ldos_") + (f == 0.9 ? "0.9" : "0.7"),
			      pe.calculateLDOS({{0}}).getData());
		}
	}

	{	// The calls of Templates/BasicChebyshev/src/main.cpp:91-116 on its
		// model (100x100 with spin, mu = -1, t = 1, open edges) with the
		// window [-4.99, 4.99] (the template's [-5, 5] at s = 5 exits in the
		// reference and here alike).
		const int SIZE_X = 100, SIZE_Y = 100;
		const complex<double> mu = -1.0, t = 1.0;
		Model model;
		model.setVerbose(false);
		for (int x = 0; x < SIZE_X; x++)
			for (int y = 0; y < SIZE_Y; y++)
				for (int s = 0; s < 2; s++) {
					model << HoppingAmplitude(-mu, {x, y, s}, {x, y, s});
					if (x + 1 < SIZE_X) model << HoppingAmplitude(-t, {x + 1, y, s}, {x, y, s}) + HC;
					if (y + 1 < SIZE_Y) model << HoppingAmplitude(-t, {x, y + 1, s}, {x, y, s}) + HC;
				}
		model.construct();
		Solver::ChebyshevExpander solver;
		solver.setVerbose(false);
		solver.setModel(model);
		solver.setScaleFactor(5.);
		solver.setCalculateCoefficientsOnGPU(gpuFlags);
		solver.setGenerateGreensFunctionsOnGPU(gpuFlags);
		solver.setUseLookupTable(true);
		solver.setNumCoefficients(500);
		PropertyExtractor::ChebyshevExpander propertyExtractor(solver);
		propertyExtractor.setEnergyWindow(-4.99, 4.99, 1000);
		Property::LDOS ldos = propertyExtractor.calculateLDOS({{_a_, SIZE_Y / 2, IDX_SUM_ALL}});
		write("basic_chebyshev_template_ldos", ldos.getData());
	}

	if (argc > 2 && std::string(argv[2]) != "gpu") {	// Green's functions of given coefficients, 4 types
		std::vector<complex<double>> crand;
		FILE *f = fopen(argv[2], "rb");
		if (!f) {
			fprintf(stderr, "cannot read %s\n", argv[2]);
			return 2;
		}
		double v[2];
		while (fread(v, 8, 2, f) == 2) crand.push_back(complex<double>(v[0], v[1]));
		fclose(f);
		struct Case {
			const char *name;
			std::vector<complex<double>> c;
			double lo, hi;
			int E;
		} cases[2] = {{"c123", {1, 2, 3}, -5, 5, 10}, {"crand", crand, -8.7, 9.1, 37}};
		const char *names[4] = {"Advanced", "Retarded", "Principal", "NonPrincipal"};
		typedef Solver::ChebyshevExpander::Type T;
		const T types[4] = {T::Advanced, T::Retarded, T::Principal, T::NonPrincipal};
		for (const Case &cs : cases)
			for (int table = 0; table < 2; table++)
				for (int t = 0; t < 4; t++) {
					Solver::ChebyshevExpander solver;
					gpu(solver);
					solver.setScaleFactor(10);
					solver.setNumCoefficients((int)cs.c.size());
					lookup(solver, table);
					solver.setEnergyWindow(Range(cs.lo, cs.hi, cs.E));
					write(std::string("gf_types_") + cs.name + "_" + names[t] + "_lookup" +
					          std::to_string(table),
					      solver.generateGreensFunction(cs.c, types[t]));
				}
	}
	printf("pe_outputs done\n");
	return 0;
}
