// Reads a JSON serialization made by the reference (file argv[2]) of kind
// argv[1] (LDOS, DOS, Density, EnergyResolvedProperty = a Green's
// function, Range, IndexTree), builds this build's object from it and writes
// the object's own serialization to stdout. Kind "Magnetization" serializes
// a small Magnetization instead (which exits: SpinMatrix elements have no
// serialization, as in the reference).
#include "TBTK/Json.h"
#include "TBTK/Property/Property.h"

#include <complex>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

using namespace TBTK;

int main(int argc, char **argv) {
	if (argc < 2) {
		fprintf(stderr, "usage: json_roundtrip <kind> [file]\n");
		return 2;
	}
	const std::string kind = argv[1];
	const Serializable::Mode json = Serializable::Mode::JSON;
	if (kind == "Magnetization") {
		Property::Magnetization m(std::vector<int>{2});
		printf("%s", m.serialize(json).c_str());
		return 0;
	}
	if (argc < 3) return 2;
	std::ifstream f(argv[2]);
	std::stringstream ss;
	ss << f.rdbuf();
	const std::string text = ss.str();
	std::string out;
	if (kind == "LDOS") out = Property::LDOS(text, json).serialize(json);
	else if (kind == "DOS") out = Property::DOS(text, json).serialize(json);
	else if (kind == "Density") out = Property::Density(text, json).serialize(json);
	else if (kind == "EnergyResolvedProperty")
		out = Property::EnergyResolvedProperty<std::complex<double>>(JsonValue::parse(text))
		          .serialize(json);
	else if (kind == "Range") out = Range(text, json).serialize(json);
	else if (kind == "IndexTree") out = IndexTree(text, json).serialize(json);
	else {
		fprintf(stderr, "unknown kind %s\n", kind.c_str());
		return 2;
	}
	fwrite(out.data(), 1, out.size(), stdout);
	return 0;
}
