/* JSON values for the Serializable::Mode::JSON serializations of the
 * property containers. dump() writes the compact form the reference's
 * serializations have (json/TBTK/json.hpp, nlohmann 2.1.1, dump() without
 * indentation): object members in sorted key order, no whitespace, doubles
 * as printf("%.15g") with ".0" appended to integral values ("0.0" / "-0.0"
 * for zeros, "null" for non-finite values), integers in decimal, strings
 * with \", \\, \b, \f, \n, \r, \t and \u00XX escapes. parse() reads any
 * JSON text of that form. */
#ifndef TBTK_AMD_JSON
#define TBTK_AMD_JSON

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace TBTK {

class JsonValue {
public:
	enum class Kind { Null, Bool, Integer, Unsigned, Float, String, Array, Object };

	JsonValue() : kind(Kind::Null) {}
	JsonValue(bool b) : kind(Kind::Bool), b(b) {}
	JsonValue(int i) : kind(Kind::Integer), i(i) {}
	JsonValue(long i) : kind(Kind::Integer), i(i) {}
	JsonValue(long long i) : kind(Kind::Integer), i(i) {}
	JsonValue(unsigned int u) : kind(Kind::Unsigned), u(u) {}
	JsonValue(unsigned long u) : kind(Kind::Unsigned), u(u) {}
	JsonValue(unsigned long long u) : kind(Kind::Unsigned), u(u) {}
	JsonValue(double d) : kind(Kind::Float), d(d) {}
	JsonValue(const char *s) : kind(Kind::String), s(s) {}
	JsonValue(const std::string &s) : kind(Kind::String), s(s) {}

	Kind getKind() const { return kind; }
	/* Object member (created, and the value made an object, if needed). */
	JsonValue &operator[](const std::string &key);
	/* Append to an array (a null value becomes an array). */
	void push_back(const JsonValue &v);

	/* Member access for reading; exits when absent or of the wrong kind. */
	const JsonValue &at(const std::string &key) const;
	bool has(const std::string &key) const;
	const std::vector<JsonValue> &elements() const;
	bool getBool() const;
	int64_t getInt() const;      // Integer or Unsigned
	double getDouble() const;    // any number
	const std::string &getString() const;

	std::string dump() const;
	static JsonValue parse(const std::string &text);

private:
	Kind kind;
	bool b = false;
	int64_t i = 0;
	uint64_t u = 0;
	double d = 0;
	std::string s;
	std::vector<JsonValue> a;
	std::map<std::string, JsonValue> o;

	void dumpTo(std::string &out) const;
	friend class JsonParser;
};

}  // namespace TBTK
#endif
