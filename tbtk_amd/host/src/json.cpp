/* JsonValue: compact dump in the form of the reference's serializations and
 * a recursive-descent parser. */
#include "TBTK/Json.h"

#include "TBTK/TBTKMacros.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace TBTK {

JsonValue &JsonValue::operator[](const std::string &key) {
	if (kind != Kind::Object) {
		TBTKAssert(kind == Kind::Null, "JsonValue::operator[]()", "Not an object.", "");
		kind = Kind::Object;
	}
	return o[key];
}

void JsonValue::push_back(const JsonValue &v) {
	if (kind != Kind::Array) {
		TBTKAssert(kind == Kind::Null, "JsonValue::push_back()", "Not an array.", "");
		kind = Kind::Array;
	}
	a.push_back(v);
}

const JsonValue &JsonValue::at(const std::string &key) const {
	TBTKAssert(kind == Kind::Object, "JsonValue::at()", "Not an object.", "");
	auto it = o.find(key);
	TBTKAssert(it != o.end(), "JsonValue::at()", "Missing member '" << key << "'.", "");
	return it->second;
}

bool JsonValue::has(const std::string &key) const {
	return kind == Kind::Object && o.count(key) != 0;
}

const std::vector<JsonValue> &JsonValue::elements() const {
	TBTKAssert(kind == Kind::Array, "JsonValue::elements()", "Not an array.", "");
	return a;
}

bool JsonValue::getBool() const {
	TBTKAssert(kind == Kind::Bool, "JsonValue::getBool()", "Not a boolean.", "");
	return b;
}

int64_t JsonValue::getInt() const {
	if (kind == Kind::Unsigned) return (int64_t)u;
	TBTKAssert(kind == Kind::Integer, "JsonValue::getInt()", "Not an integer.", "");
	return i;
}

double JsonValue::getDouble() const {
	if (kind == Kind::Integer) return (double)i;
	if (kind == Kind::Unsigned) return (double)u;
	TBTKAssert(kind == Kind::Float, "JsonValue::getDouble()", "Not a number.", "");
	return d;
}

const std::string &JsonValue::getString() const {
	TBTKAssert(kind == Kind::String, "JsonValue::getString()", "Not a string.", "");
	return s;
}

static void dumpString(const std::string &str, std::string &out) {
	out += '"';
	for (unsigned char c : str) {
		switch (c) {
		case '"': out += "\\\""; break;
		case '\\': out += "\\\\"; break;
		case '\b': out += "\\b"; break;
		case '\f': out += "\\f"; break;
		case '\n': out += "\\n"; break;
		case '\r': out += "\\r"; break;
		case '\t': out += "\\t"; break;
		default:
			if (c < 0x20) {
				char buf[8];
				snprintf(buf, sizeof buf, "\\u%04x", c);
				out += buf;
			} else {
				out += (char)c;
			}
		}
	}
	out += '"';
}

static void dumpFloat(double x, std::string &out) {
	if (!std::isfinite(x)) {
		out += "null";
		return;
	}
	if (x == 0) {
		out += std::signbit(x) ? "-0.0" : "0.0";
		return;
	}
	char buf[64];
	const int len = snprintf(buf, sizeof buf, "%.*g", 15, x);
	out.append(buf, len);
	bool intLike = true;
	for (int k = 0; k < len; k++)
		if (buf[k] == '.' || buf[k] == 'e') intLike = false;
	if (intLike) out += ".0";
}

void JsonValue::dumpTo(std::string &out) const {
	switch (kind) {
	case Kind::Null: out += "null"; break;
	case Kind::Bool: out += b ? "true" : "false"; break;
	case Kind::Integer: out += std::to_string(i); break;
	case Kind::Unsigned: out += std::to_string(u); break;
	case Kind::Float: dumpFloat(d, out); break;
	case Kind::String: dumpString(s, out); break;
	case Kind::Array:
		out += '[';
		for (size_t k = 0; k < a.size(); k++) {
			if (k) out += ',';
			a[k].dumpTo(out);
		}
		out += ']';
		break;
	case Kind::Object: {
		out += '{';
		bool first = true;
		for (const auto &kv : o) {
			if (!first) out += ',';
			first = false;
			dumpString(kv.first, out);
			out += ':';
			kv.second.dumpTo(out);
		}
		out += '}';
		break;
	}
	}
}

std::string JsonValue::dump() const {
	std::string out;
	dumpTo(out);
	return out;
}

class JsonParser {
public:
	explicit JsonParser(const std::string &t) : t(t) {}

	JsonValue document() {
		JsonValue v = value();
		ws();
		TBTKAssert(p == t.size(), "JsonValue::parse()", "Trailing characters at " << p << ".", "");
		return v;
	}

private:
	const std::string &t;
	size_t p = 0;

	void ws() {
		while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) p++;
	}
	void expect(char c) {
		ws();
		TBTKAssert(p < t.size() && t[p] == c, "JsonValue::parse()",
		           "Expected '" << c << "' at " << p << ".", "");
		p++;
	}
	bool literal(const char *w) {
		size_t n = 0;
		while (w[n]) n++;
		if (t.compare(p, n, w) == 0) {
			p += n;
			return true;
		}
		return false;
	}
	std::string str() {
		expect('"');
		std::string out;
		while (true) {
			TBTKAssert(p < t.size(), "JsonValue::parse()", "Unterminated string.", "");
			const char c = t[p++];
			if (c == '"') break;
			if (c != '\\') {
				out += c;
				continue;
			}
			TBTKAssert(p < t.size(), "JsonValue::parse()", "Unterminated escape.", "");
			const char e = t[p++];
			switch (e) {
			case '"': out += '"'; break;
			case '\\': out += '\\'; break;
			case '/': out += '/'; break;
			case 'b': out += '\b'; break;
			case 'f': out += '\f'; break;
			case 'n': out += '\n'; break;
			case 'r': out += '\r'; break;
			case 't': out += '\t'; break;
			case 'u': {
				TBTKAssert(p + 4 <= t.size(), "JsonValue::parse()", "Short \\u escape.", "");
				const unsigned cp = (unsigned)strtoul(t.substr(p, 4).c_str(), nullptr, 16);
				p += 4;
				if (cp < 0x80) {
					out += (char)cp;
				} else if (cp < 0x800) {
					out += (char)(0xc0 | (cp >> 6));
					out += (char)(0x80 | (cp & 0x3f));
				} else {
					out += (char)(0xe0 | (cp >> 12));
					out += (char)(0x80 | ((cp >> 6) & 0x3f));
					out += (char)(0x80 | (cp & 0x3f));
				}
				break;
			}
			default:
				TBTKExit("JsonValue::parse()", "Bad escape '\\" << e << "'.", "");
			}
		}
		return out;
	}
	JsonValue number() {
		const size_t start = p;
		if (t[p] == '-') p++;
		bool isFloat = false;
		while (p < t.size()) {
			const char c = t[p];
			if (c >= '0' && c <= '9') {
				p++;
			} else if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') {
				isFloat = true;
				p++;
			} else {
				break;
			}
		}
		const std::string n = t.substr(start, p - start);
		TBTKAssert(!n.empty() && n != "-", "JsonValue::parse()", "Bad number at " << start << ".", "");
		if (isFloat) return JsonValue(strtod(n.c_str(), nullptr));
		if (n[0] == '-') return JsonValue((long long)strtoll(n.c_str(), nullptr, 10));
		return JsonValue((unsigned long long)strtoull(n.c_str(), nullptr, 10));
	}
	JsonValue value() {
		ws();
		TBTKAssert(p < t.size(), "JsonValue::parse()", "Unexpected end of text.", "");
		const char c = t[p];
		if (c == '{') {
			p++;
			JsonValue v;
			v.kind = JsonValue::Kind::Object;
			ws();
			if (p < t.size() && t[p] == '}') {
				p++;
				return v;
			}
			while (true) {
				const std::string key = str();
				expect(':');
				v.o[key] = value();
				ws();
				TBTKAssert(p < t.size(), "JsonValue::parse()", "Unterminated object.", "");
				if (t[p] == ',') {
					p++;
					continue;
				}
				expect('}');
				return v;
			}
		}
		if (c == '[') {
			p++;
			JsonValue v;
			v.kind = JsonValue::Kind::Array;
			ws();
			if (p < t.size() && t[p] == ']') {
				p++;
				return v;
			}
			while (true) {
				v.a.push_back(value());
				ws();
				TBTKAssert(p < t.size(), "JsonValue::parse()", "Unterminated array.", "");
				if (t[p] == ',') {
					p++;
					continue;
				}
				expect(']');
				return v;
			}
		}
		if (c == '"') return JsonValue(str());
		if (literal("true")) return JsonValue(true);
		if (literal("false")) return JsonValue(false);
		if (literal("null")) return JsonValue();
		return number();
	}
};

JsonValue JsonValue::parse(const std::string &text) {
	JsonParser parser(text);
	return parser.document();
}

}  // namespace TBTK
