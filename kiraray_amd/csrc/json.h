// json.h - minimal JSON value + parser + writer for pass and scene configs.
//
// KiRaRay configures passes and scenes with JSON (from_json/to_json of each
// pass, src/render/wavefront/integrator.h:85-103; the scene importer,
// src/scene/krrscene.cpp:253-305).  This is a small self-contained reader for
// those documents: objects, arrays, strings (with \uXXXX escapes folded to
// UTF-8), numbers, booleans and null.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace krr {

class Json {
public:
	enum class Type { Null, Bool, Number, String, Array, Object };

	Json() = default;
	explicit Json(bool b) : type_(Type::Bool), b_(b) {}
	explicit Json(double d) : type_(Type::Number), d_(d) {}
	explicit Json(const std::string &s) : type_(Type::String), s_(s) {}
	static Json array() { Json j; j.type_ = Type::Array; return j; }
	static Json object() { Json j; j.type_ = Type::Object; return j; }

	Type type() const { return type_; }
	bool isNull() const { return type_ == Type::Null; }
	bool isBool() const { return type_ == Type::Bool; }
	bool isNumber() const { return type_ == Type::Number; }
	bool isString() const { return type_ == Type::String; }
	bool isArray() const { return type_ == Type::Array; }
	bool isObject() const { return type_ == Type::Object; }

	bool asBool() const { need(Type::Bool, "bool"); return b_; }
	double asNumber() const { need(Type::Number, "number"); return d_; }
	const std::string &asString() const { need(Type::String, "string"); return s_; }
	const std::vector<Json> &items() const { need(Type::Array, "array"); return a_; }
	const std::map<std::string, Json> &members() const { need(Type::Object, "object"); return o_; }
	size_t size() const { return isArray() ? a_.size() : (isObject() ? o_.size() : 0); }
	const Json &operator[](size_t i) const { need(Type::Array, "array"); return a_.at(i); }

	bool contains(const std::string &k) const { return isObject() && o_.count(k) > 0; }
	const Json &at(const std::string &k) const {
		need(Type::Object, "object");
		auto it = o_.find(k);
		if (it == o_.end()) throw std::runtime_error("missing key \"" + k + "\"");
		return it->second;
	}
	// j.value(key, default) with type checking of present keys.
	double value(const std::string &k, double def) const {
		if (!contains(k)) return def;
		const Json &v = o_.at(k);
		if (v.isBool()) return v.b_ ? 1.0 : 0.0;
		if (!v.isNumber()) throw std::runtime_error("key \"" + k + "\" must be a number");
		return v.d_;
	}
	bool value(const std::string &k, bool def) const {
		if (!contains(k)) return def;
		const Json &v = o_.at(k);
		if (v.isNumber()) return v.d_ != 0;
		if (!v.isBool()) throw std::runtime_error("key \"" + k + "\" must be a boolean");
		return v.b_;
	}
	std::string value(const std::string &k, const char *def) const {
		if (!contains(k)) return def;
		const Json &v = o_.at(k);
		if (!v.isString()) throw std::runtime_error("key \"" + k + "\" must be a string");
		return v.s_;
	}
	// Fixed-length numeric array (e.g. a vector3) with default.
	template <int N> void vec(const std::string &k, float (&out)[N]) const {
		if (!contains(k)) return;
		const Json &v = o_.at(k);
		if (v.isNumber()) {
			for (int i = 0; i < N; i++) out[i] = (float) v.d_;
			return;
		}
		if (!v.isArray() || v.a_.size() != (size_t) N) throw std::runtime_error("key \"" + k + "\" must be an array of " + std::to_string(N) + " numbers");
		for (int i = 0; i < N; i++) out[i] = (float) v.a_[i].asNumber();
	}

	void push(const Json &v) { need(Type::Array, "array"); a_.push_back(v); }
	void set(const std::string &k, const Json &v) { need(Type::Object, "object"); o_[k] = v; }

	static Json parse(const std::string &text) {
		Parser p{text.c_str(), text.c_str() + text.size()};
		p.ws();
		Json j = p.value();
		p.ws();
		if (p.s != p.e) p.fail("trailing characters");
		return j;
	}

	std::string dump() const {
		std::string out;
		write(out);
		return out;
	}

private:
	void need(Type t, const char *what) const {
		if (type_ != t) throw std::runtime_error(std::string("JSON value is not a ") + what);
	}
	void write(std::string &out) const {
		switch (type_) {
			case Type::Null: out += "null"; break;
			case Type::Bool: out += b_ ? "true" : "false"; break;
			case Type::Number: {
				char buf[64];
				if (std::isfinite(d_) && d_ == std::floor(d_) && std::fabs(d_) < 1e15) std::snprintf(buf, sizeof(buf), "%.0f", d_);
				else std::snprintf(buf, sizeof(buf), "%.9g", d_);
				out += buf;
				break;
			}
			case Type::String: writeString(out, s_); break;
			case Type::Array:
				out += "[";
				for (size_t i = 0; i < a_.size(); i++) {
					if (i) out += ",";
					a_[i].write(out);
				}
				out += "]";
				break;
			case Type::Object: {
				out += "{";
				bool first = true;
				for (auto &kv : o_) {
					if (!first) out += ",";
					first = false;
					writeString(out, kv.first);
					out += ":";
					kv.second.write(out);
				}
				out += "}";
				break;
			}
		}
	}
	static void writeString(std::string &out, const std::string &s) {
		out += "\"";
		for (char c : s) {
			if (c == '"' || c == '\\') { out += '\\'; out += c; }
			else if ((unsigned char) c < 0x20) {
				char buf[8];
				std::snprintf(buf, sizeof(buf), "\\u%04x", (unsigned) c);
				out += buf;
			} else out += c;
		}
		out += "\"";
	}

	struct Parser {
		const char *s, *e;
		[[noreturn]] void fail(const char *msg) { throw std::runtime_error(std::string("JSON parse error: ") + msg); }
		void ws() { while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) s++; }
		bool lit(const char *w) {
			size_t n = std::char_traits<char>::length(w);
			if ((size_t) (e - s) >= n && std::string(s, n) == w) { s += n; return true; }
			return false;
		}
		Json value() {
			if (s >= e) fail("unexpected end");
			char c = *s;
			if (c == '{') return object();
			if (c == '[') return arr();
			if (c == '"') return Json(str());
			if (lit("true")) return Json(true);
			if (lit("false")) return Json(false);
			if (lit("null")) return Json();
			if (c == '-' || (c >= '0' && c <= '9')) return num();
			fail("unexpected character");
		}
		Json num() {
			const char *b = s;
			if (*s == '-') s++;
			while (s < e && ((*s >= '0' && *s <= '9') || *s == '.' || *s == 'e' || *s == 'E' || *s == '+' || *s == '-')) s++;
			std::string t(b, s);
			char *end = nullptr;
			double d  = std::strtod(t.c_str(), &end);
			if (!end || *end) fail("bad number");
			return Json(d);
		}
		static void utf8(std::string &o, unsigned cp) {
			if (cp < 0x80) o += (char) cp;
			else if (cp < 0x800) { o += (char) (0xC0 | (cp >> 6)); o += (char) (0x80 | (cp & 0x3F)); }
			else { o += (char) (0xE0 | (cp >> 12)); o += (char) (0x80 | ((cp >> 6) & 0x3F)); o += (char) (0x80 | (cp & 0x3F)); }
		}
		std::string str() {
			s++; // opening quote
			std::string o;
			while (true) {
				if (s >= e) fail("unterminated string");
				char c = *s++;
				if (c == '"') break;
				if (c != '\\') { o += c; continue; }
				if (s >= e) fail("bad escape");
				char x = *s++;
				switch (x) {
					case '"': o += '"'; break;
					case '\\': o += '\\'; break;
					case '/': o += '/'; break;
					case 'b': o += '\b'; break;
					case 'f': o += '\f'; break;
					case 'n': o += '\n'; break;
					case 'r': o += '\r'; break;
					case 't': o += '\t'; break;
					case 'u': {
						if (e - s < 4) fail("bad \\u escape");
						unsigned cp = (unsigned) std::strtoul(std::string(s, 4).c_str(), nullptr, 16);
						s += 4;
						utf8(o, cp);
						break;
					}
					default: fail("bad escape");
				}
			}
			return o;
		}
		Json arr() {
			s++;
			Json j = Json::array();
			ws();
			if (s < e && *s == ']') { s++; return j; }
			while (true) {
				ws();
				j.a_.push_back(value());
				ws();
				if (s < e && *s == ',') { s++; continue; }
				if (s < e && *s == ']') { s++; return j; }
				fail("expected , or ]");
			}
		}
		Json object() {
			s++;
			Json j = Json::object();
			ws();
			if (s < e && *s == '}') { s++; return j; }
			while (true) {
				ws();
				if (s >= e || *s != '"') fail("expected key");
				std::string k = str();
				ws();
				if (s >= e || *s != ':') fail("expected :");
				s++;
				ws();
				j.o_[k] = value();
				ws();
				if (s < e && *s == ',') { s++; continue; }
				if (s < e && *s == '}') { s++; return j; }
				fail("expected , or }");
			}
		}
	};

	Type type_ = Type::Null;
	bool b_	   = false;
	double d_  = 0;
	std::string s_;
	std::vector<Json> a_;
	std::map<std::string, Json> o_;
};

} // namespace krr
