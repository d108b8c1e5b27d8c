// json.h -- minimal JSON reader for plans crossing the C boundary (objects, arrays, strings with escapes,
// numbers, true/false/null). Throws std::runtime_error with the byte offset on malformed input.
#pragma once

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace vgpu {
namespace json {

struct Value {
	enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
	bool b = false;
	double num = 0;
	bool is_uint = false; // the literal was a non-negative integer that fits 64 bits: `uint` holds it exactly
	uint64_t uint = 0;
	std::string str;
	std::vector<Value> arr;
	std::map<std::string, Value> obj;

	bool Has(const std::string &k) const {
		return kind == OBJ && obj.count(k) != 0;
	}
	const Value &At(const std::string &k) const {
		static const Value null_value;
		if (kind != OBJ) {
			return null_value;
		}
		auto it = obj.find(k);
		return it == obj.end() ? null_value : it->second;
	}
	std::string Str(const std::string &k, const std::string &def = "") const {
		const Value &v = At(k);
		return v.kind == STR ? v.str : def;
	}
	double Num(const std::string &k, double def = 0) const {
		const Value &v = At(k);
		return v.kind == NUM ? v.num : def;
	}
	//! Exact for integer literals up to 2^64 - 1 (a double would round them past 2^53).
	uint64_t U64(const std::string &k, uint64_t def = 0) const {
		const Value &v = At(k);
		if (v.kind != NUM) {
			return def;
		}
		if (v.is_uint) {
			return v.uint;
		}
		if (!(v.num >= 0.0)) {
			return 0;
		}
		return v.num >= 18446744073709551615.0 ? ~0ull : static_cast<uint64_t>(v.num);
	}
	bool Bool(const std::string &k, bool def = false) const {
		const Value &v = At(k);
		return v.kind == BOOL ? v.b : def;
	}
	std::vector<std::string> StrList(const std::string &k) const {
		std::vector<std::string> out;
		const Value &v = At(k);
		if (v.kind == ARR) {
			for (auto &e : v.arr) {
				out.push_back(e.kind == STR ? e.str : std::string());
			}
		}
		return out;
	}
};

class Parser {
public:
	explicit Parser(const std::string &text) : s_(text) {
	}
	Value Parse() {
		Value v = ParseValue();
		SkipWs();
		if (i_ != s_.size()) {
			Fail("trailing characters");
		}
		return v;
	}

private:
	[[noreturn]] void Fail(const char *what) {
		throw std::runtime_error(std::string("plan JSON: ") + what + " at offset " + std::to_string(i_));
	}
	void SkipWs() {
		while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\t' || s_[i_] == '\r')) {
			i_++;
		}
	}
	Value ParseValue() {
		SkipWs();
		if (i_ >= s_.size()) {
			Fail("unexpected end");
		}
		char c = s_[i_];
		Value v;
		if (c == '{') {
			v.kind = Value::OBJ;
			i_++;
			SkipWs();
			if (i_ < s_.size() && s_[i_] == '}') {
				i_++;
				return v;
			}
			while (true) {
				SkipWs();
				if (i_ >= s_.size() || s_[i_] != '"') {
					Fail("expected key");
				}
				std::string key = ParseString();
				SkipWs();
				if (i_ >= s_.size() || s_[i_] != ':') {
					Fail("expected ':'");
				}
				i_++;
				v.obj[key] = ParseValue();
				SkipWs();
				if (i_ < s_.size() && s_[i_] == ',') {
					i_++;
					continue;
				}
				if (i_ < s_.size() && s_[i_] == '}') {
					i_++;
					return v;
				}
				Fail("expected ',' or '}'");
			}
		}
		if (c == '[') {
			v.kind = Value::ARR;
			i_++;
			SkipWs();
			if (i_ < s_.size() && s_[i_] == ']') {
				i_++;
				return v;
			}
			while (true) {
				v.arr.push_back(ParseValue());
				SkipWs();
				if (i_ < s_.size() && s_[i_] == ',') {
					i_++;
					continue;
				}
				if (i_ < s_.size() && s_[i_] == ']') {
					i_++;
					return v;
				}
				Fail("expected ',' or ']'");
			}
		}
		if (c == '"') {
			v.kind = Value::STR;
			v.str = ParseString();
			return v;
		}
		if (s_.compare(i_, 4, "true") == 0) {
			v.kind = Value::BOOL;
			v.b = true;
			i_ += 4;
			return v;
		}
		if (s_.compare(i_, 5, "false") == 0) {
			v.kind = Value::BOOL;
			i_ += 5;
			return v;
		}
		if (s_.compare(i_, 4, "null") == 0) {
			i_ += 4;
			return v;
		}
		size_t start = i_;
		while (i_ < s_.size() && (std::isdigit((unsigned char)s_[i_]) || s_[i_] == '-' || s_[i_] == '+' ||
		                          s_[i_] == '.' || s_[i_] == 'e' || s_[i_] == 'E')) {
			i_++;
		}
		if (start == i_) {
			Fail("unexpected character");
		}
		v.kind = Value::NUM;
		const std::string lit = s_.substr(start, i_ - start);
		try {
			v.num = std::stod(lit);
		} catch (const std::exception &) {
			Fail("bad number");
		}
		if (lit.find_first_not_of("0123456789") == std::string::npos) {
			errno = 0;
			char *end = nullptr;
			unsigned long long u = std::strtoull(lit.c_str(), &end, 10);
			if (errno == 0 && end != nullptr && *end == '\0') {
				v.is_uint = true;
				v.uint = u;
			}
		}
		return v;
	}
	std::string ParseString() {
		std::string out;
		i_++; // opening quote
		while (i_ < s_.size() && s_[i_] != '"') {
			char c = s_[i_++];
			if (c != '\\') {
				out.push_back(c);
				continue;
			}
			if (i_ >= s_.size()) {
				Fail("bad escape");
			}
			char e = s_[i_++];
			switch (e) {
			case 'n':
				out.push_back('\n');
				break;
			case 't':
				out.push_back('\t');
				break;
			case 'r':
				out.push_back('\r');
				break;
			case 'b':
				out.push_back('\b');
				break;
			case 'f':
				out.push_back('\f');
				break;
			case 'u': {
				if (i_ + 4 > s_.size()) {
					Fail("bad \\u escape");
				}
				unsigned code = (unsigned)std::stoul(s_.substr(i_, 4), nullptr, 16);
				i_ += 4;
				if (code < 0x80) {
					out.push_back((char)code);
				} else if (code < 0x800) {
					out.push_back((char)(0xC0 | (code >> 6)));
					out.push_back((char)(0x80 | (code & 0x3F)));
				} else {
					out.push_back((char)(0xE0 | (code >> 12)));
					out.push_back((char)(0x80 | ((code >> 6) & 0x3F)));
					out.push_back((char)(0x80 | (code & 0x3F)));
				}
				break;
			}
			default:
				out.push_back(e);
			}
		}
		if (i_ >= s_.size()) {
			Fail("unterminated string");
		}
		i_++;
		return out;
	}
	const std::string &s_;
	size_t i_ = 0;
};

inline Value Parse(const std::string &text) {
	return Parser(text).Parse();
}

} // namespace json
} // namespace vgpu
