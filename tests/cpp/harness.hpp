// harness.hpp -- self-registering test cases for the C++ drop-in suites, reported in TAP form.
//
// VT_CASE(name) { ... } defines and registers a case; EXPECT(cond) records one expectation;
// EXPECT_THROWS(stmt, "fragment") expects a std::exception whose message contains the fragment ("" = any).
// A case that lets an exception escape fails, and the run continues with the next case. VT_MAIN() runs every
// case (or those whose name contains argv[1]), prints "ok N - name" / "not ok N - name" lines, then
// "# <passed>/<total> expectations met", and exits non-zero if anything failed.
#pragma once

#include <cstdio>
#include <exception>
#include <string>
#include <vector>

namespace vt {

struct Case {
	const char *name;
	void (*body)();
};

inline std::vector<Case> &Cases() {
	static std::vector<Case> cases;
	return cases;
}

struct Tally {
	long expectations = 0;
	long unmet = 0;
	std::vector<std::string> notes; // unmet expectations of the running case
};

inline Tally &Current() {
	static Tally tally;
	return tally;
}

struct Registrar {
	Registrar(const char *name, void (*body)()) {
		Cases().push_back({name, body});
	}
};

inline void Expect(bool met, const std::string &what, const char *file, int line) {
	Tally &t = Current();
	t.expectations++;
	if (!met) {
		t.unmet++;
		t.notes.push_back(std::string(file) + ":" + std::to_string(line) + ": expected " + what);
	}
}

inline int Run(int argc, char **argv) {
	const std::string filter = argc > 1 ? argv[1] : "";
	int index = 0, failed_cases = 0;
	for (const Case &c : Cases()) {
		if (!filter.empty() && std::string(c.name).find(filter) == std::string::npos) {
			continue;
		}
		index++;
		Tally &t = Current();
		t.notes.clear();
		const long before = t.unmet;
		try {
			c.body();
		} catch (const std::exception &e) {
			Expect(false, std::string("no exception, got: ") + e.what(), c.name, 0);
		} catch (...) {
			Expect(false, "no exception, got a non-standard one", c.name, 0);
		}
		const bool ok = t.unmet == before;
		failed_cases += ok ? 0 : 1;
		std::printf("%s %d - %s\n", ok ? "ok" : "not ok", index, c.name);
		for (const auto &n : t.notes) {
			std::printf("#   %s\n", n.c_str());
		}
		std::fflush(stdout);
	}
	const Tally &t = Current();
	std::printf("1..%d\n# %ld/%ld expectations met, %d case(s) failed\n", index, t.expectations - t.unmet,
	            t.expectations, failed_cases);
	return failed_cases == 0 && index > 0 ? 0 : 1;
}

} // namespace vt

#define VT_CASE(name)                                                                                                \
	static void name();                                                                                              \
	static ::vt::Registrar name##_registrar(#name, &name);                                                           \
	static void name()

#define EXPECT(cond) ::vt::Expect(static_cast<bool>(cond), #cond, __FILE__, __LINE__)

#define EXPECT_THROWS(stmt, fragment)                                                                                \
	do {                                                                                                             \
		std::string vt_message_;                                                                                     \
		bool vt_thrown_ = false;                                                                                     \
		try {                                                                                                        \
			stmt;                                                                                                    \
		} catch (const std::exception &vt_e_) {                                                                      \
			vt_thrown_ = true;                                                                                       \
			vt_message_ = vt_e_.what();                                                                              \
		}                                                                                                            \
		::vt::Expect(vt_thrown_ && vt_message_.find(fragment) != std::string::npos,                                  \
		             std::string(#stmt " to throw \"") + (fragment) + "\" (got \"" + vt_message_ + "\")", __FILE__,   \
		             __LINE__);                                                                                      \
	} while (0)

#define VT_MAIN()                                                                                                    \
	int main(int argc, char **argv) {                                                                                \
		return ::vt::Run(argc, argv);                                                                                \
	}
