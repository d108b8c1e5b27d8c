// check.hpp -- minimal assertion harness for the C++ drop-in suites (counts checks, reports failures,
// keeps running after a failed check or an escaped exception, exits non-zero when anything failed).
#pragma once

#include <exception>
#include <iostream>
#include <string>

namespace vcheck {
inline int &Failures() {
	static int n = 0;
	return n;
}
inline int &Checks() {
	static int n = 0;
	return n;
}
inline void Report(bool ok, const std::string &what, const char *file, int line) {
	Checks()++;
	if (!ok) {
		Failures()++;
		std::cerr << "[FAIL] " << file << ":" << line << "  " << what << "\n";
	}
}
} // namespace vcheck

#define CHECK(expr) vcheck::Report(static_cast<bool>(expr), "CHECK(" #expr ")", __FILE__, __LINE__)

#define CHECK_THROWS(stmt)                                                                                           \
	do {                                                                                                             \
		bool threw_ = false;                                                                                         \
		try {                                                                                                        \
			stmt;                                                                                                    \
		} catch (const std::exception &) {                                                                           \
			threw_ = true;                                                                                           \
		}                                                                                                            \
		vcheck::Report(threw_, "throws: " #stmt, __FILE__, __LINE__);                                                \
	} while (0)

#define RUN_TEST(fn)                                                                                                 \
	do {                                                                                                             \
		std::cout << "--- " #fn "\n" << std::flush;                                                                  \
		try {                                                                                                        \
			fn();                                                                                                    \
		} catch (const std::exception &e) {                                                                          \
			vcheck::Report(false, std::string("exception escaped " #fn ": ") + e.what(), __FILE__, __LINE__);        \
		} catch (...) {                                                                                              \
			vcheck::Report(false, "non-std exception escaped " #fn, __FILE__, __LINE__);                             \
		}                                                                                                            \
	} while (0)

#define TEST_MAIN_EPILOGUE()                                                                                         \
	do {                                                                                                             \
		std::cout << (vcheck::Checks() - vcheck::Failures()) << "/" << vcheck::Checks() << " checks passed\n";      \
		return vcheck::Failures() == 0 ? 0 : 1;                                                                      \
	} while (0)
