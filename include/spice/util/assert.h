// Precondition / invariant checks (API of reference util/assert.h).  A failed
// check throws std::logic_error("Assertion failed (<file>:<line>): <condition>").
// Checks are compiled in when SPICE_ASSERT_PRECONDITIONS / SPICE_ASSERT_INVARIANTS
// are defined and only evaluated otherwise.  Device code never checks.
#pragma once

#include <stdexcept>
#include <string>

namespace spice::util::detail {
[[noreturn]] inline void assert_failed(char const* file, int const line, char const* condition) {
	std::string f(file);
	auto const slash = f.find_last_of('/');
	if (slash != std::string::npos) f = f.substr(slash + 1);
	throw std::logic_error("Assertion failed (" + f + ":" + std::to_string(line) + "): " + condition);
}
}  // namespace spice::util::detail

#if defined(__HIP_DEVICE_COMPILE__)
#define SPICE_ASSERT(X) (void)(X);
#else
#define SPICE_ASSERT(X)            \
	if (__builtin_expect(!(X), 0)) \
		::spice::util::detail::assert_failed(__FILE__, __LINE__, #X);
#endif

#ifdef SPICE_ASSERT_PRECONDITIONS
#define SPICE_PRE(X) SPICE_ASSERT(X)
#else
#define SPICE_PRE(X) (void)(X);
#endif

#ifdef SPICE_ASSERT_INVARIANTS
#define SPICE_INV(X) SPICE_ASSERT(X)
#else
#define SPICE_INV(X) (void)(X);
#endif
