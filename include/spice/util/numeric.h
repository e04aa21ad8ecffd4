// Compensated (Kahan) summation of the simulation time (API of reference
// util/numeric.h).  operator+= returns the compensated increment y, which the
// engine hands to the models as that step's dt.
#pragma once

#include <concepts>

namespace spice::util {
template <std::floating_point Real>
class kahan_sum {
public:
	constexpr Real operator+=(Real delta) {
#if defined(__clang__) && !defined(__HIP_DEVICE_COMPILE__)
#pragma float_control(precise, on)
#endif
		Real const y = delta - _c;
		Real const t = _sum + y;
		_c           = (t - _sum) - y;
		_sum         = t;
		return y;
	}

	constexpr operator Real() const { return _sum; }

	constexpr void reset() { _sum = 0; }

private:
	Real _c   = 0;
	Real _sum = 0;
};
}  // namespace spice::util
