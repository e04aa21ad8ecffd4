// Runtime-bool to template-bool dispatch (API of reference util/meta.h).
#pragma once

#include <utility>

namespace spice::util {
template <class F, class... Args>
constexpr auto invoke(bool const b, F&& f, Args&&... args) {
	if (b)
		std::forward<F>(f).template operator()<true>(std::forward<Args>(args)...);
	else
		std::forward<F>(f).template operator()<false>(std::forward<Args>(args)...);
}

template <class F, class... Args>
constexpr auto invoke(bool const a, bool const b, F&& f, Args&&... args) {
	if (a)
		invoke(b, [&]<bool B>() { std::forward<F>(f).template operator()<true, B>(std::forward<Args>(args)...); });
	else
		invoke(b, [&]<bool B>() { std::forward<F>(f).template operator()<false, B>(std::forward<Args>(args)...); });
}
}  // namespace spice::util
