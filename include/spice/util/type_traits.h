// Small type utilities used by the model concepts (API of reference util/type_traits.h).
#pragma once

#include <type_traits>

namespace spice::util {
struct empty_t {};

template <class T>
constexpr bool is_empty_v = std::is_same_v<T, empty_t>;

// Converts to any lvalue reference; only used in unevaluated signature probes.
struct any_t {
	template <class T>
	constexpr operator T&() const;
};

template <class T, bool HasValue>
struct optional {
	using type = empty_t;
};
template <class T>
struct optional<T, true> {
	using type = T;
};
template <class T, bool HasValue>
using optional_t = typename optional<T, HasValue>::type;

template <bool... p>
constexpr bool none_of = (int(p) + ... + 0) == 0;
template <bool... p>
constexpr bool one_of = (int(p) + ... + 0) == 1;
template <bool... p>
constexpr bool up_to_one_of = (int(p) + ... + 0) <= 1;
template <bool... p>
constexpr bool any_of = (int(p) + ... + 0) >= 1;
template <bool... p>
constexpr bool all_of = (int(p) + ... + 0) == int(sizeof...(p));
}  // namespace spice::util
