// Integer aliases and a 128-bit seed value (API of reference util/stdint.h).
#pragma once

#include <cstdint>

using Int8   = std::int8_t;
using Int16  = std::int16_t;
using Int32  = std::int32_t;
using Int    = std::int64_t;
using UInt8  = std::uint8_t;
using UInt16 = std::uint16_t;
using UInt32 = std::uint32_t;
using UInt   = std::uint64_t;

constexpr UInt operator""_u64(unsigned long long int x) { return UInt(x); }

// 128-bit unsigned value with a 64-bit add (carry into hi), as seeds are kept.
struct UInt128 {
	UInt lo;
	UInt hi;

	constexpr UInt128 operator+(UInt const n) const {
		UInt const l = lo + n;
		return {l, hi + (l < lo ? 1u : 0u)};
	}
};
