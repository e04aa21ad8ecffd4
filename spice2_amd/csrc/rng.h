// xoroshiro128+ (rotations 24/16/37) as the reference draws it (random.h:222-234),
// with GF(2) jump-ahead so that any position of a stream can be reached without
// stepping through the positions before it.
//
// The state update s -> T(s) is linear over GF(2)^128 (the output s0 + s1 is
// not, but it is computed from the state).  With P(x) the characteristic
// polynomial of T (degree 128, found once by Berlekamp-Massey on the host) and
// x^n mod P(x) = sum_i c_i x^i, Cayley-Hamilton gives T^n = sum_i c_i T^i, so
//     T^n s = XOR over {i : c_i = 1} of T^i s,
// which costs 128 steps of the generator for any n.  The coefficient sets c are
// computed on the host (square-and-multiply modulo P) and passed to kernels.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define SPICE_HD __host__ __device__ inline
#else
#define SPICE_HD inline
#endif

namespace spice_rt {

struct xoro_state {
	uint64_t s0, s1;
};

// Polynomial over GF(2) of degree < 128: bit i of (lo, hi) is the coefficient of x^i.
struct gf2poly {
	uint64_t lo, hi;
};

SPICE_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// One generator step; returns the output of the state before the step.
SPICE_HD uint64_t xoro_next(xoro_state& s) {
	uint64_t const out = s.s0 + s.s1;
	uint64_t const t   = s.s0 ^ s.s1;
	s.s0               = rotl64(s.s0, 24) ^ t ^ (t << 16);
	s.s1               = rotl64(t, 37);
	return out;
}

// T^n s for the n whose x^n mod P(x) is `c`.
SPICE_HD xoro_state xoro_jump(xoro_state s, gf2poly c) {
	xoro_state acc{0, 0};
	for (int i = 0; i < 128; i++) {
		uint64_t const bit = (i < 64 ? (c.lo >> i) : (c.hi >> (i - 64))) & 1u;
		uint64_t const m   = 0 - bit;
		acc.s0 ^= s.s0 & m;
		acc.s1 ^= s.s1 & m;
		xoro_next(s);
	}
	return acc;
}

// Reference random.h:236-247 for Real = float: (x >> 40) * 2^-24, in [0, 1).
SPICE_HD float canonical_f32(uint64_t x) { return float(x >> 40) / 16777216.0f; }
// Reference random.h:236-247 for Real = double, LeftOpen: ((x >> 11) + 1) * 2^-53, in (0, 1].
SPICE_HD double canonical_f64_leftopen(uint64_t x) {
	return double((x >> 11) + 1) / 9007199254740992.0;
}

// ---------------------------------------------------------------- host side
// Characteristic polynomial and x^n mod P; implemented in jump.cpp.
gf2poly xoro_charpoly();                 // P(x) without its x^128 term
gf2poly xoro_jump_poly(uint64_t n);      // x^n mod P(x)
gf2poly gf2_mulmod(gf2poly a, gf2poly b); // a * b mod P(x)
xoro_state xoro_jump_n(xoro_state s, uint64_t n);  // convenience: T^n s

}  // namespace spice_rt
