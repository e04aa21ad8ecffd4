// Correctly rounded natural logarithm of a double, written for bit-exact
// reproduction of the reference's exponential draws (random.h:270-272 calls
// std::log on a left-open canonical double).  The result is computed in
// double-double arithmetic with a relative error far below 2^-100 and then
// rounded once to double, so it equals the correctly rounded log(x) except when
// log(x) lies within ~2^-100 relative of a rounding midpoint.
//
// Method (standard, e.g. Muller, "Handbook of Floating-Point Arithmetic"):
//   x = m * 2^k with m in [sqrt(1/2), sqrt(2));
//   log(m) = 2 atanh(f), f = (m - 1) / (m + 1), |f| <= 0.1716;
//   atanh(f) = f * sum_{i>=0} f^(2i) / (2i + 1), 22 terms in double-double;
//   log(x) = k * ln2 + log(m), ln2 as a double-double constant.
//
// Usable from host C++ (g++) and from HIP device code.  Must be compiled with
// floating-point contraction off; every fused operation is an explicit fma().
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define SPICE_HD __host__ __device__ inline
#else
#define SPICE_HD inline
#endif

namespace spice_crlog {

struct dd {
	double hi, lo;
};

SPICE_HD dd two_sum(double a, double b) {
	double const s  = a + b;
	double const bb = s - a;
	double const e  = (a - (s - bb)) + (b - bb);
	return {s, e};
}

SPICE_HD dd fast_two_sum(double a, double b) {  // requires |a| >= |b| or a == 0
	double const s = a + b;
	return {s, b - (s - a)};
}

SPICE_HD dd two_prod(double a, double b) {
	double const p = a * b;
	return {p, std::fma(a, b, -p)};
}

SPICE_HD dd add(dd x, dd y) {
	dd s       = two_sum(x.hi, y.hi);
	dd t       = two_sum(x.lo, y.lo);
	s.lo      += t.hi;
	s          = fast_two_sum(s.hi, s.lo);
	s.lo      += t.lo;
	return fast_two_sum(s.hi, s.lo);
}

SPICE_HD dd mul(dd x, dd y) {
	dd p  = two_prod(x.hi, y.hi);
	p.lo += x.hi * y.lo + x.lo * y.hi;
	return fast_two_sum(p.hi, p.lo);
}

SPICE_HD dd mul_d(dd x, double y) {
	dd p  = two_prod(x.hi, y);
	p.lo += x.lo * y;
	return fast_two_sum(p.hi, p.lo);
}

// x / y in double-double: three correction steps of the long division.
SPICE_HD dd div(dd x, dd y) {
	double const q1 = x.hi / y.hi;
	dd r            = add(x, mul_d(y, -q1));
	double const q2 = r.hi / y.hi;
	r               = add(r, mul_d(y, -q2));
	double const q3 = r.hi / y.hi;
	dd q            = fast_two_sum(q1, q2);
	return add(q, dd{q3, 0.0});
}

SPICE_HD double from_bits(uint64_t b) {
	double d;
	std::memcpy(&d, &b, 8);
	return d;
}
SPICE_HD uint64_t to_bits(double d) {
	uint64_t b;
	std::memcpy(&b, &d, 8);
	return b;
}

// 1/(2i+1) as double-double: hi = RN(1/d), lo = RN((1 - hi*d)/d), where the
// residual 1 - hi*d is exact (one fma).
SPICE_HD dd inv_odd(int i) {
	double const d  = 2.0 * i + 1.0;
	double const hi = 1.0 / d;
	double const r  = -std::fma(hi, d, -1.0);  // exact residual 1 - hi*d
	return {hi, r / d};
}

// log(x) as an unevaluated double-double (hi + lo, |lo| <= ulp(hi)/2) with a
// relative error far below 2^-100, for positive finite x != 1.
SPICE_HD dd log_dd(double x) {
	uint64_t bits = to_bits(x);
	int k         = 0;
	if ((bits >> 52) == 0) {  // subnormal: scale into the normal range
		x *= 18014398509481984.0;  // 2^54
		bits = to_bits(x);
		k    = -54;
	}
	k += int((bits >> 52) & 0x7ff) - 1023;
	double m = from_bits((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
	if (m > 1.4142135623730951) {
		m *= 0.5;
		k += 1;
	}

	// f = (m - 1) / (m + 1); m - 1 is exact for m in [0.5, 2].
	dd const num = {m - 1.0, 0.0};
	dd const den = two_sum(m, 1.0);
	dd const f   = div(num, den);
	dd const f2  = mul(f, f);

	dd s = inv_odd(21);
	for (int i = 20; i >= 0; i--) s = add(mul(s, f2), inv_odd(i));
	dd lm = mul(f, s);
	lm    = dd{2.0 * lm.hi, 2.0 * lm.lo};

	// ln2 = 0x3FE62E42FEFA39EF + 0x3C7ABC9E3B39803F (double-double)
	dd const ln2 = {from_bits(0x3FE62E42FEFA39EFull), from_bits(0x3C7ABC9E3B39803Full)};
	dd const kl  = mul_d(ln2, double(k));
	return add(kl, lm);
}

// Natural log, correctly rounded for positive normal and subnormal inputs.
SPICE_HD double log_cr(double x) {
	if (!(x > 0)) return x == 0 ? -INFINITY : NAN;
	if (x == 1.0) return 0.0;
	if (!(x < INFINITY)) return x;
	dd const r = log_dd(x);
	return r.hi + r.lo;
}

// The exponential draw -scale * log(u) of the reference (random.h:270-272) for
// u in (0, 1], as the interval [lo, hi] of values a C library log with error
// at most `bound` ulp (glibc 2.35: 0.519, measured max 0.5183) can produce.
// lo == hi unless log(u) lies within (bound - 0.5) ulp of a rounding midpoint
// and the two candidate logs scale to different draws; returns lo != hi.
SPICE_HD bool exp_draw_interval(double u, double scale, double bound, double& lo, double& hi) {
	if (u == 1.0) {  // log(1) is exactly 0 in every C library
		lo = hi = -scale * 0.0;
		return false;
	}
	dd const L       = log_dd(u);  // negative
	double const c   = L.hi;       // correctly rounded log(u)
	double const e1  = -scale * c;
	lo = hi = e1;
	if (L.lo == 0.0) return false;
	// neighbour of c on the side of the exact value
	uint64_t const cb = to_bits(c);
	double const n    = from_bits(L.lo > 0 ? cb - 1 : cb + 1);  // c < 0: toward zero is bits - 1
	double const sp   = std::fabs(n - c);
	double const d    = std::fabs(L.lo);
	double const up   = std::fabs(from_bits(cb + 1) - c);  // spacing above |c|
	double const ref  = sp > up ? sp : up;
	if ((sp - d) > bound * ref) return false;
	double const e2 = -scale * n;
	lo              = e1 < e2 ? e1 : e2;
	hi              = e1 < e2 ? e2 : e1;
	return lo != hi;
}

}  // namespace spice_crlog
