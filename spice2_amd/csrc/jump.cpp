// Host-side GF(2) polynomial arithmetic for xoroshiro128+ jump-ahead (see rng.h).
#include <mutex>

#include "rng.h"

namespace spice_rt {
namespace {

bool get_bit(gf2poly const& p, int i) { return ((i < 64 ? p.lo >> i : p.hi >> (i - 64)) & 1u) != 0; }

// Berlekamp-Massey over GF(2) on bit 0 of s0 along the orbit of a fixed state.
// The linear complexity of that sequence is 128 for a full-period generator,
// so the connection polynomial is the characteristic polynomial of T.
gf2poly compute_charpoly() {
	constexpr int NBITS = 512;
	unsigned char seq[NBITS];
	xoro_state s{0x0123456789abcdefull, 0xfedcba9876543210ull};
	for (int i = 0; i < NBITS; i++) {
		seq[i] = (unsigned char)(s.s0 & 1u);
		xoro_next(s);
	}
	// C(x) = 1 + c1 x + ... ; B(x) previous.  Arrays of NBITS+1 coefficients.
	static unsigned char C[NBITS + 1], B[NBITS + 1], T[NBITS + 1];
	for (int i = 0; i <= NBITS; i++) C[i] = B[i] = 0;
	C[0] = B[0] = 1;
	int L = 0, m = 1;
	for (int n = 0; n < NBITS; n++) {
		unsigned char d = seq[n];
		for (int i = 1; i <= L; i++) d ^= (unsigned char)(C[i] & seq[n - i]);
		if (d == 0) {
			m++;
		} else if (2 * L <= n) {
			for (int i = 0; i <= NBITS; i++) T[i] = C[i];
			for (int i = 0; i + m <= NBITS; i++) C[i + m] ^= B[i];
			L = n + 1 - L;
			for (int i = 0; i <= NBITS; i++) B[i] = T[i];
			m = 1;
		} else {
			for (int i = 0; i + m <= NBITS; i++) C[i + m] ^= B[i];
			m++;
		}
	}
	// Connection polynomial C(x) = sum c_i x^i (c_0 = 1) of degree L = 128; the
	// characteristic polynomial is its reciprocal x^L C(1/x) = sum c_i x^(L-i).
	gf2poly p{0, 0};
	for (int i = 0; i <= L; i++) {
		if (!C[i]) continue;
		int const e = L - i;
		if (e == 128) continue;  // leading term kept implicit
		if (e < 64)
			p.lo |= 1ull << e;
		else
			p.hi |= 1ull << (e - 64);
	}
	return p;
}

// a * b mod P, P monic of degree 128 given without its leading term.
gf2poly mulmod(gf2poly a, gf2poly b, gf2poly P) {
	gf2poly r{0, 0};
	for (int i = 127; i >= 0; i--) {
		// r = r * x mod P
		bool const top = (r.hi >> 63) & 1u;
		r.hi           = (r.hi << 1) | (r.lo >> 63);
		r.lo <<= 1;
		if (top) {
			r.lo ^= P.lo;
			r.hi ^= P.hi;
		}
		if (get_bit(a, i)) {
			r.lo ^= b.lo;
			r.hi ^= b.hi;
		}
	}
	return r;
}

}  // namespace

gf2poly xoro_charpoly() {
	static gf2poly P;
	static std::once_flag once;
	std::call_once(once, [] { P = compute_charpoly(); });
	return P;
}

gf2poly xoro_jump_poly(uint64_t n) {
	gf2poly const P = xoro_charpoly();
	gf2poly result{1, 0};  // x^0
	gf2poly base{2, 0};    // x^1
	while (n) {
		if (n & 1u) result = mulmod(result, base, P);
		base = mulmod(base, base, P);
		n >>= 1;
	}
	return result;
}

gf2poly gf2_mulmod(gf2poly a, gf2poly b) { return mulmod(a, b, xoro_charpoly()); }

xoro_state xoro_jump_n(xoro_state s, uint64_t n) { return xoro_jump(s, xoro_jump_poly(n)); }

}  // namespace spice_rt
