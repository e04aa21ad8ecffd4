// Seeds, generators and distributions (API of reference util/random.h).
//
// seed_seq hashes a list of 32-bit seeds into 128 bits with a MurmurHash3
// x64-128 variant (non-zero initial state, reference random.h:32-106) and
// derives further seeds by hashing its own value (random.h:108-139,163-167).
// The generators and distributions are usable in host and device code; the
// neuron kernels hand user models a stream object with the same call syntax.
#pragma once

#include <algorithm>
#include <cmath>
#include <concepts>
#include <cstring>
#include <initializer_list>
#include <iterator>
#include <limits>
#include <numbers>
#include <random>
#include <type_traits>

#include "spice/util/assert.h"
#include "spice/util/stdint.h"

#if defined(__HIPCC__)
#define SPICE_HOST_DEVICE __host__ __device__
#else
#define SPICE_HOST_DEVICE
#endif

namespace spice::util {
namespace detail {
template <std::integral Integer>
SPICE_HOST_DEVICE constexpr Integer rotl(Integer const x, Int const k) {
	return Integer(x << k) | Integer(x >> (sizeof(Integer) * 8 - k));
}

SPICE_HOST_DEVICE constexpr UInt fmix64(UInt k) {
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return k;
}

inline constexpr UInt kMurmurC1 = 0x87c37b91114253d5ull;
inline constexpr UInt kMurmurC2 = 0x4cf5ad432745937full;
// The reference's fixed non-zero initial hash state.
inline constexpr UInt128 kMurmurInit{0x2E4016967F18E81ull, 0x447567949F9AA86ull};

SPICE_HOST_DEVICE constexpr void murmur_mix_lo(UInt128& h, UInt k1) {
	k1 *= kMurmurC1;
	k1 = rotl(k1, 31);
	k1 *= kMurmurC2;
	h.lo ^= k1;
}
SPICE_HOST_DEVICE constexpr void murmur_mix_hi(UInt128& h, UInt k2) {
	k2 *= kMurmurC2;
	k2 = rotl(k2, 33);
	k2 *= kMurmurC1;
	h.hi ^= k2;
}
SPICE_HOST_DEVICE constexpr void murmur_block(UInt128& h, UInt k1, UInt k2) {
	murmur_mix_lo(h, k1);
	h.lo = rotl(h.lo, 27);
	h.lo += h.hi;
	h.lo = h.lo * 5 + 0x52dce729;
	murmur_mix_hi(h, k2);
	h.hi = rotl(h.hi, 31);
	h.hi += h.lo;
	h.hi = h.hi * 5 + 0x38495ab5;
}
SPICE_HOST_DEVICE constexpr UInt128 murmur_final(UInt128 h, UInt len) {
	h.lo ^= len;
	h.hi ^= len;
	h.lo += h.hi;
	h.hi += h.lo;
	h.lo = fmix64(h.lo);
	h.hi = fmix64(h.hi);
	h.lo += h.hi;
	h.hi += h.lo;
	return h;
}

// Hash of `len` bytes.
inline UInt128 murmur3(void const* ptr, UInt const len) {
	auto const* data = static_cast<unsigned char const*>(ptr);
	UInt128 h        = kMurmurInit;
	UInt const nblocks = len / 16;
	for (UInt i = 0; i < nblocks; i++) {
		UInt k1, k2;
		std::memcpy(&k1, data + 16 * i, 8);
		std::memcpy(&k2, data + 16 * i + 8, 8);
		murmur_block(h, k1, k2);
	}
	unsigned char const* tail = data + 16 * nblocks;
	UInt const rem            = len & 15;
	if (rem > 8) {  // tail bytes 8..14 form the high lane
		UInt k2 = 0;
		for (UInt b = 8; b < rem; b++) k2 |= UInt(tail[b]) << (8 * (b - 8));
		murmur_mix_hi(h, k2);
	}
	if (rem > 0) {  // tail bytes 0..7 form the low lane
		UInt k1 = 0;
		for (UInt b = 0; b < std::min<UInt>(rem, 8); b++) k1 |= UInt(tail[b]) << (8 * b);
		murmur_mix_lo(h, k1);
	}
	return murmur_final(h, len);
}

// Hash of one 16-byte block (used to advance a seed).
SPICE_HOST_DEVICE constexpr UInt128 murmur3(UInt128 k) {
	UInt128 h = kMurmurInit;
	murmur_block(h, k.lo, k.hi);
	return murmur_final(h, 16);
}
}  // namespace detail

// Copyable, fixed-size seed sequence.
class seed_seq {
public:
	seed_seq(std::seed_seq seq) {
		UInt32 words[4];
		seq.generate(words, words + 4);
		std::memcpy(&_seed, words, sizeof(_seed));
	}
	seed_seq(std::initializer_list<UInt32> il) : _seed(detail::murmur3(il.begin(), 4 * il.size())) {
		SPICE_PRE(il.size() > 0 && "Please provide at least 1 seed to seed_seq");
	}

	template <std::output_iterator<Int32> OutputIt>
	constexpr void generate(OutputIt first, OutputIt last) const {
		UInt32 words[4] = {UInt32(_seed.lo), UInt32(_seed.lo >> 32), UInt32(_seed.hi),
		                   UInt32(_seed.hi >> 32)};
		for (Int i = 0; first != last; i++, first++) *first = Int32(words[i % 4]);
	}

	constexpr UInt128 seed() const { return _seed; }

	// Returns the current seed and advances to the next one.
	constexpr seed_seq operator++(int) {
		seed_seq const old = *this;
		_seed              = detail::murmur3(_seed);
		return old;
	}

	constexpr seed_seq stream(UInt id) const { return seed_seq(detail::murmur3(_seed + (id + 1))); }

	static constexpr seed_seq from_value(UInt128 v) { return seed_seq(v); }

private:
	UInt128 _seed;

	constexpr explicit seed_seq(UInt128 const s) : _seed(s) {}
};

template <std::integral Integer, int Period>
class xoroshiro {
public:
	using result_type = Integer;

	constexpr explicit xoroshiro(seed_seq const& seq) : _state(seq.seed()) {}

	SPICE_HOST_DEVICE constexpr Integer min() { return 0; }
	SPICE_HOST_DEVICE constexpr Integer max() { return std::numeric_limits<Integer>::max(); }

protected:
	struct state2 {
		SPICE_HOST_DEVICE constexpr state2(UInt128 const s) : s0(s.lo), s1(s.hi) {}
		Integer s0;
		Integer s1;
	};
	struct state4 {
		SPICE_HOST_DEVICE constexpr state4(UInt128 const s) :
		s0(Integer(s.lo)), s1(Integer(s.lo >> 32)), s2(Integer(s.hi)), s3(Integer(s.hi >> 32)) {}
		Integer s0;
		Integer s1;
		Integer s2;
		Integer s3;
	};
	std::conditional_t<Period / (sizeof(Integer) * 8) == 2, state2, state4> _state;
};

// xoshiro128+ on four 32-bit words (used by the reference's tests only).
struct xoroshiro32_128p : public xoroshiro<UInt32, 128> {
	using xoroshiro::xoroshiro;

	SPICE_HOST_DEVICE UInt32 operator()() {
		UInt32 const out = _state.s0 + _state.s3;
		UInt32 const t   = _state.s1 << 9;
		_state.s2 ^= _state.s0;
		_state.s3 ^= _state.s1;
		_state.s1 ^= _state.s2;
		_state.s0 ^= _state.s3;
		_state.s2 ^= t;
		_state.s3 = detail::rotl(_state.s3, 11);
		return out;
	}
};

// xoroshiro128+ with rotations 24/16/37 - the engine's generator.
struct xoroshiro64_128p : public xoroshiro<UInt, 128> {
	using xoroshiro::xoroshiro;

	SPICE_HOST_DEVICE UInt operator()() {
		UInt const out = _state.s0 + _state.s1;
		UInt const t   = _state.s0 ^ _state.s1;
		_state.s0      = detail::rotl(_state.s0, 24) ^ t ^ (t << 16);
		_state.s1      = detail::rotl(t, 37);
		return out;
	}
};

// Uniform real in [0, 1) - or (0, 1] when LeftOpen - from the top `digits` bits
// of one (or, for a 32-bit generator and double, two) generator outputs.
template <std::floating_point Real, bool LeftOpen = false>
SPICE_HOST_DEVICE constexpr Real generate_canonical(auto& rng) {
	constexpr std::size_t rng_size = sizeof(decltype(rng()));
	constexpr std::size_t digits   = std::numeric_limits<Real>::digits;
	constexpr std::size_t width    = 8 * (rng_size > sizeof(Real) ? rng_size : sizeof(Real));

	UInt bits = rng();
	if constexpr (rng_size < sizeof(Real)) bits = (bits << (8 * rng_size)) | rng();
	UInt const v = (bits >> (width - digits)) + (LeftOpen ? 1u : 0u);
	// v <= 2^digits: below 2^32 it converts from 32 bits (exactly, as from 64)
	if constexpr (digits < 32)
		return Real(uint32_t(v)) / Real(1_u64 << digits);
	else
		return Real(v) / Real(1_u64 << digits);
}

template <std::floating_point Real, bool LeftOpen = false>
class uniform_real_distribution {
public:
	SPICE_HOST_DEVICE constexpr explicit uniform_real_distribution(Real const a = 0, Real const b = 1) :
	_offset(a), _scale(b - a) {}

	SPICE_HOST_DEVICE constexpr Real operator()(auto& rng) const {
		return std::fma(generate_canonical<Real, LeftOpen>(rng), _scale, _offset);
	}

private:
	Real _offset = 0;
	Real _scale  = 1;
};

template <std::floating_point Real>
class exponential_distribution {
public:
	constexpr explicit exponential_distribution(Real const scale = 1) : _scale(scale) {
		SPICE_PRE(scale >= 0);
	}
	SPICE_HOST_DEVICE constexpr Real operator()(auto& rng) const {
		return -_scale * std::log(generate_canonical<Real, true>(rng));
	}

private:
	Real _scale;
};

// Box-Muller, two values per pair of draws.
template <std::floating_point Real>
class normal_distribution {
public:
	constexpr explicit normal_distribution(Real const mu = 0, Real const sigma = 1) :
	_mu(mu), _sigma(sigma) {
		SPICE_PRE(sigma >= 0);
	}

	SPICE_HOST_DEVICE constexpr Real operator()(auto& rng) {
		if (_have) {
			_have = false;
			return _z1;
		}
		Real const R     = std::sqrt(-2 * std::log(generate_canonical<Real, true>(rng)));
		Real const theta = 2 * std::numbers::pi_v<Real> * generate_canonical<Real, false>(rng);
		Real const z0    = std::fma(R * std::cos(theta), _sigma, _mu);
		_z1              = std::fma(R * std::sin(theta), _sigma, _mu);
		_have            = true;
		return z0;
	}

private:
	bool _have = false;
	Real _z1   = 0;
	Real _mu;
	Real _sigma;
};

// Normal approximation of the binomial, clamped to [0, N].
template <std::integral Integer>
class binomial_distribution {
public:
	using Real = std::conditional_t<sizeof(Integer) == 8, double, float>;

	constexpr explicit binomial_distribution(Integer const N, Real const p) :
	_N(N), _norm(N * p, std::sqrt(N * p * (1 - p))) {
		SPICE_PRE(N >= 0);
		SPICE_PRE(0 <= p && p <= 1);
	}

	SPICE_HOST_DEVICE constexpr Integer operator()(auto& rng) {
		Real const x = std::round(_norm(rng));
		return std::min(_N, static_cast<Integer>(x > Real(0) ? x : Real(0)));
	}

private:
	Integer _N;
	normal_distribution<Real> _norm;
};
}  // namespace spice::util
