// Host-side checks of the C++ user API (include/spice/), restating the
// reference's unit tests (test/detail/csr.cpp, test/util/numeric.cpp,
// test/util/stdint.cpp, test/util/random.cpp:117-157, test/concepts.cpp) plus
// seed-chain values the reference produced.  Runs on the CPU: nothing here
// touches the device.  Prints one line per check and exits non-zero on failure.
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "spice/concepts.h"
#include "spice/detail/csr.h"
#include "spice/topology.h"
#include "spice/util/numeric.h"
#include "spice/util/random.h"
#include "spice/util/range.h"
#include "spice/util/stdint.h"

using namespace spice;
using namespace spice::util;

static int g_fail = 0;
#define CHECK(cond)                                                        \
	do {                                                                   \
		if (!(cond)) {                                                     \
			std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
			g_fail++;                                                      \
		}                                                                  \
	} while (0)

static void csr_tests() {
	{  // CSR.Empty
		adj_list adj;
		adj(1, 0);
		spice::detail::csr c(adj, {1337});
		CHECK(c.neighbors(0).size() == 0);
	}
	{  // CSR.Custom
		adj_list adj;
		CHECK(adj.size() == 0);
		adj.connect(0, 0);
		adj.connect(0, 2);
		adj.connect(0, 7);
		adj.connect(0, 9);
		adj.connect(2, 1);
		adj.connect(2, 2);
		adj(3, 10);
		spice::detail::csr c(adj, {1337});
		CHECK(c.neighbors(0).size() == 4);
		CHECK(c.neighbors(1).size() == 0);
		CHECK(c.neighbors(2).size() == 2);
		std::vector<std::pair<Int32, void*>> r0(c.neighbors(0).begin(), c.neighbors(0).end());
		CHECK(r0.size() == 4 && r0[0].first == 0 && r0[1].first == 2 && r0[2].first == 7 && r0[3].first == 9);
		std::vector<std::pair<Int32, void*>> r2(c.neighbors(2).begin(), c.neighbors(2).end());
		CHECK(r2.size() == 2 && r2[0].first == 1 && r2[1].first == 2);
	}
	{  // CSR.Payload
		adj_list adj;
		adj.connect(0, 0);
		adj(1, 1);
		spice::detail::csr<int> c(adj, {1337});
		std::vector<std::pair<Int32, int*>> r(c.neighbors(0).begin(), c.neighbors(0).end());
		CHECK(r.size() == 1 && r[0].second != nullptr);
	}
	{  // edge_stream leaves rows after the last source untouched (topology.cpp:27-32)
		std::vector<Int> off(4, -1);
		std::vector<Int32> nb(2, 0);
		edge_stream es(off, nb);
		es << std::pair<Int32, Int32>{0, 5} << std::pair<Int32, Int32>{0, 6};
		es.flush();
		CHECK(off[0] == 0 && off[1] == 2 && off[2] == -1 && nb[0] == 5 && nb[1] == 6);
	}
	{  // precondition: connecting ids beyond Int32 throws
		adj_list adj;
		bool threw = false;
		try {
			adj.connect(Int(std::numeric_limits<Int32>::max()), 0);
		} catch (std::logic_error const& e) {
			threw = std::string(e.what()).rfind("Assertion failed (topology.h:", 0) == 0;
		}
		CHECK(threw);
	}
}

static void numeric_tests() {
	// test/util/numeric.cpp:9-29
	Int const ITER    = 1000;
	float const DELTA = 0.001f;
	float sum         = 0;
	kahan_sum<float> ksum;
	float deltas = 0;
	for (Int i : range(ITER)) {
		CHECK(std::abs(float(ksum) - i * DELTA) <= 4 * std::numeric_limits<float>::epsilon() * (i * DELTA + 1));
		sum += DELTA;
		float const d = ksum += DELTA;
		deltas += d;
		CHECK(std::abs(d - DELTA) <= DELTA / 100);
	}
	CHECK(sum != 1);
	CHECK(float(ksum) == 1);
	CHECK(deltas == 1);
}

static void stdint_tests() {
	// test/util/stdint.cpp:5-21
	UInt128 a{7, 2};
	auto b = a + 123;
	CHECK(b.lo == 130 && b.hi == 2);
	UInt128 c{~0_u64, 2};
	auto d = c + 1;
	CHECK(d.lo == 0 && d.hi == 3);
}

template <std::integral I>
struct constant_rng {
	I x;
	constexpr explicit constant_rng(I v) : x(v) {}
	constexpr I operator()() { return x; }
};

template <class Real>
static void interval(auto rng) {
	uniform_real_distribution<Real, false> right;
	uniform_real_distribution<Real, true> left;
	Real const x = right(rng);
	CHECK(0 <= x && x < 1);
	Real const y = left(rng);
	CHECK(0 < y && y <= 1);
}

static void random_tests() {
	// test/util/random.cpp:117-157
	UInt32 const MAX = std::numeric_limits<UInt32>::max();
	interval<float>(constant_rng<UInt32>(0));
	interval<double>(constant_rng<UInt32>(0));
	interval<float>(constant_rng<UInt32>(MAX));
	interval<double>(constant_rng<UInt32>(MAX));
	interval<float>(constant_rng<UInt>(0));
	interval<double>(constant_rng<UInt>(0));
	interval<float>(constant_rng<UInt>(MAX));
	interval<double>(constant_rng<UInt>(MAX));
	interval<float>(constant_rng<UInt>(~UInt(0)));
	interval<double>(constant_rng<UInt>(~UInt(0)));

	// seed chain of {1337} as the reference produced it
	seed_seq s{1337};
	CHECK(s.seed().lo == 0x6748b9668c35b07bull && s.seed().hi == 0x3c011c497f4b0865ull);
	for (int i = 0; i < 8; i++) s++;
	CHECK(s.seed().lo == 0xd64e4ef75fcd078full && s.seed().hi == 0xb3d44233ff13c9b5ull);
	xoroshiro64_128p rng(s);
	CHECK(rng() == 0x8a22912b5ee0d144ull);
	CHECK(rng() == 0xb9db286a94741908ull);
	CHECK(rng() == 0x8ccbe2b2f921e3abull);
	// an empty seed list is a precondition failure (random.h:151)
	bool threw = false;
	try {
		seed_seq e{std::initializer_list<UInt32>{}};
		(void)e;
	} catch (std::logic_error const&) {
		threw = true;
	}
	CHECK(threw);
	// generate() repeats the four 32-bit words of the seed
	seed_seq t{1};
	Int32 w[6];
	t.generate(w, w + 6);
	CHECK(UInt32(w[0]) == UInt32(t.seed().lo) && UInt32(w[4]) == UInt32(t.seed().lo) &&
	      UInt32(w[3]) == UInt32(t.seed().hi >> 32));
}

// --- concepts (test/concepts.cpp restated for representative signatures)
struct n_stateless {
	bool update(float, auto&) const { return false; }
};
struct n_stateful {
	struct neuron {
		float v;
	};
	bool update(neuron&, float, auto&) const { return false; }
};
struct n_init : n_stateful {
	void init(neuron&, Int, auto&) const {}
};
struct n_popinit : n_stateful {
	void init(std::span<neuron>, auto&) {}
};
struct n_both_init : n_stateful {
	void init(neuron&, Int, auto&) const {}
	void init(std::span<neuron>, auto&) {}
};
struct n_popupdate {
	void update(float, auto&, std::vector<Int32>&) {}
};
struct n_popupdate_stateful {
	struct neuron {};
	void update(float, auto&, std::vector<Int32>&) {}
};
struct n_bad_return {
	void update(float, auto&) const {}
};
struct s_plain {
	void deliver(n_stateful::neuron&) const {}
};
struct s_stateful {
	struct synapse {
		float w;
	};
	void deliver(synapse const&, n_stateful::neuron&) const {}
};
struct s_from_to {
	void deliver(n_stateful::neuron const&, n_stateful::neuron&) const {}
};
struct s_plastic {
	struct synapse {
		float w;
	};
	void deliver(synapse const&, n_stateful::neuron&) const {}
	void update(synapse&, float, bool, bool) const {}
	void skip(synapse&, float, Int) const {}
};
struct s_update_no_skip {
	struct synapse {};
	void deliver(synapse const&, n_stateful::neuron&) const {}
	void update(synapse&, float, bool, bool) const {}
};

static_assert(Neuron<n_stateless> && StatelessNeuron<n_stateless> && !StatefulNeuron<n_stateless>);
static_assert(Neuron<n_stateful> && StatefulNeuron<n_stateful> && PerNeuronUpdate<n_stateful>);
static_assert(Neuron<n_init> && PerNeuronInit<n_init> && !PerPopulationInit<n_init>);
static_assert(Neuron<n_popinit> && PerPopulationInit<n_popinit>);
static_assert(!Neuron<n_both_init>);
static_assert(Neuron<n_popupdate> && PerPopulationUpdate<n_popupdate>);
static_assert(!Neuron<n_popupdate_stateful>);
static_assert(!Neuron<n_bad_return>);
static_assert(Synapse<s_plain, n_stateless, n_stateful> && DeliverTo<s_plain, n_stateful>);
static_assert(Synapse<s_stateful, n_stateless, n_stateful> && StatefulSynapse<s_stateful>);
static_assert(Synapse<s_from_to, n_stateful, n_stateful> && DeliverFromTo<s_from_to, n_stateful, n_stateful>);
static_assert(!Synapse<s_from_to, n_stateless, n_stateful>);  // source has no state
static_assert(PlasticSynapse<s_plastic> && !PlasticSynapse<s_stateful>);
static_assert(!PlasticSynapse<s_update_no_skip>);
static_assert(CheckNeuron<n_stateless>() && CheckNeuron<n_init>() && CheckSynapse<s_plastic>());

int main() {
	csr_tests();
	numeric_tests();
	stdint_tests();
	random_tests();
	if (g_fail)
		std::printf("api host tests: %d failures\n", g_fail);
	else
		std::printf("api host tests: all passed\n");
	return g_fail ? 1 : 0;
}
