// Test-only C ABI over the *reference* Spice library (denniskb/spice2), built from
// the reference's own sources under /root/reference by oracle/Makefile into
// oracle/_ref/.  It exists to produce golden vectors (tests/golden/make_golden.py)
// and to cross-check the C restatement in oracle/spice_oracle.c.  Nothing in the
// product links or loads it.
//
// The model structs below are the workload definitions of the reference samples
// (samples/brunel.cpp:23-70, samples/vogels.cpp:10-59), restated with the same
// floating-point expressions so that the reference produces the rasters the
// product must reproduce.  Rates and weights are parameters so that the same
// driver covers the scaled and swept Brunel configurations.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "spice/snn.h"

using namespace spice;
using namespace spice::util;

namespace {
thread_local std::string g_err;

// seed_seq can only be built from a braced list; a runtime list is hashed with
// the same murmur3 byte hash that constructor applies (random.h:149-152) and
// stored into the object, whose only member is that 128-bit value.
seed_seq make_seed(uint32_t const* seeds, int nseeds, int advance) {
	if (nseeds <= 0) throw std::logic_error("empty seed list");
	seed_seq s({0u});
	UInt128 h = spice::util::detail::murmur3(seeds, 4 * UInt(nseeds));
	static_assert(sizeof(seed_seq) == sizeof(UInt128));
	std::memcpy(static_cast<void*>(&s), &h, sizeof(h));
	for (int i = 0; i < advance; i++) s++;
	return s;
}

uint64_t fnv1a(uint64_t h, void const* p, size_t n) {
	auto const* b = static_cast<unsigned char const*>(p);
	for (size_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 0x100000001b3ull;
	}
	return h;
}
constexpr uint64_t FNV_BASIS = 0xcbf29ce484222325ull;

struct poisson_r {
	float rate;
	bool update(float dt, auto& rng) const { return util::generate_canonical<float>(rng) < (rate * dt); }
};

struct lif {
	struct neuron {
		float V   = 0;
		int Twait = 0;
	};
	bool update(neuron& n, float dt, auto) const {
		float const TmemInv = 1.0 / 0.02;
		float const Vrest   = 0.0;
		int const Tref      = 20;
		float const Vthres  = 0.02;
		if (--n.Twait <= 0) {
			if (n.V > Vthres) {
				n.V     = Vrest;
				n.Twait = Tref;
				return true;
			}
			n.V += (Vrest - n.V) * (dt * TmemInv);
		}
		return false;
	}
};

struct fixed_weight {
	float weight;
	void deliver(lif::neuron& to) const { to.V += weight; }
};

struct coba {
	int init_mode = 0;  // 0: default state, 1: V ~ U[-0.06, -0.05)
	struct neuron {
		float V     = -0.06;
		float Gex   = 0;
		float Gin   = 0;
		Int32 Twait = 0;
	};
	void init(neuron& n, Int, auto& rng) const {
		if (init_mode == 1) n.V = uniform_real_distribution<float>(-0.06f, -0.05f)(rng);
	}
	bool update(neuron& n, float const dt, auto&) const {
		Int32 const Tref    = 50;
		float const Vrest   = -0.06;
		float const Vthres  = -0.05;
		float const TmemInv = 1.0f / 0.02;
		float const Eex     = 0.0;
		float const Ein     = -0.08;
		float const Ibg     = 0.02;
		float const TexInv  = 1.0f / 0.005;
		float const TinInv  = 1.0f / 0.01;
		bool spiked         = false;
		if (--n.Twait <= 0) {
			if (n.V > Vthres) {
				n.V     = Vrest;
				n.Twait = Tref;
				spiked  = true;
			} else
				n.V += ((Vrest - n.V) + n.Gex * (Eex - n.V) + n.Gin * (Ein - n.V) + Ibg) *
				       (dt * TmemInv);
		}
		n.Gex -= n.Gex * (dt * TexInv);
		n.Gin -= n.Gin * (dt * TinInv);
		return spiked;
	}
};
struct excitatory {
	float weight;
	void deliver(coba::neuron& to) const { to.Gex += weight; }
};
// samples/brunel+.cpp:59-99
struct plastic {
	struct synapse {
		float W     = 1e-4;
		float Zpre  = 0;
		float Zpost = 0;
	};
	void deliver(synapse const& syn, lif::neuron& to) const { to.V += syn.W; }
	void update(synapse& syn, float const dt, bool const pre, bool const post) const {
		float const TstdpInv = 1.0f / 0.02f;
		float const dtInv    = 1.0f / dt;
		syn.W = std::clamp(syn.W - pre * 0.0202f * syn.W * std::exp(-syn.Zpost * dtInv) +
		                       post * 0.01f * (1.0f - syn.W) * std::exp(-syn.Zpre * dtInv),
		                   0.0f, 0.0003f);
		syn.Zpre += pre;
		syn.Zpost += post;
		syn.Zpre -= syn.Zpre * dt * TstdpInv;
		syn.Zpost -= syn.Zpost * dt * TstdpInv;
	}
	void skip(synapse& syn, float const dt, Int const n) const {
		float const TstdpInv = 1.0f / 0.02f;
		syn.Zpre *= std::pow(1 - dt * TstdpInv, n);
		syn.Zpost *= std::pow(1 - dt * TstdpInv, n);
	}
};

// Pair STDP on exponential traces (tau 20 ms, A+ 0.01, A- 0.0105 relative to w0,
// W clipped to [0, 2 w0]) in the same update()/skip() form.
struct stdp_pair {
	float w0;
	struct synapse {
		float W     = 0;
		float Zpre  = 0;
		float Zpost = 0;
	};
	void init(synapse& syn, Int, Int, auto&) const { syn.W = w0; }
	void deliver(synapse const& syn, lif::neuron& to) const { to.V += syn.W; }
	void update(synapse& syn, float const dt, bool const pre, bool const post) const {
		float const tauInv = 1.0f / 0.02f;
		float const wmax   = 2.0f * w0;
		if (pre) syn.W = std::clamp(syn.W - 0.0105f * w0 * syn.Zpost, 0.0f, wmax);
		if (post) syn.W = std::clamp(syn.W + 0.01f * w0 * syn.Zpre, 0.0f, wmax);
		syn.Zpre += pre;
		syn.Zpost += post;
		syn.Zpre -= syn.Zpre * dt * tauInv;
		syn.Zpost -= syn.Zpost * dt * tauInv;
	}
	void skip(synapse& syn, float const dt, Int const n) const {
		float const tauInv = 1.0f / 0.02f;
		syn.Zpre *= std::pow(1 - dt * tauInv, n);
		syn.Zpost *= std::pow(1 - dt * tauInv, n);
	}
};

struct inhibitory {
	float weight;
	void deliver(coba::neuron& to) const { to.Gin += weight; }
};

template <class Pop>
void record(Pop* pop, int64_t step, int slot, int32_t* counts, int32_t* raster, int64_t cap,
            int64_t& used, uint64_t& rhash) {
	auto s                  = pop->spikes(0);
	counts[step * 3 + slot] = static_cast<int32_t>(s.size());
	rhash                   = fnv1a(rhash, s.data(), sizeof(Int32) * s.size());
	for (auto id : s) {
		if (raster && used < cap) raster[used] = id;
		used++;
	}
}
}  // namespace

extern "C" {

char const* ref_last_error() { return g_err.c_str(); }

int ref_murmur3(void const* data, uint64_t len, uint64_t* out) {
	UInt128 h = spice::util::detail::murmur3(data, len);
	out[0]    = h.lo;
	out[1]    = h.hi;
	return 0;
}

int ref_seed_chain(uint32_t const* seeds, int nseeds, int count, uint64_t* out) {
	try {
		seed_seq s = make_seed(seeds, nseeds, 0);
		for (int k = 0; k < count; k++) {
			UInt128 v      = s.seed();
			out[2 * k]     = v.lo;
			out[2 * k + 1] = v.hi;
			s++;
		}
		return 0;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

int ref_xoroshiro(uint32_t const* seeds, int nseeds, int advance, int64_t n, uint64_t* out) {
	xoroshiro64_128p rng(make_seed(seeds, nseeds, advance));
	for (int64_t i = 0; i < n; i++) out[i] = rng();
	return 0;
}

int ref_canonical_f32(uint32_t const* seeds, int nseeds, int advance, int64_t n, float* out) {
	xoroshiro64_128p rng(make_seed(seeds, nseeds, advance));
	for (int64_t i = 0; i < n; i++) out[i] = generate_canonical<float>(rng);
	return 0;
}

int ref_canonical_f64_leftopen(uint32_t const* seeds, int nseeds, int advance, int64_t n,
                               double* out) {
	xoroshiro64_128p rng(make_seed(seeds, nseeds, advance));
	for (int64_t i = 0; i < n; i++) out[i] = generate_canonical<double, true>(rng);
	return 0;
}

int ref_exponential(uint32_t const* seeds, int nseeds, int advance, double scale, int64_t n,
                    double* out) {
	xoroshiro64_128p rng(make_seed(seeds, nseeds, advance));
	exponential_distribution<double> d(scale);
	for (int64_t i = 0; i < n; i++) out[i] = d(rng);
	return 0;
}

int ref_uniform_f32(uint32_t const* seeds, int nseeds, int advance, float a, float b, int64_t n,
                    float* out) {
	xoroshiro64_128p rng(make_seed(seeds, nseeds, advance));
	uniform_real_distribution<float> d(a, b);
	for (int64_t i = 0; i < n; i++) out[i] = d(rng);
	return 0;
}

int ref_kahan(float dt, int64_t n, float* out_dt, float* out_sum) {
	kahan_sum<float> k;
	for (int64_t i = 0; i < n; i++) {
		float const d = k += dt;
		if (k >= 1) k.reset();
		out_dt[i]  = d;
		out_sum[i] = k;
	}
	return 0;
}

int64_t ref_fp_size(int64_t src, int64_t dst, double p) {
	try {
		fixed_probability fp(p);
		fp(src, dst);
		return fp.size();
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

// Generates the CSR exactly as detail::csr does (csr.h:69-77): offsets sized
// src+1 (0 if src == 0), neighbors sized Topology::size(), zero-filled.
// Returns the synapse count or -1 with ref_last_error() set.
int64_t ref_fp_generate(int64_t src, int64_t dst, double p, uint32_t const* seeds, int nseeds,
                        int advance, int64_t* offsets, int32_t* neighbors, int64_t cap) {
	try {
		fixed_probability fp(p);
		fp(src, dst);
		Int const size = fp.size();
		if (size > cap) throw std::logic_error("neighbor buffer too small");
		std::vector<Int> off(src > 0 ? src + 1 : 0, 0);
		std::vector<Int32> nb(size, 0);
		seed_seq s = make_seed(seeds, nseeds, advance);
		static_cast<Topology&>(fp).generate(off, nb, s);
		for (size_t i = 0; i < off.size(); i++) offsets[i] = off[i];
		Int const n = off.empty() ? 0 : off.back();
		std::memcpy(neighbors, nb.data(), sizeof(Int32) * n);
		return n;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

// FNV-1a-64 over offsets[0..src] (int64 LE) then neighbors[0..offsets[src]) (int32 LE),
// of the CSR the reference generates, computed without keeping a copy.
int ref_fp_hash(int64_t src, int64_t dst, double p, uint32_t const* seeds, int nseeds, int advance,
                uint64_t* out_hash, int64_t* out_count, int32_t* out_row0, int row0_len) {
	try {
		fixed_probability fp(p);
		fp(src, dst);
		std::vector<Int> off(src > 0 ? src + 1 : 0, 0);
		std::vector<Int32> nb(fp.size(), 0);
		seed_seq s = make_seed(seeds, nseeds, advance);
		static_cast<Topology&>(fp).generate(off, nb, s);
		uint64_t h = FNV_BASIS;
		h          = fnv1a(h, off.data(), sizeof(Int) * off.size());
		Int const n = off.empty() ? 0 : off.back();
		h          = fnv1a(h, nb.data(), sizeof(Int32) * n);
		*out_hash  = h;
		*out_count = n;
		for (int i = 0; i < row0_len; i++)
			out_row0[i] = (!off.empty() && i < off[1]) ? nb[i] : -1;
		return 0;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

// Brunel network as in samples/brunel.cpp:75-103 with parameterised sizes,
// probability, rate and weights.  Per step: spike counts of P, E, I
// (counts[step*3 + {0,1,2}]), the spike ids (population-local, P then E then I)
// appended to raster, an FNV-1a hash of that step's ids (P, E, I lists in turn)
// and an FNV-1a hash of the E then I neuron arrays.
int ref_brunel_run(int64_t nP, int64_t nE, int64_t nI, double p, float rate, float wex, float win,
                   float dt, float delay, uint32_t const* seeds, int nseeds, int64_t steps,
                   int32_t* counts, int32_t* raster, int64_t raster_cap, int64_t* raster_used,
                   uint64_t* raster_hash, uint64_t* state_hash, float* final_V) {
	try {
		snn net(dt, delay, make_seed(seeds, nseeds, 0));
		auto P = net.add_population<poisson_r>(nP, {rate});
		auto E = net.add_population<lif>(nE);
		auto I = net.add_population<lif>(nI);
		net.connect<fixed_weight>(P, E, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(P, I, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(E, E, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(E, I, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(I, E, fixed_probability(p), delay, {win});
		net.connect<fixed_weight>(I, I, fixed_probability(p), delay, {win});
		int64_t used = 0;
		for (int64_t t = 0; t < steps; t++) {
			net.step();
			uint64_t rh = FNV_BASIS;
			record(P, t, 0, counts, raster, raster_cap, used, rh);
			record(E, t, 1, counts, raster, raster_cap, used, rh);
			record(I, t, 2, counts, raster, raster_cap, used, rh);
			if (raster_hash) raster_hash[t] = rh;
			if (state_hash) {
				uint64_t h = FNV_BASIS;
				auto en    = E->get_neurons();
				auto in    = I->get_neurons();
				h          = fnv1a(h, en.data(), sizeof(lif::neuron) * en.size());
				h          = fnv1a(h, in.data(), sizeof(lif::neuron) * in.size());
				state_hash[t] = h;
			}
		}
		if (final_V) {
			int64_t k = 0;
			for (auto& n : E->get_neurons()) final_V[k++] = n.V;
			for (auto& n : I->get_neurons()) final_V[k++] = n.V;
		}
		*raster_used = used;
		return 0;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

// Brunel+ (samples/brunel+.cpp:101-117): Brunel whose E->E projection is
// plastic (rule 1: the sample's `plastic`, rule 2: stdp_pair with w0).  Same
// outputs as ref_brunel_run.
int ref_brunel_plastic_run(int64_t nP, int64_t nE, int64_t nI, double p, float rate, float wex, float win,
                           float dt, float delay, uint32_t const* seeds, int nseeds, int rule, float w0,
                           int64_t steps, int32_t* counts, int32_t* raster, int64_t raster_cap,
                           int64_t* raster_used, uint64_t* raster_hash, uint64_t* state_hash) {
	try {
		snn net(dt, delay, make_seed(seeds, nseeds, 0));
		auto P = net.add_population<poisson_r>(nP, {rate});
		auto E = net.add_population<lif>(nE);
		auto I = net.add_population<lif>(nI);
		net.connect<fixed_weight>(P, E, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(P, I, fixed_probability(p), delay, {wex});
		if (rule == 1)
			net.connect<plastic>(E, E, fixed_probability(p), delay);
		else
			net.connect<stdp_pair>(E, E, fixed_probability(p), delay, {w0});
		net.connect<fixed_weight>(E, I, fixed_probability(p), delay, {wex});
		net.connect<fixed_weight>(I, E, fixed_probability(p), delay, {win});
		net.connect<fixed_weight>(I, I, fixed_probability(p), delay, {win});
		int64_t used = 0;
		for (int64_t t = 0; t < steps; t++) {
			net.step();
			uint64_t rh = FNV_BASIS;
			record(P, t, 0, counts, raster, raster_cap, used, rh);
			record(E, t, 1, counts, raster, raster_cap, used, rh);
			record(I, t, 2, counts, raster, raster_cap, used, rh);
			if (raster_hash) raster_hash[t] = rh;
			if (state_hash) {
				uint64_t h = FNV_BASIS;
				auto en    = E->get_neurons();
				auto in    = I->get_neurons();
				h          = fnv1a(h, en.data(), sizeof(lif::neuron) * en.size());
				h          = fnv1a(h, in.data(), sizeof(lif::neuron) * in.size());
				state_hash[t] = h;
			}
		}
		*raster_used = used;
		return 0;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}

// Conductance-based E/I network with the neuron of samples/vogels.cpp:10-46.
// Counts are stored at counts[step*3 + {1,2}] (slot 0 unused) to share the
// Brunel layout.
int ref_vogels_run(int64_t nE, int64_t nI, double p, float wex, float win, float dt, float delay,
                   int init_mode, uint32_t const* seeds, int nseeds, int64_t steps,
                   int32_t* counts, int32_t* raster, int64_t raster_cap, int64_t* raster_used,
                   uint64_t* raster_hash, uint64_t* state_hash) {
	try {
		snn net(dt, delay, make_seed(seeds, nseeds, 0));
		auto E = net.add_population<coba>(nE, {init_mode});
		auto I = net.add_population<coba>(nI, {init_mode});
		net.connect<excitatory>(E, E, fixed_probability(p), delay, {wex});
		net.connect<excitatory>(E, I, fixed_probability(p), delay, {wex});
		net.connect<inhibitory>(I, E, fixed_probability(p), delay, {win});
		net.connect<inhibitory>(I, I, fixed_probability(p), delay, {win});
		int64_t used = 0;
		for (int64_t t = 0; t < steps; t++) {
			net.step();
			counts[t * 3] = 0;
			uint64_t rh   = FNV_BASIS;
			record(E, t, 1, counts, raster, raster_cap, used, rh);
			record(I, t, 2, counts, raster, raster_cap, used, rh);
			if (raster_hash) raster_hash[t] = rh;
			if (state_hash) {
				uint64_t h = FNV_BASIS;
				auto en    = E->get_neurons();
				auto in    = I->get_neurons();
				h          = fnv1a(h, en.data(), sizeof(coba::neuron) * en.size());
				h          = fnv1a(h, in.data(), sizeof(coba::neuron) * in.size());
				state_hash[t] = h;
			}
		}
		*raster_used = used;
		return 0;
	} catch (std::exception const& e) {
		g_err = e.what();
		return -1;
	}
}
}
