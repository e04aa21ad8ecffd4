// Benchmark and test networks written against the user API (include/spice/snn.h),
// exported with a C ABI for the Python tests and bench.py.
//
// The model structs are the reference samples' workloads (samples/brunel.cpp:23-70,
// samples/vogels.cpp:10-59) with their floating-point expressions unchanged;
// the Poisson rate is a member so that swept rates reuse the same struct (at
// 20 Hz it computes exactly what the sample's constant does).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "spice/serve.h"
#include "spice/snn.h"

using namespace spice;
using namespace spice::util;

namespace wl {
struct poisson {
	float rate = 20;
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

// Pair STDP on exponential traces: tau 20 ms, A+ 0.01 and A- 0.0105 relative
// to the initial weight w0, W clipped to [0, 2 w0].
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

// A synapse with per-synapse state whose init() draws: one uniform weight per
// synapse from the projection's init generator, in CSR order.
struct random_weight {
	float scale;
	struct synapse {
		float w = 0;
	};
	void init(synapse& syn, Int, Int, auto& rng) const { syn.w = util::generate_canonical<float>(rng) * scale; }
	void deliver(synapse const& syn, lif::neuron& to) const { to.V += syn.w; }
};

struct coba {
	int init_mode = 0;  // 1: V ~ U[-0.06, -0.05)
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
				n.V += ((Vrest - n.V) + n.Gex * (Eex - n.V) + n.Gin * (Ein - n.V) + Ibg) * (dt * TmemInv);
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
struct inhibitory {
	float weight;
	void deliver(coba::neuron& to) const { to.Gin += weight; }
};
}  // namespace wl

#include "spice/host_code.h"

namespace {
thread_local std::string g_err;

struct network {
	std::unique_ptr<snn> net;
	std::vector<spice::detail::NeuronPopulation*> pops;
	int kind = 0;  // 0 Brunel (P, E, I lif), 1 conductance (E, I), 3 P -> lif with random weights
};

// Many Brunel networks in one spice::batch.
struct brunel_batch {
	std::unique_ptr<spice::batch> b;
	std::vector<std::vector<spice::detail::NeuronPopulation*>> pops;  // per instance: P, E, I
};

template <class Net>
std::vector<spice::detail::NeuronPopulation*> build_brunel(Net& s, int64_t nP, int64_t nE, int64_t nI, double p,
                                                          float rate, float wex, float win, float delay) {
	auto P = s.template add_population<wl::poisson>(nP, {rate});
	auto E = s.template add_population<wl::lif>(nE);
	auto I = s.template add_population<wl::lif>(nI);
	s.template connect<wl::fixed_weight>(P, E, fixed_probability(p), delay, {wex});
	s.template connect<wl::fixed_weight>(P, I, fixed_probability(p), delay, {wex});
	s.template connect<wl::fixed_weight>(E, E, fixed_probability(p), delay, {wex});
	s.template connect<wl::fixed_weight>(E, I, fixed_probability(p), delay, {wex});
	s.template connect<wl::fixed_weight>(I, E, fixed_probability(p), delay, {win});
	s.template connect<wl::fixed_weight>(I, I, fixed_probability(p), delay, {win});
	return {P, E, I};
}

int fail(std::exception const& e) {
	g_err = e.what();
	return dynamic_cast<std::logic_error const*>(&e) ? -3 : -1;
}

seed_seq make_seed(uint32_t const* seeds, int n) {
	if (n <= 0) throw std::logic_error("Assertion failed (random.h:151): il.size() > 0");
	UInt128 const h = util::detail::murmur3(seeds, 4 * UInt(n));
	return seed_seq::from_value(h);
}
}  // namespace

extern "C" {

char const* wl_error() { return g_err.c_str(); }

// Brunel network of samples/brunel.cpp:75-103 with parameterised sizes,
// probability, Poisson rate, weights and seed.  mode: SPICE_DELIVER_*.
void* wl_brunel_create(int64_t nP, int64_t nE, int64_t nI, double p, float rate, float wex, float win,
                       float dt, float delay, uint32_t const* seeds, int nseeds, int mode) {
	try {
		auto w  = std::make_unique<network>();
		w->net  = std::make_unique<snn>(dt, delay, make_seed(seeds, nseeds));
		auto& s = *w->net;
		s.set_delivery(mode);
		w->pops = build_brunel(s, nP, nE, nI, p, rate, wex, win, delay);
		w->kind = 0;
		return w.release();
	} catch (std::exception const& e) {
		fail(e);
		return nullptr;
	}
}

// P (Poisson, rate) -> L (LIF) over fixed_probability(p) with random_weight
// synapses (their init() draws, so it runs on the host in CSR order).
void* wl_random_weight_create(int64_t nP, int64_t nL, double p, float rate, float scale, float dt, float delay,
                              uint32_t const* seeds, int nseeds) {
	try {
		auto w  = std::make_unique<network>();
		w->net  = std::make_unique<snn>(dt, delay, make_seed(seeds, nseeds));
		auto& s = *w->net;
		auto P  = s.add_population<wl::poisson>(nP, {rate});
		auto L  = s.add_population<wl::lif>(nL);
		s.connect<wl::random_weight>(P, L, fixed_probability(p), delay, {scale});
		w->pops = {P, L};
		w->kind = 3;
		return w.release();
	} catch (std::exception const& e) {
		fail(e);
		return nullptr;
	}
}

// Conductance-based E/I network (samples/vogels.cpp:10-73), init_mode 1 draws V0.
void* wl_coba_create(int64_t nE, int64_t nI, double p, float wex, float win, float dt, float delay,
                     int init_mode, uint32_t const* seeds, int nseeds, int mode) {
	try {
		auto w  = std::make_unique<network>();
		w->net  = std::make_unique<snn>(dt, delay, make_seed(seeds, nseeds));
		auto& s = *w->net;
		s.set_delivery(mode);
		auto E = s.add_population<wl::coba>(nE, {init_mode});
		auto I = s.add_population<wl::coba>(nI, {init_mode});
		s.connect<wl::excitatory>(E, E, fixed_probability(p), delay, {wex});
		s.connect<wl::excitatory>(E, I, fixed_probability(p), delay, {wex});
		s.connect<wl::inhibitory>(I, E, fixed_probability(p), delay, {win});
		s.connect<wl::inhibitory>(I, I, fixed_probability(p), delay, {win});
		w->pops = {E, I};
		w->kind = 1;
		return w.release();
	} catch (std::exception const& e) {
		fail(e);
		return nullptr;
	}
}

// Brunel+ (samples/brunel+.cpp:101-117): the E->E projection is plastic, with
// rule 1 (the sample's `plastic`) or rule 2 (stdp_pair with w0).
void* wl_brunel_plastic_create(int64_t nP, int64_t nE, int64_t nI, double p, float rate, float wex, float win,
                               float dt, float delay, uint32_t const* seeds, int nseeds, int rule, float w0,
                               int mode) {
	try {
		auto w  = std::make_unique<network>();
		w->net  = std::make_unique<snn>(dt, delay, make_seed(seeds, nseeds));
		auto& s = *w->net;
		s.set_delivery(mode);
		auto P = s.add_population<wl::poisson>(nP, {rate});
		auto E = s.add_population<wl::lif>(nE);
		auto I = s.add_population<wl::lif>(nI);
		s.connect<wl::fixed_weight>(P, E, fixed_probability(p), delay, {wex});
		s.connect<wl::fixed_weight>(P, I, fixed_probability(p), delay, {wex});
		if (rule == 1)
			s.connect<wl::plastic>(E, E, fixed_probability(p), delay);
		else
			s.connect<wl::stdp_pair>(E, E, fixed_probability(p), delay, {w0});
		s.connect<wl::fixed_weight>(E, I, fixed_probability(p), delay, {wex});
		s.connect<wl::fixed_weight>(I, E, fixed_probability(p), delay, {win});
		s.connect<wl::fixed_weight>(I, I, fixed_probability(p), delay, {win});
		w->pops = {P, E, I};
		w->kind = 0;
		return w.release();
	} catch (std::exception const& e) {
		fail(e);
		return nullptr;
	}
}

void wl_destroy(void* h) { delete static_cast<network*>(h); }

int wl_step(void* h, int64_t n) {
	try {
		auto* w = static_cast<network*>(h);
		for (int64_t i = 0; i < n; i++) w->net->step();
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_sync(void* h) {
	try {
		static_cast<network*>(h)->net->sync();
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_spikes(void* h, int pop, int64_t age, int32_t* out, int64_t cap, int64_t* count) {
	try {
		auto* w = static_cast<network*>(h);
		auto s  = w->pops.at(size_t(pop))->spikes(age);
		*count  = int64_t(s.size());
		if (int64_t(s.size()) > cap) throw std::runtime_error("spike buffer too small");
		std::memcpy(out, s.data(), sizeof(int32_t) * s.size());
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

// Copies a stateful population's neurons (the reference's AoS layout).
int wl_neurons(void* h, int pop, void* out, int64_t bytes) {
	try {
		auto* w = static_cast<network*>(h);
		auto* p = w->pops.at(size_t(pop));
		void* src = p->neurons();
		if (!src) throw std::runtime_error("stateless population");
		int64_t const need = p->size() * (w->kind == 1 ? int64_t(sizeof(wl::coba::neuron))
		                                               : int64_t(sizeof(wl::lif::neuron)));
		if (bytes < need) throw std::runtime_error("state buffer too small");
		std::memcpy(out, src, size_t(need));
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

spice_ctx* wl_context(void* h) { return static_cast<network*>(h)->net->context(); }

int wl_set_graphs(void* h, int on) {
	try {
		static_cast<network*>(h)->net->set_graphs(on != 0);
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

// n Brunel networks in one batch: instance i has N[i] neurons split P/E/I as
// N/2, N*4/10, N/10, probability p[i], Poisson rate rate[i], weights
// +0.2/(N p) and -1/(N p), and seed {seeds[2i], seeds[2i+1]}.
void* wl_batch_brunel_create(int n, int64_t const* N, double const* p, float const* rate,
                             uint32_t const* seeds, float dt, float delay, int mode) {
	try {
		auto w = std::make_unique<brunel_batch>();
		w->b   = std::make_unique<spice::batch>(Int(std::lround(delay / dt)));
		w->b->set_delivery(mode);
		for (int i = 0; i < n; i++) {
			auto& s         = w->b->add(dt, delay, make_seed(seeds + 2 * i, 2));
			float const wex = float(0.2 / (double(N[i]) * p[i]));
			float const win = float(-1.0 / (double(N[i]) * p[i]));
			w->pops.push_back(build_brunel(s, N[i] / 2, N[i] * 4 / 10, N[i] / 10, p[i], rate[i], wex, win, delay));
		}
		return w.release();
	} catch (std::exception const& e) {
		fail(e);
		return nullptr;
	}
}

void wl_batch_destroy(void* h) { delete static_cast<brunel_batch*>(h); }

int wl_batch_step(void* h, int64_t n) {
	try {
		static_cast<brunel_batch*>(h)->b->step(n);
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_batch_sync(void* h) {
	try {
		static_cast<brunel_batch*>(h)->b->sync();
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_batch_set_graphs(void* h, int on) {
	try {
		static_cast<brunel_batch*>(h)->b->set_graphs(on != 0);
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_batch_spikes(void* h, int inst, int pop, int64_t age, int32_t* out, int64_t cap, int64_t* count) {
	try {
		auto* w = static_cast<brunel_batch*>(h);
		auto s  = w->pops.at(size_t(inst)).at(size_t(pop))->spikes(age);
		*count  = int64_t(s.size());
		if (int64_t(s.size()) > cap) throw std::runtime_error("spike buffer too small");
		std::memcpy(out, s.data(), sizeof(int32_t) * s.size());
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

spice_ctx* wl_batch_context(void* h) { return static_cast<brunel_batch*>(h)->b->context(); }

// ---- served job stream -----------------------------------------------------
// Job kinds: 0 Brunel (P/E/I = N/2, 4N/10, N/10, delay 15 steps, weights
// +0.2/(N p), -1/(N p)), 1 conductance E/I (E/I = 4N/5, N/5, delay 1 step,
// V0 ~ U[-60,-50] mV, weights 0.27 and 4.5 (x 80/(N p)) of the leak conductance,
// as in samples/vogels.cpp), 2 Brunel whose E->E projection is pair-STDP plastic.
// The result of a job is an FNV-1a hash of its stateful populations' neurons
// after its last step (AoS, population order) plus the spike counts of that step.
namespace {
uint64_t fnv1a(uint64_t h, void const* p, size_t n) {
	auto const* b = static_cast<unsigned char const*>(p);
	for (size_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 0x100000001b3ull;
	}
	return h;
}

std::vector<spice::detail::NeuronPopulation*> build_job(snn& s, int kind, int64_t N, double p, float delay) {
	if (kind == 0) {
		float const wex = float(0.2 / (double(N) * p)), win = float(-1.0 / (double(N) * p));
		return build_brunel(s, N / 2, N * 4 / 10, N / 10, p, 20.0f, wex, win, delay);
	}
	if (kind == 1) {
		float const scale = float(80.0 / (double(N) * p));
		auto E            = s.add_population<wl::coba>(N * 4 / 5, {1});
		auto I            = s.add_population<wl::coba>(N / 5, {1});
		s.connect<wl::excitatory>(E, E, fixed_probability(p), delay, {0.27f * scale});
		s.connect<wl::excitatory>(E, I, fixed_probability(p), delay, {0.27f * scale});
		s.connect<wl::inhibitory>(I, E, fixed_probability(p), delay, {4.5f * scale});
		s.connect<wl::inhibitory>(I, I, fixed_probability(p), delay, {4.5f * scale});
		return {E, I};
	}
	float const wex = float(0.2 / (double(N) * p)), win = float(-1.0 / (double(N) * p));
	auto P = s.add_population<wl::poisson>(N / 2, {20.0f});
	auto E = s.add_population<wl::lif>(N * 4 / 10);
	auto I = s.add_population<wl::lif>(N / 10);
	s.connect<wl::fixed_weight>(P, E, fixed_probability(p), delay, {wex});
	s.connect<wl::fixed_weight>(P, I, fixed_probability(p), delay, {wex});
	s.connect<wl::stdp_pair>(E, E, fixed_probability(p), delay, {wex});
	s.connect<wl::fixed_weight>(E, I, fixed_probability(p), delay, {wex});
	s.connect<wl::fixed_weight>(I, E, fixed_probability(p), delay, {win});
	s.connect<wl::fixed_weight>(I, I, fixed_probability(p), delay, {win});
	return {P, E, I};
}

uint64_t job_result(int kind, std::vector<spice::detail::NeuronPopulation*> const& pops) {
	uint64_t h = 0xcbf29ce484222325ull;
	for (auto* p : pops) {
		void* n = p->neurons();
		if (n) h = fnv1a(h, n, size_t(p->size()) * (kind == 1 ? sizeof(wl::coba::neuron) : sizeof(wl::lif::neuron)));
	}
	for (auto* p : pops) {
		int64_t const c = int64_t(p->spikes(0).size());
		h               = fnv1a(h, &c, sizeof(c));
	}
	return h;
}

float job_delay(int kind) { return kind == 1 ? 1e-4f : 15e-4f; }
}  // namespace

// Runs n jobs through spice::server (device-memory budget `budget` bytes,
// windows of `window` steps).  Per job: kind, N, p, steps, arrival (s), seed
// words seeds[2i..2i+1] and a footprint estimate in bytes.  Out per job:
// admitted/finished times (s) and the result hash; *events = synaptic events.
int wl_serve_run(int n, int const* kind, int64_t const* N, double const* p, int64_t const* steps,
                 double const* arrival, uint32_t const* seeds, int64_t const* footprint, int64_t budget,
                 int64_t window, int graphs, double* admitted, double* finished, uint64_t* result, int64_t* events,
                 double* phase_seconds) {
	try {
		spice::server srv(budget, window, 15, graphs != 0);
		srv.on_progress([](size_t done, size_t total, double t) {
			std::fprintf(stderr, "served %zu/%zu jobs after %.1f s\n", done, total, t);
			std::fflush(stderr);
		});
		for (int i = 0; i < n; i++) {
			spice::request r;
			r.max_delay   = job_delay(kind[i]);
			r.seed        = make_seed(seeds + 2 * i, 2);
			r.steps       = steps[i];
			r.footprint   = footprint[i];
			r.single_step = kind[i] != 0;
			r.arrival     = arrival[i];
			auto pops     = std::make_shared<std::vector<spice::detail::NeuronPopulation*>>();
			int const k   = kind[i];
			int64_t const Ni = N[i];
			double const pi  = p[i];
			r.build  = [=](snn& s) { *pops = build_job(s, k, Ni, pi, job_delay(k)); };
			r.finish = [=](snn&) { result[i] = job_result(k, *pops); };
			srv.submit(std::move(r));
		}
		auto recs = srv.run();
		for (int i = 0; i < n; i++) {
			admitted[i] = recs[size_t(i)].admitted;
			finished[i] = recs[size_t(i)].finished;
		}
		*events          = srv.events();
		phase_seconds[0] = srv.build_seconds();
		phase_seconds[1] = srv.round_seconds();
		phase_seconds[2] = srv.retire_seconds();
		phase_seconds[3] = srv.generation_seconds();
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

// One job of wl_serve_run on a network of its own (the reference's way):
// the result hash, for comparison with the served run.
int wl_job_alone(int kind, int64_t N, double p, int64_t steps, uint32_t const* seed2, uint64_t* result) {
	try {
		snn s(1e-4f, job_delay(kind), make_seed(seed2, 2));
		auto pops = build_job(s, kind, N, p, job_delay(kind));
		for (int64_t t = 0; t < steps; t++) s.step();
		*result = job_result(kind, pops);
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}

int wl_batch_set_window(void* h, int64_t steps) {
	try {
		static_cast<brunel_batch*>(h)->b->set_window(steps);
		return 0;
	} catch (std::exception const& e) {
		return fail(e);
	}
}
}
