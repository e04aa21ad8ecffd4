// The network (API of reference snn.h / snn.cpp) on the MI355X engine.
//
// snn(dt, max_delay, seed), add_population<Model>(n), connect<Synapse>(from,
// to, topology, delay, params) and step() keep the reference's signatures and
// semantics: the same seed consumption, the same per-step Kahan dt and RNG
// stream, updates in population order and deliveries in connect order.  step()
// queues the step; queued steps run on the device in windows when a result is
// read (population->spikes(), get_neurons()) or when a window is full.
//
// This header must be compiled with hipcc.  It ends by making every following
// function of the translation unit callable from device code, so that unchanged
// model structs (update(), deliver(), init()) compile into the kernels; a
// translation unit that defines further host-only functions after its models
// can end that region with #include "spice/host_code.h".
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <memory>
#include <span>
#include <stdexcept>
#include <utility>
#include <vector>

#include "spice/concepts.h"
#include "spice/detail/csr.h"
#include "spice/detail/kernels.h"
#include "spice/detail/neuron_population.h"
#include "spice/detail/synapse_population.h"
#include "spice/topology.h"
#include "spice/util/assert.h"
#include "spice/util/meta.h"
#include "spice/util/numeric.h"
#include "spice/util/random.h"
#include "spice/util/range.h"
#include "spice/util/scope.h"
#include "spice/util/stdint.h"
#include "spice/util/type_traits.h"
#include "spice_hip.h"

namespace spice {
class batch;

class snn {
public:
	snn(float const dt, float const max_delay, util::seed_seq seed) :
	snn(nullptr, dt, max_delay, std::move(seed)) {}
	snn(snn const&)            = delete;
	snn& operator=(snn const&) = delete;
	// Frees the network's device state; in a batch the other networks go on.
	~snn() {
		if (_eng) _eng->remove_instance(_inst);
	}

	template <Neuron Neur>
	detail::neuron_population<Neur>* add_population(Int const size, Neur neur = {}) {
		_eng->flush();  // steps queued so far ran without this population
		auto pop = std::make_unique<detail::neuron_population<Neur>>(_eng.get(), _inst, std::move(neur), size,
		                                                              _seed, _dt);
		auto* raw = pop.get();
		_eng->pops.push_back(raw);
		_eng->regroup = true;
		_neurons.push_back(std::move(pop));
		return raw;
	}

	template <class Syn, Neuron SrcNeur, StatefulNeuron DstNeur>
	requires Synapse<Syn, SrcNeur, DstNeur>
	void connect(detail::neuron_population<SrcNeur>* source,
	             detail::neuron_population<DstNeur>* target, Topology& c, float const delay,
	             Syn syn = {}) {
		Int const d = Int(std::round(delay / _dt));
		SPICE_PRE(d >= 1 && "The delay must be at least 1dt.");
		SPICE_PRE(
		    d <= _max_delay &&
		    "The delay of a synapse population may not exceed the maximum delay of the network.");
		c(source->size(), target->size());
		if (source->size() == 0)  // the reference's CSR then has no offsets (topology.cpp:106)
			throw std::logic_error("Assertion failed (topology.cpp:106): offsets.size() > src_count");
		_eng->flush();  // steps queued so far ran without this projection
		_synapses.push_back(std::make_unique<detail::synapse_population<Syn, SrcNeur, DstNeur>>(
		    _eng.get(), _inst, std::move(syn), c, _seed, d, source, target));
	}

	template <class Syn, Neuron SrcNeur, StatefulNeuron DstNeur>
	requires Synapse<Syn, SrcNeur, DstNeur>
	void connect(detail::neuron_population<SrcNeur>* source,
	             detail::neuron_population<DstNeur>* target, Topology&& c, float const delay,
	             Syn syn = {}) {
		connect<Syn, SrcNeur, DstNeur>(source, target, c, delay, std::move(syn));
	}

	void step() {
		float const dt = _simtime += _dt;
		if (_simtime >= 1) _simtime.reset();
		UInt128 const s = (_seed++).seed();
		_eng->queue(_inst, dt, _dt, s);
		_time++;
	}

	// ---- extensions of this engine (not in the reference API) -------------
	// Runs every queued step now.
	void sync() {
		_eng->flush();
		detail::check_spice(spice_ctx_sync(_eng->ctx));
	}
	// SPICE_DELIVER_PULL (default) or SPICE_DELIVER_PUSH; both exact.
	void set_delivery(int mode) {
		_eng->flush();
		detail::check_spice(spice_ctx_set_delivery(_eng->ctx, mode));
	}
	// Replays the neuron-update, compaction and delivery launches of equal
	// windows from HIP graphs.
	void set_graphs(bool on) { detail::check_spice(spice_ctx_set_graphs(_eng->ctx, on ? 1 : 0)); }
	spice_ctx* context() { return _eng->ctx; }
	int64_t instance() const { return _inst; }
	Int time() const { return _time; }

private:
	friend class batch;

	// eng == nullptr: a network with a context of its own
	snn(std::shared_ptr<detail::engine> eng, float const dt, float const max_delay, util::seed_seq seed) :
	_dt(dt), _max_delay(Int(std::round(max_delay / dt))), _seed(std::move(seed)), _eng(std::move(eng)) {
		if (!_eng) _eng = std::make_shared<detail::engine>(std::max<Int>(_max_delay, 1));
		_inst = _eng->add_instance(std::max<Int>(_max_delay, 1));
	}

	Int _time = 0;
	float _dt;
	Int _max_delay;
	util::kahan_sum<float> _simtime;
	util::seed_seq _seed;
	std::shared_ptr<detail::engine> _eng;
	int64_t _inst = 0;
	std::vector<std::unique_ptr<detail::NeuronPopulation>> _neurons;
	std::vector<std::unique_ptr<detail::SynapsePopulation>> _synapses;
};

// Many independent networks on one GPU (an extension; the reference runs one
// snn per process): parameter sweeps and seed ensembles.  Each network keeps
// the reference semantics of a stand-alone snn - its own seed chain, dt
// sequence, populations and projections - while every step launches one
// neuron-update kernel per neuron type, one compaction and one delivery for
// all of them.
class batch {
public:
	// max_delay_steps bounds every member network's max_delay / dt.
	explicit batch(Int max_delay_steps) :
	_eng(std::make_shared<detail::engine>(std::max<Int>(max_delay_steps, 1))) {}

	snn& add(float const dt, float const max_delay, util::seed_seq seed) {
		_nets.push_back(std::unique_ptr<snn>(new snn(_eng, dt, max_delay, std::move(seed))));
		return *_nets.back();
	}

	// Steps every network n times.
	void step(Int n = 1) {
		struct hold_guard {
			detail::engine& e;
			explicit hold_guard(detail::engine& x) : e(x) { e.hold = true; }
			~hold_guard() { e.hold = false; }
		} guard(*_eng);
		for (Int i = 0; i < n; i++) {
			for (auto& s : _nets) s->step();
			if (Int(_eng->most_pending()) >= _eng->window) _eng->flush();
		}
	}
	// Steps per device window (1 .. SPICE_MAX_WINDOW); with graphs enabled each
	// window length is captured once and replayed.
	void set_window(Int steps) {
		SPICE_PRE(1 <= steps && steps <= SPICE_MAX_WINDOW);
		_eng->flush();
		_eng->window = steps;
	}
	void sync() {
		_eng->flush();
		detail::check_spice(spice_ctx_sync(_eng->ctx));
	}
	// Launches every queued step (asynchronously; results are read on demand).
	void flush() { _eng->flush(); }
	// While held, stepping member networks only queues (no window is launched
	// when one network's queue fills), so steps queued network by network run
	// together in the next flush().
	void hold(bool on) { _eng->hold = on; }
	// Frees one network of the batch (its device memory is released; the others go on).
	void remove(snn& net) {
		std::erase_if(_nets, [&](std::unique_ptr<snn> const& p) { return p.get() == &net; });
	}
	void set_delivery(int mode) {
		_eng->flush();
		detail::check_spice(spice_ctx_set_delivery(_eng->ctx, mode));
	}
	void set_graphs(bool on) { detail::check_spice(spice_ctx_set_graphs(_eng->ctx, on ? 1 : 0)); }
	Int size() const { return Int(_nets.size()); }
	snn& operator[](Int i) { return *_nets[size_t(i)]; }
	spice_ctx* context() { return _eng->ctx; }

private:
	std::shared_ptr<detail::engine> _eng;
	std::vector<std::unique_ptr<snn>> _nets;
};
}  // namespace spice

// From here on every function of the translation unit is host+device, so that
// user model structs compile into the kernels above.  The device compilation
// pass hides `main` (host-only code) from this region.
#if defined(__HIP_DEVICE_COMPILE__)
#define main __attribute__((unused)) static inline spice_device_pass_main
#endif
#pragma clang force_cuda_host_device begin
