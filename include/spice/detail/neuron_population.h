// Populations of one user neuron type (API of reference detail/neuron_population.h),
// backed by device state in libspice_hip.  Also the engine that runs pending steps.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <vector>

#include "spice/concepts.h"
#include "spice/detail/kernels.h"
#include "spice/topology.h"
#include "spice/util/assert.h"
#include "spice/util/random.h"
#include "spice/util/range.h"
#include "spice/util/stdint.h"
#include "spice_hip.h"

namespace spice::detail {

inline void check_hip(hipError_t e, char const* what) {
	if (e != hipSuccess) throw std::runtime_error(std::string("spice: ") + what + ": " + hipGetErrorString(e));
}

struct NeuronPopulation;

// Owns the device context; steps are queued by snn::step() and run in blocks
// of up to min(16, smallest delay) steps when results are needed.
struct engine {
	spice_ctx* ctx = nullptr;
	std::vector<NeuronPopulation*> pops;  // context population order
	std::vector<float> pend_dt;
	std::vector<UInt> pend_seed;  // lo, hi per step
	Int min_delay = SPICE_MAX_BLOCK;

	explicit engine(Int max_delay) { check_spice(spice_ctx_create(0, max_delay, &ctx)); }
	~engine() { spice_ctx_destroy(ctx); }
	engine(engine const&)            = delete;
	engine& operator=(engine const&) = delete;

	hipStream_t stream() const { return static_cast<hipStream_t>(spice_ctx_stream(ctx)); }
	void flush();
};

struct NeuronPopulation {
	virtual ~NeuronPopulation()                       = default;
	virtual Int size() const                          = 0;
	virtual void* neurons()                           = 0;
	virtual std::span<Int32 const> spikes(Int age) const = 0;
	virtual void plastic()                            = 0;
	virtual std::span<UInt const> history() const     = 0;

	// engine hooks
	virtual void prepare_block()                                                     = 0;
	virtual void launch(spice_pop_view const& v, spice_block_view const& b, hipStream_t s) = 0;
	virtual int64_t id() const                                                       = 0;
};

inline void engine::flush() {
	Int done = 0;
	Int const total = Int(pend_dt.size());
	if (total == 0) return;
	for (auto* p : pops) p->prepare_block();
	std::vector<spice_pop_view> views(std::max<size_t>(pops.size(), 1));
	while (done < total) {
		int32_t const n = int32_t(std::min<Int>({total - done, Int(SPICE_MAX_BLOCK), min_delay}));
		spice_block_view b;
		check_spice(spice_block_begin(ctx, n, pend_dt.data() + done, pend_seed.data() + 2 * done, &b,
		                              views.data(), int32_t(views.size())));
		for (size_t p = 0; p < pops.size(); p++) pops[p]->launch(views[p], b, stream());
		check_spice(spice_block_end(ctx));
		done += n;
	}
	pend_dt.clear();
	pend_seed.clear();
}

// Incoming projection of a population, as its update kernel replays it.
struct incoming {
	std::type_index type;
	void const* obj;   // device copy of the synapse functor
	replay_fn fn;      // device address of its replay thunk
};

template <Neuron Neur>
class neuron_population : public NeuronPopulation {
	static_assert(!PerPopulationUpdate<Neur>,
	              "populations with a per-population update() are not supported by the device engine");

public:
	using neuron_t = neuron_traits_t<Neur>;
	static constexpr bool stateful = StatefulNeuron<Neur>;

	neuron_population(engine* eng, Neur neuron, Int const size, util::seed_seq& seed, float dt) :
	_eng(eng), _neuron(std::move(neuron)), _size(size) {
		SPICE_INV(size >= 0);
		int32_t words = 0;
		if constexpr (stateful) {
			static_assert(std::is_trivially_copyable_v<neuron_t>,
			              "neuron state must be trivially copyable to live in device memory");
			words = int32_t((sizeof(neuron_t) + 3) / 4);
		}
		_draws = probe_draws(dt);
		check_spice(spice_pop_create(eng->ctx, size, words, _draws, &_id));
		if constexpr (stateful) {
			// stateful adapters take a seed whether or not they initialise (neuron_population.h:60)
			util::xoroshiro64_128p rng(seed++);
			_host.assign(size_t(size), neuron_t{});
			if constexpr (PerNeuronInit<Neur>) {
				Int id = 0;
				for (auto& n : _host) _neuron.init(n, id++, rng);
			} else if constexpr (PerPopulationInit<Neur>) {
				_neuron.init(std::span<neuron_t>(_host), rng);
			}
			_host_dirty = true;
			_host_valid = true;
		}
	}

	~neuron_population() override {
		if (_slots_dev) (void)hipFree(_slots_dev);
	}

	Int size() const override { return _size; }
	int64_t id() const override { return _id; }

	void* neurons() override {
		if constexpr (stateful)
			return get_neurons().data();
		else
			return nullptr;
	}

	// Neuron state on the host.  The span stays valid until the next step();
	// edits made through it are uploaded before the next step runs.
	auto get_neurons() {
		static_assert(stateful, "Can only return collections of stateful neurons.");
		_eng->flush();
		if (!_host_valid) {
			apply_pending();
			check_spice(spice_pop_download(_eng->ctx, _id, _host.data(), int32_t(sizeof(neuron_t))));
			_host_valid = true;
		}
		_host_dirty = true;  // the caller may write through the span
		return std::span<neuron_t>(_host);
	}

	std::span<Int32 const> spikes(Int age) const override {
		_eng->flush();
		int64_t cnt = 0;
		_spike_buf.resize(size_t(std::max<Int>(_size, 1)));
		check_spice(spice_spikes(_eng->ctx, _id, age, _spike_buf.data(), Int(_spike_buf.size()), &cnt));
		return {_spike_buf.data(), size_t(cnt)};
	}

	void plastic() override {
		throw std::logic_error("plastic synapses are not supported by the device engine");
	}
	std::span<UInt const> history() const override { return {}; }

	// Registers an incoming projection (connect order).  The first synapse type
	// into a population is inlined into its update kernel; other types are
	// replayed through a device function pointer.
	template <class Syn>
	void add_incoming(Syn const& syn) {
		void* obj = nullptr;
		check_hip(hipMalloc(&obj, sizeof(Syn)), "hipMalloc");
		check_hip(hipMemcpy(obj, &syn, sizeof(Syn), hipMemcpyHostToDevice), "hipMemcpy");
		_syn_objs.emplace_back(obj, [](void* p) { (void)hipFree(p); });
		replay_fn* d_fn = nullptr;
		check_hip(hipMalloc(&d_fn, sizeof(replay_fn)), "hipMalloc");
		hipLaunchKernelGGL((replay_address_kernel<Syn, neuron_t>), dim3(1), dim3(1), 0, 0, d_fn);
		check_hip(hipGetLastError(), "replay address");
		replay_fn fn = nullptr;
		check_hip(hipMemcpy(&fn, d_fn, sizeof(fn), hipMemcpyDeviceToHost), "hipMemcpy");
		(void)hipFree(d_fn);
		_in.push_back({std::type_index(typeid(Syn)), obj, fn});
		if (_in.size() == 1) {
			_primary = std::type_index(typeid(Syn));
			_launch_primary = [this](spice_pop_view const& v, spice_block_view const& b, hipStream_t s) {
				run<Syn>(v, b, s);
			};
			_apply_primary = [this](spice_pop_view const& v, int64_t step, hipStream_t s) {
				if constexpr (stateful) {
					unsigned const grid = unsigned((_size + 255) / 256);
					hipLaunchKernelGGL((apply_pending_kernel<Neur, Syn>), dim3(grid), dim3(256), 0, s, v, step,
					                   _slots_dev);
					check_hip(hipGetLastError(), "apply pending deliveries");
				}
			};
		}
		_slots_dirty = true;
	}

	void prepare_block() override {
		if constexpr (stateful) {
			if (_host_dirty) {
				check_spice(spice_pop_upload(_eng->ctx, _id, _host.data(), int32_t(sizeof(neuron_t))));
				_host_dirty = false;
			}
			_host_valid = false;  // state changes on the device from here on
		}
		if (_slots_dirty) {
			std::vector<replay_slot> h;
			for (auto& in : _in) h.push_back({in.obj, in.type == _primary ? nullptr : in.fn});
			if (_slots_dev) (void)hipFree(_slots_dev);
			_slots_dev = nullptr;
			check_hip(hipMalloc(&_slots_dev, sizeof(replay_slot) * std::max<size_t>(h.size(), 1)), "hipMalloc");
			if (!h.empty())
				check_hip(hipMemcpy(_slots_dev, h.data(), sizeof(replay_slot) * h.size(), hipMemcpyHostToDevice),
				          "hipMemcpy");
			_slots_dirty = false;
		}
	}

	void launch(spice_pop_view const& v, spice_block_view const& b, hipStream_t s) override {
		if (_size == 0) return;
		if (_launch_primary)
			_launch_primary(v, b, s);
		else
			run<void>(v, b, s);
	}

	// Applies the deliveries already counted for the next step (see kernels.h).
	void apply_pending() {
		if constexpr (stateful) {
			if (_in.empty() || _size == 0) return;
			prepare_block();
			spice_pop_view v;
			check_spice(spice_pop_view_now(_eng->ctx, _id, &v));
			int64_t const step = spice_steps(_eng->ctx);
			_apply_primary(v, step, _eng->stream());
		}
	}

private:
	engine* _eng;
	Neur _neuron;
	Int _size;
	int64_t _id   = -1;
	int32_t _draws = 0;
	std::vector<std::conditional_t<stateful, neuron_t, char>> _host;
	bool _host_valid = false, _host_dirty = false;
	mutable std::vector<Int32> _spike_buf;
	std::vector<incoming> _in;
	std::vector<std::unique_ptr<void, void (*)(void*)>> _syn_objs;
	std::type_index _primary = std::type_index(typeid(void));
	std::function<void(spice_pop_view const&, spice_block_view const&, hipStream_t)> _launch_primary;
	std::function<void(spice_pop_view const&, int64_t, hipStream_t)> _apply_primary;
	replay_slot* _slots_dev = nullptr;
	bool _slots_dirty       = true;

	// RNG outputs update() takes per neuron, probed on the host with a default
	// neuron.  The device kernels check every neuron against it.
	int32_t probe_draws(float dt) const {
		probe_stream rng;
		if constexpr (stateful) {
			neuron_t n{};
			(void)_neuron.update(n, dt, rng);
		} else {
			(void)_neuron.update(dt, rng);
		}
		return rng.count;
	}

	template <class Syn>
	void run(spice_pop_view const& v, spice_block_view const& b, hipStream_t s) {
		if (_draws == 0) {
			unsigned const grid = unsigned((_size + 255) / 256);
			hipLaunchKernelGGL((update_lanes_kernel<Neur, Syn>), dim3(grid), dim3(256), 0, s, v, b, _neuron,
			                   _slots_dev);
		} else {
			Int const chunks    = (_size + 15) / 16;
			unsigned const grid = unsigned((chunks + 255) / 256);
			hipLaunchKernelGGL((update_chunks_kernel<Neur, Syn>), dim3(grid), dim3(256), 0, s, v, b,
			                   _neuron, _slots_dev);
		}
		check_hip(hipGetLastError(), "neuron update launch");
	}
};
}  // namespace spice::detail
