// Populations of one user neuron type (API of reference detail/neuron_population.h),
// backed by device state in libspice_hip, and the engine that runs queued steps.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <utility>
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

// Device buffer owned by the host-side engine objects.
template <class T>
struct dev_array {
	T* p     = nullptr;
	size_t n = 0;
	dev_array() = default;
	dev_array(dev_array const&)            = delete;
	dev_array& operator=(dev_array const&) = delete;
	~dev_array() { reset(); }
	void reset() {
		if (p) (void)hipFree(p);
		p = nullptr;
		n = 0;
	}
	void assign(std::vector<T> const& h) {
		reset();
		check_hip(hipMalloc(&p, sizeof(T) * std::max<size_t>(h.size(), 1)), "hipMalloc");
		n = h.size();
		if (n) check_hip(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy");
	}
};

// The populations of one neuron type (and one inlined synapse type) in every
// instance of a context: one kernel launch per block updates all of them.
struct group_base {
	virtual ~group_base()                                        = default;
	virtual void add(NeuronPopulation* p)                        = 0;
	virtual void rebuild()                                       = 0;
	virtual void launch(spice_block const& b, hipStream_t s)     = 0;
	// updated once per window, every step at once (stateless populations that draw)
	virtual bool per_window() const                              = 0;
};

// The plastic projections of one synapse type (in every instance): one sweep
// and one delivery launch per step cover all of them.
struct plastic_group_base {
	virtual ~plastic_group_base()                                 = default;
	virtual void prepare()                                        = 0;  // device tables, before a window
	virtual void launch(spice_block const& b, hipStream_t s)      = 0;
	virtual void remove_instance(int64_t inst)                    = 0;
};

// Owns the device context.  Steps are queued per instance (snn::step()) and
// run in windows when results are needed or a window is full.
struct engine {
	spice_ctx* ctx = nullptr;
	std::vector<NeuronPopulation*> pops;  // context population order
	std::vector<std::vector<spice_step>> pending;  // per instance
	std::vector<std::pair<std::type_index, std::unique_ptr<group_base>>> groups;
	bool regroup = true;
	Int min_delay = SPICE_MAX_BLOCK;
	Int window    = SPICE_MAX_WINDOW;  // steps per spice_run() (graphs are cached per length)
	bool hold     = false;             // a batch queues whole rounds before flushing
	std::string cb_error;
	// plastic projections by synapse type (synapse phase)
	std::vector<std::pair<std::type_index, std::unique_ptr<plastic_group_base>>> plastic_groups;

	explicit engine(Int max_delay) {
		check_spice(spice_ctx_create(0, max_delay, &ctx));
		check_spice(spice_ctx_set_update(ctx, &engine::launch_all, this));
	}
	~engine() { spice_ctx_destroy(ctx); }
	engine(engine const&)            = delete;
	engine& operator=(engine const&) = delete;

	int64_t add_instance(Int max_delay) {
		int64_t id = 0;
		check_spice(spice_instance_create(ctx, max_delay, &id));  // may reuse a destroyed instance's id
		if (pending.size() <= size_t(id)) pending.resize(size_t(id) + 1);
		pending[size_t(id)].clear();
		return id;
	}

	hipStream_t stream() const { return static_cast<hipStream_t>(spice_ctx_stream(ctx)); }

	void queue(int64_t inst, float dt, float dt_nominal, UInt128 seed) {
		spice_step s{};
		s.dt         = dt;
		s.dt_nominal = dt_nominal;
		s.seed       = {seed.lo, seed.hi};
		pending[size_t(inst)].push_back(s);
		if (!hold && Int(pending[size_t(inst)].size()) >= window) flush();
	}

	size_t most_pending() const {
		size_t m = 0;
		for (auto& q : pending) m = std::max(m, q.size());
		return m;
	}

	void flush();
	// Drops an instance's queued steps, populations and plastic launchers and
	// frees its device memory (a network that is destroyed).
	void remove_instance(int64_t inst) noexcept;

	static int launch_all(void* user, spice_block const* b, void* stream) {
		auto* e = static_cast<engine*>(user);
		try {
			if (b->phase == SPICE_PHASE_NEURONS || b->phase == SPICE_PHASE_WINDOW) {
				bool const window = b->phase == SPICE_PHASE_WINDOW;
				for (auto& g : e->groups)
					if (g.second->per_window() == window) g.second->launch(*b, static_cast<hipStream_t>(stream));
			} else
				for (auto& g : e->plastic_groups) g.second->launch(*b, static_cast<hipStream_t>(stream));
		} catch (std::exception const& x) {
			e->cb_error = x.what();
			return 1;
		}
		return 0;
	}
};

struct NeuronPopulation {
	virtual ~NeuronPopulation()                          = default;
	virtual Int size() const                             = 0;
	virtual void* neurons()                              = 0;
	virtual std::span<Int32 const> spikes(Int age) const = 0;
	virtual void plastic()                               = 0;
	virtual std::span<UInt const> history() const        = 0;

	// engine hooks
	virtual void prepare()                                 = 0;  // before a window
	virtual std::type_index group_key() const              = 0;
	virtual std::unique_ptr<group_base> make_group() const = 0;
	virtual int64_t id() const                             = 0;
	virtual int64_t instance() const                       = 0;
};

inline void engine::remove_instance(int64_t inst) noexcept {
	if (size_t(inst) < pending.size()) pending[size_t(inst)].clear();
	std::erase_if(pops, [&](NeuronPopulation* p) { return p->instance() == inst; });
	for (auto& g : plastic_groups) g.second->remove_instance(inst);
	groups.clear();
	regroup = true;
	(void)spice_instance_destroy(ctx, inst);
}

inline void engine::flush() {
	size_t const most = most_pending();
	if (most == 0) return;
	for (auto* p : pops) p->prepare();
	for (auto& g : plastic_groups) g.second->prepare();
	if (regroup) {
		groups.clear();
		for (auto* p : pops) {
			auto const key = p->group_key();
			auto it        = std::find_if(groups.begin(), groups.end(), [&](auto& g) { return g.first == key; });
			if (it == groups.end()) {
				groups.emplace_back(key, p->make_group());
				it = groups.end() - 1;
			}
			it->second->add(p);
		}
		for (auto& g : groups) g.second->rebuild();
		regroup = false;
	}
	size_t const ninst = pending.size();
	std::vector<spice_step> win;
	size_t done = 0;
	while (done < most) {
		int32_t const n = int32_t(std::min<size_t>(most - done, size_t(window)));
		win.assign(ninst * size_t(n), spice_step{});
		for (size_t i = 0; i < ninst; i++) {
			int64_t const base = spice_instance_steps(ctx, int64_t(i));
			for (int32_t k = 0; k < n; k++) {
				spice_step& s = win[i * size_t(n) + size_t(k)];
				if (done + size_t(k) < pending[i].size()) {
					s       = pending[i][done + size_t(k)];
					s.local = base + k;
				} else {
					s.local = -1;
				}
			}
		}
		int const r = spice_run(ctx, n, win.data());
		if (r != SPICE_OK && !cb_error.empty()) {
			std::string const m = cb_error;
			cb_error.clear();
			throw std::runtime_error("spice: " + m);
		}
		check_spice(r);
		done += size_t(n);
	}
	for (auto& q : pending) q.clear();
}

// Incoming projection of a population, as its update kernel replays it.
struct incoming {
	std::type_index type;
	void const* obj;   // device copy of the synapse functor
	replay_fn fn;      // device address of its counted replay thunk (stateless synapses)
	events_fn ev_fn;   // device address of its event thunk (stateful synapses)
};

template <Neuron Neur>
class neuron_population;

// Kernel shape of a population: one thread per neuron (no draws), per
// (16-neuron chunk, step) (stateless neurons that draw), or per 16-neuron chunk.
enum class update_kind { lanes, stateless, chunks };

template <class Neur, class Syn, update_kind Kind, int MaxIn, bool Mixed>
struct group final : group_base {
	std::vector<neuron_population<Neur>*> members;
	dev_array<group_entry> wg;
	dev_array<Neur> neurs;
	dev_array<replay_slot const*> slots;
	unsigned nwg = 0;

	void add(NeuronPopulation* p) override { members.push_back(static_cast<neuron_population<Neur>*>(p)); }

	void rebuild() override {
		std::vector<group_entry> w;
		std::vector<Neur> nh;
		std::vector<replay_slot const*> sh;
		int64_t const per = Kind == update_kind::lanes ? 256 : 256;  // elements per workgroup
		for (size_t m = 0; m < members.size(); m++) {
			auto* p = members[m];
			nh.push_back(p->functor());
			sh.push_back(p->slots_device());
			// stateless: every 16-neuron chunk of the population's whole spike words
			int64_t const elems = Kind == update_kind::lanes       ? p->size()
			                      : Kind == update_kind::stateless ? (p->size() + 63) / 64 * 4
			                                                       : (p->size() + 15) / 16;
			for (int64_t f = 0; f < elems; f += per) w.push_back({int32_t(p->id()), int32_t(m), f});
		}
		wg.assign(w);
		neurs.assign(nh);
		slots.assign(sh);
		nwg = unsigned(w.size());
	}

	bool per_window() const override { return Kind == update_kind::stateless; }

	void launch(spice_block const& b, hipStream_t s) override {
		if (nwg == 0) return;
		group_args<Neur> const a{wg.p, neurs.p, slots.p};
		if constexpr (Kind == update_kind::lanes)
			hipLaunchKernelGGL((update_lanes_kernel<Neur, Syn, MaxIn, Mixed>), dim3(nwg), dim3(256), 0, s, b, a);
		else if constexpr (Kind == update_kind::stateless)
			hipLaunchKernelGGL((update_stateless_kernel<Neur>), dim3(nwg, unsigned(b.nsteps)), dim3(256), 0, s, b,
			                   a);
		else
			hipLaunchKernelGGL((update_chunks_kernel<Neur, Syn, MaxIn, Mixed>), dim3(nwg), dim3(256), 0, s, b, a);
		check_hip(hipGetLastError(), "neuron update launch");
	}
};

template <Neuron Neur>
class neuron_population : public NeuronPopulation {
	static_assert(!PerPopulationUpdate<Neur>,
	              "populations with a per-population update() are not supported by the device engine");

public:
	using neuron_t = neuron_traits_t<Neur>;
	static constexpr bool stateful = StatefulNeuron<Neur>;

	neuron_population(engine* eng, int64_t inst, Neur neuron, Int const size, util::seed_seq& seed, float dt) :
	_eng(eng), _inst(inst), _neuron(std::move(neuron)), _size(size) {
		SPICE_INV(size >= 0);
		static_assert(std::is_trivially_copyable_v<Neur>, "neuron functors must be trivially copyable");
		int32_t words = 0;
		if constexpr (stateful) {
			static_assert(std::is_trivially_copyable_v<neuron_t>,
			              "neuron state must be trivially copyable to live in device memory");
			words = int32_t((sizeof(neuron_t) + 3) / 4);
		}
		_draws = probe_draws(dt);
		check_spice(spice_pop_create_in(eng->ctx, inst, size, words, _draws, &_id));
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
		set_group<void>();
	}

	Int size() const override { return _size; }
	int64_t id() const override { return _id; }
	int64_t instance() const override { return _inst; }
	Neur const& functor() const { return _neuron; }
	replay_slot const* slots_device() const { return _slots_dev.p; }

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

	// The reference keeps this population's spike history once it is the source
	// of a plastic projection (snn.h:46-47); the runtime keeps it on the device.
	void plastic() override { _plastic = true; }
	std::span<UInt const> history() const override { return {}; }

	// Registers an incoming projection (connect order).  The first synapse type
	// into a population is inlined into its update kernel; other types are
	// replayed through a device function pointer.
	template <class Syn>
	void add_incoming(Syn const& syn) {
		void* obj = nullptr;  // whole 16-byte units: kernels read it word by word
		check_hip(hipMalloc(&obj, (sizeof(Syn) + 15) / 16 * 16), "hipMalloc");
		check_hip(hipMemcpy(obj, &syn, sizeof(Syn), hipMemcpyHostToDevice), "hipMemcpy");
		_syn_objs.emplace_back(obj, [](void* p) { (void)hipFree(p); });
		void* d_fn = nullptr;  // the thunks' device addresses
		check_hip(hipMalloc(&d_fn, sizeof(replay_fn) + sizeof(events_fn)), "hipMalloc");
		auto* d_rf = static_cast<replay_fn*>(d_fn);
		auto* d_ef = reinterpret_cast<events_fn*>(d_rf + 1);
		hipLaunchKernelGGL((replay_address_kernel<Syn, neuron_t>), dim3(1), dim3(1), 0, 0, d_rf, d_ef);
		check_hip(hipGetLastError(), "replay address");
		replay_fn fn  = nullptr;
		events_fn efn = nullptr;
		check_hip(hipMemcpy(&fn, d_rf, sizeof(fn), hipMemcpyDeviceToHost), "hipMemcpy");
		check_hip(hipMemcpy(&efn, d_ef, sizeof(efn), hipMemcpyDeviceToHost), "hipMemcpy");
		(void)hipFree(d_fn);
		_in.push_back({std::type_index(typeid(Syn)), obj, fn, efn});
		if (_in.size() == 1) {
			_primary  = std::type_index(typeid(Syn));
			_reselect = [this] { set_group<Syn>(); };
		}
		_reselect();
		_slots_dirty = true;
		_eng->regroup = true;
	}

	void prepare() override {
		if constexpr (stateful) {
			if (_host_dirty) {
				check_spice(spice_pop_upload(_eng->ctx, _id, _host.data(), int32_t(sizeof(neuron_t))));
				_host_dirty = false;
			}
			_host_valid = false;  // state changes on the device from here on
		}
		if (_slots_dirty) {
			std::vector<replay_slot> h;
			for (auto& in : _in) {
				bool const own = in.type == _primary;
				h.push_back({in.obj, own ? nullptr : in.fn, own ? nullptr : in.ev_fn});
			}
			h.resize(std::max<size_t>(h.size(), 1), replay_slot{nullptr, nullptr, nullptr});
			_slots_dev.assign(h);
			_slots_dirty  = false;
			_eng->regroup = true;
		}
	}

	std::type_index group_key() const override { return _group_key; }
	std::unique_ptr<group_base> make_group() const override { return _make_group(); }

	// Applies the deliveries already counted for the instance's next step.
	void apply_pending() {
		if constexpr (stateful) {
			if (_in.empty() || _size == 0) return;
			prepare();
			spice_pop_view const* views = nullptr;
			check_spice(spice_ctx_views(_eng->ctx, &views));
			_apply(views, spice_instance_steps(_eng->ctx, _inst));
		}
	}

private:
	engine* _eng;
	int64_t _inst;
	Neur _neuron;
	Int _size;
	int64_t _id    = -1;
	int32_t _draws = 0;
	std::vector<std::conditional_t<stateful, neuron_t, char>> _host;
	bool _host_valid = false, _host_dirty = false;
	mutable std::vector<Int32> _spike_buf;
	bool _plastic = false;
	std::vector<incoming> _in;
	std::vector<std::unique_ptr<void, void (*)(void*)>> _syn_objs;
	std::type_index _primary   = std::type_index(typeid(void));
	std::type_index _group_key = std::type_index(typeid(void));
	std::function<std::unique_ptr<group_base>()> _make_group;
	std::function<void(spice_pop_view const*, int64_t)> _apply;
	std::function<void()> _reselect;
	dev_array<replay_slot> _slots_dev;
	bool _slots_dirty = true;

	// Picks the update kernel: its shape, the number of incoming projections it
	// unrolls (4 or SPICE_MAX_IN) and whether some of them have a synapse type
	// other than the inlined one.
	template <class Syn>
	void set_group() {
		bool mixed = false;
		for (auto& in : _in) mixed = mixed || in.type != _primary;
		bool const small = _in.size() <= 4;
		auto pick_kind   = [&]<int MaxIn, bool Mixed>() {
			auto pick = [&]<update_kind K>() {
				_group_key  = std::type_index(typeid(group<Neur, Syn, K, MaxIn, Mixed>));
				_make_group = [] { return std::unique_ptr<group_base>(new group<Neur, Syn, K, MaxIn, Mixed>()); };
			};
			if (_draws == 0)
				pick.template operator()<update_kind::lanes>();
			else if constexpr (!stateful)
				pick.template operator()<update_kind::stateless>();
			else
				pick.template operator()<update_kind::chunks>();
			_apply = [this](spice_pop_view const* views, int64_t local) {
				(void)this;
				if constexpr (stateful && !std::is_void_v<Syn>) {
					unsigned const grid = unsigned((_size + 255) / 256);
					hipLaunchKernelGGL((apply_pending_kernel<Neur, Syn, MaxIn, Mixed>), dim3(grid), dim3(256), 0,
					                   _eng->stream(), views, _id, local, _slots_dev.p);
					check_hip(hipGetLastError(), "apply pending deliveries");
				}
			};
		};
		if (small && !mixed)
			pick_kind.template operator()<4, false>();
		else if (small)
			pick_kind.template operator()<4, true>();
		else if (!mixed)
			pick_kind.template operator()<SPICE_MAX_IN, false>();
		else
			pick_kind.template operator()<SPICE_MAX_IN, true>();
	}

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
};
}  // namespace spice::detail
