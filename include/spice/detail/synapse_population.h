// Projections between populations (API of reference detail/synapse_population.h).
// The CSR and the per-step event counts live in libspice_hip; the synapse
// functor is replayed by the target population's update kernel.
#pragma once

#include <algorithm>
#include <span>
#include <stdexcept>
#include <type_traits>
#include <typeindex>
#include <vector>

#include "spice/concepts.h"
#include "spice/detail/neuron_population.h"
#include "spice/topology.h"
#include "spice/util/assert.h"
#include "spice/util/random.h"
#include "spice/util/stdint.h"
#include "spice_hip.h"

namespace spice::detail {
template <class Syn>
struct plastic_group final : plastic_group_base {
	struct member {
		int64_t inst, proj, rows;
		Syn syn;
	};
	std::vector<member> members;
	dev_array<int64_t> ids;
	dev_array<Syn> funcs;  // whole 16-byte units: the kernels read functors word by word
	int64_t most_rows = 0;
	bool dirty        = true;

	void prepare() override {
		if (!dirty) return;
		std::vector<int64_t> h;
		std::vector<Syn> f;
		most_rows = 0;
		for (auto& m : members) {
			h.push_back(m.proj);
			f.push_back(m.syn);
			most_rows = std::max(most_rows, m.rows);
		}
		f.resize(f.size() + 16 / sizeof(Syn) + 1, members.empty() ? Syn{} : members.back().syn);
		ids.assign(h);
		funcs.assign(f);
		dirty = false;
	}
	void launch(spice_block const& b, hipStream_t st) override {
		if (members.empty()) return;
		unsigned const n     = unsigned(members.size());
		unsigned const sweep = unsigned(std::clamp<int64_t>((most_rows + 3) / 4, 1, 4096));
		hipLaunchKernelGGL((plastic_sweep_kernel<Syn>), dim3(sweep, n), dim3(256), 0, st, b, ids.p, funcs.p);
		hipLaunchKernelGGL((plastic_deliver_kernel<Syn>), dim3(64, n), dim3(256), 0, st, b, ids.p, funcs.p);
		check_hip(hipGetLastError(), "plastic synapse launch");
	}
	void remove_instance(int64_t inst) override {
		if (std::erase_if(members, [&](member const& m) { return m.inst == inst; })) dirty = true;
	}
};

template <class Syn>
plastic_group<Syn>& plastic_group_of(engine* eng) {
	auto const key = std::type_index(typeid(plastic_group<Syn>));
	for (auto& g : eng->plastic_groups)
		if (g.first == key) return static_cast<plastic_group<Syn>&>(*g.second);
	eng->plastic_groups.emplace_back(key, std::make_unique<plastic_group<Syn>>());
	return static_cast<plastic_group<Syn>&>(*eng->plastic_groups.back().second);
}

struct SynapsePopulation {
	virtual ~SynapsePopulation() = default;
	virtual Int delay() const    = 0;
	virtual int64_t id() const   = 0;
};

template <class Syn, Neuron SrcNeur, StatefulNeuron DstNeur>
requires Synapse<Syn, SrcNeur, DstNeur>
class synapse_population : public SynapsePopulation {
	static_assert(DeliverTo<Syn, DstNeur>,
	              "deliver(source, target) synapses are not supported by the device engine yet");
	static_assert(std::is_trivially_copyable_v<Syn>, "synapse functors must be trivially copyable");

public:
	// Generates the connectivity (fixed_probability on the device, other
	// topologies on the host) with the network's next seed, initialises the
	// per-synapse state (with one more seed for a per-synapse init(),
	// synapse_population.h:30-41) and registers the projection with the target
	// population (connect order = replay order).
	synapse_population(engine* eng, int64_t inst, Syn syn, Topology& c, util::seed_seq& seed, Int const delay,
	                   neuron_population<SrcNeur>* src, neuron_population<DstNeur>* dst) :
	_syn(std::move(syn)), _delay(delay) {
		SPICE_PRE(delay >= 1);
		util::seed_seq const gs = seed++;  // the topology's seed (synapse_population.h:31)
		UInt128 const s          = gs.seed();
		if (auto* fp = dynamic_cast<fixed_probability*>(&c)) {
			check_spice(spice_proj_create_fixed_probability(eng->ctx, src->id(), dst->id(), fp->probability(),
			                                                s.lo, s.hi, delay, &_id));
		} else {
			std::vector<Int> offsets(size_t(c.src_count + 1), 0);
			std::vector<Int32> neighbors(size_t(std::max<Int>(c.size(), 1)), 0);
			c.generate(std::span<Int>(offsets), std::span<Int32>(neighbors.data(), size_t(c.size())), gs);
			check_spice(spice_proj_create_csr(eng->ctx, src->id(), dst->id(), offsets.data(),
			                                  neighbors.data(), delay, &_id));
		}
		if constexpr (StatefulSynapse<Syn>) init_synapses(eng, c, seed);
		if constexpr (PlasticSynapse<Syn>) {
			check_spice(spice_proj_set_plastic(eng->ctx, _id));
			src->plastic();
			auto& g = plastic_group_of<Syn>(eng);
			g.members.push_back({inst, _id, src->size(), _syn});
			g.dirty = true;
		}
		dst->template add_incoming<Syn>(_syn);
		eng->min_delay = std::min(eng->min_delay, delay);
	}



	Int delay() const override { return _delay; }
	int64_t id() const override { return _id; }

private:
	Syn _syn;
	Int _delay;
	int64_t _id = -1;

	// Every synapse starts as a default synapse; a per-synapse init() runs in
	// CSR order with its own generator (synapse_population.h:34-40): on the
	// device when it takes no draws (the calls are then independent), else on
	// the host.
	void init_synapses(engine* eng, Topology& c, util::seed_seq& seed) {
		using S = typename Syn::synapse;
		static_assert(std::is_trivially_copyable_v<S>, "synapse state must be trivially copyable");
		if constexpr (PerSynapseInit<Syn>) {
			util::xoroshiro64_128p rng(seed++);
			if (init_draws() == 0) {
				S const one{};
				check_spice(spice_proj_set_synapses(eng->ctx, _id, &one, 1, int32_t(sizeof(S))));
				init_on_device(eng);
				return;
			}
			int64_t nnz = 0;
			check_spice(spice_proj_synapses(eng->ctx, _id, &nnz));
			std::vector<Int> off(size_t(c.src_count + 1));
			std::vector<Int32> nb(size_t(std::max<int64_t>(nnz, 1)));
			check_spice(spice_proj_download(eng->ctx, _id, off.data(), nb.data()));
			std::vector<S> syns(size_t(std::max<int64_t>(nnz, 1)));
			for (Int src = 0; src < c.src_count; src++)
				for (Int e = off[size_t(src)]; e < off[size_t(src) + 1]; e++)
					_syn.init(syns[size_t(e)], src, nb[size_t(e)], rng);
			check_spice(spice_proj_set_synapses(eng->ctx, _id, syns.data(), nnz, int32_t(sizeof(S))));
		} else {
			S const one{};
			check_spice(spice_proj_set_synapses(eng->ctx, _id, &one, 1, int32_t(sizeof(S))));
		}
	}

	// Generator outputs one init() call takes (probed on the host).
	int init_draws() const {
		typename Syn::synapse s{};
		probe_stream rng;
		_syn.init(s, 0, 0, rng);
		return rng.count;
	}

	void init_on_device(engine* eng) {
		spice_proj_view const* pv = nullptr;
		check_spice(spice_ctx_proj_views(eng->ctx, &pv));
		dev_array<Syn> f;
		f.assign(std::vector<Syn>{_syn});
		dev_array<int32_t> err;
		err.assign(std::vector<int32_t>{0});
		hipStream_t const s = eng->stream();
		hipLaunchKernelGGL(init_synapses_kernel<Syn>, dim3(1024), dim3(256), 0, s, pv + _id, f.p, err.p);
		check_hip(hipGetLastError(), "synapse init launch");
		int32_t flag = 0;
		check_hip(hipMemcpyAsync(&flag, err.p, sizeof(flag), hipMemcpyDeviceToHost, s), "synapse init");
		check_hip(hipStreamSynchronize(s), "synapse init");
		if (flag != 0)
			throw std::logic_error("a synapse's init() took generator draws on the device although the host "
			                       "probe took none; per-synapse draw counts must not depend on the synapse");
	}
};
}  // namespace spice::detail
