// Projections between populations (API of reference detail/synapse_population.h).
// The CSR and the per-step event counts live in libspice_hip; the synapse
// functor is replayed by the target population's update kernel.
#pragma once

#include <span>
#include <type_traits>
#include <vector>

#include "spice/concepts.h"
#include "spice/detail/neuron_population.h"
#include "spice/topology.h"
#include "spice/util/assert.h"
#include "spice/util/random.h"
#include "spice/util/stdint.h"
#include "spice_hip.h"

namespace spice::detail {
struct SynapsePopulation {
	virtual ~SynapsePopulation() = default;
	virtual Int delay() const    = 0;
	virtual int64_t id() const   = 0;
};

template <class Syn, Neuron SrcNeur, StatefulNeuron DstNeur>
requires Synapse<Syn, SrcNeur, DstNeur>
class synapse_population : public SynapsePopulation {
	static_assert(!StatefulSynapse<Syn>,
	              "synapses with per-synapse state are not supported by the device engine yet");
	static_assert(DeliverTo<Syn, DstNeur>,
	              "deliver(source, target) synapses are not supported by the device engine yet");
	static_assert(std::is_trivially_copyable_v<Syn>, "synapse functors must be trivially copyable");

public:
	// Generates the connectivity (fixed_probability on the device, other
	// topologies on the host) with `seed` and registers the projection with
	// the target population (connect order = replay order).
	synapse_population(engine* eng, Syn syn, Topology& c, util::seed_seq const& seed, Int const delay,
	                   neuron_population<SrcNeur>* src, neuron_population<DstNeur>* dst) :
	_syn(std::move(syn)), _delay(delay) {
		SPICE_PRE(delay >= 1);
		UInt128 const s = seed.seed();
		if (auto* fp = dynamic_cast<fixed_probability*>(&c)) {
			check_spice(spice_proj_create_fixed_probability(eng->ctx, src->id(), dst->id(), fp->probability(),
			                                                s.lo, s.hi, delay, &_id));
		} else {
			std::vector<Int> offsets(size_t(c.src_count + 1), 0);
			std::vector<Int32> neighbors(size_t(std::max<Int>(c.size(), 1)), 0);
			c.generate(std::span<Int>(offsets), std::span<Int32>(neighbors.data(), size_t(c.size())), seed);
			check_spice(spice_proj_create_csr(eng->ctx, src->id(), dst->id(), offsets.data(),
			                                  neighbors.data(), delay, &_id));
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
};
}  // namespace spice::detail
