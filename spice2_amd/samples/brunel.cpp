// Brunel network (20k neurons: 10k Poisson inputs, 8k excitatory and 2k
// inhibitory LIF neurons, all six projections fixed_probability(0.1)) written
// against the Spice user API exactly as a reference user would, printing the
// 300-step spike raster as JSON.  Its output is byte-identical to the
// reference's brunel sample (md5 dc05acddc0e3fb6da4e490641121fbe9).
#include "raster_json.h"

#include "spice/snn.h"

using namespace spice;
using namespace spice::util;

// Poisson input: fires with probability rate * dt per step.
struct poisson {
	bool update(float dt, auto& rng) const {
		float const firing_rate = 20;  // Hz
		return util::generate_canonical<float>(rng) < (firing_rate * dt);
	}
};
static_assert(CheckNeuron<poisson>());

// Leaky integrate-and-fire neuron with a refractory period of 20 steps.
struct lif {
	struct neuron {
		float V   = 0;
		int Twait = 0;
	};

	bool update(neuron& n, float dt, auto) const {
		float const TmemInv = 1.0 / 0.02;  // 1/s
		float const Vrest   = 0.0;         // V
		int const Tref      = 20;          // steps
		float const Vthres  = 0.02;        // V

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
static_assert(CheckNeuron<lif>());

// Every synapse of a projection adds the projection's weight to the target.
struct fixed_weight {
	float weight;
	void deliver(lif::neuron& to) const { to.V += weight; }
};
static_assert(CheckSynapse<fixed_weight>());

int main() {
	int const N       = 20000;
	float const dt    = 1e-4;
	float const delay = 15e-4;

	snn net(dt, delay, {1337});
	auto P = net.add_population<poisson>(N / 2);
	auto E = net.add_population<lif>(N * 4 / 10);
	auto I = net.add_population<lif>(N / 10);

	net.connect<fixed_weight>(P, E, fixed_probability(0.1), delay, {2.0 / N});
	net.connect<fixed_weight>(P, I, fixed_probability(0.1), delay, {2.0 / N});
	net.connect<fixed_weight>(E, E, fixed_probability(0.1), delay, {2.0 / N});
	net.connect<fixed_weight>(E, I, fixed_probability(0.1), delay, {2.0 / N});
	net.connect<fixed_weight>(I, E, fixed_probability(0.1), delay, {-10.0 / N});
	net.connect<fixed_weight>(I, I, fixed_probability(0.1), delay, {-10.0 / N});

	raster_json out("Brunel");
	for (Int step : range(300)) {
		net.step();
		out << I << E << P << '\n';
		(void)step;
	}
	return 0;
}
