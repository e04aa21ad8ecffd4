// The model contract (API of reference concepts.h): which member functions make
// a type a Neuron or a Synapse, and static_assert diagnostics for mistakes.
#pragma once

#include <concepts>
#include <random>
#include <span>
#include <vector>

#include "spice/util/stdint.h"
#include "spice/util/type_traits.h"

namespace spice {
template <class T>
concept StatelessNeuron = std::default_initializable<T>;

template <class T>
concept StatefulNeuron = std::default_initializable<T> && requires {
	typename T::neuron;
	requires std::default_initializable<typename T::neuron>;
};

template <class T>
concept PerNeuronInit = StatefulNeuron<T> && requires(T const t, typename T::neuron& n, Int id,
                                                      std::mt19937& rng) { t.init(n, id, rng); };

template <class T>
concept PerPopulationInit =
    StatefulNeuron<T> &&
    requires(T t, std::span<typename T::neuron> n, std::mt19937& rng) { t.init(n, rng); };

namespace detail {
template <class T, bool = StatefulNeuron<T>>
struct per_neuron_update_check {
	static constexpr bool value = requires(T const t, float dt, std::mt19937& rng) {
		{ t.update(dt, rng) } -> std::same_as<bool>;
	};
};
template <class T>
struct per_neuron_update_check<T, true> {
	static constexpr bool value =
	    requires(T const t, typename T::neuron& n, float dt, std::mt19937& rng) {
		    { t.update(n, dt, rng) } -> std::same_as<bool>;
	    };
};
}  // namespace detail

template <class T>
concept PerNeuronUpdate = detail::per_neuron_update_check<T>::value;

template <class T>
concept PerPopulationUpdate =
    std::default_initializable<T> && (std::copy_constructible<T> || std::move_constructible<T>) &&
    requires(T t, float dt, std::mt19937& rng, std::vector<Int32>& out) { t.update(dt, rng, out); };

template <class T>
concept Neuron = (PerPopulationUpdate<T>
                      ? util::none_of<StatefulNeuron<T>, PerNeuronUpdate<T>, PerNeuronInit<T>,
                                      PerPopulationInit<T>>
                      : ((StatelessNeuron<T> || StatefulNeuron<T>) && PerNeuronUpdate<T> &&
                         util::up_to_one_of<PerNeuronInit<T>, PerPopulationInit<T>>));

template <class T>
concept StatelessSynapse = std::default_initializable<T>;

template <class T>
concept StatefulSynapse = std::default_initializable<T> && requires {
	typename T::synapse;
	requires std::default_initializable<typename T::synapse>;
};

namespace detail {
template <class T, class Neur, bool = StatefulSynapse<T>>
struct deliver_to_check {
	static constexpr bool value = requires(T const t, typename Neur::neuron& n) { t.deliver(n); };
};
template <class T, class Neur>
struct deliver_to_check<T, Neur, true> {
	static constexpr bool value = requires(T const t, typename T::synapse const& s,
	                                       typename Neur::neuron& n) { t.deliver(s, n); };
};
template <class T, class Src, class Dst, bool = StatefulSynapse<T>>
struct deliver_from_to_check {
	static constexpr bool value = requires(T const t, typename Src::neuron const& a,
	                                       typename Dst::neuron& b) { t.deliver(a, b); };
};
template <class T, class Src, class Dst>
struct deliver_from_to_check<T, Src, Dst, true> {
	static constexpr bool value =
	    requires(T const t, typename T::synapse const& s, typename Src::neuron const& a,
	             typename Dst::neuron& b) { t.deliver(s, a, b); };
};
}  // namespace detail

template <class T, class Neur>
concept DeliverTo = StatefulNeuron<Neur> && detail::deliver_to_check<T, Neur>::value;

template <class T, class SrcNeur, class DstNeur>
concept DeliverFromTo = StatefulNeuron<SrcNeur> && StatefulNeuron<DstNeur> &&
                        detail::deliver_from_to_check<T, SrcNeur, DstNeur>::value;

template <class T>
concept PlasticSynapse = StatefulSynapse<T> && requires(T const t, typename T::synapse& syn,
                                                        float dt, bool pre, bool post, Int n) {
	t.update(syn, dt, pre, post);
	t.skip(syn, dt, n);
};

template <class T>
concept PerSynapseInit = StatefulSynapse<T> && requires(T const t, typename T::synapse& syn,
                                                        Int src, Int dst, std::mt19937& rng) {
	t.init(syn, src, dst, rng);
};

template <class T, class SrcNeur, class DstNeur>
concept Synapse = (StatelessSynapse<T> || StatefulSynapse<T> || PlasticSynapse<T>) &&
                  util::one_of<DeliverTo<T, DstNeur>, DeliverFromTo<T, SrcNeur, DstNeur>>;

namespace detail {
template <class T>
struct neuron_traits {
	using type = void;
};
template <StatefulNeuron T>
struct neuron_traits<T> {
	using type = typename T::neuron;
};
template <class T>
using neuron_traits_t = typename neuron_traits<T>::type;

template <class T>
struct synapse_traits {
	using type = void;
};
template <StatefulSynapse T>
struct synapse_traits<T> {
	using type = typename T::synapse;
};
template <class T>
using synapse_traits_t = typename synapse_traits<T>::type;

// Signature probes with wildcard arguments, for the diagnostics below.
template <class T>
constexpr bool HasUpdate(auto... args) {
	return requires(T t) { t.update(args...); };
}
template <class T>
constexpr bool HasInit(auto... args) {
	return requires(T t) { t.init(args...); };
}
template <class T>
constexpr bool HasDeliver(auto... args) {
	return requires(T t) { t.deliver(args...); };
}
template <class T>
constexpr bool HasSkip(auto... args) {
	return requires(T t) { t.skip(args...); };
}

struct any_neuron_t {
	using neuron = util::any_t;
};
}  // namespace detail

template <class T>
constexpr bool CheckNeuron() {
	util::any_t any;

	static_assert(StatelessNeuron<T>,
	              "Every neuron must at least conform to the StatelessNeuron concept.");

	constexpr bool has_update =
	    detail::HasUpdate<T>(any, any) || detail::HasUpdate<T>(any, any, any);
	static_assert(has_update, "Every neuron must define an update() method.");
	static_assert(!has_update || (PerNeuronUpdate<T> || PerPopulationUpdate<T>),
	              "Your update() method has the wrong signautre.");
	static_assert(
	    !PerPopulationUpdate<T> || util::none_of<StatefulNeuron<T>, PerNeuronUpdate<T>,
	                                             PerNeuronInit<T>, PerPopulationInit<T>>,
	    "Defining a per-population update() method prohibits you from defining any of the following: "
	    "per-neuron update(), per-neuron init(), per-population init(), making your neuron stateful.");

	constexpr bool has_init = detail::HasInit<T>(any, any) || detail::HasInit<T>(any, any, any);
	static_assert(!has_init || StatefulNeuron<T>,
	              "You defined an init() method but your neuron has no state.");
	static_assert(!has_init || (PerNeuronInit<T> || PerPopulationInit<T>),
	              "Your neuron's init() method has the wrong signature.");
	static_assert(util::up_to_one_of<PerNeuronInit<T>, PerPopulationInit<T>>,
	              "Your neuron must define at most 1 init() method.");
	return true;
}

template <class T>
constexpr bool CheckSynapse() {
	util::any_t any;

	static_assert(StatelessSynapse<T>,
	              "Every synapse must at least conform to the StatelessSynapse concept.");

	constexpr bool has_deliver = detail::HasDeliver<T>(any) || detail::HasDeliver<T>(any, any) ||
	                             detail::HasDeliver<T>(any, any, any);
	static_assert(has_deliver, "Every synapse must define a deliver() method.");
	static_assert(!has_deliver || (DeliverTo<T, detail::any_neuron_t> ||
	                               DeliverFromTo<T, detail::any_neuron_t, detail::any_neuron_t>),
	              "Your deliver() method has the wrong signature.");

	constexpr bool has_update = detail::HasUpdate<T>(any, any, any, any);
	constexpr bool has_skip   = detail::HasSkip<T>(any, any, any);
	static_assert(!has_update || StatefulSynapse<T>,
	              "You defined an update() method but your synapse has no state.");
	static_assert(!has_update || has_skip,
	              "Your synapse defines an update() method but no skip() method.");
	static_assert(!has_skip || has_update,
	              "Your synapse defines a skip() method but no update() method.");
	static_assert(!has_update || PlasticSynapse<T>,
	              "Your synapse defines an update() method, "
	              "suggesting you intend to write a plastic synapse. "
	              "But your synapse does not conform to the PlasticSynapse concept. "
	              "Probably your update() or skip() method have the wrong signature.");

	constexpr bool has_init = detail::HasInit<T>(any, any, any, any);
	static_assert(!has_init || StatefulSynapse<T>,
	              "You defined an init() method but your synapse has no state.");
	static_assert(!has_init || PerSynapseInit<T>,
	              "Your synapse's init() method has the wrong signature.");
	return true;
}
}  // namespace spice
