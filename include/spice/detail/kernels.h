// Model-specific device code: the fused neuron-update kernel a population of a
// user neuron type runs once per block of steps.  Instantiated in the user's
// translation unit (the user's neuron and synapse functors are inlined).
//
// Per neuron and step of the block, in the reference's order (snn.cpp:7-28):
//   1. replay the deliveries of the previous step: for every incoming
//      projection in connect order, deliver() applied k times, k = that
//      projection's event count for this neuron and step (synapse_population.h:88-134
//      applies the same deliver() once per event; for a synapse without
//      per-synapse state every event of a projection is the same function);
//   2. update() with the step's dt and the neuron's position in the step's RNG
//      stream (neuron_population.h:40-45,72-77);
//   3. a wavefront ballot of the spike flags into the step's spike-bit words.
// Neuron state stays in registers for the whole block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "spice/concepts.h"
#include "spice/util/stdint.h"
#include "spice_hip.h"

namespace spice::detail {

// deliver() k times through a type-erased pointer (targets with incoming
// projections of several synapse types).
using replay_fn = void (*)(void const* syn, void* neuron, uint32_t k);

struct replay_slot {
	void const* obj;  // device copy of the synapse functor
	replay_fn fn;     // nullptr: the kernel's own synapse type, inlined
};

// The generator a neuron's update() receives: xoroshiro128+ positioned at the
// neuron's offset in the step's stream.  It counts its own outputs so that the
// kernel can check the per-neuron draw count the stream layout assumes.
struct device_stream {
	using result_type = UInt;
	UInt s0, s1;
	int count;

	__host__ __device__ static constexpr UInt min() { return 0; }
	__host__ __device__ static constexpr UInt max() { return ~UInt(0); }
	__host__ __device__ UInt operator()() {
		UInt const out = s0 + s1;
		UInt const t   = s0 ^ s1;
		s0             = ((s0 << 24) | (s0 >> 40)) ^ t ^ (t << 16);
		s1             = (t << 37) | (t >> 27);
		count++;
		return out;
	}
};

// Host-side counterpart used to probe how many outputs update() takes.
struct probe_stream {
	using result_type = UInt;
	UInt s0 = 0x9E3779B97F4A7C15ull, s1 = 0xD1B54A32D192ED03ull;
	int count = 0;
	static constexpr UInt min() { return 0; }
	static constexpr UInt max() { return ~UInt(0); }
	UInt operator()() {
		UInt const out = s0 + s1;
		UInt const t   = s0 ^ s1;
		s0             = ((s0 << 24) | (s0 >> 40)) ^ t ^ (t << 16);
		s1             = (t << 37) | (t >> 27);
		count++;
		return out;
	}
};

__device__ inline void jump_stream(device_stream& r, spice_rng_state c) {
	UInt a0 = 0, a1 = 0;
	for (int i = 0; i < 128; i++) {
		UInt const bit = (i < 64 ? (c.s0 >> i) : (c.s1 >> (i - 64))) & 1u;
		UInt const m   = UInt(0) - bit;
		a0 ^= r.s0 & m;
		a1 ^= r.s1 & m;
		r();
	}
	r.s0    = a0;
	r.s1    = a1;
	r.count = 0;
}

template <class Syn, class N>
__device__ void replay_thunk(void const* syn, void* neuron, uint32_t k) {
	Syn const& s = *static_cast<Syn const*>(syn);
	N& n         = *static_cast<N*>(neuron);
	for (uint32_t i = 0; i < k; i++) s.deliver(n);
}

template <class Syn, class N>
__global__ void replay_address_kernel(replay_fn* out) {
	*out = &replay_thunk<Syn, N>;
}

template <class N>
__device__ inline void load_state(N& n, uint32_t const* state, int64_t size, int64_t i) {
	constexpr int W = int((sizeof(N) + 3) / 4);
	uint32_t w[W];
#pragma unroll
	for (int k = 0; k < W; k++) w[k] = state[k * size + i];
	__builtin_memcpy(&n, w, sizeof(N));
}

template <class N>
__device__ inline void store_state(N const& n, uint32_t* state, int64_t size, int64_t i) {
	constexpr int W = int((sizeof(N) + 3) / 4);
	uint32_t w[W] = {};
	__builtin_memcpy(w, &n, sizeof(N));
#pragma unroll
	for (int k = 0; k < W; k++) state[k * size + i] = w[k];
}

template <class Syn, class N>
__device__ inline void apply_deliveries(N& n, spice_pop_view const& v, replay_slot const* slots,
                                        int64_t step, int64_t i) {
	for (int c = 0; c < v.nin; c++) {
		uint32_t* cp     = v.in[c].counts + (step % v.in[c].slots) * v.size + i;
		uint32_t const k = *cp;
		if (k == 0) continue;
		if (v.clear_counts) *cp = 0;
		replay_slot const s = slots[c];
		if constexpr (!std::is_void_v<Syn>) {
			if (s.fn == nullptr) {
				Syn const& syn = *static_cast<Syn const*>(s.obj);
				for (uint32_t r = 0; r < k; r++) syn.deliver(n);
				continue;
			}
		}
		s.fn(s.obj, &n, k);
	}
}

template <class Neur>
__device__ inline bool call_update(Neur const& neur, typename neuron_traits<Neur>::type* n, float dt,
                                   device_stream& rng) {
	if constexpr (StatefulNeuron<Neur>)
		return neur.update(*n, dt, rng);
	else
		return neur.update(dt, rng);
}

// One thread per neuron (populations whose neurons take no draws).
template <class Neur, class Syn>
__global__ void __launch_bounds__(256)
update_lanes_kernel(spice_pop_view v, spice_block_view b, Neur neur, replay_slot const* slots) {
	using N            = typename neuron_traits<Neur>::type;
	int64_t const i    = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	bool const active  = i < v.size;
	int const lane     = threadIdx.x % 64;
	std::conditional_t<std::is_void_v<N>, char, N> n{};
	if constexpr (!std::is_void_v<N>)
		if (active) load_state(n, v.state, v.size, i);
	for (int u = 0; u < b.nsteps; u++) {
		bool spiked = false;
		if (active) {
			if constexpr (!std::is_void_v<N>) apply_deliveries<Syn>(n, v, slots, b.t0 + u, i);
			device_stream rng{v.rng[u].s0, v.rng[u].s1, 0};
			if constexpr (std::is_void_v<N>)
				spiked = call_update<Neur>(neur, nullptr, b.dt[u], rng);
			else
				spiked = call_update<Neur>(neur, &n, b.dt[u], rng);
			if (rng.count != 0) atomicExch(v.error, 1);
		}
		uint64_t const bits = __ballot(spiked);
		if (lane == 0 && (i >> 6) < v.bitmap_stride) v.bitmap[u * v.bitmap_stride + (i >> 6)] = bits;
	}
	if constexpr (!std::is_void_v<N>)
		if (active) store_state(n, v.state, v.size, i);
}

// Applies the deliveries counted for `step` (the next step to run) to the
// stored state, so that host reads see the state the reference has after the
// previous step; the counters are cleared, so the next update does not apply
// them again.
template <class Neur, class Syn>
__global__ void __launch_bounds__(256)
apply_pending_kernel(spice_pop_view v, int64_t step, replay_slot const* slots) {
	using N         = typename neuron_traits<Neur>::type;
	int64_t const i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= v.size) return;
	N n;
	load_state(n, v.state, v.size, i);
	apply_deliveries<Syn>(n, v, slots, step, i);
	store_state(n, v.state, v.size, i);
}

// One thread per chunk of 16 consecutive neurons (neurons take `draws` draws
// each; the thread jumps once to its chunk's stream position per step).
template <class Neur, class Syn>
__global__ void __launch_bounds__(256)
update_chunks_kernel(spice_pop_view v, spice_block_view b, Neur neur, replay_slot const* slots) {
	using N             = typename neuron_traits<Neur>::type;
	constexpr int C     = 16;
	int64_t const chunk = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	int64_t const i0    = chunk * C;
	bool const active   = i0 < v.size;
	int const lane      = threadIdx.x % 64;
	spice_rng_state const jp = active ? v.chunk_jump[chunk] : spice_rng_state{1, 0};
	for (int u = 0; u < b.nsteps; u++) {
		uint32_t bits = 0;
		if (active) {
			device_stream rng{v.rng[u].s0, v.rng[u].s1, 0};
			jump_stream(rng, jp);
			int64_t const end = i0 + C < v.size ? i0 + C : v.size;
			for (int64_t i = i0; i < end; i++) {
				bool spiked;
				if constexpr (std::is_void_v<N>) {
					spiked = call_update<Neur>(neur, nullptr, b.dt[u], rng);
				} else {
					N n;
					load_state(n, v.state, v.size, i);
					apply_deliveries<Syn>(n, v, slots, b.t0 + u, i);
					spiked = call_update<Neur>(neur, &n, b.dt[u], rng);
					store_state(n, v.state, v.size, i);
				}
				if (rng.count != int((i - i0 + 1) * v.draws)) atomicExch(v.error, 1);
				bits |= uint32_t(spiked) << (i - i0);
			}
		}
		uint64_t word = uint64_t(bits) << (16 * (lane % 4));
		word |= __shfl_xor(word, 1, 64);
		word |= __shfl_xor(word, 2, 64);
		if (active && lane % 4 == 0) v.bitmap[u * v.bitmap_stride + (i0 >> 6)] = word;
	}
}

}  // namespace spice::detail
