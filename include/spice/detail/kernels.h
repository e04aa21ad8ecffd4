// Model-specific device code: the fused neuron-update kernels of a group of
// populations of one user neuron type, launched once per block of steps.
// Instantiated in the user's translation unit (the user's neuron and synapse
// functors are inlined).  One launch covers every population of the group, in
// every instance of the context.
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
// deliver(synapse, neuron) for the events ev[0..k) of a projection with
// per-synapse state, in ascending synapse id (= the reference's source order).
using events_fn = void (*)(void const* syn, void* neuron, uint32_t const* ev, uint32_t k,
                           uint32_t const* state, int64_t nnz);

struct replay_slot {
	void const* obj;  // device copy of the synapse functor
	replay_fn fn;     // counted deliveries of another synapse type (else null)
	events_fn ev_fn;  // event deliveries of another synapse type (else null)
	// both null: the kernel's own synapse type, inlined
};

// One workgroup's share of a group launch: population `pop` (context id),
// group member `member`, elements first .. first + blockDim.x - 1 (neurons or
// chunks, by kernel).
struct group_entry {
	int32_t pop;
	int32_t member;
	int64_t first;
};

template <class Neur>
struct group_args {
	group_entry const* wg;              // per workgroup (blockIdx.x)
	Neur const* neurs;                  // per member: the population's functor
	replay_slot const* const* slots;    // per member: incoming projections, connect order
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

// r <- T^n r for the jump polynomial c = x^n mod P: T^n = sum_i c_i T^i.
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

template <class S>
__device__ inline S load_synapse(uint32_t const* state, int64_t nnz, int64_t e) {
	constexpr int W = int((sizeof(S) + 3) / 4);
	uint32_t w[W];
#pragma unroll
	for (int k = 0; k < W; k++) w[k] = state[k * nnz + e];
	S s;
	__builtin_memcpy(&s, w, sizeof(S));
	return s;
}

template <class S>
__device__ inline void store_synapse(uint32_t* state, int64_t nnz, int64_t e, S const& s) {
	constexpr int W = int((sizeof(S) + 3) / 4);
	uint32_t w[W] = {};
	__builtin_memcpy(w, &s, sizeof(S));
#pragma unroll
	for (int k = 0; k < W; k++) state[k * nnz + e] = w[k];
}

// The events of one target, applied by ascending synapse id: each round picks
// the smallest id above the last one (lists are short: about p times the
// source's spikes per step).
template <class Syn, class N>
__device__ inline void apply_events(N& n, Syn const& syn, uint32_t const* ev, uint32_t k, uint32_t const* state,
                                    int64_t nnz) {
	using S       = typename Syn::synapse;
	uint32_t last = 0;
	for (uint32_t j = 0; j < k; j++) {
		uint32_t best = 0xffffffffu;
		for (uint32_t q = 0; q < k; q++) {
			uint32_t const e = ev[q];
			if ((j == 0 || e > last) && e < best) best = e;
		}
		if (int64_t(best) >= nnz) break;  // lists always hold k valid ids; never read past the state
		last = best;
		syn.deliver(load_synapse<S>(state, nnz, best), n);
	}
}

// deliver() k times, exactly as k calls in a row: in blocks of 8 calls, each on
// a copy that is kept only for the first k (a select, not a branch), so a wave
// loops over blocks of its lanes' largest count instead of over single events.
// A stateless synapse's deliver() changes nothing but the neuron passed to it.
template <class Syn, class N>
__device__ inline void deliver_times(N& n, Syn const& syn, uint32_t k) {
	for (uint32_t r = 0; r < k; r += 8) {
#pragma unroll
		for (uint32_t j = 0; j < 8; j++) {
			N t = n;
			syn.deliver(t);
			if (r + j < k) n = t;
		}
	}
}

template <class Syn, class N>
__device__ void replay_thunk(void const* syn, void* neuron, uint32_t k) {
	Syn const& s = *static_cast<Syn const*>(syn);
	N& n         = *static_cast<N*>(neuron);
	deliver_times(n, s, k);
}

template <class Syn, class N>
__device__ void events_thunk(void const* syn, void* neuron, uint32_t const* ev, uint32_t k, uint32_t const* state,
                             int64_t nnz) {
	apply_events<Syn>(*static_cast<N*>(neuron), *static_cast<Syn const*>(syn), ev, k, state, nnz);
}

// Device addresses of the thunks of a synapse type (one of them is null).
template <class Syn, class N>
__global__ void replay_address_kernel(replay_fn* out, events_fn* out_ev) {
	if constexpr (StatefulSynapse<Syn>) {
		*out    = nullptr;
		*out_ev = &events_thunk<Syn, N>;
	} else {
		*out    = &replay_thunk<Syn, N>;
		*out_ev = nullptr;
	}
}

// Accesses through global-address-space pointers.  Generic (flat) accesses also
// count against the counter that scalar and LDS loads wait on, so each such wait
// would drain every load in flight (e.g. the prefetched event counts).
__device__ inline uint32_t gload32(uint32_t const* p) {
	using gptr = __attribute__((address_space(1))) uint32_t const*;
	return *(gptr)(p);
}
__device__ inline void gstore32(uint32_t* p, uint32_t v) {
	using gptr = __attribute__((address_space(1))) uint32_t*;
	*(gptr)(p) = v;
}
__device__ inline void gstore64(uint64_t* p, uint64_t v) {
	using gptr = __attribute__((address_space(1))) uint64_t*;
	*(gptr)(p) = v;
}

__device__ inline void gstore16(uint16_t* p, uint16_t v) {
	using gptr = __attribute__((address_space(1))) uint16_t*;
	*(gptr)(p) = v;
}

template <class N>
__device__ inline void load_state(N& n, uint32_t const* state, int64_t size, int64_t i) {
	constexpr int W = int((sizeof(N) + 3) / 4);
	uint32_t w[W];
#pragma unroll
	for (int k = 0; k < W; k++) w[k] = gload32(state + k * size + i);
	__builtin_memcpy(&n, w, sizeof(N));
}

template <class N>
__device__ inline void store_state(N const& n, uint32_t* state, int64_t size, int64_t i) {
	constexpr int W = int((sizeof(N) + 3) / 4);
	uint32_t w[W] = {};
	__builtin_memcpy(w, &n, sizeof(N));
#pragma unroll
	for (int k = 0; k < W; k++) gstore32(state + k * size + i, w[k]);
}

// The step table is written only by host copies between launches, so it is read
// through the constant address space (scalar loads: the address is uniform).
__device__ inline spice_step step_of(spice_block const& b, int32_t inst, int u);

// Event counts of neuron i for application step `local`, all incoming
// projections at once (independent loads, issued together).  MaxIn bounds the
// population's incoming projections (a compile-time size keeps k in registers).
__device__ inline int64_t ring_slot(int64_t local, int64_t slots) {
	if ((uint64_t(local) | uint64_t(slots)) < (uint64_t(1) << 32)) return int64_t(uint32_t(local) % uint32_t(slots));
	return local % slots;
}

template <int MaxIn>
__device__ inline void load_counts(uint32_t (&k)[MaxIn], spice_pop_view const& v, int64_t local, int64_t i,
                                   bool clear) {
#pragma unroll
	for (int c = 0; c < MaxIn; c++) {
		k[c] = 0;
		if (c < v.nin) {
			uint32_t* cp = v.in[c].counts + ring_slot(local, v.in[c].slots) * v.size + i;
			k[c]         = gload32(cp);
			if (clear && k[c]) gstore32(cp, 0);
		}
	}
}

// Reads a trivially copyable object that kernels never write (synapse functors,
// replay slots) through the constant address space, word by word (device copies
// are allocated in whole 16-byte units).
template <class T>
__device__ inline T cload_obj(void const* p) {
	using cptr         = __attribute__((address_space(4))) uint32_t const*;
	constexpr int W    = int((sizeof(T) + 3) / 4);
	uint32_t w[W > 0 ? W : 1];
#pragma unroll
	for (int k = 0; k < W; k++) w[k] = ((cptr)(p))[k];
	T t;
	__builtin_memcpy(&t, w, sizeof(T));
	return t;
}

__device__ inline spice_step step_of(spice_block const& b, int32_t inst, int u) {
	spice_step const* steps = cload_obj<spice_window>(b.win).steps;
	return cload_obj<spice_step>(steps + int64_t(inst) * SPICE_MAX_WINDOW + b.rel + u);
}

// The block's steps of one instance, loaded once by the first threads of the
// workgroup into LDS (a kernel that walks the steps then reads each from LDS
// instead of waiting on a table load per step).  Every thread must call it.
__device__ inline void stage_steps(spice_block const& b, int32_t inst, spice_step* sts) {
	if (int(threadIdx.x) < b.nsteps) sts[threadIdx.x] = step_of(b, inst, int(threadIdx.x));
	__syncthreads();
}

// The incoming projections of a population, loaded once per kernel: replay
// slots and, for the kernel's own synapse type, a register copy of the functor.
template <class Syn, int MaxIn>
struct synapse_set {
	replay_slot slot[MaxIn];
	struct empty {};
	std::conditional_t<std::is_void_v<Syn>, empty, Syn> syn[MaxIn];

	__device__ void load(replay_slot const* slots, int nin) {
#pragma unroll
		for (int c = 0; c < MaxIn; c++) {
			if (c < nin) {
				slot[c] = cload_obj<replay_slot>(slots + c);
				if constexpr (!std::is_void_v<Syn>)
					if (slot[c].fn == nullptr && slot[c].ev_fn == nullptr) syn[c] = cload_obj<Syn>(slot[c].obj);
			}
		}
	}
};

// deliver() k[c] times per incoming projection c, in connect order.  Without
// Mixed every projection has the kernel's own synapse type and is inlined; with
// Mixed the others are called through their device function pointer on a copy
// of the neuron, so that the neuron itself never has its address taken (it
// stays in registers).
template <class Syn, int MaxIn, bool Mixed, class N>
__device__ inline void replay(N& n, uint32_t const (&k)[MaxIn], spice_pop_view const& v, int64_t i,
                              synapse_set<Syn, MaxIn> const& ss) {
#pragma unroll
	for (int c = 0; c < MaxIn; c++) {
		if (c < v.nin && k[c] != 0) {
			replay_slot const& sl = ss.slot[c];
			bool done             = false;
			if constexpr (!std::is_void_v<Syn>) {
				if (!Mixed || (sl.fn == nullptr && sl.ev_fn == nullptr)) {
					Syn const& syn = ss.syn[c];
					if constexpr (StatefulSynapse<Syn>) {
						spice_in_view const& in = v.in[c];
						apply_events<Syn>(n, syn, in.events + in.ev_offsets[i], k[c], in.syn_state, in.nnz);
					} else {
						deliver_times(n, syn, k[c]);
					}
					done = true;
				}
			}
			if constexpr (Mixed || std::is_void_v<Syn>) {
				if (!done) {
					N tmp = n;
					if (sl.ev_fn) {
						spice_in_view const& in = v.in[c];
						sl.ev_fn(sl.obj, &tmp, in.events + in.ev_offsets[i], k[c], in.syn_state, in.nnz);
					} else {
						sl.fn(sl.obj, &tmp, k[c]);
					}
					n = tmp;
				}
			}
		}
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

// One thread per neuron (populations whose neurons take no draws).  Every count
// a block's steps read was delivered before the block began (blocks are no
// longer than the shortest delay), so the counts of step u + A are loaded while
// step u is computed, A = counts_ahead steps ahead.  Every wave issues the same
// memory operations in the same order whatever its lanes hold (loads past the
// block or past nin read valid addresses and are ignored; waves past the end of
// the population store their spike word to the row's scratch word), so that a
// step waits only for its own counts, not for the loads issued after them.
template <int MaxIn>
constexpr int counts_ahead = MaxIn <= 4 ? 4 : 2;

template <class Neur, class Syn, int MaxIn, bool Mixed, bool Clear, class N>
__device__ inline void update_lanes_steps(spice_block const& b, spice_pop_view const& v, Neur const& neur,
                                          synapse_set<Syn, MaxIn> const& ss, spice_step const* sts, N& n,
                                          int64_t i, bool active) {
	constexpr int A   = counts_ahead<MaxIn>;
	constexpr bool Nv = std::is_void_v<typename neuron_traits<Neur>::type>;
	// The block's steps are consecutive: the count slot of step u of projection
	// c is slot0[c] + u, wrapped once (rings are longer than a block plus A).
	uint32_t* crow[MaxIn];
	int64_t slot0[MaxIn], nslot[MaxIn];
	int64_t const local0 = sts[0].local;
#pragma unroll
	for (int c = 0; c < MaxIn; c++) {
		int const cc = c < v.nin ? c : 0;
		if (v.nin > 0) {
			crow[c]  = v.in[cc].counts + (active ? i : 0);
			nslot[c] = v.in[cc].slots;
			slot0[c] = local0 >= 0 ? ring_slot(local0, nslot[c]) : 0;
		} else {  // nothing to count: any valid address
			crow[c]  = v.state;
			nslot[c] = 1;
			slot0[c] = 0;
		}
	}
	auto counts_of = [&](int u, int c) {
		int64_t s = slot0[c] + u;
		if (s >= nslot[c]) s -= nslot[c];
		return crow[c] + (v.nin > 0 ? s * v.size : 0);
	};
	uint32_t kb[A][MaxIn] = {};  // counts of steps u0 + q, q < A
	if constexpr (!Nv) {
#pragma unroll
		for (int q = 0; q < A; q++)
#pragma unroll
			for (int c = 0; c < MaxIn; c++) kb[q][c] = gload32(counts_of(q, c));
	}
	int64_t const word = (i >> 6) < v.bitmap_stride - 1 ? (i >> 6) : v.bitmap_stride - 1;
	bool done = false;  // an instance's steps in a window form a prefix
	for (int u0 = 0; u0 < b.nsteps && !done; u0 += A) {
#pragma unroll
		for (int q = 0; q < A; q++) {
			int const u = u0 + q;
			if (u >= b.nsteps || done) continue;
			spice_step const st = sts[u];
			if (st.local < 0) {
				done = true;
				continue;
			}
			uint32_t k[MaxIn];
#pragma unroll
			for (int c = 0; c < MaxIn; c++) k[c] = c < v.nin && active ? kb[q][c] : 0;
			if constexpr (!Nv) {
#pragma unroll
				for (int c = 0; c < MaxIn; c++) kb[q][c] = gload32(counts_of(u + A, c));
				if constexpr (Clear)  // push delivery accumulates: consumed counts restart at zero
#pragma unroll
					for (int c = 0; c < MaxIn; c++)
						if (c < v.nin && active) gstore32(counts_of(u, c), 0u);
			}
			bool spiked = false;
			if (active) {
				if constexpr (!Nv) replay<Syn, MaxIn, Mixed>(n, k, v, i, ss);
				device_stream rng{st.seed.s0, st.seed.s1, 0};
				if constexpr (Nv)
					spiked = call_update<Neur>(neur, nullptr, st.dt, rng);
				else
					spiked = call_update<Neur>(neur, &n, st.dt, rng);
				if (rng.count != 0) atomicExch(v.error, 1);
			}
			uint64_t const bits = __ballot(spiked);
			gstore64(v.bitmap + u * v.bitmap_stride + word, bits);
		}
	}
}

template <class Neur, class Syn, int MaxIn, bool Mixed>
__global__ void __launch_bounds__(256) update_lanes_kernel(spice_block b, group_args<Neur> g) {
	using N                 = typename neuron_traits<Neur>::type;
	group_entry const e     = g.wg[blockIdx.x];
	spice_pop_view const v  = b.pops[e.pop];  // a local copy: stores cannot alias it
	Neur const neur         = g.neurs[e.member];
	synapse_set<Syn, MaxIn> ss;
	if constexpr (!std::is_void_v<typename neuron_traits<Neur>::type>) ss.load(g.slots[e.member], v.nin);
	int64_t const i         = e.first + threadIdx.x;
	bool const active       = i < v.size;
	std::conditional_t<std::is_void_v<N>, char, N> n{};
	if constexpr (!std::is_void_v<N>)
		if (active) load_state(n, v.state, v.size, i);
	__shared__ spice_step sts[SPICE_MAX_BLOCK];
	stage_steps(b, v.inst, sts);
	if (v.clear_counts != 0)
		update_lanes_steps<Neur, Syn, MaxIn, Mixed, true>(b, v, neur, ss, sts, n, i, active);
	else
		update_lanes_steps<Neur, Syn, MaxIn, Mixed, false>(b, v, neur, ss, sts, n, i, active);
	if constexpr (!std::is_void_v<N>)
		if (active) store_state(n, v.state, v.size, i);
}

// Stateless neurons that draw: their spikes depend only on the step's stream,
// so every (16-neuron chunk, step) pair is independent.  One thread per pair
// jumps to the chunk's stream position and writes the chunk's 16 spike bits.
// A jump is the XOR of the step stream's states at the set bits of the
// chunk's jump polynomial; the workgroup first builds, in LDS, the XOR of each
// 4-bit group of states for all 16 bit patterns (32 x 16 entries from the
// window's stored states), so a jump is 32 table reads.  blockIdx.y = step.
template <class Neur>
__global__ void __launch_bounds__(256) update_stateless_kernel(spice_block b, group_args<Neur> g) {
	group_entry const e     = g.wg[blockIdx.x];
	spice_pop_view const v  = b.pops[e.pop];  // a local copy: stores cannot alias it
	int const u             = blockIdx.y;
	if (u >= b.nsteps) return;
	spice_step const st = step_of(b, v.inst, u);
	if (st.local < 0) return;
	__shared__ spice_rng_state tab[SPICE_STREAM_STATES / 4 * 16];
	{
		spice_rng_state const* S = cload_obj<spice_window>(b.win).streams +
		                           (int64_t(v.inst) * SPICE_MAX_WINDOW + b.rel + u) * SPICE_STREAM_STATES;
		for (int q = threadIdx.x; q < SPICE_STREAM_STATES / 4 * 16; q += blockDim.x) {
			int const j = q >> 4, x = q & 15;
			UInt a0 = 0, a1 = 0;
#pragma unroll
			for (int k = 0; k < 4; k++)
				if ((x >> k) & 1) {
					spice_rng_state const s = S[4 * j + k];
					a0 ^= s.s0;
					a1 ^= s.s1;
				}
			tab[q] = {a0, a1};
		}
	}
	__syncthreads();
	int64_t const c  = e.first + threadIdx.x;  // 16-neuron chunk; 4 chunks per spike word
	int64_t const i0 = c * 16;
	if (i0 >= (v.size + 63) / 64 * 64) return;
	int const n   = int(v.size - i0 < 16 ? (v.size - i0 > 0 ? v.size - i0 : 0) : 16);
	uint32_t bits = 0;
	if (n > 0) {
		spice_rng_state const J = cload_obj<spice_rng_state>(v.chunk_jump + c);
		UInt a0 = 0, a1 = 0;
#pragma unroll
		for (int j = 0; j < 32; j++) {
			int const x             = int(((j < 16 ? J.s0 : J.s1) >> (4 * (j & 15))) & 15);
			spice_rng_state const t = tab[j * 16 + x];
			a0 ^= t.s0;
			a1 ^= t.s1;
		}
		device_stream rng{a0, a1, 0};
		Neur const neur = g.neurs[e.member];
		if (n == 16) {
#pragma unroll
			for (int t = 0; t < 16; t++) bits |= uint32_t(call_update<Neur>(neur, nullptr, st.dt, rng)) << t;
		} else {
			for (int t = 0; t < n; t++) bits |= uint32_t(call_update<Neur>(neur, nullptr, st.dt, rng)) << t;
		}
		if (rng.count != n * v.draws) atomicExch(v.error, 1);
	}
	// one bitmap row per step of the window (launched once per window, rel 0)
	gstore16(reinterpret_cast<uint16_t*>(v.bitmap + (b.rel + u) * v.bitmap_stride) + c, uint16_t(bits));
}

// One thread per chunk of 16 consecutive neurons (stateful neurons that draw:
// each takes `draws` draws; the thread jumps once to its chunk's stream
// position per step).
template <class Neur, class Syn, int MaxIn, bool Mixed>
__global__ void __launch_bounds__(256) update_chunks_kernel(spice_block b, group_args<Neur> g) {
	using N                 = typename neuron_traits<Neur>::type;
	constexpr int C         = 16;
	group_entry const e     = g.wg[blockIdx.x];
	spice_pop_view const v  = b.pops[e.pop];  // a local copy: stores cannot alias it
	Neur const neur         = g.neurs[e.member];
	synapse_set<Syn, MaxIn> ss;
	if constexpr (!std::is_void_v<N>) ss.load(g.slots[e.member], v.nin);
	int64_t const chunk     = e.first + threadIdx.x;
	int64_t const i0        = chunk * C;
	bool const active       = i0 < v.size;
	int const lane          = threadIdx.x % 64;
	bool const clear        = v.clear_counts != 0;
	spice_rng_state const jp = active ? v.chunk_jump[chunk] : spice_rng_state{1, 0};
	__shared__ spice_step sts[SPICE_MAX_BLOCK];
	stage_steps(b, v.inst, sts);
	for (int u = 0; u < b.nsteps; u++) {
		spice_step const st = sts[u];
		if (st.local < 0) break;
		uint32_t bits = 0;
		if (active) {
			device_stream rng{st.seed.s0, st.seed.s1, 0};
			jump_stream(rng, jp);
			int64_t const end = i0 + C < v.size ? i0 + C : v.size;
			for (int64_t i = i0; i < end; i++) {
				bool spiked;
				if constexpr (std::is_void_v<N>) {
					spiked = call_update<Neur>(neur, nullptr, st.dt, rng);
				} else {
					N n;
					load_state(n, v.state, v.size, i);
					uint32_t k[MaxIn];
					load_counts<MaxIn>(k, v, st.local, i, clear);
					replay<Syn, MaxIn, Mixed>(n, k, v, i, ss);
					spiked = call_update<Neur>(neur, &n, st.dt, rng);
					store_state(n, v.state, v.size, i);
				}
				if (rng.count != int((i - i0 + 1) * v.draws)) atomicExch(v.error, 1);
				bits |= uint32_t(spiked) << (i - i0);
			}
		}
		uint64_t word = uint64_t(bits) << (16 * (lane % 4));
		word |= __shfl_xor(word, 1, 64);
		word |= __shfl_xor(word, 2, 64);
		if (active && lane % 4 == 0) gstore64(v.bitmap + u * v.bitmap_stride + (i0 >> 6), word);
	}
}

// Applies the deliveries counted for application step `local` (the next step
// of the population's instance) to the stored state and clears the counters,
// so that host reads see the state the reference has after the previous step
// and the next update does not apply them again.
template <class Neur, class Syn, int MaxIn, bool Mixed>
__global__ void __launch_bounds__(256)
apply_pending_kernel(spice_pop_view const* pops, int64_t pop, int64_t local, replay_slot const* slots) {
	using N                 = typename neuron_traits<Neur>::type;
	spice_pop_view const v  = pops[pop];
	int64_t const i         = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= v.size) return;
	N n;
	load_state(n, v.state, v.size, i);
	synapse_set<Syn, MaxIn> ss;
	ss.load(slots, v.nin);
	uint32_t k[MaxIn];
	load_counts<MaxIn>(k, v, local, i, true);
	replay<Syn, MaxIn, Mixed>(n, k, v, i, ss);
	store_state(n, v.state, v.size, i);
}

// ---------------------------------------------------------------- plastic synapses
// The reference updates plastic synapses lazily (synapse_population.h:87-115):
// a synapse is brought up to the current step only when its source's spike is
// delivered or at the sweep every 64 steps (snn.cpp:17-19), replaying the
// target's 64-step spike history with update() at the target's spikes and
// skip() over the quiet stretches in between.  Same arithmetic here, per synapse
// in parallel; shift counts are taken mod 64 as the reference's x86-64 code does.
template <class Syn>
__device__ inline void catch_up(Syn const& f, typename Syn::synapse& s, uint64_t ages, int64_t time, uint64_t hist,
                                float dt) {
	int const pre     = int(ages >> 63);
	int64_t const age = int64_t(ages & ~(uint64_t(1) << 63));
	if (!(time >= age)) return;
	int64_t const prefix = 63 + pre - time + age;
	if (pre) f.update(s, dt, true, ((hist >> (time - age)) & 1) != 0);
	hist &= ~uint64_t(0) >> (prefix & 63);
	int64_t p = prefix;
	while (hist) {
		int64_t const lz = __clzll(static_cast<long long>(hist));
		f.skip(s, dt, Int(lz - p));
		f.update(s, dt, false, true);
		hist ^= uint64_t(1) << (63 - lz);
		p = lz + 1;
	}
	f.skip(s, dt, Int(64 - p));
}

// Brings `row` (a source's synapses) up to `time`; one wave per row.
template <class Syn>
__device__ inline void catch_up_row(Syn const& f, spice_proj_view const& P, int64_t src, int64_t time, float dt,
                                    uint64_t new_age, int lane) {
	using S           = typename Syn::synapse;
	uint64_t const ag = P.ages[src];
	int64_t const a = P.offsets[src], b = P.offsets[src + 1];
	for (int64_t e = a + lane; e < b; e += 64) {
		S s = load_synapse<S>(P.syn_state, P.nnz, e);
		catch_up(f, s, ag, time, P.dst_hist[P.neighbors[e]], dt);
		store_synapse<S>(P.syn_state, P.nnz, e, s);
	}
	if (lane == 0) P.ages[src] = new_age;  // every lane has read the old value above
}

// Per-synapse init() of a synapse type whose init takes no draws: every synapse
// is independent, so one wave per source row initialises its synapses in
// parallel (the same calls, in any order, as the host's CSR-order loop).  A
// call that draws sets `error` (the host then reports the model unsupported).
template <class Syn>
__global__ void __launch_bounds__(256) init_synapses_kernel(spice_proj_view const* pv, Syn const* f, int32_t* error) {
	using S                 = typename Syn::synapse;
	spice_proj_view const P = *pv;
	Syn const syn           = cload_obj<Syn>(f);
	int const lane          = threadIdx.x % 64;
	int64_t const gw        = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
	int64_t const nw        = int64_t(gridDim.x) * blockDim.x / 64;
	for (int64_t src = gw; src < P.src_size; src += nw) {
		int64_t const a = P.offsets[src], b = P.offsets[src + 1];
		for (int64_t e = a + lane; e < b; e += 64) {
			S s{};
			device_stream rng{0, 0, 0};
			syn.init(s, Int(src), Int(P.neighbors[e]), rng);
			if (rng.count != 0) atomicExch(error, 1);
			store_synapse<S>(P.syn_state, P.nnz, e, s);
		}
	}
}

// The sweep at steps that are multiples of 64 (snn.cpp:17-19): every source.
// blockIdx.y: the projection (projs[y], functor funcs[y]) of a group of plastic
// projections with one synapse type.
template <class Syn>
__global__ void __launch_bounds__(256) plastic_sweep_kernel(spice_block b, int64_t const* projs, Syn const* funcs) {
	spice_proj_view const P = b.projs[projs[blockIdx.y]];
	if (P.inst < 0) return;
	spice_step const st     = step_of(b, P.inst, 0);
	if (st.local < 0 || st.local % 64 != 0) return;
	Syn const f       = cload_obj<Syn>(funcs + blockIdx.y);
	int const lane    = threadIdx.x % 64;
	int64_t const gw  = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
	int64_t const nw  = int64_t(gridDim.x) * blockDim.x / 64;
	for (int64_t src = gw; src < P.src_size; src += nw)
		catch_up_row(f, P, src, st.local, st.dt_nominal, uint64_t(st.local + 1), lane);
}

// The plastic deliveries of the step (snn.cpp:21-25): the sources that spiked
// delay - 1 steps ago are brought up to the step; their events were listed by
// the runtime and are applied by the targets at the next step.
template <class Syn>
__global__ void __launch_bounds__(256) plastic_deliver_kernel(spice_block b, int64_t const* projs,
                                                              Syn const* funcs) {
	spice_proj_view const P = b.projs[projs[blockIdx.y]];
	if (P.inst < 0) return;
	spice_step const st     = step_of(b, P.inst, 0);
	int64_t const t         = st.local;
	if (t < 0 || t < P.delay - 1) return;
	Syn const f          = cload_obj<Syn>(funcs + blockIdx.y);
	int64_t const slot   = (t - P.delay + 1) % P.ring;
	int64_t const nsp    = P.src_ring_cnt[slot];
	int32_t const* ids   = P.src_ring_ids + slot * P.src_size;
	int const lane       = threadIdx.x % 64;
	int64_t const gw     = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
	int64_t const nw     = int64_t(gridDim.x) * blockDim.x / 64;
	uint64_t const aged  = uint64_t(t + 1) | (uint64_t(1) << 63);
	for (int64_t k = gw; k < nsp; k += nw) catch_up_row(f, P, ids[k], t, st.dt_nominal, aged, lane);
}

}  // namespace spice::detail
