// Runtime of the MI355X Spice engine: contexts, populations, projections, the
// delay ring, spike compaction and spike delivery.  C ABI in include/spice_hip.h.
//
// Data layout in HBM (per context):
//   population state   word-planar SoA, uint32 [words][size]
//   spike bits         uint64 [SPICE_MAX_BLOCK][ceil(size/64)] (one block of steps)
//   spike ring         int32 [ring][size] ascending ids + int32 [ring] counts,
//                      slot = step % ring, ring = max_delay + SPICE_MAX_BLOCK
//   projection CSR     int64 offsets [src+1], int32 neighbors [nnz]
//   event counts       uint32 [ring][dst_size] per projection, slot = application step % ring
// Steps are the population's instance's own step indices (spice_step::local).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "crlog.h"
#include "internal.h"

using namespace spice_rt;

namespace spice_rt {
thread_local std::string g_error;

int set_error(int code, std::string const& msg) {
	g_error = msg;
	return code;
}

// csr_gen.hip: fixed_probability projections are queued and generated together
int queue_fixed_probability(spice_ctx* ctx, int64_t proj, int64_t src_size, int64_t dst_size, double p,
                            xoro_state seed);
int flush_fixed_probability(spice_ctx* ctx);
int build_blocked(spice_ctx* ctx, Proj& Q, int64_t tile);  // blocked.hip
}  // namespace spice_rt

namespace {

constexpr int kWave = 64;

template <class T>
int dev_alloc(T** p, size_t count, bool zero = true) {
	*p = nullptr;
	if (count == 0) count = 1;
	hipError_t e = dmalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
	if (e != hipSuccess)
		return set_error(SPICE_ENOMEM, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
		                                   " bytes failed: " + hipGetErrorString(e));
	if (zero) {
		e = hipMemset(*p, 0, sizeof(T) * count);
		if (e != hipSuccess) return set_error(SPICE_EHIP, hipGetErrorString(e));
	}
	return SPICE_OK;
}

int precondition(char const* file, int line, char const* cond) {
	return set_error(SPICE_EPRECONDITION, std::string("Assertion failed (") + file + ":" +
	                                          std::to_string(line) + "): " + cond);
}

// ------------------------------------------------------------------ kernels

__device__ inline int64_t local_step(spice_window const* win, int64_t inst, int rel) {
	return win->steps[inst * SPICE_MAX_WINDOW + rel].local;
}

// The first SPICE_STREAM_STATES states of every active step's stream of the
// window (one thread per instance and step).
__global__ void __launch_bounds__(64)
stream_states_kernel(spice_step const* steps, spice_rng_state* streams, int64_t ninst, int nsteps) {
	int64_t const q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (q >= ninst * nsteps) return;
	int64_t const inst = q / nsteps;
	int const k        = int(q % nsteps);
	spice_step const st = steps[inst * SPICE_MAX_WINDOW + k];
	if (st.local < 0) return;
	uint64_t s0 = st.seed.s0, s1 = st.seed.s1;
	spice_rng_state* out = streams + (inst * SPICE_MAX_WINDOW + k) * SPICE_STREAM_STATES;
	for (int i = 0; i < SPICE_STREAM_STATES; i++) {
		out[i]           = {s0, s1};
		uint64_t const t = s0 ^ s1;  // xoroshiro128+ transition (rotl 24, shl 16, rotl 37)
		s0               = ((s0 << 24) | (s0 >> 40)) ^ t ^ (t << 16);
		s1               = (t << 37) | (t >> 27);
	}
}

// Spike bits of one (population, step) -> ascending id list in the ring slot.
// Each thread owns a run of consecutive bitmap words; a workgroup scan of the
// threads' spike counts places every thread's ids.
constexpr int kCompactThreads = 1024;
__global__ void __launch_bounds__(kCompactThreads)
compact_kernel(PopDev const* pops, spice_window const* win, int rel, int64_t ring) {
	PopDev const P  = pops[blockIdx.x];
	int const k     = blockIdx.y;
	if (P.inst < 0) return;  // destroyed
	int64_t const t = local_step(win, P.inst, rel + k);
	if (t < 0) return;
	int64_t const nw = (P.size + 63) / 64;
	uint64_t const* bm = P.bitmap + (P.window_bits ? rel + k : k) * P.bitmap_stride;
	int64_t const slot = t % ring;
	int32_t* out       = P.ring_ids + slot * P.size;

	int const tid      = threadIdx.x, lane = tid % 64, wave = tid / 64;
	constexpr int kWaves = kCompactThreads / 64;
	int64_t const per  = (nw + kCompactThreads - 1) / kCompactThreads;
	int64_t const w0   = min<int64_t>(nw, tid * per);
	int64_t const w1   = min<int64_t>(nw, w0 + per);
	int cnt            = 0;
	for (int64_t w = w0; w < w1; w++) cnt += __popcll(bm[w]);

	int x = cnt;  // inclusive scan within the wave, then over the wave totals
	for (int off = 1; off < 64; off <<= 1) {
		int const y = __shfl_up(x, off, 64);
		if (lane >= off) x += y;
	}
	__shared__ int wsum[kWaves];
	if (lane == 63) wsum[wave] = x;
	__syncthreads();
	if (wave == 0) {
		int y = lane < kWaves ? wsum[lane] : 0;
		for (int off = 1; off < kWaves; off <<= 1) {
			int const z = __shfl_up(y, off, 64);
			if (lane >= off) y += z;
		}
		if (lane < kWaves) wsum[lane] = y;
	}
	__syncthreads();
	int pos = (wave > 0 ? wsum[wave - 1] : 0) + x - cnt;
	for (int64_t w = w0; w < w1; w++) {
		uint64_t m = bm[w];
		while (m) {
			int const b = __ffsll((unsigned long long)m) - 1;
			out[pos++]  = int32_t(w * 64 + b);
			m &= m - 1;
		}
	}
	if (tid == 0) P.ring_cnt[slot] = wsum[kWaves - 1];
}

// Spike history of populations with plastic projections (neuron_population.h:122-128):
// per step of the block, hist = hist << 1 | spiked.
__global__ void __launch_bounds__(256) history_kernel(PopDev const* pops, spice_window const* win, int rel, int n) {
	PopDev const P = pops[blockIdx.y];
	if (!P.hist || P.inst < 0) return;
	int64_t const i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= P.size) return;
	uint64_t h = P.hist[i];
	int64_t const row0 = P.window_bits ? rel : 0;
	for (int u = 0; u < n; u++) {
		if (local_step(win, P.inst, rel + u) < 0) break;
		h = (h << 1) | ((P.bitmap[(row0 + u) * P.bitmap_stride + (i >> 6)] >> (i & 63)) & 1u);
	}
	P.hist[i] = h;
}

// Tiled delivery.  Pull: one workgroup per (target tile, step).  The workgroup
// counts, in LDS, the events that the step's spikes of its incoming projections
// put on each target of its tile, then writes the tile's counters once: no
// global atomics.  Push (shared-then-global): one workgroup per (target tile,
// step, chunk of the step's spikes); it counts its chunk's events in LDS the
// same way and adds its nonzero counters to the global ones with device-scope
// atomics, in tile order (16 counters of one cache line per request), so a
// launch issues about one global atomic request per 16 targets and chunk
// instead of one per event; the neuron update clears the counters it consumes.  Counters are 32 bits (a target receives at most one event per
// spiking source of a projection); the padding slots of the runs count into
// kPadSlots spare counters past the tile, which are never written out.  Counts
// are integers, so the result does not depend on the order of the LDS atomics
// and equals the push result exactly.  Projections are counted `pass` at a time
// (as many as fit the LDS share that keeps kPullResident workgroups per compute
// unit).
//
// Work split: a spike's run in the tile (a slice of its CSR row, about p * tile
// targets, padded to whole 16-byte vectors of 8 tile-local targets) is read by
// a group of kGroup lanes, lane l loading vectors l, l + kGroup, ...; a group
// keeps kSpikesInFlight runs x kRunUnroll vectors per lane in flight.  Loads
// are unconditional (a lane past the end of its run reloads the run's first
// vector and does not count it), so they issue back to back and are consumed in
// order; counting a vector is 8 LDS atomics with no per-target checks.  The run
// descriptors of a batch of spikes (address of the first vector, vector count,
// counter base) are computed first, one spike per thread, and staged in LDS.
//
// Workgroup L handles step L % 16 and (tile, chunk) L / 16: the steps of one block land on
// fixed XCDs (workgroups are dealt to the 8 XCDs round-robin), so the spike
// lists and run bounds that all workgroups of a step read are cached in one L2.
struct PullArgs {
	ProjDev const* projs;
	PullTile const* tiles;
	PullTarget const* targets;
	spice_window const* win;
	int32_t rel;
	int32_t nsteps;
	int64_t ring;
	int64_t max_tile;  // largest tile of the launch (sets the LDS layout)
	int32_t pass;      // projections counted together (<= kPassProj)
	int32_t chunks;    // push: spike chunks per (tile, step)
	unsigned long long* events;
};

constexpr int kPullThreads    = 512;
constexpr int kGroup          = 8;             // lanes per spike run
constexpr int kRunUnroll      = 4;             // vectors per lane per run in flight
constexpr int kSpikesInFlight = 2;             // runs per group in flight
constexpr int kDescBatch      = kPullThreads;  // spike descriptors staged in LDS at once
constexpr int kPassProj       = 4;             // most projections counted in one pass
constexpr int kPadSlots       = 8;             // spare counters taking the padding slots
constexpr int kPullResident   = 3;             // workgroups per compute unit the pass size aims for

struct RunDesc {
	uint64_t addr;  // first 16-byte vector of the run
	uint32_t nv;    // vectors in the run
	uint32_t cnt;   // LDS word offset of the run's projection counters
};

// An incoming projection as seen from one (tile, step) workgroup.
struct PassProj {
	int32_t const* ids;      // the step's spiking sources
	uint32_t const* rv;      // run bounds of the tile: rv[src * rvs], rv[src * rvs + 1]
	uint16_t const* btgt;
	int64_t const* offsets;  // CSR rows (event counts)
	uint32_t* out;           // the tile's counters in the delivery slot
	int64_t nsp;             // spikes
	int64_t rvs;
};

constexpr size_t kLdsMax = 160 * 1024;  // LDS of a CDNA4 compute unit

// Loads through global-address-space pointers: generic (flat) loads also count
// against the LDS wait counter, so every LDS wait would drain them as well.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ inline T gload(T const* p) {
	using gptr = __attribute__((address_space(1))) T const*;
	return *(gptr)(p);
}

template <class T>
__device__ inline void gstore(T* p, T v) {
	using gptr = __attribute__((address_space(1))) T*;
	*(gptr)(p) = v;
}

// LDS bytes of one pull workgroup counting `pass` projections of a tile together.
size_t pull_lds_bytes(int64_t tile, int pass) {
	return sizeof(uint32_t) * size_t(pass) * size_t(tile + kPadSlots) + sizeof(RunDesc) * kDescBatch +
	       sizeof(PassProj) * SPICE_MAX_IN;
}

// Projections per pass for a tile: as many as the incoming projections need
// (at most kPassProj) while kPullResident workgroups fit a compute unit's LDS,
// at least one (0: one projection's counters do not fit).
int pull_pass(int64_t tile, int max_in) {
	int p = std::min(kPassProj, std::max(max_in, 1));
	while (p > 1 && pull_lds_bytes(tile, p) > kLdsMax / kPullResident) p--;
	return pull_lds_bytes(tile, p) <= kLdsMax ? p : 0;
}

template <bool Push>
__device__ __forceinline__ void deliver_tiles(PullArgs const& a) {
	extern __shared__ uint32_t lds[];  // [pass][max_tile + kPadSlots] counters, then descriptors
	int const k = int(blockIdx.x % SPICE_MAX_BLOCK);
	if (k >= a.nsteps) return;
	unsigned const tc    = blockIdx.x / SPICE_MAX_BLOCK;
	int const chunk      = Push ? int(tc % unsigned(a.chunks)) : 0;
	PullTile const pt    = a.tiles[Push ? tc / unsigned(a.chunks) : tc];
	PullTarget const& tg = a.targets[pt.target];
	int64_t const step   = local_step(a.win, tg.inst, a.rel + k);
	if (step < 0) return;
	int64_t const T0     = int64_t(pt.tau) * tg.tile;
	int const Tn         = int(min<int64_t>(tg.tile, tg.dst_size - T0));
	int const stride     = int(tg.tile) + kPadSlots;  // counters per projection
	RunDesc* desc        = reinterpret_cast<RunDesc*>(lds + a.pass * (a.max_tile + kPadSlots));
	PassProj* pp         = reinterpret_cast<PassProj*>(desc + kDescBatch);
	int const tid = threadIdx.x, lane = tid % kWave, group = tid / kGroup, gl = tid % kGroup;
	constexpr int kGroups = kPullThreads / kGroup;
	int64_t const slot_in = step % a.ring;
	unsigned long long ev = 0, rows = 0;
	int const nin = tg.nin;
	if (tid < nin) {  // the incoming projections as seen from this tile and step
		ProjDev const& P = a.projs[tg.proj[tid]];
		pp[tid] = {P.src_ring_ids + slot_in * P.src_size, P.rv + pt.tau, P.btgt, P.offsets,
		           P.counts + ((step + P.delay) % P.slots) * P.dst_size + T0, gload(P.src_ring_cnt + slot_in),
		           P.ntile + 1};
	}
	__syncthreads();

	for (int j0 = 0; j0 < nin; j0 += a.pass) {
		int const np = min(a.pass, nin - j0);
		int64_t pre[kPassProj + 1];  // spikes of the pass's projections, prefix
		pre[0] = 0;
#pragma unroll
		for (int j = 0; j < kPassProj; j++) pre[j + 1] = pre[j] + (j < np ? pp[j0 + j].nsp : 0);
		int64_t const M  = pre[kPassProj];
		int64_t const lo = Push ? M * chunk / a.chunks : 0;  // this workgroup's spikes of the pass
		int64_t const hi = Push ? M * (chunk + 1) / a.chunks : M;
		if (pt.tau == 0 && tid == 0) rows += uint64_t(hi - lo);
		if (Push && lo == hi) continue;  // nothing to add (uniform over the workgroup)
		for (int i = tid; i < np * stride; i += kPullThreads) lds[i] = 0;
		for (int64_t d0 = lo; d0 < hi; d0 += kDescBatch) {
			int const nd = int(min<int64_t>(kDescBatch, hi - d0));
			__syncthreads();  // counters cleared / previous descriptors consumed
			if (tid < nd) {
				int64_t const g = d0 + tid;
				int j           = 0;
#pragma unroll
				for (int q = 1; q < kPassProj; q++) j += g >= pre[q] ? 1 : 0;
				PassProj const& q = pp[j0 + j];
				int64_t const src = gload(q.ids + (g - pre[j]));
				uint32_t const* r = q.rv + src * q.rvs;
				uint32_t const lo = gload(r), hi = gload(r + 1);
				desc[tid] = {reinterpret_cast<uint64_t>(q.btgt + uint64_t(lo) * 8), hi - lo, uint32_t(j * stride)};
				if (pt.tau == 0) ev += uint64_t(gload(q.offsets + src + 1) - gload(q.offsets + src));
			}
			__syncthreads();
			for (int r0 = group; r0 < nd; r0 += kGroups * kSpikesInFlight) {
				u32x4 const* v[kSpikesInFlight];
				uint32_t nv[kSpikesInFlight], cb[kSpikesInFlight], most = 0;
#pragma unroll
				for (int r = 0; r < kSpikesInFlight; r++) {
					int const idx   = r0 + r * kGroups;
					RunDesc const d = desc[idx < nd ? idx : r0];
					v[r]            = reinterpret_cast<u32x4 const*>(d.addr);
					nv[r]           = idx < nd ? d.nv : 0;
					cb[r]           = d.cnt;
					most            = max(most, nv[r]);
				}
				for (uint32_t vb = 0; vb < most; vb += kGroup * kRunUnroll) {
					u32x4 w[kSpikesInFlight][kRunUnroll];
#pragma unroll
					for (int r = 0; r < kSpikesInFlight; r++)
#pragma unroll
						for (int u = 0; u < kRunUnroll; u++) {
							uint32_t const vi = vb + uint32_t(u * kGroup + gl);
							w[r][u]           = gload(v[r] + (vi < nv[r] ? vi : 0));
						}
					// keep every load issued here, ahead of the counting (the compiler
					// would otherwise move a load into the branch that consumes it)
					asm volatile("" ::: "memory");
#pragma unroll
					for (int r = 0; r < kSpikesInFlight; r++)
#pragma unroll
						for (int u = 0; u < kRunUnroll; u++) {
							uint32_t const vi = vb + uint32_t(u * kGroup + gl);
							if (vi < nv[r]) {
								uint32_t* c = lds + cb[r];
								atomicAdd(c + (w[r][u].x & 0xffffu), 1u);
								atomicAdd(c + (w[r][u].x >> 16), 1u);
								atomicAdd(c + (w[r][u].y & 0xffffu), 1u);
								atomicAdd(c + (w[r][u].y >> 16), 1u);
								atomicAdd(c + (w[r][u].z & 0xffffu), 1u);
								atomicAdd(c + (w[r][u].z >> 16), 1u);
								atomicAdd(c + (w[r][u].w & 0xffffu), 1u);
								atomicAdd(c + (w[r][u].w >> 16), 1u);
							}
						}
				}
			}
		}
		__syncthreads();
		for (int j = 0; j < np; j++) {
			uint32_t* out = pp[j0 + j].out;
			if (Push) {
				for (int i = tid; i < Tn; i += kPullThreads) {
					uint32_t const c = lds[j * stride + i];
					if (c) atomicAdd(out + i, c);
				}
			} else {
				for (int i = tid; i < Tn; i += kPullThreads) gstore(out + i, lds[j * stride + i]);
			}
		}
		__syncthreads();  // counters written before the next pass clears them
	}
	for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off, kWave);
	if (lane == 0 && ev) atomicAdd(a.events, ev);
	if (tid == 0 && rows) atomicAdd(a.events + 1, rows);
}

__global__ void __launch_bounds__(kPullThreads) deliver_pull_kernel(PullArgs a) { deliver_tiles<false>(a); }
__global__ void __launch_bounds__(kPullThreads) deliver_push_kernel(PullArgs a) { deliver_tiles<true>(a); }

// Event lists of projections with per-synapse state (one-step blocks): the
// events a target receives at the next step are laid out contiguously, target
// by target.  The offsets are an exclusive scan of the targets' event counts of
// the next step (written by the counting delivery), in two launches over chunks
// of kEvChunk targets: chunk sums, then each chunk scans its counts from the
// sum of the chunks before it.
constexpr int kEvChunk   = 4096;
constexpr int kEvThreads = 256;  // 16 consecutive targets per thread

__device__ inline uint32_t const* next_counts(spice_proj_view const& P, ProjDev const& Q, int64_t t) {
	return Q.counts + ((t + 1) % Q.slots) * P.dst_size;
}

__global__ void __launch_bounds__(kEvThreads)
event_chunk_sums_kernel(spice_proj_view const* pv, int64_t const* ids, ProjDev const* projs, spice_window const* win,
                        int rel, uint32_t* sums, int64_t max_chunks) {
	int64_t const c         = ids[blockIdx.y];
	spice_proj_view const P = pv[c];
	int64_t const t         = local_step(win, P.inst, rel);
	int64_t const i0        = int64_t(blockIdx.x) * kEvChunk;
	if (t < 0 || i0 >= P.dst_size) return;
	uint32_t const* counts = next_counts(P, projs[c], t);
	int64_t const a = min<int64_t>(P.dst_size, i0 + int64_t(threadIdx.x) * 16);
	int64_t const b = min<int64_t>(P.dst_size, a + 16);
	uint32_t sum    = 0;
	for (int64_t i = a; i < b; i++) sum += counts[i];
	for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
	__shared__ uint32_t ws[kEvThreads / 64];
	if (threadIdx.x % 64 == 0) ws[threadIdx.x / 64] = sum;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t s = 0;
		for (int w = 0; w < kEvThreads / 64; w++) s += ws[w];
		sums[blockIdx.y * max_chunks + blockIdx.x] = s;
	}
}

__global__ void __launch_bounds__(kEvThreads)
event_offsets_kernel(spice_proj_view const* pv, int64_t const* ids, ProjDev const* projs, spice_window const* win,
                     int rel, uint32_t const* sums, int64_t max_chunks) {
	int64_t const c         = ids[blockIdx.y];
	spice_proj_view const P = pv[c];
	int64_t const t         = local_step(win, P.inst, rel);
	int64_t const i0        = int64_t(blockIdx.x) * kEvChunk;
	if (t < 0 || i0 >= P.dst_size) return;
	uint32_t const* counts = next_counts(P, projs[c], t);
	int const tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
	constexpr int kWaves = kEvThreads / 64;
	__shared__ uint32_t ws[kWaves];
	__shared__ uint32_t base;
	// events of the chunks before this one
	uint32_t pre = 0;
	for (int64_t q = tid; q < int64_t(blockIdx.x); q += kEvThreads) pre += sums[blockIdx.y * max_chunks + q];
	for (int off = 32; off > 0; off >>= 1) pre += __shfl_xor(pre, off, 64);
	if (lane == 0) ws[wave] = pre;
	__syncthreads();
	if (tid == 0) {
		uint32_t s = 0;
		for (int w = 0; w < kWaves; w++) s += ws[w];
		base = s;
	}
	__syncthreads();
	int64_t const a = min<int64_t>(P.dst_size, i0 + int64_t(tid) * 16);
	int64_t const b = min<int64_t>(P.dst_size, a + 16);
	uint32_t k[16];
	uint32_t sum = 0;
#pragma unroll
	for (int j = 0; j < 16; j++) {
		k[j] = a + j < b ? counts[a + j] : 0u;
		sum += k[j];
	}
	uint32_t x = sum;  // inclusive scan over the wave, then over the waves
	for (int off = 1; off < 64; off <<= 1) {
		uint32_t const y = __shfl_up(x, off, 64);
		if (lane >= off) x += y;
	}
	__syncthreads();  // every thread has read ws
	if (lane == 63) ws[wave] = x;
	__syncthreads();
	uint32_t pos = base + x - sum;
	for (int w = 0; w < wave; w++) pos += ws[w];
	uint32_t* offs = projs[c].ev_offsets;
#pragma unroll
	for (int j = 0; j < 16; j++)
		if (a + j < b) {
			offs[a + j]        = pos;
			P.ev_cursor[a + j] = pos;
			pos += k[j];
		}
}

// Appends the synapse ids of the spikes delivered this step (spikes of step
// t - delay + 1, snn.cpp:21-25) to their targets' event lists: one wave per
// spiking source, lanes over its row.  The order inside a list is arbitrary;
// the target's kernel applies them by ascending synapse id.
__global__ void __launch_bounds__(256)
event_scatter_kernel(spice_proj_view const* pv, int64_t const* ids, spice_window const* win, int rel) {
	spice_proj_view const P = pv[ids[blockIdx.y]];
	int64_t const t         = local_step(win, P.inst, rel);
	if (t < 0 || t < P.delay - 1) return;
	int64_t const slot = (t - P.delay + 1) % P.ring;
	int64_t const nsp  = P.src_ring_cnt[slot];
	int32_t const* sp  = P.src_ring_ids + slot * P.src_size;
	int const lane     = threadIdx.x % kWave;
	int64_t const gw   = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
	int64_t const nw   = int64_t(gridDim.x) * blockDim.x / kWave;
	for (int64_t k = gw; k < nsp; k += nw) {
		int64_t const src = sp[k];
		int64_t const a = P.offsets[src], b = P.offsets[src + 1];
		for (int64_t e = a + lane; e < b; e += kWave) {
			uint32_t const pos = atomicAdd(P.ev_cursor + P.neighbors[e], 1u);
			P.events[pos]      = uint32_t(e);
		}
	}
}

int64_t pop_tile(spice_ctx const* ctx, Pop const& P) { return P.tile ? P.tile : ctx->tile; }

// Drops the cached window graphs (they hold the old descriptors and grids).
void drop_graphs(spice_ctx* ctx) {
	for (auto& g : ctx->graph_cache) (void)hipGraphExecDestroy(g.second);
	ctx->graph_cache.clear();
}

int refresh_device_descriptors(spice_ctx* ctx) {
	if (int r = flush_fixed_probability(ctx)) return r;
	if (!ctx->dev_dirty) return SPICE_OK;
	drop_graphs(ctx);
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	for (auto& Q : ctx->projs) {  // both modes deliver over the tile view
		if (Q.dead) continue;
		int64_t const t = pop_tile(ctx, ctx->pops[Q.dst]);
		if (Q.tile != t)
			if (int r = build_blocked(ctx, Q, t)) return r;
	}
	std::vector<PopDev> pd;
	std::vector<spice_pop_view> views;
	ctx->max_pop_size = 0;
	for (auto& P : ctx->pops) {
		pd.push_back({P.dead ? -1 : int64_t(P.inst), P.bitmap, P.bitmap_stride, P.window_bits ? 1 : 0, P.size,
		              P.ring_ids, P.ring_cnt, P.hist});
		ctx->max_pop_size = std::max(ctx->max_pop_size, P.size);
		spice_pop_view v;
		std::memset(&v, 0, sizeof(v));
		v.state         = P.state;
		v.size          = P.size;
		v.words         = P.words;
		v.draws         = P.draws;
		v.bitmap        = P.bitmap;
		v.bitmap_stride = P.bitmap_stride;
		v.chunk_jump    = reinterpret_cast<spice_rng_state const*>(P.chunk_jump);
		v.inst          = P.inst;
		v.nin           = int32_t(P.incoming.size());
		for (size_t j = 0; j < P.incoming.size(); j++) {
			Proj& Q = ctx->projs[P.incoming[j]];
			v.in[j] = {Q.counts, Q.slots, P.incoming[j], Q.ev_offsets, Q.events, Q.syn_state, Q.nnz};
		}
		v.error        = ctx->error;
		v.clear_counts = ctx->mode == SPICE_DELIVER_PUSH ? 1 : 0;
		v.hist         = P.hist;
		views.push_back(v);
	}
	std::vector<ProjDev> qd;
	for (auto& Q : ctx->projs) {
		Pop const& S = ctx->pops[Q.src];
		qd.push_back({Q.dead ? -1 : int64_t(S.inst), Q.offsets, Q.neighbors, Q.counts, Q.slots, ctx->pops[Q.dst].size, Q.delay,
		              S.ring_ids, S.ring_cnt, S.size, Q.rv, Q.btgt, Q.ntile, Q.ev_offsets});
	}
	// projection views for the synapse kernels; ids of stateful projections
	std::vector<spice_proj_view> pvs;
	std::vector<int64_t> stateful;
	ctx->any_plastic = false;
	for (size_t c = 0; c < ctx->projs.size(); c++) {
		Proj& Q      = ctx->projs[c];
		Pop const& S = ctx->pops[Q.src];
		Pop const& D = ctx->pops[Q.dst];
		spice_proj_view v;
		std::memset(&v, 0, sizeof(v));
		v.offsets      = Q.offsets;
		v.neighbors    = Q.neighbors;
		v.nnz          = Q.nnz;
		v.src_size     = S.size;
		v.dst_size     = D.size;
		v.delay        = Q.delay;
		v.ring         = ctx->ring;
		v.src_ring_ids = S.ring_ids;
		v.src_ring_cnt = S.ring_cnt;
		v.syn_state    = Q.syn_state;
		v.syn_words    = Q.syn_words;
		v.inst         = Q.dead ? -1 : S.inst;
		v.ages         = Q.ages;
		v.dst_hist     = D.hist;
		v.ev_cursor    = Q.ev_cursor;
		v.events       = Q.events;
		pvs.push_back(v);
		if (Q.dead) continue;
		if (Q.syn_state) stateful.push_back(int64_t(c));
		ctx->any_plastic = ctx->any_plastic || Q.plastic;
	}
	ctx->nstateful = int64_t(stateful.size());
	int64_t chunks = 1;  // event-list scan chunks of the largest stateful projection
	for (int64_t c : stateful) chunks = std::max(chunks, (ctx->pops[ctx->projs[c].dst].size + kEvChunk - 1) / kEvChunk);
	if (ctx->nstateful > 0 && ctx->nstateful * chunks > ctx->ev_sums_cap) {
		dfree(ctx->stream, ctx->ev_sums);
		ctx->ev_sums     = nullptr;
		ctx->ev_sums_cap = 0;
		if (int r = dev_alloc(&ctx->ev_sums, size_t(ctx->nstateful * chunks))) return r;
		ctx->ev_sums_cap = ctx->nstateful * chunks;
	}
	ctx->ev_max_chunks = chunks;
	// pull tiles: every population with incoming projections, split into tiles
	std::vector<PullTarget> pt;
	std::vector<PullTile> tiles;
	int max_in        = 0;
	ctx->pull_max_tile = 64;
	for (auto& P : ctx->pops) {
		if (P.dead) continue;
		max_in = std::max(max_in, int(P.incoming.size()));
		if (!P.incoming.empty()) ctx->pull_max_tile = std::max(ctx->pull_max_tile, pop_tile(ctx, P));
	}
	ctx->pull_pass = pull_pass(ctx->pull_max_tile, max_in);
	if (ctx->pull_pass == 0) return set_error(SPICE_EINVAL, "tile too large for the LDS of a compute unit");
	void const* kern = ctx->mode == SPICE_DELIVER_PUSH ? reinterpret_cast<void const*>(&deliver_push_kernel)
	                                                   : reinterpret_cast<void const*>(&deliver_pull_kernel);
	SPICE_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
	                              int(pull_lds_bytes(ctx->pull_max_tile, ctx->pull_pass))));
	for (auto& P : ctx->pops) {
		if (P.dead || P.incoming.empty() || P.size == 0) continue;
		PullTarget t{};
		t.inst     = P.inst;
		t.tile     = pop_tile(ctx, P);
		t.dst_size = P.size;
		t.nin      = int32_t(P.incoming.size());
		for (size_t j = 0; j < P.incoming.size(); j++) t.proj[j] = int32_t(P.incoming[j]);
		int32_t const id = int32_t(pt.size());
		pt.push_back(t);
		for (int64_t tau = 0; tau * t.tile < P.size; tau++) tiles.push_back({id, int32_t(tau)});
	}
	dfree(ctx->stream, ctx->pull_tiles);
	dfree(ctx->stream, ctx->pull_targets);
	dfree(ctx->stream, ctx->pops_dev);
	dfree(ctx->stream, ctx->projs_dev);
	dfree(ctx->stream, ctx->views_dev);
	dfree(ctx->stream, ctx->proj_views_dev);
	dfree(ctx->stream, ctx->stateful_dev);
	ctx->proj_views_dev = nullptr;
	ctx->stateful_dev   = nullptr;
	ctx->pull_tiles   = nullptr;
	ctx->pull_targets = nullptr;
	ctx->pops_dev     = nullptr;
	ctx->projs_dev    = nullptr;
	ctx->views_dev    = nullptr;
	ctx->pull_ntiles  = int64_t(tiles.size());
	if (int r = dev_alloc(&ctx->pull_tiles, tiles.size(), false)) return r;
	if (int r = dev_alloc(&ctx->pull_targets, pt.size(), false)) return r;
	if (int r = dev_alloc(&ctx->pops_dev, pd.size(), false)) return r;
	if (int r = dev_alloc(&ctx->projs_dev, qd.size(), false)) return r;
	if (int r = dev_alloc(&ctx->views_dev, views.size(), false)) return r;
	if (int r = dev_alloc(&ctx->proj_views_dev, pvs.size(), false)) return r;
	if (int r = dev_alloc(&ctx->stateful_dev, stateful.size(), false)) return r;
	if (!pvs.empty())
		SPICE_HIP(hipMemcpy(ctx->proj_views_dev, pvs.data(), sizeof(spice_proj_view) * pvs.size(),
		                    hipMemcpyHostToDevice));
	if (!stateful.empty())
		SPICE_HIP(hipMemcpy(ctx->stateful_dev, stateful.data(), sizeof(int64_t) * stateful.size(),
		                    hipMemcpyHostToDevice));
	if (!tiles.empty())
		SPICE_HIP(hipMemcpy(ctx->pull_tiles, tiles.data(), sizeof(PullTile) * tiles.size(), hipMemcpyHostToDevice));
	if (!pt.empty())
		SPICE_HIP(hipMemcpy(ctx->pull_targets, pt.data(), sizeof(PullTarget) * pt.size(), hipMemcpyHostToDevice));
	if (!pd.empty())
		SPICE_HIP(hipMemcpy(ctx->pops_dev, pd.data(), sizeof(PopDev) * pd.size(), hipMemcpyHostToDevice));
	if (!qd.empty())
		SPICE_HIP(hipMemcpy(ctx->projs_dev, qd.data(), sizeof(ProjDev) * qd.size(), hipMemcpyHostToDevice));
	if (!views.empty())
		SPICE_HIP(hipMemcpy(ctx->views_dev, views.data(), sizeof(spice_pop_view) * views.size(),
		                    hipMemcpyHostToDevice));
	// window tables sized for every instance
	int64_t const ninst = std::max<int64_t>(int64_t(ctx->insts.size()), 1);
	if (ninst > ctx->win_cap) {
		int64_t const cap = std::max<int64_t>(ninst, 2 * ctx->win_cap);
		for (int b = 0; b < 2; b++) {
			if (ctx->win_copied[b]) SPICE_HIP(hipEventSynchronize(ctx->win_copied[b]));
			(void)hipHostFree(ctx->win_host[b]);
			ctx->win_host[b] = nullptr;
			SPICE_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->win_host[b]),
			                        sizeof(spice_step) * size_t(cap) * SPICE_MAX_WINDOW));
		}
		dfree(ctx->stream, ctx->win_steps_dev);
		ctx->win_steps_dev = nullptr;
		if (int r = dev_alloc(&ctx->win_steps_dev, size_t(cap) * SPICE_MAX_WINDOW)) return r;
		ctx->win_cap = cap;
	}
	ctx->any_draws = false;
	for (auto& P : ctx->pops) ctx->any_draws = ctx->any_draws || (!P.dead && P.draws > 0 && P.size > 0);
	if (ctx->any_draws && ctx->win_streams_cap < ctx->win_cap) {
		dfree(ctx->stream, ctx->win_streams_dev);
		ctx->win_streams_dev = nullptr;
		ctx->win_streams_cap = 0;
		if (int r = dev_alloc(&ctx->win_streams_dev, size_t(ctx->win_cap) * SPICE_MAX_WINDOW * SPICE_STREAM_STATES,
		                      false))
			return r;
		ctx->win_streams_cap = ctx->win_cap;
	}
	spice_window const w{ctx->win_steps_dev, ctx->any_draws ? ctx->win_streams_dev : nullptr, 0,
	                     int32_t(ctx->insts.size())};
	if (!ctx->win_dev)
		if (int r = dev_alloc(&ctx->win_dev, 1, false)) return r;
	SPICE_HIP(hipMemcpy(ctx->win_dev, &w, sizeof(w), hipMemcpyHostToDevice));
	ctx->dev_dirty = false;
	return SPICE_OK;
}

int check_device_error(spice_ctx* ctx) {
	int32_t flag = 0;
	SPICE_HIP(hipMemcpyAsync(&flag, ctx->error, sizeof(flag), hipMemcpyDeviceToHost, ctx->stream));
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	if (flag == 0) return SPICE_OK;
	(void)hipMemsetAsync(ctx->error, 0, sizeof(int32_t), ctx->stream);
	if (flag == 1)
		return set_error(SPICE_EDEVICE,
		                 "a neuron's update() took a different number of RNG draws than the first "
		                 "neuron of its population did when probed on the host; per-neuron draw "
		                 "counts must not depend on the neuron's state");
	return set_error(SPICE_EDEVICE, "device-side check failed (code " + std::to_string(flag) + ")");
}

// Spike chunks per (tile, step) of a push launch: the context's setting, else
// one.  Every chunk past the first adds a global atomic per counter it touches;
// splitting was slower at every size measured (N = 20k to 500k), also where the
// (tile, step) workgroups leave compute units idle.
int32_t push_chunks(spice_ctx const* ctx) { return ctx->push_chunks > 0 ? ctx->push_chunks : 1; }

// Delivers the spikes of steps rel .. rel+n-1 of the staged window.
int launch_delivery(spice_ctx* ctx, int32_t rel, int32_t n) {
	if (ctx->projs.empty() || ctx->pull_ntiles == 0) return SPICE_OK;
	bool const push    = ctx->mode == SPICE_DELIVER_PUSH;
	int32_t const ch   = push ? push_chunks(ctx) : 1;
	PullArgs const a{ctx->projs_dev, ctx->pull_tiles, ctx->pull_targets, ctx->win_dev, rel, n,
	                 ctx->ring, ctx->pull_max_tile, ctx->pull_pass, ch, ctx->events};
	dim3 const grid(unsigned(ctx->pull_ntiles * ch * SPICE_MAX_BLOCK));
	size_t const lds = pull_lds_bytes(ctx->pull_max_tile, ctx->pull_pass);
	if (push)
		hipLaunchKernelGGL(deliver_push_kernel, grid, dim3(kPullThreads), lds, ctx->stream, a);
	else
		hipLaunchKernelGGL(deliver_pull_kernel, grid, dim3(kPullThreads), lds, ctx->stream, a);
	SPICE_HIP(hipGetLastError());
	return SPICE_OK;
}

// Launches the blocks of the staged window: per block the caller's neuron
// updates, spike compaction and delivery.  Nothing here synchronises or
// allocates, so a window can be captured into a graph.
int run_blocks(spice_ctx* ctx, int32_t nsteps, int32_t L, bool timed) {
	for (int32_t rel = 0; rel < nsteps; rel += L) {
		int32_t const n = std::min(L, nsteps - rel);
		TimedBlock tb{};
		if (timed) {
			for (auto& e : tb.ev) SPICE_HIP(hipEventCreate(&e));
			SPICE_HIP(hipEventRecord(tb.ev[0], ctx->stream));
		}
		if (rel == 0 && ctx->update_fn) {  // stateless populations: every step of the window
			spice_block const blk{ctx->win_dev, ctx->views_dev, ctx->proj_views_dev, 0, nsteps, SPICE_PHASE_WINDOW};
			if (ctx->update_fn(ctx->update_user, &blk, ctx->stream) != 0)
				return set_error(SPICE_EINVAL, "neuron-update launch failed");
		}
		if (ctx->update_fn) {
			spice_block const blk{ctx->win_dev, ctx->views_dev, ctx->proj_views_dev, rel, n, SPICE_PHASE_NEURONS};
			if (ctx->update_fn(ctx->update_user, &blk, ctx->stream) != 0)
				return set_error(SPICE_EINVAL, "neuron-update launch failed");
		}
		if (timed) SPICE_HIP(hipEventRecord(tb.ev[1], ctx->stream));
		if (!ctx->pops.empty()) {
			hipLaunchKernelGGL(compact_kernel, dim3(unsigned(ctx->pops.size()), unsigned(n)), dim3(kCompactThreads), 0,
			                   ctx->stream, ctx->pops_dev, ctx->win_dev, rel, ctx->ring);
			SPICE_HIP(hipGetLastError());
		}
		if (ctx->any_plastic && ctx->max_pop_size > 0) {
			hipLaunchKernelGGL(history_kernel, dim3(unsigned((ctx->max_pop_size + 255) / 256), unsigned(ctx->pops.size())),
			                   dim3(256), 0, ctx->stream, ctx->pops_dev, ctx->win_dev, rel, n);
			SPICE_HIP(hipGetLastError());
		}
		if (timed) SPICE_HIP(hipEventRecord(tb.ev[2], ctx->stream));
		if (int r = launch_delivery(ctx, rel, n)) return r;
		ctx->last_rel = rel;
		ctx->last_n   = n;
		if (ctx->nstateful > 0) {  // one-step blocks: event lists of the next step
			dim3 const grid(unsigned(ctx->ev_max_chunks), unsigned(ctx->nstateful));
			hipLaunchKernelGGL(event_chunk_sums_kernel, grid, dim3(kEvThreads), 0, ctx->stream, ctx->proj_views_dev,
			                   ctx->stateful_dev, ctx->projs_dev, ctx->win_dev, rel, ctx->ev_sums, ctx->ev_max_chunks);
			hipLaunchKernelGGL(event_offsets_kernel, grid, dim3(kEvThreads), 0, ctx->stream, ctx->proj_views_dev,
			                   ctx->stateful_dev, ctx->projs_dev, ctx->win_dev, rel, ctx->ev_sums, ctx->ev_max_chunks);
			hipLaunchKernelGGL(event_scatter_kernel, dim3(64, unsigned(ctx->nstateful)), dim3(256), 0,
			                   ctx->stream, ctx->proj_views_dev, ctx->stateful_dev, ctx->win_dev, rel);
			SPICE_HIP(hipGetLastError());
		}
		if (ctx->any_plastic && ctx->update_fn) {
			spice_block const blk{ctx->win_dev, ctx->views_dev, ctx->proj_views_dev, rel, n, SPICE_PHASE_SYNAPSES};
			if (ctx->update_fn(ctx->update_user, &blk, ctx->stream) != 0)
				return set_error(SPICE_EINVAL, "synapse-update launch failed");
		}
		if (timed) {
			SPICE_HIP(hipEventRecord(tb.ev[3], ctx->stream));
			ctx->timed.push_back(tb);
		}
	}
	return SPICE_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

char const* spice_last_error(void) { return g_error.c_str(); }

int spice_ctx_create(int device, int64_t max_delay, spice_ctx** out) {
	*out = nullptr;
	if (max_delay < 1) return precondition("neuron_population.h", 108, "max_delay >= 1");
	SPICE_HIP(hipSetDevice(device));
	auto* ctx      = new spice_ctx();
	ctx->device    = device;
	ctx->max_delay = max_delay;
	ctx->ring      = max_delay + SPICE_MAX_BLOCK;
	for (auto& e : ctx->win_copied) {
		if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
			delete ctx;
			return set_error(SPICE_EHIP, "hipEventCreate failed");
		}
	}
	hipError_t e   = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete ctx;
		return set_error(SPICE_EHIP, hipGetErrorString(e));
	}
	// generation temporaries come from the device's stream-ordered pool: keep up
	// to 8 GiB of released blocks cached for the next generation
	hipMemPool_t pool = nullptr;
	if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
		uint64_t keep = uint64_t(8) << 30;
		(void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
	}
	if (int r = dev_alloc(&ctx->error, 1)) return r;
	if (int r = dev_alloc(&ctx->events, 2)) return r;
	*out = ctx;
	return SPICE_OK;
}

void spice_ctx_destroy(spice_ctx* ctx) {
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	for (auto& P : ctx->pops) {
		dfree(ctx->stream, P.state);
		dfree(ctx->stream, P.bitmap);
		dfree(ctx->stream, P.ring_ids);
		dfree(ctx->stream, P.ring_cnt);
		dfree(ctx->stream, P.chunk_jump);
		dfree(ctx->stream, P.hist);
	}
	for (auto& Q : ctx->projs) {
		dfree(ctx->stream, Q.offsets);
		dfree(ctx->stream, Q.neighbors);
		dfree(ctx->stream, Q.counts);
		dfree(ctx->stream, Q.rv);
		dfree(ctx->stream, Q.btgt);
		dfree(ctx->stream, Q.syn_state);
		dfree(ctx->stream, Q.ev_offsets);
		dfree(ctx->stream, Q.ev_cursor);
		dfree(ctx->stream, Q.events);
		dfree(ctx->stream, Q.ages);
	}
	dfree(ctx->stream, ctx->pull_tiles);
	dfree(ctx->stream, ctx->pull_targets);
	for (auto& t : ctx->timed)
		for (auto e : t.ev) (void)hipEventDestroy(e);
	drop_graphs(ctx);
	for (int b = 0; b < 2; b++) {
		(void)hipHostFree(ctx->win_host[b]);
		(void)hipEventDestroy(ctx->win_copied[b]);
	}
	dfree(ctx->stream, ctx->win_steps_dev);
	dfree(ctx->stream, ctx->win_streams_dev);
	dfree(ctx->stream, ctx->win_dev);
	dfree(ctx->stream, ctx->views_dev);
	dfree(ctx->stream, ctx->proj_views_dev);
	dfree(ctx->stream, ctx->stateful_dev);
	dfree(ctx->stream, ctx->ev_sums);
	dfree(ctx->stream, ctx->pops_dev);
	dfree(ctx->stream, ctx->projs_dev);
	dfree(ctx->stream, ctx->error);
	dfree(ctx->stream, ctx->events);
	(void)hipStreamSynchronize(ctx->stream);  // the releases complete before the stream goes
	(void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

int spice_ctx_set_delivery(spice_ctx* ctx, int mode) {
	if (mode != SPICE_DELIVER_PUSH && mode != SPICE_DELIVER_PULL)
		return set_error(SPICE_EINVAL, "unknown delivery mode");
	// push accumulates into counters that the neuron kernel clears; pull overwrites
	// them and leaves them set, so the mode is fixed once steps have run
	if (ctx->steps > 0 && mode != ctx->mode)
		return set_error(SPICE_EINVAL, "the delivery mode can only be changed before the first step");
	ctx->mode      = mode;
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int spice_ctx_set_push_chunks(spice_ctx* ctx, int32_t chunks) {
	if (chunks < 0 || chunks > 64) return set_error(SPICE_EINVAL, "push chunks must be in [0, 64]");
	if (ctx->push_chunks != chunks) drop_graphs(ctx);  // captured launches hold the grid
	ctx->push_chunks = chunks;
	return SPICE_OK;
}

int spice_ctx_set_tile(spice_ctx* ctx, int64_t tile) {
	if (tile < 64 || tile > kMaxPullTile || tile % 64)
		return set_error(SPICE_EINVAL, "tile must be a multiple of 64 in [64, " + std::to_string(kMaxPullTile) + "]");
	ctx->tile      = tile;
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int64_t spice_ctx_tile(spice_ctx* ctx) { return ctx->tile; }

int spice_pop_set_tile(spice_ctx* ctx, int64_t pop, int64_t tile) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "bad population");
	if (tile != 0 && (tile < 64 || tile > kMaxPullTile || tile % 64))
		return set_error(SPICE_EINVAL, "tile must be 0 or a multiple of 64 in [64, " + std::to_string(kMaxPullTile) + "]");
	if (ctx->pops[pop].tile != tile) {
		ctx->pops[pop].tile = tile;
		ctx->dev_dirty      = true;
	}
	return SPICE_OK;
}

int64_t spice_pop_tile(spice_ctx* ctx, int64_t pop) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return -1;
	return pop_tile(ctx, ctx->pops[pop]);
}

int spice_ctx_set_timing(spice_ctx* ctx, int enabled) {
	ctx->timing = enabled != 0;
	return SPICE_OK;
}

void* spice_ctx_stream(spice_ctx* ctx) { return ctx->stream; }

int spice_ctx_sync(spice_ctx* ctx) {
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = flush_fixed_probability(ctx)) return r;
	return check_device_error(ctx);
}

int spice_ctx_stats(spice_ctx* ctx, spice_stats* out) {
	if (int r = spice_ctx_sync(ctx)) return r;
	for (auto& t : ctx->timed) {
		float ms[3] = {0, 0, 0};
		for (int k = 0; k < 3; k++) SPICE_HIP(hipEventElapsedTime(&ms[k], t.ev[k], t.ev[k + 1]));
		ctx->stats.update_ms += ms[0];
		ctx->stats.update_launches++;
		ctx->stats.compact_ms += ms[1];
		ctx->stats.deliver_ms += ms[2];
		ctx->stats.deliver_launches++;
		for (auto e : t.ev) (void)hipEventDestroy(e);
	}
	ctx->timed.clear();
	unsigned long long ev[2] = {0, 0};
	SPICE_HIP(hipMemcpy(ev, ctx->events, sizeof(ev), hipMemcpyDeviceToHost));
	ctx->stats.events = ev[0];
	ctx->stats.rows   = ev[1];
	ctx->stats.steps  = ctx->steps;
	*out              = ctx->stats;
	return SPICE_OK;
}

int spice_instance_create(spice_ctx* ctx, int64_t max_delay, int64_t* out_inst) {
	if (max_delay < 1) return precondition("neuron_population.h", 108, "max_delay >= 1");
	if (max_delay > ctx->max_delay)
		return set_error(SPICE_EINVAL, "an instance's max_delay may not exceed its context's");
	Instance I;
	I.max_delay = max_delay;
	ctx->dev_dirty = true;
	for (size_t i = 0; i < ctx->insts.size(); i++)  // reuse a destroyed instance's slot
		if (!ctx->insts[i].live) {
			ctx->insts[i] = I;
			*out_inst     = int64_t(i);
			return SPICE_OK;
		}
	ctx->insts.push_back(I);
	*out_inst = int64_t(ctx->insts.size()) - 1;
	return SPICE_OK;
}

int spice_instance_destroy(spice_ctx* ctx, int64_t inst) {
	if (inst < 0 || inst >= int64_t(ctx->insts.size()) || !ctx->insts[inst].live)
		return set_error(SPICE_EINVAL, "bad instance");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = flush_fixed_probability(ctx)) return r;
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	drop_graphs(ctx);
	for (auto& Q : ctx->projs) {
		if (Q.dead || ctx->pops[Q.src].inst != inst) continue;
		for (void* p : {static_cast<void*>(Q.offsets), static_cast<void*>(Q.neighbors), static_cast<void*>(Q.counts),
		                static_cast<void*>(Q.rv), static_cast<void*>(Q.btgt), static_cast<void*>(Q.syn_state),
		                static_cast<void*>(Q.ev_offsets), static_cast<void*>(Q.ev_cursor),
		                static_cast<void*>(Q.events), static_cast<void*>(Q.ages)})
			dfree(ctx->stream, p);
		Proj dead;
		dead.dead = true;
		dead.src  = Q.src;
		dead.dst  = Q.dst;
		Q         = dead;
	}
	for (auto& P : ctx->pops) {
		if (P.dead || P.inst != inst) continue;
		for (void* p : {static_cast<void*>(P.state), static_cast<void*>(P.bitmap), static_cast<void*>(P.ring_ids),
		                static_cast<void*>(P.ring_cnt), static_cast<void*>(P.chunk_jump), static_cast<void*>(P.hist)})
			dfree(ctx->stream, p);
		Pop dead;
		dead.dead = true;
		dead.inst = P.inst;
		P         = dead;
	}
	ctx->insts[inst].live = false;
	ctx->dev_dirty        = true;
	return SPICE_OK;
}

int64_t spice_instance_steps(spice_ctx* ctx, int64_t inst) {
	if (inst < 0 || inst >= int64_t(ctx->insts.size())) return -1;
	return ctx->insts[inst].steps;
}

int64_t spice_instance_count(spice_ctx* ctx) { return int64_t(ctx->insts.size()); }

int spice_pop_create_in(spice_ctx* ctx, int64_t inst, int64_t size, int32_t state_words, int32_t draws,
                        int64_t* out_pop) {
	if (inst < 0 || inst >= int64_t(ctx->insts.size()) || !ctx->insts[inst].live)
		return set_error(SPICE_EINVAL, "bad instance");
	if (size < 0) return precondition("neuron_population.h", 35, "size >= 0");
	if (size >= INT32_MAX) return precondition("topology.cpp", 35, "size < INT32_MAX");
	if (state_words < 0 || draws < 0) return set_error(SPICE_EINVAL, "negative state words or draws");
	SPICE_HIP(hipSetDevice(ctx->device));
	Instance& I = ctx->insts[inst];
	Pop P;
	P.inst          = int32_t(inst);
	P.size          = size;
	P.words         = state_words;
	P.draws         = draws;
	P.bitmap_stride = (size + 63) / 64 + 1;  // one spare word per row: scratch for waves past the end
	// stateless populations that draw are updated once per window (every step's spikes at once)
	P.window_bits = state_words == 0 && draws > 0;
	if (int r = dev_alloc(&P.state, size_t(size) * size_t(std::max(state_words, 0)))) return r;
	if (int r = dev_alloc(&P.bitmap, size_t(P.window_bits ? SPICE_MAX_WINDOW : SPICE_MAX_BLOCK) * P.bitmap_stride))
		return r;
	if (int r = dev_alloc(&P.ring_ids, size_t(ctx->ring) * size_t(size), false)) return r;
	if (int r = dev_alloc(&P.ring_cnt, size_t(ctx->ring))) return r;
	if (draws > 0) {
		// per chunk of 16 neurons: x^(offset of the chunk's first draw in the
		// instance's per-step stream) mod P; the instance's populations draw in
		// creation order (reference snn.cpp:12-15)
		constexpr int64_t C   = 16;
		int64_t const nchunks = (size + C - 1) / C;
		std::vector<xoro_state> host(std::max<int64_t>(nchunks, 1));
		gf2poly const step = xoro_jump_poly(uint64_t(C) * uint64_t(draws));
		gf2poly cur        = xoro_jump_poly(I.draws);
		for (int64_t j = 0; j < nchunks; j++) {
			host[j] = {cur.lo, cur.hi};
			cur     = gf2_mulmod(cur, step);
		}
		if (int r = dev_alloc(&P.chunk_jump, host.size(), false)) return r;
		SPICE_HIP(hipMemcpy(P.chunk_jump, host.data(), sizeof(xoro_state) * host.size(),
		                    hipMemcpyHostToDevice));
	}
	I.draws += uint64_t(size) * uint64_t(draws);
	ctx->pops.push_back(P);
	ctx->dev_dirty = true;
	*out_pop       = int64_t(ctx->pops.size()) - 1;
	return SPICE_OK;
}

int spice_pop_create(spice_ctx* ctx, int64_t size, int32_t state_words, int32_t draws,
                     int64_t* out_pop) {
	if (ctx->insts.empty()) {
		int64_t inst = 0;
		if (int r = spice_instance_create(ctx, ctx->max_delay, &inst)) return r;
	}
	return spice_pop_create_in(ctx, 0, size, state_words, draws, out_pop);
}

int64_t spice_pop_size(spice_ctx* ctx, int64_t pop) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return -1;
	return ctx->pops[pop].size;
}

int spice_pop_upload(spice_ctx* ctx, int64_t pop, void const* aos, int32_t struct_bytes) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "bad population");
	Pop& P = ctx->pops[pop];
	if (struct_bytes <= 0 || (struct_bytes + 3) / 4 != P.words)
		return set_error(SPICE_EINVAL, "neuron struct size does not match the population");
	if (P.size == 0) return SPICE_OK;
	SPICE_HIP(hipSetDevice(ctx->device));
	std::vector<uint32_t> soa(size_t(P.words) * P.size, 0);
	auto const* src = static_cast<unsigned char const*>(aos);
	for (int64_t i = 0; i < P.size; i++) {
		uint32_t w[64] = {};
		if (P.words > 64) return set_error(SPICE_EINVAL, "neuron struct larger than 256 bytes");
		std::memcpy(w, src + size_t(i) * struct_bytes, size_t(struct_bytes));
		for (int k = 0; k < P.words; k++) soa[size_t(k) * P.size + i] = w[k];
	}
	SPICE_HIP(hipMemcpyAsync(P.state, soa.data(), sizeof(uint32_t) * soa.size(), hipMemcpyHostToDevice,
	                         ctx->stream));
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	return SPICE_OK;
}

int spice_pop_download(spice_ctx* ctx, int64_t pop, void* aos, int32_t struct_bytes) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "bad population");
	Pop& P = ctx->pops[pop];
	if (struct_bytes <= 0 || (struct_bytes + 3) / 4 != P.words)
		return set_error(SPICE_EINVAL, "neuron struct size does not match the population");
	if (P.size == 0) return SPICE_OK;
	SPICE_HIP(hipSetDevice(ctx->device));
	std::vector<uint32_t> soa(size_t(P.words) * P.size);
	SPICE_HIP(hipMemcpyAsync(soa.data(), P.state, sizeof(uint32_t) * soa.size(), hipMemcpyDeviceToHost,
	                         ctx->stream));
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	auto* dst = static_cast<unsigned char*>(aos);
	for (int64_t i = 0; i < P.size; i++) {
		uint32_t w[64];
		for (int k = 0; k < P.words; k++) w[k] = soa[size_t(k) * P.size + i];
		std::memcpy(dst + size_t(i) * struct_bytes, w, size_t(struct_bytes));
	}
	return SPICE_OK;
}

static int proj_common(spice_ctx* ctx, int64_t src, int64_t dst, int64_t delay, Proj& Q) {
	if (src < 0 || dst < 0 || src >= int64_t(ctx->pops.size()) || dst >= int64_t(ctx->pops.size()) ||
	    ctx->pops[src].dead || ctx->pops[dst].dead)
		return set_error(SPICE_EINVAL, "bad population id");
	if (!(delay >= 1)) return precondition("snn.h", 35, "d >= 1 && \"The delay must be at least 1dt.\"");
	if (!(delay <= ctx->insts[ctx->pops[src].inst].max_delay))
		return precondition("snn.h", 36,
		                    "d <= _max_delay && \"The delay of a synapse population may not exceed "
		                    "the maximum delay of the network.\"");
	if (ctx->pops[dst].words == 0) return set_error(SPICE_EINVAL, "target population has no state");
	if (ctx->pops[dst].incoming.size() >= SPICE_MAX_IN)
		return set_error(SPICE_EINVAL, "too many incoming projections for one population");
	if (ctx->pops[src].inst != ctx->pops[dst].inst)
		return set_error(SPICE_EINVAL, "a projection must connect populations of one instance");
	Q.src   = src;
	Q.dst   = dst;
	Q.delay = delay;
	Q.slots = ctx->ring;
	return dev_alloc(&Q.counts, size_t(Q.slots) * size_t(ctx->pops[dst].size));
}

static int64_t finish_proj(spice_ctx* ctx, Proj& Q) {
	ctx->projs.push_back(Q);
	int64_t const id = int64_t(ctx->projs.size()) - 1;
	ctx->pops[Q.dst].incoming.push_back(id);
	ctx->dev_dirty = true;
	return id;
}

int spice_proj_create_csr(spice_ctx* ctx, int64_t src, int64_t dst, int64_t const* offsets,
                          int32_t const* neighbors, int64_t delay, int64_t* out_proj) {
	SPICE_HIP(hipSetDevice(ctx->device));
	Proj Q;
	if (int r = proj_common(ctx, src, dst, delay, Q)) return r;
	int64_t const S = ctx->pops[src].size, D = ctx->pops[dst].size;
	if (offsets[0] != 0) return set_error(SPICE_EINVAL, "offsets[0] must be 0");
	for (int64_t i = 0; i < S; i++)
		if (offsets[i + 1] < offsets[i]) return precondition("csr.h", 77, "std::is_sorted(_offsets)");
	Q.nnz = offsets[S];
	for (int64_t e = 0; e < Q.nnz; e++)
		if (neighbors[e] < 0 || neighbors[e] >= D)
			return precondition("synapse_population.h", 119, "edge.first < dst_neurons.size()");
	if (int r = dev_alloc(&Q.offsets, size_t(S + 1), false)) return r;
	if (int r = dev_alloc(&Q.neighbors, size_t(Q.nnz), false)) return r;
	SPICE_HIP(hipMemcpy(Q.offsets, offsets, sizeof(int64_t) * (S + 1), hipMemcpyHostToDevice));
	if (Q.nnz) SPICE_HIP(hipMemcpy(Q.neighbors, neighbors, sizeof(int32_t) * Q.nnz, hipMemcpyHostToDevice));
	*out_proj = finish_proj(ctx, Q);
	return SPICE_OK;
}

int spice_proj_create_fixed_probability(spice_ctx* ctx, int64_t src, int64_t dst, double p,
                                        uint64_t seed_lo, uint64_t seed_hi, int64_t delay,
                                        int64_t* out_proj) {
	SPICE_HIP(hipSetDevice(ctx->device));
	if (!(0 <= p && p <= 1)) return precondition("topology.cpp", 73, "0 <= p && p <= 1");
	Proj Q;
	if (int r = proj_common(ctx, src, dst, delay, Q)) return r;
	int64_t const S = ctx->pops[src].size, D = ctx->pops[dst].size;
	if (S == 0) {
		dfree(ctx->stream, Q.counts);
		return precondition("topology.cpp", 106, "offsets.size() > src_count");
	}
	int64_t const id = finish_proj(ctx, Q);
	if (int r = queue_fixed_probability(ctx, id, S, D, p, {seed_lo, seed_hi})) {
		Proj& bad = ctx->projs.back();
		dfree(ctx->stream, bad.counts);
		dfree(ctx->stream, bad.offsets);
		ctx->pops[bad.dst].incoming.pop_back();
		ctx->projs.pop_back();
		return r;
	}
	*out_proj = id;
	return SPICE_OK;
}

int spice_proj_synapses(spice_ctx* ctx, int64_t proj, int64_t* out) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = flush_fixed_probability(ctx)) return r;
	*out = ctx->projs[proj].nnz;
	return SPICE_OK;
}

int spice_proj_download(spice_ctx* ctx, int64_t proj, int64_t* offsets, int32_t* neighbors) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size()) || ctx->projs[proj].dead)
		return set_error(SPICE_EINVAL, "bad projection");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = flush_fixed_probability(ctx)) return r;
	Proj& Q         = ctx->projs[proj];
	int64_t const S = ctx->pops[Q.src].size;
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	if (offsets) SPICE_HIP(hipMemcpy(offsets, Q.offsets, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost));
	if (neighbors && Q.nnz)
		SPICE_HIP(hipMemcpy(neighbors, Q.neighbors, sizeof(int32_t) * Q.nnz, hipMemcpyDeviceToHost));
	return SPICE_OK;
}

__global__ void __launch_bounds__(256)
fill_synapses_kernel(uint32_t* state, int64_t nnz, int32_t words, uint32_t const* one) {
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz; e += stride)
		for (int w = 0; w < words; w++) state[w * nnz + e] = one[w];
}

int spice_proj_set_synapses(spice_ctx* ctx, int64_t proj, void const* aos, int64_t count, int32_t struct_bytes) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	if (struct_bytes <= 0 || struct_bytes > 256) return set_error(SPICE_EINVAL, "synapse struct of 1 to 256 bytes");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = flush_fixed_probability(ctx)) return r;
	Proj& Q = ctx->projs[proj];
	if (count != 1 && count != Q.nnz) return set_error(SPICE_EINVAL, "give one synapse struct or one per synapse");
	if (Q.nnz >= (int64_t(1) << 32)) return set_error(SPICE_EINVAL, "stateful projections hold fewer than 2^32 synapses");
	SPICE_HIP(hipSetDevice(ctx->device));
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	int32_t const words = (struct_bytes + 3) / 4;
	int64_t const D     = ctx->pops[Q.dst].size;
	if (!Q.syn_state) {
		if (int r = dev_alloc(&Q.syn_state, size_t(words) * size_t(std::max<int64_t>(Q.nnz, 1)), false)) return r;
		if (int r = dev_alloc(&Q.ev_offsets, size_t(std::max<int64_t>(D, 1)))) return r;
		if (int r = dev_alloc(&Q.ev_cursor, size_t(std::max<int64_t>(D, 1)))) return r;
		if (int r = dev_alloc(&Q.events, size_t(std::max<int64_t>(Q.nnz, 1)), false)) return r;
	} else if (Q.syn_words != words) {
		return set_error(SPICE_EINVAL, "synapse struct size changed");
	}
	Q.syn_words = words;
	auto const* src = static_cast<unsigned char const*>(aos);
	if (count == 1) {
		uint32_t one[64] = {};
		std::memcpy(one, src, size_t(struct_bytes));
		uint32_t* d_one = nullptr;
		if (int r = dev_alloc(&d_one, 64, false)) return r;
		SPICE_HIP(hipMemcpy(d_one, one, sizeof(one), hipMemcpyHostToDevice));
		if (Q.nnz > 0)
			hipLaunchKernelGGL(fill_synapses_kernel, dim3(unsigned(std::min<int64_t>((Q.nnz + 255) / 256, 65536))),
			                   dim3(256), 0, ctx->stream, Q.syn_state, Q.nnz, words, d_one);
		SPICE_HIP(hipGetLastError());
		SPICE_HIP(hipStreamSynchronize(ctx->stream));
		dfree(ctx->stream, d_one);
	} else if (Q.nnz > 0) {
		std::vector<uint32_t> soa(size_t(words) * size_t(Q.nnz), 0);
		for (int64_t e = 0; e < Q.nnz; e++) {
			uint32_t w[64] = {};
			std::memcpy(w, src + size_t(e) * size_t(struct_bytes), size_t(struct_bytes));
			for (int k = 0; k < words; k++) soa[size_t(k) * size_t(Q.nnz) + size_t(e)] = w[k];
		}
		SPICE_HIP(hipMemcpy(Q.syn_state, soa.data(), sizeof(uint32_t) * soa.size(), hipMemcpyHostToDevice));
	}
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int spice_proj_get_synapses(spice_ctx* ctx, int64_t proj, void* aos, int32_t struct_bytes) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	Proj& Q = ctx->projs[proj];
	if (!Q.syn_state) return set_error(SPICE_EINVAL, "projection without per-synapse state");
	if ((struct_bytes + 3) / 4 != Q.syn_words) return set_error(SPICE_EINVAL, "synapse struct size mismatch");
	SPICE_HIP(hipSetDevice(ctx->device));
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	std::vector<uint32_t> soa(size_t(Q.syn_words) * size_t(std::max<int64_t>(Q.nnz, 1)));
	if (Q.nnz)
		SPICE_HIP(hipMemcpy(soa.data(), Q.syn_state, sizeof(uint32_t) * size_t(Q.syn_words) * size_t(Q.nnz),
		                    hipMemcpyDeviceToHost));
	auto* dst = static_cast<unsigned char*>(aos);
	for (int64_t e = 0; e < Q.nnz; e++) {
		uint32_t w[64];
		for (int k = 0; k < Q.syn_words; k++) w[k] = soa[size_t(k) * size_t(Q.nnz) + size_t(e)];
		std::memcpy(dst + size_t(e) * size_t(struct_bytes), w, size_t(struct_bytes));
	}
	return SPICE_OK;
}

int spice_proj_set_plastic(spice_ctx* ctx, int64_t proj) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	Proj& Q = ctx->projs[proj];
	if (!Q.syn_state) return set_error(SPICE_EINVAL, "plastic synapses need per-synapse state");
	if (Q.plastic) return SPICE_OK;
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = dev_alloc(&Q.ages, size_t(std::max<int64_t>(ctx->pops[Q.src].size, 1)))) return r;
	// the reference keeps the source's history (snn.h:46-47) and reads the
	// target's (synapse_population.h:97-100); both are kept here
	for (int64_t pop : {Q.src, Q.dst}) {
		Pop& P = ctx->pops[pop];
		if (!P.hist)
			if (int r = dev_alloc(&P.hist, size_t(std::max<int64_t>(P.size, 1)))) return r;
	}
	Q.plastic      = true;
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int spice_ctx_proj_views(spice_ctx* ctx, spice_proj_view const** out) {
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = refresh_device_descriptors(ctx)) return r;
	*out = ctx->proj_views_dev;
	return SPICE_OK;
}

int spice_ctx_set_update(spice_ctx* ctx, spice_update_fn fn, void* user) {
	ctx->update_fn   = fn;
	ctx->update_user = user;
	drop_graphs(ctx);
	return SPICE_OK;
}

int spice_ctx_set_graphs(spice_ctx* ctx, int enabled) {
	ctx->graphs = enabled != 0;
	if (!ctx->graphs) drop_graphs(ctx);
	return SPICE_OK;
}

// Size bucket of a target population for tile tuning.
static int tile_bucket(int64_t size) { return size < 4096 ? 0 : size < 16384 ? 1 : size < 65536 ? 2 : 3; }

int spice_ctx_autotune(spice_ctx* ctx, int64_t const* candidates, int32_t ncand, int32_t reps,
                       int64_t* chosen) {
	if (ctx->mode != SPICE_DELIVER_PULL) return set_error(SPICE_EINVAL, "tile tuning needs pull delivery");
	if (ctx->last_rel < 0) return set_error(SPICE_EINVAL, "tile tuning replays the last block: run steps first");
	if (ncand < 1 || reps < 1) return set_error(SPICE_EINVAL, "no tile candidates");
	for (int32_t c = 0; c < ncand; c++)
		if (candidates[c] < 64 || candidates[c] > kMaxPullTile || candidates[c] % 64)
			return set_error(SPICE_EINVAL, "tile candidates must be multiples of 64 in [64, " +
			                                   std::to_string(kMaxPullTile) + "]");
	SPICE_HIP(hipSetDevice(ctx->device));
	// pull delivery overwrites the counts of the block it delivers, so replaying
	// the last block is idempotent; only the event counters are restored
	unsigned long long saved[2];
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	SPICE_HIP(hipMemcpy(saved, ctx->events, sizeof(saved), hipMemcpyDeviceToHost));
	hipEvent_t e0, e1;
	SPICE_HIP(hipEventCreate(&e0));
	SPICE_HIP(hipEventCreate(&e1));
	auto time_tiles = [&](float& ms) -> int {
		if (int r = refresh_device_descriptors(ctx)) return r;
		if (int r = launch_delivery(ctx, ctx->last_rel, ctx->last_n)) return r;  // warm
		SPICE_HIP(hipEventRecord(e0, ctx->stream));
		for (int32_t k = 0; k < reps; k++)
			if (int r = launch_delivery(ctx, ctx->last_rel, ctx->last_n)) return r;
		SPICE_HIP(hipEventRecord(e1, ctx->stream));
		SPICE_HIP(hipEventSynchronize(e1));
		SPICE_HIP(hipEventElapsedTime(&ms, e0, e1));
		return SPICE_OK;
	};
	int r = SPICE_OK;
	for (int b = 0; b < 4 && r == SPICE_OK; b++) {
		chosen[b] = 0;
		bool any  = false;
		for (auto& P : ctx->pops) any = any || (!P.incoming.empty() && tile_bucket(P.size) == b);
		if (!any) continue;
		float best = 0;
		for (int32_t c = 0; c < ncand && r == SPICE_OK; c++) {
			for (auto& P : ctx->pops)
				if (!P.incoming.empty() && tile_bucket(P.size) == b) P.tile = candidates[c];
			ctx->dev_dirty = true;
			float ms       = 0;
			r              = time_tiles(ms);
			if (r == SPICE_OK && (chosen[b] == 0 || ms < best)) {
				best      = ms;
				chosen[b] = candidates[c];
			}
		}
		for (auto& P : ctx->pops)
			if (!P.incoming.empty() && tile_bucket(P.size) == b) P.tile = chosen[b];
		ctx->dev_dirty = true;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (r) return r;
	if (int r2 = refresh_device_descriptors(ctx)) return r2;
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	SPICE_HIP(hipMemcpy(ctx->events, saved, sizeof(saved), hipMemcpyHostToDevice));
	return SPICE_OK;
}

int spice_ctx_views(spice_ctx* ctx, spice_pop_view const** out) {
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = refresh_device_descriptors(ctx)) return r;
	*out = ctx->views_dev;
	return SPICE_OK;
}

int spice_run(spice_ctx* ctx, int32_t nsteps, spice_step const* steps) {
	if (nsteps < 1 || nsteps > SPICE_MAX_WINDOW)
		return set_error(SPICE_EINVAL, "a window has 1 to SPICE_MAX_WINDOW steps");
	if (ctx->insts.empty()) return set_error(SPICE_EINVAL, "the context has no instance");
	if (!ctx->update_fn && !ctx->pops.empty())
		return set_error(SPICE_EINVAL, "no neuron-update callback set (spice_ctx_set_update)");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = refresh_device_descriptors(ctx)) return r;
	int64_t const ninst = int64_t(ctx->insts.size());
	std::vector<int64_t> active(size_t(ninst), 0);
	for (int64_t i = 0; i < ninst; i++) {
		bool on = true;
		for (int k = 0; k < nsteps; k++) {
			spice_step const& st = steps[i * nsteps + k];
			if (st.local < 0) {
				on = false;
				continue;
			}
			if (!on) return set_error(SPICE_EINVAL, "an instance's steps in a window must form a prefix");
			if (st.local != ctx->insts[i].steps + k)
				return set_error(SPICE_EINVAL, "spice_step::local must continue the instance's step count");
			active[i]++;
		}
	}
	// stage the window: pinned buffer b (its previous copy has completed) -> device
	int const b = ctx->win_next;
	ctx->win_next ^= 1;
	SPICE_HIP(hipEventSynchronize(ctx->win_copied[b]));
	spice_step* h = ctx->win_host[b];
	for (int64_t i = 0; i < ninst; i++)
		std::memcpy(h + i * SPICE_MAX_WINDOW, steps + i * nsteps, sizeof(spice_step) * size_t(nsteps));
	for (int64_t i = 0; i < ninst; i++)
		SPICE_HIP(hipMemcpyAsync(ctx->win_steps_dev + i * SPICE_MAX_WINDOW, h + i * SPICE_MAX_WINDOW,
		                         sizeof(spice_step) * size_t(nsteps), hipMemcpyHostToDevice, ctx->stream));
	SPICE_HIP(hipEventRecord(ctx->win_copied[b], ctx->stream));
	if (ctx->any_draws) {  // the window's step streams for the jump tables
		int64_t const n = ninst * nsteps;
		hipLaunchKernelGGL(stream_states_kernel, dim3(unsigned((n + 63) / 64)), dim3(64), 0, ctx->stream,
		                   ctx->win_steps_dev, ctx->win_streams_dev, ninst, nsteps);
		SPICE_HIP(hipGetLastError());
	}

	int32_t L = ctx->nstateful > 0 ? 1 : SPICE_MAX_BLOCK;  // event lists are built step by step
	for (auto& Q : ctx->projs) L = int32_t(std::min<int64_t>(L, Q.delay));
	if (ctx->graphs && !ctx->timing) {
		auto it = ctx->graph_cache.find(nsteps);
		if (it == ctx->graph_cache.end()) {
			SPICE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
			int const r = run_blocks(ctx, nsteps, L, false);
			hipGraph_t g = nullptr;
			hipError_t const e = hipStreamEndCapture(ctx->stream, &g);
			if (r) {
				if (g) (void)hipGraphDestroy(g);
				return r;
			}
			if (e != hipSuccess) return set_error(SPICE_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
			hipGraphExec_t x = nullptr;
			hipError_t const e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			if (e2 != hipSuccess) return set_error(SPICE_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
			it = ctx->graph_cache.emplace(nsteps, x).first;
		}
		SPICE_HIP(hipGraphLaunch(it->second, ctx->stream));
	} else {
		if (int r = run_blocks(ctx, nsteps, L, ctx->timing)) return r;
	}
	for (int64_t i = 0; i < ninst; i++) ctx->insts[i].steps += active[i];
	ctx->steps += nsteps;
	return SPICE_OK;
}

int64_t spice_history_len(spice_ctx* ctx, int64_t pop) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size()) || ctx->pops[pop].dead) return 0;
	Instance const& I = ctx->insts[ctx->pops[pop].inst];
	return std::min<int64_t>(I.steps, I.max_delay);
}

int spice_spikes(spice_ctx* ctx, int64_t pop, int64_t age, int32_t* out, int64_t cap,
                 int64_t* out_count) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size()) || ctx->pops[pop].dead)
		return set_error(SPICE_EINVAL, "bad population");
	if (!(0 <= age && age < spice_history_len(ctx, pop)))
		return precondition("neuron_population.h", 148, "0 <= age && age < _spike_counts.size()");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = check_device_error(ctx)) return r;
	Pop& P             = ctx->pops[pop];
	int64_t const slot = (ctx->insts[P.inst].steps - 1 - age) % ctx->ring;
	int32_t cnt        = 0;
	SPICE_HIP(hipMemcpy(&cnt, P.ring_cnt + slot, sizeof(cnt), hipMemcpyDeviceToHost));
	*out_count = cnt;
	if (out && cnt > 0) {
		if (cap < cnt) return set_error(SPICE_EINVAL, "spike buffer too small");
		SPICE_HIP(hipMemcpy(out, P.ring_ids + slot * P.size, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
	}
	return SPICE_OK;
}

void spice_rng_jump(uint64_t s0, uint64_t s1, uint64_t n, uint64_t* o0, uint64_t* o1) {
	xoro_state const s = xoro_jump_n({s0, s1}, n);
	*o0                = s.s0;
	*o1                = s.s1;
}

void spice_rng_jump_poly(uint64_t n, uint64_t* lo, uint64_t* hi) {
	gf2poly const p = xoro_jump_poly(n);
	*lo             = p.lo;
	*hi             = p.hi;
}

double spice_log_cr(double x) { return spice_crlog::log_cr(x); }

int spice_exp_draw_interval(double u, double scale, double* lo, double* hi) {
	return spice_crlog::exp_draw_interval(u, scale, 0.55, *lo, *hi) ? 1 : 0;
}

}  // extern "C"
