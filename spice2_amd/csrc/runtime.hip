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

int generate_fixed_probability(spice_ctx* ctx, Proj& pr, int64_t src_size, int64_t dst_size,
                               double p, xoro_state seed);  // csr_gen.hip
int build_blocked(spice_ctx* ctx, Proj& Q, int64_t tile);  // blocked.hip
}  // namespace spice_rt

namespace {

constexpr int kWave = 64;

template <class T>
int dev_alloc(T** p, size_t count, bool zero = true) {
	*p = nullptr;
	if (count == 0) count = 1;
	hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
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

// Spike bits of one (population, step) -> ascending id list in the ring slot.
__global__ void __launch_bounds__(256) compact_kernel(PopDev const* pops, int64_t t0, int64_t ring) {
	PopDev const P  = pops[blockIdx.x];
	int const k     = blockIdx.y;
	int64_t const nw = (P.size + 63) / 64;
	uint64_t const* bm = P.bitmap + k * P.bitmap_stride;
	int64_t const slot = (t0 + k) % ring;
	int32_t* out       = P.ring_ids + slot * P.size;

	int const tid      = threadIdx.x;
	int64_t const per  = (nw + blockDim.x - 1) / blockDim.x;
	int64_t const w0   = min<int64_t>(nw, tid * per);
	int64_t const w1   = min<int64_t>(nw, w0 + per);
	int cnt            = 0;
	for (int64_t w = w0; w < w1; w++) cnt += __popcll(bm[w]);

	__shared__ int scan[256];
	scan[tid] = cnt;
	__syncthreads();
	for (int off = 1; off < 256; off <<= 1) {
		int const v = tid >= off ? scan[tid - off] : 0;
		__syncthreads();
		scan[tid] += v;
		__syncthreads();
	}
	int pos = scan[tid] - cnt;
	for (int64_t w = w0; w < w1; w++) {
		uint64_t m = bm[w];
		while (m) {
			int const b = __ffsll((unsigned long long)m) - 1;
			out[pos++]  = int32_t(w * 64 + b);
			m &= m - 1;
		}
	}
	if (tid == 255) P.ring_cnt[slot] = scan[255];
}

// Push delivery: waves walk (projection, step, spike) items; each wave streams
// the spike's CSR row and increments the target's event counter for the
// application step with a device-scope atomic.  Counts are integers, so the
// result does not depend on the order of the atomics.
__global__ void __launch_bounds__(256) deliver_push_kernel(ProjDev const* projs, int nproj,
                                                           int64_t t0, int nsteps, int64_t ring,
                                                           unsigned long long* events) {
	__shared__ int64_t pref[SPICE_MAX_IN * 4 * SPICE_MAX_BLOCK + 1];
	int const nitems = nproj * nsteps;
	if (threadIdx.x == 0) {
		int64_t acc = 0;
		for (int j = 0; j < nitems; j++) {
			int const c = j / nsteps, k = j % nsteps;
			pref[j]     = acc;
			acc += projs[c].src_ring_cnt[(t0 + k) % ring];
		}
		pref[nitems] = acc;
	}
	__syncthreads();
	int64_t const total  = pref[nitems];
	int const lane       = threadIdx.x % kWave;
	int64_t const gwave  = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
	int64_t const nwaves = int64_t(gridDim.x) * blockDim.x / kWave;
	unsigned long long ev = 0, rows = 0;
	for (int64_t item = gwave; item < total; item += nwaves) {
		int lo = 0, hi = nitems;  // pref[lo] <= item < pref[hi]
		while (hi - lo > 1) {
			int const mid = (lo + hi) / 2;
			if (pref[mid] <= item)
				lo = mid;
			else
				hi = mid;
		}
		int const c = lo / nsteps, k = lo % nsteps;
		ProjDev const& P  = projs[c];
		int64_t const idx = item - pref[lo];
		int32_t const src = P.src_ring_ids[((t0 + k) % ring) * P.src_size + idx];
		int64_t const a = P.offsets[src], b = P.offsets[src + 1];
		uint32_t* cnt   = P.counts + ((t0 + k + P.delay) % P.slots) * P.dst_size;
		for (int64_t e = a + lane; e < b; e += kWave) atomicAdd(cnt + P.neighbors[e], 1u);
		ev += uint64_t(b - a);
		rows++;
	}
	if (lane == 0 && rows) {
		atomicAdd(events, ev);
		atomicAdd(events + 1, rows);
	}
}

// Pull delivery: one workgroup per (target tile, step).  The workgroup owns the
// event counters of its tile for every incoming projection in LDS, streams the
// tile's run of each spiking source's row (16-bit local targets, 16-byte vector
// loads, spread over the lanes by a wave-wide scan of the run lengths), counts
// with LDS atomics and writes its counters once, without global atomics.
// Counts are integers, so the result is independent of the order of the atomics
// and equals the push result exactly.
// Workgroup L handles tile L / 16 and step L % 16: the steps of one block land on
// fixed XCDs (blocks are dealt to the 8 XCDs round-robin), so the spike lists and
// row offsets a step's workgroups all read are fetched into one XCD's L2.
struct PullArgs {
	ProjDev const* projs;
	PullTile const* tiles;
	PullTarget const* targets;
	int64_t t0;
	int32_t nsteps;
	int64_t ring;
	int64_t tile;
	unsigned long long* events;
};

__global__ void __launch_bounds__(512) deliver_pull_kernel(PullArgs a) {
	extern __shared__ uint32_t lds_cnt[];  // [nin][tile]
	int const k = int(blockIdx.x % SPICE_MAX_BLOCK);
	if (k >= a.nsteps) return;
	PullTile const pt   = a.tiles[blockIdx.x / SPICE_MAX_BLOCK];
	PullTarget const& tg = a.targets[pt.target];
	int64_t const step  = a.t0 + k;
	int64_t const T0    = int64_t(pt.tau) * a.tile;
	int64_t const Tn    = min<int64_t>(a.tile, tg.dst_size - T0);
	int const tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave, nwaves = blockDim.x / kWave;
	for (int64_t i = tid; i < int64_t(tg.nin) * a.tile; i += blockDim.x) lds_cnt[i] = 0;
	__syncthreads();
	unsigned long long ev = 0, rows = 0;
	int64_t const slot_in = step % a.ring;
	for (int j = 0; j < tg.nin; j++) {
		ProjDev const& P      = a.projs[tg.proj[j]];
		int32_t const nsp     = P.src_ring_cnt[slot_in];
		int32_t const* ids    = P.src_ring_ids + slot_in * P.src_size;
		uint32_t* cnt         = lds_cnt + int64_t(j) * a.tile;
		int64_t const rtw     = P.ntile + 1;
		if (pt.tau == 0 && tid == 0) rows += uint64_t(nsp);
		for (int64_t b0 = int64_t(wave) * kWave; b0 < nsp; b0 += int64_t(nwaves) * kWave) {
			int64_t const sp = b0 + lane;
			uint64_t lo = 0, hi = 0;
			if (sp < nsp) {
				int64_t const src  = ids[sp];
				int64_t const base = P.offsets[src];
				lo                 = uint64_t(base) + P.rt[src * rtw + pt.tau];
				hi                 = uint64_t(base) + P.rt[src * rtw + pt.tau + 1];
			}
			ev += hi - lo;
			uint64_t const v0 = lo >> 3;
			uint32_t const nv = hi > lo ? uint32_t(((hi + 7) >> 3) - v0) : 0u;
			uint32_t incl     = nv;
			for (int off = 1; off < kWave; off <<= 1) {
				uint32_t const t = __shfl_up(incl, off, kWave);
				if (lane >= off) incl += t;
			}
			uint32_t const total = __shfl(incl, kWave - 1, kWave);
			uint32_t const excl  = incl - nv;
			// wave-uniform loop: every lane takes part in the shuffles, also lanes
			// past the last vector (a shuffle from an inactive lane is undefined)
			for (uint32_t vb = 0; vb < total; vb += kWave) {
				uint32_t const v = vb + lane;
				int q = 0;  // first lane whose inclusive count exceeds v
				for (int s = 32; s > 0; s >>= 1)
					if (__shfl(incl, q + s - 1, kWave) <= v) q += s;
				q                  = min(q, kWave - 1);
				uint64_t const vq  = __shfl(v0, q, kWave) + (v - __shfl(excl, q, kWave));
				uint64_t const lq  = __shfl(lo, q, kWave);
				uint64_t const hq  = __shfl(hi, q, kWave);
				if (v < total) {
					uint4 const w           = *reinterpret_cast<uint4 const*>(P.btgt + vq * 8);
					uint32_t const words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
					for (int e = 0; e < 8; e++) {
						uint64_t const idx = vq * 8 + e;
						if (idx >= lq && idx < hq)
							atomicAdd(&cnt[(words[e >> 1] >> (16 * (e & 1))) & 0xffffu], 1u);
					}
				}
			}
		}
	}
	__syncthreads();
	for (int j = 0; j < tg.nin; j++) {
		ProjDev const& P = a.projs[tg.proj[j]];
		uint32_t* out    = P.counts + ((step + P.delay) % P.slots) * P.dst_size + T0;
		uint32_t const* c = lds_cnt + int64_t(j) * a.tile;
		for (int64_t i = tid; i < Tn; i += blockDim.x) out[i] = c[i];
	}
	for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off, kWave);
	if (lane == 0 && ev) atomicAdd(a.events, ev);
	if (tid == 0 && rows) atomicAdd(a.events + 1, rows);
}

int refresh_device_descriptors(spice_ctx* ctx) {
	if (!ctx->dev_dirty) return SPICE_OK;
	if (ctx->mode == SPICE_DELIVER_PULL) {
		// tile size: counters of all incoming projections of a tile fit 64 KB of LDS
		int64_t max_in = 1;
		for (auto& P : ctx->pops) max_in = std::max<int64_t>(max_in, int64_t(P.incoming.size()));
		int64_t tile = std::min<int64_t>(ctx->tile, 65536 / (4 * max_in) / 64 * 64);
		tile         = std::max<int64_t>(tile, 64);
		ctx->tile    = tile;
		for (auto& Q : ctx->projs)
			if (Q.tile != tile)
				if (int r = build_blocked(ctx, Q, tile)) return r;
	}
	std::vector<PopDev> pd;
	for (auto& P : ctx->pops)
		pd.push_back({P.bitmap, P.bitmap_stride, P.size, P.ring_ids, P.ring_cnt});
	std::vector<ProjDev> qd;
	for (auto& Q : ctx->projs) {
		Pop const& S = ctx->pops[Q.src];
		qd.push_back({Q.offsets, Q.neighbors, Q.counts, Q.slots, ctx->pops[Q.dst].size, Q.delay,
		              S.ring_ids, S.ring_cnt, S.size, Q.rt, Q.btgt, Q.ntile});
	}
	// pull tiles: every population with incoming projections, split into tiles
	std::vector<PullTarget> pt;
	std::vector<PullTile> tiles;
	ctx->pull_max_in = 0;
	for (auto& P : ctx->pops) {
		if (P.incoming.empty() || P.size == 0) continue;
		PullTarget t{};
		t.dst_size = P.size;
		t.nin      = int32_t(P.incoming.size());
		for (size_t j = 0; j < P.incoming.size(); j++) t.proj[j] = int32_t(P.incoming[j]);
		int32_t const id = int32_t(pt.size());
		pt.push_back(t);
		ctx->pull_max_in = std::max(ctx->pull_max_in, t.nin);
		for (int64_t tau = 0; tau * ctx->tile < P.size; tau++) tiles.push_back({id, int32_t(tau)});
	}
	if (ctx->pull_tiles) (void)hipFree(ctx->pull_tiles);
	if (ctx->pull_targets) (void)hipFree(ctx->pull_targets);
	ctx->pull_tiles   = nullptr;
	ctx->pull_targets = nullptr;
	ctx->pull_ntiles  = int64_t(tiles.size());
	if (int r = dev_alloc(&ctx->pull_tiles, tiles.size(), false)) return r;
	if (int r = dev_alloc(&ctx->pull_targets, pt.size(), false)) return r;
	if (!tiles.empty())
		SPICE_HIP(hipMemcpy(ctx->pull_tiles, tiles.data(), sizeof(PullTile) * tiles.size(), hipMemcpyHostToDevice));
	if (!pt.empty())
		SPICE_HIP(hipMemcpy(ctx->pull_targets, pt.data(), sizeof(PullTarget) * pt.size(), hipMemcpyHostToDevice));
	if (ctx->pops_dev) (void)hipFree(ctx->pops_dev);
	if (ctx->projs_dev) (void)hipFree(ctx->projs_dev);
	ctx->pops_dev  = nullptr;
	ctx->projs_dev = nullptr;
	if (int r = dev_alloc(&ctx->pops_dev, pd.size(), false)) return r;
	if (int r = dev_alloc(&ctx->projs_dev, qd.size(), false)) return r;
	if (!pd.empty())
		SPICE_HIP(hipMemcpy(ctx->pops_dev, pd.data(), sizeof(PopDev) * pd.size(), hipMemcpyHostToDevice));
	if (!qd.empty())
		SPICE_HIP(hipMemcpy(ctx->projs_dev, qd.data(), sizeof(ProjDev) * qd.size(), hipMemcpyHostToDevice));
	// stream offsets: population p's first draw follows all earlier populations' draws
	uint64_t base = 0;
	for (auto& P : ctx->pops) {
		P.base_poly = xoro_jump_poly(base);
		base += uint64_t(P.size) * uint64_t(P.draws);
	}
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
	hipError_t e   = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete ctx;
		return set_error(SPICE_EHIP, hipGetErrorString(e));
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
		(void)hipFree(P.state);
		(void)hipFree(P.bitmap);
		(void)hipFree(P.ring_ids);
		(void)hipFree(P.ring_cnt);
		(void)hipFree(P.chunk_jump);
	}
	for (auto& Q : ctx->projs) {
		(void)hipFree(Q.offsets);
		(void)hipFree(Q.neighbors);
		(void)hipFree(Q.counts);
		(void)hipFree(Q.rt);
		(void)hipFree(Q.btgt);
	}
	(void)hipFree(ctx->pull_tiles);
	(void)hipFree(ctx->pull_targets);
	for (auto& t : ctx->timed)
		for (auto e : t.ev) (void)hipEventDestroy(e);
	if (ctx->upd_begin) (void)hipEventDestroy(ctx->upd_begin);
	(void)hipFree(ctx->pops_dev);
	(void)hipFree(ctx->projs_dev);
	(void)hipFree(ctx->error);
	(void)hipFree(ctx->events);
	(void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

int spice_ctx_set_delivery(spice_ctx* ctx, int mode) {
	if (mode != SPICE_DELIVER_PUSH && mode != SPICE_DELIVER_PULL)
		return set_error(SPICE_EINVAL, "unknown delivery mode");
	if (ctx->pending) return set_error(SPICE_EINVAL, "cannot change delivery mode inside a block");
	// push accumulates into counters that the neuron kernel clears; pull overwrites
	// them and leaves them set, so the mode is fixed once steps have run
	if (ctx->steps > 0 && mode != ctx->mode)
		return set_error(SPICE_EINVAL, "the delivery mode can only be changed before the first step");
	ctx->mode      = mode;
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int spice_ctx_set_tile(spice_ctx* ctx, int64_t tile) {
	if (tile < 64 || tile > 65536 || tile % 64) return set_error(SPICE_EINVAL, "tile must be a multiple of 64 in [64, 65536]");
	if (ctx->pending) return set_error(SPICE_EINVAL, "cannot change the tile inside a block");
	ctx->tile      = tile;
	ctx->dev_dirty = true;
	return SPICE_OK;
}

int64_t spice_ctx_tile(spice_ctx* ctx) { return ctx->tile; }

int spice_ctx_set_timing(spice_ctx* ctx, int enabled) {
	ctx->timing = enabled != 0;
	return SPICE_OK;
}

void* spice_ctx_stream(spice_ctx* ctx) { return ctx->stream; }

int spice_ctx_sync(spice_ctx* ctx) {
	SPICE_HIP(hipSetDevice(ctx->device));
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

int spice_pop_create(spice_ctx* ctx, int64_t size, int32_t state_words, int32_t draws,
                     int64_t* out_pop) {
	if (size < 0) return precondition("neuron_population.h", 35, "size >= 0");
	if (size >= INT32_MAX) return precondition("topology.cpp", 35, "size < INT32_MAX");
	if (state_words < 0 || draws < 0) return set_error(SPICE_EINVAL, "negative state words or draws");
	if (ctx->pending) return set_error(SPICE_EINVAL, "cannot add a population inside a block");
	SPICE_HIP(hipSetDevice(ctx->device));
	Pop P;
	P.size          = size;
	P.words         = state_words;
	P.draws         = draws;
	P.bitmap_stride = (size + 63) / 64;
	if (int r = dev_alloc(&P.state, size_t(size) * size_t(std::max(state_words, 0)))) return r;
	if (int r = dev_alloc(&P.bitmap, size_t(SPICE_MAX_BLOCK) * P.bitmap_stride)) return r;
	if (int r = dev_alloc(&P.ring_ids, size_t(ctx->ring) * size_t(size), false)) return r;
	if (int r = dev_alloc(&P.ring_cnt, size_t(ctx->ring))) return r;
	if (draws > 0) {
		int64_t const nchunks = (size + P.chunk - 1) / P.chunk;
		std::vector<xoro_state> host(std::max<int64_t>(nchunks, 1));
		gf2poly const step = xoro_jump_poly(uint64_t(P.chunk) * uint64_t(draws));
		gf2poly cur{1, 0};
		for (int64_t j = 0; j < nchunks; j++) {
			host[j] = {cur.lo, cur.hi};
			cur     = gf2_mulmod(cur, step);
		}
		if (int r = dev_alloc(&P.chunk_jump, host.size(), false)) return r;
		SPICE_HIP(hipMemcpy(P.chunk_jump, host.data(), sizeof(xoro_state) * host.size(),
		                    hipMemcpyHostToDevice));
	}
	ctx->pops.push_back(P);
	ctx->dev_dirty = true;
	*out_pop       = int64_t(ctx->pops.size()) - 1;
	return SPICE_OK;
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
	if (src < 0 || dst < 0 || src >= int64_t(ctx->pops.size()) || dst >= int64_t(ctx->pops.size()))
		return set_error(SPICE_EINVAL, "bad population id");
	if (!(delay >= 1)) return precondition("snn.h", 35, "d >= 1 && \"The delay must be at least 1dt.\"");
	if (!(delay <= ctx->max_delay))
		return precondition("snn.h", 36,
		                    "d <= _max_delay && \"The delay of a synapse population may not exceed "
		                    "the maximum delay of the network.\"");
	if (ctx->pops[dst].words == 0) return set_error(SPICE_EINVAL, "target population has no state");
	if (ctx->pops[dst].incoming.size() >= SPICE_MAX_IN)
		return set_error(SPICE_EINVAL, "too many incoming projections for one population");
	if (ctx->pending) return set_error(SPICE_EINVAL, "cannot connect inside a block");
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
		(void)hipFree(Q.counts);
		return precondition("topology.cpp", 106, "offsets.size() > src_count");
	}
	if (int r = generate_fixed_probability(ctx, Q, S, D, p, {seed_lo, seed_hi})) {
		(void)hipFree(Q.counts);
		return r;
	}
	*out_proj = finish_proj(ctx, Q);
	return SPICE_OK;
}

int spice_proj_synapses(spice_ctx* ctx, int64_t proj, int64_t* out) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	*out = ctx->projs[proj].nnz;
	return SPICE_OK;
}

int spice_proj_download(spice_ctx* ctx, int64_t proj, int64_t* offsets, int32_t* neighbors) {
	if (proj < 0 || proj >= int64_t(ctx->projs.size())) return set_error(SPICE_EINVAL, "bad projection");
	SPICE_HIP(hipSetDevice(ctx->device));
	Proj& Q         = ctx->projs[proj];
	int64_t const S = ctx->pops[Q.src].size;
	SPICE_HIP(hipStreamSynchronize(ctx->stream));
	if (offsets) SPICE_HIP(hipMemcpy(offsets, Q.offsets, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost));
	if (neighbors && Q.nnz)
		SPICE_HIP(hipMemcpy(neighbors, Q.neighbors, sizeof(int32_t) * Q.nnz, hipMemcpyDeviceToHost));
	return SPICE_OK;
}

int spice_block_begin(spice_ctx* ctx, int32_t nsteps, float const* dt, uint64_t const* seeds,
                      spice_block_view* block, spice_pop_view* views, int32_t max_pops) {
	if (ctx->pending) return set_error(SPICE_EINVAL, "a block is already open");
	int32_t limit = SPICE_MAX_BLOCK;
	for (auto& Q : ctx->projs) limit = int32_t(std::min<int64_t>(limit, Q.delay));
	if (nsteps < 1 || nsteps > limit)
		return set_error(SPICE_EINVAL, "block length must be in [1, min(16, smallest delay)]");
	if (max_pops < int32_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "view array too small");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = refresh_device_descriptors(ctx)) return r;

	block->t0     = ctx->steps;
	block->nsteps = nsteps;
	for (int k = 0; k < SPICE_MAX_BLOCK; k++) block->dt[k] = k < nsteps ? dt[k] : 0.0f;
	for (size_t p = 0; p < ctx->pops.size(); p++) {
		Pop& P           = ctx->pops[p];
		spice_pop_view& v = views[p];
		std::memset(&v, 0, sizeof(v));
		v.state         = P.state;
		v.size          = P.size;
		v.words         = P.words;
		v.draws         = P.draws;
		v.bitmap        = P.bitmap;
		v.bitmap_stride = P.bitmap_stride;
		v.chunk_jump    = reinterpret_cast<spice_rng_state const*>(P.chunk_jump);
		v.chunk         = P.chunk;
		v.nin           = int32_t(P.incoming.size());
		for (size_t j = 0; j < P.incoming.size(); j++) {
			Proj& Q     = ctx->projs[P.incoming[j]];
			v.in[j]     = {Q.counts, Q.slots, P.incoming[j]};
		}
		for (int k = 0; k < nsteps; k++) {
			xoro_state const s = xoro_jump({seeds[2 * k], seeds[2 * k + 1]}, P.base_poly);
			v.rng[k]           = {s.s0, s.s1};
		}
		v.error        = ctx->error;
		v.clear_counts = ctx->mode == SPICE_DELIVER_PUSH ? 1 : 0;
	}
	ctx->pending    = nsteps;
	ctx->pending_t0 = ctx->steps;
	if (ctx->timing) {
		SPICE_HIP(hipEventCreate(&ctx->upd_begin));
		SPICE_HIP(hipEventRecord(ctx->upd_begin, ctx->stream));
	}
	return SPICE_OK;
}

int spice_block_end(spice_ctx* ctx) {
	if (!ctx->pending) return set_error(SPICE_EINVAL, "no open block");
	SPICE_HIP(hipSetDevice(ctx->device));
	int const n      = ctx->pending;
	int64_t const t0 = ctx->pending_t0;
	TimedBlock tb{};
	bool const timed = ctx->timing && ctx->upd_begin;
	if (timed) {
		tb.ev[0]       = ctx->upd_begin;
		ctx->upd_begin = nullptr;
		for (int k = 1; k < 4; k++) SPICE_HIP(hipEventCreate(&tb.ev[k]));
		SPICE_HIP(hipEventRecord(tb.ev[1], ctx->stream));
	}
	if (!ctx->pops.empty()) {
		hipLaunchKernelGGL(compact_kernel, dim3(unsigned(ctx->pops.size()), unsigned(n)), dim3(256), 0,
		                   ctx->stream, ctx->pops_dev, t0, ctx->ring);
		SPICE_HIP(hipGetLastError());
	}
	if (timed) SPICE_HIP(hipEventRecord(tb.ev[2], ctx->stream));
	if (!ctx->projs.empty() && ctx->mode == SPICE_DELIVER_PUSH) {
		if (ctx->projs.size() * size_t(n) > SPICE_MAX_IN * 4 * SPICE_MAX_BLOCK)
			return set_error(SPICE_EINVAL, "too many projections for one delivery launch");
		hipLaunchKernelGGL(deliver_push_kernel, dim3(2048), dim3(256), 0, ctx->stream, ctx->projs_dev,
		                   int(ctx->projs.size()), t0, n, ctx->ring, ctx->events);
		SPICE_HIP(hipGetLastError());
	} else if (!ctx->projs.empty() && ctx->pull_ntiles > 0) {
		PullArgs a{ctx->projs_dev, ctx->pull_tiles, ctx->pull_targets, t0, n, ctx->ring, ctx->tile,
		           ctx->events};
		size_t const lds = sizeof(uint32_t) * size_t(ctx->pull_max_in) * size_t(ctx->tile);
		hipLaunchKernelGGL(deliver_pull_kernel, dim3(unsigned(ctx->pull_ntiles * SPICE_MAX_BLOCK)), dim3(512),
		                   lds, ctx->stream, a);
		SPICE_HIP(hipGetLastError());
	}
	if (timed) {
		SPICE_HIP(hipEventRecord(tb.ev[3], ctx->stream));
		ctx->timed.push_back(tb);
	}
	ctx->steps += n;
	ctx->pending = 0;
	return SPICE_OK;
}

int64_t spice_history_len(spice_ctx* ctx) { return std::min<int64_t>(ctx->steps, ctx->max_delay); }

int64_t spice_steps(spice_ctx* ctx) { return ctx->steps; }

int spice_pop_view_now(spice_ctx* ctx, int64_t pop, spice_pop_view* v) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "bad population");
	if (ctx->pending) return set_error(SPICE_EINVAL, "block still open");
	Pop& P = ctx->pops[pop];
	std::memset(v, 0, sizeof(*v));
	v->state         = P.state;
	v->size          = P.size;
	v->words         = P.words;
	v->draws         = P.draws;
	v->bitmap        = P.bitmap;
	v->bitmap_stride = P.bitmap_stride;
	v->chunk_jump    = reinterpret_cast<spice_rng_state const*>(P.chunk_jump);
	v->chunk         = P.chunk;
	v->nin           = int32_t(P.incoming.size());
	for (size_t j = 0; j < P.incoming.size(); j++) {
		Proj& Q = ctx->projs[P.incoming[j]];
		v->in[j] = {Q.counts, Q.slots, P.incoming[j]};
	}
	v->error        = ctx->error;
	v->clear_counts = 1;  // deliveries applied out of band must not be applied again
	return SPICE_OK;
}

int spice_spikes(spice_ctx* ctx, int64_t pop, int64_t age, int32_t* out, int64_t cap,
                 int64_t* out_count) {
	if (pop < 0 || pop >= int64_t(ctx->pops.size())) return set_error(SPICE_EINVAL, "bad population");
	if (ctx->pending) return set_error(SPICE_EINVAL, "block still open");
	if (!(0 <= age && age < spice_history_len(ctx)))
		return precondition("neuron_population.h", 148, "0 <= age && age < _spike_counts.size()");
	SPICE_HIP(hipSetDevice(ctx->device));
	if (int r = check_device_error(ctx)) return r;
	Pop& P             = ctx->pops[pop];
	int64_t const slot = (ctx->steps - 1 - age) % ctx->ring;
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
