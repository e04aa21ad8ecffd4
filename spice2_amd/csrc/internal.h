// Internal state of a spice_ctx shared by the runtime translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/spice_hip.h"
#include "rng.h"

namespace spice_rt {

// Sets the thread's last error and returns `code`.
int set_error(int code, std::string const& msg);

// Device memory of a context comes from the device's stream-ordered pool: it is
// allocated in order on the null stream and waited for (so any stream may use
// it at once), and released in order on the context's stream, so that a release
// waits for the context's own work only, not for every stream of the device.
inline hipError_t dmalloc(void** p, size_t bytes) {
	hipError_t e = hipMallocAsync(p, bytes, nullptr);
	if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
	return e;
}
inline void dfree(hipStream_t s, void* p) {
	if (p) (void)hipFreeAsync(p, s);
}

// Largest pull tile: one projection's 32-bit counters for the tile must fit in
// the LDS of a compute unit beside the run descriptors.
constexpr int64_t kMaxPullTile = 32768;

#define SPICE_HIP(expr)                                                                    \
	do {                                                                                   \
		hipError_t const e_ = (expr);                                                      \
		if (e_ != hipSuccess)                                                              \
			return ::spice_rt::set_error(SPICE_EHIP, std::string(#expr) + ": " +            \
			                                             hipGetErrorString(e_));            \
	} while (0)

// An independent network inside a context: its own step count (the ring
// slots of its populations are indexed by it), spike-history bound and RNG
// stream offsets.
struct Instance {
	int64_t max_delay = 1;
	int64_t steps     = 0;  // steps run
	uint64_t draws    = 0;  // stream outputs taken per step by its populations so far
	bool live         = true;
};

struct Pop {
	bool dead      = false;  // its instance was destroyed; the id is not reused
	int32_t inst   = 0;
	int64_t tile   = 0;  // pull tile of its incoming projections (0: the context's)
	int64_t size   = 0;
	int32_t words  = 0;
	int32_t draws  = 0;
	uint32_t* state = nullptr;     // [words][size]
	uint64_t* bitmap = nullptr;    // [bitmap rows][bitmap_stride]
	int64_t bitmap_stride = 0;
	bool window_bits = false;      // stateless and draws: one bitmap row per window step
	int32_t* ring_ids = nullptr;   // [ring][size]
	int32_t* ring_cnt = nullptr;   // [ring]
	xoro_state* chunk_jump = nullptr; // per 16 neurons: x^(stream offset in the instance) mod P
	uint64_t* hist = nullptr;      // spike history (populations with plastic projections)
	std::vector<int64_t> incoming; // projection ids, connect order
};

struct Proj {
	bool dead = false;
	int64_t src = 0, dst = 0;
	int64_t delay = 1;
	int64_t* offsets = nullptr;    // [src_size + 1]
	int32_t* neighbors = nullptr;  // [nnz]
	int64_t nnz = 0;
	uint32_t* counts = nullptr;    // [slots][dst_size]
	int64_t slots = 0;
	// Tile view for pull delivery (blocked.hip): 16-bit tile-local targets in
	// runs padded to whole 16-byte vectors and, per row, the first vector of
	// each tile's run.
	int64_t tile = 0;              // 0: not built
	int64_t ntile = 0;
	int64_t nvec = 0;              // 16-byte vectors of btgt (one spare vector follows)
	uint32_t* rv = nullptr;        // [src_size][ntile + 1]
	uint16_t* btgt = nullptr;      // [nvec + 1][8]
	// per-synapse state and event lists (stateful synapses)
	int32_t syn_words = 0;
	uint32_t* syn_state = nullptr; // [syn_words][nnz]
	uint32_t* ev_offsets = nullptr;// [dst_size]
	uint32_t* ev_cursor = nullptr; // [dst_size]
	uint32_t* events = nullptr;    // [nnz]
	bool plastic = false;
	uint64_t* ages = nullptr;      // [src_size]
};

// A fixed_probability projection waiting for generation (projections are
// generated together, see flush_fixed_probability).
struct PendingFp {
	int64_t proj;
	int64_t S, D;
	double p;
	xoro_state seed;
};

// Device-side descriptors, rebuilt when populations/projections change.
struct PopDev {
	int64_t inst;
	uint64_t const* bitmap;
	int64_t bitmap_stride;
	int64_t window_bits;  // rows indexed by the window step (else by the block step)
	int64_t size;
	int32_t* ring_ids;
	int32_t* ring_cnt;
	uint64_t* hist;
};

struct ProjDev {
	int64_t inst;
	int64_t const* offsets;
	int32_t const* neighbors;
	uint32_t* counts;
	int64_t slots;
	int64_t dst_size;
	int64_t delay;
	int32_t const* src_ring_ids;
	int32_t const* src_ring_cnt;
	int64_t src_size;
	uint32_t const* rv;
	uint16_t const* btgt;
	int64_t ntile;
	uint32_t* ev_offsets;
};

// Pull delivery: one workgroup per (target tile, step).
struct PullTile {
	int32_t target;  // index into PullTarget array
	int32_t tau;     // tile within the target population
};
struct PullTarget {
	int64_t inst;
	int64_t tile;
	int64_t dst_size;
	int32_t nin;
	int32_t proj[SPICE_MAX_IN];  // projection ids, connect order
};

// Events of one timed block: [0] update start, [1] update end = compaction
// start, [2] compaction end = delivery start, [3] delivery end.
struct TimedBlock {
	hipEvent_t ev[4];
};

}  // namespace spice_rt

struct spice_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	int64_t max_delay = 1;
	int64_t ring = 0;  // spike-ring and count-ring slots
	int64_t steps = 0; // steps run (windows), all instances
	int mode = SPICE_DELIVER_PULL;
	bool timing = false;
	std::vector<spice_rt::Instance> insts;
	std::vector<spice_rt::Pop> pops;
	std::vector<spice_rt::Proj> projs;
	spice_rt::PopDev* pops_dev = nullptr;
	spice_rt::ProjDev* projs_dev = nullptr;
	spice_rt::PullTile* pull_tiles = nullptr;
	spice_rt::PullTarget* pull_targets = nullptr;
	int64_t pull_ntiles = 0;
	int32_t pull_pass = 1;  // projections a pull workgroup counts together
	int64_t pull_max_tile = 64;
	int32_t last_rel = -1, last_n = 0;  // last block run (autotuning replays its delivery)
	int64_t tile = 8192;  // targets per pull tile
	int32_t push_chunks = 0;  // push: spike chunks per (tile, step) (0: one)
	bool dev_dirty = true;
	int32_t* error = nullptr;                 // device flag
	unsigned long long* events = nullptr;     // device counters: events, rows
	spice_pop_view* views_dev = nullptr;    // [pops]
	spice_proj_view* proj_views_dev = nullptr; // [projs]
	int64_t* stateful_dev = nullptr;        // ids of projections with per-synapse state
	int64_t nstateful = 0;
	bool any_draws = false;                 // a population draws: window stream states are kept
	spice_rng_state* win_streams_dev = nullptr;  // [win_streams_cap][SPICE_MAX_WINDOW][SPICE_STREAM_STATES]
	int64_t win_streams_cap = 0;
	uint32_t* ev_sums = nullptr;            // [nstateful][ev_max_chunks] event-list scan chunk sums
	int64_t ev_sums_cap = 0, ev_max_chunks = 1;
	bool any_plastic = false;
	std::vector<spice_rt::PendingFp> pending_fp;  // projections to generate
	int64_t max_pop_size = 0;
	// the running window: device table [insts][SPICE_MAX_WINDOW] filled from one
	// of two pinned staging buffers (the other may still be in flight)
	spice_window* win_dev = nullptr;
	spice_step* win_steps_dev = nullptr;
	int64_t win_cap = 0;                    // instances the tables hold
	spice_step* win_host[2] = {nullptr, nullptr};
	hipEvent_t win_copied[2] = {nullptr, nullptr};
	int win_next = 0;
	// neuron updates: launched by the caller's callback once per block
	spice_update_fn update_fn = nullptr;
	void* update_user = nullptr;
	// graphs of whole windows, by window length; dropped when the network changes
	bool graphs = false;
	std::map<int32_t, hipGraphExec_t> graph_cache;
	std::vector<spice_rt::TimedBlock> timed;
	spice_stats stats{};
};
