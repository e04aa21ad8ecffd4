// Internal state of a spice_ctx shared by the runtime translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spice_hip.h"
#include "rng.h"

namespace spice_rt {

// Sets the thread's last error and returns `code`.
int set_error(int code, std::string const& msg);

#define SPICE_HIP(expr)                                                                    \
	do {                                                                                   \
		hipError_t const e_ = (expr);                                                      \
		if (e_ != hipSuccess)                                                              \
			return ::spice_rt::set_error(SPICE_EHIP, std::string(#expr) + ": " +            \
			                                             hipGetErrorString(e_));            \
	} while (0)

struct Pop {
	int64_t size   = 0;
	int32_t words  = 0;
	int32_t draws  = 0;
	uint32_t* state = nullptr;     // [words][size]
	uint64_t* bitmap = nullptr;    // [SPICE_MAX_BLOCK][bitmap_stride]
	int64_t bitmap_stride = 0;
	int32_t* ring_ids = nullptr;   // [ring][size]
	int32_t* ring_cnt = nullptr;   // [ring]
	xoro_state* chunk_jump = nullptr;
	int32_t chunk = 16;
	gf2poly base_poly{1, 0};       // x^(stream offset of neuron 0) mod P
	std::vector<int64_t> incoming; // projection ids, connect order
};

struct Proj {
	int64_t src = 0, dst = 0;
	int64_t delay = 1;
	int64_t* offsets = nullptr;    // [src_size + 1]
	int32_t* neighbors = nullptr;  // [nnz]
	int64_t nnz = 0;
	uint32_t* counts = nullptr;    // [slots][dst_size]
	int64_t slots = 0;
	// Tile view for pull delivery (blocked.hip): 16-bit tile-local targets in
	// CSR order and, per row, the in-row start of each tile's run.
	int64_t tile = 0;              // 0: not built
	int64_t ntile = 0;
	uint16_t* rt = nullptr;        // [src_size][ntile + 1]
	uint16_t* btgt = nullptr;      // [nnz], padded to whole 16-byte vectors
};

// Device-side descriptors, rebuilt when populations/projections change.
struct PopDev {
	uint64_t const* bitmap;
	int64_t bitmap_stride;
	int64_t size;
	int32_t* ring_ids;
	int32_t* ring_cnt;
};

struct ProjDev {
	int64_t const* offsets;
	int32_t const* neighbors;
	uint32_t* counts;
	int64_t slots;
	int64_t dst_size;
	int64_t delay;
	int32_t const* src_ring_ids;
	int32_t const* src_ring_cnt;
	int64_t src_size;
	uint16_t const* rt;
	uint16_t const* btgt;
	int64_t ntile;
};

// Pull delivery: one workgroup per (target tile, step).
struct PullTile {
	int32_t target;  // index into PullTarget array
	int32_t tau;     // tile within the target population
};
struct PullTarget {
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
	int64_t steps = 0; // completed steps
	int mode = SPICE_DELIVER_PULL;
	bool timing = false;
	std::vector<spice_rt::Pop> pops;
	std::vector<spice_rt::Proj> projs;
	spice_rt::PopDev* pops_dev = nullptr;
	spice_rt::ProjDev* projs_dev = nullptr;
	spice_rt::PullTile* pull_tiles = nullptr;
	spice_rt::PullTarget* pull_targets = nullptr;
	int64_t pull_ntiles = 0;
	int32_t pull_max_in = 0;
	int64_t tile = 8192;  // targets per pull tile
	bool dev_dirty = true;
	int32_t* error = nullptr;                 // device flag
	unsigned long long* events = nullptr;     // device counters: events, rows
	// pending block
	int32_t pending = 0;
	int64_t pending_t0 = 0;
	hipEvent_t upd_begin = nullptr;
	std::vector<spice_rt::TimedBlock> timed;
	spice_stats stats{};
};
