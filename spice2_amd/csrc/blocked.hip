// Tile view of a projection's CSR for delivery (pull and push).
//
// Targets are split into tiles of `tile` consecutive neurons.  Rows are sorted,
// so a row's targets in one tile form a contiguous run.  Kept beside the
// canonical int32 CSR:
//   btgt          16-bit tile-local targets (neighbors[j] mod tile), row by row
//                 and tile by tile in CSR order; every run starts on a 16-byte
//                 vector and is padded to whole vectors with the values
//                 tile + (position mod 8), which name spare counters past the
//                 tile, so a reader handles whole vectors without bounds checks
//   rv[src][t]    index of the first 16-byte vector of row src's run in tile t,
//                 t in [0, ntile]; the run is rv[src][t] .. rv[src][t+1]
//                 (rv[src][ntile] == rv[src+1][0])
// A workgroup that owns tile t reads, per spiking source, the two run bounds
// and the run: 2 bytes per synaptic event plus the padding of the run's last
// vector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "internal.h"
#include "scan.h"

namespace spice_rt {
namespace {

unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536))); }

// pos[src][t] = index within row src of its first target >= t * tile.
__global__ void __launch_bounds__(256)
tile_runs_kernel(int64_t const* offsets, int32_t const* nb, int64_t S, int64_t ntile, int64_t tile, uint32_t* pos) {
	int64_t const n      = (ntile + 1) * S;
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
		int64_t const src = k / (ntile + 1), t = k % (ntile + 1);
		int64_t const a = offsets[src], b = offsets[src + 1];
		int64_t const key = t * tile;
		int64_t lo = a, hi = b;  // first index in [a, b) with nb >= key
		while (lo < hi) {
			int64_t const mid = (lo + hi) >> 1;
			if (nb[mid] < key)
				lo = mid + 1;
			else
				hi = mid;
		}
		pos[k] = uint32_t(lo - a);
	}
}

// Vectors of each row (sum over its tiles of ceil(run / 8)).
__global__ void __launch_bounds__(256)
row_vectors_kernel(uint32_t const* pos, int64_t S, int64_t ntile, uint64_t* rowv) {
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t src = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; src < S; src += stride) {
		uint32_t const* p = pos + src * (ntile + 1);
		uint64_t v        = 0;
		for (int64_t t = 0; t < ntile; t++) v += (p[t + 1] - p[t] + 7) / 8;
		rowv[src] = v;
	}
}

// rv[src][t] from the scanned row starts.
__global__ void __launch_bounds__(256)
run_starts_kernel(uint32_t const* pos, int64_t S, int64_t ntile, uint64_t const* rowv, uint32_t* rv) {
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t src = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; src < S; src += stride) {
		uint32_t const* p = pos + src * (ntile + 1);
		uint32_t* r       = rv + src * (ntile + 1);
		uint64_t v        = rowv[src];
		for (int64_t t = 0; t < ntile; t++) {
			r[t] = uint32_t(v);
			v += (p[t + 1] - p[t] + 7) / 8;
		}
		r[ntile] = uint32_t(v);
	}
}

// One wave per row: targets to their runs, then the padding of every run.
__global__ void __launch_bounds__(256)
fill_runs_kernel(int64_t const* offsets, int32_t const* nb, int64_t S, int64_t ntile, int64_t tile,
                 uint32_t const* pos, uint32_t const* rv, uint16_t* btgt) {
	int const lane       = threadIdx.x % 64;
	int64_t const nw     = int64_t(gridDim.x) * blockDim.x / 64;
	for (int64_t src = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64; src < S; src += nw) {
		int64_t const a   = offsets[src], n = offsets[src + 1] - a;
		uint32_t const* p = pos + src * (ntile + 1);
		uint32_t const* r = rv + src * (ntile + 1);
		for (int64_t i = lane; i < n; i += 64) {
			int32_t const g = nb[a + i];
			int64_t const t = g / tile;
			btgt[int64_t(r[t]) * 8 + (i - p[t])] = uint16_t(g - t * tile);
		}
		for (int64_t t = lane; t < ntile; t += 64) {
			int64_t const len = p[t + 1] - p[t];
			int64_t const end = int64_t(r[t + 1] - r[t]) * 8;
			for (int64_t e = len; e < end; e++) btgt[int64_t(r[t]) * 8 + e] = uint16_t(tile + (e & 7));
		}
	}
}

}  // namespace

int build_blocked(spice_ctx* ctx, Proj& Q, int64_t tile) {
	hipStream_t const st = ctx->stream;
	int64_t const S      = ctx->pops[Q.src].size;
	int64_t const D      = ctx->pops[Q.dst].size;
	if (tile <= 0 || tile > kMaxPullTile || tile % 64)
		return set_error(SPICE_EINVAL, "tile must be a multiple of 64 in [64, " + std::to_string(kMaxPullTile) + "]");
	int64_t const ntile = std::max<int64_t>((D + tile - 1) / tile, 1);
	int64_t const nrv   = (ntile + 1) * S;
	if (nrv > (int64_t(1) << 32))
		return set_error(SPICE_EINVAL, "tile too small for this projection (run table over 2^32 entries)");
	dfree(st, Q.rv);
	dfree(st, Q.btgt);
	Q.rv   = nullptr;
	Q.btgt = nullptr;
	Q.tile = 0;
	uint32_t* pos  = nullptr;
	uint64_t* rowv = nullptr;
	uint64_t* tsum = nullptr;
	int64_t const nscan = S + 1;
	int r               = SPICE_OK;
	auto fail = [&](int code, std::string const& msg) {
		r = set_error(code, msg);
		return r;
	};
	do {
		// temporaries from the stream-ordered allocator (released without a device-wide wait)
		if (hipMallocAsync(reinterpret_cast<void**>(&pos), sizeof(uint32_t) * std::max<int64_t>(nrv, 1), st) != hipSuccess ||
		    hipMallocAsync(reinterpret_cast<void**>(&rowv), sizeof(uint64_t) * nscan, st) != hipSuccess ||
		    hipMallocAsync(reinterpret_cast<void**>(&tsum), sizeof(uint64_t) * ((nscan + 4095) / 4096 + 1), st) !=
		        hipSuccess) {
			fail(SPICE_EDEVICE, "out of device memory building the tile view");
			break;
		}
		SPICE_HIP(hipMemsetAsync(rowv, 0, sizeof(uint64_t) * nscan, st));
		if (S > 0) {
			hipLaunchKernelGGL(tile_runs_kernel, dim3(grid_for(nrv)), dim3(256), 0, st, Q.offsets, Q.neighbors, S,
			                   ntile, tile, pos);
			hipLaunchKernelGGL(row_vectors_kernel, dim3(grid_for(S)), dim3(256), 0, st, pos, S, ntile, rowv);
		}
		SPICE_HIP(exclusive_scan_u64(rowv, nscan, tsum, tsum + (nscan + 4095) / 4096, st));
		uint64_t nvec = 0;
		SPICE_HIP(hipMemcpyAsync(&nvec, rowv + S, sizeof(nvec), hipMemcpyDeviceToHost, st));
		SPICE_HIP(hipStreamSynchronize(st));
		if (nvec >= (uint64_t(1) << 32) - 1) {
			fail(SPICE_EINVAL, "projection too large for tiled delivery (over 2^32 16-byte target vectors)");
			break;
		}
		// one spare vector past the end: readers may load it (and never count it)
		if (dmalloc(reinterpret_cast<void**>(&Q.rv), sizeof(uint32_t) * std::max<int64_t>(nrv, 1)) != hipSuccess ||
		    dmalloc(reinterpret_cast<void**>(&Q.btgt), 16 * (nvec + 1)) != hipSuccess) {
			fail(SPICE_EDEVICE, "out of device memory building the tile view");
			break;
		}
		SPICE_HIP(hipMemsetAsync(Q.btgt + nvec * 8, 0, 16, st));
		if (S > 0) {
			hipLaunchKernelGGL(run_starts_kernel, dim3(grid_for(S)), dim3(256), 0, st, pos, S, ntile, rowv, Q.rv);
			hipLaunchKernelGGL(fill_runs_kernel, dim3(grid_for(S * 64)), dim3(256), 0, st, Q.offsets, Q.neighbors, S,
			                   ntile, tile, pos, Q.rv, Q.btgt);
		}
		SPICE_HIP(hipGetLastError());
		SPICE_HIP(hipStreamSynchronize(st));
		Q.tile  = tile;
		Q.ntile = ntile;
		Q.nvec  = int64_t(nvec);
	} while (false);
	if (pos) (void)hipFreeAsync(pos, st);
	if (rowv) (void)hipFreeAsync(rowv, st);
	if (tsum) (void)hipFreeAsync(tsum, st);
	if (r) {
		dfree(st, Q.rv);
		dfree(st, Q.btgt);
		Q.rv   = nullptr;
		Q.btgt = nullptr;
	}
	return r;
}

}  // namespace spice_rt
