// Tile view of a projection's CSR for pull delivery.
//
// Targets are split into tiles of `tile` consecutive neurons.  Rows are sorted,
// so a row's targets in one tile form a contiguous run.  Kept beside the
// canonical int32 CSR:
//   btgt[j]       = neighbors[j] mod tile  (16-bit tile-local target, same order)
//   rt[src][t]    = index within row src of its first target >= t * tile,
//                   t in [0, ntile] (16-bit: rows are shorter than 65536)
// A workgroup that owns tile t reads, per spiking source, one line holding the
// source's rt row (shared by all tiles) and the source's run in btgt: 2 bytes
// per synaptic event.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "internal.h"

namespace spice_rt {
namespace {

__global__ void __launch_bounds__(256)
tile_runs_kernel(int64_t const* offsets, int32_t const* nb, int64_t S, int64_t ntile, int64_t tile,
                 uint16_t* rt) {
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
		rt[k] = uint16_t(lo - a);
	}
}

__global__ void __launch_bounds__(256)
tile_local_kernel(int32_t const* nb, int64_t nnz, int64_t tile, uint16_t* btgt) {
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nnz; j += stride)
		btgt[j] = uint16_t(nb[j] % tile);
}

unsigned grid_for(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536))); }

}  // namespace

int build_blocked(spice_ctx* ctx, Proj& Q, int64_t tile) {
	hipStream_t const st = ctx->stream;
	int64_t const S      = ctx->pops[Q.src].size;
	int64_t const D      = ctx->pops[Q.dst].size;
	if (tile <= 0 || tile > 65536 || tile % 64)
		return set_error(SPICE_EINVAL, "pull tile must be a multiple of 64 in [64, 65536]");
	int64_t const ntile = std::max<int64_t>((D + tile - 1) / tile, 1);
	// rows must be shorter than 65536 for 16-bit run starts
	int64_t maxdeg = 0;
	{
		std::vector<int64_t> off(size_t(S + 1));
		SPICE_HIP(hipMemcpy(off.data(), Q.offsets, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost));
		for (int64_t i = 0; i < S; i++) maxdeg = std::max(maxdeg, off[i + 1] - off[i]);
	}
	if (maxdeg >= 65536)
		return set_error(SPICE_EINVAL, "pull delivery needs rows shorter than 65536 synapses; use push");
	(void)hipFree(Q.rt);
	(void)hipFree(Q.btgt);
	Q.rt   = nullptr;
	Q.btgt = nullptr;
	SPICE_HIP(hipMalloc(reinterpret_cast<void**>(&Q.rt), sizeof(uint16_t) * std::max<int64_t>((ntile + 1) * S, 1)));
	// padded to whole 16-byte vectors so that vector loads of a row's last run stay in bounds
	int64_t const padded = (Q.nnz + 7) / 8 * 8 + 8;
	SPICE_HIP(hipMalloc(reinterpret_cast<void**>(&Q.btgt), sizeof(uint16_t) * padded));
	SPICE_HIP(hipMemsetAsync(Q.btgt, 0, sizeof(uint16_t) * padded, st));
	if (S > 0)
		hipLaunchKernelGGL(tile_runs_kernel, dim3(grid_for((ntile + 1) * S)), dim3(256), 0, st, Q.offsets,
		                   Q.neighbors, S, ntile, tile, Q.rt);
	if (Q.nnz > 0)
		hipLaunchKernelGGL(tile_local_kernel, dim3(grid_for(Q.nnz)), dim3(256), 0, st, Q.neighbors, Q.nnz,
		                   tile, Q.btgt);
	SPICE_HIP(hipGetLastError());
	SPICE_HIP(hipStreamSynchronize(st));
	Q.tile  = tile;
	Q.ntile = ntile;
	return SPICE_OK;
}

}  // namespace spice_rt
