// Exclusive prefix sum (mod 2^64) of n uint64 values in place, in three
// launches: per-4096 tile sums, a scan of the tile sums, and the in-tile scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spice_rt {
namespace {
__global__ void __launch_bounds__(256) scan_tile_sums(uint64_t const* v, int64_t n, uint64_t* tiles) {
	int64_t const base = int64_t(blockIdx.x) * 4096;
	uint64_t s         = 0;
	for (int i = threadIdx.x; i < 4096; i += 256)
		if (base + i < n) s += v[base + i];
	__shared__ uint64_t red[256];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0) tiles[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) scan_tiles(uint64_t* tiles, int64_t ntiles, uint64_t* total) {
	// sequential chunks per thread + block scan of chunk sums
	int64_t const per = (ntiles + 255) / 256;
	int64_t const a = min<int64_t>(ntiles, threadIdx.x * per), b = min<int64_t>(ntiles, a + per);
	uint64_t s = 0;
	for (int64_t i = a; i < b; i++) s += tiles[i];
	__shared__ uint64_t sc[256];
	sc[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < 256; off <<= 1) {
		uint64_t const v = threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
		__syncthreads();
		sc[threadIdx.x] += v;
		__syncthreads();
	}
	uint64_t run = sc[threadIdx.x] - s;
	for (int64_t i = a; i < b; i++) {
		uint64_t const t = tiles[i];
		tiles[i]         = run;
		run += t;
	}
	if (threadIdx.x == 255) *total = sc[255];
}

__global__ void __launch_bounds__(256) scan_apply(uint64_t* v, int64_t n, uint64_t const* tiles) {
	int64_t const base = int64_t(blockIdx.x) * 4096;
	// each thread owns 16 consecutive values
	int64_t const a = base + threadIdx.x * 16;
	uint64_t loc[16];
	uint64_t s = 0;
	for (int i = 0; i < 16; i++) {
		loc[i] = (a + i < n) ? v[a + i] : 0;
		s += loc[i];
	}
	__shared__ uint64_t sc[256];
	sc[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < 256; off <<= 1) {
		uint64_t const t = threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
		__syncthreads();
		sc[threadIdx.x] += t;
		__syncthreads();
	}
	uint64_t run = tiles[blockIdx.x] + sc[threadIdx.x] - s;
	for (int i = 0; i < 16; i++) {
		if (a + i < n) v[a + i] = run;
		run += loc[i];
	}
}


// Scans v[0..n) (n includes a trailing slot whose result is the total); tiles
// must hold (n + 4095) / 4096 entries.
inline hipError_t exclusive_scan_u64(uint64_t* v, int64_t n, uint64_t* tiles, uint64_t* total,
                                     hipStream_t st) {
	int64_t const ntiles = (n + 4095) / 4096;
	hipLaunchKernelGGL(scan_tile_sums, dim3(unsigned(ntiles)), dim3(256), 0, st, v, n, tiles);
	hipLaunchKernelGGL(scan_tiles, dim3(1), dim3(256), 0, st, tiles, ntiles, total);
	hipLaunchKernelGGL(scan_apply, dim3(unsigned(ntiles)), dim3(256), 0, st, v, n, tiles);
	return hipGetLastError();
}
}  // namespace
}  // namespace spice_rt
