// On-device fixed_probability generation, bit-exact with the reference
// (topology.cpp:80-112): one xoroshiro128+ stream per projection; row `src`
// draws e = -scale * log(u) (scale = 1/p - 1), accumulates them sequentially in
// double (`noise`) and emits dst = index + round(noise) until dst >= dst_count or
// index >= max_degree; every row consumes deg + 1 draws, so where a row starts
// in the stream depends on all rows before it.
//
// Three phases, all on the device:
//   A  draw:  for every block of 64 stream positions, the generator state at
//             its start (jump-ahead) and the fixed-point sum of its draws; then
//             an exclusive scan gives C(k), the fixed-point prefix sum of draws.
//   B  chain: one wavefront per projection walks the rows in order.  The row
//             starting at s ends at the first m with
//                 (C(m) - C(s)) + (m - 1 - s) >= D - 1/2   (or at the degree cap),
//             found with one coarse probe over the block prefixes of the next
//             max_degree + 1 positions and one fine probe inside a block.  The
//             fixed-point sums differ from the reference's sequential double
//             sums by a bounded amount; a row whose end lies within that bound
//             of the threshold is decided exactly (as in phase C).
//   C  emit:  one wavefront per row regenerates its draws and forms the
//             reference's sequential double sum exactly, writing the targets.
// Exactness: each draw's log is computed in double-double.  The reference
// calls glibc 2.35's log, which is within 0.519 ulp but not always correctly
// rounded; a draw whose exact log lies within 0.05 ulp of a rounding midpoint
// is carried as the interval of both candidate doubles, and the sequential sum
// is carried as [lower, upper] trajectories (floating-point addition is
// monotone).  A row whose targets or length differ between the trajectories -
// it could depend on how glibc rounded - is regenerated on the host with the C
// library's own log.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "crlog.h"
#include "internal.h"
#include "scan.h"

namespace spice_rt {
namespace {

constexpr int kWave       = 64;
constexpr int kGroup      = 4096;  // blocks of 64 positions per jump group
constexpr double kGlibcUlp = 0.55; // bound used for glibc's log error (published 0.519)
constexpr int kMaxUncertain = 1 << 14;

struct FpParams {
	int64_t S, D, M;
	double scale;
	int F;              // fractional bits of the fixed-point draw sums
	double two_F;
	uint64_t theta;     // (D - 1/2) * 2^F, or UINT64_MAX when unreachable
	uint64_t tol;       // decision margin in fixed point
	int64_t nblk;       // blocks of 64 positions covering the worst-case stream length
};

__device__ __forceinline__ double fast_draw(uint64_t x, double scale) {
	return -scale * log(canonical_f64_leftopen(x));
}
__device__ __forceinline__ uint64_t quantize(double e, double two_F) {
	return uint64_t(e * two_F + 0.5);
}

// Generator states are kept every kSub positions (8 per block of 64), so any
// position is at most kSub - 1 steps away from a stored state.
constexpr int kSub = 8;

__device__ __forceinline__ xoro_state state_at(xoro_state const* sub_state, int64_t pos) {
	xoro_state st = sub_state[pos / kSub];
	for (int i = 0, n = int(pos % kSub); i < n; i++) xoro_next(st);
	return st;
}

// Value of lane `lane` (uniform across the wave) read into a scalar register.
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int lane) {
	uint32_t const lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), lane));
	uint32_t const hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), lane));
	return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ double wave_bcast(double v, int lane) {
	return __builtin_bit_cast(double, lane_u64(__builtin_bit_cast(uint64_t, v), lane));
}

// Inclusive scan of a 64-bit value over the wave with DPP row shifts and row
// broadcasts (lane l receives the sum of lanes 0..l).
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
	uint32_t const lo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(v)), Ctrl, RowMask, 0xf, true));
	uint32_t const hi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(v >> 32)), Ctrl, RowMask, 0xf, true));
	return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t x) {
	x += dpp_u64<0x111, 0xf>(x);  // row_shr:1
	x += dpp_u64<0x112, 0xf>(x);  // row_shr:2
	x += dpp_u64<0x114, 0xf>(x);  // row_shr:4
	x += dpp_u64<0x118, 0xf>(x);  // row_shr:8
	x += dpp_u64<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
	x += dpp_u64<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
	return x;
}

// Phase A: generator states (every kSub positions) and per-block fixed-point
// draw sums.
__global__ void __launch_bounds__(256)
fp_draw_kernel(FpParams P, xoro_state const* group_state, gf2poly const* local_poly,
               xoro_state* sub_state, uint64_t* block_sum) {
	int64_t const stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < P.nblk; b += stride) {
		xoro_state st = xoro_jump(group_state[b / kGroup], local_poly[b % kGroup]);
		uint64_t sum  = 0;
		for (int i = 0; i < 64; i++) {
			if (i % kSub == 0) sub_state[b * (64 / kSub) + i / kSub] = st;
			sum += quantize(fast_draw(xoro_next(st), P.scale), P.two_F);
		}
		block_sum[b] = sum;
	}
}

// Exact row evaluation shared by phases B and C (one wavefront).  Walks the row
// starting at stream position s with lower/upper draw intervals and sequential
// double sums, in rounds of 1024 positions staged through LDS.
//   len_lo / len_hi: row lengths under the lower / upper trajectory (upper
//                    crosses first, so len_hi <= len_lo);
//   values_certain:  targets agree between trajectories for j < min length;
//   out:             if non-null, targets j < out_len are written (lower trajectory).
struct RowResult {
	int64_t len_lo, len_hi;
	bool values_certain;
	uint64_t ambiguous;
};

__device__ RowResult row_exact(FpParams const& P, xoro_state const* block_state, int64_t s,
                               int32_t* out, int64_t out_len, double* lds_lo, double* lds_hi) {
	int const lane = threadIdx.x % kWave;
	RowResult R{-1, -1, true, 0};
	double nlo = 0.0, nhi = 0.0;  // running sums (uniform across lanes)
	for (int64_t j0 = 0;; j0 += 1024) {
		// generate 16 consecutive draws per lane
		{
			xoro_state st = state_at(block_state, s + j0 + 16 * lane);
			for (int t = 0; t < 16; t++) {
				double lo, hi;
				double const u = canonical_f64_leftopen(xoro_next(st));
				if (spice_crlog::exp_draw_interval(u, P.scale, kGlibcUlp, lo, hi)) R.ambiguous++;
				lds_lo[16 * lane + t] = lo;
				lds_hi[16 * lane + t] = hi;
			}
		}
		__syncthreads();
		for (int i = 0; i < 16; i++) {
			double const xlo = lds_lo[64 * i + lane];
			double const xhi = lds_hi[64 * i + lane];
			double mylo = 0.0, myhi = 0.0;
#pragma unroll
			for (int jj = 0; jj < kWave; jj++) {
				nlo += wave_bcast(xlo, jj);
				nhi += wave_bcast(xhi, jj);
				if (lane == jj) {
					mylo = nlo;
					myhi = nhi;
				}
			}
			int64_t const j   = j0 + 64 * i + lane;
			int64_t const dlo = j + int64_t(round(mylo));
			int64_t const dhi = j + int64_t(round(myhi));
			bool const tlo    = (dlo >= P.D) || (j >= P.M);
			bool const thi    = (dhi >= P.D) || (j >= P.M);
			uint64_t const blo = __ballot(tlo);
			uint64_t const bhi = __ballot(thi);
			if (R.len_hi < 0 && bhi) R.len_hi = j0 + 64 * i + (__ffsll((unsigned long long)bhi) - 1);
			int64_t const first_lo =
			    blo ? j0 + 64 * i + (__ffsll((unsigned long long)blo) - 1) : INT64_MAX;
			int64_t const lim = min(first_lo, R.len_hi >= 0 ? R.len_hi : INT64_MAX);
			if (j < lim && dlo != dhi) R.values_certain = false;
			if (out && j < out_len && j < first_lo) out[j] = int32_t(dlo);
			if (blo) {
				R.len_lo = first_lo;
				break;
			}
		}
		__syncthreads();
		if (R.len_lo >= 0) break;
	}
	// reduce per-lane ambiguity counts and certainty
	for (int off = 32; off > 0; off >>= 1) R.ambiguous += __shfl_xor(R.ambiguous, off, kWave);
	R.values_certain = __all(R.values_certain);
	return R;
}

// Phase B: walk the row chain.  One 64-thread workgroup per projection.
struct ChainJob {
	FpParams P;
	xoro_state const* block_state;  // generator state every kSub positions
	uint64_t const* block_prefix;  // nblk + 1 entries
	int64_t* row_start;            // S + 1 entries
	int64_t r0, s0;                // resume point
	int32_t active;                // walk this launch
	int32_t status;                // 0 done, 1 host needed at row `status_row`, 2 error
	int64_t status_row;
	unsigned long long ambiguous;
	int64_t exact_rows;
};

__device__ uint64_t prefix_at(FpParams const& P, xoro_state const* block_state,
                              uint64_t const* block_prefix, int64_t pos) {
	// C(pos): block prefix plus the draws of pos's block before pos (wave-parallel)
	int const lane     = threadIdx.x % kWave;
	int64_t const b    = pos >> 6;
	int const r        = int(pos & 63);
	uint64_t q         = 0;
	if (lane < r) {
		xoro_state st = state_at(block_state, (b << 6) + lane);
		q             = quantize(fast_draw(xoro_next(st), P.scale), P.two_F);
	}
	for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, kWave);
	return block_prefix[b] + q;
}

// Blocks of the prefix and state tables the walk keeps in LDS: a row spans
// about (M + 1) / 64 blocks, so a window serves many consecutive rows and the
// walk reads global memory once per window instead of twice per row.
constexpr int kChainWin = 256;

__global__ void __launch_bounds__(64) fp_chain_kernel(ChainJob* jobs) {
	__shared__ double lds_lo[1024], lds_hi[1024];
	__shared__ uint64_t wpre[kChainWin], ws0[kChainWin * 8], ws1[kChainWin * 8];
	ChainJob& J      = jobs[blockIdx.x];
	if (!J.active) return;
	FpParams const P = J.P;
	int const lane   = threadIdx.x;
	int64_t wb = 0, wn = 0;  // window: blocks [wb, wb + wn)
	bool win   = false;      // the current row's blocks are in the window
	// makes blocks [lo, hi] resident if they fit (uniform across the wave)
	auto stage = [&](int64_t lo, int64_t hi) {
		if (hi - lo + 1 > kChainWin) {
			win = false;
			return;
		}
		win = true;
		if (wn > 0 && lo >= wb && hi < wb + wn) return;
		__syncthreads();
		wb = lo;
		wn = min<int64_t>(kChainWin, P.nblk + 1 - lo);
		// eight loads per lane in flight (clamped indices: every load is issued),
		// then their LDS writes
		constexpr int kBurst = 8;
		for (int64_t i0 = lane; i0 < wn; i0 += kBurst * kWave) {
			uint64_t v[kBurst];
#pragma unroll
			for (int k = 0; k < kBurst; k++) v[k] = J.block_prefix[wb + min<int64_t>(i0 + k * kWave, wn - 1)];
#pragma unroll
			for (int k = 0; k < kBurst; k++)
				if (i0 + k * kWave < wn) wpre[i0 + k * kWave] = v[k];
		}
		int64_t const ns = min<int64_t>(wn, P.nblk - wb) * (64 / kSub);
		for (int64_t i0 = lane; i0 < ns; i0 += kBurst * kWave) {
			xoro_state st[kBurst];
#pragma unroll
			for (int k = 0; k < kBurst; k++)
				st[k] = J.block_state[wb * (64 / kSub) + min<int64_t>(i0 + k * kWave, ns - 1)];
#pragma unroll
			for (int k = 0; k < kBurst; k++)
				if (i0 + k * kWave < ns) {
					ws0[i0 + k * kWave] = st[k].s0;
					ws1[i0 + k * kWave] = st[k].s1;
				}
		}
		__syncthreads();
	};
	auto pre = [&](int64_t b) -> uint64_t { return win ? wpre[b - wb] : J.block_prefix[b]; };
	auto sst = [&](int64_t q) -> xoro_state {  // state at position kSub * q
		return win ? xoro_state{ws0[q - wb * (64 / kSub)], ws1[q - wb * (64 / kSub)]} : J.block_state[q];
	};
	int64_t s        = J.s0;
	uint64_t Cs      = prefix_at(P, J.block_state, J.block_prefix, s);
	unsigned long long amb = 0;
	int64_t exact = 0;
	for (int64_t r = J.r0; r < P.S; r++) {
		if (lane == 0) J.row_start[r] = s;
		int64_t const last = s + P.M + 1;  // cap position for m
		// coarse: block boundaries q in [s+1, last]
		int64_t const b_lo = (s >> 6) + 1, b_hi = last >> 6;
		stage(s >> 6, min<int64_t>(b_hi, P.nblk - 1));
		int64_t fb = -1;  // block holding m - 1
		for (int64_t b0 = b_lo; b0 <= b_hi && fb < 0; b0 += kWave) {
			int64_t const b = b0 + lane;
			bool hit        = false;
			if (b <= b_hi) {
				uint64_t const G = (pre(b) - Cs) + (uint64_t((b << 6) - 1 - s) << P.F);
				hit              = G >= P.theta;
			}
			uint64_t const m = __ballot(hit);
			if (m) fb = b0 + (__ffsll((unsigned long long)m) - 1) - 1;
		}
		if (fb < 0) fb = (last - 1) >> 6;  // no crossing before the cap: block of position last-1
		// fine: positions P_j = 64 fb + j, G at P_j + 1
		int64_t const pj = (fb << 6) + lane;
		uint64_t qv      = 0;
		{
			xoro_state st = sst(pj / kSub);
			for (int i = 0; i < int(pj % kSub); i++) xoro_next(st);
			qv = quantize(fast_draw(xoro_next(st), P.scale), P.two_F);
		}
		uint64_t const incl = wave_inclusive_scan(qv);
		uint64_t const Cnext = pre(fb) + incl;  // C(pj + 1)
		bool const valid     = pj >= s && pj <= last - 1;
		uint64_t const G     = (Cnext - Cs) + (uint64_t(pj - s) << P.F);
		uint64_t const cross = __ballot(valid && G >= P.theta);
		int64_t m;
		int jl;  // lane of position m - 1
		if (cross) {
			jl = __ffsll((unsigned long long)cross) - 1;
			m  = (fb << 6) + jl + 1;
		} else {
			m  = last;
			jl = int(last - 1 - (fb << 6));
		}
		// margin checks on the decisions that fix the row length
		bool uncertain = false;
		uint64_t const G_m  = lane_u64(G, jl);
		uint64_t const G_pm = jl > 0 ? lane_u64(G, jl - 1) : 0;
		if (cross && m < last && G_m - P.theta < P.tol) uncertain = true;
		if (m - 1 > s) {  // position m-2 did not cross
			uint64_t gp;
			if (jl > 0)
				gp = G_pm;
			else  // m - 1 is the first position of block fb: boundary value
				gp = (pre(fb) - Cs) + (uint64_t((fb << 6) - 1 - s) << P.F);
			if (!(gp < P.theta && P.theta - gp >= P.tol)) uncertain = true;
		}
		uint64_t Cm = lane_u64(Cnext, jl);
		if (uncertain) {
			RowResult const R = row_exact(P, J.block_state, s, nullptr, 0, lds_lo, lds_hi);
			amb += R.ambiguous;
			exact++;
			if (R.len_lo != R.len_hi) {
				if (lane == 0) {
					J.status     = 1;
					J.status_row = r;
				}
				return;
			}
			m  = s + R.len_lo + 1;
			Cm = prefix_at(P, J.block_state, J.block_prefix, m);
		}
		s  = m;
		Cs = Cm;
	}
	if (lane == 0) {
		J.row_start[P.S] = s;
		J.status         = 0;
		J.ambiguous += amb;
		J.exact_rows += exact;
	}
}

// Phase C: emit every row.  One 64-thread workgroup per row (grid-stride).
__global__ void __launch_bounds__(64)
fp_emit_kernel(FpParams P, xoro_state const* block_state, int64_t const* row_start,
               int64_t* offsets, int32_t* neighbors, int64_t* uncertain, int32_t* n_uncertain,
               int32_t* error, unsigned long long* ambiguous) {
	__shared__ double lds_lo[1024], lds_hi[1024];
	unsigned long long amb = 0;
	for (int64_t r = blockIdx.x; r < P.S; r += gridDim.x) {
		int64_t const s   = row_start[r];
		int64_t const len = row_start[r + 1] - s - 1;
		int64_t const off = s - r;
		if (threadIdx.x == 0) offsets[r] = off;
		RowResult const R = row_exact(P, block_state, s, neighbors + off, len, lds_lo, lds_hi);
		amb += R.ambiguous;
		if (R.len_lo != len || R.len_hi != len || !R.values_certain) {
			if (threadIdx.x == 0) {
				if (R.len_lo == R.len_hi && R.len_lo != len) {
					atomicExch(error, 3);  // chain length disagrees with an exact row
				} else {
					int const k = atomicAdd(n_uncertain, 1);
					if (k < kMaxUncertain)
						uncertain[k] = r;
					else
						atomicExch(error, 4);
				}
			}
		}
	}
	if (threadIdx.x == 0) {
		if (blockIdx.x == 0) offsets[P.S] = row_start[P.S] - P.S;
		atomicAdd(ambiguous, amb);
	}
}

// ------------------------------------------------------------------- host

// The reference row computation with the C library's log (topology.cpp:93-110);
// used for rows the device proved could depend on glibc's rounding.
int64_t host_row(xoro_state st, FpParams const& P, std::vector<int32_t>& out) {
	out.clear();
	int32_t index = 0;
	double noise  = 0;
	for (;;) {
		double const u = canonical_f64_leftopen(xoro_next(st));
		noise += -P.scale * std::log(u);
		int32_t const dst = index + static_cast<int32_t>(std::round(noise));
		if ((dst >= P.D) | (index >= P.M)) break;
		out.push_back(dst);
		index++;
	}
	return index;
}

std::vector<gf2poly> const& local_polys() {
	static std::vector<gf2poly> v;
	static std::once_flag once;
	std::call_once(once, [] {
		v.resize(kGroup);
		gf2poly const step = xoro_jump_poly(64);
		gf2poly cur{1, 0};
		for (int j = 0; j < kGroup; j++) {
			v[j] = cur;
			cur  = gf2_mulmod(cur, step);
		}
	});
	return v;
}

double ulp_of(double x) {
	x = std::fabs(x);
	return std::nextafter(x, INFINITY) - x;
}

// A temporary from the stream-ordered allocator: allocation and release are
// ordered on the generating stream and neither waits for the rest of the device
// (other instances keep stepping on their own streams).
template <class T>
struct DevBuf {
	T* p          = nullptr;
	hipStream_t s = nullptr;
	~DevBuf() {
		if (p) (void)hipFreeAsync(p, s);
	}
	hipError_t alloc(size_t n, hipStream_t st) {
		s = st;
		return hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1), st);
	}
};

}  // namespace

// One projection being generated: parameters and device temporaries.
struct GenJob {
	int64_t proj = -1;
	int64_t S = 0, D = 0;
	FpParams P{};
	xoro_state seed{};
	int64_t L = 0;
	DevBuf<xoro_state> groups, block_state;
	DevBuf<uint64_t> prefix, tiles, total;
	DevBuf<int64_t> row_start, unc;
	DevBuf<int32_t> nunc;
	DevBuf<unsigned long long> amb;
	ChainJob job{};
	int64_t host_rows = 0;
};

// The parameters of topology.cpp:73-112 for one projection, checked on the host.
static int fp_params(int64_t S, int64_t D, double p, FpParams& P, int64_t& L) {
	// max_degree exactly as topology.cpp:74-75 (double arithmetic, truncation)
	int64_t const M = int64_t(double(D) * p + 3 * std::sqrt(double(D) * p * (1 - p)));
	P       = FpParams{};
	P.S     = S;
	P.D     = D;
	P.M     = M;
	P.scale = 1 / p - 1;
	// fixed point: differences over M + 1 positions stay below 2^62
	double const emax  = P.scale * 37.5 + 1.0;  // -log(2^-53) = 36.74
	double const span  = double(M + 2) * emax;
	P.F                = std::min(52, int(std::floor(62.0 - std::log2(span))));
	if (P.F < 8)
		return set_error(SPICE_EINVAL, "fixed_probability: connection probability too small for "
		                               "this target population size");
	P.two_F            = std::ldexp(1.0, P.F);
	double const th    = (double(D) - 0.5) * P.two_F;  // exact: F >= 1
	// every G value is below span * 2^F < 2^62, so a threshold at or above 2^62 is never reached
	P.theta            = th >= std::ldexp(1.0, 62) ? UINT64_MAX : uint64_t(th);
	// decision margin: sequential double rounding + log error + quantisation, x2
	double const nmax  = double(D) + emax;
	double const tol   = 2.0 * double(M + 2) *
	                   (0.5 * ulp_of(nmax) + 8.0 * ulp_of(emax) + std::ldexp(1.0, -P.F)) + 1e-12;
	P.tol              = uint64_t(std::ceil(tol * P.two_F)) + 2;
	// SPICE_CSR_EXACT_ROWS=1 decides every row's length with the exact interval
	// evaluation instead of the fixed-point probe (used to test that path).
	if (char const* e = std::getenv("SPICE_CSR_EXACT_ROWS"); e && e[0] == '1') P.tol = UINT64_MAX / 2;
	L = S * (M + 1);  // worst-case stream length
	// every kernel reads positions below L + 1024 (a row evaluation may run one
	// 1024-position round past the last row's end)
	P.nblk = (L + 1024 + 63) / 64 + 1;
	return SPICE_OK;
}

int queue_fixed_probability(spice_ctx* ctx, int64_t proj, int64_t S, int64_t D, double p, xoro_state seed) {
	Proj& pr = ctx->projs[proj];
	SPICE_HIP(dmalloc(reinterpret_cast<void**>(&pr.offsets), sizeof(int64_t) * (S + 1)));
	if (D == 0 || p == 0) {  // topology.cpp:84-85: all rows empty
		SPICE_HIP(hipMemsetAsync(pr.offsets, 0, sizeof(int64_t) * (S + 1), ctx->stream));
		SPICE_HIP(dmalloc(reinterpret_cast<void**>(&pr.neighbors), sizeof(int32_t)));
		pr.nnz = 0;
		return SPICE_OK;
	}
	FpParams P;
	int64_t L;
	if (int r = fp_params(S, D, p, P, L)) return r;
	ctx->pending_fp.push_back({proj, S, D, p, seed});
	pr.nnz = -1;  // not generated yet
	return SPICE_OK;
}

// Generates every queued fixed_probability projection.  Phases A and C run
// projection after projection (each fills the device); the row-chain walks of
// phase B, one workgroup each, run side by side in one launch.
int flush_fixed_probability(spice_ctx* ctx) {
	if (ctx->pending_fp.empty()) return SPICE_OK;
	auto const t_begin = std::chrono::steady_clock::now();
	hipStream_t const st = ctx->stream;
	std::vector<PendingFp> const reqs = std::move(ctx->pending_fp);
	ctx->pending_fp.clear();
	std::vector<std::unique_ptr<GenJob>> jobs;
	DevBuf<gf2poly> d_local;
	SPICE_HIP(d_local.alloc(kGroup, st));
	SPICE_HIP(hipMemcpyAsync(d_local.p, local_polys().data(), sizeof(gf2poly) * kGroup, hipMemcpyHostToDevice, st));
	std::vector<std::vector<xoro_state>> group_host;
	// phase A
	for (auto const& q : reqs) {
		auto J  = std::make_unique<GenJob>();
		J->proj = q.proj;
		J->S    = q.S;
		J->D    = q.D;
		J->seed = q.seed;
		if (int r = fp_params(q.S, q.D, q.p, J->P, J->L)) return r;
		FpParams const& P     = J->P;
		int64_t const ngroups = (P.nblk + kGroup - 1) / kGroup;
		group_host.emplace_back(ngroups);
		auto& groups = group_host.back();
		{
			gf2poly const gstep = xoro_jump_poly(uint64_t(64) * kGroup);
			xoro_state g        = q.seed;
			for (int64_t i = 0; i < ngroups; i++) {
				groups[i] = g;
				g         = xoro_jump(g, gstep);
			}
		}
		int64_t const ntiles = (P.nblk + 1 + 4095) / 4096;
		SPICE_HIP(J->groups.alloc(ngroups, st));
		SPICE_HIP(J->block_state.alloc(P.nblk * (64 / kSub), st));
		SPICE_HIP(J->prefix.alloc(P.nblk + 1, st));
		SPICE_HIP(J->tiles.alloc(ntiles, st));
		SPICE_HIP(J->total.alloc(1, st));
		SPICE_HIP(J->row_start.alloc(q.S + 1, st));
		SPICE_HIP(hipMemcpyAsync(J->groups.p, groups.data(), sizeof(xoro_state) * ngroups, hipMemcpyHostToDevice, st));
		SPICE_HIP(hipMemsetAsync(J->prefix.p, 0, sizeof(uint64_t) * (P.nblk + 1), st));
		int64_t const grid = std::min<int64_t>((P.nblk + 255) / 256, 65536);
		hipLaunchKernelGGL(fp_draw_kernel, dim3(unsigned(grid)), dim3(256), 0, st, P, J->groups.p, d_local.p,
		                   J->block_state.p, J->prefix.p);
		int64_t const n = P.nblk + 1;  // includes a trailing zero: its prefix is the total
		hipLaunchKernelGGL(scan_tile_sums, dim3(unsigned(ntiles)), dim3(256), 0, st, J->prefix.p, n, J->tiles.p);
		hipLaunchKernelGGL(scan_tiles, dim3(1), dim3(256), 0, st, J->tiles.p, ntiles, J->total.p);
		hipLaunchKernelGGL(scan_apply, dim3(unsigned(ntiles)), dim3(256), 0, st, J->prefix.p, n, J->tiles.p);
		SPICE_HIP(hipGetLastError());
		J->job.P            = P;
		J->job.block_state  = J->block_state.p;
		J->job.block_prefix = J->prefix.p;
		J->job.row_start    = J->row_start.p;
		J->job.active       = 1;
		jobs.push_back(std::move(J));
	}
	// phase B: all chains at once (one workgroup each); rows the device cannot
	// settle are settled on the host (glibc's rounding) and the chains resumed
	size_t const nj = jobs.size();
	std::vector<ChainJob> cj(nj);
	std::vector<int32_t> row;
	for (size_t i = 0; i < nj; i++) cj[i] = jobs[i]->job;
	DevBuf<ChainJob> d_jobs;
	SPICE_HIP(d_jobs.alloc(nj, st));
	for (;;) {
		SPICE_HIP(hipMemcpyAsync(d_jobs.p, cj.data(), sizeof(ChainJob) * nj, hipMemcpyHostToDevice, st));
		hipLaunchKernelGGL(fp_chain_kernel, dim3(unsigned(nj)), dim3(64), 0, st, d_jobs.p);
		SPICE_HIP(hipGetLastError());
		SPICE_HIP(hipMemcpyAsync(cj.data(), d_jobs.p, sizeof(ChainJob) * nj, hipMemcpyDeviceToHost, st));
		SPICE_HIP(hipStreamSynchronize(st));
		bool again = false;
		for (size_t i = 0; i < nj; i++) {
			ChainJob& job = cj[i];
			if (!job.active) continue;
			if (job.status == 0) {
				job.active = 0;
				continue;
			}
			if (job.status != 1) return set_error(SPICE_EDEVICE, "fixed_probability chain walk failed");
			// the row's length depends on glibc's rounding: settle it with the C library
			GenJob& J       = *jobs[i];
			int64_t const r = job.status_row;
			int64_t s       = 0;
			SPICE_HIP(hipMemcpy(&s, J.row_start.p + r, sizeof(s), hipMemcpyDeviceToHost));
			xoro_state bs;
			SPICE_HIP(hipMemcpy(&bs, J.block_state.p + s / kSub, sizeof(bs), hipMemcpyDeviceToHost));
			for (int k = 0; k < int(s % kSub); k++) xoro_next(bs);
			int64_t const len = host_row(bs, J.P, row);
			J.host_rows++;
			job.r0     = r + 1;
			job.s0     = s + len + 1;
			job.status = 0;
			if (r + 1 == J.S) {  // last row: record the end and finish
				int64_t const end = job.s0;
				SPICE_HIP(hipMemcpy(J.row_start.p + J.S, &end, sizeof(end), hipMemcpyHostToDevice));
				job.active = 0;
			} else {
				again = true;
			}
		}
		if (!again) break;
	}
	// phase C
	SPICE_HIP(hipMemsetAsync(ctx->error, 0, sizeof(int32_t), st));
	for (size_t i = 0; i < nj; i++) {
		GenJob& J = *jobs[i];
		Proj& pr  = ctx->projs[J.proj];
		int64_t end = 0;
		SPICE_HIP(hipMemcpy(&end, J.row_start.p + J.S, sizeof(end), hipMemcpyDeviceToHost));
		pr.nnz = end - J.S;
		if (pr.nnz < 0 || end > J.L) return set_error(SPICE_EDEVICE, "fixed_probability chain out of range");
		SPICE_HIP(dmalloc(reinterpret_cast<void**>(&pr.neighbors), sizeof(int32_t) * std::max<int64_t>(pr.nnz, 1)));
		SPICE_HIP(J.unc.alloc(kMaxUncertain, st));
		SPICE_HIP(J.nunc.alloc(1, st));
		SPICE_HIP(J.amb.alloc(1, st));
		SPICE_HIP(hipMemsetAsync(J.nunc.p, 0, sizeof(int32_t), st));
		SPICE_HIP(hipMemsetAsync(J.amb.p, 0, sizeof(unsigned long long), st));
		int64_t const grid = std::min<int64_t>(J.S, 65536);
		hipLaunchKernelGGL(fp_emit_kernel, dim3(unsigned(grid)), dim3(64), 0, st, J.P, J.block_state.p, J.row_start.p,
		                   pr.offsets, pr.neighbors, J.unc.p, J.nunc.p, ctx->error, J.amb.p);
		SPICE_HIP(hipGetLastError());
	}
	int32_t err = 0;
	SPICE_HIP(hipMemcpyAsync(&err, ctx->error, sizeof(err), hipMemcpyDeviceToHost, st));
	SPICE_HIP(hipStreamSynchronize(st));
	if (err) {
		(void)hipMemset(ctx->error, 0, sizeof(int32_t));
		return set_error(SPICE_EDEVICE, err == 3 ? "fixed_probability: row length disagreement"
		                                         : "fixed_probability: too many undecided rows");
	}
	for (size_t i = 0; i < nj; i++) {
		GenJob& J = *jobs[i];
		Proj& pr  = ctx->projs[J.proj];
		int32_t nunc = 0;
		unsigned long long amb = 0;
		SPICE_HIP(hipMemcpy(&nunc, J.nunc.p, sizeof(nunc), hipMemcpyDeviceToHost));
		SPICE_HIP(hipMemcpy(&amb, J.amb.p, sizeof(amb), hipMemcpyDeviceToHost));
		if (nunc > 0) {
			std::vector<int64_t> rows(nunc), starts(2);
			SPICE_HIP(hipMemcpy(rows.data(), J.unc.p, sizeof(int64_t) * nunc, hipMemcpyDeviceToHost));
			for (int64_t r : rows) {
				SPICE_HIP(hipMemcpy(starts.data(), J.row_start.p + r, sizeof(int64_t) * 2, hipMemcpyDeviceToHost));
				xoro_state bs;
				SPICE_HIP(hipMemcpy(&bs, J.block_state.p + starts[0] / kSub, sizeof(bs), hipMemcpyDeviceToHost));
				for (int k = 0; k < int(starts[0] % kSub); k++) xoro_next(bs);
				int64_t const len = host_row(bs, J.P, row);
				if (len != starts[1] - starts[0] - 1)
					return set_error(SPICE_EDEVICE, "fixed_probability: host row length disagreement");
				if (len)
					SPICE_HIP(hipMemcpy(pr.neighbors + (starts[0] - r), row.data(), sizeof(int32_t) * len,
					                    hipMemcpyHostToDevice));
				J.host_rows++;
			}
		}
		ctx->stats.csr_rows_host_resolved += J.host_rows;
		ctx->stats.csr_ambiguous_draws += int64_t(amb + cj[i].ambiguous);
	}
	ctx->stats.csr_ms +=
	    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count();
	return SPICE_OK;
}

}  // namespace spice_rt
