/* spice_hip.h - C ABI of the MI355X simulation runtime (libspice_hip.so).
 *
 * The C++20 user API (include/spice/snn.h and friends) is a header layer over
 * these entry points.  The layer instantiates the model-specific neuron-update
 * kernel for each user neuron type in the user's translation unit; everything
 * model independent - device memory, on-device fixed_probability generation,
 * spike compaction, the delay ring and spike delivery - is behind this ABI.
 *
 * Plain C types only.  Every call returns SPICE_OK (0) or a negative status;
 * spice_last_error() then holds a message.  SPICE_EPRECONDITION carries the
 * reference's "Assertion failed (<file>:<line>): <condition>" text so the C++
 * layer can rethrow it as std::logic_error (reference util/assert.cpp:8-15).
 *
 * Each entry point names the reference interface it replaces.  One context per
 * host thread; a context is not thread-safe (reference: snn is single-threaded).
 */
#ifndef SPICE_HIP_H
#define SPICE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPICE_OK 0
#define SPICE_EINVAL (-1)       /* bad argument or id */
#define SPICE_EHIP (-2)         /* HIP runtime error */
#define SPICE_EPRECONDITION (-3) /* reference SPICE_PRE/SPICE_INV failure */
#define SPICE_EDEVICE (-4)      /* a device-side check failed */
#define SPICE_ENOMEM (-5)

#define SPICE_MAX_BLOCK 16 /* steps per fused block (bounded by the smallest delay) */
#define SPICE_MAX_IN 16    /* incoming projections per population */

/* Delivery algorithms.  Both are exact for synapses whose deliver() does not
 * depend on per-synapse state: they produce per-(projection, target, step)
 * event counts and the neuron kernel replays deliver() count times in connect
 * order (the order the reference applies them, synapse_population.h:88-134). */
#define SPICE_DELIVER_PULL 0 /* target-tile owned, LDS counters, no global atomics */
#define SPICE_DELIVER_PUSH 1 /* source-row parallel, global atomic counters */

typedef struct spice_ctx spice_ctx;

typedef struct {
	uint64_t s0, s1;
} spice_rng_state;

/* Per incoming projection, as seen by a neuron-update kernel. */
typedef struct {
	uint32_t* counts; /* ring [slots][dst_size]: events for application step u at slot u % slots */
	int64_t slots;
	int64_t proj;
} spice_in_view;

/* Everything a neuron-update kernel needs for one population and one block of
 * steps t0 .. t0+nsteps-1.  Filled by spice_block_begin(). */
typedef struct {
	uint32_t* state;        /* SoA: 32-bit word w of neuron i at state[w * size + i] */
	int64_t size;
	int32_t words;          /* 32-bit words per neuron (sizeof(neuron) rounded up) */
	int32_t draws;          /* RNG outputs each neuron takes per step (stream offset stride) */
	uint64_t* bitmap;       /* [nsteps][bitmap_stride] spike bits, bit i%64 of word i/64 */
	int64_t bitmap_stride;
	spice_rng_state const* chunk_jump; /* per chunk of `chunk` neurons: x^(first*draws) mod P */
	int32_t chunk;
	int32_t nin;
	spice_in_view in[SPICE_MAX_IN]; /* connect order */
	spice_rng_state rng[SPICE_MAX_BLOCK]; /* per step: stream state at this population's first draw */
	int32_t* error;         /* device flag: nonzero after a failed device-side check */
	int32_t clear_counts;   /* the update kernel clears the counters it consumes (push mode) */
} spice_pop_view;

typedef struct {
	int64_t t0;
	int32_t nsteps;
	float dt[SPICE_MAX_BLOCK]; /* Kahan-compensated dt of each step (numeric.h:9-15) */
} spice_block_view;

typedef struct {
	int64_t steps;          /* simulated steps completed on the device */
	uint64_t events;        /* synaptic events delivered (sum of delivered row lengths) */
	uint64_t rows;          /* CSR rows delivered (spikes times outgoing projections) */
	double deliver_ms;      /* time in delivery kernels (timing enabled only) */
	int64_t deliver_launches;
	double update_ms;       /* time in neuron-update kernels (timing enabled only) */
	int64_t update_launches;
	double compact_ms;
	int64_t csr_rows_host_resolved; /* see spice_proj_create_fixed_probability */
	int64_t csr_ambiguous_draws;
	double csr_ms;
} spice_stats;

char const* spice_last_error(void);

/* ---- context (replaces the snn container, reference snn.h:16-75) ---------- */
/* max_delay: the network's maximum delay in steps (snn.h:19); it bounds every
 * projection's delay and the spike history kept for spice_spikes(). */
int spice_ctx_create(int device, int64_t max_delay, spice_ctx** out);
void spice_ctx_destroy(spice_ctx* ctx);
/* Delivery algorithm (default SPICE_DELIVER_PULL); fixed once steps have run. */
int spice_ctx_set_delivery(spice_ctx* ctx, int mode);
/* Targets per pull tile (multiple of 64, <= 65536); the runtime lowers it so that
 * the counters of a tile's incoming projections fit 64 KB of LDS. */
int spice_ctx_set_tile(spice_ctx* ctx, int64_t tile);
int64_t spice_ctx_tile(spice_ctx* ctx);
int spice_ctx_set_timing(spice_ctx* ctx, int enabled);
void* spice_ctx_stream(spice_ctx* ctx); /* hipStream_t all work of the context runs on */
int spice_ctx_sync(spice_ctx* ctx);     /* waits and reports device-side errors */
int spice_ctx_stats(spice_ctx* ctx, spice_stats* out);

/* ---- populations (replaces detail::neuron_population, reference
 *      neuron_population.h:103-171 and its adapters :29-101) ---------------- */
/* state_words = 0 for stateless neurons.  draws = RNG outputs per neuron per
 * step (the stream offset of neuron i is base + i * draws, reference
 * snn.cpp:12-15 plus the in-order loops of neuron_population.h:40-45,72-77). */
int spice_pop_create(spice_ctx* ctx, int64_t size, int32_t state_words, int32_t draws,
                     int64_t* out_pop);
/* AoS <-> SoA transfer of the neuron state (get_neurons(), neuron_population.h:142-145). */
int spice_pop_upload(spice_ctx* ctx, int64_t pop, void const* aos, int32_t struct_bytes);
int spice_pop_download(spice_ctx* ctx, int64_t pop, void* aos, int32_t struct_bytes);
int64_t spice_pop_size(spice_ctx* ctx, int64_t pop);

/* ---- projections (replaces detail::synapse_population + detail::csr +
 *      fixed_probability::generate, synapse_population.h:26-45, csr.h:69-77,
 *      topology.cpp:73-112) ------------------------------------------------ */
/* Generates the fixed_probability CSR on the device from the stream seeded with
 * (seed_lo, seed_hi), bit-exact with the reference (which draws with glibc's
 * log).  The device computes each draw's logarithm in double-double; a draw
 * whose logarithm lies within 0.05 ulp of a rounding midpoint - where glibc's
 * <= 0.519 ulp result may be either neighbour - is carried as an interval, and
 * a row whose targets or length could depend on such a draw is regenerated on
 * the host with the C library's log (counted in spice_stats). */
int spice_proj_create_fixed_probability(spice_ctx* ctx, int64_t src_pop, int64_t dst_pop,
                                        double p, uint64_t seed_lo, uint64_t seed_hi,
                                        int64_t delay, int64_t* out_proj);
/* A caller-generated CSR (adj_list / edge_stream topologies, topology.cpp:12-71). */
int spice_proj_create_csr(spice_ctx* ctx, int64_t src_pop, int64_t dst_pop,
                          int64_t const* offsets, int32_t const* neighbors, int64_t delay,
                          int64_t* out_proj);
int spice_proj_synapses(spice_ctx* ctx, int64_t proj, int64_t* out_count);
int spice_proj_download(spice_ctx* ctx, int64_t proj, int64_t* offsets, int32_t* neighbors);

/* ---- stepping (replaces snn::step, snn.cpp:7-28) ----------------------------
 * A block runs nsteps <= min(SPICE_MAX_BLOCK, smallest projection delay) steps:
 *   1. spice_block_begin() fills one spice_pop_view per population;
 *   2. the caller launches each population's neuron-update kernel on
 *      spice_ctx_stream() (it applies pending deliveries, updates, and sets
 *      spike bits);
 *   3. spice_block_end() compacts spike bits into ordered spike lists (the
 *      delay ring) and delivers them.
 * seeds: 2 per step (lo, hi) of the step's RNG stream; dt: per step. */
int spice_block_begin(spice_ctx* ctx, int32_t nsteps, float const* dt, uint64_t const* seeds,
                      spice_block_view* block, spice_pop_view* pops, int32_t max_pops);
int spice_block_end(spice_ctx* ctx);
/* The view of a population outside a block (no RNG fields): used to apply the
 * deliveries already counted for the next step before the host reads neuron
 * state, since the reference applies each step's deliveries at the end of that
 * step (snn.cpp:21-25) while the device applies them at the start of the next. */
int spice_pop_view_now(spice_ctx* ctx, int64_t pop, spice_pop_view* out);
/* Steps completed on the device. */
int64_t spice_steps(spice_ctx* ctx);
/* Spike ids of `pop` from `age` steps ago (age 0 = last step), ascending.
 * Reference neuron_population.h:147-153. */
int spice_spikes(spice_ctx* ctx, int64_t pop, int64_t age, int32_t* out, int64_t cap,
                 int64_t* out_count);
/* Number of steps of spike history kept (the delay window). */
int64_t spice_history_len(spice_ctx* ctx);

/* ---- host utilities (no device work) ------------------------------------ */
/* xoroshiro128+ state advanced by n outputs (random.h:222-234). */
void spice_rng_jump(uint64_t s0, uint64_t s1, uint64_t n, uint64_t* out_s0, uint64_t* out_s1);
/* x^n mod P(x), P the characteristic polynomial of the generator. */
void spice_rng_jump_poly(uint64_t n, uint64_t* out_lo, uint64_t* out_hi);
/* Correctly rounded natural log (the device generator's log, on the host). */
double spice_log_cr(double x);
/* The exponential draw -scale*log(u) of random.h:270-272 as the interval [lo, hi]
 * of values a C library log within 0.55 ulp can give; returns 1 if lo != hi. */
int spice_exp_draw_interval(double u, double scale, double* lo, double* hi);

#ifdef __cplusplus
}
#endif
#endif
