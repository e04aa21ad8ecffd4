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
#define SPICE_DELIVER_PUSH 1 /* spike chunks counted in LDS, then global atomic counters */

typedef struct spice_ctx spice_ctx;

typedef struct {
	uint64_t s0, s1;
} spice_rng_state;

/* Per incoming projection, as seen by a neuron-update kernel. */
typedef struct {
	uint32_t* counts; /* ring [slots][dst_size]: events for application step u at slot u % slots */
	int64_t slots;
	int64_t proj;
	/* projections with per-synapse state: the next step's events of each target
	 * are synapse ids events[ev_offsets[i] .. ev_offsets[i] + counts], in any
	 * order (ascending synapse id = the reference's source order); null otherwise */
	uint32_t const* ev_offsets;
	uint32_t const* events;
	uint32_t const* syn_state; /* SoA [syn_words][nnz] */
	int64_t nnz;
} spice_in_view;

/* One population as the neuron-update kernels see it (device-resident, one per
 * population, indexed by population id; rebuilt when the network changes). */
typedef struct {
	uint32_t* state;        /* SoA: 32-bit word w of neuron i at state[w * size + i] */
	int64_t size;
	int32_t words;          /* 32-bit words per neuron (sizeof(neuron) rounded up) */
	int32_t draws;          /* RNG outputs each neuron takes per step (stream offset stride) */
	uint64_t* bitmap;       /* spike bits, row per step: [SPICE_MAX_BLOCK] rows indexed by the
	                           step in the running block, or for stateless populations that draw
	                           (no state words) [SPICE_MAX_WINDOW] rows indexed by the step in the
	                           window; the last word of a row is scratch (waves past the end) */
	int64_t bitmap_stride;
	/* per chunk of 16 neurons: x^(stream offset of the chunk's first draw) mod P,
	 * P the generator's characteristic polynomial (offset counted from the
	 * instance's first population) */
	spice_rng_state const* chunk_jump;
	int32_t inst;           /* instance the population belongs to */
	int32_t nin;
	spice_in_view in[SPICE_MAX_IN]; /* connect order */
	int32_t* error;         /* device flag: nonzero after a failed device-side check */
	int32_t clear_counts;   /* the update kernel clears the counters it consumes (push mode) */
	uint64_t* hist;         /* populations with plastic projections: bit b = spiked b steps ago */
} spice_pop_view;

/* One projection as the synapse kernels see it (device-resident, one per
 * projection, indexed by projection id). */
typedef struct {
	int64_t const* offsets;      /* CSR [src_size + 1] */
	int32_t const* neighbors;    /* CSR [nnz] */
	int64_t nnz;
	int64_t src_size, dst_size;
	int64_t delay;
	int64_t ring;                /* spike-ring slots of the source */
	int32_t const* src_ring_ids; /* [ring][src_size] ascending */
	int32_t const* src_ring_cnt; /* [ring] */
	uint32_t* syn_state;         /* SoA [syn_words][nnz]; null without per-synapse state */
	int32_t syn_words;
	int32_t inst;
	uint64_t* ages;              /* plastic: per source, (last touch + 1) | delivered << 63 */
	uint64_t const* dst_hist;    /* plastic: target population's spike history */
	uint32_t* ev_cursor;         /* [dst_size]: next free event slot of each target */
	uint32_t* events;            /* [nnz] */
} spice_proj_view;

/* One step of one instance inside a window. */
typedef struct {
	float dt;                /* Kahan-compensated dt of the step (numeric.h:9-15) */
	float dt_nominal;        /* the network's dt, which synapse updates receive (snn.cpp:18,22) */
	int64_t local;           /* the instance's own step index; -1: the instance does not step */
	spice_rng_state seed;    /* xoroshiro128+ seed of the step's stream (snn.cpp:12) */
} spice_step;

#define SPICE_MAX_WINDOW 256 /* steps per spice_run() call */

#define SPICE_STREAM_STATES 128 /* states kept per step stream (jump tables) */

/* The steps of a window, device-resident: steps[inst * SPICE_MAX_WINDOW + k].
 * streams (null when no population draws): the first SPICE_STREAM_STATES states
 * of every step's stream, streams[(inst * SPICE_MAX_WINDOW + k) * SPICE_STREAM_STATES + i]
 * = the state after i outputs.  A jump by n is the XOR of the states at the set
 * bits of x^n mod P(x) (spice_pop_view::chunk_jump holds those bits). */
typedef struct {
	spice_step const* steps;
	spice_rng_state const* streams;
	int32_t nsteps;
	int32_t ninst;
} spice_window;

/* What the caller's kernel launches receive for one block of steps
 * rel .. rel+nsteps-1 of the running window (device pointers). */
typedef struct {
	spice_window const* win;
	spice_pop_view const* pops;
	spice_proj_view const* projs;
	int32_t rel;
	int32_t nsteps;
	int32_t phase; /* SPICE_PHASE_WINDOW, SPICE_PHASE_NEURONS or SPICE_PHASE_SYNAPSES */
} spice_block;

#define SPICE_PHASE_NEURONS 0  /* neuron updates (first in a block) */
#define SPICE_PHASE_SYNAPSES 1 /* plastic-synapse sweep and deliveries (last in a block) */
#define SPICE_PHASE_WINDOW 2   /* stateless populations that draw, every step of the window at
                                  once (first in a window; rel = 0, nsteps = the window's) */

/* Launches the caller's kernels of one phase of one block on `stream` (a
 * hipStream_t), without synchronising or allocating: blocks may be captured
 * into a HIP graph and replayed.  Returns 0, or nonzero after an error (the
 * callback keeps its own message). */
typedef int (*spice_update_fn)(void* user, spice_block const* block, void* stream);

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
/* Spike chunks per (tile, step) of push delivery, 1..64 (0, the default: one).  Any
 * value gives the same counts. */
int spice_ctx_set_push_chunks(spice_ctx* ctx, int32_t chunks);
/* Targets per delivery tile (multiple of 64, <= 32768; LDS use is 4 bytes per target and
 * projection counted together). */
int spice_ctx_set_tile(spice_ctx* ctx, int64_t tile);
int64_t spice_ctx_tile(spice_ctx* ctx);
/* Pull tile of one target population (0: the context's); exact for any tile. */
int spice_pop_set_tile(spice_ctx* ctx, int64_t pop, int64_t tile);
int64_t spice_pop_tile(spice_ctx* ctx, int64_t pop);
/* Tile autotuner (pull delivery): target populations are bucketed by size
 * (< 4096, < 16384, < 65536, larger); for each bucket in turn every candidate
 * tile is timed by replaying the delivery of the last block `reps` times (pull
 * delivery rewrites the same counts, so the simulation is unchanged), and the
 * fastest is kept.  chosen[4] receives the tile per bucket (0: empty bucket). */
int spice_ctx_autotune(spice_ctx* ctx, int64_t const* candidates, int32_t ncand, int32_t reps,
                       int64_t* chosen);
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
/* Per-synapse state (stateful synapses, synapse_population.h:30-41 and csr.h's
 * edge payload): `count` AoS structs of `struct_bytes`, either one (every synapse
 * starts from it) or one per synapse in CSR order.  The state lives SoA,
 * [words][nnz].  The projection's deliveries then become per-target event lists
 * that the target's update kernel applies in source order. */
int spice_proj_set_synapses(spice_ctx* ctx, int64_t proj, void const* aos, int64_t count,
                            int32_t struct_bytes);
int spice_proj_get_synapses(spice_ctx* ctx, int64_t proj, void* aos, int32_t struct_bytes);
/* Plastic synapses (update()/skip(), synapse_population.h:41-45,68-115): keeps
 * per-source ages and the spike history of both populations, runs the
 * SPICE_PHASE_SYNAPSES callback each step, and limits blocks to one step (a
 * plastic synapse's weight at delivery depends on the target's spikes up to
 * that step). */
int spice_proj_set_plastic(spice_ctx* ctx, int64_t proj);
int spice_ctx_proj_views(spice_ctx* ctx, spice_proj_view const** out);

/* ---- instances (independent networks sharing one context: batched
 *      parameter sweeps and served jobs; no reference counterpart - the
 *      reference runs one snn per process) ------------------------------------
 * spice_pop_create() adds to instance 0, created on first use with the
 * context's max_delay; spice_instance_create() adds instances with their own
 * max_delay (<= the context's).
 * Each instance has its own RNG seed chain, dt sequence and step count; its
 * populations and projections never interact with other instances'. */
int spice_instance_create(spice_ctx* ctx, int64_t max_delay, int64_t* out_inst);
int spice_pop_create_in(spice_ctx* ctx, int64_t inst, int64_t size, int32_t state_words,
                        int32_t draws, int64_t* out_pop);
/* Frees an instance's populations and projections (served jobs that finished);
 * their ids become invalid and the instance id is reused by a later
 * spice_instance_create(). */
int spice_instance_destroy(spice_ctx* ctx, int64_t inst);
int64_t spice_instance_steps(spice_ctx* ctx, int64_t inst); /* steps the instance has run */
int64_t spice_instance_count(spice_ctx* ctx);

/* ---- stepping (replaces snn::step, snn.cpp:7-28) ----------------------------
 * spice_run() runs a window of nsteps <= SPICE_MAX_WINDOW steps.  steps is
 * [instance][nsteps]: for each instance and step, its dt and stream seed, and
 * whether it steps (an instance's active steps within a window form a prefix).
 * First the update callback (SPICE_PHASE_WINDOW) launches the update of the
 * stateless populations that draw (no state words: their spikes depend only on
 * the steps' streams) for every step of the window.  The window is then split
 * into blocks of at most min(SPICE_MAX_BLOCK, smallest projection delay) steps;
 * per block:
 *   1. the update callback (SPICE_PHASE_NEURONS) launches every other
 *      population's neuron-update kernel (each applies the deliveries due,
 *      updates, and sets spike bits);
 *   2. the runtime compacts spike bits into ordered spike lists (the delay
 *      ring) and delivers them (event counts per target);
 *   3. with plastic projections (one-step blocks), the runtime lays out each
 *      target's event list and the callback (SPICE_PHASE_SYNAPSES) runs the
 *      synapse sweep and the plastic deliveries.
 * The reference order of snn.cpp:7-28 holds exactly: a spike of step s is
 * delivered at step s+d-1 and applied before the update of step s+d, and a
 * block never spans a step that depends on a delivery made inside it.
 * With graphs enabled, the blocks of a window are captured once per window
 * shape and replayed. */
int spice_ctx_set_update(spice_ctx* ctx, spice_update_fn fn, void* user);
int spice_ctx_set_graphs(spice_ctx* ctx, int enabled);
int spice_run(spice_ctx* ctx, int32_t nsteps, spice_step const* steps);
/* Device array of population views (valid until the network changes).  Used by
 * the kernels that apply the deliveries already counted for the next step before
 * the host reads neuron state: the reference applies each step's deliveries at
 * the end of that step (snn.cpp:21-25), the device at the start of the next. */
int spice_ctx_views(spice_ctx* ctx, spice_pop_view const** out);
/* Spike ids of `pop` from `age` steps ago (age 0 = its instance's last step),
 * ascending.  Reference neuron_population.h:147-153. */
int spice_spikes(spice_ctx* ctx, int64_t pop, int64_t age, int32_t* out, int64_t cap,
                 int64_t* out_count);
/* Number of steps of spike history kept for the population's instance. */
int64_t spice_history_len(spice_ctx* ctx, int64_t pop);

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
