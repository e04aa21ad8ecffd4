/* CPU oracle: plain-C restatement of the reference Spice (denniskb/spice2)
 * hot path.  TEST INFRASTRUCTURE ONLY - see spice_oracle.h.  Compiled with
 * strict IEEE semantics (-fno-fast-math -ffp-contract=off); every float
 * expression keeps the operand order of the reference source it cites.
 */
#include "spice_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[256];
char const* orc_last_error(void) { return g_err; }
static int fail(char const* msg) {
	snprintf(g_err, sizeof g_err, "%s", msg);
	return -1;
}

/* ---------------------------------------------------------------- RNG ---- */

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t fmix64(uint64_t k) { /* random.h:23-30 */
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return k;
}

static const uint64_t MM_C1 = 0x87c37b91114253d5ull, MM_C2 = 0x4cf5ad432745937full;
static const uint64_t MM_H0 = 0x2E4016967F18E81ull, MM_H1 = 0x447567949F9AA86ull;

static void mm_finish(uint64_t* lo, uint64_t* hi, uint64_t len) {
	*lo ^= len;
	*hi ^= len;
	*lo += *hi;
	*hi += *lo;
	*lo = fmix64(*lo);
	*hi = fmix64(*hi);
	*lo += *hi;
	*hi += *lo;
}

orc_u128 orc_murmur3_bytes(void const* data, uint64_t len) { /* random.h:32-106 */
	unsigned char const* d = (unsigned char const*)data;
	uint64_t lo = MM_H0, hi = MM_H1;
	uint64_t nblocks = len / 16;
	for (uint64_t i = 0; i < nblocks; i++) {
		uint64_t k1, k2;
		memcpy(&k1, d + 16 * i, 8);
		memcpy(&k2, d + 16 * i + 8, 8);
		k1 *= MM_C1;
		k1 = rotl64(k1, 31);
		k1 *= MM_C2;
		lo ^= k1;
		lo = rotl64(lo, 27);
		lo += hi;
		lo = lo * 5 + 0x52dce729;
		k2 *= MM_C2;
		k2 = rotl64(k2, 33);
		k2 *= MM_C1;
		hi ^= k2;
		hi = rotl64(hi, 31);
		hi += lo;
		hi = hi * 5 + 0x38495ab5;
	}
	unsigned char const* tail = d + 16 * nblocks;
	uint64_t rem = len & 15, k1 = 0, k2 = 0;
	/* bytes 8..14 feed k2, bytes 0..7 feed k1 (little-endian lanes) */
	if (rem > 8) {
		for (uint64_t b = rem; b-- > 8;) k2 ^= (uint64_t)tail[b] << (8 * (b - 8));
		k2 *= MM_C2;
		k2 = rotl64(k2, 33);
		k2 *= MM_C1;
		hi ^= k2;
	}
	if (rem > 0) {
		uint64_t top = rem > 8 ? 8 : rem;
		for (uint64_t b = top; b-- > 0;) k1 ^= (uint64_t)tail[b] << (8 * b);
		k1 *= MM_C1;
		k1 = rotl64(k1, 31);
		k1 *= MM_C2;
		lo ^= k1;
	}
	mm_finish(&lo, &hi, len);
	orc_u128 r = {lo, hi};
	return r;
}

orc_u128 orc_murmur3_u128(orc_u128 k) { /* random.h:108-139 */
	uint64_t lo = MM_H0, hi = MM_H1;
	k.lo *= MM_C1;
	k.lo = rotl64(k.lo, 31);
	k.lo *= MM_C2;
	lo ^= k.lo;
	lo = rotl64(lo, 27);
	lo += hi;
	lo = lo * 5 + 0x52dce729;
	k.hi *= MM_C2;
	k.hi = rotl64(k.hi, 33);
	k.hi *= MM_C1;
	hi ^= k.hi;
	hi = rotl64(hi, 31);
	hi += lo;
	hi = hi * 5 + 0x38495ab5;
	mm_finish(&lo, &hi, 16);
	orc_u128 r = {lo, hi};
	return r;
}

int orc_seed_chain(uint32_t const* seeds, int nseeds, int count, uint64_t* out) {
	if (nseeds <= 0) return fail("Please provide at least 1 seed to seed_seq");
	orc_u128 s = orc_murmur3_bytes(seeds, 4 * (uint64_t)nseeds);
	for (int k = 0; k < count; k++) {
		out[2 * k]     = s.lo;
		out[2 * k + 1] = s.hi;
		s              = orc_murmur3_u128(s);
	}
	return 0;
}

uint64_t orc_rng_next(orc_rng* r) { /* random.h:225-233 */
	uint64_t const result = r->s0 + r->s1;
	uint64_t const t      = r->s0 ^ r->s1;
	r->s0                 = rotl64(r->s0, 24) ^ t ^ (t << 16);
	r->s1                 = rotl64(t, 37);
	return result;
}

float orc_canonical_f32(orc_rng* r) { /* random.h:236-247, Real=float */
	return (float)(orc_rng_next(r) >> 40) / 16777216.0f;
}

double orc_canonical_f64_leftopen(orc_rng* r) { /* random.h:236-247, Real=double, LeftOpen */
	return (double)((orc_rng_next(r) >> 11) + 1) / 9007199254740992.0;
}

double orc_exponential(orc_rng* r, double scale) { /* random.h:270-272 */
	return -scale * log(orc_canonical_f64_leftopen(r));
}

float orc_uniform_f32(orc_rng* r, float a, float b) { /* random.h:252-257 */
	float const scale = b - a;
	return fmaf(orc_canonical_f32(r), scale, a);
}

float orc_kahan_add(orc_kahan* k, float dt) { /* numeric.h:9-15 */
	float const y = dt - k->c;
	float const t = k->sum + y;
	k->c          = (t - k->sum) - y;
	k->sum        = t;
	return y;
}

/* ------------------------------------------------------ fixed_probability */

int64_t orc_fp_max_degree(int64_t dst_count, double p) { /* topology.cpp:73-76 */
	double const dp = (double)dst_count * p;
	return (int64_t)(dp + 3 * sqrt(dp * (1 - p)));
}

int64_t orc_fp_size(int64_t src_count, int64_t dst_count, double p) {
	return src_count * orc_fp_max_degree(dst_count, p);
}

int64_t orc_fp_generate(int64_t src_count, int64_t dst_count, double p, orc_u128 seed,
                        int64_t* offsets, int32_t* neighbors) { /* topology.cpp:80-112 */
	if (src_count == 0 || dst_count == 0 || p == 0) {
		for (int64_t i = 0; i <= src_count; i++) offsets[i] = 0;
		return 0;
	}
	orc_rng rng           = {seed.lo, seed.hi};
	double const scale    = 1 / p - 1;
	int64_t const max_deg = orc_fp_size(src_count, dst_count, p) / src_count;
	int64_t count         = 0;
	for (int64_t src = 0; src < src_count; src++) {
		offsets[src]  = count;
		int32_t index = 0;
		double noise  = 0;
		for (;;) {
			noise += orc_exponential(&rng, scale);
			int32_t const dst = index + (int32_t)round(noise);
			if ((dst >= dst_count) | (index >= max_deg)) break;
			neighbors[count++] = dst;
			index++;
		}
	}
	offsets[src_count] = count;
	return count;
}

uint64_t orc_fnv1a(uint64_t h, void const* p, uint64_t n) {
	unsigned char const* b = (unsigned char const*)p;
	for (uint64_t i = 0; i < n; i++) {
		h ^= b[i];
		h *= 0x100000001b3ull;
	}
	return h;
}

uint64_t orc_csr_hash(int64_t src_count, int64_t const* offsets, int32_t const* neighbors) {
	uint64_t h = 0xcbf29ce484222325ull;
	if (src_count == 0) return h; /* the reference CSR has no offsets at all then */
	h = orc_fnv1a(h, offsets, 8 * (uint64_t)(src_count + 1));
	h = orc_fnv1a(h, neighbors, 4 * (uint64_t)offsets[src_count]);
	return h;
}

/* ------------------------------------------------------------- network --- */

typedef struct {
	int model;
	int64_t size;
	float rate;
	int init;
	float *V, *Gex, *Gin;
	int32_t* Twait;
	/* delay window: ring of max_delay per-step lists (neuron_population.h:116-134) */
	int32_t* ring;     /* max_delay * size */
	int64_t* ring_len; /* max_delay */
	int64_t nsteps;    /* steps recorded so far */
	uint64_t* hist;    /* plastic populations: bit b = spiked b steps ago */
} opop;

typedef struct {
	int src, dst;
	int64_t delay;
	float weight;
	int field;
	int64_t *offsets;
	int32_t* neighbors;
	int owned;
	int rule;          /* 0: fixed weight */
	float w0;
	float *W, *Zpre, *Zpost; /* per synapse, CSR order */
	uint64_t* ages;    /* per source: (last touch + 1) | delivered << 63 */
} oproj;

struct orc_net {
	float dt;
	int64_t max_delay;
	int64_t time;
	orc_kahan simtime;
	orc_u128 seed;
	opop pops[16];
	int npops;
	oproj projs[64];
	int nprojs;
	int64_t events;
};

static orc_u128 seed_take(orc_net* n) { /* seed_seq::operator++(int): returns the old value */
	orc_u128 old = n->seed;
	n->seed      = orc_murmur3_u128(n->seed);
	return old;
}

orc_net* orc_net_create(float dt, float max_delay, uint32_t const* seeds, int nseeds) {
	if (nseeds <= 0) {
		fail("Please provide at least 1 seed to seed_seq");
		return NULL;
	}
	orc_net* n   = (orc_net*)calloc(1, sizeof(orc_net));
	n->dt        = dt;
	n->max_delay = (int64_t)roundf(max_delay / dt); /* snn.h:19 */
	n->seed      = orc_murmur3_bytes(seeds, 4 * (uint64_t)nseeds);
	return n;
}

void orc_net_destroy(orc_net* n) {
	if (!n) return;
	for (int i = 0; i < n->npops; i++) {
		opop* p = &n->pops[i];
		free(p->V);
		free(p->Gex);
		free(p->Gin);
		free(p->Twait);
		free(p->ring);
		free(p->ring_len);
		free(p->hist);
	}
	for (int i = 0; i < n->nprojs; i++) {
		if (n->projs[i].owned) {
			free(n->projs[i].offsets);
			free(n->projs[i].neighbors);
		}
		free(n->projs[i].W);
		free(n->projs[i].Zpre);
		free(n->projs[i].Zpost);
		free(n->projs[i].ages);
	}
	free(n);
}

int orc_add_population(orc_net* n, int model, int64_t size, float rate, int init) {
	if (n->npops >= 16 || size < 0) return fail("bad population");
	opop* p  = &n->pops[n->npops];
	p->model = model;
	p->size  = size;
	p->rate  = rate;
	p->init  = init;
	if (model != ORC_POISSON) {
		/* stateful_neuron_adapter consumes a seed even without init (neuron_population.h:60) */
		orc_u128 s = seed_take(n);
		orc_rng rng = {s.lo, s.hi};
		p->V        = (float*)calloc(size ? size : 1, 4);
		p->Gex      = (float*)calloc(size ? size : 1, 4);
		p->Gin      = (float*)calloc(size ? size : 1, 4);
		p->Twait    = (int32_t*)calloc(size ? size : 1, 4);
		if (model == ORC_COBA) {
			for (int64_t i = 0; i < size; i++) p->V[i] = -0.06f; /* vogels.cpp:12 */
			if (init == 1)
				for (int64_t i = 0; i < size; i++) p->V[i] = orc_uniform_f32(&rng, -0.06f, -0.05f);
		}
	}
	int64_t md  = n->max_delay > 0 ? n->max_delay : 1;
	p->ring     = (int32_t*)malloc(sizeof(int32_t) * (size_t)(md * (size ? size : 1)));
	p->ring_len = (int64_t*)calloc(md, sizeof(int64_t));
	return n->npops++;
}

static int connect_common(orc_net* n, int src, int dst, float delay, int64_t* d_out) {
	int64_t const d = (int64_t)roundf(delay / n->dt); /* snn.h:34 (std::round of a float) */
	if (!(d >= 1)) return fail("Assertion failed (snn.h:35): d >= 1 && \"The delay must be at least 1dt.\"");
	if (!(d <= n->max_delay))
		return fail("Assertion failed (snn.h:36): d <= _max_delay");
	if (src < 0 || src >= n->npops || dst < 0 || dst >= n->npops || n->nprojs >= 64)
		return fail("bad projection");
	if (n->pops[dst].model == ORC_POISSON) return fail("target must be stateful");
	*d_out = d;
	return 0;
}

int orc_connect(orc_net* n, int src, int dst, double p, float delay, float weight, int field) {
	int64_t d;
	if (connect_common(n, src, dst, delay, &d)) return -1;
	if (!(0 <= p && p <= 1)) return fail("Assertion failed (topology.cpp:73): 0 <= p && p <= 1");
	int64_t const S = n->pops[src].size, D = n->pops[dst].size;
	if (S == 0) return fail("Assertion failed (topology.cpp:106): offsets.size() > src_count");
	orc_u128 s     = seed_take(n); /* synapse_population.h:31 */
	oproj* pr      = &n->projs[n->nprojs];
	pr->offsets    = (int64_t*)calloc(S + 1, 8);
	int64_t sz     = orc_fp_size(S, D, p);
	pr->neighbors  = (int32_t*)calloc(sz ? sz : 1, 4);
	pr->owned      = 1;
	orc_fp_generate(S, D, p, s, pr->offsets, pr->neighbors);
	pr->src    = src;
	pr->dst    = dst;
	pr->delay  = d;
	pr->weight = weight;
	pr->field  = field;
	return n->nprojs++;
}

int orc_connect_csr(orc_net* n, int src, int dst, int64_t const* offsets, int32_t const* neighbors,
                    float delay, float weight, int field) {
	int64_t d;
	if (connect_common(n, src, dst, delay, &d)) return -1;
	(void)seed_take(n);
	oproj* pr     = &n->projs[n->nprojs];
	pr->offsets   = (int64_t*)offsets;
	pr->neighbors = (int32_t*)neighbors;
	pr->owned     = 0;
	pr->src       = src;
	pr->dst       = dst;
	pr->delay     = d;
	pr->weight    = weight;
	pr->field     = field;
	return n->nprojs++;
}

int orc_connect_plastic(orc_net* n, int src, int dst, double p, float delay, int rule, float w0) {
	if (rule != ORC_RULE_BRUNELPLUS && rule != ORC_RULE_STDP_PAIR) return fail("unknown plastic rule");
	int const c = orc_connect(n, src, dst, p, delay, 0.0f, ORC_V);
	if (c < 0) return c;
	oproj* pr       = &n->projs[c];
	int64_t const S = n->pops[src].size;
	int64_t const E = pr->offsets[S];
	pr->rule        = rule;
	pr->w0          = w0;
	pr->W           = (float*)calloc(E ? E : 1, 4);
	pr->Zpre        = (float*)calloc(E ? E : 1, 4);
	pr->Zpost       = (float*)calloc(E ? E : 1, 4);
	pr->ages        = (uint64_t*)calloc(S ? S : 1, 8); /* synapse_population.h:45 */
	float const W0  = rule == ORC_RULE_BRUNELPLUS ? 1e-4f : w0;
	for (int64_t e = 0; e < E; e++) pr->W[e] = W0;
	if (rule == ORC_RULE_STDP_PAIR) (void)seed_take(n); /* PerSynapseInit: synapse_population.h:34-35 */
	opop* sp = &n->pops[src];                           /* snn.h:46-47: source->plastic() */
	if (!sp->hist) sp->hist = (uint64_t*)calloc(sp->size ? sp->size : 1, 8);
	return c;
}

int orc_get_synapses(orc_net* n, int proj, float* W, float* Zpre, float* Zpost) {
	oproj* pr = &n->projs[proj];
	if (!pr->rule) return fail("not a plastic projection");
	int64_t const E = pr->offsets[n->pops[pr->src].size];
	memcpy(W, pr->W, 4 * (size_t)E);
	memcpy(Zpre, pr->Zpre, 4 * (size_t)E);
	memcpy(Zpost, pr->Zpost, 4 * (size_t)E);
	return 0;
}

/* The rules' update() and skip() (brunel+.cpp:71-99 and the pair rule above). */
static void syn_update(oproj const* pr, float* W, float* Zpre, float* Zpost, float dt, int pre, int post) {
	if (pr->rule == ORC_RULE_COUNT) {
		*W += 1.0f;
	} else if (pr->rule == ORC_RULE_BRUNELPLUS) {
		float const TstdpInv = 1.0f / 0.02f;
		float const dtInv    = 1.0f / dt;
		float const a  = (float)pre * 0.0202f * *W * expf(-*Zpost * dtInv);
		float const b  = (float)post * 0.01f * (1.0f - *W) * expf(-*Zpre * dtInv);
		float w        = *W - a + b;
		*W             = w < 0.0f ? 0.0f : w > 0.0003f ? 0.0003f : w;
		*Zpre += (float)pre;
		*Zpost += (float)post;
		*Zpre -= *Zpre * dt * TstdpInv;
		*Zpost -= *Zpost * dt * TstdpInv;
	} else {
		float const tauInv = 1.0f / 0.02f;
		float const wmax   = 2.0f * pr->w0;
		if (pre) {
			float const w = *W - 0.0105f * pr->w0 * *Zpost;
			*W            = w < 0.0f ? 0.0f : w > wmax ? wmax : w;
		}
		if (post) {
			float const w = *W + 0.01f * pr->w0 * *Zpre;
			*W            = w < 0.0f ? 0.0f : w > wmax ? wmax : w;
		}
		*Zpre += (float)pre;
		*Zpost += (float)post;
		*Zpre -= *Zpre * dt * tauInv;
		*Zpost -= *Zpost * dt * tauInv;
	}
}

static void syn_skip(oproj const* pr, float* W, float* Zpre, float* Zpost, float dt, int64_t steps) {
	if (pr->rule == ORC_RULE_COUNT) {
		*W += (float)steps;
		return;
	}
	float const tauInv = 1.0f / 0.02f;
	double const f     = pow((double)(1 - dt * tauInv), (double)steps); /* std::pow(float, Int) */
	*Zpre  = (float)((double)*Zpre * f);
	*Zpost = (float)((double)*Zpost * f);
}

/* Brings synapse e of a spiking (or swept) source up to `time` with the target's
 * spike history (synapse_population.h:97-115).  Shift counts are taken mod 64, as
 * the reference's x86-64 code does for ~0 >> 64. */
static int catch_up_hist(oproj* pr, uint64_t const* dst_hist, int64_t e, uint64_t ages, int64_t time,
                         float dt) {
	int const pre      = (int)(ages >> 63);
	int64_t const age  = (int64_t)(ages & ~(1ull << 63));
	if (!(time >= age)) return 0;
	if (!dst_hist) return fail("Assertion failed (synapse_population.h:97): edge.first < dst_history.size()");
	int64_t const prefix = 63 + pre - time + age;
	uint64_t hist        = dst_hist[pr->neighbors[e]];
	if (pre) syn_update(pr, &pr->W[e], &pr->Zpre[e], &pr->Zpost[e], dt, 1, (int)((hist >> (time - age)) & 1));
	hist &= ~0ull >> (prefix & 63);
	int64_t p = prefix;
	while (hist) {
		int64_t const lz = __builtin_clzll(hist);
		syn_skip(pr, &pr->W[e], &pr->Zpre[e], &pr->Zpost[e], dt, lz - p);
		syn_update(pr, &pr->W[e], &pr->Zpre[e], &pr->Zpost[e], dt, 0, 1);
		hist ^= 1ull << (63 - lz);
		p = lz + 1;
	}
	syn_skip(pr, &pr->W[e], &pr->Zpre[e], &pr->Zpost[e], dt, 64 - p);
	return 0;
}

static int catch_up(orc_net* n, oproj* pr, int64_t e, uint64_t ages, int64_t time, float dt) {
	return catch_up_hist(pr, n->pops[pr->dst].hist, e, ages, time, dt);
}

/* ---- one projection on its own (the reference's unit tests of the plastic
 *      path, test/detail/synapse_population.cpp:84-167) ---------------------- */
struct orc_syn {
	oproj pr;
	int64_t S;
};

orc_syn* orc_syn_create(int64_t S, int64_t D, int64_t const* offsets, int32_t const* neighbors, int rule,
                        float w0) {
	(void)D;
	orc_syn* s      = (orc_syn*)calloc(1, sizeof(orc_syn));
	int64_t const E = offsets[S];
	s->S            = S;
	s->pr.offsets   = (int64_t*)malloc(8 * (size_t)(S + 1));
	s->pr.neighbors = (int32_t*)malloc(4 * (size_t)(E ? E : 1));
	memcpy(s->pr.offsets, offsets, 8 * (size_t)(S + 1));
	memcpy(s->pr.neighbors, neighbors, 4 * (size_t)E);
	s->pr.rule  = rule;
	s->pr.w0    = w0;
	s->pr.W     = (float*)calloc(E ? E : 1, 4);
	s->pr.Zpre  = (float*)calloc(E ? E : 1, 4);
	s->pr.Zpost = (float*)calloc(E ? E : 1, 4);
	s->pr.ages  = (uint64_t*)calloc(S ? S : 1, 8);
	float const W0 = rule == ORC_RULE_BRUNELPLUS ? 1e-4f : rule == ORC_RULE_STDP_PAIR ? w0 : 0.0f;
	for (int64_t e = 0; e < E; e++) s->pr.W[e] = W0;
	return s;
}

void orc_syn_destroy(orc_syn* s) {
	if (!s) return;
	free(s->pr.offsets);
	free(s->pr.neighbors);
	free(s->pr.W);
	free(s->pr.Zpre);
	free(s->pr.Zpost);
	free(s->pr.ages);
	free(s);
}

int orc_syn_update(orc_syn* s, int64_t time, float dt, uint64_t const* dst_hist) {
	for (int64_t src = 0; src < s->S; src++) {
		for (int64_t e = s->pr.offsets[src]; e < s->pr.offsets[src + 1]; e++)
			if (catch_up_hist(&s->pr, dst_hist, e, s->pr.ages[src], time, dt)) return -1;
		s->pr.ages[src] = (uint64_t)(time + 1);
	}
	return 0;
}

/* deliver(syn, to): the count rule assigns the synapse's update count to the
 * target (received_count = update_count), the others add W. */
int orc_syn_deliver(orc_syn* s, int64_t time, float dt, int32_t const* spikes, int64_t nspikes,
                    uint64_t const* dst_hist, float* dst) {
	for (int64_t k = 0; k < nspikes; k++) {
		int64_t const src = spikes[k];
		for (int64_t e = s->pr.offsets[src]; e < s->pr.offsets[src + 1]; e++) {
			if (catch_up_hist(&s->pr, dst_hist, e, s->pr.ages[src], time, dt)) return -1;
			if (s->pr.rule == ORC_RULE_COUNT)
				dst[s->pr.neighbors[e]] = s->pr.W[e];
			else
				dst[s->pr.neighbors[e]] += s->pr.W[e];
		}
		s->pr.ages[src] = (uint64_t)(time + 1) | (1ull << 63);
	}
	return 0;
}

static void pop_update(orc_net* n, opop* p, float dt, orc_rng* rng) {
	int64_t const md = n->max_delay;
	int64_t const slot = p->nsteps % md;
	int32_t* out = p->ring + slot * p->size;
	int64_t cnt  = 0;
	if (p->model == ORC_POISSON) { /* brunel.cpp:27-30 */
		for (int64_t i = 0; i < p->size; i++)
			if (orc_canonical_f32(rng) < (p->rate * dt)) out[cnt++] = (int32_t)i;
	} else if (p->model == ORC_LIF) { /* brunel.cpp:45-61 */
		float const TmemInv = (float)(1.0 / 0.02);
		float const Vrest   = 0.0f;
		int32_t const Tref  = 20;
		float const Vthres  = (float)0.02;
		for (int64_t i = 0; i < p->size; i++) {
			if (--p->Twait[i] <= 0) {
				if (p->V[i] > Vthres) {
					p->V[i]     = Vrest;
					p->Twait[i] = Tref;
					out[cnt++]  = (int32_t)i;
					continue;
				}
				float const a = Vrest - p->V[i];
				float const b = dt * TmemInv;
				p->V[i]       = p->V[i] + a * b;
			}
		}
	} else { /* vogels.cpp:18-45 */
		int32_t const Tref  = 50;
		float const Vrest   = (float)-0.06;
		float const Vthres  = (float)-0.05;
		float const TmemInv = (float)(1.0f / 0.02);
		float const Eex     = 0.0f;
		float const Ein     = (float)-0.08;
		float const Ibg     = (float)0.02;
		float const TexInv  = (float)(1.0f / 0.005);
		float const TinInv  = (float)(1.0f / 0.01);
		for (int64_t i = 0; i < p->size; i++) {
			float V = p->V[i], Gex = p->Gex[i], Gin = p->Gin[i];
			if (--p->Twait[i] <= 0) {
				if (V > Vthres) {
					V           = Vrest;
					p->Twait[i] = Tref;
					out[cnt++]  = (int32_t)i;
				} else {
					float s = (Vrest - V) + Gex * (Eex - V);
					s       = s + Gin * (Ein - V);
					s       = s + Ibg;
					V       = V + s * (dt * TmemInv);
				}
			}
			Gex      = Gex - Gex * (dt * TexInv);
			Gin      = Gin - Gin * (dt * TinInv);
			p->V[i]  = V;
			p->Gex[i] = Gex;
			p->Gin[i] = Gin;
		}
	}
	if (p->hist) { /* neuron_population.h:122-128 */
		for (int64_t i = 0; i < p->size; i++) p->hist[i] <<= 1;
		for (int64_t k = 0; k < cnt; k++) p->hist[out[k]] |= 1;
	}
	p->ring_len[slot] = cnt;
	p->nsteps++;
}

int64_t orc_spikes(orc_net* n, int pop, int64_t age, int32_t const** ids) {
	opop* p          = &n->pops[pop];
	int64_t const md = n->max_delay;
	int64_t const have = p->nsteps < md ? p->nsteps : md;
	if (!(0 <= age && age < have)) {
		fail("Assertion failed (neuron_population.h:148): 0 <= age && age < _spike_counts.size()");
		return -1;
	}
	int64_t const slot = (p->nsteps - 1 - age) % md;
	*ids               = p->ring + slot * p->size;
	return p->ring_len[slot];
}

int orc_step(orc_net* n) { /* snn.cpp:7-28 */
	float const dt = orc_kahan_add(&n->simtime, n->dt);
	if (n->simtime.sum >= 1) n->simtime.sum = 0;
	orc_u128 s  = seed_take(n);
	orc_rng rng = {s.lo, s.hi};
	for (int i = 0; i < n->npops; i++) pop_update(n, &n->pops[i], dt, &rng);
	if (n->time % 64 == 0) /* snn.cpp:17-19: catch every plastic synapse up (nominal dt) */
		for (int c = 0; c < n->nprojs; c++) {
			oproj* pr = &n->projs[c];
			if (!pr->rule) continue;
			for (int64_t src = 0; src < n->pops[pr->src].size; src++) {
				for (int64_t e = pr->offsets[src]; e < pr->offsets[src + 1]; e++)
					if (catch_up(n, pr, e, pr->ages[src], n->time, n->dt)) return -1;
				pr->ages[src] = (uint64_t)(n->time + 1);
			}
		}
	for (int c = 0; c < n->nprojs; c++) {
		oproj* pr = &n->projs[c];
		if (n->time < pr->delay - 1) continue;
		int32_t const* ids;
		int64_t const cnt = orc_spikes(n, pr->src, pr->delay - 1, &ids);
		opop* dst         = &n->pops[pr->dst];
		float* f = pr->field == ORC_V ? dst->V : pr->field == ORC_GEX ? dst->Gex : dst->Gin;
		for (int64_t k = 0; k < cnt; k++) { /* synapse_population.h:88-134 */
			int64_t const src = ids[k];
			int64_t const a = pr->offsets[src], b = pr->offsets[src + 1];
			if (pr->rule) {
				for (int64_t e = a; e < b; e++) {
					if (catch_up(n, pr, e, pr->ages[src], n->time, n->dt)) return -1;
					f[pr->neighbors[e]] += pr->W[e]; /* deliver(syn, to): to.V += syn.W */
				}
				pr->ages[src] = (uint64_t)(n->time + 1) | (1ull << 63);
			} else {
				for (int64_t e = a; e < b; e++) f[pr->neighbors[e]] += pr->weight;
			}
			n->events += b - a;
		}
	}
	n->time++;
	return 0;
}

int orc_get_field(orc_net* n, int pop, int field, void* out) {
	opop* p = &n->pops[pop];
	if (p->model == ORC_POISSON) return fail("stateless population");
	float* f = field == ORC_V ? p->V : field == ORC_GEX ? p->Gex : p->Gin;
	memcpy(out, f, 4 * (size_t)p->size);
	return 0;
}

int orc_get_twait(orc_net* n, int pop, int32_t* out) {
	opop* p = &n->pops[pop];
	if (p->model == ORC_POISSON) return fail("stateless population");
	memcpy(out, p->Twait, 4 * (size_t)p->size);
	return 0;
}

int64_t orc_events(orc_net* n) { return n->events; }

uint64_t orc_state_hash(orc_net* n) {
	uint64_t h = 0xcbf29ce484222325ull;
	for (int i = 0; i < n->npops; i++) {
		opop* p = &n->pops[i];
		for (int64_t k = 0; k < p->size && p->model != ORC_POISSON; k++) {
			if (p->model == ORC_LIF) { /* lif::neuron {float V; int Twait;} */
				h = orc_fnv1a(h, &p->V[k], 4);
				h = orc_fnv1a(h, &p->Twait[k], 4);
			} else { /* vogels lif::neuron {V, Gex, Gin, Twait} */
				h = orc_fnv1a(h, &p->V[k], 4);
				h = orc_fnv1a(h, &p->Gex[k], 4);
				h = orc_fnv1a(h, &p->Gin[k], 4);
				h = orc_fnv1a(h, &p->Twait[k], 4);
			}
		}
	}
	return h;
}

int64_t orc_csr_synapses(orc_net* n, int proj) {
	oproj* pr = &n->projs[proj];
	return pr->offsets[n->pops[pr->src].size];
}

uint64_t orc_proj_hash(orc_net* n, int proj) {
	oproj* pr = &n->projs[proj];
	return orc_csr_hash(n->pops[pr->src].size, pr->offsets, pr->neighbors);
}
