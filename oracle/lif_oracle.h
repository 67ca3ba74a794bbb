/*
 * CPU oracle for the whole-brain LIF step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's time loop, used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
 * Nothing in the shipped library calls it.
 *
 * Semantics follow flyzero/brain.py:32-58 (`_kernel`, the numba path, which
 * the reference GPU test asserts equal to the GPU path, tests/test_gpu.py:39-47),
 * with the batched extras of flyzero/gpu.py:33-118 (per-fly ext slots for
 * Poisson probability and bias, the counter-hash RNG, per-fly plastic rows).
 * Deliberate, documented differences from the reference:
 *   - synaptic input is accumulated as signed integers (synapse counts in
 *     fixed point, `frac_bits` fractional bits) in an int32 ring and scaled
 *     once on arrival: in = (float)acc * in_scale.  The reference sums float32
 *     mV terms (gpu.py:108, brain.py:56); integer sums are order-independent.
 *   - no snapping of tiny g / near-rest v (gpu.py:52,61).
 *   - the membrane potential is held relative to rest, u = v - v_rest, in
 *     float32 (the model is unchanged: u' = u + (g + bias - u) * decay_m,
 *     spike when u >= v_th - v_rest).  Near threshold |u| ~ 7 mV instead of
 *     |v| ~ 45 mV, so every rounding is 8x finer.
 * All float arithmetic is evaluated exactly as written (build with
 * -ffp-contract=off): no FMA contraction, no reassociation.
 */
#ifndef FLYHIP_LIF_ORACLE_H
#define FLYHIP_LIF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;                  /* neurons */
    int n_delay;            /* ring slots (steps of synaptic delay) */
    int n_ref;              /* refractory steps */
    float decay_syn, decay_m;
    float u_th;             /* f32(v_th - v_rest): threshold relative to rest */
    float u_reset;          /* f32(v_reset - v_rest) */
    float kick;
    float in_scale;         /* mV per fixed-point unit = w_syn * 2^-frac_bits */
    float pw_qscale;        /* fixed-point units per mV of plastic weight */
    const long long* indptr;    /* fixed CSR, n+1 */
    const int* col;             /* post neuron per edge */
    const int* w;               /* fixed-point signed weight per edge */
    int n_pl;                   /* plastic synapses */
    const long long* pl_indptr; /* plastic CSR by pre neuron, n+1 (NULL if n_pl == 0) */
    const int* pl_post;
    int n_ext;                  /* ext slots */
    const int* ext_idx;         /* n entries: slot or -1 */
} orc_net;

typedef struct {
    int B;                  /* flies */
    float* u;               /* [B][n] membrane potential relative to rest, mV (u = v - v_rest) */
    float* g;               /* [B][n] */
    signed char* ref;       /* [B][n] */
    int* ring;              /* [B][n_delay][n] */
    long long t;            /* steps taken (ring phase) */
    unsigned step_no;       /* RNG step counter */
    /* optional FP64 shadow that follows the FP32 spike decisions */
    double* u64;            /* [B][n] or NULL */
    double* g64;            /* [B][n] or NULL */
} orc_state;

typedef struct {
    const float* p_ext;         /* [B][n_ext] per-step Poisson probability (NULL: none) */
    const float* bias_ext;      /* [B][n_ext] mV (NULL: none) */
    const float* pw;            /* [B][n_pl] mV (NULL if n_pl == 0) */
    const unsigned* hits;       /* NULL (hash RNG) or [steps][B][ceil(n_ext/32)] bits */
    const unsigned char* alive; /* NULL (all) or [B] */
    unsigned seed;
    int nthreads;               /* OpenMP threads for the neuron sweep (<=1: serial) */
} orc_input;

typedef struct {
    double max_dv;          /* max |u32 - u64| over the run */
    double max_dv_sub;      /* same, over neuron-steps with -12 <= u64 < u_th mV */
    double max_abs_v;       /* largest |u64| seen */
    double max_dg;          /* max |g32 - g64| over the run */
    long long flips;        /* neuron-steps where the FP64 spike test disagrees */
    double* dv_step;        /* per-step max |v32 - v64| ([steps]) or NULL */
    /* per-step reading: every step advanced once more in FP64 from that step's FP32 state */
    int one_step;           /* in: 1 = measure it */
    double step_max_dv;     /* max |u32 - u64| after one step, before the spike test (mV) */
    double step_max_dg;     /* max |g32 - g64| after one step (mV) */
    long long step_flips;   /* neuron-steps whose one-step FP64 spike test disagrees */
    double* step_dv_step;   /* per-step max of the one-step |u32 - u64| ([steps]) or NULL */
} orc_shadow_report;

unsigned orc_mix(unsigned x);
unsigned orc_hash(unsigned seed, unsigned b, unsigned step, unsigned tid);

/* Advance `steps` steps.  counts[B][n] is incremented (caller zeroes it);
 * raster, when given, receives bit i%32 of word i/32 of [step][b] per spike.
 * Returns the number of spikes, or <0 on bad arguments. */
long long orc_lif_run(const orc_net* net, orc_state* st, const orc_input* in, int steps,
                      int* counts, unsigned* raster, orc_shadow_report* rep);

#ifdef __cplusplus
}
#endif
#endif
