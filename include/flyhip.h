/*
 * flyhip -- MI355X-native whole-brain leaky integrate-and-fire step (C ABI).
 *
 * The reference reaches its GPU kernels through CuPy's RawModule with
 * positional arguments (flyzero/gpu.py:124-128, 242-251) behind the duck-typed
 * Python classes `Brain` (flyzero/brain.py:79-172) and `BatchBrain`
 * (flyzero/gpu.py:131-254).  This header is the boundary that replaces those
 * launches: plain pointers and sizes, no framework types.  All array pointers
 * are DEVICE pointers owned by the caller; the engine keeps only its own
 * scratch (spike work lists, counters, degree table, graphs).
 *
 * Every function returns 0 on success or a negative FZ_E* code; the message
 * for the calling thread is in fz_last_error().  A handle is not re-entrant:
 * use it from one host thread.  All work is enqueued on the given stream
 * (a hipStream_t passed as void*; NULL = the default stream).
 */
#ifndef FLYHIP_H
#define FLYHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FZ_ABI_VERSION 4

#define FZ_OK 0
#define FZ_E_ARG (-1)      /* invalid argument (null pointer, size out of range) */
#define FZ_E_STATE (-2)    /* call out of order (e.g. fz_run before fz_bind) */
#define FZ_E_HIP (-3)      /* a HIP runtime call failed */
#define FZ_E_OVERFLOW (-4) /* a device-side capacity check fired */

typedef struct fz_engine fz_engine;
typedef struct fz_graph fz_graph;

/* Model constants, derived on the host exactly as flyzero/brain.py:92-96
 * (n_delay, n_ref, decay_m = f32(dt/tau_m), decay_syn = f32(exp(-dt/tau_syn)),
 * kick = f32(f_poisson*w_syn)).  The membrane potential is held relative to
 * rest, u = v - v_rest: u_th = f32(v_th - v_rest), u_reset = f32(v_reset - v_rest).
 * Synaptic input is integer: a ring value `acc` arrives as (float)acc * in_scale
 * mV, in_scale = w_syn * 2^-frac_bits.  Plastic weights are given in mV and
 * quantised per use as rintf(pw * pw_qscale) fixed-point units. */
typedef struct fz_params {
    float decay_syn, decay_m, u_th, u_reset, kick, in_scale, pw_qscale;
    int n_ref, n_delay;
} fz_params;

/* Caller-owned per-fly state (flyzero/gpu.py:186-199 holds the same arrays). */
typedef struct fz_state {
    int B;              /* flies */
    float* u;           /* [B][n] membrane potential relative to rest, mV */
    float* g;           /* [B][n] synaptic conductance, mV */
    signed char* ref;   /* [B][n] refractory steps left */
    int* ring;          /* [B][n_delay][n] int32 delay ring, fixed-point synapse counts */
    int* counts;        /* [B][n] spike counts of the last window */
    long long* clock;   /* [4]: t (ring phase), step_no (RNG step), block parity, reserved */
    int ring_slots;     /* n_delay (push engine) or 2 * n_delay (tile engine) */
    int* spk_idx;       /* tile engine: [B][3 n_delay][n] spiking neurons per step (else NULL) */
    int* spk_cnt;       /* tile engine: [B][3 n_delay] list lengths, zero at rest (else NULL) */
} fz_state;

/* Per-window inputs (flyzero/gpu.py:174-179, 201-226; plastic weights gpu.py:162-163). */
typedef struct fz_inputs {
    int n_ext;                  /* ext slots */
    const int* ext_idx;         /* [n] slot or -1 */
    const float* p_ext;         /* [B][n_ext] Poisson probability per step (Hz*dt/1000) */
    const float* bias_ext;      /* [B][n_ext] steady bias, mV */
    const float* pw;            /* [B][n_pl] plastic weights, mV (NULL if n_pl == 0) */
    const unsigned char* alive; /* [B] 0 = retired fly (skipped), or NULL = all alive */
    unsigned seed;              /* RNG seed (flyzero/gpu.py:172) */
    unsigned fly_base;          /* the draw numbers fly b as b + fly_base (0 in the reference): a
                                 * brain holding flies [fly_base, fly_base + B) of a larger batch
                                 * draws the same noise as that batch does for them */
} fz_inputs;

/* Launch shape (chosen by the autotuner; zero fields take defaults). */
typedef struct fz_tune {
    int upd_threads;   /* LIF-update workgroup size, multiple of 64, <= 1024 */
    int push_threads;  /* push workgroup size, multiple of 64, <= 1024 */
    int push_chunk;    /* edges per push work item */
    int push_grid;     /* push workgroups (grid-stride) */
    int block_steps;   /* steps per temporal block, 1..min(n_delay, 32) */
    int tile_stage;    /* tile engine: spike-list entries staged per thread, 1 or 4 */
    int tile_source;   /* tile engine: 0 = every tile looks its segments up in the segment table
                        * (one 8-byte (first edge, length) pair per spike and tile: an engine copy
                        * of the table, built by fz_set_tiles when it takes at most 1 GiB, else
                        * as in 2); 1 = a prologue kernel per block expands each spike once into
                        * per-tile buckets of pieces (first edge, length <= 2, step): engine
                        * scratch, allocated on first use, 2 x 4 B x pieces bound per fly; needs
                        * fewer than 2^24 fixed edges; 2 = as 0 from the caller's int32 table (or
                        * its uint16 row-offset copy): three loads per spike and tile */
    int plastic_pull;  /* tile engine with plastic synapses: 0 = every spiking plastic row is pushed
                        * at spike time into the arrival ring (global atomics); 1 = a spike whose
                        * arrival lies in the window it was emitted in is pulled by the target
                        * tiles with the fly's weights (plastic edges ride in the segment lookups
                        * and buckets), only the others are pushed.  Arrivals are the same; pushes
                        * are global atomics per plastic edge, pulls cost a table load per spike
                        * and tile.  1 needs packed edges and a plastic CSR sorted by target within
                        * rows (checked by fz_set_tiles); otherwise plastic rows are pushed. */
    int block_parts;   /* tile engine: 1 (or 0) = a block's arrivals of all its steps sit in LDS at
                        * once; 2 = the block runs as two parts (steps [0, ceil(K/2)) and the rest),
                        * each part's arrivals pulled into LDS in turn while the neuron state stays
                        * in registers: half the LDS per workgroup, so more workgroups per CU.  The
                        * same results either way. */
    int acc_bits;      /* tile engine: bits per block step's arrival sum in LDS.  32 = one int per
                        * step; 16 = two steps per int (16-bit sums, the low one biased): about half
                        * the LDS per workgroup, so more workgroups per CU.  16 is allowed with
                        * packed edges, tile_source 1, one-part blocks and no plastic synapses, when
                        * every neuron's summed |input weight| (in the packed integer units) is
                        * below 32768, which keeps every partial sum inside int16: the same results
                        * either way.  0 = 16 where allowed, else 32. */
} fz_tune;

int fz_version(void);
const char* fz_last_error(void);
/* Which build is loaded: "src <hash>; <compiler version>", the hash covering the HIP sources and
 * headers the library was compiled from (no counterpart in the reference: for callers and
 * harnesses that must show which native code ran). */
const char* fz_build_info(void);

/* Engine over a device CSR of the FIXED (non-plastic) synapses.  Replaces the
 * device upload in flyzero/gpu.py:143-163.  indptr int32[n+1], col int32[nnz],
 * w int16[nnz] (w_bits = 16: whole synapse counts, shifted by frac_bits on use)
 * or int32[nnz] (w_bits = 32: fixed-point values).  Arrays must outlive the engine. */
int fz_create(int n, long long nnz, const int* indptr, const int* col, const void* w, int w_bits,
              int frac_bits, const fz_params* p, fz_engine** out);
int fz_destroy(fz_engine* e);

/* Plastic synapses as a CSR by pre neuron (flyzero/gpu.py:150-161): pl_indptr
 * int32[n+1], pl_post int32[n_pl]; entry k uses weight pw[b][k]. n_pl may be 0. */
int fz_set_plastic(fz_engine* e, int n_pl, const int* pl_indptr, const int* pl_post, void* stream);

/* Bind caller state for B flies and size the scratch for it. */
int fz_bind(fz_engine* e, const fz_state* st);

/* Switch the engine to arrivals pulled per neuron tile (call before fz_bind).  tile: neurons
 * per workgroup (multiple of 64); tb int32[n][n_tiles+1]: first fixed edge of row i whose
 * post neuron is >= t*tile (columns sorted within rows); ptg_ptr int32[n_tiles+1] / ptg_idx:
 * the plastic synapses' post neurons of each tile (NULL without plastic synapses).  The ring
 * then has 2*n_delay slots and carries plastic arrivals only; fixed arrivals come from the
 * per-step spike lists in fz_state.  The engine keeps a packed copy of the fixed edges
 * (integer value and target offset inside its tile in one int32 per edge) when every value
 * fits beside the offset; fixed weights must not change after this call. */
int fz_set_tiles(fz_engine* e, int tile, int n_tiles, const int* tb, const int* ptg_ptr, const int* ptg_idx);

/* Advance every live fly `steps` steps (one window; flyzero/gpu.py:228-254,
 * flyzero/brain.py:133-172).  counts are overwritten with this window's spike
 * counts.  hits: NULL (counter-hash RNG, flyzero/gpu.py:66-67) or precomputed
 * Poisson hits uint32[steps][B][ceil(n_ext/32)], bit j = ext slot j kicked.
 * raster: NULL or uint32[steps][B][ceil(n/32)] receiving every spike. */
int fz_run(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
           const fz_tune* tune, void* stream);

/* fz_run without device memory of the engine's own, for callers that cannot allocate device
 * memory or take memory the library allocated for them (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It has no counterpart in the reference.
 * Push engine: the per-fly spike work list that fz_bind otherwise allocates is carved out of a
 * fixed device array of FZ_STATIC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends (about 3.7 MB per fly of the 138,639-neuron, 15.1 M-connection fly, so
 * up to 8 such flies); a batch that needs more is refused with FZ_E_ARG.  Tile engine: only the
 * segment-table source (tile_source 0), which uses caller state only; the bucket source is
 * refused.  The call allocates nothing and calls no allocating routine; the caller still passes
 * (and binds) the state, inputs and outputs.  Every call on one GPU shares that GPU's array and
 * leaves the last call's contents in it (its counters are cleared at the start of each call):
 * calls that overlap on one GPU race, so serialise them (one stream, or events).  The stream
 * must belong to the device the engine was created on. */
#define FZ_STATIC_SCRATCH_BYTES (32ll << 20)
int fz_run_static(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
                  const fz_tune* tune, void* stream);

/* As fz_run, timing every kernel with a start and a stop HIP event that the
 * runtime stamps from the kernel's own dispatch (hipExtLaunchKernel) on
 * `stream`; waits for the stream.  Out: summed milliseconds and launch counts
 * of the update kernels (the tile kernel with the tile engine) and of the push
 * kernels (the bucket prologue with the tile engine's bucket source). */
int fz_run_timed(fz_engine* e, const fz_inputs* in, int steps, const fz_tune* tune, void* stream,
                 float* upd_ms, int* upd_launches, float* push_ms, int* push_launches);

/* One window of `steps` steps as a replayable HIP graph.  The graph reads the
 * clock, inputs and weights from device memory at every replay, so callers
 * update them in place between replays. */
int fz_graph_create(fz_engine* e, const fz_inputs* in, int steps, const fz_tune* tune, fz_graph** out);
int fz_graph_launch(fz_graph* gr, void* stream);
int fz_graph_destroy(fz_graph* gr);

/* acc[col] += weight (fixed-point units) for every fixed edge of the n_list
 * neurons in `list` (int32, device): the spike scatter of flyzero/brain.py:123-131
 * (`Brain._propagate`) as an integer sum.  acc is int32[n]. */
int fz_propagate(fz_engine* e, int n_list, const int* list, int* acc, void* stream);

/* ---- plasticity and readout (stand-alone: they act on the counts of a window) ---- */

/* One frame of reward-modulated plasticity on the plastic synapses (the exploratory-Hebbian
 * rule of flyzero/learning.py:116-131, batched as flyzero/batch.py:482-495 with the per-fly
 * averaging of batch.py:158-172).  Float64 throughout, weights in synapse units; the simulation
 * uses count = rint(w) and pw = count * w_syn.  Every array is a device pointer. */
typedef struct fz_reward_args {
    int B;                      /* slots (simulated flies) */
    int n;                      /* neurons */
    int n_pl;                   /* plastic synapses */
    int n_post;                 /* distinct post neurons of plastic synapses */
    int n_flies;                /* learners; a learner averages its slots' changes */
    int any_learn;              /* 0: traces and means only (flyzero/batch.py:493) */
    const int* counts;          /* [B][n] spike counts of the window */
    const int* pre;             /* [n_pl] pre neuron */
    const int* post;            /* [n_pl] post neuron */
    const int* post_k;          /* [n_pl] index of post in upost */
    const int* upost;           /* [n_post] */
    const signed char* sign;    /* [n_pl] Dale sign, +1 / -1 */
    const double* wmax;         /* [n_pl] cap, synapses */
    float* elig;                /* [B][n_pl] eligibility trace */
    double* w;                  /* [B][n_pl] weight, synapses */
    int* count;                 /* [B][n_pl] rint(w) */
    float* pw;                  /* [B][n_pl] count * w_syn, mV (the engine's plastic weights) */
    double* post_mean;          /* [B][n_post] running mean of post counts */
    double* r_mean;             /* [B] running mean reward */
    const double* reward;       /* [B] this frame's reward */
    const unsigned char* learn; /* [B] slot learns this frame */
    const int* fly_ptr;         /* [n_flies + 1] slots of each learner ... */
    const int* fly_slots;       /* ... ascending slot indices */
    double ae, ap, ar;          /* exp(-w/tau_e), 1-exp(-w/tau_post), 1-exp(-w/tau_R) */
    double eta;                 /* synapses per unit of (reward x eligibility) */
    float w_syn;                /* mV per synapse */
    double* adv_out;            /* [B] advantage R - <R> before the update */
    const double* eta_slot;     /* [B] per-slot eta in synapses per unit (the per-slot learning rate of
                                 * flyzero/school.py:289-304, `_step_scaled`), or NULL = `eta` for
                                 * every slot */
} fz_reward_args;

int fz_reward_step(const fz_reward_args* a, void* stream);

/* rates[b][g] += alpha * (mean count of group g * hz_scale - rates[b][g]) for the
 * descending-neuron groups (flyzero/batch.py:107-111; groups flyzero/instruct.py:43-48).
 * Group g's neurons are gidx[gptr[g] .. gptr[g+1]). */
int fz_dn_rates(int B, int n, const int* counts, int n_groups, const int* gptr, const int* gidx, float hz_scale,
                float alpha, float* rates, void* stream);

/* ---- instructed learning (replaces the rule of flyzero/instruct.py:124-148 as batched by
 * BatchInstruct / BatchInstructDeep / BatchInstructDeep2, flyzero/batch.py:226-451, and the
 * deep_update kernel of flyzero/batch.py:27-52).  Float32 as written, fixed summation orders;
 * slot -> fly grouping as a CSR (fly_ptr, fly_slots ascending).  All pointers are device
 * pointers. ---- */

/* trace[b][k] += a_pre * (counts[b][upre[k]] * hz * pre_gain - trace[b][k]) and
 * rate[b][u] += a_rate * (counts[b][upost[u]] * hz - rate[b][u])
 * (flyzero/batch.py:289-293; pre_gain = tau_pre / 1000). */
int fz_instruct_traces(int B, int n, const int* counts, int n_upre, const int* upre, float* trace, float a_pre,
                       float pre_gain, int n_upost, const int* upost, float* rate, float a_rate, float hz,
                       void* stream);

/* Retrograde error of the L1 neurons (flyzero/batch.py:395-399): delta[b][l] = sum over
 * q in [lptr[l], lptr[l+1]) of (pw[b][lsyn[q]] / w_syn) * (tgt[b][lgroup[q]] - rate[b][lpostk[q]]),
 * in ascending q.  Read pw before the DN synapses are updated. */
int fz_instruct_delta(int B, int n_l1, const int* lptr, const int* lsyn, const int* lgroup, const int* lpostk,
                      const float* tgt, int n_groups, const float* rate, int n_upost, const float* pw, int n_pl,
                      float w_syn, float* delta, void* stream);

typedef struct fz_instruct_update_args {
    int mode;                   /* 0: delta rule onto the DNs (flyzero/batch.py:158-172,320-322,392-393);
                                 * 1: a deeper layer (the deep_update kernel, flyzero/batch.py:27-52) */
    int n_flies, n_sel, n_pl, n_upre, n_upost, n_groups, n_l;
    const int* sel;             /* [n_sel] plastic index of each listed synapse, or NULL = identity */
    const int* pre_k;           /* [n_sel] trace column of its pre neuron */
    const int* group;           /* mode 0: [n_sel] instructed group of its post neuron */
    const int* post_k;          /* mode 0: [n_sel] rate column of its post neuron */
    const unsigned char* mask;  /* mode 0: [n_sel] 1 = learns, or NULL = all */
    const int* post_l;          /* mode 1: [n_sel] delta column of its post neuron */
    const int* fly_ptr;         /* [n_flies + 1] */
    const int* fly_slots;       /* [B] */
    const unsigned char* learn; /* mode 0: [B] */
    const float* coef;          /* [B] mode 0: eta per slot; mode 1: eta_deep / learning slots of
                                 * the fly, 0 for slots that do not learn */
    const float* trace;         /* [B][n_upre] */
    const float* rate;          /* mode 0: [B][n_upost] */
    const float* tgt;           /* mode 0: [B][n_groups] target rates */
    const float* delta;         /* mode 1: [B][n_l] */
    const float* sign;          /* [n_pl] sign of the initial weight */
    const float* wmax;          /* [n_pl] cap */
    float* pw;                  /* [B][n_pl] plastic weights, mV, updated in place */
} fz_instruct_update_args;

/* w = sign * clip(sign * (w + d), 0, wmax) for every slot of each fly, d the fly's mean change
 * (mode 0: sum over learning slots of (1 / n) * eta * trace * err * mask in ascending slot
 * order, applied to every slot; mode 1: sum over slots of coef * trace * delta, applied where
 * nonzero). */
int fz_instruct_update(const fz_instruct_update_args* a, void* stream);

/* Intrinsic plasticity of the instructed DNs (flyzero/batch.py:302-310): per (fly, DN) the
 * error averaged over the fly's learning slots, then for every slot ib' = clip(ib + eta_bias *
 * mean * ratio[b], -limit, limit) and bias_ext[b][dn_slot] += ib' - ib. */
int fz_instruct_bias(int n_flies, int n_dn, const int* fly_ptr, const int* fly_slots, const unsigned char* learn,
                     const int* dn_group, const int* dn_rate_k, const float* tgt, int n_groups, const float* rate,
                     int n_upost, float eta_bias, const float* ratio, float limit, float* ib, float* bias_ext,
                     int n_ext, const int* dn_slot, void* stream);

/* ---- batched vision front end (replaces MotionEye.see, flyzero/vision.py:137-191, run per emulator
 * worker on the CPU, flyzero/pool.py:41-44): the T4/T5 motion channels, the medulla transient and
 * sustained-contrast channels and the R7/R8 colour channels ---- */
#define FZ_EYE_ON_TRANSIENT 8   /* tr_gain * max(hp, 0) / gain      (Mi1, Tm3) */
#define FZ_EYE_OFF_TRANSIENT 9  /* tr_gain * max(-hp, 0) / gain     (Tm1, Tm2, Tm4) */
#define FZ_EYE_ON_SUSTAINED 10  /* su_gain * max(con, 0) / gain     (Mi4) */
#define FZ_EYE_OFF_SUSTAINED 11 /* su_gain * max(-con, 0) / gain    (Mi9, Tm9) */
#define FZ_EYE_BLUE 12          /* ch_gain * blue cell mean / gain  (R7) */
#define FZ_EYE_GREEN 13         /* ch_gain * green cell mean / gain (R8) */
typedef struct fz_eye_args {
    int B;                       /* flies (one frame each) */
    int h, w;                    /* frame size, RGB uint8 [B][h][w][3] */
    int cell, gh, gw;            /* pixels per grid point; gh = h / cell, gw = w / cell */
    int shift;                   /* grid points between correlator arms */
    int n_cells;                 /* T4/T5 neurons */
    const unsigned char* first;  /* [B] 1: the fly's filters start from this frame (lp = lum,
                                  * delayed = 0), or NULL */
    const unsigned char* frames; /* [B][h][w][3] */
    float* lp;                   /* [B][gh][gw] low-pass of the luminance (state) */
    float* d_on;                 /* [B][gh][gw] delayed ON signal (state) */
    float* d_off;                /* [B][gh][gw] delayed OFF signal (state) */
    const int* key;              /* [n_cells] 0-7: polarity * 4 + screen direction (0 right, 1 left, 2 down,
                                  * 3 up); 8-13: the FZ_EYE_* feature channels */
    const int* gy;               /* [n_cells] grid row */
    const int* gx;               /* [n_cells] grid column */
    const int* py;               /* [n_cells] screen row, pixels */
    float y_cut;                 /* rows above y_min * h give 0 */
    float a_hp, a_d;             /* 1 / tau_hp, 1 / tau_delay (frames) */
    float gain, max_rate;        /* rate = max_rate * clip(gain * motion, 0, 1) */
    float* rates;                /* [B][n_cells] Hz */
    float* p_ext;                /* optional [B][n_ext]: p_ext[b][slot[n]] = rate * p_scale */
    int n_ext;
    const int* slot;             /* [n_cells] ext slot of each neuron */
    float p_scale;               /* dt / 1000 */
    float tr_gain, su_gain, ch_gain; /* transient, sustained and chromatic gains (0: channel off, its
                                  * cells give 0) */
    int surround;                /* grid points, radius of the contrast surround: con = lum - box mean of
                                  * lum over the window clipped to the grid */
} fz_eye_args;

/* One frame per fly: grid luminance, ON / OFF split, Hassenstein-Reichardt correlators, the
 * feature channels and every cell's rate (and, with p_ext, the brain's per-step input
 * probabilities, flyzero/fleet.py:76-80). */
int fz_eye_step(const fz_eye_args* a, void* stream);

/* ---- learned rate surrogate: BF16 MFMA GEMM and training / inference element-wise kernels ---- */

#define FZ_SG_STORE32 0 /* C32 = A B (+ bias) */
#define FZ_SG_ACC32 1   /* C32 += A B (+ bias) */
#define FZ_SG_STORE16 2 /* C16 = bf16(A B (+ bias)) */
#define FZ_SG_GELU 3    /* C32 = A B + bias (pre-activation), C16 = bf16(gelu(C32)) */
#define FZ_SG_GELU16 4  /* C16 = bf16(gelu(A B + bias)) (inference: no pre-activation kept) */

/* C[M][N] = sum_k A[m][k] B[k][n]; A, B bf16 (raw uint16 bit patterns), FP32 accumulation on
 * MFMA.  trans_a: A is stored [K][M] (lda = row stride); trans_b: B is stored [N][K].
 * K must be a multiple of 32; M and N are free.  No counterpart in the reference (its
 * nearest ancestor is the PyTorch MLP of flyzero/implant.py:136-154). */
int fz_sg_gemm(int M, int N, int K, const void* A, int lda, int trans_a, const void* B, int ldb, int trans_b,
               int mode, float* C32, void* C16, int ldc, const float* bias, void* stream);
/* `count` products of one shape in one launch (the ensemble's members): product z reads A + z
 * stride_a, B + z stride_b, bias + z stride_bias and writes C + z stride_c (element strides; 0
 * shares an operand, e.g. the input every member reads). */
int fz_sg_gemm_batched(int count, int M, int N, int K, const void* A, int lda, long long stride_a, int trans_a,
                       const void* B, int ldb, long long stride_b, int trans_b, int mode, float* C32, void* C16,
                       int ldc, long long stride_c, const float* bias, long long stride_bias, void* stream);
/* Output layer of `members` ensemble members and the ensemble's statistics for B frames:
 * y_k = ([h2_k, s_k] Wo_k^T + bo_k) * rate_scale (h2 bf16 [members][B][hid], s bf16
 * [members][B][H]; member k's bf16 weights at w16 + k wstride, Wo at offset wo_off as
 * [n_pad][hid + H], its FP32 bias at w32 + k wstride + bo_off), mean[B][n_out] over members, and
 * spread[B] = ||population std|| / (||mean|| + 1 Hz), the disagreement that sends a frame to the
 * full simulation.  members <= 8, n_out <= 8. */
int fz_sg_ens_head(int B, int members, int hid, int H, int n_out, const void* h2, const void* s, const void* w16,
                   const float* w32, long long wstride, long long wo_off, long long bo_off, float rate_scale,
                   float* mean, float* spread, void* stream);
/* dpre = dact * gelu'(pre) (tanh-approximated GELU), as FP32 and/or BF16. */
int fz_sg_gelu_bwd(long long n, const float* dact, const float* pre, float* dpre32, void* dpre16, void* stream);
/* db[n] (+)= sum_m d[m][n] over a row-major [M][N] array. */
int fz_sg_colsum(int M, int N, const float* d, float* db, int accumulate, void* stream);
/* One GRU step (gate order r, z, n): gx = x W_ih^T + b_ih, gh = h W_hh^T + b_hh, [B][3H];
 * hn = new state (FP32, optionally BF16); gates [B][4H] = (r, z, n, gh_n) for the backward. */
int fz_sg_gru_fwd(int B, int H, const float* gx, const float* gh, const float* h, float* hn, void* hn16,
                  float* gates, void* stream);
/* Backward of one GRU step: dgx, dgh [B][3H] (FP32, optionally BF16) and dh += dhn * z (the
 * direct term is added into dh; the caller adds dgh W_hh). */
int fz_sg_gru_bwd(int B, int H, const float* dhn, const float* h, const float* gates, float* dgx, float* dgh,
                  void* dgx16, void* dgh16, float* dh, void* stream);
/* loss += scale * sum (y - t)^2 (device float); dy = 2 scale (y - t). */
int fz_sg_mse(long long n, const float* y, const float* t, float* dy, float scale, float* loss, void* stream);
/* Adam with bias correction factors c1 = 1 - b1^t, c2 = 1 - b2^t and decoupled-from-nothing
 * L2 weight decay wd; w16 receives the BF16 copy. */
int fz_sg_adam(long long n, float* w, const float* grad, float* m, float* v, void* w16, float lr, float b1, float b2,
               float eps, float c1, float c2, float wd, void* stream);
/* y = bf16(x). */
int fz_sg_cast(long long n, const float* x, void* y, void* stream);

/* Reads (and clears) the device-side error flag; waits for `stream`. */
int fz_check(fz_engine* e, void* stream);

#ifdef __cplusplus
}
#endif
#endif
