/*
 * CPU oracle for the whole-brain LIF step -- TEST INFRASTRUCTURE ONLY
 * (the checker for tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline; the shipped library never calls it).  Semantics and the
 * deliberate differences from the reference are listed in lif_oracle.h.
 *
 * Step order restated from flyzero/brain.py:36-57 (numba `_kernel`):
 *   slot = t % n_delay                                     brain.py:37
 *   g = g*decay_syn + buf[i]; buf[i] = 0                   brain.py:40-41
 *   if ref > 0: ref -= 1  else: v += (g+bias-(v-v_rest))*decay_m   brain.py:42-45
 *     (here with u = v - v_rest:  u += (g + bias - u) * decay_m)
 *   Poisson kick v += kick, also while refractory          brain.py:46-48
 *   spike if v >= v_th and ref <= 0 (after the decrement)  brain.py:50
 *     -> v = v_reset, ref = n_ref, count++, row pushed into the same slot
 *        (read again n_delay steps later)                  brain.py:51-56
 * Batched extras from flyzero/gpu.py: ext-slot bias/probability per fly
 * (gpu.py:55-57,63-69), the counter hash RNG (gpu.py:29-31,66-67), plastic
 * rows with per-fly weights (gpu.py:109-115).
 */
#include "lif_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 32-bit integer finaliser ("lowbias32", C. Wellons' hash-prospector), the
 * same mixing function as flyzero/gpu.py:29-31. */
unsigned orc_mix(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

/* gpu.py:66: h = mix(tid ^ mix(step * 0x9e3779b9 ^ mix(seed + b))) */
unsigned orc_hash(unsigned seed, unsigned b, unsigned step, unsigned tid) {
    return orc_mix(tid ^ orc_mix(step * 0x9e3779b9U ^ orc_mix(seed + b)));
}

typedef struct {
    int* idx;
    int n;
} spk_list;

/* Shadow statistics of one thread (merged after each sweep). */
typedef struct {
    double dv, dg, dv_sub, vabs;  /* FP64 trajectory vs FP32 */
    long long flips;
    double sdv, sdg;              /* one FP64 step from the FP32 state vs the FP32 step */
    long long sflips;
} shadow_acc;

/* One fly, one step, neurons [i0, i1): phase 1 of brain.py:39-50. */
static void sweep(const orc_net* net, orc_state* st, const orc_input* in, int b, int s, int slot,
                  int i0, int i1, int* counts, unsigned* raster, spk_list* out, int one_step, shadow_acc* sa) {
    const int n = net->n;
    const size_t bn = (size_t)b * n;
    int* ring = st->ring + ((size_t)b * net->n_delay + slot) * n;
    float* v = st->u + bn;
    float* g = st->g + bn;
    signed char* ref = st->ref + bn;
    double* v64 = st->u64 ? st->u64 + bn : NULL;
    double* g64 = st->g64 ? st->g64 + bn : NULL;
    const int hw = (net->n_ext + 31) / 32;
    const int rw = (n + 31) / 32;
    const unsigned step = st->step_no + (unsigned)s;
    for (int i = i0; i < i1; i++) {
        int acc = ring[i];
        ring[i] = 0;
        const float g_prev = g[i], u_prev = v[i];
        float inj = (float)acc * net->in_scale;
        float gi = g[i] * net->decay_syn + inj;
        g[i] = gi;
        int j = net->ext_idx ? net->ext_idx[i] : -1;
        float bias = (j >= 0 && in->bias_ext) ? in->bias_ext[(size_t)b * net->n_ext + j] : 0.0f;
        int r = ref[i];
        float vi = v[i];
        if (r > 0) {
            r -= 1;
        } else {
            vi += (gi + bias - vi) * net->decay_m;
        }
        int hit = 0;
        if (j >= 0) {
            if (in->hits) {
                hit = (in->hits[((size_t)s * st->B + b) * hw + (j >> 5)] >> (j & 31)) & 1u;
            } else if (in->p_ext) {
                float p = in->p_ext[(size_t)b * net->n_ext + j];
                if (p > 0.0f) {
                    unsigned tid = (unsigned)((long long)b * n + i);
                    unsigned h = orc_hash(in->seed, (unsigned)b, step, tid);
                    hit = ((float)(h >> 8) * (1.0f / 16777216.0f)) < p;
                }
            }
        }
        if (hit) vi += net->kick;
        int spike = (vi >= net->u_th) && (r <= 0);
        if (one_step) {
            /* The same step in FP64 from this step's FP32 state (g, u before it; same integer
             * arrivals, kick and refractory state; the model's float32 constants), against the
             * FP32 result before the spike test. */
            double gd = (double)g_prev * (double)net->decay_syn + (double)acc * (double)net->in_scale;
            double vd = (double)u_prev;
            if (ref[i] <= 0) vd += (gd + (double)bias - vd) * (double)net->decay_m;
            if (hit) vd += (double)net->kick;
            const double dv = fabs(vd - (double)vi), dg = fabs(gd - (double)gi);
            if (dv > sa->sdv) sa->sdv = dv;
            if (dg > sa->sdg) sa->sdg = dg;
            if (((vd >= (double)net->u_th) && (r <= 0)) != spike) sa->sflips++;
        }
        if (v64) {
            /* FP64 shadow: same integer arrivals, same kicks, spike decisions
             * forced to the FP32 ones so the two trajectories stay comparable. */
            double gd = g64[i] * (double)net->decay_syn + (double)acc * (double)net->in_scale;
            g64[i] = gd;
            double vd = v64[i];
            if (ref[i] <= 0) vd += (gd + (double)bias - vd) * (double)net->decay_m;
            if (hit) vd += (double)net->kick;
            int spike64 = (vd >= (double)net->u_th) && (r <= 0);
            if (spike64 != spike) sa->flips++;
            double dv = fabs(vd - (double)vi), dg = fabs(gd - (double)gi);
            if (dv > sa->dv) sa->dv = dv;
            if (vd < (double)net->u_th && vd >= -12.0 && dv > sa->dv_sub) sa->dv_sub = dv;
            if (fabs(vd) > sa->vabs) sa->vabs = fabs(vd);
            if (dg > sa->dg) sa->dg = dg;
            v64[i] = spike ? (double)net->u_reset : vd;
        }
        if (spike) {
            vi = net->u_reset;
            r = net->n_ref;
            counts[bn + i] += 1;
            if (raster) raster[((size_t)s * st->B + b) * rw + (i >> 5)] |= 1u << (i & 31);
            out->idx[out->n++] = i;
        }
        v[i] = vi;
        ref[i] = (signed char)r;
    }
}

/* Phase 2 (brain.py:55-56, gpu.py:98-116): rows of this step's spikes land in
 * the slot just consumed, n_delay steps ahead.  Integer adds: order-free. */
static void push(const orc_net* net, orc_state* st, const orc_input* in, int b, int slot, const spk_list* sp) {
    const int n = net->n;
    int* ring = st->ring + ((size_t)b * net->n_delay + slot) * n;
    for (int q = 0; q < sp->n; q++) {
        int i = sp->idx[q];
        for (long long k = net->indptr[i]; k < net->indptr[i + 1]; k++) ring[net->col[k]] += net->w[k];
        if (net->n_pl > 0) {
            const float* pw = in->pw + (size_t)b * net->n_pl;
            for (long long k = net->pl_indptr[i]; k < net->pl_indptr[i + 1]; k++)
                ring[net->pl_post[k]] += (int)rintf(pw[k] * net->pw_qscale);
        }
    }
}

long long orc_lif_run(const orc_net* net, orc_state* st, const orc_input* in, int steps, int* counts,
                      unsigned* raster, orc_shadow_report* rep) {
    if (!net || !st || !in || !counts || steps < 0 || net->n < 0 || net->n_delay < 1 || st->B < 0)
        return -1;
    if (net->n_pl > 0 && (!net->pl_indptr || !net->pl_post || !in->pw)) return -2;
    const int n = net->n;
    int nthreads = in->nthreads > 1 ? in->nthreads : 1;
    if (nthreads > 64) nthreads = 64;
#ifndef _OPENMP
    nthreads = 1;
#endif
    spk_list* lists = (spk_list*)calloc((size_t)nthreads, sizeof(spk_list));
    for (int t = 0; t < nthreads; t++) lists[t].idx = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    const int one_step = rep && rep->one_step;
    shadow_acc run = {0};
    long long total = 0;
    for (int s = 0; s < steps; s++) {
        const int slot = (int)((st->t + s) % net->n_delay);
        double dv_step = 0.0, sdv_step = 0.0;
        for (int b = 0; b < st->B; b++) {
            if (in->alive && !in->alive[b]) continue;
            /* neuron chunks are multiples of 32 so raster words are not shared */
            int chunk = ((n + nthreads - 1) / nthreads + 31) / 32 * 32;
            shadow_acc part[64];
            memset(part, 0, sizeof(part));
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) if (nthreads > 1)
#endif
            for (int t = 0; t < nthreads; t++) {
                int i0 = t * chunk, i1 = i0 + chunk;
                if (i1 > n) i1 = n;
                lists[t].n = 0;
                if (i0 < i1) sweep(net, st, in, b, s, slot, i0, i1, counts, raster, &lists[t], one_step, &part[t]);
            }
            for (int t = 0; t < nthreads; t++) {
                const shadow_acc* a = &part[t];
                if (a->dv > dv_step) dv_step = a->dv;
                if (a->sdv > sdv_step) sdv_step = a->sdv;
                if (a->dg > run.dg) run.dg = a->dg;
                if (a->dv_sub > run.dv_sub) run.dv_sub = a->dv_sub;
                if (a->vabs > run.vabs) run.vabs = a->vabs;
                if (a->sdg > run.sdg) run.sdg = a->sdg;
                run.flips += a->flips;
                run.sflips += a->sflips;
                push(net, st, in, b, slot, &lists[t]);
                total += lists[t].n;
            }
        }
        if (dv_step > run.dv) run.dv = dv_step;
        if (sdv_step > run.sdv) run.sdv = sdv_step;
        if (rep && rep->dv_step) rep->dv_step[s] = dv_step;
        if (rep && rep->step_dv_step) rep->step_dv_step[s] = sdv_step;
    }
    st->t += steps;
    st->step_no += (unsigned)steps;
    if (rep) {
        rep->max_dv = run.dv;
        rep->max_dg = run.dg;
        rep->max_dv_sub = run.dv_sub;
        rep->max_abs_v = run.vabs;
        rep->flips = run.flips;
        rep->step_max_dv = run.sdv;
        rep->step_max_dg = run.sdg;
        rep->step_flips = run.sflips;
    }
    for (int t = 0; t < nthreads; t++) free(lists[t].idx);
    free(lists);
    return total;
}
