/*
 * The reference's own arithmetic for the LIF step -- TEST INFRASTRUCTURE ONLY (used by tests/
 * to measure how far this build's arithmetic takes the spike trains from the reference's; the
 * shipped library never calls it).
 *
 * Unlike lif_oracle.c (this build's numeric contract: integer synaptic sums, u = v - v_rest),
 * this restates the reference as it computes:
 *   - v in mV, float32 (flyzero/brain.py:105);
 *   - synaptic weights data = f32(f32(count) * f32(w_syn)) mV per connection (brain.py:84-89),
 *     added into a float32 delay ring one spike at a time, spikes in ascending neuron order,
 *     each row in CSR order (brain.py:49-56);
 *   - g = g * decay_syn + buf, v += ((g + bias) - (v - v_rest)) * decay_m (brain.py:40-45),
 *     optionally with fused multiply-adds where numba's fastmath (brain.py:62) may contract;
 *   - kicks from an input list that may repeat a neuron (brain.py:46-48).
 * numpy = 1 selects the numpy backend's semantics (brain.py:147-170): the refractory test
 * `active = ref <= 0` is taken before the decrement and gates the spike test (brain.py:154,
 * 163); a neuron listed twice in the input is kicked once per step (fancy-index add,
 * brain.py:161); a step's arrivals are summed per target in float64 in the flattened row order
 * and rounded to float32 once (np.bincount, brain.py:130-131).
 * Kicks come from precomputed hits (bit j of word j / 32 of a step = input entry j), the
 * reference's own way to reproduce a trial without its RNG.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n, n_delay, n_ref;
    float decay_syn, decay_m, v_rest, v_th, v_reset, kick;
    const long long* indptr;      /* CSR of all synapses, n + 1 */
    const int* indices;           /* post neuron */
    const float* data;            /* mV per connection */
    int n_in;                     /* input list length */
    const long long* input_idx;   /* neuron of each input entry (repeats allowed) */
    int numpy;                    /* 1: numpy backend semantics */
    int fused;                    /* 1: fmaf for g * decay_syn + buf and v + t * decay_m (numba path only) */
} orc_ref_net;

typedef struct {
    float* v;                     /* [n] mV */
    float* g;                     /* [n] */
    short* ref;                   /* [n] int16 like brain.py:107 */
    float* pending;               /* [n_delay][n] */
    const float* bias;            /* [n] mV */
    long long t;
} orc_ref_state;

long long orc_ref_run(const orc_ref_net* net, orc_ref_state* st, int steps, const unsigned* hits, int* counts,
                      unsigned* raster) {
    if (!net || !st || !counts || steps < 0 || net->n < 1 || net->n_delay < 1) return -1;
    const int n = net->n, hw = (net->n_in + 31) / 32, rw = (n + 31) / 32;
    int* spk = (int*)malloc(sizeof(int) * (size_t)n);
    unsigned char* active = (unsigned char*)malloc((size_t)n);
    double* acc64 = net->numpy ? (double*)calloc((size_t)n, sizeof(double)) : NULL;
    unsigned char* kicked = net->numpy ? (unsigned char*)calloc((size_t)n, 1) : NULL;
    long long total = 0;
    for (int s = 0; s < steps; s++) {
        const int slot = (int)(st->t % net->n_delay);
        float* buf = st->pending + (size_t)slot * n;
        for (int i = 0; i < n; i++) {
            float gi;
            if (net->fused && !net->numpy) gi = fmaf(st->g[i], net->decay_syn, buf[i]);
            else gi = st->g[i] * net->decay_syn + buf[i];
            st->g[i] = gi;
            buf[i] = 0.0f;
            active[i] = st->ref[i] <= 0;
            if (st->ref[i] > 0) {
                st->ref[i] -= 1;
            } else {
                const float t1 = st->v[i] - net->v_rest, t2 = gi + st->bias[i], t3 = t2 - t1;
                if (net->fused && !net->numpy) st->v[i] = fmaf(t3, net->decay_m, st->v[i]);
                else st->v[i] = st->v[i] + t3 * net->decay_m;
            }
        }
        if (hits) {
            const unsigned* hs = hits + (size_t)s * hw;
            for (int j = 0; j < net->n_in; j++) {
                if (!((hs[j >> 5] >> (j & 31)) & 1u)) continue;
                const long long i = net->input_idx[j];
                if (net->numpy) {  /* v[idx[hit]] += kick: once per listed neuron */
                    if (kicked[i]) continue;
                    kicked[i] = 1;
                }
                st->v[i] += net->kick;
            }
            if (net->numpy)
                for (int j = 0; j < net->n_in; j++) kicked[net->input_idx[j]] = 0;
        }
        int ns = 0;
        for (int i = 0; i < n; i++) {
            const int may = net->numpy ? active[i] : (st->ref[i] <= 0);
            if (st->v[i] >= net->v_th && may) {
                st->v[i] = net->v_reset;
                st->ref[i] = (short)net->n_ref;
                counts[i] += 1;
                if (raster) raster[(size_t)s * rw + (i >> 5)] |= 1u << (i & 31);
                spk[ns++] = i;
            }
        }
        if (net->numpy) {  /* float64 sums in flattened order, one float32 rounding per target */
            for (int q = 0; q < ns; q++)
                for (long long k = net->indptr[spk[q]]; k < net->indptr[spk[q] + 1]; k++)
                    acc64[net->indices[k]] += (double)net->data[k];
            /* the slot was cleared this step, so buf[j] += f32(sum) stores the rounded sum; a target
             * reached twice adds 0 the second time */
            for (int q = 0; q < ns; q++)
                for (long long k = net->indptr[spk[q]]; k < net->indptr[spk[q] + 1]; k++) {
                    const int j = net->indices[k];
                    buf[j] += (float)acc64[j];
                    acc64[j] = 0.0;
                }
        } else {
            for (int q = 0; q < ns; q++)
                for (long long k = net->indptr[spk[q]]; k < net->indptr[spk[q] + 1]; k++)
                    buf[net->indices[k]] += net->data[k];
        }
        total += ns;
        st->t += 1;
    }
    free(spk);
    free(active);
    free(acc64);
    free(kicked);
    return total;
}
