// Reward-modulated plasticity and descending-neuron rate readout on the plastic CSR slice.
//
// The rule is the exploratory-Hebbian one of flyzero/learning.py:116-131 in its batched form
// (flyzero/batch.py:482-495 with the per-fly averaging of batch.py:158-172):
//
//   dev_k   = c_post - <post>                     (running mean per post neuron)
//   e_k     = ae * e_k + c_pre * dev_k            (eligibility, stored float32)
//   <post> += ap * (c_post - <post>)
//   adv     = R - <R>,  <R> += ar * adv
//   w_k    += mean over the fly's learning slots of eta * adv * e_k
//   w_k     = sign * clip(sign * w_k, 0, wmax_k)  (Dale's law and cap)
//
// Weights are kept in synapse units in float64 and the simulation uses the integer count
// rint(w) (pw = count * w_syn mV).  Every operation is float64 evaluated as written
// (-ffp-contract=off), with sums over slots in ascending slot order, so the integer edits are
// identical to the FP64 CPU restatement.  One thread per (fly, plastic synapse) reads the
// window's spike counts directly (no gathered copies) and walks the fly's slots.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "errors.h"
#include "flyhip.h"

namespace {

__global__ void reward_synapses(fz_reward_args a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)a.n_flies * a.n_pl;
    if (t >= total) return;
    const int f = (int)(t / a.n_pl);
    const int k = (int)(t - (long long)f * a.n_pl);
    const int pre = a.pre[k], post = a.post[k], pk = a.post_k[k];
    const int s0 = a.fly_ptr[f], s1 = a.fly_ptr[f + 1];
    double acc = 0.0;
    int nact = 0;
    for (int q = s0; q < s1; q++) {
        const int b = a.fly_slots[q];
        const size_t bn = (size_t)b * a.n;
        const double c_pre = (double)a.counts[bn + pre];
        const double c_post = (double)a.counts[bn + post];
        const double dev = c_post - a.post_mean[(size_t)b * a.n_post + pk];
        const size_t bk = (size_t)b * a.n_pl + k;
        const double e = a.ae * (double)a.elig[bk] + c_pre * dev;
        const float ef = (float)e;
        a.elig[bk] = ef;
        if (a.learn[b]) {
            const double adv = a.reward[b] - a.r_mean[b];
            acc += a.eta * adv * (double)ef;
            nact++;
        }
    }
    if (!a.any_learn || nact == 0) return;
    const double d = acc / (double)nact;
    const double sg = (double)a.sign[k], mx = a.wmax[k];
    for (int q = s0; q < s1; q++) {
        const int b = a.fly_slots[q];
        const size_t bk = (size_t)b * a.n_pl + k;
        double w = a.w[bk] + d;
        double m = sg * w;
        m = m < 0.0 ? 0.0 : (m > mx ? mx : m);
        w = sg * m;
        a.w[bk] = w;
        const double c = rint(w);
        a.count[bk] = (int)c;
        a.pw[bk] = (float)c * a.w_syn;
    }
}

__global__ void reward_means(fz_reward_args a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)a.B * a.n_post;
    if (t < total) {
        const int b = (int)(t / a.n_post);
        const int u = (int)(t - (long long)b * a.n_post);
        const size_t bu = (size_t)b * a.n_post + u;
        const double c_post = (double)a.counts[(size_t)b * a.n + a.upost[u]];
        a.post_mean[bu] += a.ap * (c_post - a.post_mean[bu]);
    }
    if (t < a.B) {
        const double adv = a.reward[t] - a.r_mean[t];
        a.adv_out[t] = adv;
        a.r_mean[t] += a.ar * adv;
    }
}

// rates[b][gidx] += a * (mean count of the group's neurons * 1000 / window - rates)
// (flyzero/batch.py:107-111 for the 8 descending-neuron groups of flyzero/instruct.py:43-48).
__global__ void dn_rates(int B, int n, const int* counts, int n_groups, const int* gptr, const int* gidx,
                         float hz_scale, float alpha, float* rates) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * n_groups) return;
    const int b = t / n_groups, gi = t - b * n_groups;
    const int m0 = gptr[gi], m1 = gptr[gi + 1];
    int s = 0;
    for (int q = m0; q < m1; q++) s += counts[(size_t)b * n + gidx[q]];
    const float inst = m1 > m0 ? ((float)s / (float)(m1 - m0)) * hz_scale : 0.0f;
    rates[t] = rates[t] + alpha * (inst - rates[t]);
}

}  // namespace

extern "C" {

int fz_reward_step(const fz_reward_args* a, void* stream) {
    if (!a) return fzi_fail(FZ_E_ARG, "null reward arguments");
    if (a->B < 1 || a->n < 1 || a->n_pl < 0 || a->n_post < 0 || a->n_flies < 1)
        return fzi_fail(FZ_E_ARG, "bad sizes (B %d, n %d, n_pl %d, n_post %d, flies %d)", a->B, a->n, a->n_pl,
                        a->n_post, a->n_flies);
    if (!a->counts || !a->reward || !a->learn || !a->r_mean || !a->adv_out)
        return fzi_fail(FZ_E_ARG, "null counts, reward, learn, r_mean or adv_out");
    if (a->n_pl > 0 && (!a->pre || !a->post || !a->post_k || !a->sign || !a->wmax || !a->elig || !a->w ||
                        !a->count || !a->pw || !a->fly_ptr || !a->fly_slots || !a->post_mean))
        return fzi_fail(FZ_E_ARG, "null plastic-synapse array");
    if (a->n_post > 0 && (!a->upost || !a->post_mean)) return fzi_fail(FZ_E_ARG, "null upost or post_mean");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long syn = (long long)a->n_flies * a->n_pl;
    if (syn > 0) {
        hipLaunchKernelGGL(reward_synapses, dim3((unsigned)((syn + 255) / 256)), dim3(256), 0, s, *a);
        if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "reward_synapses launch failed");
    }
    long long mt = (long long)a->B * a->n_post;
    if (mt < a->B) mt = a->B;
    hipLaunchKernelGGL(reward_means, dim3((unsigned)((mt + 255) / 256)), dim3(256), 0, s, *a);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "reward_means launch failed");
    return FZ_OK;
}

int fz_dn_rates(int B, int n, const int* counts, int n_groups, const int* gptr, const int* gidx, float hz_scale,
                float alpha, float* rates, void* stream) {
    if (B < 1 || n < 1 || n_groups < 1) return fzi_fail(FZ_E_ARG, "bad sizes (B %d, n %d, groups %d)", B, n, n_groups);
    if (!counts || !gptr || !gidx || !rates) return fzi_fail(FZ_E_ARG, "null counts, groups or rates");
    const int t = B * n_groups;
    hipLaunchKernelGGL(dn_rates, dim3((t + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), B, n,
                       counts, n_groups, gptr, gidx, hz_scale, alpha, rates);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "dn_rates launch failed");
    return FZ_OK;
}

}  // extern "C"
