// Reward-modulated plasticity and descending-neuron rate readout on the plastic CSR slice.
//
// The rule is the exploratory-Hebbian one of flyzero/learning.py:116-131 in its batched form
// (flyzero/batch.py:482-495 with the per-fly averaging of batch.py:158-172):
//
//   dev_k   = c_post - <post>                     (running mean per post neuron)
//   e_k     = ae * e_k + c_pre * dev_k            (eligibility, stored float32)
//   <post> += ap * (c_post - <post>)
//   adv     = R - <R>,  <R> += ar * adv
//   w_k    += mean over the fly's learning slots of eta_b * adv * e_k   (eta_b: one eta for every
//             slot, flyzero/batch.py:494, or a per-slot rate, flyzero/school.py:289-304)
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
            const double eta = a.eta_slot ? a.eta_slot[b] : a.eta;  // school.py:300-302 scales per slot
            acc += eta * adv * (double)ef;
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


// ---- instructed learning (flyzero/instruct.py:124-148, batched as flyzero/batch.py:226-451) ----
//
//   trace_k <- trace_k + a_pre  * (c_pre  * hz * tau_pre_s - trace_k)   (per slot, unique pre)
//   rate_u  <- rate_u  + a_rate * (c_post * hz - rate_u)                 (per slot, unique post)
//   DN synapses:   dw_b = eta_b * trace * (target - rate) * mask,  w += mean over the fly's
//                  learning slots (each term scaled by 1 / n, summed in ascending slot order)
//   retrograde:    delta_b[l] = sum over the L1 neuron's DN synapses of (pw / w_syn) * err
//   deep synapses: acc = sum over the fly's slots of coef_b * trace * delta_b (coef_b = 0 when
//                  not learning), applied only where acc != 0 (flyzero/batch.py:27-52)
//   every update:  w = sign * clip(sign * (w + d), 0, wmax)                (Dale's law and cap)
//
// Float32 evaluated as written (-ffp-contract=off) in fixed orders, so the CPU restatement in
// the same order gives the same bits.

__global__ void instruct_traces(int B, int n, const int* counts, int n_upre, const int* upre, float* trace,
                                float a_pre, float pre_gain, int n_upost, const int* upost, float* rate,
                                float a_rate, float hz) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)n_upre + n_upost;
    if (t >= (long long)B * per) return;
    const int b = (int)(t / per);
    const int k = (int)(t - (long long)b * per);
    const size_t bn = (size_t)b * n;
    if (k < n_upre) {
        const float c = (float)counts[bn + upre[k]];
        float* x = trace + (size_t)b * n_upre + k;
        *x = *x + a_pre * (c * hz * pre_gain - *x);
    } else {
        const int u = k - n_upre;
        const float c = (float)counts[bn + upost[u]];
        float* x = rate + (size_t)b * n_upost + u;
        *x = *x + a_rate * (c * hz - *x);
    }
}

// delta[b][l] for the L1 neurons: entries q in [lptr[l], lptr[l+1]) name a DN synapse by its
// plastic index lsyn[q], its group lgroup[q] and its post neuron's rate column lpostk[q].
__global__ void instruct_delta(int B, int n_l1, const int* lptr, const int* lsyn, const int* lgroup,
                               const int* lpostk, const float* tgt, int n_groups, const float* rate, int n_upost,
                               const float* pw, int n_pl, float w_syn, float* delta) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)B * n_l1) return;
    const int b = (int)(t / n_l1);
    const int l = (int)(t - (long long)b * n_l1);
    float acc = 0.0f;
    for (int q = lptr[l]; q < lptr[l + 1]; q++) {
        const float err = tgt[(size_t)b * n_groups + lgroup[q]] - rate[(size_t)b * n_upost + lpostk[q]];
        acc += (pw[(size_t)b * n_pl + lsyn[q]] / w_syn) * err;
    }
    delta[t] = acc;
}

// One thread per (fly, listed synapse s -> plastic index sel[s]).  mode 0 (delta rule onto the
// DNs): term_b = inv_n * (eta_b * trace * err * mask), the fly's learning slots only, then every
// slot of the fly is updated.  mode 1 (deeper layers): term_b = coef_b * trace * delta_b over
// slots with coef_b != 0, and slots are updated only where the sum is nonzero.
__global__ void instruct_update(fz_instruct_update_args a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.n_flies * a.n_sel) return;
    const int f = (int)(t / a.n_sel);
    const int s = (int)(t - (long long)f * a.n_sel);
    const int j = a.sel ? a.sel[s] : s;
    const int pk = a.pre_k[s];
    const int q0 = a.fly_ptr[f], q1 = a.fly_ptr[f + 1];
    float acc = 0.0f;
    if (a.mode == 0) {
        const int g = a.group[s], uk = a.post_k[s];
        const float m = a.mask ? (float)a.mask[s] : 1.0f;
        int nl = 0;
        for (int q = q0; q < q1; q++) nl += a.learn[a.fly_slots[q]] ? 1 : 0;
        const float inv_n = 1.0f / (float)(nl > 1 ? nl : 1);
        for (int q = q0; q < q1; q++) {
            const int b = a.fly_slots[q];
            if (!a.learn[b]) continue;
            const float err = a.tgt[(size_t)b * a.n_groups + g] - a.rate[(size_t)b * a.n_upost + uk];
            const float dw = a.coef[b] * a.trace[(size_t)b * a.n_upre + pk] * err * m;
            acc += inv_n * dw;
        }
    } else {
        const int l = a.post_l[s];
        for (int q = q0; q < q1; q++) {
            const int b = a.fly_slots[q];
            const float c = a.coef[b];
            if (c != 0.0f) acc += c * a.trace[(size_t)b * a.n_upre + pk] * a.delta[(size_t)b * a.n_l + l];
        }
        if (acc == 0.0f) return;
    }
    const float sg = a.sign[j], mx = a.wmax[j];
    for (int q = q0; q < q1; q++) {
        const size_t bj = (size_t)a.fly_slots[q] * a.n_pl + j;
        const float w = sg * (a.pw[bj] + acc);
        a.pw[bj] = sg * fminf(fmaxf(w, 0.0f), mx);
    }
}

// Intrinsic plasticity of the instructed DNs (flyzero/batch.py:302-310): one thread per (fly,
// DN); err averaged over the fly's learning slots as in mode 0, then for every slot of the fly
// ib' = clip(ib + eta_bias * mean * ratio_b, -limit, limit), bias_ext += ib' - ib.
__global__ void instruct_bias(int n_flies, int n_dn, const int* fly_ptr, const int* fly_slots,
                              const unsigned char* learn, const int* dn_group, const int* dn_rate_k,
                              const float* tgt, int n_groups, const float* rate, int n_upost, float eta_bias,
                              const float* ratio, float limit, float* ib, float* bias_ext, int n_ext,
                              const int* dn_slot) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n_flies * n_dn) return;
    const int f = (int)(t / n_dn);
    const int d = (int)(t - (long long)f * n_dn);
    const int q0 = fly_ptr[f], q1 = fly_ptr[f + 1];
    int nl = 0;
    for (int q = q0; q < q1; q++) nl += learn[fly_slots[q]] ? 1 : 0;
    const float inv_n = 1.0f / (float)(nl > 1 ? nl : 1);
    float mean = 0.0f;
    for (int q = q0; q < q1; q++) {
        const int b = fly_slots[q];
        if (learn[b]) mean += inv_n * (tgt[(size_t)b * n_groups + dn_group[d]] - rate[(size_t)b * n_upost + dn_rate_k[d]]);
    }
    for (int q = q0; q < q1; q++) {
        const int b = fly_slots[q];
        const size_t bd = (size_t)b * n_dn + d;
        const float step = eta_bias * mean * ratio[b];
        const float nw = fminf(fmaxf(ib[bd] + step, -limit), limit);
        bias_ext[(size_t)b * n_ext + dn_slot[d]] += nw - ib[bd];
        ib[bd] = nw;
    }
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

int fz_instruct_traces(int B, int n, const int* counts, int n_upre, const int* upre, float* trace, float a_pre,
                       float pre_gain, int n_upost, const int* upost, float* rate, float a_rate, float hz,
                       void* stream) {
    if (B < 1 || n < 1 || n_upre < 0 || n_upost < 0)
        return fzi_fail(FZ_E_ARG, "bad sizes (B %d, n %d, pre %d, post %d)", B, n, n_upre, n_upost);
    if (!counts || (n_upre && (!upre || !trace)) || (n_upost && (!upost || !rate)))
        return fzi_fail(FZ_E_ARG, "null counts, neuron lists, trace or rate");
    const long long t = (long long)B * ((long long)n_upre + n_upost);
    if (t == 0) return FZ_OK;
    hipLaunchKernelGGL(instruct_traces, dim3((unsigned)((t + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), B, n, counts, n_upre, upre, trace, a_pre, pre_gain,
                       n_upost, upost, rate, a_rate, hz);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "instruct_traces launch failed");
    return FZ_OK;
}

int fz_instruct_delta(int B, int n_l1, const int* lptr, const int* lsyn, const int* lgroup, const int* lpostk,
                      const float* tgt, int n_groups, const float* rate, int n_upost, const float* pw, int n_pl,
                      float w_syn, float* delta, void* stream) {
    if (B < 1 || n_l1 < 0 || n_groups < 1 || n_upost < 0 || n_pl < 0)
        return fzi_fail(FZ_E_ARG, "bad sizes (B %d, L1 %d, groups %d)", B, n_l1, n_groups);
    if (n_l1 == 0) return FZ_OK;
    if (!lptr || !lsyn || !lgroup || !lpostk || !tgt || !rate || !pw || !delta)
        return fzi_fail(FZ_E_ARG, "null delta argument");
    const long long t = (long long)B * n_l1;
    hipLaunchKernelGGL(instruct_delta, dim3((unsigned)((t + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), B, n_l1, lptr, lsyn, lgroup, lpostk, tgt, n_groups, rate,
                       n_upost, pw, n_pl, w_syn, delta);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "instruct_delta launch failed");
    return FZ_OK;
}

int fz_instruct_update(const fz_instruct_update_args* a, void* stream) {
    if (!a) return fzi_fail(FZ_E_ARG, "null update arguments");
    if (a->mode != 0 && a->mode != 1) return fzi_fail(FZ_E_ARG, "mode %d: must be 0 or 1", a->mode);
    if (a->n_flies < 1 || a->n_sel < 0 || a->n_pl < 0 || a->n_upre < 0)
        return fzi_fail(FZ_E_ARG, "bad sizes (flies %d, synapses %d)", a->n_flies, a->n_sel);
    if (a->n_sel == 0) return FZ_OK;
    if (!a->pre_k || !a->fly_ptr || !a->fly_slots || !a->coef || !a->trace || !a->sign || !a->wmax || !a->pw)
        return fzi_fail(FZ_E_ARG, "null update array");
    if (a->mode == 0 && (!a->group || !a->post_k || !a->tgt || !a->rate || !a->learn))
        return fzi_fail(FZ_E_ARG, "delta rule needs group, post_k, tgt, rate and learn");
    if (a->mode == 1 && (!a->post_l || !a->delta)) return fzi_fail(FZ_E_ARG, "deep rule needs post_l and delta");
    const long long t = (long long)a->n_flies * a->n_sel;
    hipLaunchKernelGGL(instruct_update, dim3((unsigned)((t + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), *a);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "instruct_update launch failed");
    return FZ_OK;
}

int fz_instruct_bias(int n_flies, int n_dn, const int* fly_ptr, const int* fly_slots, const unsigned char* learn,
                     const int* dn_group, const int* dn_rate_k, const float* tgt, int n_groups, const float* rate,
                     int n_upost, float eta_bias, const float* ratio, float limit, float* ib, float* bias_ext,
                     int n_ext, const int* dn_slot, void* stream) {
    if (n_flies < 1 || n_dn < 0 || n_groups < 1 || n_ext < 0)
        return fzi_fail(FZ_E_ARG, "bad sizes (flies %d, DNs %d, groups %d)", n_flies, n_dn, n_groups);
    if (n_dn == 0) return FZ_OK;
    if (!fly_ptr || !fly_slots || !learn || !dn_group || !dn_rate_k || !tgt || !rate || !ratio || !ib ||
        !bias_ext || !dn_slot)
        return fzi_fail(FZ_E_ARG, "null intrinsic-bias argument");
    const long long t = (long long)n_flies * n_dn;
    hipLaunchKernelGGL(instruct_bias, dim3((unsigned)((t + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), n_flies, n_dn, fly_ptr, fly_slots, learn, dn_group,
                       dn_rate_k, tgt, n_groups, rate, n_upost, eta_bias, ratio, limit, ib, bias_ext, n_ext,
                       dn_slot);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "instruct_bias launch failed");
    return FZ_OK;
}

}  // extern "C"
