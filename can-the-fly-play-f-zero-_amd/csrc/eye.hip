// Batched vision front end: game frames -> Poisson rates of the visual input neurons
// (flyzero/vision.py:137-191, MotionEye.see): T4/T5 motion channels, the medulla transient and
// sustained-contrast channels and the R7/R8 colour channels.
//
// One workgroup per fly.  The frame is reduced to a grid of cell x cell luminance means, high-
// passed against a running low-pass (the ON / OFF split), and each T4/T5 neuron's rate is a
// Hassenstein-Reichardt correlator of its polarity and screen direction at its grid point:
//
//   lum      = sum over the cell of (r + g + b) / 3 / 255 (rows, then columns) / cell^2
//   lp      += a_hp * (lum - lp),   hp = lum - lp
//   sig_on   = max(hp, 0),  sig_off = max(-hp, 0)
//   right    = d * sig[x + k] - sig * d[x + k]      (0 in the last k columns)
//   down     = d * sig[y + k] - sig * d[y + k]      (0 in the last k rows)
//   motion   = max(right, 0) | max(-right, 0) | max(down, 0) | max(-down, 0)   (right/left/down/up)
//   d       += a_d * (sig - d)                       (after every correlator has read it)
//
// The feature channels give m = feature / gain with
//
//   transient  = tr_gain * sig_on | tr_gain * sig_off
//   sustained  = su_gain * max(con, 0) | su_gain * max(-con, 0),
//                con = lum - (sum of lum over the (2 r + 1)^2 window clipped to the grid, row-major)
//                      / (cells in the clipped window)
//   colour     = ch_gain * (sum over the cell of blue (green) / 255) / cell^2
//
// and every cell's rate = max_rate * clip(gain * m, 0, 1), 0 above y_min.
//
// The grid's signals and delayed signals of the frame sit in LDS (4 floats per grid point, plus
// the luminance with the sustained channel and the two colour means with the colour channel),
// the filter state (lp, d_on, d_off) in HBM between frames.  Float32 as written
// (-ffp-contract=off); the CPU restatement in the same order gives the same bits.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "errors.h"
#include "flyhip.h"

namespace {

constexpr int EYE_THREADS = 1024;

__global__ void __launch_bounds__(EYE_THREADS) eye_step(fz_eye_args a) {
    extern __shared__ float sm[];
    const int b = blockIdx.x;
    const int G = a.gh * a.gw;
    const bool sust = a.su_gain != 0.0f, chrom = a.ch_gain != 0.0f;
    float* s_on = sm;          // [G] this frame's ON signal
    float* s_off = sm + G;     // [G] OFF signal
    float* d_on = sm + 2 * G;  // [G] delayed ON (before this frame's update)
    float* d_off = sm + 3 * G; // [G] delayed OFF
    float* s_lum = sm + 4 * G;                   // [G] luminance (sustained channel)
    float* s_blue = sm + (sust ? 5 : 4) * G;     // [G] blue cell mean (colour channel)
    float* s_green = s_blue + G;                 // [G] green cell mean
    const unsigned char* fr = a.frames + (size_t)b * a.h * a.w * 3;
    float* lp = a.lp + (size_t)b * G;
    float* dg_on = a.d_on + (size_t)b * G;
    float* dg_off = a.d_off + (size_t)b * G;
    const int c = a.cell;
    const float c2 = (float)(c * c);
    const bool first = a.first && a.first[b];
    for (int q = threadIdx.x; q < G; q += blockDim.x) {
        const int gy = q / a.gw, gx = q - gy * a.gw;
        float sum = 0.0f, sb = 0.0f, sg = 0.0f;
        for (int dy = 0; dy < c; dy++) {
            const unsigned char* row = fr + ((size_t)(gy * c + dy) * a.w + (size_t)gx * c) * 3;
            if (chrom) {
                for (int dx = 0; dx < c; dx++) {
                    const float r = (float)row[3 * dx], g = (float)row[3 * dx + 1], bl = (float)row[3 * dx + 2];
                    sum += (r + g + bl) / 3.0f / 255.0f;
                    sg += g / 255.0f;
                    sb += bl / 255.0f;
                }
            } else {
                for (int dx = 0; dx < c; dx++) {
                    const float r = (float)row[3 * dx], g = (float)row[3 * dx + 1], bl = (float)row[3 * dx + 2];
                    sum += (r + g + bl) / 3.0f / 255.0f;
                }
            }
        }
        const float lum = sum / c2;
        float l = first ? lum : lp[q];
        l = l + a.a_hp * (lum - l);
        lp[q] = l;
        const float hp = lum - l;
        s_on[q] = fmaxf(hp, 0.0f);
        s_off[q] = fmaxf(-hp, 0.0f);
        d_on[q] = first ? 0.0f : dg_on[q];
        d_off[q] = first ? 0.0f : dg_off[q];
        if (sust) s_lum[q] = lum;
        if (chrom) {
            s_blue[q] = sb / c2;
            s_green[q] = sg / c2;
        }
    }
    __syncthreads();
    const int k = a.shift;
    for (int n = threadIdx.x; n < a.n_cells; n += blockDim.x) {
        const int key = a.key[n];
        const int gy = min(max(a.gy[n], 0), a.gh - 1), gx = min(max(a.gx[n], 0), a.gw - 1);
        const int q = gy * a.gw + gx;
        float m = 0.0f;
        if (key < 8) {  // polarity * 4 + direction (0 right, 1 left, 2 down, 3 up)
            const float* s = key < 4 ? s_on : s_off;
            const float* d = key < 4 ? d_on : d_off;
            const int dir = key & 3;
            if (dir < 2) {
                if (gx < a.gw - k) {
                    const float right = d[q] * s[q + k] - s[q] * d[q + k];
                    m = dir == 0 ? fmaxf(right, 0.0f) : fmaxf(-right, 0.0f);
                }
            } else if (gy < a.gh - k) {
                const float down = d[q] * s[q + k * a.gw] - s[q] * d[q + k * a.gw];
                m = dir == 2 ? fmaxf(down, 0.0f) : fmaxf(-down, 0.0f);
            }
        } else {
            float f = 0.0f;
            if (key == FZ_EYE_ON_TRANSIENT) {
                f = a.tr_gain * s_on[q];
            } else if (key == FZ_EYE_OFF_TRANSIENT) {
                f = a.tr_gain * s_off[q];
            } else if ((key == FZ_EYE_ON_SUSTAINED || key == FZ_EYE_OFF_SUSTAINED) && sust) {
                const int r = a.surround;
                const int y0 = max(gy - r, 0), y1 = min(gy + r + 1, a.gh);
                const int x0 = max(gx - r, 0), x1 = min(gx + r + 1, a.gw);
                float box = 0.0f;
                for (int y = y0; y < y1; y++)
                    for (int x = x0; x < x1; x++) box += s_lum[y * a.gw + x];
                const float con = s_lum[q] - box / (float)((y1 - y0) * (x1 - x0));
                f = a.su_gain * (key == FZ_EYE_ON_SUSTAINED ? fmaxf(con, 0.0f) : fmaxf(-con, 0.0f));
            } else if (key == FZ_EYE_BLUE && chrom) {
                f = a.ch_gain * s_blue[q];
            } else if (key == FZ_EYE_GREEN && chrom) {
                f = a.ch_gain * s_green[q];
            }
            m = f / a.gain;
        }
        float rate = a.max_rate * fminf(fmaxf(a.gain * m, 0.0f), 1.0f);
        if ((float)a.py[n] < a.y_cut) rate = 0.0f;
        a.rates[(size_t)b * a.n_cells + n] = rate;
        if (a.p_ext) a.p_ext[(size_t)b * a.n_ext + a.slot[n]] = rate * a.p_scale;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < G; q += blockDim.x) {
        dg_on[q] = d_on[q] + a.a_d * (s_on[q] - d_on[q]);
        dg_off[q] = d_off[q] + a.a_d * (s_off[q] - d_off[q]);
    }
}

}  // namespace

extern "C" {

int fz_eye_step(const fz_eye_args* a, void* stream) {
    if (!a) return fzi_fail(FZ_E_ARG, "null eye arguments");
    if (a->B < 1 || a->h < 1 || a->w < 1 || a->cell < 1 || a->n_cells < 0 || a->shift < 1)
        return fzi_fail(FZ_E_ARG, "bad eye sizes (B %d, frame %d x %d, cell %d, shift %d)", a->B, a->h, a->w,
                        a->cell, a->shift);
    if (a->gh != a->h / a->cell || a->gw != a->w / a->cell || a->gh < 1 || a->gw < 1)
        return fzi_fail(FZ_E_ARG, "grid %d x %d does not match frame %d x %d / cell %d", a->gh, a->gw, a->h, a->w,
                        a->cell);
    if (!a->frames || !a->lp || !a->d_on || !a->d_off || !a->rates)
        return fzi_fail(FZ_E_ARG, "null frames, filter state or rates");
    if (a->n_cells > 0 && (!a->key || !a->gy || !a->gx || !a->py))
        return fzi_fail(FZ_E_ARG, "null per-cell arrays");
    if (a->p_ext && (!a->slot || a->n_ext < 1)) return fzi_fail(FZ_E_ARG, "p_ext needs slots and n_ext");
    if (a->surround < 0) return fzi_fail(FZ_E_ARG, "negative surround radius %d", a->surround);
    if ((a->tr_gain != 0.0f || a->su_gain != 0.0f || a->ch_gain != 0.0f) && !(a->gain > 0.0f))
        return fzi_fail(FZ_E_ARG, "the feature channels need a positive gain");
    const int planes = 4 + (a->su_gain != 0.0f ? 1 : 0) + (a->ch_gain != 0.0f ? 2 : 0);
    const size_t lds = (size_t)planes * a->gh * a->gw * sizeof(float);
    // the device's LDS limit is queried, and the kernel's dynamic-LDS cap raised to it, once
    // per device (the attribute call is slow; doing it every frame cost more than the kernel)
    constexpr int MAX_DEV = 64;
    static int max_lds_of[MAX_DEV] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) return fzi_fail(FZ_E_HIP, "device query failed");
    if (max_lds_of[dev] == 0) {
        int max_lds = 0;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return fzi_fail(FZ_E_HIP, "device query failed");
        if (max_lds > 65536 && hipFuncSetAttribute(reinterpret_cast<const void*>(&eye_step),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, max_lds) != hipSuccess) {
            (void)hipGetLastError();
            return fzi_fail(FZ_E_HIP, "hipFuncSetAttribute(eye LDS %d) failed", max_lds);
        }
        max_lds_of[dev] = max_lds;
    }
    if (lds > (size_t)max_lds_of[dev])
        return fzi_fail(FZ_E_ARG, "grid %d x %d needs %zu B of LDS (device: %d)", a->gh, a->gw, lds, max_lds_of[dev]);
    hipLaunchKernelGGL(eye_step, dim3(a->B), dim3(EYE_THREADS), lds, static_cast<hipStream_t>(stream), *a);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "eye_step launch failed");
    return FZ_OK;
}

}  // extern "C"
