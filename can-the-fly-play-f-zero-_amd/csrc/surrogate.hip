// Learned rate surrogate: BF16 MFMA GEMMs (FP32 accumulate) and the element-wise training and
// inference kernels of a small network that maps pooled visual drive to descending-neuron rates.
//
// Network (per frame; batch of independent sequences, one GRU state each):
//   h1 = gelu(X W1^T + b1)            2048 -> 1024
//   h2 = gelu(h1 W2^T + b2)           1024 -> 1024
//   s' = GRU(h2, s)                   1024 -> 256 (PyTorch gate order r, z, n)
//   y  = [h2, s'] Wo^T + bo           1280 -> 8
//
// sg_gemm: C[M][N] = sum_k A[m][k] B[k][n], A and B bf16 (each optionally stored transposed),
// FP32 accumulation on v_mfma_f32_16x16x32_bf16.  Workgroup tile 64 x 64 x 32 (128 x 128 in
// sg_gemm_big for the large, aligned products) staged in LDS
// (k-contiguous rows for both operands), four wavefronts of 32 x 32, bounds-checked
// epilogues: FP32 store, FP32 accumulate (gradient sums), BF16 store, bias + GELU (FP32
// pre-activation and BF16 activation).  K must be a multiple of 32.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "errors.h"
#include "flyhip.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even (NaN kept quiet)
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) {  // tanh approximation
    const float c = 0.7978845608028654f;
    return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
}
__device__ __forceinline__ float sigm(float x) { return 1.0f / (1.0f + expf(-x)); }

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS row stride padded against bank conflicts

struct GemmArgs {
    int M, N, K;
    const uint16_t* A;  // [M][K] (lda) or, transposed, [K][M]
    const uint16_t* B;  // [K][N] (ldb) or, transposed, [N][K]
    int lda, ldb, ldc;
    int mode;           // FZ_SG_*
    float* C32;         // modes STORE32, ACC32; pre-activation for GELU
    uint16_t* C16;      // modes STORE16, GELU, GELU16
    const float* bias;  // [N] or null
    long long sa, sb, sc, sbias;  // batched: element offsets between the products of blockIdx.z
};

// The product of blockIdx.z in a batched launch (every operand offset by z times its stride).
__device__ __forceinline__ void batch_offsets(GemmArgs& g) {
    const long long z = blockIdx.z;
    if (z == 0) return;
    g.A += z * g.sa;
    g.B += z * g.sb;
    if (g.C32) g.C32 += z * g.sc;
    if (g.C16) g.C16 += z * g.sc;
    if (g.bias) g.bias += z * g.sbias;
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) sg_gemm(GemmArgs g) {
    batch_offsets(g);
    __shared__ uint16_t As[BM][LDK];
    __shared__ uint16_t Bs[BN][LDK];  // B staged as [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        // A tile: 64 x 32 bf16 = 256 threads x 8
        if (!TA) {
            const int r = tid >> 2, c = (tid & 3) * 8, m = m0 + r;
#pragma unroll
            for (int q = 0; q < 8; q++) As[r][c + q] = m < g.M ? g.A[(size_t)m * g.lda + k0 + c + q] : 0;
        } else {
            const int kr = tid >> 3, c = (tid & 7) * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int m = m0 + c + q;
                As[c + q][kr] = m < g.M ? g.A[(size_t)(k0 + kr) * g.lda + m] : 0;
            }
        }
        if (!TB) {
            const int kr = tid >> 3, c = (tid & 7) * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int n = n0 + c + q;
                Bs[c + q][kr] = n < g.N ? g.B[(size_t)(k0 + kr) * g.ldb + n] : 0;
            }
        } else {
            const int r = tid >> 2, c = (tid & 3) * 8, n = n0 + r;
#pragma unroll
            for (int q = 0; q < 8; q++) Bs[r][c + q] = n < g.N ? g.B[(size_t)n * g.ldb + k0 + c + q] : 0;
        }
        __syncthreads();
        // fragments: lane holds row (lane % 16), k = 8 * (lane / 16) + 0..7
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 af[2], bfm[2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int q = 0; q < 8; q++) af[a][q] = (short)As[wm + a * 16 + fr][fk + q];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int q = 0; q < 8; q++) bfm[b][q] = (short)Bs[wn + b * 16 + fr][fk + q];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfm[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // output layout: lane holds C[4 * (lane / 16) + i][lane % 16]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + i;
                const int n = n0 + wn + b * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[a][b][i];
                if (g.bias) v += g.bias[n];
                const size_t o = (size_t)m * g.ldc + n;
                switch (g.mode) {
                    case FZ_SG_STORE32: g.C32[o] = v; break;
                    case FZ_SG_ACC32: g.C32[o] += v; break;
                    case FZ_SG_STORE16: g.C16[o] = f2bf(v); break;
                    case FZ_SG_GELU16: g.C16[o] = f2bf(gelu(v)); break;
                    default:  // FZ_SG_GELU
                        g.C32[o] = v;
                        g.C16[o] = f2bf(gelu(v));
                }
            }
}

// The same product with 128 x 128 workgroup tiles for the large GEMMs (M, N >= 128; operands
// 16-byte aligned with leading dimensions a multiple of 8): four wavefronts of 64 x 64 (4 x 4
// MFMA tiles), 16-byte global loads staged through registers (the next k-slice is loaded while
// the current one is multiplied), 16-byte LDS fragment reads.  Operands stored along k (A not
// transposed, B transposed) go to LDS as 16-byte rows; the others are scattered into k-rows.
constexpr int GB_M = 128, GB_N = 128;

typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

template <bool KROWS>  // KROWS: the operand is stored [rows][K]; else [K][rows]
struct TileLoad {
    u32x4 v[2];
    // rows r0 .. r0+127 of the operand, k0 .. k0+31
    __device__ __forceinline__ void load(const uint16_t* p, int ld, int rows, int r0, int k0, int tid) {
        if (KROWS) {
            const int r = tid >> 1, kc = (tid & 1) * 16;
            if (r0 + r < rows) {
                const u32x4* q = reinterpret_cast<const u32x4*>(p + (size_t)(r0 + r) * ld + k0 + kc);
                v[0] = q[0];
                v[1] = q[1];
            } else {
                v[0] = v[1] = u32x4{0u, 0u, 0u, 0u};
            }
        } else {
            const int kr = tid >> 3, rc = (tid & 7) * 16;
            const uint16_t* row = p + (size_t)(k0 + kr) * ld + r0 + rc;
            if (r0 + rc + 16 <= rows) {
                const u32x4* q = reinterpret_cast<const u32x4*>(row);
                v[0] = q[0];
                v[1] = q[1];
            } else {
                uint16_t t[16];
#pragma unroll
                for (int j = 0; j < 16; j++) t[j] = r0 + rc + j < rows ? row[j] : 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    v[0][j] = (unsigned)t[2 * j] | ((unsigned)t[2 * j + 1] << 16);
                    v[1][j] = (unsigned)t[8 + 2 * j] | ((unsigned)t[8 + 2 * j + 1] << 16);
                }
            }
        }
    }
    __device__ __forceinline__ void store(uint16_t (*S)[LDK], int tid) const {
        if (KROWS) {
            const int r = tid >> 1, kc = (tid & 1) * 16;
            u32x4* q = reinterpret_cast<u32x4*>(&S[r][kc]);
            q[0] = v[0];
            q[1] = v[1];
        } else {
            const int kr = tid >> 3, rc = (tid & 7) * 16;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                S[rc + 2 * j][kr] = (uint16_t)(v[0][j] & 0xffffu);
                S[rc + 2 * j + 1][kr] = (uint16_t)(v[0][j] >> 16);
                S[rc + 8 + 2 * j][kr] = (uint16_t)(v[1][j] & 0xffffu);
                S[rc + 8 + 2 * j + 1][kr] = (uint16_t)(v[1][j] >> 16);
            }
        }
    }
};

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) sg_gemm_big(GemmArgs g) {
    batch_offsets(g);
    __shared__ __attribute__((aligned(16))) uint16_t As[GB_M][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GB_N][LDK];  // B staged as [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    TileLoad<!TA> la;
    TileLoad<TB> lb;
    la.load(g.A, g.lda, g.M, m0, 0, tid);
    lb.load(g.B, g.ldb, g.N, n0, 0, tid);
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        la.store(As, tid);
        lb.store(Bs, tid);
        __syncthreads();
        if (k0 + BK < g.K) {  // next k-slice in flight during the MFMAs
            la.load(g.A, g.lda, g.M, m0, k0 + BK, tid);
            lb.load(g.B, g.ldb, g.N, n0, k0 + BK, tid);
        }
        bf16x8 af[4], bfm[4];
#pragma unroll
        for (int a = 0; a < 4; a++) af[a] = *reinterpret_cast<const bf16x8*>(&As[wm + a * 16 + fr][fk]);
#pragma unroll
        for (int b = 0; b < 4; b++) bfm[b] = *reinterpret_cast<const bf16x8*>(&Bs[wn + b * 16 + fr][fk]);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfm[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + i;
                const int n = n0 + wn + b * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[a][b][i];
                if (g.bias) v += g.bias[n];
                const size_t o = (size_t)m * g.ldc + n;
                switch (g.mode) {
                    case FZ_SG_STORE32: g.C32[o] = v; break;
                    case FZ_SG_ACC32: g.C32[o] += v; break;
                    case FZ_SG_STORE16: g.C16[o] = f2bf(v); break;
                    case FZ_SG_GELU16: g.C16[o] = f2bf(gelu(v)); break;
                    default:
                        g.C32[o] = v;
                        g.C16[o] = f2bf(gelu(v));
                }
            }
}

// dpre = dact * gelu'(pre); written as FP32 and as BF16 (the next GEMM operand)
__global__ void sg_gelu_bwd_k(long long n, const float* dact, const float* pre, float* dpre32, uint16_t* dpre16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dact[i] * gelu_grad(pre[i]);
    if (dpre32) dpre32[i] = d;
    if (dpre16) dpre16[i] = f2bf(d);
}

// column sums: db[n] (+)= sum_m d[m][n]
__global__ void sg_colsum_k(int M, int N, const float* d, float* db, int accumulate) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; m++) s += d[(size_t)m * N + n];
    db[n] = accumulate ? db[n] + s : s;
}

// GRU step (PyTorch gate order r, z, n); gx = x W_ih^T + b_ih, gh = h W_hh^T + b_hh, [B][3H]
__global__ void sg_gru_fwd_k(int B, int H, const float* gx, const float* gh, const float* h, float* hn,
                             uint16_t* hn16, float* gates) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * H) return;
    const int b = t / H, j = t - b * H;
    const float* x = gx + (size_t)b * 3 * H;
    const float* y = gh + (size_t)b * 3 * H;
    const float r = sigm(x[j] + y[j]);
    const float z = sigm(x[H + j] + y[H + j]);
    const float nn = tanhf(x[2 * H + j] + r * y[2 * H + j]);
    const float hv = (1.0f - z) * nn + z * h[t];
    hn[t] = hv;
    if (hn16) hn16[t] = f2bf(hv);
    if (gates) {  // saved for the backward pass
        float* gs = gates + (size_t)b * 4 * H;
        gs[j] = r;
        gs[H + j] = z;
        gs[2 * H + j] = nn;
        gs[3 * H + j] = y[2 * H + j];
    }
}

// Output layer of every ensemble member and the ensemble's statistics:
// y_k = ([h2_k, s_k] Wo_k^T + bo_k) * rate_scale for members k < E (bf16 inputs and weights,
// FP32 sums), then per rate the mean over members and their population variance (Welford), and
// the relative disagreement ||std|| / (||mean|| + 1 Hz).  h2 [E][B][hid], s [E][B][H] (bf16);
// Wo_k at w16 + k * wstride + wo_off ([n_pad >= n_out][hid + H]); bo_k at w32 + k * wstride + bo_off.
// A workgroup serves 16 frames with one wavefront per member: the wavefront multiplies the
// frames' [h2_k, s_k] rows (A fragments straight from global memory, 16 bytes per lane) by
// Wo_k^T (B fragments: lane column n < n_out, zero beyond) on v_mfma_f32_16x16x32_bf16 along
// k = hid + H, four k-chunks of loads in flight, and leaves its 16 x n_out outputs in LDS;
// sixteen threads then fold the members in order.  hid, H, wstride and wo_off are multiples of
// 8 and the pointers 16-byte aligned (checked by the host).
constexpr int ENS_MAX = 8, OUT_MAX = 8, ENS_ROWS = 16;
__global__ void __launch_bounds__(64 * ENS_MAX) sg_ens_head_k(int B, int hid, int H, int n_out, const uint16_t* h2,
                                                              const uint16_t* s, const uint16_t* w16, const float* w32,
                                                              long long wstride, long long wo_off, long long bo_off,
                                                              float rate_scale, float* mean_out, float* spread_out) {
    __shared__ float ys[ENS_MAX][ENS_ROWS][OUT_MAX];
    const int E = blockDim.x >> 6, k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int F = hid + H, b0 = blockIdx.x * ENS_ROWS;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const int b = b0 + fr;
    const uint16_t* xh = h2 + ((size_t)k * B + b) * hid;
    const uint16_t* xs = s + ((size_t)k * B + b) * H - hid;  // indexed by g >= hid
    const uint16_t* wr = w16 + k * wstride + wo_off + (size_t)fr * F;
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < F; kk += 4 * 32) {
        bf16x8 af[4], wf[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int g = kk + 32 * u + fk;  // a whole 8-group lies in h2 or in s (hid % 8 == 0)
            af[u] = b < B && g < F ? *reinterpret_cast<const bf16x8*>(g < hid ? xh + g : xs + g) : zero;
            wf[u] = fr < n_out && g < F ? *reinterpret_cast<const bf16x8*>(wr + g) : zero;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[u], wf[u], acc, 0, 0, 0);
    }
    // lane holds C[4 * (lane / 16) + i][lane % 16]: frame row, output column
    if (fr < n_out) {
        const float bo = w32[k * wstride + bo_off + fr];
#pragma unroll
        for (int i = 0; i < 4; i++) ys[k][(lane >> 4) * 4 + i][fr] = (acc[i] + bo) * rate_scale;
    }
    __syncthreads();
    if (threadIdx.x < ENS_ROWS && b0 + (int)threadIdx.x < B) {
        const int r = threadIdx.x, bb = b0 + r;
        float mn2 = 0.f, sd2 = 0.f;
        for (int o = 0; o < n_out; o++) {
            float mean = 0.f, m2 = 0.f;
            for (int q = 0; q < E; q++) {
                const float y = ys[q][r][o];
                const float dlt = y - mean;
                mean += dlt / (float)(q + 1);
                m2 += dlt * (y - mean);
            }
            mean_out[(size_t)bb * n_out + o] = mean;
            mn2 += mean * mean;
            sd2 += m2 / (float)E;
        }
        spread_out[bb] = sqrtf(sd2) / (sqrtf(mn2) + 1.0f);
    }
}

// GRU step backward: from dh' (dhn) to dgx, dgh (FP32 and BF16); the direct term dhn*z is
// added into dh (which already holds the previous step's gradient from the output layer)
__global__ void sg_gru_bwd_k(int B, int H, const float* dhn, const float* h, const float* gates, float* dgx,
                             float* dgh, uint16_t* dgx16, uint16_t* dgh16, float* dh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * H) return;
    const int b = t / H, j = t - b * H;
    const float* gs = gates + (size_t)b * 4 * H;
    const float r = gs[j], z = gs[H + j], nn = gs[2 * H + j], hh = gs[3 * H + j];
    const float d = dhn[t];
    const float dn = d * (1.0f - z) * (1.0f - nn * nn);
    const float dz = d * (h[t] - nn) * z * (1.0f - z);
    const float dr = dn * hh * r * (1.0f - r);
    const size_t o = (size_t)b * 3 * H;
    const float vx[3] = {dr, dz, dn};
    const float vh[3] = {dr, dz, dn * r};
    for (int q = 0; q < 3; q++) {
        dgx[o + q * H + j] = vx[q];
        dgh[o + q * H + j] = vh[q];
        if (dgx16) dgx16[o + q * H + j] = f2bf(vx[q]);
        if (dgh16) dgh16[o + q * H + j] = f2bf(vh[q]);
    }
    dh[t] += d * z;
}

// mean squared error over valid outputs; dy = 2 (y - t) / count; loss accumulated by atomics
__global__ void sg_mse_k(long long n, const float* y, const float* t, float* dy, float scale, float* loss) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float e = y[i] - t[i];
        dy[i] = 2.0f * e * scale;
        l = e * e * scale;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// Adam with bias correction; FP32 master weights, BF16 copy for the GEMMs
__global__ void sg_adam_k(long long n, float* w, const float* grad, float* m, float* v, uint16_t* w16, float lr,
                          float b1, float b2, float eps, float c1, float c2, float wd) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

__global__ void sg_cast_k(long long n, const float* x, uint16_t* y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

int gemm_launch(int count, int M, int N, int K, const void* A, int lda, long long sa, int trans_a, const void* B,
                int ldb, long long sb, int trans_b, int mode, float* C32, void* C16, int ldc, long long sc,
                const float* bias, long long sbias, void* stream) {
    if (count < 1 || count > 65535) return fzi_fail(FZ_E_ARG, "gemm: batch count %d outside [1, 65535]", count);
    if (M < 1 || N < 1 || K < 1 || K % BK) return fzi_fail(FZ_E_ARG, "gemm %dx%dx%d: K must be a positive multiple of %d", M, N, K, BK);
    if (!A || !B) return fzi_fail(FZ_E_ARG, "gemm: null operand");
    if ((mode == FZ_SG_STORE32 || mode == FZ_SG_ACC32 || mode == FZ_SG_GELU) && !C32)
        return fzi_fail(FZ_E_ARG, "gemm: mode %d needs C32", mode);
    if ((mode == FZ_SG_STORE16 || mode == FZ_SG_GELU || mode == FZ_SG_GELU16) && !C16)
        return fzi_fail(FZ_E_ARG, "gemm: mode %d needs C16", mode);
    if (mode < FZ_SG_STORE32 || mode > FZ_SG_GELU16) return fzi_fail(FZ_E_ARG, "gemm: unknown mode %d", mode);
    if (sa < 0 || sb < 0 || sc < 0 || sbias < 0) return fzi_fail(FZ_E_ARG, "gemm: negative batch stride");
    GemmArgs g{M, N, K, static_cast<const uint16_t*>(A), static_cast<const uint16_t*>(B), lda, ldb, ldc, mode, C32,
               static_cast<uint16_t*>(C16), bias, sa, sb, sc, sbias};
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool big = M >= GB_M && N >= GB_N && lda % 8 == 0 && ldb % 8 == 0 && sa % 8 == 0 && sb % 8 == 0 &&
                     (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0;
    if (big) {
        const dim3 gridb((N + GB_N - 1) / GB_N, (M + GB_M - 1) / GB_M, count);
        if (!trans_a && !trans_b) hipLaunchKernelGGL((sg_gemm_big<false, false>), gridb, dim3(256), 0, s, g);
        else if (!trans_a && trans_b) hipLaunchKernelGGL((sg_gemm_big<false, true>), gridb, dim3(256), 0, s, g);
        else if (trans_a && !trans_b) hipLaunchKernelGGL((sg_gemm_big<true, false>), gridb, dim3(256), 0, s, g);
        else hipLaunchKernelGGL((sg_gemm_big<true, true>), gridb, dim3(256), 0, s, g);
        if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "sg_gemm launch failed");
        return FZ_OK;
    }
    const dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, count);
    if (!trans_a && !trans_b) hipLaunchKernelGGL((sg_gemm<false, false>), grid, dim3(256), 0, s, g);
    else if (!trans_a && trans_b) hipLaunchKernelGGL((sg_gemm<false, true>), grid, dim3(256), 0, s, g);
    else if (trans_a && !trans_b) hipLaunchKernelGGL((sg_gemm<true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((sg_gemm<true, true>), grid, dim3(256), 0, s, g);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "sg_gemm launch failed");
    return FZ_OK;
}

}  // namespace

extern "C" {

int fz_sg_gemm(int M, int N, int K, const void* A, int lda, int trans_a, const void* B, int ldb, int trans_b,
               int mode, float* C32, void* C16, int ldc, const float* bias, void* stream) {
    return gemm_launch(1, M, N, K, A, lda, 0, trans_a, B, ldb, 0, trans_b, mode, C32, C16, ldc, 0, bias, 0, stream);
}

int fz_sg_gemm_batched(int count, int M, int N, int K, const void* A, int lda, long long stride_a, int trans_a,
                       const void* B, int ldb, long long stride_b, int trans_b, int mode, float* C32, void* C16,
                       int ldc, long long stride_c, const float* bias, long long stride_bias, void* stream) {
    return gemm_launch(count, M, N, K, A, lda, stride_a, trans_a, B, ldb, stride_b, trans_b, mode, C32, C16, ldc,
                       stride_c, bias, stride_bias, stream);
}

int fz_sg_ens_head(int B, int members, int hid, int H, int n_out, const void* h2, const void* s, const void* w16,
                   const float* w32, long long wstride, long long wo_off, long long bo_off, float rate_scale,
                   float* mean, float* spread, void* stream) {
    if (B < 1 || members < 1 || members > ENS_MAX || n_out < 1 || n_out > OUT_MAX || hid < 1 || H < 1)
        return fzi_fail(FZ_E_ARG, "ens_head: B %d, members %d (1..%d), n_out %d (1..%d)", B, members, ENS_MAX, n_out,
                        OUT_MAX);
    if (!h2 || !s || !w16 || !w32 || !mean || !spread) return fzi_fail(FZ_E_ARG, "ens_head: null pointer");
    if (hid % 8 || H % 8 || wstride % 8 || wo_off % 8 || ((reinterpret_cast<uintptr_t>(h2) | reinterpret_cast<uintptr_t>(s) |
                                                          reinterpret_cast<uintptr_t>(w16)) & 15))
        return fzi_fail(FZ_E_ARG, "ens_head: widths, strides and offsets must be multiples of 8 elements, pointers "
                        "16-byte aligned");
    hipLaunchKernelGGL(sg_ens_head_k, dim3(blocks(B, ENS_ROWS)), dim3(64 * members), 0,
                       static_cast<hipStream_t>(stream), B, hid, H, n_out, static_cast<const uint16_t*>(h2),
                       static_cast<const uint16_t*>(s), static_cast<const uint16_t*>(w16), w32, wstride, wo_off,
                       bo_off, rate_scale, mean, spread);
    if (hipGetLastError() != hipSuccess) return fzi_fail(FZ_E_HIP, "ens_head launch failed");
    return FZ_OK;
}

int fz_sg_gelu_bwd(long long n, const float* dact, const float* pre, float* dpre32, void* dpre16, void* stream) {
    if (n < 1 || !dact || !pre || (!dpre32 && !dpre16)) return fzi_fail(FZ_E_ARG, "gelu_bwd: bad arguments");
    hipLaunchKernelGGL(sg_gelu_bwd_k, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, dact,
                       pre, dpre32, static_cast<uint16_t*>(dpre16));
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "gelu_bwd launch failed");
}

int fz_sg_colsum(int M, int N, const float* d, float* db, int accumulate, void* stream) {
    if (M < 1 || N < 1 || !d || !db) return fzi_fail(FZ_E_ARG, "colsum: bad arguments");
    hipLaunchKernelGGL(sg_colsum_k, dim3(blocks(N, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), M, N, d, db,
                       accumulate);
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "colsum launch failed");
}

int fz_sg_gru_fwd(int B, int H, const float* gx, const float* gh, const float* h, float* hn, void* hn16,
                  float* gates, void* stream) {
    if (B < 1 || H < 1 || !gx || !gh || !h || !hn) return fzi_fail(FZ_E_ARG, "gru_fwd: bad arguments");
    hipLaunchKernelGGL(sg_gru_fwd_k, dim3(blocks((long long)B * H, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       B, H, gx, gh, h, hn, static_cast<uint16_t*>(hn16), gates);
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "gru_fwd launch failed");
}

int fz_sg_gru_bwd(int B, int H, const float* dhn, const float* h, const float* gates, float* dgx, float* dgh,
                  void* dgx16, void* dgh16, float* dh, void* stream) {
    if (B < 1 || H < 1 || !dhn || !h || !gates || !dgx || !dgh || !dh) return fzi_fail(FZ_E_ARG, "gru_bwd: bad arguments");
    hipLaunchKernelGGL(sg_gru_bwd_k, dim3(blocks((long long)B * H, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       B, H, dhn, h, gates, dgx, dgh, static_cast<uint16_t*>(dgx16), static_cast<uint16_t*>(dgh16), dh);
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "gru_bwd launch failed");
}

int fz_sg_mse(long long n, const float* y, const float* t, float* dy, float scale, float* loss, void* stream) {
    if (n < 1 || !y || !t || !dy || !loss) return fzi_fail(FZ_E_ARG, "mse: bad arguments");
    hipLaunchKernelGGL(sg_mse_k, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, y, t, dy,
                       scale, loss);
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "mse launch failed");
}

int fz_sg_adam(long long n, float* w, const float* grad, float* m, float* v, void* w16, float lr, float b1, float b2,
               float eps, float c1, float c2, float wd, void* stream) {
    if (n < 1 || !w || !grad || !m || !v) return fzi_fail(FZ_E_ARG, "adam: bad arguments");
    hipLaunchKernelGGL(sg_adam_k, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, w, grad, m,
                       v, static_cast<uint16_t*>(w16), lr, b1, b2, eps, c1, c2, wd);
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "adam launch failed");
}

int fz_sg_cast(long long n, const float* x, void* y, void* stream) {
    if (n < 1 || !x || !y) return fzi_fail(FZ_E_ARG, "cast: bad arguments");
    hipLaunchKernelGGL(sg_cast_k, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, x,
                       static_cast<uint16_t*>(y));
    return hipGetLastError() == hipSuccess ? FZ_OK : fzi_fail(FZ_E_HIP, "cast launch failed");
}

}  // extern "C"
