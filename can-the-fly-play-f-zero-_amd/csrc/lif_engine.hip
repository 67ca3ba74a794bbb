// flyhip LIF engine: CDNA4 (gfx950) kernels and the C ABI declared in include/flyhip.h.
//
// Hot path of flyzero/brain.py:32-58 and flyzero/gpu.py:33-118, re-designed for MI355X:
//
//  * Temporal blocking.  A spike at step s lands exactly n_delay steps later
//    (brain.py:54-56), so for K <= n_delay consecutive steps every neuron's
//    input is already sitting in the ring.  `lif_update` advances each neuron
//    through a whole block of K steps with its state in registers: v, g, ref
//    are read and written once per block instead of once per step.
//  * Spikes of the block are compacted with a wavefront ballot + prefix sum and
//    one 64-bit atomic per wavefront that reserves both the list entries and
//    the range of outgoing edges they own, so the edge space of the block is
//    one contiguous, ordered range.
//  * `lif_push` splits that edge range into equal chunks (perfect load balance
//    for any row-length distribution), stages the row segments of a chunk in
//    LDS and scatters signed integer synapse counts into the int32 ring with
//    no-return atomics.  Integer adds commute, so the result is bit-exact and
//    independent of scheduling.
//  * One window is a chain of (update, push) pairs plus a one-thread clock
//    kernel; it can be instantiated once as a HIP graph and replayed.
//
// The membrane potential is held relative to rest (u = v - v_rest): the model
// is unchanged, and near threshold every float32 rounding is 8x finer than
// with |v| ~ 45 mV.  Must be compiled with -ffp-contract=off: the float update
// is evaluated exactly as written so that it matches the CPU oracle bit for bit.

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "flyhip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(FZ_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

constexpr int KMAX = 32;                          // steps per block held in registers
constexpr int ENT_SHIFT = 36;                     // packed counter: entries << 36 | edges
constexpr unsigned long long EDGE_MASK = (1ull << ENT_SHIFT) - 1;
constexpr long long ENT_LIMIT = 1ll << (64 - ENT_SHIFT);
constexpr int PUSH_MAX_THREADS = 1024;

struct Consts {
    float decay_syn, decay_m, u_th, u_reset, kick, in_scale, pw_qscale;
    int n_ref, n_delay, frac_bits;
};

// lowbias32 integer finaliser (C. Wellons, hash-prospector), the same mixing
// function as the reference's counter RNG (flyzero/gpu.py:29-31).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

struct UpdArgs {
    const long long* clock;
    int N, B, K, kb, blk, first;
    float* u;              // membrane potential relative to rest, mV
    float* g;
    signed char* ref;
    int* ring;
    int* counts;
    const int* ext_idx;
    int n_ext;
    const float* p_ext;
    const float* bias_ext;
    const unsigned* hits;  // [steps][B][hw] or null
    int hw;
    unsigned* raster;      // [steps][B][rw] or null
    int rw;
    unsigned seed;
    const unsigned char* alive;
    const int* deg;        // fixed + plastic out-degree per neuron
    unsigned long long* ctr;  // [2], parity-selected
    int* ent_nrn;
    int* ent_meta;            // fly << 8 | ring slot
    unsigned long long* ent_off;
    long long cap;
    int* err;
    Consts c;
};

// One thread per (fly, neuron); advances K steps of flyzero/brain.py:37-56.
__global__ void __launch_bounds__(1024) lif_update(UpdArgs a) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < a.N;
    const size_t bn = (size_t)b * a.N;
    if (a.alive && !a.alive[b]) {  // retired fly: uniform per workgroup
        if (a.first && valid) a.counts[bn + i] = 0;
        return;
    }
    const long long t0 = a.clock[0] + a.kb;
    const uint32_t step0 = (uint32_t)a.clock[1] + (uint32_t)a.kb;
    const int nd = a.c.n_delay;
    const int slot0 = (int)(((t0 % nd) + nd) % nd);
    int* ringb = a.ring + (size_t)b * nd * a.N;

    // Issue every ring-slot load of the block up front (K independent loads).
    int acc[KMAX];
    float u = 0.f, g = 0.f, bias = 0.f, p = 0.f;
    int r = 0, j = -1;
    if (valid) {
        int s = slot0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k < a.K) {
                acc[k] = ringb[(size_t)s * a.N + i];
                s = (s + 1 == nd) ? 0 : s + 1;
            } else {
                acc[k] = 0;
            }
        }
        u = a.u[bn + i];
        g = a.g[bn + i];
        r = a.ref[bn + i];
        j = a.ext_idx ? a.ext_idx[i] : -1;
        if (j >= 0) {
            const size_t e = (size_t)b * a.n_ext + j;
            if (a.bias_ext) bias = a.bias_ext[e];
            if (a.p_ext) p = a.p_ext[e];
        }
        // consume the slots (brain.py:41); untouched entries are already zero
        s = slot0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k < a.K) {
                if (acc[k] != 0) ringb[(size_t)s * a.N + i] = 0;
                s = (s + 1 == nd) ? 0 : s + 1;
            }
        }
    }
    const float u_in = u, g_in = g;
    const int r_in = r;

    const uint32_t hseed = mix32(a.seed + (uint32_t)b);
    const uint32_t tid32 = (uint32_t)(bn + (size_t)i);
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < a.K) {
            const float inj = (float)acc[k] * a.c.in_scale;
            g = g * a.c.decay_syn + inj;                                   // brain.py:40
            if (r > 0) {
                r -= 1;                                                    // brain.py:42-43
            } else {
                u += (g + bias - u) * a.c.decay_m;                         // brain.py:45 with u = v - v_rest
            }
            bool hit = false;
            if (j >= 0) {
                if (a.hits) {
                    hit = (a.hits[((size_t)(a.kb + k) * a.B + b) * a.hw + (j >> 5)] >> (j & 31)) & 1u;
                } else if (p > 0.f) {                                      // gpu.py:65-67
                    const uint32_t h = mix32(tid32 ^ mix32((step0 + (uint32_t)k) * 0x9e3779b9U ^ hseed));
                    hit = (float)(h >> 8) * (1.0f / 16777216.0f) < p;
                }
            }
            if (hit) u += a.c.kick;                                        // brain.py:46-48
            if (u >= a.c.u_th && r <= 0) {                                 // brain.py:50-53
                u = a.c.u_reset;
                r = a.c.n_ref;
                mask |= 1u << k;
            }
        }
    }

    if (valid) {
        if (u != u_in) a.u[bn + i] = u;
        if (g != g_in) a.g[bn + i] = g;
        if (r != r_in) a.ref[bn + i] = (signed char)r;
        const int ns = __popc(mask);
        if (a.first) a.counts[bn + i] = ns;
        else if (ns) a.counts[bn + i] += ns;
    }

    if (a.raster) {
        const int wbase = (blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> 5;
        for (int k = 0; k < a.K; k++) {
            const unsigned long long bal = __ballot(valid && ((mask >> k) & 1u));
            unsigned* row = a.raster + ((size_t)(a.kb + k) * a.B + b) * a.rw;
            if (lane == 0 && wbase < a.rw) row[wbase] = (unsigned)bal;
            if (lane == 32 && wbase + 1 < a.rw) row[wbase + 1] = (unsigned)(bal >> 32);
        }
    }

    // Spike work list: wavefront prefix sum, one packed atomic per wavefront.
    int d = 0, nsp = 0;
    if (valid && mask) {
        d = a.deg[i];
        if (d > 0) nsp = __popc(mask);
    }
    if (__ballot(nsp > 0)) {
        int xn = nsp;
        unsigned long long xe = (unsigned long long)nsp * (unsigned)d;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int yn = __shfl_up(xn, o);
            const unsigned long long ye = __shfl_up(xe, o);
            if (lane >= o) {
                xn += yn;
                xe += ye;
            }
        }
        unsigned long long old = 0;
        if (lane == 63) {
            const int par = (int)((a.clock[2] + a.blk) & 1);
            old = atomicAdd(&a.ctr[par], ((unsigned long long)xn << ENT_SHIFT) + xe);
        }
        old = __shfl(old, 63);
        if (nsp) {
            long long pos = (long long)(old >> ENT_SHIFT) + xn - nsp;
            unsigned long long e = (old & EDGE_MASK) + xe - (unsigned long long)nsp * (unsigned)d;
            if (pos + nsp > a.cap) {
                atomicOr(a.err, 1);
            } else {
                for (int k = 0; k < a.K; k++) {
                    if ((mask >> k) & 1u) {
                        const int slot = (slot0 + k) % nd;
                        a.ent_nrn[pos] = i;
                        a.ent_meta[pos] = (b << 8) | slot;
                        a.ent_off[pos] = e;
                        pos++;
                        e += (unsigned)d;
                    }
                }
            }
        }
    }
}

template <typename WT>
struct PushArgs {
    const long long* clock;
    int N, n_delay, blk, chunk, frac_bits;
    const int* indptr;
    const int* col;
    const WT* w;
    const int* pl_indptr;
    const int* pl_post;
    const float* pw;
    int n_pl;
    float pw_qscale;
    int* ring;
    unsigned long long* ctr;
    const int* ent_nrn;
    const int* ent_meta;
    const unsigned long long* ent_off;
    long long cap;
};

__device__ __forceinline__ int wval(int16_t x, int s) { return (int)x << s; }
__device__ __forceinline__ int wval(int32_t x, int) { return x; }

// Scatter the block's spikes (flyzero/gpu.py:98-116) over equal edge chunks.
template <typename WT>
__global__ void __launch_bounds__(PUSH_MAX_THREADS) lif_push(PushArgs<WT> a) {
    __shared__ unsigned long long L_off[PUSH_MAX_THREADS];
    __shared__ int L_fbeg[PUSH_MAX_THREADS];
    __shared__ int L_fdeg[PUSH_MAX_THREADS];
    __shared__ int L_pbeg[PUSH_MAX_THREADS];
    __shared__ int L_meta[PUSH_MAX_THREADS];
    __shared__ unsigned long long S_end;
    const int nt = blockDim.x, tid = threadIdx.x;
    const int par = (int)((a.clock[2] + a.blk) & 1);
    const unsigned long long c = a.ctr[par];
    if (blockIdx.x == 0 && tid == 0) a.ctr[par ^ 1] = 0;  // the list the next block fills
    const long long nent = (long long)(c >> ENT_SHIFT);
    const unsigned long long T = c & EDGE_MASK;
    if (nent > a.cap) return;  // the update kernel flagged the overflow and wrote no entries
    const size_t ring_fly = (size_t)a.n_delay * a.N;

    for (unsigned long long e0 = (unsigned long long)blockIdx.x * a.chunk; e0 < T;
         e0 += (unsigned long long)gridDim.x * a.chunk) {
        const unsigned long long e1 = e0 + (unsigned long long)a.chunk < T ? e0 + a.chunk : T;
        // last entry whose edge range starts at or before e0 (offsets ascend)
        long long lo = 0, hi = nent;
        while (hi - lo > 1) {
            const long long m = hi - lo - 1;
            auto posf = [&](long long t) -> long long {
                return lo + 1 + (m >= nt ? (t * m) / nt : (t < m ? t : m - 1));
            };
            const int f = a.ent_off[posf(tid)] <= e0;
            const int cnt = __syncthreads_count(f);
            const long long nlo = cnt > 0 ? posf(cnt - 1) : lo;
            const long long nhi = cnt < nt ? posf(cnt) : hi;
            lo = nlo;
            hi = nhi;
        }
        long long s = lo;
        while (true) {
            const int mc = (int)((nent - s) < nt ? (nent - s) : nt);
            if (tid < mc) {
                const long long q = s + tid;
                const int nrn = a.ent_nrn[q];
                const int fb = a.indptr[nrn];
                L_off[tid] = a.ent_off[q];
                L_fbeg[tid] = fb;
                L_fdeg[tid] = a.indptr[nrn + 1] - fb;
                L_pbeg[tid] = a.pl_indptr ? a.pl_indptr[nrn] : 0;
                L_meta[tid] = a.ent_meta[q];
            }
            if (tid == 0) S_end = (s + mc < nent) ? a.ent_off[s + mc] : T;
            __syncthreads();
            const unsigned long long st_end = S_end;
            const unsigned long long lo_e = e0 > L_off[0] ? e0 : L_off[0];
            const unsigned long long hi_e = e1 < st_end ? e1 : st_end;
            for (unsigned long long e = lo_e + tid; e < hi_e; e += nt) {
                int jl = 0, jh = mc;
                while (jh - jl > 1) {
                    const int mid = (jl + jh) >> 1;
                    if (L_off[mid] <= e) jl = mid;
                    else jh = mid;
                }
                const int rr = (int)(e - L_off[jl]);
                const int meta = L_meta[jl];
                const int b = meta >> 8, slot = meta & 255;
                int post, val;
                if (rr < L_fdeg[jl]) {
                    const int k = L_fbeg[jl] + rr;
                    post = a.col[k];
                    val = wval(a.w[k], a.frac_bits);
                } else {
                    const int k = L_pbeg[jl] + rr - L_fdeg[jl];
                    post = a.pl_post[k];
                    val = (int)rintf(a.pw[(size_t)b * a.n_pl + k] * a.pw_qscale);
                }
                if (val != 0)
                    __hip_atomic_fetch_add(a.ring + (size_t)b * ring_fly + (size_t)slot * a.N + post, val,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            if (st_end >= e1) break;
            s += mc;
        }
    }
}

__global__ void lif_clock(long long* clock, int steps, int nblocks) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        clock[0] += steps;
        clock[1] = (clock[1] + steps) & 0xFFFFFFFFll;
        clock[2] += nblocks;
    }
}

// Integer SpMV with a 0/1 vector: acc[col] += w for every fixed edge of the
// listed neurons (flyzero/brain.py:123-131 with integer sums).  One wavefront
// per listed neuron.
template <typename WT>
__global__ void lif_scatter_rows(int n_list, const int* list, const int* indptr, const int* col, const WT* w,
                                 int frac_bits, int* acc) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= n_list) return;
    const int i = list[wave];
    for (int k = indptr[i] + lane; k < indptr[i + 1]; k += 64) {
        const int val = wval(w[k], frac_bits);
        if (val != 0) atomicAdd(acc + col[k], val);
    }
}

__global__ void lif_degree(int n, const int* indptr, const int* pl_indptr, int* deg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) deg[i] = indptr[i + 1] - indptr[i] + (pl_indptr ? pl_indptr[i + 1] - pl_indptr[i] : 0);
}

}  // namespace

struct fz_engine {
    int n = 0;
    long long nnz = 0;
    const int* indptr = nullptr;
    const int* col = nullptr;
    const void* w = nullptr;
    int w_bits = 16;
    Consts c{};
    int n_pl = 0;
    const int* pl_indptr = nullptr;
    const int* pl_post = nullptr;
    int* deg = nullptr;
    fz_state st{};
    bool bound = false;
    unsigned long long* ctr = nullptr;
    int* ent_nrn = nullptr;
    int* ent_meta = nullptr;
    unsigned long long* ent_off = nullptr;
    long long cap = 0;
    int* err = nullptr;
    int kmax = 1;
    int n_cu = 256;
};

struct fz_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

struct Shape {
    int ut, pt, chunk, grid, K;
};

int resolve(const fz_engine* e, const fz_tune* t, Shape* s) {
    s->ut = (t && t->upd_threads) ? t->upd_threads : 256;
    s->pt = (t && t->push_threads) ? t->push_threads : 256;
    s->chunk = (t && t->push_chunk) ? t->push_chunk : 2048;
    s->grid = (t && t->push_grid) ? t->push_grid : e->n_cu * 8;
    s->K = (t && t->block_steps) ? t->block_steps : e->kmax;
    if (s->ut < 64 || s->ut > 1024 || s->ut % 64) return fail(FZ_E_ARG, "upd_threads %d: need a multiple of 64 in [64, 1024]", s->ut);
    if (s->pt < 64 || s->pt > PUSH_MAX_THREADS || s->pt % 64)
        return fail(FZ_E_ARG, "push_threads %d: need a multiple of 64 in [64, %d]", s->pt, PUSH_MAX_THREADS);
    if (s->chunk < 1) return fail(FZ_E_ARG, "push_chunk %d < 1", s->chunk);
    if (s->grid < 1) return fail(FZ_E_ARG, "push_grid %d < 1", s->grid);
    if (s->K < 1 || s->K > e->kmax) return fail(FZ_E_ARG, "block_steps %d outside [1, %d]", s->K, e->kmax);
    return FZ_OK;
}

int check_inputs(const fz_engine* e, const fz_inputs* in) {
    if (!e) return fail(FZ_E_ARG, "null engine");
    if (!e->bound) return fail(FZ_E_STATE, "fz_bind has not been called");
    if (!in) return fail(FZ_E_ARG, "null inputs");
    if (in->n_ext < 0) return fail(FZ_E_ARG, "n_ext %d < 0", in->n_ext);
    if (in->n_ext > 0 && (!in->ext_idx || !in->p_ext || !in->bias_ext))
        return fail(FZ_E_ARG, "n_ext > 0 needs ext_idx, p_ext and bias_ext");
    if (e->n_pl > 0 && !in->pw) return fail(FZ_E_ARG, "plastic synapses bound but pw is null");
    return FZ_OK;
}

UpdArgs make_upd(const fz_engine* e, const fz_inputs* in, int K, int kb, int blk, int first,
                 const unsigned* hits, unsigned* raster) {
    UpdArgs a{};
    a.clock = e->st.clock;
    a.N = e->n;
    a.B = e->st.B;
    a.K = K;
    a.kb = kb;
    a.blk = blk;
    a.first = first;
    a.u = e->st.u;
    a.g = e->st.g;
    a.ref = e->st.ref;
    a.ring = e->st.ring;
    a.counts = e->st.counts;
    a.ext_idx = in->n_ext > 0 ? in->ext_idx : nullptr;
    a.n_ext = in->n_ext;
    a.p_ext = in->n_ext > 0 ? in->p_ext : nullptr;
    a.bias_ext = in->n_ext > 0 ? in->bias_ext : nullptr;
    a.hits = hits;
    a.hw = (in->n_ext + 31) / 32;
    a.raster = raster;
    a.rw = (e->n + 31) / 32;
    a.seed = in->seed;
    a.alive = in->alive;
    a.deg = e->deg;
    a.ctr = e->ctr;
    a.ent_nrn = e->ent_nrn;
    a.ent_meta = e->ent_meta;
    a.ent_off = e->ent_off;
    a.cap = e->cap;
    a.err = e->err;
    a.c = e->c;
    return a;
}

template <typename WT>
PushArgs<WT> make_push(const fz_engine* e, const fz_inputs* in, int blk, int chunk) {
    PushArgs<WT> a{};
    a.clock = e->st.clock;
    a.N = e->n;
    a.n_delay = e->c.n_delay;
    a.blk = blk;
    a.chunk = chunk;
    a.frac_bits = e->c.frac_bits;
    a.indptr = e->indptr;
    a.col = e->col;
    a.w = static_cast<const WT*>(e->w);
    a.pl_indptr = e->n_pl > 0 ? e->pl_indptr : nullptr;
    a.pl_post = e->pl_post;
    a.pw = in->pw;
    a.n_pl = e->n_pl;
    a.pw_qscale = e->c.pw_qscale;
    a.ring = e->st.ring;
    a.ctr = e->ctr;
    a.ent_nrn = e->ent_nrn;
    a.ent_meta = e->ent_meta;
    a.ent_off = e->ent_off;
    a.cap = e->cap;
    return a;
}

// Enqueue one window.  With `ev` non-null, events bracket every kernel.
int enqueue(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
            const Shape& s, hipStream_t stream, std::vector<hipEvent_t>* ev_upd, std::vector<hipEvent_t>* ev_push) {
    const dim3 ugrid((e->n + s.ut - 1) / s.ut, e->st.B);
    int kb = 0, blk = 0;
    while (kb < steps) {
        const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
        UpdArgs ua = make_upd(e, in, K, kb, blk, kb == 0, hits, raster);
        if (ev_upd) {
            hipEvent_t a0, a1;
            HIPCHK(hipEventCreate(&a0));
            HIPCHK(hipEventCreate(&a1));
            ev_upd->push_back(a0);
            ev_upd->push_back(a1);
            HIPCHK(hipEventRecord(a0, stream));
        }
        hipLaunchKernelGGL(lif_update, ugrid, dim3(s.ut), 0, stream, ua);
        HIPCHK(hipGetLastError());
        if (ev_upd) HIPCHK(hipEventRecord(ev_upd->back(), stream));
        if (ev_push) {
            hipEvent_t a0, a1;
            HIPCHK(hipEventCreate(&a0));
            HIPCHK(hipEventCreate(&a1));
            ev_push->push_back(a0);
            ev_push->push_back(a1);
            HIPCHK(hipEventRecord(a0, stream));
        }
        if (e->w_bits == 16) {
            hipLaunchKernelGGL(lif_push<int16_t>, dim3(s.grid), dim3(s.pt), 0, stream,
                               make_push<int16_t>(e, in, blk, s.chunk));
        } else {
            hipLaunchKernelGGL(lif_push<int32_t>, dim3(s.grid), dim3(s.pt), 0, stream,
                               make_push<int32_t>(e, in, blk, s.chunk));
        }
        HIPCHK(hipGetLastError());
        if (ev_push) HIPCHK(hipEventRecord(ev_push->back(), stream));
        kb += K;
        blk++;
    }
    hipLaunchKernelGGL(lif_clock, dim3(1), dim3(64), 0, stream, e->st.clock, steps, blk);
    HIPCHK(hipGetLastError());
    return FZ_OK;
}

}  // namespace

extern "C" {

int fz_version(void) { return FZ_ABI_VERSION; }

const char* fz_last_error(void) { return g_err.c_str(); }

int fz_create(int n, long long nnz, const int* indptr, const int* col, const void* w, int w_bits, int frac_bits,
              const fz_params* p, fz_engine** out) {
    if (!out) return fail(FZ_E_ARG, "null output handle");
    *out = nullptr;
    if (n < 0 || n >= (1 << 30)) return fail(FZ_E_ARG, "n = %d out of range", n);
    if (nnz < 0 || nnz >= (1ll << 31)) return fail(FZ_E_ARG, "nnz = %lld out of int32 range", nnz);
    if (!p) return fail(FZ_E_ARG, "null params");
    if (n > 0 && !indptr) return fail(FZ_E_ARG, "null indptr");
    if (nnz > 0 && (!col || !w)) return fail(FZ_E_ARG, "null col or w");
    if (w_bits != 16 && w_bits != 32) return fail(FZ_E_ARG, "w_bits must be 16 or 32, got %d", w_bits);
    if (frac_bits < 0 || frac_bits > 24) return fail(FZ_E_ARG, "frac_bits %d outside [0, 24]", frac_bits);
    if (p->n_delay < 1 || p->n_delay > 256) return fail(FZ_E_ARG, "n_delay %d outside [1, 256]", p->n_delay);
    if (p->n_ref < 0 || p->n_ref > 127) return fail(FZ_E_ARG, "n_ref %d outside [0, 127] (int8 counter)", p->n_ref);
    fz_engine* e = new fz_engine();
    e->n = n;
    e->nnz = nnz;
    e->indptr = indptr;
    e->col = col;
    e->w = w;
    e->w_bits = w_bits;
    e->c = Consts{p->decay_syn, p->decay_m, p->u_th, p->u_reset, p->kick, p->in_scale, p->pw_qscale,
                  p->n_ref, p->n_delay, frac_bits};
    e->kmax = p->n_delay < KMAX ? p->n_delay : KMAX;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        e->n_cu = prop.multiProcessorCount;
    if (hipMalloc(&e->deg, sizeof(int) * (size_t)(n > 0 ? n : 1)) != hipSuccess ||
        hipMalloc(&e->ctr, sizeof(unsigned long long) * 2) != hipSuccess ||
        hipMalloc(&e->err, sizeof(int)) != hipSuccess) {
        fz_destroy(e);
        return fail(FZ_E_HIP, "hipMalloc of engine tables failed");
    }
    if (hipMemset(e->ctr, 0, sizeof(unsigned long long) * 2) != hipSuccess ||
        hipMemset(e->err, 0, sizeof(int)) != hipSuccess) {
        fz_destroy(e);
        return fail(FZ_E_HIP, "hipMemset of engine tables failed");
    }
    if (n > 0) {
        hipLaunchKernelGGL(lif_degree, dim3((n + 255) / 256), dim3(256), 0, nullptr, n, indptr, nullptr, e->deg);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            fz_destroy(e);
            return fail(FZ_E_HIP, "degree table kernel failed");
        }
    }
    *out = e;
    return FZ_OK;
}

int fz_destroy(fz_engine* e) {
    if (!e) return FZ_OK;
    hipFree(e->deg);
    hipFree(e->ctr);
    hipFree(e->err);
    hipFree(e->ent_nrn);
    hipFree(e->ent_meta);
    hipFree(e->ent_off);
    delete e;
    return FZ_OK;
}

int fz_set_plastic(fz_engine* e, int n_pl, const int* pl_indptr, const int* pl_post, void* stream) {
    if (!e) return fail(FZ_E_ARG, "null engine");
    if (n_pl < 0) return fail(FZ_E_ARG, "n_pl %d < 0", n_pl);
    if (n_pl > 0 && (!pl_indptr || !pl_post)) return fail(FZ_E_ARG, "n_pl > 0 needs pl_indptr and pl_post");
    e->n_pl = n_pl;
    e->pl_indptr = n_pl > 0 ? pl_indptr : nullptr;
    e->pl_post = n_pl > 0 ? pl_post : nullptr;
    if (e->n > 0) {
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(lif_degree, dim3((e->n + 255) / 256), dim3(256), 0, s, e->n, e->indptr, e->pl_indptr,
                           e->deg);
        HIPCHK(hipGetLastError());
    }
    if (e->bound) {  // re-check capacity limits with the new degree
        fz_state st = e->st;
        return fz_bind(e, &st);
    }
    return FZ_OK;
}

int fz_bind(fz_engine* e, const fz_state* st) {
    if (!e || !st) return fail(FZ_E_ARG, "null engine or state");
    if (st->B < 1 || st->B > 65535) return fail(FZ_E_ARG, "B = %d outside [1, 65535]", st->B);
    if (st->B >= (1 << 23)) return fail(FZ_E_ARG, "B too large for the work-list encoding");
    if (!st->u || !st->g || !st->ref || !st->ring || !st->counts || !st->clock)
        return fail(FZ_E_ARG, "null state array");
    const int per = e->c.n_ref > 1 ? e->c.n_ref : 1;
    const long long maxspk = (e->kmax + per - 1) / per;  // spikes per neuron per block
    const long long cap = (long long)st->B * e->n * maxspk;
    if (cap >= ENT_LIMIT) return fail(FZ_E_ARG, "B*n too large for one work list (%lld entries)", cap);
    const double edges = (double)st->B * (double)(e->nnz + e->n_pl) * (double)maxspk;
    if (edges >= (double)EDGE_MASK) return fail(FZ_E_ARG, "B*nnz too large for one work list");
    if (cap > e->cap) {
        hipFree(e->ent_nrn);
        hipFree(e->ent_meta);
        hipFree(e->ent_off);
        e->ent_nrn = nullptr;
        e->ent_meta = nullptr;
        e->ent_off = nullptr;
        e->cap = 0;
        const size_t c1 = (size_t)(cap > 0 ? cap : 1);
        HIPCHK(hipMalloc(&e->ent_nrn, sizeof(int) * c1));
        HIPCHK(hipMalloc(&e->ent_meta, sizeof(int) * c1));
        HIPCHK(hipMalloc(&e->ent_off, sizeof(unsigned long long) * c1));
        e->cap = cap;
    }
    e->st = *st;
    e->bound = true;
    return FZ_OK;
}

int fz_run(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
           const fz_tune* tune, void* stream) {
    int rc = check_inputs(e, in);
    if (rc) return rc;
    if (steps < 0) return fail(FZ_E_ARG, "steps %d < 0", steps);
    if (steps == 0) return FZ_OK;
    Shape s;
    if ((rc = resolve(e, tune, &s))) return rc;
    return enqueue(e, in, steps, hits, raster, s, static_cast<hipStream_t>(stream), nullptr, nullptr);
}

int fz_run_timed(fz_engine* e, const fz_inputs* in, int steps, const fz_tune* tune, void* stream, float* upd_ms,
                 int* upd_launches, float* push_ms, int* push_launches) {
    int rc = check_inputs(e, in);
    if (rc) return rc;
    if (steps < 1) return fail(FZ_E_ARG, "steps %d < 1", steps);
    Shape s;
    if ((rc = resolve(e, tune, &s))) return rc;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    std::vector<hipEvent_t> eu, ep;
    rc = enqueue(e, in, steps, nullptr, nullptr, s, hs, &eu, &ep);
    float tu = 0.f, tp = 0.f;
    if (rc == FZ_OK) {
        if (hipStreamSynchronize(hs) != hipSuccess) rc = fail(FZ_E_HIP, "stream synchronize failed");
        for (size_t k = 0; rc == FZ_OK && k + 1 < eu.size(); k += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, eu[k], eu[k + 1]) != hipSuccess) rc = fail(FZ_E_HIP, "event timing failed");
            tu += ms;
        }
        for (size_t k = 0; rc == FZ_OK && k + 1 < ep.size(); k += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ep[k], ep[k + 1]) != hipSuccess) rc = fail(FZ_E_HIP, "event timing failed");
            tp += ms;
        }
    }
    for (auto ev : eu) hipEventDestroy(ev);
    for (auto ev : ep) hipEventDestroy(ev);
    if (upd_ms) *upd_ms = tu;
    if (push_ms) *push_ms = tp;
    if (upd_launches) *upd_launches = (int)(eu.size() / 2);
    if (push_launches) *push_launches = (int)(ep.size() / 2);
    return rc;
}

int fz_graph_create(fz_engine* e, const fz_inputs* in, int steps, const fz_tune* tune, fz_graph** out) {
    if (!out) return fail(FZ_E_ARG, "null output handle");
    *out = nullptr;
    int rc = check_inputs(e, in);
    if (rc) return rc;
    if (steps < 1) return fail(FZ_E_ARG, "steps %d < 1", steps);
    Shape s;
    if ((rc = resolve(e, tune, &s))) return rc;
    fz_graph* gr = new fz_graph();
    auto bail = [&](const char* what, hipError_t err) {
        if (gr->graph) hipGraphDestroy(gr->graph);
        delete gr;
        return fail(FZ_E_HIP, "%s failed: %s", what, hipGetErrorString(err));
    };
    hipError_t err = hipGraphCreate(&gr->graph, 0);
    if (err != hipSuccess) return bail("hipGraphCreate", err);
    hipGraphNode_t prev = nullptr;
    auto add = [&](void* fn, dim3 grid, dim3 block, void* arg) -> hipError_t {
        hipKernelNodeParams kp{};
        void* args[] = {arg};
        kp.func = fn;
        kp.gridDim = grid;
        kp.blockDim = block;
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node;
        hipError_t r = hipGraphAddKernelNode(&node, gr->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
        if (r == hipSuccess) prev = node;
        return r;
    };
    const dim3 ugrid((e->n + s.ut - 1) / s.ut, e->st.B);
    int kb = 0, blk = 0;
    while (kb < steps) {
        const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
        UpdArgs ua = make_upd(e, in, K, kb, blk, kb == 0, nullptr, nullptr);
        if ((err = add(reinterpret_cast<void*>(&lif_update), ugrid, dim3(s.ut), &ua)) != hipSuccess)
            return bail("hipGraphAddKernelNode(update)", err);
        if (e->w_bits == 16) {
            PushArgs<int16_t> pa = make_push<int16_t>(e, in, blk, s.chunk);
            err = add(reinterpret_cast<void*>(&lif_push<int16_t>), dim3(s.grid), dim3(s.pt), &pa);
        } else {
            PushArgs<int32_t> pa = make_push<int32_t>(e, in, blk, s.chunk);
            err = add(reinterpret_cast<void*>(&lif_push<int32_t>), dim3(s.grid), dim3(s.pt), &pa);
        }
        if (err != hipSuccess) return bail("hipGraphAddKernelNode(push)", err);
        kb += K;
        blk++;
    }
    long long* clock = e->st.clock;
    int st = steps, nb = blk;
    {
        hipKernelNodeParams kp{};
        void* args[] = {&clock, &st, &nb};
        kp.func = reinterpret_cast<void*>(&lif_clock);
        kp.gridDim = dim3(1);
        kp.blockDim = dim3(64);
        kp.kernelParams = args;
        hipGraphNode_t node;
        err = hipGraphAddKernelNode(&node, gr->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
        if (err != hipSuccess) return bail("hipGraphAddKernelNode(clock)", err);
    }
    err = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if (err != hipSuccess) return bail("hipGraphInstantiate", err);
    *out = gr;
    return FZ_OK;
}

int fz_graph_launch(fz_graph* gr, void* stream) {
    if (!gr || !gr->exec) return fail(FZ_E_ARG, "null graph");
    HIPCHK(hipGraphLaunch(gr->exec, static_cast<hipStream_t>(stream)));
    return FZ_OK;
}

int fz_graph_destroy(fz_graph* gr) {
    if (!gr) return FZ_OK;
    if (gr->exec) hipGraphExecDestroy(gr->exec);
    if (gr->graph) hipGraphDestroy(gr->graph);
    delete gr;
    return FZ_OK;
}

int fz_propagate(fz_engine* e, int n_list, const int* list, int* acc, void* stream) {
    if (!e) return fail(FZ_E_ARG, "null engine");
    if (n_list < 0) return fail(FZ_E_ARG, "n_list %d < 0", n_list);
    if (n_list == 0) return FZ_OK;
    if (!list || !acc) return fail(FZ_E_ARG, "null list or acc");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)(((long long)n_list * 64 + 255) / 256));
    if (e->w_bits == 16)
        hipLaunchKernelGGL(lif_scatter_rows<int16_t>, grid, dim3(256), 0, s, n_list, list, e->indptr, e->col,
                           static_cast<const int16_t*>(e->w), e->c.frac_bits, acc);
    else
        hipLaunchKernelGGL(lif_scatter_rows<int32_t>, grid, dim3(256), 0, s, n_list, list, e->indptr, e->col,
                           static_cast<const int32_t*>(e->w), e->c.frac_bits, acc);
    HIPCHK(hipGetLastError());
    return FZ_OK;
}

int fz_check(fz_engine* e, void* stream) {
    if (!e) return fail(FZ_E_ARG, "null engine");
    HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    int h = 0;
    HIPCHK(hipMemcpy(&h, e->err, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        HIPCHK(hipMemset(e->err, 0, sizeof(int)));
        return fail(FZ_E_OVERFLOW, "device capacity check fired (flag %d)", h);
    }
    return FZ_OK;
}

}  // extern "C"
