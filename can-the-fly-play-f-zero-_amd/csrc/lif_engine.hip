// flyhip LIF engine: CDNA4 (gfx950) kernels and the C ABI declared in include/flyhip.h.
//
// Hot path of flyzero/brain.py:32-58 and flyzero/gpu.py:33-118, re-designed for MI355X:
//
//  * Temporal blocking.  A spike at step s lands exactly n_delay steps later
//    (brain.py:54-56), so for K <= n_delay consecutive steps every neuron's
//    input is already sitting in the ring.  `lif_update` advances each neuron
//    through a whole block of K steps with its state in registers: u, g, ref
//    are read and written once per block instead of once per step.
//  * Spikes of the block go to a per-fly work list.  Each wavefront compacts
//    its spikes with a prefix sum over its lanes, the waves of a workgroup
//    combine in LDS, and one 64-bit atomic per workgroup reserves both the
//    list entries and the range of outgoing edges they own, so each fly's
//    edges of the block form one contiguous, ordered range.  Every entry
//    carries its row offsets, and the update kernel records which entry holds
//    the first edge of every fixed-size edge chunk.
//  * `lif_push` gives every workgroup whole edge chunks (perfect load balance
//    for any row-length distribution), stages the chunk's row segments in LDS
//    and scatters signed integer synapse counts into the int32 ring with
//    no-return atomics.  Integer adds commute, so the result is bit-exact and
//    independent of scheduling.
//  * One window is a chain of (update, push) pairs plus a one-thread clock
//    kernel; it can be instantiated once as a HIP graph and replayed.
//
// The membrane potential is held relative to rest (u = v - v_rest): the model
// is unchanged, and near threshold every float32 rounding is 8x finer than
// with |v| ~ 45 mV.  Must be compiled with -ffp-contract=off: the float update
// is evaluated exactly as written so that it matches the CPU oracle bit for bit.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "errors.h"
#include "flyhip.h"

namespace {
thread_local std::string g_err;
}  // namespace

int fzi_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

#define fail fzi_fail

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(FZ_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

constexpr int KMAX = 32;          // steps per block held in registers
constexpr int MAX_THREADS = 1024;
constexpr int MAX_WAVES = MAX_THREADS / 64;
constexpr int QMIN = 64;          // smallest push chunk (sizes the chunk-start table)
constexpr int BK_PIECE = 2;  // edges per tile-bucket entry at most (longer segments are split)
constexpr int BK_ROUND = 4;  // tile-bucket entries per thread and round in lif_tile
constexpr size_t TILE_STATIC_LDS = 1536;  // room left for lif_tile's static shared arrays (at most 1,248 B)
constexpr size_t TP_MAX_BYTES = 1ull << 30;  // largest (first edge, length) pair table kept
constexpr int PL_SERIAL = 2;  // plastic rows up to this length are pushed by their own thread
constexpr int UNROLL = 2;         // tile engine: edge loads in flight per thread (each also pays a search)

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

// Per-fly spike work list of one block (entries of fly b start at b * cap).
struct WorkList {
    unsigned long long* ctr;  // [2][B] packed (entries << 32 | edges), parity-selected
    unsigned* off;            // first edge of the entry within the fly's edge range
    int* fbeg;                // fixed CSR row start
    int* fdeg;                // fixed row length
    int* pbeg;                // plastic CSR row start
    int* slot;                // ring slot the row lands in
    unsigned* cs;             // [B][ncs] entry holding edge c * chunk
    long long cap;            // entries per fly
    long long ncs;            // chunk-start slots per fly
};

// t mod m, the slot of step t in a ring of m < 2^16 slots, in 32-bit arithmetic (the 64-bit
// remainder is a long scalar sequence that every wavefront would run at its start).  Step
// counts stay below 2^32 for 4.9 days of model time at dt = 0.1 ms; past that the high word is
// folded in with 2^32 mod m.
__device__ __forceinline__ int step_slot(long long t, int m) {
    if (t < 0) return (int)(((t % m) + m) % m);
    const unsigned mm = (unsigned)m, lo = (unsigned)t, hi = (unsigned)((unsigned long long)t >> 32);
    if (hi == 0) return (int)(lo % mm);
    const unsigned p32 = (0xffffffffu % mm + 1u) % mm;  // 2^32 mod m
    return (int)(((hi % mm) * p32 + lo % mm) % mm);     // < m^2 + m < 2^32
}

// Division by a divisor that is fixed for a launch (tile count, ring lengths, block length),
// as one multiply-high and two shifts instead of the ~12-instruction reciprocal sequence the
// compiler emits for a runtime divisor (on uniform operands the multiply and shifts are scalar
// instructions).  Granlund and Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), the unsigned round-up multiplier: with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(n, m), n / d = (t + ((n - t) >> min(l, 1))) >>
// max(l - 1, 0), exact for every 32-bit n and 1 <= d < 2^31.  `div32_for` builds it on the host.
struct Div32 {
    unsigned m;
    unsigned s1, s2;
    unsigned d;
};

inline Div32 div32_for(unsigned d) {
    unsigned l = 0;
    while (l < 32 && (1ull << l) < d) l++;
    Div32 r;
    r.m = (unsigned)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    r.s1 = l < 1 ? l : 1;
    r.s2 = l > 1 ? l - 1 : 0;
    r.d = d;
    return r;
}

__device__ __forceinline__ unsigned div32(unsigned n, const Div32& v) {
    const unsigned t = __umulhi(n, v.m);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}

// step_slot with the ring length's divider: the common case (0 <= t < 2^32) takes no division.
__device__ __forceinline__ int step_slot(long long t, const Div32& v) {
    if (t >= 0 && ((unsigned long long)t >> 32) == 0) {
        const unsigned lo = (unsigned)t;
        return (int)(lo - div32(lo, v) * v.d);
    }
    return step_slot(t, (int)v.d);
}

// x mod m for 0 <= x < 2 m: a ring slot a few steps past a slot, without a division.
__device__ __forceinline__ int wrap2(int x, int m) { return x >= m ? x - m : x; }

struct UpdArgs {
    const long long* clock;
    int N, B, K, kb, blk, first, chunk;
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
    unsigned fly_base;     // RNG numbering of fly b is b + fly_base
    const unsigned char* alive;
    const int* indptr;
    const int* pl_indptr;  // null without plastic synapses
    WorkList wl;
    int* err;
    Consts c;
};

// One thread per (fly, neuron); advances K steps of flyzero/brain.py:37-56.
__global__ void __launch_bounds__(MAX_THREADS) lif_update(UpdArgs a) {
    __shared__ unsigned long long s_wave[MAX_WAVES];
    __shared__ unsigned long long s_base;
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool valid = i < a.N;
    const size_t bn = (size_t)b * a.N;
    if (a.alive && !a.alive[b]) {  // retired fly: uniform per workgroup
        if (a.first && valid) a.counts[bn + i] = 0;
        return;
    }
    const long long t0 = a.clock[0] + a.kb;
    const uint32_t step0 = (uint32_t)a.clock[1] + (uint32_t)a.kb;
    const int nd = a.c.n_delay;
    const int slot0 = step_slot(t0, nd);
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

    const uint32_t hb = (uint32_t)b + a.fly_base;               // the fly's number in the draw
    const uint32_t hseed = mix32(a.seed + hb);
    const uint32_t tid32 = hb * (uint32_t)a.N + (uint32_t)i;     // (b * N + i) mod 2^32
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

    // ---- spike work list: lane prefix sums, wave totals combined in LDS, one
    // packed atomic per workgroup on this fly's counter.
    int fb = 0, fd = 0, pb = 0, d = 0, nsp = 0;
    if (valid && mask) {
        fb = a.indptr[i];
        fd = a.indptr[i + 1] - fb;
        d = fd;
        if (a.pl_indptr) {
            pb = a.pl_indptr[i];
            d += a.pl_indptr[i + 1] - pb;
        }
        if (d > 0) nsp = __popc(mask);
    }
    unsigned long long x = ((unsigned long long)nsp << 32) + (unsigned long long)nsp * (unsigned)d;
    const unsigned long long mine = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;  // wave total
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; w++) {
            const unsigned long long t = s_wave[w];
            s_wave[w] = tot;  // exclusive prefix over waves
            tot += t;
        }
        unsigned long long base = 0;
        if (tot) {
            const int par = (int)((a.clock[2] + a.blk) & 1);
            base = atomicAdd(&a.wl.ctr[(size_t)par * a.B + b], tot);
        }
        s_base = base;
    }
    __syncthreads();
    if (nsp) {
        const unsigned long long start = s_base + s_wave[wave] + x - mine;
        long long pos = (long long)(start >> 32);
        unsigned e = (unsigned)(start & 0xffffffffull);
        if (pos + nsp > a.wl.cap) {
            atomicOr(a.err, 1);
        } else {
            const size_t fo = (size_t)b * a.wl.cap;
            unsigned* cs = a.wl.cs + (size_t)b * a.wl.ncs;
            const unsigned q = (unsigned)a.chunk;
            for (int k = 0; k < a.K; k++) {
                if ((mask >> k) & 1u) {
                    a.wl.off[fo + pos] = e;
                    a.wl.fbeg[fo + pos] = fb;
                    a.wl.fdeg[fo + pos] = fd;
                    a.wl.pbeg[fo + pos] = pb;
                    a.wl.slot[fo + pos] = (slot0 + k) % nd;
                    // chunks whose first edge lies in [e, e + d)
                    for (unsigned c = (e + q - 1) / q; c <= (e + (unsigned)d - 1) / q; c++) cs[c] = (unsigned)pos;
                    pos++;
                    e += (unsigned)d;
                }
            }
        }
    }
}

template <typename WT>
struct PushArgs {
    const long long* clock;
    int N, B, n_delay, blk, chunk, frac_bits;
    const int* col;
    const WT* w;
    const int* pl_post;
    const float* pw;
    int n_pl;
    float pw_qscale;
    int* ring;
    WorkList wl;
};

__device__ __forceinline__ int wval(int16_t x, int s) { return (int)x << s; }
__device__ __forceinline__ int wval(int32_t x, int) { return x; }

// Tag for the tile engine's packed fixed edges: one int32 per edge holding the integer
// synaptic value in the high (32 - lbits) bits and the target's offset inside its tile
// (post % TS < 2^lbits) in the low bits, so an edge costs one 4-byte load instead of a
// column load and a weight load from two arrays.
struct PackedEdge {};

template <typename WT>
__global__ void tile_pack(long long nnz, const int* col, const WT* w, int frac_bits, int ts, int lbits, int* pk,
                          int* over) {
    const long long lim = 1ll << (31 - lbits);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (long long)gridDim.x * blockDim.x) {
        const int v = wval(w[e], frac_bits);
        if (v >= lim || v < -lim) *over = 1;
        pk[e] = (int)(((unsigned)v << lbits) | (unsigned)(col[e] % ts));
    }
}

// Summed |packed weight| of each neuron's fixed input edges into s (64-bit, zeroed), then their
// maximum into *mx: the bound under which two block steps' sums share one LDS word.
__global__ void in_strength(long long nnz, const int* col, const int* pk, int lbits, unsigned long long* s) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (long long)gridDim.x * blockDim.x) {
        const int v = pk[e] >> lbits;
        atomicAdd(&s[col[e]], (unsigned long long)(v < 0 ? -(long long)v : (long long)v));
    }
}
__global__ void max_u64(int n, const unsigned long long* s, unsigned long long* mx) {
    unsigned long long m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = s[i] > m ? s[i] : m;
    atomicMax(mx, m);
}

// Scatter fly blockIdx.y's spikes of the block (flyzero/gpu.py:98-116), one
// edge chunk per workgroup iteration.
template <typename WT>
__global__ void __launch_bounds__(MAX_THREADS) lif_push(PushArgs<WT> a) {
    __shared__ unsigned L_off[MAX_THREADS];
    __shared__ int L_fbeg[MAX_THREADS];
    __shared__ int L_fdeg[MAX_THREADS];
    __shared__ int L_pbeg[MAX_THREADS];
    __shared__ int L_slot[MAX_THREADS];
    __shared__ unsigned S_end;
    const int nt = blockDim.x, tid = threadIdx.x;
    const int b = blockIdx.y;
    const int par = (int)((a.clock[2] + a.blk) & 1);
    const unsigned long long c = a.wl.ctr[(size_t)par * a.B + b];
    if (blockIdx.x == 0 && tid == 0) a.wl.ctr[(size_t)(par ^ 1) * a.B + b] = 0;  // the list the next block fills
    const long long nent = (long long)(c >> 32);
    const unsigned T = (unsigned)(c & 0xffffffffull);
    if (nent > a.wl.cap) return;  // the update kernel flagged the overflow and wrote no entries
    const size_t fo = (size_t)b * a.wl.cap;
    const unsigned* off = a.wl.off + fo;
    const unsigned* cs = a.wl.cs + (size_t)b * a.wl.ncs;
    int* ring = a.ring + (size_t)b * a.n_delay * a.N;
    const float* pw = a.pw ? a.pw + (size_t)b * a.n_pl : nullptr;
    const unsigned q = (unsigned)a.chunk;
    const unsigned nch = (T + q - 1) / q;

    for (unsigned ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const unsigned e0 = ch * q;
        const unsigned e1 = (T - e0) < q ? T : e0 + q;
        const long long s_lo = cs[ch];
        const long long s_hi = (ch + 1 < nch) ? (long long)cs[ch + 1] : nent - 1;  // last entry touching the chunk
        for (long long s = s_lo; s <= s_hi; s += nt) {
            const int mc = (int)((s_hi - s + 1) < nt ? (s_hi - s + 1) : nt);
            if (tid < mc) {
                const long long qq = s + tid;
                L_off[tid] = off[qq];
                L_fbeg[tid] = a.wl.fbeg[fo + qq];
                L_fdeg[tid] = a.wl.fdeg[fo + qq];
                L_pbeg[tid] = a.wl.pbeg[fo + qq];
                L_slot[tid] = a.wl.slot[fo + qq];
            }
            if (tid == 0) S_end = (s + mc < nent) ? off[s + mc] : T;
            __syncthreads();
            const unsigned st_end = S_end;
            const unsigned lo_e = e0 > L_off[0] ? e0 : L_off[0];
            const unsigned hi_e = e1 < st_end ? e1 : st_end;
            for (unsigned e = lo_e + tid; e < hi_e; e += nt) {
                int jl = 0, jh = mc;
                while (jh - jl > 1) {
                    const int mid = (jl + jh) >> 1;
                    if (L_off[mid] <= e) jl = mid;
                    else jh = mid;
                }
                const int rr = (int)(e - L_off[jl]);
                int post, val;
                if (rr < L_fdeg[jl]) {
                    const int k = L_fbeg[jl] + rr;
                    post = a.col[k];
                    val = wval(a.w[k], a.frac_bits);
                } else {
                    const int k = L_pbeg[jl] + rr - L_fdeg[jl];
                    post = a.pl_post[k];
                    val = (int)rintf(pw[k] * a.pw_qscale);
                }
                if (val != 0)
                    __hip_atomic_fetch_add(ring + (size_t)L_slot[jl] * a.N + post, val, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Tile engine: arrivals pulled into LDS per neuron tile (no global atomics for
// fixed synapses, no dense ring traffic).
//
// A workgroup owns TS neurons of one fly.  For a block of K <= n_delay steps
// the inputs arriving at steps [t0, t0+K) come from the spikes of steps
// [t0-n_delay, t0+K-n_delay), all already on the fly's per-step spike lists.
// For each listed spike the workgroup reads the segment of the spiking row
// that targets its tile (a per-(row, tile) start table; columns are sorted
// within rows), and adds the signed counts into an LDS accumulator
// acc[TS][K|1] with LDS atomics, after a block-wide prefix sum that spreads the
// segments' edges evenly over the threads.  Plastic synapses are still pushed
// at spike time (their weights may change between windows) into a ring of
// 2*n_delay slots that is read only for the tile's plastic targets; pushes go
// to slots n_delay ahead, so reads and writes in one launch never meet.
// Then each thread advances its NPT neurons K steps from LDS and appends its
// spikes to the per-step lists (one LDS count per step, one global atomic per
// step and workgroup).  Lists have 3*n_delay slots: the launch for [t0, t0+K)
// reads slots t0-n_delay.., writes t0.., and clears the counters of the next
// n_delay steps, three disjoint ranges.

struct TileArgs {
    const long long* clock;
    int N, B, K, KP, kb, blk, first, TS, n_tiles, nd, M, ring_slots;
    int kh;                // steps of the block's first part (K: the block runs as one part)
    float* u;
    float* g;
    signed char* ref;
    int* ring;             // [B][ring_slots][N]: plastic arrivals only
    int* counts;
    const int* ext_idx;
    int n_ext;
    const float* p_ext;
    const float* bias_ext;
    const unsigned* hits;
    int hw;
    unsigned* raster;
    int rw;
    unsigned seed;
    unsigned fly_base;     // RNG numbering of fly b is b + fly_base
    const unsigned char* alive;
    const int* col;
    const void* w;
    int frac_bits;
    const int* pk;         // packed fixed edges (value << lbits | post % TS), or null
    int lbits;
    const int* tb;         // [N][n_tiles + 1] first fixed edge of row i with post >= tile * TS
    const int* pl_indptr;  // null without plastic synapses
    const int* pl_post;
    const float* pw;
    int n_pl;
    float pw_qscale;
    const int* ptg_ptr;    // [n_tiles + 1] plastic target neurons of each tile ...
    const int* ptg_idx;    // ... (null without plastic synapses)
    int* spk_idx;          // [B][M][N]
    int* spk_cnt;          // [B][M]
    const uint16_t* tb16;  // segment table as offsets from the row start, or null
    const int2* tp;        // [N][n_tiles] (first edge, length) of each segment, or null
    const int2* ptp;       // [N][n_tiles] the same over the plastic CSR: plastic edges are pulled
                           // for spikes that arrive in their window; null: all are pushed
    long long nnz;         // fixed edges: a segment or bucket entry nnz + q is plastic edge q
    int win;               // steps of this window
    const int* indptr;
    const int* bk;         // tile buckets (source 1): [2][B][bk_stride] entries (first edge << 7 |
                           // (len - 1) << 5 | step), two buffers alternating with the block parity;
                           // entries of steps < kh from a bucket's front, the others from its back
    const long long* bk_off;  // [n_tiles + 1] offset of each tile's bucket in a fly's arena
    long long bk_stride;
    const int* bk_cnt;     // [2][B][n_tiles][2] front and back entries of each bucket
    int b0, nb;            // this launch covers flies [b0, b0 + nb) (grid n_tiles x nb)
    Div32 dv_tiles, dv_M, dv_ring, dv_K;  // dividers by n_tiles, M, ring_slots and K
    Consts c;
};

// Workgroups are handed to the 8 XCDs round-robin by linear id, and each XCD has its
// own L2.  Map linear id p to the (tile, fly) pair so that XCD p % 8 receives one
// contiguous run of the fly-major order: the tiles of a fly then share an L2, so a
// spike's segment-table row and its row of edges are fetched once per XCD instead of
// by every XCD.  The map is a bijection onto [0, n_tiles * B) for any dispatch order.
constexpr int NXCD = 8;
// (With the tile count's divider: no division instruction sequence.)
__device__ __forceinline__ void xcd_order(const Div32& dt, int B, int* T, int* b) {
    const int n_tiles = (int)dt.d;
    const int total = n_tiles * B;
    const int p = blockIdx.y * gridDim.x + blockIdx.x;
    const int x = p % NXCD, i = p / NXCD;
    const int q = total / NXCD, r = total % NXCD;
    const int L = x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    *b = (int)div32((unsigned)L, dt);
    *T = L - *b * n_tiles;
}

// Offsets of the per-step spike lists a block's arrivals come from (the spikes of steps
// t0 + k - n_delay, k < K): s_off[k] = entries of the source steps before k, s_off[K] = all.
// The first wavefront's lanes load the K list lengths together and scan them with shuffles (K <=
// KMAX < 64), instead of one thread walking them.  Every thread of the workgroup calls it: it
// ends with a barrier.
__device__ __forceinline__ void source_offsets(const int* cnt_b, long long t0, int lt0, int K, int nd, int M, int N,
                                               int* s_off) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int c = 0;
        if (lane < K && t0 + lane - nd >= 0) {
            c = cnt_b[wrap2(lt0 + lane - nd + M, M)];  // lane < K <= nd: below 2 M
            c = c < N ? c : N;
        }
        int x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane < K) s_off[lane] = x - c;
        if (lane == K - 1) s_off[K] = x;
    }
    __syncthreads();
}

// The source step of entry e (0 <= e < s_off[K]) of those lists: the last k with s_off[k] <= e
// (steps without spikes repeat an offset), by bisection.
__device__ __forceinline__ int entry_step(const int* s_off, int K, int e) {
    int lo = 0, hi = K;  // s_off[lo] <= e < s_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Inclusive prefix sum over the workgroup; wsum needs MAX_WAVES + 1 ints.
__device__ __forceinline__ int block_incl_scan(int v, int* wsum, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; w++) {
            const int t = wsum[w];
            wsum[w] = acc;
            acc += t;
        }
        wsum[nw] = acc;
    }
    __syncthreads();
    *total = wsum[nw];
    return x + wsum[wave];
}

// The arrival sum of block step k in a neuron's LDS row.  H16: two steps per word, step 2j in
// the low half (biased by 0x8000, so that a negative addend never borrows from the high half)
// and step 2j + 1 in the high half; exact while every partial sum stays inside int16, which the
// host guarantees by allowing the mode only when each neuron's summed |input weight| is below
// 32768.
template <bool H16>
__device__ __forceinline__ int arrivals(const int* row, int k) {
    if constexpr (H16) {
        const int w = row[k >> 1];
        return (k & 1) ? (w >> 16) : (int)(short)(w ^ 0x8000);
    } else {
        return row[k];
    }
}

// The step loop of a block for the case n_ref >= K (at most one spike per block, the refractory
// counter carried as the block step rend at which it runs out; see the kernel).  KC > 0 is the
// block length known at compile time: the loop unrolls, the LDS reads take immediate offsets
// and the step index is a constant in every compare.  The float operations and their order are
// those of the runtime loop (KC = 0), so the two give the same bits.  KICK = false drops the
// Poisson kick (bit test, add and select) for a wavefront none of whose neurons has a hit in
// the block: without a hit the select keeps u as it is, so the result is the same.  With KICK,
// wh (wavefront-uniform) has bit k set when some lane of the wavefront has a hit at step k; the
// kick of the other steps is skipped by a scalar branch, for the same reason.
template <int KC, bool KICK = true, bool H16 = false>
__device__ __forceinline__ int once_steps(const int* row, int K, float& u_io, float& g_io, int rend, float bias,
                                          uint32_t hits, uint32_t wh, const Consts& c) {
    float u = u_io, g = g_io;
    auto step = [&](int k) {
        const float inj = (float)arrivals<H16>(row, k) * c.in_scale;
        g = g * c.decay_syn + inj;
        const float u_int = u + (g + bias - u) * c.decay_m;
        u = k < rend ? u : u_int;
        const float u_kick = u + c.kick;
        u = ((hits >> k) & 1u) ? u_kick : u;
        const bool sp = u >= c.u_th && k + 1 >= rend;
        u = sp ? c.u_reset : u;
        rend = sp ? k + 1 + c.n_ref : rend;
    };
    if constexpr (KC > 0) {
        // The same steps with fewer VALU operations each.  "held" (k < rend) is carried from the
        // previous step, where it was computed for the spike test: k < rend_new is (k < rend_old)
        // unless a spike just moved rend_new to k + n_ref > k (n_ref >= K >= 1).  After a spike
        // the neuron is held to the end of the block (k' + 1 < k + 1 + n_ref for every k' < K),
        // so "k + 1 < rend" is one compare against the incoming rend or'ed with the lane mask
        // "has spiked", and the spike's step is kept by one select.  The empty asm hides the
        // range of rend (a refractory count that fits a byte), which would otherwise turn the
        // compare into a 16-bit one that cannot take the step as an inline constant.  The kick
        // body holds an empty volatile asm so that it stays behind its scalar branch instead of
        // being turned into selects that run on every step; the block's LDS row is read before
        // the first step (those branches would otherwise hold each read back to its own step)
        // and the reset and kick values sit in VGPRs once, not moved or loaded on every step.
        asm("" : "+v"(rend));
        float u_reset = c.u_reset, kick = c.kick;
        asm("" : "+v"(u_reset));
        asm("" : "+v"(kick));
        int rv[KC];
#pragma unroll
        for (int k = 0; k < KC; k++) rv[k] = arrivals<H16>(row, k);
        bool held = 0 < rend, done = false;
        int ks = 0;
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const float inj = (float)rv[k] * c.in_scale;
            g = g * c.decay_syn + inj;
            const float u_int = u + (g + bias - u) * c.decay_m;
            u = held ? u : u_int;
            if constexpr (KICK) {
                if ((wh >> k) & 1u) {
                    asm volatile("");
                    const float u_kick = u + kick;
                    u = ((hits >> k) & 1u) ? u_kick : u;
                }
            }
            const bool held_next = k + 1 < rend || done;
            const bool ge = u >= c.u_th;
            const bool sp = ge && !held_next;
            u = sp ? u_reset : u;
            ks = sp ? k + 1 : ks;
            done = done || sp;
            held = held_next || ge;  // = held_next || sp
        }
        rend = done ? ks + c.n_ref : rend;
    } else {
        for (int k = 0; k < K; k++) step(k);
    }
    u_io = u;
    g_io = g;
    return rend;
}

// Fills the LDS accumulator, KP x TPB x NPT ints (a multiple of 4: TPB is a multiple of 64),
// with v, in 16-byte stores: thread tid stores int4 number tid + j TPB for j < ceil(KP NPT / 4),
// a trip count without a division.
__device__ __forceinline__ void clear_acc(int* acc, int KP, int NPT, int TPB, int tid, int v = 0) {
    const int nq = (KP * NPT * TPB) >> 2, per = (KP * NPT + 3) >> 2;
    for (int j = 0; j < per; j++) {
        const int q = tid + j * TPB;
        if (q < nq) reinterpret_cast<int4*>(acc)[q] = make_int4(v, v, v, v);
    }
}

// IPT: spike-list entries staged per thread (more loads in flight, more LDS).
// BK: segments come from the tile's bucket (filled by tile_bucket) instead of being looked up
// in the segment table for every spike of the fly.
// PL: plastic synapses are pulled here, with the fly's own weights, for spikes that arrive in
// the window they were emitted in (a.ptp); only the others are pushed at spike time.
// PARTS: 2 runs each block as two parts (half the arrival columns in LDS), 1 as one.
// H16: two block steps' arrival sums per LDS word (see arrivals): a row of ceil(K / 2) + 1
// words instead of K + 1, so more workgroups fit a CU (bucket source, one part, no plastic
// synapses; the host checks the int16 bound).
template <typename WT, int NPT, int IPT, bool BK, bool PL, int PARTS, bool H16 = false>
__global__ void __launch_bounds__(MAX_THREADS) lif_tile(TileArgs a) {
    extern __shared__ __attribute__((aligned(16))) int lds[];
    __shared__ int s_off[KMAX + 1];
    __shared__ int s_lcnt[KMAX];
    __shared__ int s_wsum[MAX_WAVES + 1];
    const int TPB = blockDim.x, tid = threadIdx.x, lane = tid & 63;
    const int TS = TPB * NPT;
    // acc[neuron][KP]: a neuron's arrivals of the block part are adjacent (two steps per ds_read2
    // with immediate offsets in the step loop), and column kh holds its Poisson hits as a K-bit
    // mask.  KP >= kh + 1 is odd when that fits, so the 64 lanes of a wavefront hit 64 different
    // banks.  Offsets are 24-bit products (neuron < 8192, KP <= 33): full-rate multiplies.
    const int KP = a.KP;
    int* acc = lds;                      // [TS][KP]
    int* st_seg = acc + KP * TS;         // [TPB * IPT] (segment-table source only)
    int* st_end = st_seg + TPB * IPT;    // [TPB * IPT] inclusive prefix of segment lengths
    int* st_k = st_end + TPB * IPT;      // [TPB * IPT]
    int b, T;
    xcd_order(a.dv_tiles, a.nb, &T, &b);
    b += a.b0;
    const int i0 = T * TS;
    const size_t bn = (size_t)b * a.N;
    // Every load the set-up needs is issued before any of them is used -- the clock, the fly's
    // alive flag, the bucket's place and both parities' entry counts, and the tile's neuron state
    // -- so their latencies overlap (a retired fly's state loads are wasted).  The state loads
    // come from valid (clamped) addresses and nothing waits for them until the step loop; the
    // loads that need a neuron's ext slot (its Poisson probability and bias) follow the next
    // barrier.  The values of neurons past the end of the last tile are never used.
    const long long clk0 = a.clock[0], clk2 = a.clock[2];
    const uint32_t clk1 = (uint32_t)a.clock[1];
    const bool live = a.alive ? a.alive[b] != 0 : true;
    long long bko0 = 0, bko1 = 0;
    int2 bkc[2] = {make_int2(0, 0), make_int2(0, 0)};
    if constexpr (BK) {
        bko0 = a.bk_off[T];
        bko1 = a.bk_off[T + 1];
        const int2* bc = reinterpret_cast<const int2*>(a.bk_cnt);
        bkc[0] = bc[(size_t)b * a.n_tiles + T];
        bkc[1] = bc[((size_t)a.B + b) * a.n_tiles + T];
    }
    float u_[NPT], g_[NPT], bias_[NPT], pp_[NPT];
    int r_[NPT], j_[NPT];
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const int i = i0 + p * TPB + tid;
        const int ic = i < a.N ? i : a.N - 1;
        j_[p] = a.ext_idx ? a.ext_idx[ic] : -1;
        u_[p] = a.u[bn + ic];
        g_[p] = a.g[bn + ic];
        r_[p] = a.ref[bn + ic];
    }
    const long long t0 = clk0 + a.kb;
    const uint32_t step0 = clk1 + (uint32_t)a.kb;
    const int nd = a.nd, M = a.M;
    const int lt0 = step_slot(t0, a.dv_M);                 // list slot of step t0
    const int rs0 = step_slot(t0, a.dv_ring);
    int* cnt_b = a.spk_cnt + (size_t)b * M;
    int* idx_b = a.spk_idx + (size_t)b * M * a.N;
    if (T == 0 && tid < nd) cnt_b[wrap2(lt0 + a.K + tid, M)] = 0;  // lists of the next n_delay steps (K <= nd)
    if (!live) {                                                // retired fly
        if (a.first)
            for (int p = 0; p < NPT; p++) {
                const int i = i0 + p * TPB + tid;
                if (i < a.N) a.counts[bn + i] = 0;
            }
        return;
    }
    if (tid < KMAX) s_lcnt[tid] = 0;  // spikes per block step (read after the step loop)
    // steps with a Poisson hit in each wavefront's neurons (bit k), per neuron slot
    __shared__ uint32_t s_wh[MAX_WAVES * NPT];
    if (tid < MAX_WAVES * NPT) s_wh[tid] = 0u;
    // KP * TS is a multiple of 4 ints (TS is a multiple of 64): 16-byte stores
    clear_acc(acc, KP, NPT, TPB, tid, H16 ? 0x8000 : 0);
    // per-step keys of the Poisson draw (read after the next barrier)
    __shared__ uint32_t s_sk[KMAX];
    const uint32_t hb = (uint32_t)b + a.fly_base;  // the fly's number in the draw
    const uint32_t hseed = mix32(a.seed + hb);
    if (tid < a.K) s_sk[tid] = mix32((step0 + (uint32_t)tid) * 0x9e3779b9U ^ hseed);
    // Source 1: the tile's bucket holds the entries of the block's first part at its front and
    // those of its second part at its back (an undivided block has only the front part).
    const int* bkp4 = nullptr;  // this tile's bucket: entry e at bkp4[e]
    int bcap = 0, bcf = 0, bcb = 0;  // bucket capacity, front and back entries
    if constexpr (BK) {
        const int par = (int)((clk2 + a.blk) & 1);
        bcap = (int)(bko1 - bko0);
        const int2 c = par ? bkc[1] : bkc[0];
        bcf = c.x < bcap ? c.x : bcap;
        bcb = c.y < bcap - bcf ? c.y : bcap - bcf;
        bkp4 = a.bk + ((size_t)par * a.B + b) * a.bk_stride + bko0;
        __syncthreads();
    } else {
        source_offsets(cnt_b, t0, lt0, a.K, nd, M, a.N, s_off);
    }
    // the ext-slot loads, also from valid addresses (a neuron without a slot reads slot 0); the
    // slot test is applied where the values are used
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        bias_[p] = pp_[p] = 0.f;
        if (a.ext_idx && a.n_ext > 0) {
            const size_t e = (size_t)b * a.n_ext + (j_[p] >= 0 ? j_[p] : 0);
            if (a.bias_ext) bias_[p] = a.bias_ext[e];
            if (a.p_ext) pp_[p] = a.p_ext[e];
        }
    }
    const WT* w = static_cast<const WT*>(a.w);  // unused with packed edges
    const int lmask = (1 << a.lbits) - 1;
    const float* pwf = PL ? a.pw + (size_t)b * a.n_pl : nullptr;  // the fly's plastic weights
    auto edge = [&](int ek, int* loc, int* val) {
        if constexpr (PL) {
            if (ek >= a.nnz) {  // plastic edge q = ek - nnz, the fly's weight in this window
                const int q = ek - (int)a.nnz;
                *loc = a.pl_post[q] - i0;
                *val = (int)rintf(pwf[q] * a.pw_qscale);
                return;
            }
        }
        if constexpr (std::is_same<WT, PackedEdge>::value) {
            const int x = a.pk[ek];  // one load: value and tile-local target
            *loc = x & lmask;
            *val = x >> a.lbits;
        } else {
            *loc = a.col[ek] - i0;
            *val = wval(w[ek], a.frac_bits);
        }
    };
    // adds v to the arrival sum of neuron loc (tile-local) at column col (block step - k0)
    auto acc_add = [&](unsigned loc, int col, int v) {
        if constexpr (H16) atomicAdd(&acc[__umul24(loc, (unsigned)KP) + (col >> 1)], (col & 1) ? (int)((unsigned)v << 16) : v);
        else atomicAdd(&acc[__umul24(loc, (unsigned)KP) + col], v);
    };
    // The arrivals of block steps [k0, k1) into acc columns k - k0 (the LDS atomics of the last
    // round may still be in flight when it returns: a barrier orders them before the step loop).
    auto pull = [&](int k0, int k1) {
        if constexpr (BK) {
            // entries nf of the front part (first part), then nbk of the back part (second part)
            const int nf = k0 == 0 ? bcf : 0, nbk = k1 == a.K ? bcb : 0;
            const int total = nf + nbk, boff = bcap - bcb - nf;  // entry e >= nf sits at e + boff
            if constexpr (std::is_same<WT, PackedEdge>::value && BK_PIECE == 2) {
                // Pieces hold at most two edges (tile segments are short: ~1 edge per (row, tile) on a
                // FlyWire-shaped graph), so each thread takes whole pieces: no scan and no search.
                // Rounds of BK_ROUND entries per thread: every entry load first, then both edges of
                // every piece (the second from the next packed edge, a valid address even past the
                // last fixed edge: the packed array ends with one padding entry, which plastic pieces
                // also read), then the LDS atomics -- two memory round trips per round and no wait in
                // between.  PL: a round with plastic pieces (first edge nnz + q) then loads their
                // targets and the fly's weights of this window.
                const int nnzi = (int)a.nnz;
                for (int e0 = tid; e0 < total; e0 += BK_ROUND * TPB) {
                    int v[BK_ROUND];
#pragma unroll
                    for (int r = 0; r < BK_ROUND; r++) {
                        const int e = e0 + r * TPB < total ? e0 + r * TPB : e0;
                        v[r] = bkp4[e < nf ? e : e + boff];  // first edge << 7 | (len - 1) << 5 | step
                    }
                    int x0[BK_ROUND], x1[BK_ROUND];
#pragma unroll
                    for (int r = 0; r < BK_ROUND; r++) {
                        const int f = v[r] >> 7;
                        const bool fx = !PL || f < nnzi;
                        x0[r] = a.pk[fx ? f : nnzi];
                        x1[r] = a.pk[fx ? f + 1 : nnzi];
                    }
                    int l0[BK_ROUND], w0[BK_ROUND], l1[BK_ROUND], w1[BK_ROUND];
#pragma unroll
                    for (int r = 0; r < BK_ROUND; r++) {
                        l0[r] = x0[r] & lmask;
                        w0[r] = x0[r] >> a.lbits;
                        l1[r] = x1[r] & lmask;
                        w1[r] = (v[r] & (1 << 5)) ? x1[r] >> a.lbits : 0;
                    }
                    if constexpr (PL) {
                        bool pr[BK_ROUND], anyp = false;
#pragma unroll
                        for (int r = 0; r < BK_ROUND; r++) {
                            pr[r] = e0 + r * TPB < total && (v[r] >> 7) >= nnzi;
                            anyp = anyp || pr[r];
                        }
                        if (__ballot(anyp)) {
#pragma unroll
                            for (int r = 0; r < BK_ROUND; r++) {
                                if (!pr[r]) continue;
                                const int q = (v[r] >> 7) - nnzi;
                                l0[r] = a.pl_post[q] - i0;
                                w0[r] = (int)rintf(pwf[q] * a.pw_qscale);
                                if (v[r] & (1 << 5)) {
                                    l1[r] = a.pl_post[q + 1] - i0;
                                    w1[r] = (int)rintf(pwf[q + 1] * a.pw_qscale);
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int r = 0; r < BK_ROUND; r++) {
                        if (e0 + r * TPB >= total) continue;
                        const int col = (v[r] & 31) - k0;
                        if (w0[r] != 0) acc_add((unsigned)l0[r], col, w0[r]);
                        if (w1[r] != 0) acc_add((unsigned)l1[r], col, w1[r]);
                    }
                }
            } else {
                // The general form (unpacked formats): the two entry loads and the
                // first edge of each piece are issued unconditionally from clamped, valid indices and
                // masked afterwards, so they are in flight together; the further edges of a piece are
                // loaded on demand.
                for (int e0 = tid; e0 < total; e0 += 2 * TPB) {
                    int v[2];
                    bool ok[2];
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int e = e0 + r * TPB;
                        ok[r] = e < total;
                        const int ec = ok[r] ? e : e0;
                        v[r] = bkp4[ec < nf ? ec : ec + boff];
                    }
                    int loc[2][BK_PIECE], val[2][BK_PIECE];
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        edge(v[r] >> 7, &loc[r][0], &val[r][0]);
                        if (!ok[r]) val[r][0] = 0;
                    }
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int seg = v[r] >> 7, len = ok[r] ? ((v[r] >> 5) & 3) + 1 : 0;
#pragma unroll
                        for (int j = 1; j < BK_PIECE; j++) {
                            loc[r][j] = val[r][j] = 0;
                            if (j < len) edge(seg + j, &loc[r][j], &val[r][j]);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int col = (v[r] & 31) - k0;
#pragma unroll
                        for (int j = 0; j < BK_PIECE; j++)
                            if (val[r][j] != 0) acc_add((unsigned)loc[r][j], col, val[r][j]);
                    }
                }
            }
        } else {
            // The list entries of source steps k0 .. k1 - 1 in chunks of TPB * IPT; each thread owns
            // IPT consecutive entries so that all of its list and table loads are in flight
            // together; each wavefront stages its segments (any length) with their prefix sums and
            // splits their edges evenly over its lanes (no workgroup barrier in the chunk loop).
            const int lo = s_off[k0], hi = s_off[k1];
            for (int base = lo; base < hi; base += TPB * IPT) {
                int seg[IPT], len[IPT], kk[IPT];
                int sum = 0;
                int ids[IPT];
                // loads from clamped, valid positions (results masked afterwards) so that the IPT
                // list reads, then the IPT table reads, of a thread are in flight together
                bool ok[IPT];
                size_t li_[IPT];
#pragma unroll
                for (int r = 0; r < IPT; r++) {  // list positions first (LDS only) ...
                    const int e = base + tid * IPT + r;
                    ok[r] = e < hi;
                    const int ec = ok[r] ? e : base;
                    kk[r] = entry_step(s_off, a.K, ec);
                    li_[r] = (size_t)wrap2(lt0 + kk[r] - nd + M, M) * a.N + (ec - s_off[kk[r]]);
                }
#pragma unroll
                for (int r = 0; r < IPT; r++) ids[r] = idx_b[li_[r]];  // ... then all list loads at once
                const unsigned nt1 = (unsigned)(a.n_tiles + 1);
                int2 pq[PL ? IPT : 1];  // plastic segments, loaded beside the fixed ones
                if constexpr (PL) {
#pragma unroll
                    for (int r = 0; r < IPT; r++) {
                        const int idc = ok[r] && (unsigned)ids[r] < (unsigned)a.N ? ids[r] : 0;
                        pq[r] = a.ptp[__umul24((unsigned)idc, (unsigned)a.n_tiles) + (unsigned)T];
                    }
                }
                if (a.tp) {  // one 8-byte load per (spike, tile): the L2 request rate bounds this phase
#pragma unroll
                    for (int r = 0; r < IPT; r++) {
                        ok[r] = ok[r] && (unsigned)ids[r] < (unsigned)a.N;
                        const int idc = ok[r] ? ids[r] : 0;
                        const int2 q = a.tp[__umul24((unsigned)idc, (unsigned)a.n_tiles) + (unsigned)T];  // < 2^31
                        seg[r] = q.x;
                        len[r] = q.y & -(int)ok[r];
                    }
                } else if (a.tb16) {
#pragma unroll
                    for (int r = 0; r < IPT; r++) {
                        ok[r] = ok[r] && (unsigned)ids[r] < (unsigned)a.N;
                        const int idc = ok[r] ? ids[r] : 0;
                        const unsigned at = __umul24((unsigned)idc, nt1) + (unsigned)T;  // < 2^31 (checked)
                        const int r0 = a.indptr[idc];
                        const int l16 = a.tb16[at], h16 = a.tb16[at + 1];
                        seg[r] = r0 + l16;
                        len[r] = (h16 - l16) & -(int)ok[r];
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < IPT; r++) {
                        ok[r] = ok[r] && (unsigned)ids[r] < (unsigned)a.N;
                        const int idc = ok[r] ? ids[r] : 0;
                        const unsigned at = __umul24((unsigned)idc, nt1) + (unsigned)T;
                        seg[r] = a.tb[at];
                        len[r] = (a.tb[at + 1] - seg[r]) & -(int)ok[r];  // masked, not branched on
                    }
                }
                // pass 0: the fixed segments; pass 1: the plastic ones of spikes emitted in this window
                // (emitted at window step kb - n_delay + k >= 0), staged the same way
#pragma unroll
                for (int pass = 0; pass < (PL ? 2 : 1); pass++) {
                    if (PL && pass == 1) {
#pragma unroll
                        for (int r = 0; r < IPT; r++) {
                            seg[r] = (int)a.nnz + pq[PL ? r : 0].x;
                            len[r] = ok[r] && a.kb + kk[r] - nd >= 0 ? pq[PL ? r : 0].y : 0;
                        }
                    }
                    sum = 0;
#pragma unroll
                    for (int r = 0; r < IPT; r++) sum += len[r];
                    // Each wavefront stages its own 64 x IPT segments with their prefix sums (a
                    // shuffle scan) in its own region of the staging arrays and spreads their edges
                    // over its own 64 lanes: no workgroup barrier, a search over 64 x IPT entries.
                    // The region is rewritten by the next chunk only after this wavefront's reads
                    // (LDS operations of a wavefront complete in order).
                    int x = sum;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const int y = __shfl_up(x, o);
                        if (lane >= o) x += y;
                    }
                    const int tot = __shfl(x, 63);
                    int run = x - sum;
                    const int wb = (tid & ~63) * IPT;
#pragma unroll
                    for (int r = 0; r < IPT; r++) {
                        run += len[r];
                        st_seg[tid * IPT + r] = seg[r];
                        st_end[tid * IPT + r] = run;
                        st_k[tid * IPT + r] = kk[r] - k0;  // acc column
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const int ns = 64 * IPT;
                    for (int q0 = lane; q0 < tot; q0 += 64 * UNROLL) {
                        int ek[UNROLL], slot[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) {
                            const int q = q0 + u * 64;
                            ek[u] = -1;
                            slot[u] = 0;
                            if (q < tot) {
                                int l = 0, h = ns - 1;  // first j with st_end[wb + j] > q
                                while (l < h) {
                                    const int mid = (l + h) >> 1;
                                    if (st_end[wb + mid] > q) h = mid;
                                    else l = mid + 1;
                                }
                                ek[u] = st_seg[wb + l] + (q - (l ? st_end[wb + l - 1] : 0));
                                slot[u] = st_k[wb + l];
                            }
                        }
                        int loc[UNROLL], val[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; u++) {  // all UNROLL edge loads in flight
                            edge(ek[u] >= 0 ? ek[u] : ek[0], &loc[u], &val[u]);
                            if (ek[u] < 0) val[u] = 0;
                        }
#pragma unroll
                        for (int u = 0; u < UNROLL; u++)
                            if (val[u] != 0) acc_add((unsigned)loc[u], slot[u], val[u]);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        // plastic arrivals pushed at spike time (all of them, or with plastic pulls only those from
        // earlier windows, which land in a window's first n_delay steps), tile's targets only
        if (a.ptg_ptr && (!PL || a.kb < nd)) {
            const int p0 = a.ptg_ptr[T], np_ = a.ptg_ptr[T + 1] - p0, kn = k1 - k0;
            for (int q = tid; q < np_ * kn; q += TPB) {
                const int kq = q / np_;
                const int i = a.ptg_idx[p0 + (q - kq * np_)];
                int* r = a.ring + ((size_t)b * a.ring_slots + (rs0 + k0 + kq) % a.ring_slots) * a.N + i;
                const int v = *r;
                if (v != 0) {
                    acc_add((unsigned)(i - i0), kq, v);
                    *r = 0;
                }
            }
        }
    };
    // A block runs as one part, or as two (steps [0, kh) and [kh, K)) when its arrivals of all K
    // steps would not leave LDS for enough workgroups per CU: acc then holds one part's columns
    // (and the hit masks, read into registers before the first part's steps).
    const int KH = PARTS == 2 ? a.kh : a.K;
    const int HC = H16 ? (a.K + 1) >> 1 : KH;  // the hit-mask column
    pull(0, KH);

    // Poisson hits of the tile's input neurons for all K steps, as K-bit masks in column HC of
    // acc.  Input neurons are scattered, so hashing inside the step loop would run the hash on
    // every lane of almost every wavefront; instead each wavefront spreads the (neuron, step)
    // pairs of its input neurons over its 64 lanes, in registers: a lane permutation lists the
    // input lanes in order (src[q] = lane of the q-th input neuron) and each pair fetches its
    // neuron's payload from that lane.  No LDS staging and no barrier; the column-HC atomics
    // touch no word the pull adds into.  The draw is flyzero/gpu.py:66-67: lowbias32 of
    // (neuron, step, seed + fly), or the precomputed hit bits.
    if constexpr (H16) {  // hit words start at zero (the fill put the arrival bias there); only
                          // lanes of the owning wavefront set their bits, after this store (a
                          // wavefront's LDS operations complete in order)
#pragma unroll
        for (int p = 0; p < NPT; p++) acc[__umul24((unsigned)(p * TPB + tid), (unsigned)KP) + HC] = 0;
    }
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const bool want = i0 + p * TPB + tid < a.N && j_[p] >= 0 && (a.hits || pp_[p] > 0.f);
        const unsigned long long wm = __ballot(want);
        const int ne = __popcll(wm);  // wavefront-uniform
        if (ne == 0) continue;
        const int below = __popcll(wm & ((1ull << lane) - 1ull));
        // input lanes go to positions 0 .. ne - 1 in lane order, the others after them: a permutation
        const int src = __builtin_amdgcn_ds_permute((want ? below : ne + lane - below) << 2, lane);
        const int payload = a.hits ? j_[p] : __float_as_int(pp_[p]);
        const int lbase = p * TPB + (tid & ~63);
        // pairs x = q * K + k, x = lane, lane + 64, ...: one division, then increments
        int q = (int)div32((unsigned)lane, a.dv_K), k = lane - q * a.K;
        const int dq = (int)div32(64u, a.dv_K), dk = 64 - dq * a.K;
        for (int x0 = 0; x0 < ne * a.K; x0 += 64) {  // uniform trip count: every lane permutes
            const int sl = __builtin_amdgcn_ds_bpermute((q < ne ? q : 0) << 2, src);
            const int pl = __builtin_amdgcn_ds_bpermute(sl << 2, payload);
            if (x0 + lane < ne * a.K) {
                const int lq = lbase + sl;
                bool hit;
                if (a.hits) {
                    hit = (a.hits[((size_t)(a.kb + k) * a.B + b) * a.hw + (pl >> 5)] >> (pl & 31)) & 1u;
                } else {
                    const uint32_t h = mix32((hb * (uint32_t)a.N + (uint32_t)(i0 + lq)) ^ s_sk[k]);
                    hit = (float)(h >> 8) * (1.0f / 16777216.0f) < __int_as_float(pl);
                }
                if (hit) {
                    atomicOr(&acc[__umul24(lq, KP) + HC], 1 << k);
                    atomicOr(&s_wh[(tid >> 6) * NPT + p], 1u << k);
                }
            }
            q += dq;
            k += dk;
            if (k >= a.K) {
                k -= a.K;
                q++;
            }
        }
    }
    __syncthreads();

    // K steps of flyzero/brain.py:37-53 from LDS, NPT neurons per thread, without branches:
    // every select keeps the float operations and their order exactly as written there.
    // When n_ref >= K (the model's 2.2 ms refractory period against a block of at most
    // n_delay = 1.8 ms) a neuron fires at most once per block, and the refractory counter is
    // carried as the block step at which it runs out, rend: the counter before step k is
    // max(rend - k, 0), so "r > 0" is k < rend, "r <= 0 after the decrement" is k + 1 >= rend,
    // and a spike at step k sets rend = k + 1 + n_ref (strictly larger than before, so the
    // spike and its step are read back from rend after the loop, without a per-step mask).
    // State stays in registers across the parts of a block.
    const bool once = a.c.n_ref >= a.K;
    uint32_t masks[NPT], hm[NPT];
    float uu[NPT], gg[NPT];
    int rr[NPT];  // once: the block step at which the refractory period ends; else the counter
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        masks[p] = hm[p] = 0;
        uu[p] = u_[p];
        gg[p] = g_[p];
        rr[p] = r_[p];
    }
    auto steps = [&](int k0, int k1) {  // block steps [k0, k1) from acc columns k - k0
        const int kn = k1 - k0;
#pragma unroll
        for (int p = 0; p < NPT; p++) {
            const int li = p * TPB + tid, i = i0 + li;
            const int* row = acc + __umul24(li, KP);
            if (i < a.N) {
                if (k0 == 0) hm[p] = (uint32_t)row[HC];
                const uint32_t hits = hm[p] >> k0;
                float u = uu[p], g = gg[p];
                const float bias = j_[p] >= 0 ? bias_[p] : 0.f;
                if (once) {
                    const int r0 = rr[p] - k0;  // relative to the part's first step
                    int rend;
                    // the block lengths of dt = 0.1 ms and 0.25 ms (n_delay 18 and 7), and the parts
                    // of the former, unrolled
                    // (a wavefront whose neurons have no Poisson hit in these steps takes the
                    // unrolled loop without the kick; the others skip the steps without a hit)
                    const uint32_t wh = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_wh[(tid >> 6) * NPT + p]) >> k0;
                    const bool kick = (kn < 32 ? wh & ((1u << kn) - 1u) : wh) != 0u;
                    if (PARTS == 1 && kn == 18)
                        rend = kick ? once_steps<18, true, H16>(row, kn, u, g, r0, bias, hits, wh, a.c)
                                    : once_steps<18, false, H16>(row, kn, u, g, r0, bias, hits, wh, a.c);
                    else if (PARTS == 1 && kn == 7)
                        rend = kick ? once_steps<7, true, H16>(row, kn, u, g, r0, bias, hits, wh, a.c)
                                    : once_steps<7, false, H16>(row, kn, u, g, r0, bias, hits, wh, a.c);
                    else if (PARTS == 2 && kn == 9)
                        rend = kick ? once_steps<9, true, H16>(row, kn, u, g, r0, bias, hits, wh, a.c)
                                    : once_steps<9, false, H16>(row, kn, u, g, r0, bias, hits, wh, a.c);
                    else rend = once_steps<0, true, H16>(row, kn, u, g, r0, bias, hits, wh, a.c);
                    if (rend != r0) masks[p] |= 1u << (rend - 1 - a.c.n_ref + k0);
                    rr[p] = rend + k0;
                } else {
                    int r = rr[p];
                    uint32_t mask = 0;
                    for (int k = 0; k < kn; k++) {
                        const float inj = (float)arrivals<H16>(row, k) * a.c.in_scale;
                        g = g * a.c.decay_syn + inj;
                        const bool refr = r > 0;
                        const float u_int = u + (g + bias - u) * a.c.decay_m;
                        u = refr ? u : u_int;
                        r = refr ? r - 1 : r;
                        const float u_kick = u + a.c.kick;
                        u = ((hits >> k) & 1u) ? u_kick : u;
                        const bool sp = u >= a.c.u_th && r <= 0;
                        u = sp ? a.c.u_reset : u;
                        r = sp ? a.c.n_ref : r;
                        mask |= (uint32_t)sp << k;
                    }
                    masks[p] |= mask << k0;
                    rr[p] = r;
                }
                uu[p] = u;
                gg[p] = g;
            }
        }
    };
    steps(0, KH);
    if constexpr (PARTS == 2) {
        if (KH < a.K) {  // the second part: acc cleared and refilled with the arrivals of [KH, K)
            __syncthreads();
            clear_acc(acc, KP, NPT, TPB, tid);
            __syncthreads();
            pull(KH, a.K);
            __syncthreads();
            steps(KH, a.K);
        }
    }
    // Spikes per block step (LDS), then one global atomic per step reserves their places in the
    // per-step lists; its latency overlaps the state stores that follow.
#pragma unroll
    for (int p = 0; p < NPT; p++)
        for (uint32_t m = masks[p]; m; m &= m - 1) atomicAdd(&s_lcnt[__ffs(m) - 1], 1);
    __syncthreads();
    int lpos = 0;
    if (tid < a.K && s_lcnt[tid] > 0) lpos = atomicAdd(&cnt_b[wrap2(lt0 + tid, M)], s_lcnt[tid]);
    // state and counts (a window's later blocks add theirs with no-return atomics: no load waits)
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const int i = i0 + p * TPB + tid;
        if (i < a.N) {
            const int r = once ? rr[p] - min(max(rr[p], 0), a.K) : rr[p];  // the counter after the block
            if (uu[p] != u_[p]) a.u[bn + i] = uu[p];
            if (gg[p] != g_[p]) a.g[bn + i] = gg[p];
            if (r != r_[p]) a.ref[bn + i] = (signed char)r;
            const int ns = __popc(masks[p]);
            if (a.first) a.counts[bn + i] = ns;
            else if (ns) atomicAdd(&a.counts[bn + i], ns);
        }
    }

    if (a.raster) {
#pragma unroll
        for (int p = 0; p < NPT; p++) {
            const int wbase = (i0 + p * TPB + (tid & ~63)) >> 5;
            for (int k = 0; k < a.K; k++) {
                const unsigned long long bal = __ballot((masks[p] >> k) & 1u);
                unsigned* row = a.raster + ((size_t)(a.kb + k) * a.B + b) * a.rw;
                if (lane == 0 && wbase < a.rw) row[wbase] = (unsigned)bal;
                if (lane == 32 && wbase + 1 < a.rw) row[wbase + 1] = (unsigned)(bal >> 32);
            }
        }
    }

    // append the spikes to the per-step lists at the reserved places, then push the spikers'
    // plastic rows n_delay steps ahead
    if (tid < a.K) s_lcnt[tid] = lpos;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NPT; p++) {
        const int i = i0 + p * TPB + tid;
        for (uint32_t m = masks[p]; m; m &= m - 1) {
            const int k = __ffs(m) - 1;
            const int pos = atomicAdd(&s_lcnt[k], 1);
            if (pos < a.N) idx_b[(size_t)wrap2(lt0 + k, M) * a.N + pos] = i;
        }
    }
    if (a.pl_indptr) {
        // Plastic rows of the block's spikes go n_delay steps ahead.  A short row is pushed by its
        // own thread; longer rows are pushed by the whole workgroup: staged with their prefix
        // sums (in acc, free after the step loop) and the edges split evenly over the threads by
        // a binary search, so a long row costs no more than many short ones.  Rounds take at
        // most one spike per neuron (one round when neurons fire once per block).
        int* p_beg = acc;           // [TS] first plastic edge of the staged row (thread-major)
        int* p_end = acc + TS;      // [TS] inclusive prefix of row lengths
        int* p_k = acc + 2 * TS;    // [TS] block step of the spike
        const float* pwb = a.pw + (size_t)b * a.n_pl;
        // with plastic pulls only the spikes whose arrival (block step k + n_delay) lies past this
        // window are pushed: block steps k >= win - kb - n_delay
        uint32_t cross = 0xffffffffu;
        if constexpr (PL) {
            const int first = a.win - a.kb - nd;
            cross = first <= 0 ? 0xffffffffu : (first >= 32 ? 0u : ~((1u << first) - 1u));
        }
        uint32_t rem[NPT];
        bool more = false;
#pragma unroll
        for (int p = 0; p < NPT; p++) {
            rem[p] = 0;
            const int i = i0 + p * TPB + tid;
            const uint32_t pm = masks[p] & cross;
            if (pm == 0) continue;
            const int q0 = a.pl_indptr[i], q1 = a.pl_indptr[i + 1];
            if (q1 - q0 > PL_SERIAL) {
                rem[p] = pm;
                more = true;
                continue;
            }
            for (uint32_t m = pm; m; m &= m - 1) {
                int* r = a.ring + ((size_t)b * a.ring_slots + (rs0 + __ffs(m) - 1 + nd) % a.ring_slots) * a.N;
                for (int q = q0; q < q1; q++) {
                    const int val = (int)rintf(pwb[q] * a.pw_qscale);
                    if (val != 0) atomicAdd(r + a.pl_post[q], val);
                }
            }
        }
        while (__syncthreads_or(more)) {  // also: the previous round's staged rows are read
            int beg[NPT], len[NPT], kk[NPT];
            int sum = 0;
#pragma unroll
            for (int p = 0; p < NPT; p++) {
                beg[p] = len[p] = kk[p] = 0;
                if (rem[p]) {
                    const int i = i0 + p * TPB + tid;
                    kk[p] = __ffs(rem[p]) - 1;
                    rem[p] &= rem[p] - 1;
                    beg[p] = a.pl_indptr[i];
                    len[p] = a.pl_indptr[i + 1] - beg[p];
                }
                sum += len[p];
            }
            int tot;
            int run = block_incl_scan(sum, s_wsum, &tot) - sum;
#pragma unroll
            for (int p = 0; p < NPT; p++) {  // thread-major, the order of the prefix sum
                run += len[p];
                p_beg[tid * NPT + p] = beg[p];
                p_end[tid * NPT + p] = run;
                p_k[tid * NPT + p] = kk[p];
            }
            __syncthreads();
            for (int q = tid; q < tot; q += TPB) {
                int lo = 0, hi = TS - 1;  // first staged row with p_end > q
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (p_end[mid] > q) hi = mid;
                    else lo = mid + 1;
                }
                const int j = p_beg[lo] + (q - (lo ? p_end[lo - 1] : 0));
                const int val = (int)rintf(pwb[j] * a.pw_qscale);
                if (val != 0) {
                    int* r = a.ring + ((size_t)b * a.ring_slots + (rs0 + p_k[lo] + nd) % a.ring_slots) * a.N;
                    atomicAdd(r + a.pl_post[j], val);
                }
            }
            more = false;
#pragma unroll
            for (int p = 0; p < NPT; p++) more = more || rem[p] != 0;
        }
    }
}

// Bucket prologue of a tile-engine block (tile source 1): the spikes whose arrivals fall in
// the block [t0, t0 + K) (emitted at steps [t0 - n_delay, t0 - n_delay + K)) are expanded once
// into per-tile buckets of (first edge, length << 5 | step in block), segments longer than
// BK_PIECE edges split into pieces, so that each (fly, tile)
// workgroup reads only the segments that reach it, contiguously, instead of looking every
// spike of the fly up in the segment table.  Grid (ceil(n_tiles / 64), B) workgroups, one per
// (tile group g, fly): lane t of every wavefront owns tile 64 g + t; the wavefronts split the
// spikes, staged through LDS in chunks; positions come from LDS counters, so a bucket's order
// depends on timing while its contents (and the integer sums made from them) do not.  A group
// of at most 32 tiles (the last one, when n_tiles is not a multiple of 64) would leave most
// lanes idle in every pass, so there lane so * tw + t takes tile t of the so-th of 64 / tw
// spikes (tw = the group's tile count rounded up to a power of two): at 136 tiles the third
// group's 8 tiles cost an eighth of a full group instead of as much.
constexpr int BK_CHUNK = 2048;
constexpr int BK_UNROLL = 8;         // spikes per batch: table loads in flight per lane
constexpr int BK_THREADS = 1024;
constexpr int BK_STAGE = 256;        // entries staged in LDS per tile between flushes
constexpr int BK_SYNC = 4;           // rounds between the barriers that may flush the stages
constexpr int BK_SP = BK_STAGE + 1;  // odd row stride: the 64 lanes' rows start in different banks

struct BucketArgs {
    const long long* clock;
    int N, B, K, kb, blk, nd, M, n_tiles;
    int kh;           // entries of steps < kh go to a bucket's front, the others to its back
    const int* spk_idx;
    const int* spk_cnt;
    const int* tb;
    const uint16_t* tb16;
    const int2* ptp;  // plastic (first edge, length) pairs, or null: no plastic entries
    long long nnz;
    const int* indptr;
    const long long* bk_off;
    long long bk_stride;
    int* bk;
    int* bk_cnt;
    const unsigned char* alive;
    int* err;
    int b0;  // this launch covers flies [b0, b0 + gridDim.y)
    Div32 dv_M, dv_grp;  // dividers by M and by the tile-group count (gridDim.x)
};

// T16: read the uint16 copy of the segment table (offsets from the row start).  At most 64
// VGPRs (8 wavefronts per SIMD), so that two 1024-thread workgroups share a CU.
//
// Entries are staged per tile in LDS and written out in flushes, one tile per wavefront with
// the lanes on consecutive entries: a tile gains only a few entries per batch, so writing
// them straight to memory would cost one memory transaction per 4-byte entry.  A flush runs
// at the barrier after every BK_SYNC rounds (every wavefront one batch per round) when some
// stage has passed its mark; entries past a full stage go straight to memory.
// PL: also the plastic segments (entries nnz + first plastic edge) of the spikes emitted in this
// window, which lif_tile pulls with the fly's own weights.
template <bool T16, bool PL>
__global__ void __launch_bounds__(BK_THREADS, 8) tile_bucket(BucketArgs a) {
    // spikes per batch: with plastic pieces fewer, so that the kernel keeps to its 64 VGPRs
    // without spilling
    constexpr int BUS = PL ? BK_UNROLL / 2 : BK_UNROLL;
    __shared__ int s_off[KMAX + 1];
    __shared__ int s_cnt[64];          // entries written to the front of each tile's bucket
    __shared__ int s_cntb[64];         // entries written to its back
    __shared__ int s_fill[64];         // entries reserved in each lane's stage since the last flush
    __shared__ int s_cap[64];          // each tile's bucket capacity
    __shared__ long long s_base[64];   // each tile's bucket in the arena
    __shared__ int s_flush;
    __shared__ int s_id[BK_CHUNK];
    __shared__ int s_stage[64 * BK_SP];  // one stage row per lane (the lane's tile: lane & (tw - 1))
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, W = blockDim.x >> 6;
    // the tile groups of one fly run on one XCD: they read the same spike lists, and the
    // table-row spans of neighbouring groups share cache lines
    int grp, b;
    xcd_order(a.dv_grp, gridDim.y, &grp, &b);
    b += a.b0;
    // tiles in this group (1 .. 64); lanes per spike tw (64 unless the group has at most 32
    // tiles: then its count rounded up to a power of two) and spikes per lane pass spl = 64 / tw
    const int w = a.n_tiles - grp * 64 < 64 ? a.n_tiles - grp * 64 : 64;
    int lsh = 6;
    while (lsh > 0 && (1 << (lsh - 1)) >= w) lsh--;
    const int tw = 1 << lsh, spl = 64 >> lsh;
    const int tl = lane & (tw - 1), so = lane >> lsh;  // this lane's tile in the group, its spike slot
    const int T = grp * 64 + tl;
    if (a.alive && !a.alive[b]) return;
    const long long t0 = a.clock[0] + a.kb;
    const int M = a.M, nd = a.nd;
    const int lt0 = step_slot(t0, a.dv_M);
    const int* cnt_b = a.spk_cnt + (size_t)b * M;
    const int* idx_b = a.spk_idx + (size_t)b * M * a.N;
    // buffers alternate with the block parity, so the bucket of block j can be written while
    // the tile kernel of block j - 1 reads the other one
    const int par = (int)((a.clock[2] + a.blk) & 1);
    const bool own = tl < w;
    if (tid < 64) {  // per tile of the group (tile index tid)
        long long base = 0, cap = 0;
        if (tid < w) {
            const int Ti = grp * 64 + tid;
            base = ((long long)par * a.B + b) * a.bk_stride + a.bk_off[Ti];
            cap = a.bk_off[Ti + 1] - a.bk_off[Ti];
        }
        s_base[tid] = base;
        s_cap[tid] = cap < 0x7fffffff ? (int)cap : 0x7fffffff;
        s_cnt[tid] = 0;
        s_cntb[tid] = 0;
        s_fill[tid] = 0;
    }
    if (tid == 0) s_flush = 0;
    source_offsets(cnt_b, t0, lt0, a.K, nd, M, a.N, s_off);
    const int total = s_off[a.K];
    int* const dst = a.bk + s_base[tl];  // this lane's bucket: entry p at dst[p]
    const int capi = s_cap[tl];
    const int nt1 = a.n_tiles + 1;
    const int Tc = own ? T : 0;
    // every thread, between barriers: stage rows to memory (entries of steps < kh from the front
    // of the bucket up, the others from its back down, positions by ballots; a block of one part
    // has only front entries, copied as they are), counters advanced.  With one row per tile
    // (tw = 64) nothing else moves a tile's counters during a flush; with several, each row
    // reserves its run of the tile's front and back with LDS atomics first.
    const bool split = a.kh < a.K;
    auto flush = [&]() {
        for (int r = wave; r < 64; r += W) {
            const int n = s_fill[r] < BK_STAGE ? s_fill[r] : BK_STAGE;
            if (n == 0) continue;  // wavefront-uniform
            const int t = r & (tw - 1), cp = s_cap[t];
            const int* row = s_stage + r * BK_SP;
            int* d = a.bk + s_base[t];
            if (!split) {
                int f0 = s_cnt[t];
                if (spl > 1) {
                    int g0 = 0;
                    if (lane == 0) g0 = atomicAdd(&s_cnt[t], n);
                    f0 = __shfl(g0, 0);
                }
                for (int j = lane; j < n; j += 64)
                    if (f0 + j < cp) d[f0 + j] = row[j];
                if (spl == 1 && lane == 0) s_cnt[t] = f0 + n;
                continue;
            }
            int f0 = s_cnt[t], b0 = s_cntb[t];
            if (spl > 1) {
                int nfr = 0;  // front entries of the row (wavefront-uniform)
                for (int j0 = 0; j0 < n; j0 += 64) {
                    const int j = j0 + lane;
                    nfr += __popcll(__ballot(j < n && (row[j < n ? j : 0] & 31) < a.kh));
                }
                int g0 = 0, g1 = 0;
                if (lane == 0) {
                    g0 = atomicAdd(&s_cnt[t], nfr);
                    g1 = atomicAdd(&s_cntb[t], n - nfr);
                }
                f0 = __shfl(g0, 0);
                b0 = __shfl(g1, 0);
            }
            int nf = 0, nbk = 0;  // wavefront-uniform
            for (int j0 = 0; j0 < n; j0 += 64) {
                const int j = j0 + lane;
                const int x = j < n ? row[j] : 0;
                const bool back = j < n && (x & 31) >= a.kh, front = j < n && !back;
                const unsigned long long mf = __ballot(front), mb = __ballot(back);
                const unsigned long long below = (1ull << lane) - 1ull;
                if (front) {
                    const int pos = f0 + nf + __popcll(mf & below);
                    if (pos < cp) d[pos] = x;
                }
                if (back) {
                    const int pos = cp - 1 - (b0 + nbk + __popcll(mb & below));
                    if (pos >= 0) d[pos] = x;
                }
                nf += __popcll(mf);
                nbk += __popcll(mb);
            }
            if (spl == 1 && lane == 0) {
                s_cnt[t] = f0 + nf;
                s_cntb[t] = b0 + nbk;
            }
        }
        __syncthreads();
        if (tid < 64) s_fill[tid] = 0;
        if (tid == 0) s_flush = 0;
        __syncthreads();
    };
    for (int c0 = 0; c0 < total; c0 += BK_CHUNK) {
        const int nc = total - c0 < BK_CHUNK ? total - c0 : BK_CHUNK;
        for (int q = tid; q < nc; q += blockDim.x) {  // stage the chunk: id << 5 | step, or -1
            const int e = c0 + q;
            const int k = entry_step(s_off, a.K, e);
            const int id = idx_b[(size_t)wrap2(lt0 + k - nd + M, M) * a.N + (e - s_off[k])];  // k < K <= nd
            s_id[q] = (unsigned)id < (unsigned)a.N ? id << 5 | k : -1;  // id < 2^22 (table < 2^31 entries)
        }
        __syncthreads();
        // One batch of BU spikes per lane: spikes q0 + u (every lane of the wavefront the same
        // spike) when a lane pass covers one spike (SCALAR), else q0 + u * spl + so.
        auto batch = [&](const int q0, auto scalar) {
            constexpr bool SCALAR = decltype(scalar)::value;
            constexpr int BU = SCALAR ? BUS : BUS / 2;  // per-lane ids take more registers
            {
                // SCALAR: the BU spikes are the same for every lane of the wavefront (lane = tile):
                // their ids are made scalar, so the row addresses are scalar arithmetic and each
                // lane's table reads are one scalar base plus its fixed tile offset, all in flight
                // together.  Past the chunk end a lane rereads a valid spike and masks the result.
                int lo[BU], len[BU], kk[BU], id[BU];
                bool use[BU];
#pragma unroll
                for (int u = 0; u < BU; u++) {
                    const int q = SCALAR ? q0 + u : q0 + u * spl + so;
                    const int sv = s_id[q < nc ? q : q0];
                    const int v = SCALAR ? __builtin_amdgcn_readfirstlane(sv) : sv;
                    use[u] = own && q < nc && v >= 0;
                    id[u] = v >= 0 ? v >> 5 : 0;
                    kk[u] = v & 31;
                }
                int l16[BU], h16[BU], r0_[BU];
                int2 pq[PL ? BU : 1];
#pragma unroll
                for (int u = 0; u < BU; u++) {  // every load first ...
                    if constexpr (PL) pq[u] = a.ptp[(unsigned)id[u] * (unsigned)a.n_tiles + (unsigned)Tc];
                    if constexpr (T16) {
                        const uint16_t* row = a.tb16 + (unsigned)id[u] * (unsigned)nt1;  // < 2^31 (checked)
                        r0_[u] = a.indptr[id[u]];
                        l16[u] = row[Tc];
                        h16[u] = row[Tc + 1];
                    } else {
                        const int* row = a.tb + (unsigned)id[u] * (unsigned)nt1;
                        r0_[u] = 0;
                        l16[u] = row[Tc];
                        h16[u] = row[Tc + 1];
                    }
                }
#pragma unroll
                for (int u = 0; u < BU; u++) {  // ... then the arithmetic
                    lo[u] = r0_[u] + l16[u];
                    len[u] = use[u] && h16[u] > l16[u] ? h16[u] - l16[u] : 0;
                }
                // pieces of at most BK_PIECE edges (almost every segment is one piece): one stage
                // reservation per lane for the batch; what does not fit takes bucket positions now
                int tot = 0;
#pragma unroll
                for (int u = 0; u < BU; u++) tot += (int)(((unsigned)len[u] + (BK_PIECE - 1)) / BK_PIECE);
                int plo[PL ? BU : 1], plen[PL ? BU : 1];
                if constexpr (PL) {
#pragma unroll
                    for (int u = 0; u < BU; u++) {  // emitted at window step kb - n_delay + k
                        plo[u] = (int)a.nnz + pq[u].x;
                        plen[u] = use[u] && a.kb + kk[u] - nd >= 0 ? pq[u].y : 0;
                        tot += (int)(((unsigned)plen[u] + (BK_PIECE - 1)) / BK_PIECE);
                    }
                }
                if (tot > 0) {
                    // each stage row gains at most one batch of one wavefront per round
                    int slot = atomicAdd(&s_fill[lane], tot);
                    if (slot + tot > BK_STAGE - W * BUS) s_flush = 1;
                    // past a full stage: straight to memory (one front reservation for the batch when
                    // the block has one part, else one per entry and part)
                    const int over = slot + tot - BK_STAGE;
                    int g = !split && over > 0 ? atomicAdd(&s_cnt[tl], over < tot ? over : tot) : 0;
                    auto put = [&](int entry) {
                        if (slot < BK_STAGE) {
                            s_stage[lane * BK_SP + slot] = entry;
                        } else if (!split) {
                            if (g < capi) dst[g] = entry;  // past the capacity: dropped, flagged below
                            g++;
                        } else if ((entry & 31) < a.kh) {
                            const int gf = atomicAdd(&s_cnt[tl], 1);
                            if (gf < capi) dst[gf] = entry;
                        } else {
                            const int gb = capi - 1 - atomicAdd(&s_cntb[tl], 1);
                            if (gb >= 0) dst[gb] = entry;
                        }
                        slot++;
                    };
                    auto pieces = [&](int first, int n, int k) {
                        put((first << 7) | (((n < BK_PIECE ? n : BK_PIECE) - 1) << 5) | k);
                        for (int c = BK_PIECE; c < n; c += BK_PIECE)  // rare: longer segments
                            put(((first + c) << 7) | (((n - c < BK_PIECE ? n - c : BK_PIECE) - 1) << 5) | k);
                    };
#pragma unroll
                    for (int u = 0; u < BU; u++)
                        if (len[u] > 0) pieces(lo[u], len[u], kk[u]);
                    if constexpr (PL) {
#pragma unroll
                        for (int u = 0; u < BU; u++)
                            if (plen[u] > 0) pieces(plo[u], plen[u], kk[u]);
                    }
                }
            }
        };
        // A barrier (and a flush if a stage passed its mark) after every BK_SYNC rounds and
        // after the chunk's last round, so that the wavefronts' rounds between them overlap
        // instead of each round waiting for the slowest wavefront.  Between barriers a stage row
        // can gain more than its room; such entries take bucket positions at once (put).
        const int bu = spl == 1 ? BUS : BUS / 2;  // spikes per lane and round
        for (int r0 = 0, rr = 1; r0 < nc; r0 += W * bu * spl, rr++) {
            const int q0 = r0 + wave * bu * spl;
            if (q0 < nc) {
                if (spl == 1) batch(q0, std::true_type{});
                else batch(q0, std::false_type{});
            }
            if (rr % BK_SYNC == 0 || r0 + W * bu * spl >= nc) {  // uniform
                __syncthreads();
                if (s_flush) flush();
            }
        }
    }
    flush();
    if (tid < w) {  // per tile of the group (tile index tid)
        int c = s_cnt[tid], cb = s_cntb[tid];
        const int cp = s_cap[tid];
        if (c + cb > cp) {  // cannot happen: the capacities bound every block
            atomicOr(a.err, 2);
            c = c < cp ? c : cp;
            cb = cb < cp - c ? cb : cp - c;
        }
        reinterpret_cast<int2*>(a.bk_cnt)[((size_t)par * a.B + b) * a.n_tiles + grp * 64 + tid] = make_int2(c, cb);
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

__global__ void zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

// Device scratch of fz_run_static: the push engine's spike work list, carved per call out of
// this fixed array instead of memory allocated by fz_bind.  Each GPU the library runs on holds
// its own copy for the life of the process.  In host code the name is only a stand-in for the
// symbol: its device address (per device) comes from hipGetSymbolAddress at fz_create.
static __device__ __align__(256) unsigned char scratch[FZ_STATIC_SCRATCH_BYTES];

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
    fz_state st{};
    bool bound = false;
    WorkList wl{};
    int wl_B = 0;  // flies the counters are sized for
    int* err = nullptr;
    int kmax = 1;
    int n_cu = 256;
    size_t lds_max = 65536;
    // tile engine (mode 2)
    int mode = 1;
    int ts = 0, n_tiles = 0;
    int* pk = nullptr;  // packed fixed edges (null: the kernel reads col and w)
    int lbits = 0;
    long long in_max = -1;  // largest summed |packed weight| into one neuron (-1: not known)
    // tile buckets (tile source 1): per-tile offsets in a fly's arena (entries), arena, counts
    uint16_t* tb16 = nullptr;  // segment table as uint16 offsets from each row's start
    int2* tp = nullptr;        // segment table as (first edge, length) pairs, when it fits TP_MAX_BYTES
    int2* ptp = nullptr;       // the plastic CSR's (first edge, length) per (row, tile): plastic pulls
    int bk4 = 0;
    std::vector<long long> bk_off_h;
    long long* bk_off = nullptr;
    long long bk_stride = 0;
    int* bk = nullptr;
    int* bk_cnt = nullptr;
    int bk_B = 0;
    const int* tb = nullptr;
    const int* ptg_ptr = nullptr;
    const int* ptg_idx = nullptr;
    // fz_run_static: the device the engine was created on and that device's scratch address
    int dev = 0;
    unsigned char* scr = nullptr;
};

struct fz_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

struct Shape {
    int ut, pt, chunk, grid, K, ipt, src;
    int pl;     // plastic pulls in effect (requested, and the engine has the plastic pair table)
    int parts;  // tile engine: a block runs as 1 part or as 2 (half the arrival columns in LDS)
    int h16;    // tile engine: two block steps' arrival sums per LDS word
};

// Steps of the first part of a block of K steps.
int part_kh(const Shape& s, int K) { return s.parts == 2 && K >= 2 ? (K + 1) / 2 : K; }

int resolve(const fz_engine* e, const fz_tune* t, Shape* s) {
    s->ut = (t && t->upd_threads) ? t->upd_threads : 256;
    s->pt = (t && t->push_threads) ? t->push_threads : 256;
    s->chunk = (t && t->push_chunk) ? t->push_chunk : 256;
    s->grid = (t && t->push_grid) ? t->push_grid : e->n_cu * 8;
    s->K = (t && t->block_steps) ? t->block_steps : e->kmax;
    s->ipt = (t && t->tile_stage) ? t->tile_stage : 1;
    if (s->ipt != 1 && s->ipt != 4) return fail(FZ_E_ARG, "tile_stage %d: must be 1 or 4", s->ipt);
    s->src = t ? t->tile_source : 0;
    s->pl = 0;
    s->parts = (t && t->block_parts) ? t->block_parts : 1;
    if (s->parts != 1 && s->parts != 2) return fail(FZ_E_ARG, "block_parts %d: must be 1 or 2", s->parts);
    if (!(e->mode == 2 && e->pk)) s->parts = 1;  // two parts are compiled for packed edges
    if (t && t->plastic_pull != 0 && t->plastic_pull != 1)
        return fail(FZ_E_ARG, "plastic_pull %d: must be 0 or 1", t->plastic_pull);
    if (t && t->plastic_pull && e->mode == 2 && e->n_pl > 0 && e->ptp && e->pk) s->pl = 1;
    if (s->src < 0 || s->src > 2) return fail(FZ_E_ARG, "tile_source %d: must be 0, 1 or 2", s->src);
    if (s->src == 1 && (e->mode != 2 || !e->bk_off)) return fail(FZ_E_ARG, "tile_source 1 needs the tile engine");
    if (s->src == 1 && !e->bk4)
        return fail(FZ_E_ARG, "tile_source 1 needs fewer than 2^24 fixed edges (bucket entries hold a 24-bit edge "
                    "index); use tile_source 0");
    const int ab = t ? t->acc_bits : 0;
    if (ab != 0 && ab != 16 && ab != 32) return fail(FZ_E_ARG, "acc_bits %d: must be 0, 16 or 32", ab);
    const bool h16_ok = e->mode == 2 && e->pk && s->src == 1 && s->parts == 1 && !s->pl && e->n_pl == 0 &&
                        e->in_max >= 0 && e->in_max < 32768;
    if (ab == 16 && !h16_ok)
        return fail(FZ_E_ARG, "acc_bits 16 needs the tile engine with packed edges, tile_source 1, one-part blocks, no "
                    "plastic synapses and every neuron's summed |input weight| below 32768 (%lld)", e->in_max);
    s->h16 = (ab == 16 || (ab == 0 && h16_ok)) ? 1 : 0;
    if (e->mode == 2) {
        if (!(t && t->upd_threads)) s->ut = e->ts < 256 ? e->ts : 256;
        const int npt = s->ut > 0 ? e->ts / s->ut : 0;
        if (s->ut <= 0 || e->ts % s->ut || (npt != 1 && npt != 2 && npt != 4 && npt != 8))
            return fail(FZ_E_ARG, "tile engine: upd_threads %d must divide the tile (%d neurons) into 1, 2, 4 or 8",
                        s->ut, e->ts);
        // the first part's arrival columns and the hit-mask column per neuron, plus the
        // segment-table staging
        const size_t lds = (size_t)(part_kh(*s, s->K) + 1) * e->ts * 4 + (s->src != 1 ? (size_t)3 * s->ipt * s->ut * 4 : 0);
        if (lds > e->lds_max - TILE_STATIC_LDS)
            return fail(FZ_E_ARG, "tile engine: %zu B of LDS for %d steps x %d neurons exceeds %zu", lds, s->K,
                        e->ts, e->lds_max);
    }
    if (s->ut < 64 || s->ut > MAX_THREADS || s->ut % 64)
        return fail(FZ_E_ARG, "upd_threads %d: need a multiple of 64 in [64, %d]", s->ut, MAX_THREADS);
    if (s->pt < 64 || s->pt > MAX_THREADS || s->pt % 64)
        return fail(FZ_E_ARG, "push_threads %d: need a multiple of 64 in [64, %d]", s->pt, MAX_THREADS);
    if (s->chunk < QMIN || s->chunk > (1 << 20))
        return fail(FZ_E_ARG, "push_chunk %d outside [%d, 2^20]", s->chunk, QMIN);
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

UpdArgs make_upd(const fz_engine* e, const fz_inputs* in, const Shape& s, int K, int kb, int blk, int first,
                 const unsigned* hits, unsigned* raster, const WorkList& wl) {
    UpdArgs a{};
    a.clock = e->st.clock;
    a.N = e->n;
    a.B = e->st.B;
    a.K = K;
    a.kb = kb;
    a.blk = blk;
    a.first = first;
    a.chunk = s.chunk;
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
    a.fly_base = in->fly_base;
    a.alive = in->alive;
    a.indptr = e->indptr;
    a.pl_indptr = e->n_pl > 0 ? e->pl_indptr : nullptr;
    a.wl = wl;
    a.err = e->err;
    a.c = e->c;
    return a;
}

template <typename WT>
PushArgs<WT> make_push(const fz_engine* e, const fz_inputs* in, const Shape& s, int blk, const WorkList& wl) {
    PushArgs<WT> a{};
    a.clock = e->st.clock;
    a.N = e->n;
    a.B = e->st.B;
    a.n_delay = e->c.n_delay;
    a.blk = blk;
    a.chunk = s.chunk;
    a.frac_bits = e->c.frac_bits;
    a.col = e->col;
    a.w = static_cast<const WT*>(e->w);
    a.pl_post = e->n_pl > 0 ? e->pl_post : nullptr;
    a.pw = e->n_pl > 0 ? in->pw : nullptr;
    a.n_pl = e->n_pl;
    a.pw_qscale = e->c.pw_qscale;
    a.ring = e->st.ring;
    a.wl = wl;
    return a;
}

dim3 push_grid(const fz_engine* e, const Shape& s) {
    int gx = (s.grid + e->st.B - 1) / e->st.B;
    if (gx < 1) gx = 1;
    return dim3(gx, e->st.B);
}

TileArgs make_tile(const fz_engine* e, const fz_inputs* in, int K, int KP, int kb, int blk, int first, int win,
                   const unsigned* hits, unsigned* raster) {
    TileArgs a{};
    a.win = win;
    a.nnz = e->nnz;
    a.ptp = nullptr;  // set by the caller when the shape pulls plastic synapses
    a.blk = blk;
    a.clock = e->st.clock;
    a.N = e->n;
    a.B = e->st.B;
    a.b0 = 0;
    a.nb = e->st.B;
    a.K = K;
    a.KP = KP;
    a.kh = K;  // one part; the caller sets the shape's
    a.kb = kb;
    a.first = first;
    a.TS = e->ts;
    a.n_tiles = e->n_tiles;
    a.nd = e->c.n_delay;
    a.M = 3 * e->c.n_delay;
    a.ring_slots = e->st.ring_slots;
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
    a.fly_base = in->fly_base;
    a.alive = in->alive;
    a.col = e->col;
    a.w = e->w;
    a.frac_bits = e->c.frac_bits;
    a.pk = e->pk;
    a.lbits = e->lbits;
    a.tb = e->tb;
    a.tb16 = e->tb16;
    a.tp = e->tp;
    a.indptr = e->indptr;
    a.bk = e->bk;
    a.bk_off = e->bk_off;
    a.bk_stride = e->bk_stride;
    a.bk_cnt = e->bk_cnt;
    a.pl_indptr = e->n_pl > 0 ? e->pl_indptr : nullptr;
    a.pl_post = e->n_pl > 0 ? e->pl_post : nullptr;
    a.pw = e->n_pl > 0 ? in->pw : nullptr;
    a.n_pl = e->n_pl;
    a.pw_qscale = e->c.pw_qscale;
    a.ptg_ptr = e->n_pl > 0 ? e->ptg_ptr : nullptr;
    a.ptg_idx = e->n_pl > 0 ? e->ptg_idx : nullptr;
    a.spk_idx = e->st.spk_idx;
    a.spk_cnt = e->st.spk_cnt;
    a.dv_tiles = div32_for((unsigned)(e->n_tiles > 0 ? e->n_tiles : 1));
    a.dv_M = div32_for((unsigned)a.M);
    a.dv_ring = div32_for((unsigned)(a.ring_slots > 0 ? a.ring_slots : 1));
    a.dv_K = div32_for((unsigned)K);
    a.c = e->c;
    return a;
}

template <typename WT, int IPT, bool BK, bool PL, int PARTS>
void* tile_kernel_npt(int npt) {
    switch (npt) {
        case 1: return reinterpret_cast<void*>(&lif_tile<WT, 1, IPT, BK, PL, PARTS>);
        case 2: return reinterpret_cast<void*>(&lif_tile<WT, 2, IPT, BK, PL, PARTS>);
        case 4: return reinterpret_cast<void*>(&lif_tile<WT, 4, IPT, BK, PL, PARTS>);
        default: return reinterpret_cast<void*>(&lif_tile<WT, 8, IPT, BK, PL, PARTS>);
    }
}

// (the staging depth only shapes the segment-table source: the bucket source has one kernel)
template <typename WT, bool PL, int PARTS>
void* tile_kernel_src(int npt, int ipt, int src) {
    if (src == 1) return tile_kernel_npt<WT, 1, true, PL, PARTS>(npt);
    return ipt == 1 ? tile_kernel_npt<WT, 1, false, PL, PARTS>(npt) : tile_kernel_npt<WT, 4, false, PL, PARTS>(npt);
}

// Plastic pulls are compiled for packed edges only (the format of every connectome whose tile
// indices fit beside the weights); with the others plastic rows are pushed.
// Blocks in two parts likewise (with the other formats a block runs as one part).
template <typename WT>
void* tile_kernel_fmt(int npt, int ipt, int src, int pl, int parts) {
    if constexpr (std::is_same<WT, PackedEdge>::value) {
        if (parts == 2) return pl ? tile_kernel_src<WT, true, 2>(npt, ipt, src) : tile_kernel_src<WT, false, 2>(npt, ipt, src);
        if (pl) return tile_kernel_src<WT, true, 1>(npt, ipt, src);
    }
    return tile_kernel_src<WT, false, 1>(npt, ipt, src);
}

// The lif_tile instantiation for an edge format (0: int16 col/w, 1: int32 col/w, 2: packed),
// neurons per thread, staging depth, segment source, plastic pulls, parts and two steps per LDS
// word (compiled for packed edges, the bucket source, one part and no plastic pulls).
void* tile_kernel(int fmt, int npt, int ipt, int src, int pl, int parts, int h16) {
    if (fmt == 2 && h16 && src == 1 && !pl && parts == 1) {
        switch (npt) {
            case 1: return reinterpret_cast<void*>(&lif_tile<PackedEdge, 1, 1, true, false, 1, true>);
            case 2: return reinterpret_cast<void*>(&lif_tile<PackedEdge, 2, 1, true, false, 1, true>);
            case 4: return reinterpret_cast<void*>(&lif_tile<PackedEdge, 4, 1, true, false, 1, true>);
            default: return reinterpret_cast<void*>(&lif_tile<PackedEdge, 8, 1, true, false, 1, true>);
        }
    }
    if (fmt == 2) return tile_kernel_fmt<PackedEdge>(npt, ipt, src, pl, parts);
    if (fmt == 0) return tile_kernel_fmt<int16_t>(npt, ipt, src, 0, 1);
    return tile_kernel_fmt<int32_t>(npt, ipt, src, 0, 1);
}
int tile_fmt(const fz_engine* e) { return e->pk ? 2 : (e->w_bits == 16 ? 0 : 1); }

// LDS staging of the segment-table source (segment, prefix end, step per staged entry).
size_t tile_staging(const Shape& s) { return s.src != 1 ? (size_t)3 * s.ipt * s.ut * 4 : 0; }

// Row stride of the LDS accumulator: the first part's arrival columns and the hit-mask column,
// rounded up to odd (no bank conflicts in the step loop) when that fits the LDS budget.
int tile_kp(const fz_engine* e, const Shape& s, int K) {
    const int kh = s.h16 ? (K + 1) / 2 : part_kh(s, K);  // arrival words per neuron
    const int odd = (kh + 1) | 1;
    return (size_t)odd * e->ts * 4 + tile_staging(s) <= e->lds_max - TILE_STATIC_LDS ? odd : kh + 1;
}

size_t tile_lds(const fz_engine* e, const Shape& s, int K) {
    return (size_t)tile_kp(e, s, K) * e->ts * 4 + tile_staging(s);
}

__global__ void tile_seg_count(int n, int n_tiles, const int* tb, int* cnt) {
    __shared__ int part[MAX_WAVES];
    const int T = blockIdx.x;
    int c = 0;
    for (int i = blockIdx.y * blockDim.x + threadIdx.x; i < n; i += gridDim.y * blockDim.x) {
        const int* tr = tb + (size_t)i * (n_tiles + 1) + T;
        c += (tr[1] - tr[0] + BK_PIECE - 1) / BK_PIECE;  // bucket entries of this segment
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t) atomicAdd(&cnt[T], t);
    }
}

// Checks a caller's segment table: row i's entries run from indptr[i] to indptr[i+1] without
// decreasing, and every edge of segment t targets a neuron of tile t.  Both tile sources index
// edges through this table, so a bad table is rejected here instead of being read out of range.
__global__ void tile_validate(int n, int n_tiles, int tile, const int* indptr, const int* col, const int* tb,
                              int* bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int* tr = tb + (size_t)i * (n_tiles + 1);
        if (tr[0] != indptr[i] || tr[n_tiles] != indptr[i + 1]) {
            atomicOr(bad, 1);
            continue;
        }
        const int r0 = indptr[i], r1 = indptr[i + 1];
        for (int t = 0; t < n_tiles; t++) {
            const int lo = tr[t], hi = tr[t + 1];
            if (hi < lo || lo < r0 || hi > r1) {  // checked before any col[] read
                atomicOr(bad, 1);
                break;
            }
            for (int e = lo; e < hi; e++)
                if (col[e] < t * tile || col[e] >= (t + 1) * tile) {
                    atomicOr(bad, 1);
                    break;
                }
        }
    }
}

int validate_tiles(fz_engine* e, const int* tb, int n_tiles, int tile) {
    int* bad = nullptr;
    HIPCHK(hipMalloc(&bad, sizeof(int)));
    HIPCHK(hipMemset(bad, 0, sizeof(int)));
    const int gv = e->n / 256 + 1 < 4096 ? e->n / 256 + 1 : 4096;
    hipLaunchKernelGGL(tile_validate, dim3(gv), dim3(256), 0, 0, e->n, n_tiles, tile, e->indptr, e->col, tb, bad);
    int h_bad = 0;
    hipError_t rv = hipGetLastError();
    if (rv == hipSuccess) rv = hipMemcpy(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(bad);
    if (rv != hipSuccess) {
        (void)hipGetLastError();
        return fail(FZ_E_HIP, "checking the segment table failed: %s", hipGetErrorString(rv));
    }
    if (h_bad) return fail(FZ_E_ARG, "segment table does not match the fixed CSR and tile size %d", tile);
    return FZ_OK;
}

// (first edge, length) of every (row, tile) segment, [n][n_tiles]: the table source of lif_tile
// reads one pair per spike instead of two table entries and the row start.
__global__ void tile_pairs(int n, int n_tiles, const int* tb, int2* tp) {
    const long long total = (long long)n * n_tiles;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q / n_tiles), t = (int)(q - (long long)i * n_tiles);
        const long long at = (long long)i * (n_tiles + 1) + t;
        tp[q] = make_int2(tb[at], tb[at + 1] - tb[at]);
    }
}

// Plastic CSR checks for the pair table: row pointers from 0 to n_pl without decreasing, targets
// in [0, n) and non-decreasing within each row (so that each tile's targets form one segment).
__global__ void pl_check(int n, int n_pl, const int* pl_indptr, const int* pl_post, int* bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r0 = pl_indptr[i], r1 = pl_indptr[i + 1];
        if ((i == 0 && r0 != 0) || (i == n - 1 && r1 != n_pl) || r1 < r0 || r0 < 0 || r1 > n_pl) {
            *bad = 1;
            continue;
        }
        for (int q = r0; q < r1; q++)
            if ((unsigned)pl_post[q] >= (unsigned)n || (q > r0 && pl_post[q] < pl_post[q - 1])) *bad = 1;
    }
}

// (first plastic edge, length) of every (row, tile) segment of the plastic CSR, [n][n_tiles].
__global__ void tile_pl_pairs(int n, int n_tiles, int ts, const int* pl_indptr, const int* pl_post, int2* ptp) {
    const long long total = (long long)n * n_tiles;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q / n_tiles), t = (int)(q - (long long)i * n_tiles);
        const int r0 = pl_indptr[i], r1 = pl_indptr[i + 1];
        auto first_ge = [&](int v) {  // first edge of the row with target >= v
            int lo = r0, hi = r1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pl_post[mid] < v) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        const int lo = first_ge(t * ts), hi = first_ge((t + 1) * ts);
        ptp[q] = make_int2(lo, hi - lo);
    }
}

// Bucket entries that plastic segments add to each tile (as tile_seg_count for the fixed ones).
__global__ void tile_pl_count(int n, int n_tiles, const int2* ptp, int* cnt) {
    __shared__ int part[MAX_WAVES];
    const int T = blockIdx.x;
    int c = 0;
    for (int i = blockIdx.y * blockDim.x + threadIdx.x; i < n; i += gridDim.y * blockDim.x)
        c += (ptp[(size_t)i * n_tiles + T].y + BK_PIECE - 1) / BK_PIECE;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t) atomicAdd(&cnt[T], t);
    }
}

__global__ void tile_rel16(int n, int n_tiles, const int* indptr, const int* tb, uint16_t* tb16, int* over) {
    const long long total = (long long)n * (n_tiles + 1);
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q / (n_tiles + 1));
        const int d = tb[q] - indptr[i];
        if (d < 0 || d > 65535) *over = 1;
        tb16[q] = (uint16_t)d;
    }
}

// The plastic pair table, for plastic pulls: when the plastic CSR is sorted by target within rows
// and the table takes at most TP_MAX_BYTES.  Otherwise there is none and every plastic row is
// pushed at spike time.
int plastic_pairs(fz_engine* e, int n_tiles, int tile) {
    (void)hipFree(e->ptp);
    e->ptp = nullptr;
    const size_t np = (size_t)e->n * (size_t)n_tiles;
    if (e->n_pl <= 0 || e->nnz + e->n_pl >= (1ll << 31) || np * sizeof(int2) > TP_MAX_BYTES) return FZ_OK;
    int* bad = nullptr;
    hipError_t r = hipMalloc(&bad, sizeof(int));
    if (r == hipSuccess) r = hipMemset(bad, 0, sizeof(int));
    int h = 1;
    if (r == hipSuccess) {
        const unsigned g = (unsigned)(e->n / 256 + 1 < 4096 ? e->n / 256 + 1 : 4096);
        hipLaunchKernelGGL(pl_check, dim3(g), dim3(256), 0, 0, e->n, e->n_pl, e->pl_indptr, e->pl_post, bad);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipMemcpy(&h, bad, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(bad);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        return fail(FZ_E_HIP, "checking the plastic CSR failed: %s", hipGetErrorString(r));
    }
    if (h) return FZ_OK;  // not sorted by target within rows: plastic rows are pushed
    if (hipMalloc(&e->ptp, np * sizeof(int2)) != hipSuccess) {
        (void)hipGetLastError();
        e->ptp = nullptr;
        return FZ_OK;
    }
    const unsigned g = (unsigned)(np / 256 + 1 < 65536 ? np / 256 + 1 : 65536);
    hipLaunchKernelGGL(tile_pl_pairs, dim3(g), dim3(256), 0, 0, e->n, n_tiles, tile, e->pl_indptr, e->pl_post, e->ptp);
    if ((r = hipGetLastError()) == hipSuccess) r = hipDeviceSynchronize();
    if (r != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(e->ptp);
        e->ptp = nullptr;
        return fail(FZ_E_HIP, "building the plastic pair table failed: %s", hipGetErrorString(r));
    }
    return FZ_OK;
}

// Bucket capacities: tile t can receive, per block, ceil(segment length / BK_PIECE) entries per
// row with an edge into t, times the spikes a neuron can emit in one block (ceil(K / n_ref),
// K <= kmax).
int bucket_layout(fz_engine* e, const int* tb, int n_tiles) {
    (void)hipFree(e->bk_off);
    (void)hipFree(e->bk);
    (void)hipFree(e->bk_cnt);
    e->bk_off = nullptr;
    e->bk = nullptr;
    e->bk_cnt = nullptr;
    e->bk_B = 0;
    e->bk_stride = 0;
    e->bk_off_h.assign((size_t)n_tiles + 1, 0);
    int* cnt = nullptr;
    HIPCHK(hipMalloc(&cnt, sizeof(int) * (size_t)n_tiles));
    HIPCHK(hipMemset(cnt, 0, sizeof(int) * (size_t)n_tiles));
    const int gy = e->n / 256 + 1 < 64 ? e->n / 256 + 1 : 64;
    hipLaunchKernelGGL(tile_seg_count, dim3(n_tiles, gy), dim3(256), 0, 0, e->n, n_tiles, tb, cnt);
    if (e->n_pl > 0 && e->ptp)  // the plastic segments that buckets carry for plastic pulls
        hipLaunchKernelGGL(tile_pl_count, dim3(n_tiles, gy), dim3(256), 0, 0, e->n, n_tiles, e->ptp, cnt);
    std::vector<int> h((size_t)n_tiles);
    hipError_t r = hipGetLastError();
    if (r == hipSuccess) r = hipMemcpy(h.data(), cnt, sizeof(int) * (size_t)n_tiles, hipMemcpyDeviceToHost);
    (void)hipFree(cnt);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        return fail(FZ_E_HIP, "counting tile segments failed: %s", hipGetErrorString(r));
    }
    const int per = e->c.n_ref > 1 ? e->c.n_ref : 1;
    const long long maxspk = (e->kmax + per - 1) / per;
    for (int t = 0; t < n_tiles; t++) e->bk_off_h[t + 1] = e->bk_off_h[t] + (long long)h[t] * maxspk;
    e->bk_stride = e->bk_off_h[n_tiles];
    // first edge (plastic ones after the nnz fixed ones) fits 24 bits beside length and step
    e->bk4 = e->nnz + (e->n_pl > 0 && e->ptp ? e->n_pl : 0) < (1ll << 24) && e->kmax <= 32;
    // uint16 copy of the table (row offsets) when every row has fewer than 65,536 fixed edges
    (void)hipFree(e->tb16);
    e->tb16 = nullptr;
    {
        uint16_t* t16 = nullptr;
        int* over = nullptr;
        const size_t nt = (size_t)e->n * ((size_t)n_tiles + 1);
        if (hipMalloc(&t16, nt * sizeof(uint16_t)) == hipSuccess && hipMalloc(&over, sizeof(int)) == hipSuccess &&
            hipMemset(over, 0, sizeof(int)) == hipSuccess) {
            const unsigned g = (unsigned)(nt / 256 + 1 < 65536 ? nt / 256 + 1 : 65536);
            hipLaunchKernelGGL(tile_rel16, dim3(g), dim3(256), 0, 0, e->n, n_tiles, e->indptr, tb, t16, over);
            int h = 1;
            if (hipGetLastError() == hipSuccess && hipMemcpy(&h, over, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
                h == 0) {
                e->tb16 = t16;
                t16 = nullptr;
            }
        }
        (void)hipGetLastError();
        (void)hipFree(t16);
        (void)hipFree(over);
    }
    // the pair table, when it takes at most TP_MAX_BYTES (the lookups fall back to the tables above)
    (void)hipFree(e->tp);
    e->tp = nullptr;
    {
        const size_t np = (size_t)e->n * (size_t)n_tiles;
        if (np * sizeof(int2) <= TP_MAX_BYTES && hipMalloc(&e->tp, np * sizeof(int2)) == hipSuccess) {
            const unsigned g = (unsigned)(np / 256 + 1 < 65536 ? np / 256 + 1 : 65536);
            hipLaunchKernelGGL(tile_pairs, dim3(g), dim3(256), 0, 0, e->n, n_tiles, tb, e->tp);
            if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
                (void)hipFree(e->tp);
                e->tp = nullptr;
            }
        }
        (void)hipGetLastError();
    }
    HIPCHK(hipMalloc(&e->bk_off, sizeof(long long) * ((size_t)n_tiles + 1)));
    HIPCHK(hipMemcpy(e->bk_off, e->bk_off_h.data(), sizeof(long long) * ((size_t)n_tiles + 1),
                     hipMemcpyHostToDevice));
    return FZ_OK;
}

// The bucket arena for the bound batch, allocated on first use of tile source 1.
int ensure_buckets(fz_engine* e) {
    if (e->bk && e->bk_B == e->st.B) return FZ_OK;
    (void)hipFree(e->bk);
    (void)hipFree(e->bk_cnt);
    e->bk = nullptr;
    e->bk_cnt = nullptr;
    e->bk_B = 0;
    // two parity buffers of B x stride 4-byte entries; (front, back) counts per bucket
    const size_t ne = 2 * (size_t)e->st.B * (size_t)(e->bk_stride > 0 ? e->bk_stride : 1);
    if (hipMalloc(&e->bk, sizeof(int) * ne) != hipSuccess ||
        hipMalloc(&e->bk_cnt, sizeof(int) * 4 * (size_t)e->st.B * (size_t)e->n_tiles) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(e->bk);
        e->bk = nullptr;
        return fail(FZ_E_HIP, "allocating %zu tile-bucket entries failed", ne);
    }
    HIPCHK(hipMemset(e->bk_cnt, 0, sizeof(int) * 4 * (size_t)e->st.B * (size_t)e->n_tiles));
    e->bk_B = e->st.B;
    return FZ_OK;
}

void* bucket_kernel(const fz_engine* e, int pl) {
    if (e->tb16)
        return pl ? reinterpret_cast<void*>(&tile_bucket<true, true>) : reinterpret_cast<void*>(&tile_bucket<true, false>);
    return pl ? reinterpret_cast<void*>(&tile_bucket<false, true>) : reinterpret_cast<void*>(&tile_bucket<false, false>);
}

BucketArgs make_bucket(const fz_engine* e, const fz_inputs* in, int K, int kb, int blk) {
    BucketArgs a{};
    a.blk = blk;
    a.clock = e->st.clock;
    a.N = e->n;
    a.B = e->st.B;
    a.K = K;
    a.kh = K;  // one part; the caller sets the shape's
    a.kb = kb;
    a.nd = e->c.n_delay;
    a.M = 3 * e->c.n_delay;
    a.n_tiles = e->n_tiles;
    a.spk_idx = e->st.spk_idx;
    a.spk_cnt = e->st.spk_cnt;
    a.tb = e->tb;
    a.tb16 = e->tb16;
    a.ptp = nullptr;  // set by the caller when the shape pulls plastic synapses
    a.nnz = e->nnz;
    a.indptr = e->indptr;
    a.bk_off = e->bk_off;
    a.bk_stride = e->bk_stride;
    a.bk = e->bk;
    a.bk_cnt = e->bk_cnt;
    a.alive = in->alive;
    a.err = e->err;
    a.dv_M = div32_for((unsigned)a.M);
    a.dv_grp = div32_for((unsigned)((e->n_tiles + 63) / 64));  // the launch's gridDim.x
    return a;
}


// One window as a sequence of kernel launches (the graph builder mirrors it).
// With `ev_*` non-null, every kernel is timed by a start and a stop event that the runtime
// stamps from the kernel's own dispatch (hipExtLaunchKernel), so the pair measures the kernel
// and not the event packets around it (about 2 us on a 15 us launch).
int enqueue(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
            const Shape& s, hipStream_t stream, std::vector<hipEvent_t>* ev_upd, std::vector<hipEvent_t>* ev_push,
            const WorkList& wl) {
    const dim3 ugrid((e->n + s.ut - 1) / s.ut, e->st.B);
    const dim3 pgrid = push_grid(e, s);
    int kb = 0, blk = 0;
    auto launch = [&](const void* fn, dim3 grid, dim3 block, void** args, size_t lds,
                      std::vector<hipEvent_t>* ev) -> int {
        if (!ev) {  // untimed (and capturable): the plain launch
            HIPCHK(hipLaunchKernel(fn, grid, block, args, lds, stream));
            HIPCHK(hipGetLastError());
            return FZ_OK;
        }
        hipEvent_t x[2];
        for (auto& y : x) {
            HIPCHK(hipEventCreate(&y));
            ev->push_back(y);
        }
        HIPCHK(hipExtLaunchKernel(fn, grid, block, args, lds, stream, x[0], x[1], 0));
        HIPCHK(hipGetLastError());
        return FZ_OK;
    };
    int rc;
    if (e->mode == 2) {
        const dim3 tgrid(e->n_tiles, e->st.B);
        const dim3 bgrid((e->n_tiles + 63) / 64, e->st.B);
        void* fn = tile_kernel(tile_fmt(e), e->ts / s.ut, s.ipt, s.src, s.pl, s.parts, s.h16);
        if (s.src == 1 && (rc = ensure_buckets(e))) return rc;
        while (kb < steps) {
            const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
            if (s.src == 1) {  // timed with the push timers: it is the spike expansion
                BucketArgs ba = make_bucket(e, in, K, kb, blk);
                ba.kh = part_kh(s, K);
                if (s.pl) ba.ptp = e->ptp;
                void* bargs[] = {&ba};
                if ((rc = launch(bucket_kernel(e, s.pl), bgrid, dim3(BK_THREADS), bargs, 0, ev_push))) return rc;
            }
            TileArgs ta = make_tile(e, in, K, tile_kp(e, s, K), kb, blk, kb == 0, steps, hits, raster);
            ta.kh = part_kh(s, K);
            if (s.src == 2) ta.tp = nullptr;
            if (s.pl) ta.ptp = e->ptp;
            void* args[] = {&ta};
            if ((rc = launch(fn, tgrid, dim3(s.ut), args, tile_lds(e, s, K), ev_upd))) return rc;
            kb += K;
            blk++;
        }
        hipLaunchKernelGGL(lif_clock, dim3(1), dim3(64), 0, stream, e->st.clock, steps, blk);
        HIPCHK(hipGetLastError());
        return FZ_OK;
    }
    while (kb < steps) {
        const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
        UpdArgs ua = make_upd(e, in, s, K, kb, blk, kb == 0, hits, raster, wl);
        void* uargs[] = {&ua};
        if ((rc = launch(reinterpret_cast<const void*>(&lif_update), ugrid, dim3(s.ut), uargs, 0, ev_upd))) return rc;
        if (e->w_bits == 16) {
            auto pa = make_push<int16_t>(e, in, s, blk, wl);
            void* pargs[] = {&pa};
            if ((rc = launch(reinterpret_cast<const void*>(&lif_push<int16_t>), pgrid, dim3(s.pt), pargs, 0, ev_push)))
                return rc;
        } else {
            auto pa = make_push<int32_t>(e, in, s, blk, wl);
            void* pargs[] = {&pa};
            if ((rc = launch(reinterpret_cast<const void*>(&lif_push<int32_t>), pgrid, dim3(s.pt), pargs, 0, ev_push)))
                return rc;
        }
        kb += K;
        blk++;
    }
    hipLaunchKernelGGL(lif_clock, dim3(1), dim3(64), 0, stream, e->st.clock, steps, blk);
    HIPCHK(hipGetLastError());
    return FZ_OK;
}

void free_worklist(fz_engine* e) {
    (void)hipFree(e->wl.ctr);
    (void)hipFree(e->wl.off);
    (void)hipFree(e->wl.fbeg);
    (void)hipFree(e->wl.fdeg);
    (void)hipFree(e->wl.pbeg);
    (void)hipFree(e->wl.slot);
    (void)hipFree(e->wl.cs);
    e->wl = WorkList{};
    e->wl_B = 0;
}

// Work-list capacity of the push engine: entries per fly (a neuron fires at most
// ceil(K / n_ref) times in a block of K <= kmax steps) and chunk-start slots per fly.
void worklist_dims(const fz_engine* e, long long* cap, long long* ncs) {
    const int per = e->c.n_ref > 1 ? e->c.n_ref : 1;
    const long long maxspk = (e->kmax + per - 1) / per;
    *cap = (long long)e->n * maxspk;
    *ncs = (e->nnz + e->n_pl) * maxspk / QMIN + 2;
}

// The work list of B flies laid out from `base` (256-byte aligned arrays); returns the bytes it
// spans.  base = nullptr only measures.
size_t static_worklist(const fz_engine* e, int B, unsigned char* base, WorkList* wl) {
    long long cap, ncs;
    worklist_dims(e, &cap, &ncs);
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* q = base ? static_cast<void*>(base + off) : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const size_t ne = (size_t)B * (size_t)cap;
    wl->ctr = static_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * (size_t)B));
    wl->off = static_cast<unsigned*>(take(sizeof(unsigned) * ne));
    wl->fbeg = static_cast<int*>(take(sizeof(int) * ne));
    wl->fdeg = static_cast<int*>(take(sizeof(int) * ne));
    wl->pbeg = static_cast<int*>(take(sizeof(int) * ne));
    wl->slot = static_cast<int*>(take(sizeof(int) * ne));
    wl->cs = static_cast<unsigned*>(take(sizeof(unsigned) * (size_t)B * (size_t)ncs));
    wl->cap = cap;
    wl->ncs = ncs;
    return off;
}

// Device address of `scratch` on device `dev` (current), looked up once per device.
constexpr int MAX_DEVICES = 64;
std::mutex g_scr_mu;
unsigned char* g_scr[MAX_DEVICES] = {};

int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(FZ_E_ARG, "device %d outside [0, %d)", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scr_mu);
    if (!g_scr[dev]) {
        void* q = nullptr;
        const hipError_t r = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !q) {
            (void)hipGetLastError();
            return fail(FZ_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(r));
        }
        g_scr[dev] = static_cast<unsigned char*>(q);
    }
    *out = g_scr[dev];
    return FZ_OK;
}

}  // namespace

extern "C" {

int fz_version(void) { return FZ_ABI_VERSION; }

#ifndef FZ_SRC_HASH
#define FZ_SRC_HASH "unknown"
#endif
const char* fz_build_info(void) { return "src " FZ_SRC_HASH "; " __VERSION__; }

const char* fz_last_error(void) { return g_err.c_str(); }

int fz_create(int n, long long nnz, const int* indptr, const int* col, const void* w, int w_bits, int frac_bits,
              const fz_params* p, fz_engine** out) {
    if (!out) return fail(FZ_E_ARG, "null output handle");
    *out = nullptr;
    if (n < 1 || n >= (1 << 30)) return fail(FZ_E_ARG, "n = %d outside [1, 2^30)", n);
    if (nnz < 0 || nnz >= (1ll << 31)) return fail(FZ_E_ARG, "nnz = %lld out of int32 range", nnz);
    if (!p) return fail(FZ_E_ARG, "null params");
    if (!indptr) return fail(FZ_E_ARG, "null indptr");
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
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        delete e;
        return fail(FZ_E_HIP, "hipGetDevice failed");
    }
    e->dev = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        e->n_cu = prop.multiProcessorCount;
        e->lds_max = prop.sharedMemPerBlock;
    }
    if (hipMalloc(&e->err, sizeof(int)) != hipSuccess || hipMemset(e->err, 0, sizeof(int)) != hipSuccess) {
        fz_destroy(e);
        return fail(FZ_E_HIP, "allocating the engine's error flag failed");
    }
    int rc = scratch_address(dev, &e->scr);  // set-up of fz_run_static, with the device current
    if (rc) {
        fz_destroy(e);
        return rc;
    }
    *out = e;
    return FZ_OK;
}

int fz_destroy(fz_engine* e) {
    if (!e) return FZ_OK;
    free_worklist(e);
    (void)hipFree(e->err);
    (void)hipFree(e->pk);
    (void)hipFree(e->bk_off);
    (void)hipFree(e->bk);
    (void)hipFree(e->bk_cnt);
    (void)hipFree(e->tb16);
    (void)hipFree(e->tp);
    (void)hipFree(e->ptp);
    delete e;
    return FZ_OK;
}

int fz_set_plastic(fz_engine* e, int n_pl, const int* pl_indptr, const int* pl_post, void* stream) {
    (void)stream;
    if (!e) return fail(FZ_E_ARG, "null engine");
    if (n_pl < 0) return fail(FZ_E_ARG, "n_pl %d < 0", n_pl);
    if (n_pl > 0 && (!pl_indptr || !pl_post)) return fail(FZ_E_ARG, "n_pl > 0 needs pl_indptr and pl_post");
    e->n_pl = n_pl;
    e->pl_indptr = n_pl > 0 ? pl_indptr : nullptr;
    e->pl_post = n_pl > 0 ? pl_post : nullptr;
    if (e->mode == 2) {  // the tile engine's plastic pair table and bucket capacities follow the new set
        int rc = plastic_pairs(e, e->n_tiles, e->ts);
        if (!rc) rc = bucket_layout(e, e->tb, e->n_tiles);
        if (rc) {
            e->mode = 1;
            e->bound = false;
            return rc;
        }
    }
    if (e->bound) {  // re-check capacity limits with the new edge count
        fz_state st = e->st;
        e->bound = false;
        return fz_bind(e, &st);
    }
    return FZ_OK;
}

int fz_bind(fz_engine* e, const fz_state* st) {
    if (!e || !st) return fail(FZ_E_ARG, "null engine or state");
    if (st->B < 1 || st->B > 65535) return fail(FZ_E_ARG, "B = %d outside [1, 65535]", st->B);
    if (!st->u || !st->g || !st->ref || !st->ring || !st->counts || !st->clock)
        return fail(FZ_E_ARG, "null state array");
    const int want_slots = e->mode == 2 ? 2 * e->c.n_delay : e->c.n_delay;
    if (st->ring_slots != want_slots)
        return fail(FZ_E_ARG, "ring has %d slots, the %s engine needs %d", st->ring_slots,
                    e->mode == 2 ? "tile" : "push", want_slots);
    if (e->mode == 2 && (!st->spk_idx || !st->spk_cnt))
        return fail(FZ_E_ARG, "the tile engine needs the spike lists spk_idx and spk_cnt");
    const int per = e->c.n_ref > 1 ? e->c.n_ref : 1;
    const long long maxspk = (e->kmax + per - 1) / per;  // spikes per neuron per block
    const long long edges = (e->nnz + e->n_pl) * maxspk;  // edges per fly and block, at most
    long long cap, ncs;                                   // entries and chunk-start slots per fly
    worklist_dims(e, &cap, &ncs);
    if (cap >= (1ll << 31)) return fail(FZ_E_ARG, "n too large for the per-fly work list");
    if (edges >= (1ll << 31)) return fail(FZ_E_ARG, "nnz too large for the per-fly work list");
    if (st->B != e->wl_B || cap != e->wl.cap || ncs != e->wl.ncs) {
        free_worklist(e);
        const size_t ne = (size_t)st->B * (size_t)cap;
        WorkList& w = e->wl;
        HIPCHK(hipMalloc(&w.ctr, sizeof(unsigned long long) * 2 * (size_t)st->B));
        HIPCHK(hipMemset(w.ctr, 0, sizeof(unsigned long long) * 2 * (size_t)st->B));
        HIPCHK(hipMalloc(&w.off, sizeof(unsigned) * ne));
        HIPCHK(hipMalloc(&w.fbeg, sizeof(int) * ne));
        HIPCHK(hipMalloc(&w.fdeg, sizeof(int) * ne));
        HIPCHK(hipMalloc(&w.pbeg, sizeof(int) * ne));
        HIPCHK(hipMalloc(&w.slot, sizeof(int) * ne));
        HIPCHK(hipMalloc(&w.cs, sizeof(unsigned) * (size_t)st->B * (size_t)ncs));
        w.cap = cap;
        w.ncs = ncs;
        e->wl_B = st->B;
    }
    e->st = *st;
    e->bound = true;
    return FZ_OK;
}

namespace {

// Packed copy of the fixed edges for tile size `tile` (see PackedEdge).  Left null, and the
// tile kernel reads col and w, when a value does not fit the bits left beside the offset.
int pack_edges(fz_engine* e, int tile) {
    (void)hipFree(e->pk);
    e->pk = nullptr;
    e->lbits = 0;
    e->in_max = -1;
    if (e->nnz == 0) return FZ_OK;
    int lbits = 0;
    while ((1 << lbits) < tile) lbits++;
    int* pk = nullptr;
    int* over = nullptr;
    // one padding entry (zero) after the last edge: the tile kernel reads the edge after a
    // piece's first one unconditionally
    if (hipMalloc(&pk, ((size_t)e->nnz + 1) * sizeof(int)) != hipSuccess ||
        hipMemset(pk + e->nnz, 0, sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(pk);
        return fail(FZ_E_HIP, "allocating %lld packed edges failed", e->nnz);
    }
    if (hipMalloc(&over, sizeof(int)) != hipSuccess || hipMemset(over, 0, sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(pk);
        (void)hipFree(over);
        return fail(FZ_E_HIP, "allocating the pack flag failed");
    }
    const unsigned grid = (unsigned)((e->nnz + 255) / 256 < 65536 ? (e->nnz + 255) / 256 : 65536);
    if (e->w_bits == 16)
        hipLaunchKernelGGL(tile_pack<int16_t>, dim3(grid), dim3(256), 0, 0, e->nnz, e->col,
                           static_cast<const int16_t*>(e->w), e->c.frac_bits, tile, lbits, pk, over);
    else
        hipLaunchKernelGGL(tile_pack<int32_t>, dim3(grid), dim3(256), 0, 0, e->nnz, e->col,
                           static_cast<const int32_t*>(e->w), e->c.frac_bits, tile, lbits, pk, over);
    int h_over = 0;
    hipError_t r = hipGetLastError();
    if (r == hipSuccess) r = hipMemcpy(&h_over, over, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(over);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(pk);
        return fail(FZ_E_HIP, "packing edges failed: %s", hipGetErrorString(r));
    }
    if (h_over) {
        (void)hipFree(pk);
        return FZ_OK;
    }
    e->pk = pk;
    e->lbits = lbits;
    // the input-strength bound (a failure here only leaves the two-steps-per-word mode off)
    e->in_max = -1;
    unsigned long long* ins = nullptr;
    if (hipMalloc(&ins, sizeof(unsigned long long) * ((size_t)e->n + 1)) == hipSuccess &&
        hipMemset(ins, 0, sizeof(unsigned long long) * ((size_t)e->n + 1)) == hipSuccess) {
        hipLaunchKernelGGL(in_strength, dim3(grid), dim3(256), 0, 0, e->nnz, e->col, pk, lbits, ins);
        const unsigned g2 = (unsigned)((e->n + 255) / 256 < 1024 ? (e->n + 255) / 256 : 1024);
        hipLaunchKernelGGL(max_u64, dim3(g2), dim3(256), 0, 0, e->n, ins, ins + e->n);
        unsigned long long mx = 0;
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(&mx, ins + e->n, sizeof(mx), hipMemcpyDeviceToHost) == hipSuccess)
            e->in_max = mx < (1ull << 62) ? (long long)mx : (1ll << 62);
    }
    (void)hipGetLastError();
    (void)hipFree(ins);
    return FZ_OK;
}


}  // namespace

int fz_set_tiles(fz_engine* e, int tile, int n_tiles, const int* tb, const int* ptg_ptr, const int* ptg_idx) {
    if (!e) return fail(FZ_E_ARG, "null engine");
    if (tile < 64 || tile > 8192 || tile % 64) return fail(FZ_E_ARG, "tile %d: need a multiple of 64 in [64, 8192]", tile);
    if (n_tiles != (e->n + tile - 1) / tile) return fail(FZ_E_ARG, "n_tiles %d != ceil(n / tile)", n_tiles);
    if (!tb) return fail(FZ_E_ARG, "null tile table");
    if ((long long)e->n * (n_tiles + 1) >= (1ll << 31))
        return fail(FZ_E_ARG, "segment table of %d x %d entries is too large (limit 2^31)", e->n, n_tiles + 1);
    if (e->n_pl > 0 && (!ptg_ptr || !ptg_idx)) return fail(FZ_E_ARG, "plastic synapses need the per-tile target lists");
    // kmax arrival columns and the hit-mask column per neuron, plus the smallest staging
    const size_t lds = (size_t)(e->kmax + 1) * tile * 4 + (size_t)3 * 64 * 4;
    if (lds > e->lds_max - TILE_STATIC_LDS)
        return fail(FZ_E_ARG, "tile %d needs %zu B of LDS for %d steps; the device allows %zu", tile, lds, e->kmax,
                    e->lds_max - TILE_STATIC_LDS);
    for (int fmt : {0, 1, 2})
        for (int npt : {1, 2, 4, 8})
            for (int ipt : {1, 4})
                for (int src : {0, 1})
                    for (int pl : {0, 1})
                        for (int parts : {1, 2})
                            for (int h16 : {0, 1}) {
                                hipError_t r = hipFuncSetAttribute(tile_kernel(fmt, npt, ipt, src, pl, parts, h16),
                                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                                   (int)(e->lds_max - TILE_STATIC_LDS));
                                if (r != hipSuccess) {
                                    (void)hipGetLastError();  // no sticky error for the caller's next HIP call
                                    return fail(FZ_E_HIP, "hipFuncSetAttribute(max dynamic LDS %zu) failed: %s",
                                                e->lds_max - TILE_STATIC_LDS, hipGetErrorString(r));
                                }
                            }
    int rc = validate_tiles(e, tb, n_tiles, tile);
    if (rc) return rc;
    // from here on a failure leaves the engine unbound, in push mode, until set up again
    if ((rc = pack_edges(e, tile)) || (rc = plastic_pairs(e, n_tiles, tile)) || (rc = bucket_layout(e, tb, n_tiles))) {
        e->mode = 1;
        e->bound = false;
        return rc;
    }
    e->mode = 2;
    e->ts = tile;
    e->n_tiles = n_tiles;
    e->tb = tb;
    e->ptg_ptr = ptg_ptr;
    e->ptg_idx = ptg_idx;
    e->bound = false;  // the state must be re-bound (ring slots, spike lists)
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
    return enqueue(e, in, steps, hits, raster, s, static_cast<hipStream_t>(stream), nullptr, nullptr, e->wl);
}

int fz_run_static(fz_engine* e, const fz_inputs* in, int steps, const unsigned* hits, unsigned* raster,
                  const fz_tune* tune, void* stream) {
    int rc = check_inputs(e, in);
    if (rc) return rc;
    if (steps < 0) return fail(FZ_E_ARG, "steps %d < 0", steps);
    if (steps == 0) return FZ_OK;
    Shape s;
    if ((rc = resolve(e, tune, &s))) return rc;
    hipStream_t hs = static_cast<hipStream_t>(stream);
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (hs) HIPCHK(hipStreamGetDevice(hs, &dev));
    if (dev != e->dev)
        return fail(FZ_E_STATE, "fz_run_static: the stream's device %d is not the engine's device %d", dev, e->dev);
    if (e->mode == 2) {  // the segment-table source reads only caller state: no engine scratch
        if (s.src == 1)
            return fail(FZ_E_ARG, "fz_run_static: the tile engine's bucket source (tile_source 1) needs the bucket "
                        "arena; use tile_source 0 or fz_run");
        return enqueue(e, in, steps, hits, raster, s, hs, nullptr, nullptr, e->wl);
    }
    WorkList wl{};
    const size_t need = static_worklist(e, e->st.B, nullptr, &wl);
    if (need > (size_t)FZ_STATIC_SCRATCH_BYTES)
        return fail(FZ_E_ARG, "fz_run_static: %d flies of %d neurons need %zu bytes of work list; the static scratch "
                    "holds %lld (use fz_run)", e->st.B, e->n, need, (long long)FZ_STATIC_SCRATCH_BYTES);
    static_worklist(e, e->st.B, e->scr, &wl);
    // the work-list counters start at zero (a fill kernel, so that a captured call has no memset node);
    // the other arrays are written before they are read in every block
    hipLaunchKernelGGL(zero_u64, dim3((2 * e->st.B + 255) / 256), dim3(256), 0, hs, wl.ctr, 2 * e->st.B);
    HIPCHK(hipGetLastError());
    return enqueue(e, in, steps, hits, raster, s, hs, nullptr, nullptr, wl);
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
    rc = enqueue(e, in, steps, nullptr, nullptr, s, hs, &eu, &ep, e->wl);
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
    for (auto ev : eu) (void)hipEventDestroy(ev);
    for (auto ev : ep) (void)hipEventDestroy(ev);
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
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        delete gr;
        return fail(FZ_E_HIP, "%s failed: %s", what, hipGetErrorString(err));
    };
    hipError_t err = hipGraphCreate(&gr->graph, 0);
    if (err != hipSuccess) return bail("hipGraphCreate", err);
    hipGraphNode_t prev = nullptr;
    auto add = [&](void* fn, dim3 grid, dim3 block, void** args, size_t lds = 0) -> hipError_t {
        hipKernelNodeParams kp{};
        kp.func = fn;
        kp.gridDim = grid;
        kp.blockDim = block;
        kp.sharedMemBytes = (unsigned)lds;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node;
        hipError_t r = hipGraphAddKernelNode(&node, gr->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
        if (r == hipSuccess) prev = node;
        return r;
    };
    const dim3 ugrid((e->n + s.ut - 1) / s.ut, e->st.B);
    const dim3 pgrid = push_grid(e, s);
    int kb = 0, blk = 0;
    std::vector<hipGraphNode_t> lasts;  // tile engine: the last node of each chain
    if (e->mode == 2) {
        void* fn = tile_kernel(tile_fmt(e), e->ts / s.ut, s.ipt, s.src, s.pl, s.parts, s.h16);
        if (s.src == 1 && (rc = ensure_buckets(e))) {
            (void)hipGraphDestroy(gr->graph);
            delete gr;
            return rc;
        }
        // With the bucket source the flies are split into two halves with independent chains of
        // nodes, so that one half's bucket prologue (latency-bound) can run beside the other
        // half's tile kernel (ALU-bound).  Within a chain, the bucket node of block j needs the
        // spikes of steps [kb - n_delay, kb - n_delay + K) and its parity buffer free (last read
        // by block j - 2): with K <= n_delay / 2 both come from block j - 2 or earlier, so it
        // also runs beside the chain's tile kernel of block j - 1.
        struct Chain {
            int b0, nb;
            hipGraphNode_t prev = nullptr;
            std::vector<hipGraphNode_t> tiles;
            std::vector<int> tile_end;  // window step after each tile block
        };
        const int B = e->st.B, H = (s.src == 1 && B >= 2) ? 2 : 1;
        std::vector<Chain> chains(H);
        for (int h = 0; h < H; h++) {
            chains[h].b0 = h * B / H;
            chains[h].nb = (h + 1) * B / H - chains[h].b0;
        }
        while (kb < steps) {
            const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
            for (Chain& c : chains) {
                hipGraphNode_t bnode = nullptr;
                if (s.src == 1) {
                    BucketArgs ba = make_bucket(e, in, K, kb, blk);
                    ba.kh = part_kh(s, K);
                    ba.b0 = c.b0;
                    if (s.pl) ba.ptp = e->ptp;
                    void* bargs[] = {&ba};
                    const int need = kb - e->c.n_delay + K - 1;  // last spike step it reads
                    int dep = blk - 2;
                    for (int i = 0; i < (int)c.tiles.size(); i++)
                        if (need >= 0 && c.tile_end[i] > need) {
                            dep = i > dep ? i : dep;
                            break;
                        }
                    if (dep >= blk) dep = blk - 1;
                    hipKernelNodeParams kp{};
                    kp.func = bucket_kernel(e, s.pl);
                    kp.gridDim = dim3((e->n_tiles + 63) / 64, c.nb);
                    kp.blockDim = dim3(BK_THREADS);
                    kp.sharedMemBytes = 0;
                    kp.kernelParams = bargs;
                    kp.extra = nullptr;
                    hipGraphNode_t d = dep >= 0 ? c.tiles[dep] : nullptr;
                    if ((err = hipGraphAddKernelNode(&bnode, gr->graph, d ? &d : nullptr, d ? 1 : 0, &kp)) != hipSuccess)
                        return bail("hipGraphAddKernelNode(bucket)", err);
                }
                TileArgs ta = make_tile(e, in, K, tile_kp(e, s, K), kb, blk, kb == 0, steps, nullptr, nullptr);
                ta.kh = part_kh(s, K);
                if (s.src == 2) ta.tp = nullptr;
                if (s.pl) ta.ptp = e->ptp;
                ta.b0 = c.b0;
                ta.nb = c.nb;
                void* targs[] = {&ta};
                hipKernelNodeParams kp{};
                kp.func = fn;
                kp.gridDim = dim3(e->n_tiles, c.nb);
                kp.blockDim = dim3(s.ut);
                kp.sharedMemBytes = (unsigned)tile_lds(e, s, K);
                kp.kernelParams = targs;
                kp.extra = nullptr;
                hipGraphNode_t deps[2];
                int nd = 0;
                if (c.prev) deps[nd++] = c.prev;
                if (bnode) deps[nd++] = bnode;
                hipGraphNode_t tnode;
                if ((err = hipGraphAddKernelNode(&tnode, gr->graph, nd ? deps : nullptr, nd, &kp)) != hipSuccess)
                    return bail("hipGraphAddKernelNode(tile)", err);
                c.prev = tnode;
                c.tiles.push_back(tnode);
                c.tile_end.push_back(kb + K);
            }
            kb += K;
            blk++;
        }
        for (Chain& c : chains) lasts.push_back(c.prev);
    }
    while (kb < steps) {
        const int K = (steps - kb) < s.K ? (steps - kb) : s.K;
        UpdArgs ua = make_upd(e, in, s, K, kb, blk, kb == 0, nullptr, nullptr, e->wl);
        void* uargs[] = {&ua};
        if ((err = add(reinterpret_cast<void*>(&lif_update), ugrid, dim3(s.ut), uargs)) != hipSuccess)
            return bail("hipGraphAddKernelNode(update)", err);
        if (e->w_bits == 16) {
            PushArgs<int16_t> pa = make_push<int16_t>(e, in, s, blk, e->wl);
            void* pargs[] = {&pa};
            err = add(reinterpret_cast<void*>(&lif_push<int16_t>), pgrid, dim3(s.pt), pargs);
        } else {
            PushArgs<int32_t> pa = make_push<int32_t>(e, in, s, blk, e->wl);
            void* pargs[] = {&pa};
            err = add(reinterpret_cast<void*>(&lif_push<int32_t>), pgrid, dim3(s.pt), pargs);
        }
        if (err != hipSuccess) return bail("hipGraphAddKernelNode(push)", err);
        kb += K;
        blk++;
    }
    long long* clock = e->st.clock;
    int st = steps, nb = blk;
    void* cargs[] = {&clock, &st, &nb};
    if (!lasts.empty()) {  // the clock advances after every chain
        hipKernelNodeParams kp{};
        kp.func = reinterpret_cast<void*>(&lif_clock);
        kp.gridDim = dim3(1);
        kp.blockDim = dim3(64);
        kp.sharedMemBytes = 0;
        kp.kernelParams = cargs;
        kp.extra = nullptr;
        hipGraphNode_t node;
        if ((err = hipGraphAddKernelNode(&node, gr->graph, lasts.data(), lasts.size(), &kp)) != hipSuccess)
            return bail("hipGraphAddKernelNode(clock)", err);
    } else if ((err = add(reinterpret_cast<void*>(&lif_clock), dim3(1), dim3(64), cargs)) != hipSuccess) {
        return bail("hipGraphAddKernelNode(clock)", err);
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
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
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
