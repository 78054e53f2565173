// pp_kernels.h -- gfx950 kernels for one Pauli-rotation step (branch, tail
// permutation, merge with truncation), Pauli noise, and overlapwithzero.
//
// One rotation with generator mask m on a sorted, duplicate-free sum S:
//   k_branch   compacts the indices of the terms that anticommute with m, in
//              order (src/Base/xorbranch.jl:214-240: created terms are appended
//              in parent order), with the sign of the created term in bit 31;
//   k_xorperm  moves that list into ascending order of term ^ m in one pass.
//              For a run of set mask bits [lo, hi] the XOR reverses the order
//              of the sub-blocks that agree from bit lo up inside every block
//              that agrees above hi (src/Base/Merge/xortailmerge.jl:1-11). The
//              rank of a ^ m is therefore j + sum over runs of
//              (s_{hi+1} + e_{hi+1} - e_lo - s_lo), with s_k, e_k the bounds of
//              the block of parent-order entries that agree with a from bit k up,
//              so all runs are handled at once from the parent order;
//   k_merge    merges S (kept coefficients scaled by cos) with the permuted
//              tail (coefficients sin * sign) along merge paths, sums equal keys
//              as head + tail, truncates every output and compacts
//              (src/Base/Merge/sortedtailmerge.jl:140-202).
//
// Sizes live on the device. Gate number `seq` reads its input state from
// ctl->st[seq & 1] and leaves the next gate's state in ctl->st[(seq & 1) ^ 1],
// so the host enqueues whole circuits without reading anything back. A gate
// whose output would not fit sets ctl->overflow; every later kernel then
// returns at once and the host grows the buffers and resumes at that gate.
// The kernels are persistent: a fixed grid loops over tiles, taking tile ids
// from an atomic counter in start order (look-back kernels) or by stride.
#pragma once
#include <climits>

#include "pp_device.h"

namespace pp {

constexpr uint32_t SRC_MASK = 0x7fffffffu;
constexpr uint32_t SRC_NEG = 0x80000000u;

struct Bufs {
    uint64_t* keys[2];   // W limb planes of `cap` words each
    double* coeffs[2];   // cap x B
    uint32_t* tail[2];   // cap entries each
    uint64_t* tkeys;     // W planes of cap words: the permuted tail's keys (term ^ mask)
    double* tcoef;       // B = 1: the permuted tail's values, source coefficient x tscale x sign (null for B > 1)
    uint64_t* tl_keys;   // W planes of cap words: keys of the terms listed in tail[0], in list order
    double* tl_coef;     // B = 1: their coefficients (null for B > 1)
    uint64_t* states;    // look-back words: two chains of lb_tiles words, one word per tile, then the
                         // one-column merge's round-window words (lb_tiles more)
    uint64_t* gstates;   // look-back words: two chains of lb_groups words, one per 64 tiles
    uint32_t* splits;    // merge-path split of every merge tile boundary
    uint64_t* ssamp;     // W words per sample: sum entries 0, 64, 128, ...
    uint64_t* tsamp;     // W words per sample: permuted tail entries 63, 127, ...
    uint32_t cap;
    uint32_t lb_tiles;   // words per chain in states
    uint32_t lb_groups;  // words per chain in gstates
    // Recording for the adjoint sweep (all null when not recording):
    uint32_t* tpj;       // permuted tail entry -> its index in parent order
    uint32_t* rec;       // per-gate record words, from the gate's roff
    double* rcoef;       // per-gate coefficient rows of the touched terms, from the gate's coff
    uint32_t rec_words;  // capacity of rec
    uint32_t rec_rows;   // capacity of rcoef, in rows of B
    uint32_t batch;      // B, the coefficient columns
};

// A gate's records. Rotation touching na > 0 of n terms, at rec + roff:
//   [0, n)          term i: REC_FLAG if it anticommutes, | its output position
//   [n, n + na)     touched term j (parent order): its index | SRC_NEG sign bit
//   [n + na, +na)   touched term j: output position of its created term
// and its coefficient row at rcoef + coff * B. Noise on n terms, at rec + roff:
//   [0, n)          term i: REC_FLAG if damped, | its output position.
// A position is REC_NONE when that output was truncated or never existed.
constexpr uint32_t REC_FLAG = 0x80000000u;
constexpr uint32_t REC_NONE = 0x7fffffffu;

struct GateState {
    uint32_t n;     // live terms entering the gate
    uint32_t cur;   // buffer holding them
    uint32_t na;    // touched (anticommuting) terms, written by k_branch
    uint32_t roff;  // where the gate's records start (recording only)
    uint32_t coff;  // where its coefficient rows start
    uint32_t run;   // amplitude damping: some term has X, Y or Z on the qubit (merge runs even when na = 0)
    uint32_t pad[2];
};

struct Ctl {
    GateState st[2];
    uint32_t overflow;       // sticky: a gate needed more room than cap
    uint32_t overflow_seq;   // that gate's sequence number
    uint32_t overflow_need;  // terms it needed (overflow 1), or record words (overflow 3)
    uint32_t overflow_rows;  // record rows it needed (overflow 3)
    uint32_t err;            // internal consistency bits
    uint32_t counter[2];     // tile counters; a look-back launch uses counter[slot], clears the other
    // A merge may also list the touched terms of the gate after it (its output
    // tested against that gate's mask): tail[0] and st[].na then hold gate
    // tail_seq's touched terms when tail_ready is set and the gate's mask
    // equals tail_mask.
    uint32_t tail_ready;
    uint32_t tail_seq;
    uint64_t tail_mask[4];
    // ssamp holds every 64th entry of the input of gate ssamp_for - 1 (0: of
    // no gate's): a merge writes them for the next gate as it writes its output
    uint32_t ssamp_for;
    uint32_t lds_done;       // gates the last shared-memory run (k_lds_run) applied
};
static_assert(sizeof(Ctl) % 16 == 0, "the control block is read in 16-byte pieces");

// The control block as a kernel sees it when it starts, read in one round trip
// (all pieces requested before any is used). The sizes, buffers and lists in
// it change only between launches; the counters, err and the overflow record
// are read again where they matter.
struct CtlView {
    uint32_t overflow;
    GateState s;  // st[p] of the launching gate's parity p
    uint32_t tail_ready, tail_seq;
    uint64_t tail_mask[4];
    uint32_t ssamp_for;
};

__device__ __forceinline__ CtlView load_ctl(const Ctl* ctl, int p) {
    constexpr int NQ = (int)(sizeof(Ctl) / 16);
    const uint4* src = reinterpret_cast<const uint4*>(ctl);
    uint4 q[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) q[i] = src[i];
    Ctl c;
#pragma unroll
    for (int i = 0; i < NQ; ++i) __builtin_memcpy(reinterpret_cast<char*>(&c) + 16 * i, &q[i], 16);
    CtlView v;
    v.overflow = c.overflow;
    v.s.n = p ? c.st[1].n : c.st[0].n;
    v.s.cur = p ? c.st[1].cur : c.st[0].cur;
    v.s.na = p ? c.st[1].na : c.st[0].na;
    v.s.roff = p ? c.st[1].roff : c.st[0].roff;
    v.s.coff = p ? c.st[1].coff : c.st[0].coff;
    v.s.run = p ? c.st[1].run : c.st[0].run;
    v.tail_ready = c.tail_ready;
    v.tail_seq = c.tail_seq;
    v.ssamp_for = c.ssamp_for;
#pragma unroll
    for (int l = 0; l < 4; ++l) v.tail_mask[l] = c.tail_mask[l];
    // all fields are needed here, so their loads go out together instead of
    // being sunk past the first branch one round trip at a time
    asm volatile("" ::"s"(v.overflow), "s"(v.s.n), "s"(v.s.cur), "s"(v.s.na), "s"(v.s.roff), "s"(v.s.coff),
                 "s"(v.s.run), "s"(v.tail_ready), "s"(v.tail_seq), "s"(v.ssamp_for));
    return v;
}

struct Launch {
    uint32_t seq;         // gate sequence number (parity selects the state)
    uint32_t epoch;       // look-back epoch
    uint32_t slot;        // tile counter slot (dynamic tiles)
    uint32_t dynamic;     // 1: tiles from the atomic counter; 0: tile = blockIdx.x + k * gridDim.x
    uint32_t spin_limit;  // bound of one look-back wait with static tiles
    uint32_t flags;       // bit 0: report a static-tile wait as run out (tests the rerun path)
    uint32_t* glog;       // n_in, n_touched, n_out, roff, coff of this gate, or null
    uint64_t* ktime;      // profiling: earliest workgroup start, latest workgroup end (device clock), or null
};

// With L.ktime set, every workgroup of a kernel stores its start and its end
// on the device's constant-rate clock (s_memrealtime) at L.ktime[2 * blockIdx.x],
// and k_ktime_reduce takes the earliest start and the latest end: the kernel's
// execution time, free of the host-visible completion that event timestamps
// add to every timed dispatch (plain stores to separate words, so the timing
// adds no contention). Every return of a kernel that holds one is
// workgroup-uniform (the end stamp waits at a barrier).
struct KTimer {
    uint64_t* kt;
    __device__ __forceinline__ explicit KTimer(uint64_t* k) : kt(k) {
        if (kt && threadIdx.x == 0) kt[2 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();
    }
    __device__ __forceinline__ ~KTimer() {
        if (kt) {
            __syncthreads();
            if (threadIdx.x == 0) kt[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
        }
    }
};

// out[0] = earliest start, out[1] = latest end over the grid's stamps (one workgroup)
__global__ __launch_bounds__(WG) void k_ktime_reduce(const uint64_t* __restrict__ st, unsigned grid,
                                                     uint64_t* __restrict__ out) {
    __shared__ uint64_t s_lo[NWAVES], s_hi[NWAVES];
    uint64_t lo = ~0ull, hi = 0ull;
    for (unsigned b = threadIdx.x; b < grid; b += WG) {
        lo = min(lo, st[2 * b]);
        hi = max(hi, st[2 * b + 1]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, (uint64_t)__shfl_xor((unsigned long long)lo, o, 64));
        hi = max(hi, (uint64_t)__shfl_xor((unsigned long long)hi, o, 64));
    }
    if (lane_id() == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NWAVES; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        out[0] = lo;
        out[1] = hi;
    }
}

// p[0 .. n) = v (grid-stride): clears look-back words inside a captured graph,
// where a memset node is not used
__global__ __launch_bounds__(WG) void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) p[i] = v;
}

__global__ void k_ktime_init(uint64_t* kt, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kt[2 * i] = ~0ull;
        kt[2 * i + 1] = 0ull;
    }
}

constexpr int LOG_WORDS = 5;

// Tile ids of a look-back kernel. Static tiles (tile = blockIdx.x + k * gridDim.x)
// need no atomics; they rely on every workgroup of the grid being resident at
// once, which the host arranges by sizing the grid from the occupancy of the
// kernel. Each wait is bounded: if it runs out (workgroups were not resident),
// the gate is abandoned and the host runs it again with dynamic tiles, which
// are taken from an atomic counter in the order workgroups start and so always
// find their predecessors running.
struct TileSource {
    uint32_t cur = 0;
    bool started = false;
    __device__ __forceinline__ uint32_t take(const Launch& L, uint32_t* ctr, uint32_t* s_slot) {
        __syncthreads();  // the previous tile's shared memory is free again
        if (!L.dynamic) {
            cur = started ? cur + gridDim.x : blockIdx.x;
            started = true;
            return cur;
        }
        if (threadIdx.x == 0) *s_slot = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        return *s_slot;
    }
};

// Exclusive prefix of this tile through the look-back; a wait that runs out
// stops every later kernel and records the gate for the host to run again.
template <int NV>
__device__ __forceinline__ void lb_prefix_n(Ctl* ctl, const Launch& L, const Bufs& bf, uint32_t tile,
                                            const uint64_t (&agg)[NV], uint64_t (&excl)[NV]) {
    bool aborted = false;
    lb_exclusive<NV>(bf.states, bf.lb_tiles, bf.gstates, bf.lb_groups, tile, L.epoch, agg,
                     L.dynamic ? 0u : L.spin_limit, &aborted, excl);
    if (aborted && lane_id() == 0) {
        ctl->overflow_seq = L.seq;
        __threadfence();
        atomicMax(&ctl->overflow, 2u);
    }
}

// lb_prefix_n in two halves, for a kernel that loads its next tile between them
template <int NV>
__device__ __forceinline__ void lb_publish_n(const Launch& L, const Bufs& bf, uint32_t tile,
                                             const uint64_t (&agg)[NV]) {
    lb_publish<NV>(bf.states, bf.lb_tiles, tile, L.epoch, agg);
}

template <int NV>
__device__ __forceinline__ void lb_resolve_n(Ctl* ctl, const Launch& L, const Bufs& bf, uint32_t tile,
                                             const uint64_t (&agg)[NV], uint64_t (&excl)[NV]) {
    bool aborted = false;
    lb_resolve<NV>(bf.states, bf.lb_tiles, bf.gstates, bf.lb_groups, tile, L.epoch, agg,
                   L.dynamic ? 0u : L.spin_limit, &aborted, excl);
    if (aborted && lane_id() == 0) {
        ctl->overflow_seq = L.seq;
        __threadfence();
        atomicMax(&ctl->overflow, 2u);
    }
}

__device__ __forceinline__ uint64_t lb_prefix(Ctl* ctl, const Launch& L, const Bufs& bf, uint32_t tile,
                                              uint64_t agg) {
    const uint64_t a[1] = {agg};
    uint64_t e[1];
    lb_prefix_n<1>(ctl, L, bf, tile, a, e);
    return e[0];
}

// Prefix of a tile's counts for look-back kernels, called by every thread of
// the workgroup; counts are packed (second chain << 32 | first chain), each
// tile's below 1 << 16.
//
// Static tiles (tile = blockIdx.x + k * G, G = gridDim.x): the workgroup keeps
// its inclusive prefix after its previous tile in `incl`; the exclusive prefix
// of `tile` is that plus the counts of the G - 1 tiles between, which run at
// the same time: every thread reads up to four of their words per poll (one
// round trip for G <= 1025), where a chained look-back needs one per level.
// A word carries the launch epoch, so words of older launches read as not
// ready. Dynamic tiles use the two-level look-back (NV chains).
__device__ __forceinline__ uint64_t* win_words(const Bufs& bf) { return bf.states + 2 * (size_t)bf.lb_tiles; }

template <int NV>
__device__ __forceinline__ void tile_publish(const Launch& L, const Bufs& bf, uint32_t tile, uint64_t agg) {
    const int t = threadIdx.x;
    if (L.dynamic) {
        if (t < 64) {
            uint64_t a[NV];
            a[0] = agg & 0xffffffffull;
            if (NV > 1) a[NV - 1] = agg >> 32;
            lb_publish_n<NV>(L, bf, tile, a);
        }
    } else if (t == 0) {
        const uint64_t ep = (uint64_t)(L.epoch & LB_EPOCH_MASK) << 40;
        lb_store(win_words(bf) + tile, (1ull << 63) | ep | ((agg >> 32) << 16) | (agg & 0xffffull));
    }
}

template <int NV>
__device__ __forceinline__ uint64_t tile_resolve(Ctl* ctl, const Launch& L, const Bufs& bf, uint32_t tile, uint64_t agg,
                                 uint64_t& incl, uint64_t* s_wacc) {
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    if (L.dynamic) {
        if (wave == 0) {
            uint64_t a[NV], e[NV];
            a[0] = agg & 0xffffffffull;
            if (NV > 1) a[NV - 1] = agg >> 32;
            lb_resolve_n<NV>(ctl, L, bf, tile, a, e);
            if (lane == 0) s_wacc[0] = NV > 1 ? (e[0] | (e[NV - 1] << 32)) : e[0];
        }
        __syncthreads();
        const uint64_t ex = s_wacc[0];
        __syncthreads();
        return ex;
    }
    const uint32_t ep32 = L.epoch & LB_EPOCH_MASK;
    const uint32_t G = gridDim.x;
    const uint32_t base = tile >= G ? tile - G + 1 : 0u;
    const uint32_t cnt = tile - base;
    const uint64_t* __restrict__ wb = win_words(bf) + base;
    uint64_t acc = 0;
    uint32_t spins = 0;
    bool stop = false;
    for (uint32_t b0 = 0; b0 < cnt && !stop; b0 += 4 * WG) {
        uint32_t pend = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) pend |= (b0 + t + (uint32_t)u * WG < cnt) ? (1u << u) : 0u;
        while (true) {
            uint64_t w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) w[u] = ((pend >> u) & 1u) ? lb_load(wb + (b0 + t + (uint32_t)u * WG)) : 0ull;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (((pend >> u) & 1u) && (w[u] >> 63) && ((uint32_t)(w[u] >> 40) & LB_EPOCH_MASK) == ep32) {
                    acc += (((w[u] >> 16) & 0xffffull) << 32) | (w[u] & 0xffffull);
                    pend &= ~(1u << u);
                }
            }
            if (!__syncthreads_or(pend != 0u)) break;
            if (L.spin_limit && ++spins > L.spin_limit) {
                // stop; the host runs the gate again with tiles claimed in start order
                if (t == 0) {
                    ctl->overflow_seq = L.seq;
                    __threadfence();
                    atomicMax(&ctl->overflow, 2u);
                }
                stop = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if ((L.flags & 1u) && t == 0) {
        ctl->overflow_seq = L.seq;
        __threadfence();
        atomicMax(&ctl->overflow, 2u);
    }
    acc = wave_sum_u64_dpp(acc);
    if (lane == 0) s_wacc[wave] = acc;
    __syncthreads();
    uint64_t tot = 0;
#pragma unroll
    for (int v = 0; v < NWAVES; ++v) tot += s_wacc[v];
    __syncthreads();  // s_wacc is free again
    const uint64_t ex = incl + tot;
    incl = ex + agg;
    return ex;
}

template <int NV>
__device__ __forceinline__ uint64_t tile_prefix(Ctl* ctl, const Launch& L, const Bufs& bf, uint32_t tile,
                                                uint64_t agg, uint64_t& incl, uint64_t* s_wacc) {
    tile_publish<NV>(L, bf, tile, agg);
    return tile_resolve<NV>(ctl, L, bf, tile, agg, incl, s_wacc);
}

// a[i] for a small register array and a runtime index, without spilling it
template <int N>
__device__ __forceinline__ double pick(const double (&a)[N], int i) {
    double r = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (i == k) ? a[k] : r;
    return r;
}

// common start of a look-back kernel: skip everything after an overflow;
// clear the counter the next look-back launch will use
__device__ __forceinline__ bool lb_start(Ctl* ctl, const Launch& L, const CtlView& c) {
    if (c.overflow) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->counter[L.slot ^ 1u] = 0u;
    return true;
}

__device__ __forceinline__ bool lb_start(Ctl* ctl, const Launch& L) {
    if (ctl->overflow) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->counter[L.slot ^ 1u] = 0u;
    return true;
}

__device__ __forceinline__ void set_state(Ctl* ctl, int p, uint32_t n, uint32_t cur) {
    ctl->st[p].n = n;
    ctl->st[p].cur = cur;
    ctl->st[p].na = 0u;
    ctl->st[p].run = 0u;
}

// Whether the merge of the gate in state p has work: created terms, or (for
// amplitude damping) terms whose coefficients change.
__device__ __forceinline__ bool merge_runs(const Ctl* ctl, int p) { return ctl->st[p].na || ctl->st[p].run; }

// The records of a gate do not fit: stop, and let the host grow them and run
// the gate again.
__device__ __forceinline__ void record_overflow(Ctl* ctl, const Launch& L, uint64_t words, uint64_t rows) {
    ctl->overflow_seq = L.seq;
    ctl->overflow_need = words > 0xffffffffull ? 0xffffffffu : (uint32_t)words;
    ctl->overflow_rows = rows > 0xffffffffull ? 0xffffffffu : (uint32_t)rows;
    __threadfence();
    atomicMax(&ctl->overflow, 3u);
}

// The next gate's records start where this gate's start (it records nothing).
__device__ __forceinline__ void carry_records(Ctl* ctl, const Launch& L, int p) {
    ctl->st[p ^ 1].roff = ctl->st[p].roff;
    ctl->st[p ^ 1].coff = ctl->st[p].coff;
    if (L.glog) {
        L.glog[3] = ctl->st[p].roff;
        L.glog[4] = ctl->st[p].coff;
    }
}

__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t incl = wave_incl_sum_u32(v);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; ++w) {
        pre += (w < wave) ? s_w[w] : 0u;
        tot += s_w[w];
    }
    total = tot;
    __syncthreads();
    return pre + incl - v;
}

// block_excl_sum for a packed pair of small counts (each below 1 << BITS:
// low half and high half of v): the wave's prefix is read off one ballot per
// bit instead of a shuffle scan.
template <int BITS>
__device__ __forceinline__ uint32_t block_excl_sum_small(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint64_t m = __ballot((v >> (16 * h + b)) & 1u);
            pre += (uint32_t)__popcll(m & lt) << (16 * h + b);
            tot += (uint32_t)__popcll(m) << (16 * h + b);
        }
    }
    if (lane == 0) s_w[wave] = tot;
    __syncthreads();
    uint32_t base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; ++w) {
        base += (w < wave) ? s_w[w] : 0u;
        all += s_w[w];
    }
    total = all;
    __syncthreads();
    return base + pre;
}

// ============================================================ k_branch

constexpr int K1_ITEMS = 4;  // 8 and 16 measured slower (most on batched sums, where every rotation branches)
constexpr int K1_TILE = WG * K1_ITEMS;

// End of the branch step, once the touched count na is known: check room for
// the na created terms, or hand the sum on unchanged when nothing is touched
// (with its samples, when they are this gate's: ssamp_for as read at start).
__device__ __forceinline__ void branch_finish(Ctl* ctl, const Launch& L, const Bufs& bf, int p, uint32_t n,
                                              uint32_t cur, uint32_t na, const GateState& sp, uint32_t ssamp_for) {
    ctl->st[p].na = na;
    const uint32_t roff = sp.roff, coff = sp.coff;
    // a rotation touching nothing is the identity: no records
    const uint64_t words = (bf.rec && na) ? (uint64_t)n + 2ull * na : 0ull;
    const uint64_t rows = (bf.rec && na) ? (uint64_t)na : 0ull;
    if (L.glog) {
        L.glog[0] = n;
        L.glog[1] = na;
        L.glog[2] = n;
        L.glog[3] = roff;
        L.glog[4] = coff;
    }
    if ((uint64_t)n + na > (uint64_t)bf.cap) {
        ctl->overflow_seq = L.seq;
        const uint64_t need = (uint64_t)n + na;
        ctl->overflow_need = need > 0xffffffffull ? 0xffffffffu : (uint32_t)need;
        __threadfence();
        ctl->overflow = 1u;
    } else if (roff + words > bf.rec_words || coff + rows > bf.rec_rows) {
        record_overflow(ctl, L, roff + words, coff + rows);
    } else {
        set_state(ctl, p ^ 1, n, cur);  // final when nothing anticommutes (or is damped)
        ctl->st[p ^ 1].roff = (uint32_t)(roff + words);
        ctl->st[p ^ 1].coff = (uint32_t)(coff + rows);
        if (na == 0 && ssamp_for == L.seq + 1u) ctl->ssamp_for = L.seq + 2u;
    }
}

// Compaction of the terms that anticommute with `mask` (DAMP: that hold Z
// where the mask does; COMM: that commute with it) into tail[0], in parent order, with SRC_NEG where the
// created term's coefficient is negated, and the terms themselves into
// tl_keys / tl_coef; tiles of K1_TILE terms through the look-back. fin(count)
// runs once, on the last tile, after every earlier tile's count is in.
template <int W, bool DAMP, bool COMM, typename Fin>
__device__ __forceinline__ void branch_scan(const Bufs& bf, Ctl* ctl, const Launch& L, const Key<W>& mask, int p,
                                            uint32_t n, uint32_t cur, Fin fin) {
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const size_t stride = bf.cap;
    uint32_t* __restrict__ tail = bf.tail[0];
    const uint32_t ntiles = (n + K1_TILE - 1) / K1_TILE;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[K1_ITEMS][NWAVES];
    __shared__ uint64_t s_excl;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t base = tile * (uint32_t)K1_TILE;
        uint64_t ballots[K1_ITEMS];
        uint32_t negbits = 0;
        // Commutation parity and sign exponent add up limb by limb, so only the
        // limbs the mask touches are read, one limb of all items at a time (all
        // loads of a limb in flight together, at clamped indices).
        uint32_t antibits = 0;
        int ex[K1_ITEMS];
#pragma unroll
        for (int k = 0; k < K1_ITEMS; ++k) ex[k] = 0;
#pragma unroll
        for (int l = 0; l < W; ++l) {
            if (mask.l[l] == 0) continue;
            uint64_t v[K1_ITEMS];
#pragma unroll
            for (int k = 0; k < K1_ITEMS; ++k) v[k] = keys[l * stride + min(base + k * WG + t, n - 1)];
            const uint64_t m = mask.l[l], sm = swap_pairs(m);
            if (DAMP) {
                uint32_t any = 0;
#pragma unroll
                for (int k = 0; k < K1_ITEMS; ++k) {
                    const bool in = base + k * WG + t < n;
                    antibits |= (uint32_t)(in && (v[k] & m) == m) << k;  // Z on the qubit
                    any |= (uint32_t)(in && (v[k] & m) != 0);
                }
                if (__ballot(any != 0) && lane == 0) atomicOr(&ctl->st[p].run, 1u);
                continue;
            }
#pragma unroll
            for (int k = 0; k < K1_ITEMS; ++k) {
                antibits ^= (uint32_t)(__popcll(sm & v[k]) & 1) << k;
                ex[k] += limb_sign_exponent(m, v[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < K1_ITEMS; ++k) {
            const uint32_t i = base + k * WG + t;
            // COMM (imaginary rotations): the commuting terms branch
            const bool f = (i < n) && (((antibits >> k) & 1u) != (COMM ? 1u : 0u));
            if (!DAMP && f && (ex[k] & 2) == 0) negbits |= 1u << k;
            ballots[k] = __ballot(f);
            if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(ballots[k]);
        }
        __syncthreads();
        uint32_t agg = 0;
        uint32_t before[K1_ITEMS];
#pragma unroll
        for (int k = 0; k < K1_ITEMS; ++k) {
            uint32_t row_before = 0, row = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; ++w) {
                row_before += (w < wave) ? s_cnt[k][w] : 0u;
                row += s_cnt[k][w];
            }
            before[k] = agg + row_before;
            agg += row;
        }
        {
            const uint64_t excl = tile_prefix<1>(ctl, L, bf, tile, agg, lb_incl, s_wacc);
            if (t == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) fin((uint32_t)(excl + agg));
            }
        }
        __syncthreads();
        const uint64_t excl = s_excl;
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        for (int k = 0; k < K1_ITEMS; ++k) {  // unrolled by the compiler (no scratch)
            if ((ballots[k] >> lane) & 1ull) {
                const uint32_t i = base + k * WG + t;
                const uint32_t pos = (uint32_t)excl + before[k] + (uint32_t)__popcll(ballots[k] & lt);
                tail[pos] = i | (((negbits >> k) & 1u) ? SRC_NEG : 0u);
                // the listed term itself, for the permutation
#pragma unroll
                for (int l = 0; l < W; ++l) bf.tl_keys[l * stride + pos] = keys[l * stride + i];
                if (bf.tl_coef) bf.tl_coef[pos] = bf.coeffs[cur][i];
            }
        }
    }
}

// DAMP: amplitude damping on the qubit whose Z the mask holds (the rule of
// src/Propagation/specializations.jl:258-276): the listed terms are those with
// Z there (they branch into I with gamma * c), and any term with X, Y or Z
// there sets st[p].run, since its coefficient changes.
//
// COMM: an imaginary rotation (src/Propagation/specializations.jl:111-131):
// the terms that commute with the generator are listed, and st[p].run is set
// so that the merge runs (it normalizes and truncates) even when none does.
template <int W, bool DAMP, bool COMM = false>
__global__ __launch_bounds__(WG) void k_branch(Bufs bf, Ctl* ctl, Launch L, Key<W> mask) {
    // every argument the early exits read, loaded with the control block
#pragma unroll
    for (int l = 0; l < W; ++l) asm volatile("" ::"s"(mask.l[l]));
    asm volatile("" ::"s"(bf.cap), "s"(bf.rec_words), "s"(bf.rec_rows), "s"(bf.rec), "s"(L.glog));
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (!lb_start(ctl, L, c)) return;
    const int p = L.seq & 1;
    const uint32_t n = c.s.n, cur = c.s.cur;
    const int t = threadIdx.x;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) {
            ctl->st[p].na = 0u;
            set_state(ctl, p ^ 1, 0u, cur);
            if (L.glog) L.glog[0] = L.glog[1] = L.glog[2] = 0u;
            carry_records(ctl, L, p);
        }
        return;
    }
    if (COMM && blockIdx.x == 0 && t == 0) ctl->st[p].run = 1u;
    // the touched terms may already be listed, by the previous gate's merge
    // or (when that gate touched nothing) by its permutation kernel
    bool listed = !DAMP && !COMM && c.tail_ready && c.tail_seq == L.seq;
#pragma unroll
    for (int l = 0; l < W; ++l) listed = listed && c.tail_mask[l] == mask.l[l];
    if (listed) {
        if (blockIdx.x == 0 && t == 0) branch_finish(ctl, L, bf, p, n, cur, c.s.na, c.s, c.ssamp_for);
        return;
    }
    branch_scan<W, DAMP, COMM>(bf, ctl, L, mask, p, n, cur,
                               [&](uint32_t na) { branch_finish(ctl, L, bf, p, n, cur, na, c.s, c.ssamp_for); });
}

// ============================================================ k_xorperm

// list entries per thread of the tail permutation: the default 2 measured
// fastest on the kicked Ising workload (1, 3, 4 and 8 slower); 1 and 4 are
// compiled too, for the tuning table
constexpr int K2_ITEMS = 2;
constexpr int K2_TILE = WG * K2_ITEMS;
constexpr int INF_POS = 0x7fffffff;

struct Levels {
    int n;                   // number of levels
    int top_runs;            // runs reaching the top bit: each adds na to every rank
    int from_limb;           // lowest limb any level compares
    int bit[MAX_LEVELS];     // compare from this bit up
    int sgn[MAX_LEVELS];     // +1 for a run's hi+1 level, -1 for its lo level
};

// smallest x in [lo, hi] with fn(x) true, where fn is false...false true...true
// on [lo, hi) and fn(hi) is taken as true. All 64 lanes call it together.
template <typename F>
__device__ __forceinline__ uint32_t wave_first_true(uint32_t lo, uint32_t hi, F fn) {
    const int lane = lane_id();
    while (lo < hi) {
        const uint32_t span = hi - lo;
        const uint32_t step = (span + 63) / 64;
        const uint64_t xl = (uint64_t)lo + (uint64_t)lane * step;
        // evaluated at a clamped point on every lane so the loads are not branched around
        const bool pc = fn(xl < hi ? (uint32_t)xl : hi - 1);
        const bool p = (xl >= hi) ? true : pc;
        const uint64_t m = __ballot(p);
        if (m == 0) {
            lo = lo + 63 * step + 1;
        } else {
            const int f = __ffsll((unsigned long long)m) - 1;
            if (f == 0) {
                hi = lo;
            } else {
                const uint64_t xf = (uint64_t)lo + (uint64_t)f * step;
                lo = lo + (uint32_t)(f - 1) * step + 1;
                hi = (uint32_t)(xf < hi ? xf : hi);
            }
        }
    }
    return lo;
}

// First x in [lo, hi) with fn(x) true, for fn false-then-true on [lo, hi]
// and fn(hi - 1) known true. Probes backward from hi - 1 with strides that
// grow 64-fold per round, so an edge d entries away costs about log64(d) + 1
// rounds of probes near hi rather than a search over all of [lo, hi).
template <typename F>
__device__ __forceinline__ uint32_t wave_first_true_back(uint32_t lo, uint32_t hi, F fn) {
    const int lane = lane_id();
    int64_t known = (int64_t)hi - 1;  // fn(known) is true
    int64_t stride = 1;
    while (true) {
        const int64_t x = known - stride * (lane + 1);
        const bool inr = x >= (int64_t)lo;
        const bool pc = fn(inr ? (uint32_t)x : lo);
        const uint64_t m = __ballot(!(inr && pc));
        if (m) {
            const int f = __ffsll((unsigned long long)m) - 1;
            const int64_t xf = known - stride * (f + 1);  // false, or below lo
            const int64_t xt = known - stride * f;        // true
            const uint32_t a = (uint32_t)(xf + 1 > (int64_t)lo ? xf + 1 : (int64_t)lo);
            return wave_first_true(a, (uint32_t)xt, fn);
        }
        known -= stride * 64;
        stride *= 64;
    }
}

// First x in [lo, hi) with fn(x) true, or hi, for fn false-then-true and
// fn(lo) known false: the forward counterpart of wave_first_true_back.
template <typename F>
__device__ __forceinline__ uint32_t wave_first_true_fwd(uint32_t lo, uint32_t hi, F fn) {
    const int lane = lane_id();
    int64_t known = lo;  // fn(known) is false
    int64_t stride = 1;
    while (true) {
        const int64_t x = known + stride * (lane + 1);
        const bool inr = x < (int64_t)hi;
        const bool pc = fn(inr ? (uint32_t)x : hi - 1);
        const uint64_t m = __ballot(!inr || pc);
        if (m) {
            const int f = __ffsll((unsigned long long)m) - 1;
            const int64_t xf = known + stride * f;        // false
            const int64_t xt = known + stride * (f + 1);  // true, or at/after hi
            const uint32_t b = (uint32_t)(xt < (int64_t)hi ? xt : (int64_t)hi);
            return wave_first_true((uint32_t)xf + 1, b, fn);
        }
        known += stride * 64;
        stride *= 64;
    }
}

// The permutation of the touched terms into ascending order of term ^ mask
// (rank formula above), the tail keys, values and partition samples. With one
// coefficient column a created term's value (source coefficient times tscale,
// the gate's sin, sinh or gamma, negated for a negative sign) is stored in
// tcoef and its index is not needed; with B columns the index goes to tail[1].
//
// BRANCH: the gate's k_branch was not launched, because the touched terms are
// listed already (by the previous gate's merge, or by this kernel of the
// previous gate); this kernel then does the branch step's end (room check,
// state of an identity gate), on a look-back launch. If the gate touches
// nothing and the next gate is a rotation (has_next, its mask nmask), the
// kernel lists that rotation's touched terms instead, so that its k_branch
// is not needed either. check: report (err bit 4) a list that is not this
// gate's.
template <int W, bool BRANCH, int K2_ITEMS = pp::K2_ITEMS>
__global__ __launch_bounds__(WG) void k_xorperm(Bufs bf, Ctl* ctl, Launch L, Levels lv, Key<W> mask, Key<W> nmask,
                                                uint32_t has_next, uint32_t check, double tscale) {
    constexpr int K2_TILE = WG * K2_ITEMS;
    // The list is read only once its length is known: a first tile loaded
    // with the control block, as the merge loads its splits, makes every
    // workgroup of the capacity-sized grid read a full tile, which on the
    // many gates with short lists is tens of MB of stale entries.
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (c.overflow) return;
    const int p = L.seq & 1;
    const uint32_t na = c.s.na;
    if (BRANCH) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->counter[L.slot ^ 1u] = 0u;
            if (check) {
                bool listed = c.tail_ready && c.tail_seq == L.seq;
#pragma unroll
                for (int l = 0; l < W; ++l) listed = listed && c.tail_mask[l] == mask.l[l];
                if (!listed) atomicOr(&ctl->err, 4u);
            }
        }
        const uint32_t n = c.s.n, cur = c.s.cur;
        if (na == 0 && has_next) {
            // the identity: the state passes on, with the next rotation's list
            auto fin = [&](uint32_t nt) {
                set_state(ctl, p ^ 1, n, cur);
                ctl->st[p ^ 1].na = nt;
                ctl->st[p ^ 1].roff = c.s.roff;
                ctl->st[p ^ 1].coff = c.s.coff;
#pragma unroll
                for (int l = 0; l < W; ++l) ctl->tail_mask[l] = nmask.l[l];
                ctl->tail_seq = L.seq + 1;
                ctl->tail_ready = 1u;
                if (c.ssamp_for == L.seq + 1u) ctl->ssamp_for = L.seq + 2u;
                if (L.glog) {
                    L.glog[0] = n;
                    L.glog[1] = 0u;
                    L.glog[2] = n;
                    L.glog[3] = c.s.roff;
                    L.glog[4] = c.s.coff;
                }
            };
            if (n == 0) {
                if (blockIdx.x == 0 && threadIdx.x == 0) fin(0u);
            } else {
                branch_scan<W, false, false>(bf, ctl, L, nmask, p, n, cur, fin);
            }
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) branch_finish(ctl, L, bf, p, n, cur, na, c.s, c.ssamp_for);
    }
    if (na == 0) return;
    const uint64_t* __restrict__ keys = bf.keys[c.s.cur];
    const size_t stride = bf.cap;
    const uint32_t* __restrict__ tail_in = bf.tail[0];
    uint32_t* __restrict__ tail_out = bf.tail[1];
    const int from_limb = lv.from_limb;
    const int64_t const_off = (int64_t)lv.top_runs * na;
    const uint32_t ntiles = (na + K2_TILE - 1) / K2_TILE;
    __shared__ Key<W> s_last[WG];
    __shared__ Key<W> s_kb, s_kl, s_kp, s_kn;
    __shared__ int s_edge_s[MAX_LEVELS], s_edge_e[MAX_LEVELS];
    __shared__ int s_wmax[MAX_LEVELS][NWAVES], s_wmin[MAX_LEVELS][NWAVES];  // per level and wave
    __shared__ int s_lvbit[MAX_LEVELS], s_lvsgn[MAX_LEVELS];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    // the levels go to shared memory with constant indices: indexing the kernel
    // argument by a loop variable would copy it to scratch memory
#pragma unroll
    for (int v = 0; v < MAX_LEVELS; ++v)
        if (t == v) {
            s_lvbit[v] = lv.bit[v];
            s_lvsgn[v] = lv.sgn[v];
        }
    __syncthreads();

    // list entry x's key (the list holds the terms themselves in list order)
    const uint64_t* __restrict__ lkeys = bf.tl_keys;
    // recording: the list's records (its entries and their source coefficient
    // rows) are written here, when they fit, under the same bound the branch
    // step checks. With one column the list carries the coefficients and each
    // tile writes its entries'; with B columns every workgroup of the grid,
    // most of which have no tile, copies a grid-strided share of the rows
    // from the sum (after its tile work, if any)
    const bool rec = bf.rec && (uint64_t)c.s.roff + c.s.n + 2ull * na <= bf.rec_words &&
                     (uint64_t)c.s.coff + na <= bf.rec_rows;
    const bool rec1 = rec && bf.tcoef;
    uint32_t* __restrict__ rsrc = rec ? bf.rec + (size_t)c.s.roff + c.s.n : nullptr;
    double* __restrict__ rcf = rec ? bf.rcoef + (size_t)c.s.coff * bf.batch : nullptr;
    auto key_at = [&](uint32_t x) -> Key<W> { return load_key_from<W>(lkeys, stride, x, from_limb); };

    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t base = tile * (uint32_t)K2_TILE;
        const uint32_t cnt = min((uint32_t)K2_TILE, na - base);
        const uint32_t last = base + cnt - 1;
        uint32_t src[K2_ITEMS];
        Key<W> key[K2_ITEMS];
        double cf[K2_ITEMS];  // the source coefficients the merge will scale
        const uint32_t j0 = base + t * K2_ITEMS;
        int nvalid = 0;
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) {
            const uint32_t x = min(j0 + i, last);
            src[i] = tail_in[x];
            key[i] = load_key<W>(lkeys, stride, x);
            cf[i] = bf.tl_coef ? bf.tl_coef[x] : 0.0;
        }
        if (rec1) {
#pragma unroll
            for (int i = 0; i < K2_ITEMS; ++i) {
                if (j0 + i <= last) {
                    rsrc[j0 + i] = src[i];
                    rcf[j0 + i] = cf[i];
                }
            }
        }
        // the entries just before and after the tile, loaded with it for every level's edges
        const bool has_prev = base > 0, has_after = last + 1 < na;
        Key<W> kout{};
        if (t == 0 && has_prev) kout = key_at(base - 1);
        else if (t == WG - 1 && has_after) kout = key_at(last + 1);
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) nvalid += (j0 + i <= last) ? 1 : 0;
        // the last valid item, selected with constant indices (a variable index
        // into the register array would put it in scratch memory)
        Key<W> klast = key[0];
#pragma unroll
        for (int i = 1; i < K2_ITEMS; ++i)
            if (i < nvalid) klast = key[i];
        if (nvalid > 0) s_last[t] = klast;
        // tile edges: the block around `base` and the block around `last`, per level
        if (t == 0) s_kb = key[0];
        if (nvalid > 0 && j0 + nvalid - 1 == last) s_kl = klast;
        if (t == 0 && has_prev) s_kp = kout;
        if (t == WG - 1 && has_after) s_kn = kout;
        __syncthreads();
        const Key<W> kb = s_kb, kl = s_kl;
        for (int task = wave; task < 2 * lv.n; task += NWAVES) {
            const int v = task >> 1;
            const int bit = s_lvbit[v];
            if ((task & 1) == 0) {
                uint32_t s = 0;
                if (has_prev) {
                    const Key<W> kp = s_kp;
                    if (!eq_from<W>(kp, kb, bit)) s = base;
                    else s = wave_first_true_back(0u, base, [&](uint32_t x) { return eq_from<W>(key_at(x), kb, bit); });
                }
                if (lane == 0) s_edge_s[v] = (int)s;
            } else {
                uint32_t e = na;
                if (has_after) {
                    const Key<W> kn = s_kn;
                    if (!eq_from<W>(kn, kl, bit)) e = last + 1;
                    else e = wave_first_true_fwd(last + 1, na, [&](uint32_t x) { return !eq_from<W>(key_at(x), kl, bit); });
                }
                if (lane == 0) s_edge_e[v] = (int)e;
            }
        }
        // previous entry of this thread's first item
        Key<W> prev;
        if (t > 0) prev = s_last[t - 1];
        else if (has_prev) prev = s_kp;
        else prev = key[0];
        __syncthreads();

        int64_t dest[K2_ITEMS];
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) dest[i] = (int64_t)(j0 + i) + const_off;

        for (int v = 0; v < lv.n; ++v) {
            const int bit = s_lvbit[v];
            bool h[K2_ITEMS];
#pragma unroll
            for (int i = 0; i < K2_ITEMS; ++i) {
                const uint32_t j = j0 + i;
                const Key<W>& pk = (i == 0) ? prev : key[i - 1];
                h[i] = (i < nvalid) && (j == 0 || !eq_from<W>(key[i], pk, bit));
            }
            int sl[K2_ITEMS];
            int run = -1;
#pragma unroll
            for (int i = 0; i < K2_ITEMS; ++i) {
                if (h[i]) run = (int)(j0 + i);
                sl[i] = run;
            }
            int el[K2_ITEMS];
            int nxt = INF_POS;
#pragma unroll
            for (int i = K2_ITEMS - 1; i >= 0; --i) {
                el[i] = nxt;
                if (h[i]) nxt = (int)(j0 + i);
            }
            // exclusive max-scan of `run` over threads, exclusive suffix-min of
            // `nxt`. Head positions grow with the lane, so the max over earlier
            // lanes is the `run` of the nearest earlier lane holding a head, and
            // the min over later lanes the `nxt` of the nearest later one: one
            // ballot and one lane read each. Both waves' totals go out together,
            // in this level's own slots, so one barrier serves both and none is
            // needed after.
            const uint64_t hm = __ballot(run >= 0);
            const uint64_t below = hm & ((1ull << lane) - 1ull);
            const uint64_t above = lane == 63 ? 0ull : (hm & (~0ull << (lane + 1)));
            const int lo_src = below ? 63 - __clzll((long long)below) : 0;
            const int hi_src = above ? __ffsll((long long)above) - 1 : 0;
            int wex = __shfl(run, lo_src, 64);
            int sex = __shfl(nxt, hi_src, 64);
            if (!below) wex = -1;
            if (!above) sex = INF_POS;
            // the wave's totals, read with every lane active
            const int wmax = __shfl(run, hm ? 63 - __clzll((long long)hm) : 0, 64);
            const int wmin = __shfl(nxt, hm ? __ffsll((long long)hm) - 1 : 0, 64);
            if (lane == 0) {
                s_wmax[v][wave] = hm ? wmax : -1;
                s_wmin[v][wave] = hm ? wmin : INF_POS;
            }
            __syncthreads();
            int pre = wex;
            for (int w = 0; w < wave; ++w) pre = max(pre, s_wmax[v][w]);
            int suf = sex;
            for (int w = wave + 1; w < NWAVES; ++w) suf = min(suf, s_wmin[v][w]);
            const int es = s_edge_s[v], ee = s_edge_e[v];
            const int sg = s_lvsgn[v];
#pragma unroll
            for (int i = 0; i < K2_ITEMS; ++i) {
                const int s = sl[i] >= 0 ? sl[i] : (pre >= 0 ? pre : es);
                const int e = el[i] < INF_POS ? el[i] : (suf < INF_POS ? suf : ee);
                dest[i] += (int64_t)sg * ((int64_t)s + (int64_t)e);
            }
        }
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) {
            if (i < nvalid) {
                // a rank outside [0, na) would be a bug; it is reported, never written
                if (dest[i] >= 0 && dest[i] < (int64_t)na) {
                    if (bf.tcoef) {
                        // one column: the created term's value itself, so the
                        // merge reads neither the index nor the sign
                        double v = cf[i] * tscale;
                        if (src[i] & SRC_NEG) v = -v;
                        bf.tcoef[dest[i]] = v;
                    } else {
                        tail_out[dest[i]] = src[i];
                    }
                    if (bf.tpj) bf.tpj[dest[i]] = j0 + i;
#pragma unroll
                    for (int l = 0; l < W; ++l) bf.tkeys[l * stride + dest[i]] = key[i].l[l] ^ mask.l[l];
                    if ((dest[i] & 63) == 63) {
#pragma unroll
                        for (int l = 0; l < W; ++l) bf.tsamp[(dest[i] >> 6) * W + l] = key[i].l[l] ^ mask.l[l];
                    }
                } else {
                    atomicOr(&ctl->err, 1u);
                }
            }
        }
        __syncthreads();
    }
    if (rec && !rec1) {
        const uint32_t B = bf.batch;
        const double* __restrict__ sc = bf.coeffs[c.s.cur];
        const uint64_t total = (uint64_t)na * B;
        for (uint64_t e = (uint64_t)blockIdx.x * WG + t; e < total; e += (uint64_t)gridDim.x * WG) {
            const uint32_t j = (uint32_t)(e / B), b = (uint32_t)(e - (uint64_t)j * B);
            const uint32_t sj = tail_in[j];
            if (b == 0) rsrc[j] = sj;
            rcf[e] = sc[(size_t)(sj & SRC_MASK) * B + b];
        }
    }
    // every 64th entry of the sum, for the merge partition, unless the merge
    // that wrote the sum left them
    if (c.ssamp_for == L.seq + 1u) return;
    const uint32_t n = c.s.n;
    for (uint32_t k = blockIdx.x * WG + t; 64ull * k < n; k += gridDim.x * WG) {
        const Key<W> kk = load_key<W>(keys, stride, 64ull * k);
#pragma unroll
        for (int l = 0; l < W; ++l) bf.ssamp[(size_t)k * W + l] = kk.l[l];
    }
}

// ============================================================ k_merge (one coefficient column)

// Values of one branching gate. Rotation: heads that anticommute times cosv,
// created terms times sign * sinv. Amplitude damping on qubit damp_q >= 0:
// heads with Z there times cosv = 1 - gamma, with X or Y times fxy =
// sqrt(1 - gamma); created terms (Z -> I) times sinv = gamma.
struct RotParams {
    double cosv, sinv;
    int max_weight;   // < 0: no weight truncation
    double min_abs;
    int damp_q = -1;
    double fxy = 1.0;
    // imaginary rotation: commuting heads take cosv (cosh), and every output
    // is multiplied by *norm (1 / the identity's merged coefficient) when
    // norm is set, before it is truncated
    int comm = 0;
    const double* norm = nullptr;
};

// number of head entries among the first d merged entries; heads go first on ties
template <int W>
__device__ __forceinline__ uint32_t merge_split(const uint64_t* __restrict__ keys, size_t stride, uint32_t n,
                                                const uint64_t* __restrict__ tkeys, uint32_t na, uint32_t d) {
    const uint32_t xlo = d > na ? d - na : 0u;
    const uint32_t xhi = min(d, n);
    return wave_first_true(xlo, xhi, [&](uint32_t x) {
        const Key<W> a = load_key<W>(keys, stride, x);
        const Key<W> b = load_key<W>(tkeys, stride, d - 1 - x);
        return !key_le<W>(a, b);
    });
}

// A batched merge tile's head range S[i0-1 .. i1) and tail range T[j0 .. j1]
// (keys, tail sources, recorded parent positions) held in registers between
// their loads and their shared-memory staging: slot x = threadIdx.x + r * WG.
template <int W, int TILE>
struct LoadRegs {
    static constexpr int R = (TILE + 1 + WG - 1) / WG;
    uint64_t hv[R][W], tv[R][W];
    uint32_t sv[R], pv[R];
};

// Issues those loads without waiting for them; indices are clamped so every
// load is issued unconditionally.
template <int W, int TILE>
__device__ __forceinline__ void merge_load_fetch(const uint64_t* __restrict__ keys, size_t stride,
                                                 const uint32_t* __restrict__ tail,
                                                 const uint64_t* __restrict__ tkeys, uint32_t na, uint32_t i0,
                                                 uint32_t j0, int ns, const uint32_t* __restrict__ tpj,
                                                 LoadRegs<W, TILE>& f) {
    const int t = threadIdx.x;
    // head slots x = 0 .. ns (global index i0 + x - 1) and tail slots y = 0 .. nt
    // (global index j0 + y)
    const size_t hmax = (size_t)i0 + ns > 0 ? (size_t)i0 + ns - 1 : 0;
    const uint32_t tmax = na > 0 ? na - 1 : 0;
#pragma unroll
    for (int r = 0; r < LoadRegs<W, TILE>::R; ++r) {
        const int x = t + r * WG;
        size_t gi = (size_t)i0 + x;
        gi = gi > 0 ? gi - 1 : 0;
        gi = gi < hmax ? gi : hmax;
        const uint32_t gj = min(j0 + (uint32_t)x, tmax);
#pragma unroll
        for (int l = 0; l < W; ++l) {
            f.hv[r][l] = keys[l * stride + gi];
            f.tv[r][l] = tkeys[l * stride + gj];
        }
        f.sv[r] = tail[gj];
        if (tpj) f.pv[r] = tpj[gj];
    }
}

// Stages fetched entries in LDS key slots: [0] = S[i0-1], [1 .. ns] =
// S[i0 .. i1), [ns+1 .. ns+1+nt] = T[j0 .. j1].
template <int W, int TILE>
__device__ __forceinline__ void merge_load_put(const LoadRegs<W, TILE>& f, uint32_t na, uint32_t i0, uint32_t j0,
                                               int ns, int nt, uint64_t (*s_key)[TILE + 2], uint32_t* s_src,
                                               const uint32_t* __restrict__ tpj, uint32_t* s_pj) {
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < LoadRegs<W, TILE>::R; ++r) {
        const int x = t + r * WG;
        if (x < ns + 1 && (x > 0 || i0 > 0)) {
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][x] = f.hv[r][l];
        }
        if (x < nt + 1 && j0 + (uint32_t)x < na) {
            s_src[x] = f.sv[r];
            if (tpj) s_pj[x] = f.pv[r];
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][ns + 1 + x] = f.tv[r][l];
        }
    }
}

// Tile of the one-column merge: by default 768 diagonals for 256-bit keys,
// 1024 otherwise, so that the tile's keys and values fit shared memory four
// times per CU; 512, 768 and 1024 are compiled for every key width, for the
// tuning table.
template <int W>
struct K3Cfg {
    static constexpr int ITEMS = W == 4 ? 3 : 4;
    static constexpr int TILE = WG * ITEMS;
};
constexpr int MIN_MERGE_ITEMS = 2;  // the narrowest merge tile any table may choose

// Merge-path split of tile boundary b of this gate's merge, by one wave: the
// head entries among the first min(b * TILE, n + na) merged entries.
//
// The split at diagonal d is the first x in [max(0, d - na), min(d, n)) with
// sum[x] > tail[d - 1 - x], or min(d, n). For d a multiple of 64 and x a
// multiple of 64, d - 1 - x is 63 modulo 64, so the sampled entries alone
// (every 64th sum entry, tail entries 63 modulo 64) locate the split to
// within 64 positions; one more probe of 64 consecutive entries finds it.
// The samples are small and shared by all boundaries, so the search reads
// little beyond them.
template <int W, int TILE>
__device__ __forceinline__ uint32_t tile_split(const Bufs& bf, const uint64_t* __restrict__ keys, uint32_t n,
                                               uint32_t na, uint32_t b) {
    static_assert(TILE % 64 == 0, "tile boundaries must fall on sampled diagonals");
    const uint32_t ntot = n + na;
    if (b == 0) return 0;
    if ((uint64_t)b * TILE >= ntot) return n;
    const uint64_t* __restrict__ ss = bf.ssamp;
    const uint64_t* __restrict__ ts = bf.tsamp;
    const size_t stride = bf.cap;
    const uint32_t d = b * (uint32_t)TILE;
    const uint32_t xlo = d > na ? d - na : 0u;
    const uint32_t xhi = min(d, n);
    // sampled points 64a in [xlo, xhi): tail partner d - 1 - 64a = 64c + 63
    const uint32_t aL = (xlo + 63) / 64, aR = (xhi + 63) / 64;
    const uint32_t cbase = d / 64 - 1;
    const uint32_t aT = wave_first_true(aL, aR, [&](uint32_t a) {
        Key<W> x, y;
#pragma unroll
        for (int l = 0; l < W; ++l) {
            x.l[l] = ss[(size_t)a * W + l];
            y.l[l] = ts[(size_t)(cbase - a) * W + l];
        }
        return !key_le<W>(x, y);
    });
    const uint32_t lo = aT > aL ? 64 * (aT - 1) + 1 : xlo;
    const uint32_t hi = aT < aR ? 64 * aT : xhi;
    return wave_first_true(lo, hi, [&](uint32_t x) {
        return !key_le<W>(load_key<W>(keys, stride, x), load_key<W>(bf.tkeys, stride, d - 1 - x));
    });
}

// Every tile boundary's split, one wave per boundary, into splits[].
template <int W, int TILE>
__global__ __launch_bounds__(WG) void k_partition(Bufs bf, Ctl* ctl, Launch L) {
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (c.overflow) return;
    const uint32_t na = c.s.na;
    if (!(c.s.na || c.s.run)) return;
    const uint32_t n = c.s.n;
    const uint64_t* __restrict__ keys = bf.keys[c.s.cur];
    const uint32_t nb = (n + na + TILE - 1) / TILE + 1;
    for (uint32_t b = blockIdx.x * NWAVES + (threadIdx.x >> 6); b < nb; b += gridDim.x * NWAVES) {
        const uint32_t split = tile_split<W, TILE>(bf, keys, n, na, b);
        if (lane_id() == 0) bf.splits[b] = split;
    }
}

// One merge tile's inputs held in registers between their loads and their
// shared-memory staging: slot s = threadIdx.x + r * WG of the staging layout
// (0 = head entry i0-1, 1..ns = head entries i0..i1-1, ns+1..ns+1+nt = tail
// entries j0..j1).
template <int W, int TILE, bool REC>
struct MergeRegs {
    static constexpr int R = (TILE + 2 + WG - 1) / WG;
    uint64_t k[R][W];
    double c[R];
    uint32_t pj[REC ? R : 1];
};

// Issues the loads of one tile's head and tail entries (head keys and
// coefficients of the sum, tail keys and stored values of the permuted list)
// without waiting for them. Every lane loads from a valid address; slots past
// the tile read a clamped entry that merge_put ignores.
template <int W, int TILE, bool REC>
__device__ __forceinline__ void merge_fetch(const Bufs& bf, const uint64_t* __restrict__ keys,
                                            const double* __restrict__ coeffs, size_t stride, uint32_t na,
                                            uint32_t i0, uint32_t j0, int ns, MergeRegs<W, TILE, REC>& f) {
    const int t = threadIdx.x;
    const size_t hmax = (size_t)i0 + ns > 0 ? (size_t)i0 + ns - 1 : 0;
    const uint32_t tmax = na > 0 ? na - 1 : 0;
#pragma unroll
    for (int r = 0; r < MergeRegs<W, TILE, REC>::R; ++r) {
        const int s = t + r * WG;
        const bool tl = s > ns && na > 0;
        size_t gi = (size_t)i0 + s;
        gi = gi > 0 ? gi - 1 : 0;
        gi = gi < hmax ? gi : hmax;
        const uint32_t gj = tl ? min(j0 + (uint32_t)(s - ns - 1), tmax) : 0u;
        const uint64_t* __restrict__ kp = tl ? bf.tkeys : keys;
        const size_t idx = tl ? (size_t)gj : gi;
#pragma unroll
        for (int l = 0; l < W; ++l) f.k[r][l] = kp[l * stride + idx];
        f.c[r] = tl ? bf.tcoef[gj] : coeffs[gi];
        if (REC) f.pj[r] = tl ? bf.tpj[gj] : 0u;
    }
}

// Stages fetched entries in shared memory with their merge values: a head
// value is the coefficient times cos when the term anticommutes (amplitude
// damping: its own factors); a tail value, the source coefficient times sin
// and negated for a negative sign (src/Propagation/specializations.jl:40), was
// stored by the permutation kernel. The merge then only adds head + tail.
template <int W, int TILE, bool REC>
__device__ __forceinline__ void merge_put(const MergeRegs<W, TILE, REC>& f, uint32_t na, uint32_t i0, uint32_t j0,
                                          int ns, int nt, const Key<W>& mask, double cosv, int damp_q, double fxy,
                                          int comm, uint64_t (*s_key)[TILE + 2], double* s_val, uint32_t* s_pj) {
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < MergeRegs<W, TILE, REC>::R; ++r) {
        const int s = t + r * WG;
        Key<W> k;
#pragma unroll
        for (int l = 0; l < W; ++l) k.l[l] = f.k[r][l];
        const double c = f.c[r];
        if (s <= ns) {
            if (s > 0 || i0 > 0) {
#pragma unroll
                for (int l = 0; l < W; ++l) s_key[l][s] = k.l[l];
                double v;
                if (damp_q >= 0) {
                    const uint32_t code = code_at<W>(k, damp_q);
                    v = code == 3u ? c * cosv : (code ? c * fxy : c);
                } else {
                    v = (anticommutes<W>(mask, k) != (comm != 0)) ? c * cosv : c;
                }
                s_val[s] = v;
            }
        } else {
            const int y = s - ns - 1;
            if (y <= nt && j0 + (uint32_t)y < na) {
#pragma unroll
                for (int l = 0; l < W; ++l) s_key[l][s] = k.l[l];
                s_val[s] = c;
                if (REC) s_pj[y] = f.pj[r];
            }
        }
    }
}

// With EMIT, the kept output terms are also tested against `nmask`, the mask
// of the next gate (a rotation), and the touched ones are listed in tail[0] in
// output order, as k_branch of that gate would list them: the next k_branch
// then only checks room. The list's offsets come from a second prefix chain
// that shares every poll with the output offsets.
//
// With REC, every head term and every touched term's created term records its
// output position (REC_NONE when truncated, merged away or absent).
//
// Tiles are software-pipelined: a workgroup orders a tile's kept outputs in
// shared memory, publishes its counts, issues the loads of its next tile, and
// only then waits for the tile's prefix and stores it, so the wait for the
// tiles before it overlaps the next tile's loads. With static tiles (tile =
// blockIdx.x + k * G) the workgroup keeps its inclusive prefix after each
// tile: the exclusive prefix of its next tile is that plus the counts of the
// G - 1 tiles in between, which every thread of the workgroup reads at once
// (up to 8 words each) -- one round trip, where a chained look-back takes one
// per level. Dynamic tiles (the fallback after a wait runs out) use the
// two-level look-back.
template <int W, bool EMIT, bool REC, int ITEMS = K3Cfg<W>::ITEMS, bool SELF = false>
// (at least three waves per SIMD for tiles of up to 768 diagonals, the register
// budget of three resident workgroups per compute unit)
__global__ __launch_bounds__(WG, (ITEMS <= 3 ? 3 : 2)) void k_merge1(Bufs bf, Ctl* ctl, Launch L, Key<W> mask, RotParams rp,
                                               Key<W> nmask) {
    const KTimer ktimer(L.ktime);
    constexpr int TILE = WG * ITEMS;
    static_assert(TILE < 32768, "kept positions carry a sign in bit 15; counts are packed in 16 bits");
    // the first static tile's splits go out with the control block (tile =
    // blockIdx.x; the index is clamped into the array, and used only if the
    // tile is that one)
    const uint32_t pt = min((uint32_t)blockIdx.x, bf.lb_tiles - 2);
    const uint32_t pre0 = bf.splits[pt], pre1 = bf.splits[pt + 1];
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (!lb_start(ctl, L, c)) return;
    const int p = L.seq & 1;
    const uint32_t na = c.s.na;
    if (!(na || c.s.run)) return;
    const uint32_t n = c.s.n, cur = c.s.cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const double inv = rp.norm ? *rp.norm : 1.0;
    const size_t stride = bf.cap;
    const uint32_t ntot = n + na;
    const uint32_t ntiles = (ntot + TILE - 1) / TILE;
    __shared__ uint64_t s_key[W][TILE + 2];
    __shared__ double s_val[TILE + 2];
    __shared__ uint32_t s_pj[REC ? TILE + 2 : 1];
    __shared__ uint16_t s_slot[TILE];
    __shared__ uint16_t s_tlp[EMIT ? TILE : 1];  // kept position of each touched output, bit 15: negative sign
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_wacc[NWAVES];
    __shared__ uint64_t s_excl, s_excl_t;
    // SELF (static tiles only): no partition kernel ran; the tiles' splits are
    // searched here, by tile_split as that kernel searches them
    constexpr bool self = SELF;
    __shared__ uint32_t s_split[SELF ? 4 : 1];
    uint32_t* __restrict__ rhead = REC ? bf.rec + c.s.roff : nullptr;
    uint32_t* __restrict__ rcreated = REC ? bf.rec + c.s.roff + n + na : nullptr;
    const int t = threadIdx.x;
    uint32_t* __restrict__ next_tail = bf.tail[0];
    // a tile's bounds from its splits; an inconsistent partition (a bug) is
    // reported, and the tile then writes nothing but still publishes
    auto bounds = [&](uint32_t tl, uint32_t a0, uint32_t a1, uint32_t& j0, int& ns, int& nt) {
        const uint32_t d0 = tl * (uint32_t)TILE;
        const uint32_t d1 = min(d0 + (uint32_t)TILE, ntot);
        j0 = d0 - a0;
        const uint32_t j1 = d1 - a1;
        ns = (int)(a1 - a0);
        nt = (int)(j1 - j0);
        if (ns < 0 || nt < 0 || ns + nt != (int)(d1 - d0) || a1 > n || j1 > na) {
            if (t == 0) atomicOr(&ctl->err, 2u);
            ns = 0;
            nt = 0;
        }
    };
    uint32_t tile;
    if (L.dynamic) {
        if (t == 0) s_tile = __hip_atomic_fetch_add(&ctl->counter[L.slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        tile = s_tile;
    } else {
        tile = blockIdx.x;
    }
    if (tile >= ntiles) return;
    const uint32_t G = gridDim.x;
    if (self) {
        // this tile's and the next static tile's boundaries, a wave each
        const int wv = t >> 6;
        const uint32_t bd = (wv < 2 ? tile : tile + G) + (uint32_t)(wv & 1);
        const uint32_t sp = tile_split<W, TILE>(bf, keys, n, na, bd);
        if ((t & 63) == 0) s_split[wv] = sp;
        __syncthreads();
    }
    uint32_t i0, j0;
    int ns, nt;
    {
        const bool first = tile == pt && !L.dynamic;
        i0 = self ? s_split[0] : first ? pre0 : bf.splits[tile];
        const uint32_t i1 = self ? s_split[1] : first ? pre1 : bf.splits[tile + 1];
        bounds(tile, i0, i1, j0, ns, nt);
    }
    MergeRegs<W, TILE, REC> F;
    merge_fetch<W, TILE, REC>(bf, keys, coeffs, stride, na, i0, j0, ns, F);
    // static tiles: splits loaded two tiles ahead (s1: tile + G), and the
    // workgroup's inclusive prefix, packed as (touched << 32) | kept
    uint64_t incl = 0;
    uint32_t s1a = 0, s1b = 0;
    if (self) {
        s1a = s_split[2];
        s1b = s_split[3];
    } else if (!L.dynamic) {
        const bool ok1 = tile + G < ntiles;
        s1a = bf.splits[ok1 ? tile + G : 0u];
        s1b = bf.splits[ok1 ? tile + G + 1 : 0u];
    }
    while (true) {
        __syncthreads();  // the previous tile's shared memory is free again
        uint32_t s2a = 0, s2b = 0;
        if (L.dynamic) {
            if (t == 0)
                s_tile = __hip_atomic_fetch_add(&ctl->counter[L.slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (!self) {
            const bool ok2 = tile + 2 * G < ntiles;
            s2a = bf.splits[ok2 ? tile + 2 * G : 0u];
            s2b = bf.splits[ok2 ? tile + 2 * G + 1 : 0u];
        } else if (tile + 2 * G < ntiles) {
            // searched splits of the static tile after next (a sum that grew
            // past one round of tiles since the host chose this path)
            const int wv = t >> 6;
            if (wv < 2) {
                const uint32_t sp = tile_split<W, TILE>(bf, keys, n, na, tile + 2 * G + (uint32_t)wv);
                if ((t & 63) == 0) s_split[wv] = sp;
            }
        }
        merge_put<W, TILE, REC>(F, na, i0, j0, ns, nt, mask, rp.cosv, rp.damp_q, rp.fxy, rp.comm, s_key, s_val,
                                s_pj);
        __syncthreads();
        if (self && !L.dynamic && tile + 2 * G < ntiles) {
            s2a = s_split[0];
            s2b = s_split[1];
        }
        // the next tile and its splits (dynamic tiles: loaded while this tile merges)
        const uint32_t nxt = L.dynamic ? s_tile : tile + G;
        const bool more = nxt < ntiles;
        uint32_t ni0 = s1a, ni1 = s1b;
        if (L.dynamic) {
            ni0 = bf.splits[more ? nxt : 0u];
            ni1 = bf.splits[more ? nxt + 1 : 0u];
        }
        auto S = [&](int x) {
            Key<W> k;
#pragma unroll
            for (int l = 0; l < W; ++l) k.l[l] = s_key[l][1 + x];
            return k;
        };
        auto T = [&](int y) {
            Key<W> k;
#pragma unroll
            for (int l = 0; l < W; ++l) k.l[l] = s_key[l][ns + 1 + y];
            return k;
        };
        const int dl = t * ITEMS;
        const int dtot = ns + nt;
        int slot[ITEMS];
        double oc[ITEMS];
        uint32_t who[ITEMS];  // REC: head x + 1 (low half), consumed tail y + 1 (high half)
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) who[k] = 0;
        uint32_t keep = 0, touch = 0, neg = 0;
        if (dl < dtot) {
            int lo = dl > nt ? dl - nt : 0, hi = min(dl, ns);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (key_le<W>(S(mid), T(dl - 1 - mid))) lo = mid + 1;
                else hi = mid;
            }
            int x = lo, y = dl - lo;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                slot[k] = 0;
                oc[k] = 0.0;
                if (x + y >= dtot) continue;
                const bool hs = x < ns, ht = y < nt;
                const Key<W> ks = S(hs ? x : 0), kt = T(y);
                const bool take_s = hs && (!ht || key_le<W>(ks, kt));
                bool live;
                Key<W> ko;
                if (take_s) {
                    double cv = s_val[1 + x];
                    const bool merged = j0 + y < na && key_eq<W>(kt, ks);
                    if (merged) cv = cv + s_val[ns + 1 + y];
                    slot[k] = 1 + x;
                    oc[k] = cv;
                    ko = ks;
                    live = true;
                    if (REC) who[k] = (uint32_t)(x + 1) | (merged ? (uint32_t)(y + 1) << 16 : 0u);
                    ++x;
                } else {
                    live = !((i0 + x > 0) && key_eq<W>(S(x - 1), kt));
                    slot[k] = ns + 1 + y;
                    oc[k] = s_val[ns + 1 + y];
                    ko = kt;
                    if (REC && live) who[k] = (uint32_t)(y + 1) << 16;
                    ++y;
                }
                if (rp.norm) oc[k] = oc[k] * inv;
                if (live && !((rp.max_weight >= 0 && weight<W>(ko) > rp.max_weight) || fabs(oc[k]) < rp.min_abs)) {
                    keep |= 1u << k;
                    if (EMIT && anticommutes<W>(nmask, ko)) {
                        touch |= 1u << k;
                        if ((sign_exponent<W>(nmask, ko) & 2) == 0) neg |= 1u << k;
                    }
                }
            }
        }
        // kept count in the low half, touched count in the high half
        uint32_t agg;
        const uint32_t mine = (uint32_t)__popc(keep) | ((uint32_t)__popc(touch) << 16);
        static_assert(ITEMS < 8, "per-thread counts fit 3 bits");
        const uint32_t pre = block_excl_sum_small<3>(mine, s_w, agg);
        // The kept outputs go to LDS in output order (their slots, and their
        // values over s_val, whose staged inputs every thread has read by the
        // block sum's barriers), to be stored in consecutive addresses once
        // the tile's offset is known: a thread's own outputs are ITEMS apart
        // in the wave, which would split every store over 3x the lines.
        {
            uint32_t lp = pre & 0xffffu, lt = pre >> 16;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                if ((keep >> k) & 1u) {
                    s_slot[lp] = (uint16_t)slot[k];
                    s_val[lp] = oc[k];
                    if (EMIT && ((touch >> k) & 1u)) {
                        s_tlp[lt] = (uint16_t)(lp | (((neg >> k) & 1u) ? 0x8000u : 0u));
                        ++lt;
                    }
                    ++lp;
                }
            }
        }
        const uint64_t a2[2] = {agg & 0xffffu, agg >> 16};
        const uint64_t aggp = a2[0] | (a2[1] << 32);
        tile_publish<EMIT ? 2 : 1>(L, bf, tile, aggp);
        // The next tile's loads go out before this tile's wait (always, from
        // clamped addresses after the last tile: a conditional fetch would
        // keep the consumed registers live through the merge).
        const uint32_t ci0 = i0;
        if (more) {
            i0 = ni0;
            bounds(nxt, ni0, ni1, j0, ns, nt);
        } else {
            i0 = 0;
            j0 = 0;
            ns = 0;
            nt = 0;
        }
        merge_fetch<W, TILE, REC>(bf, keys, coeffs, stride, na, i0, j0, ns, F);
        const uint64_t ex = tile_resolve<EMIT ? 2 : 1>(ctl, L, bf, tile, aggp, incl, s_wacc);
        const uint64_t ek = ex & 0xffffffffull, et = ex >> 32;
        if (t == 0) {
            s_excl = ek;
            s_excl_t = et;
            if (tile == ntiles - 1) {
                const uint32_t nout = (uint32_t)(ek + a2[0]);
                set_state(ctl, p ^ 1, nout, cur ^ 1);
                if (EMIT) {
                    ctl->st[p ^ 1].na = (uint32_t)(et + a2[1]);
#pragma unroll
                    for (int l = 0; l < W; ++l) ctl->tail_mask[l] = nmask.l[l];
                    ctl->tail_seq = L.seq + 1;
                    ctl->tail_ready = 1u;
                }
                ctl->ssamp_for = L.seq + 2u;
                if (L.glog) L.glog[2] = nout;
            }
        }
        __syncthreads();
        const uint32_t obase = (uint32_t)s_excl, tbase = EMIT ? (uint32_t)s_excl_t : 0u;
        if (REC) {
            uint32_t lp = pre & 0xffffu;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const bool kept = (keep >> k) & 1u;
                if (who[k]) {
                    const uint32_t at = kept ? obase + lp : REC_NONE;
                    const uint32_t hx = who[k] & 0xffffu, ty = who[k] >> 16;
                    if (hx) {
                        Key<W> hk;
#pragma unroll
                        for (int l = 0; l < W; ++l) hk.l[l] = s_key[l][hx];
                        rhead[ci0 + hx - 1] = (anticommutes<W>(mask, hk) ? REC_FLAG : 0u) | at;
                    }
                    if (ty) rcreated[s_pj[ty - 1]] = at;
                }
                lp += kept ? 1u : 0u;
            }
        }
        // every 64th output is also the next gate's partition sample
        const uint32_t nkept = agg & 0xffffu;
#pragma unroll 1
        for (uint32_t j = t; j < nkept; j += WG) {
            const int sl = s_slot[j];
            const uint32_t o = obase + j;
#pragma unroll
            for (int l = 0; l < W; ++l) out_keys[l * stride + o] = s_key[l][sl];
            out_coeffs[o] = s_val[j];
            if ((o & 63u) == 0) {
#pragma unroll
                for (int l = 0; l < W; ++l) bf.ssamp[(size_t)(o >> 6) * W + l] = s_key[l][sl];
            }
        }
        if (EMIT) {
            // the next rotation's list: positions, and the terms themselves
            const uint32_t ntouch = agg >> 16;
#pragma unroll 1
            for (uint32_t j = t; j < ntouch; j += WG) {
                const uint32_t e = s_tlp[j];
                const uint32_t lpj = e & 0x7fffu;
                const int sl = s_slot[lpj];
                next_tail[tbase + j] = (obase + lpj) | ((e & 0x8000u) ? SRC_NEG : 0u);
#pragma unroll
                for (int l = 0; l < W; ++l) bf.tl_keys[l * stride + tbase + j] = s_key[l][sl];
                bf.tl_coef[tbase + j] = s_val[lpj];
            }
        }
        if (!more) break;
        tile = nxt;
        s1a = s2a;
        s1b = s2b;
    }
}

// The one-column merge with stores one tile behind (not recording). A tile's
// kept outputs wait in a second shared-memory buffer while the workgroup
// stages its next tile, issues the loads of the tile after, merges and
// publishes; only then does it read the waiting tile's prefix window and
// store it. The window then has a whole tile of slack. Counts never depend on
// prefixes, so publishing a tile before resolving the one before it adds no
// wait to any chain (static windows and the dynamic two-level look-back
// alike). Outputs, the next rotation's list and the partition samples are
// formed exactly as in k_merge1 and land at the same positions.
template <int W, bool EMIT, int ITEMS>
__global__ __launch_bounds__(WG, (ITEMS <= 2 ? 3 : (W == 4 && ITEMS == 4 ? 1 : 2))) void k_merge1d(Bufs bf, Ctl* ctl, Launch L, Key<W> mask,
                                                                      RotParams rp, Key<W> nmask) {
    const KTimer ktimer(L.ktime);
    constexpr int TILE = WG * ITEMS;
    constexpr bool REC = false;
    static_assert(TILE < 32768, "kept positions carry a sign in bit 15; counts are packed in 16 bits");
    const uint32_t pt = min((uint32_t)blockIdx.x, bf.lb_tiles - 2);
    const uint32_t pre0 = bf.splits[pt], pre1 = bf.splits[pt + 1];
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (!lb_start(ctl, L, c)) return;
    const int p = L.seq & 1;
    const uint32_t na = c.s.na;
    if (!(na || c.s.run)) return;
    const uint32_t n = c.s.n, cur = c.s.cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const double inv = rp.norm ? *rp.norm : 1.0;
    const size_t stride = bf.cap;
    const uint32_t ntot = n + na;
    const uint32_t ntiles = (ntot + TILE - 1) / TILE;
    // staged inputs of the tile being merged
    __shared__ uint64_t s_key[W][TILE + 2];
    __shared__ double s_val[TILE + 2];
    // kept outputs of the tile waiting for its prefix, in output order
    __shared__ uint64_t o_key[W][TILE];
    __shared__ double o_val[TILE];
    __shared__ uint16_t o_tlp[EMIT ? TILE : 1];  // kept position of each touched output, bit 15: negative sign
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_wacc[NWAVES];
    __shared__ uint64_t s_excl, s_excl_t;
    const int t = threadIdx.x;
    uint32_t* __restrict__ next_tail = bf.tail[0];
    auto bounds = [&](uint32_t tl, uint32_t a0, uint32_t a1, uint32_t& j0, int& ns, int& nt) {
        const uint32_t d0 = tl * (uint32_t)TILE;
        const uint32_t d1 = min(d0 + (uint32_t)TILE, ntot);
        j0 = d0 - a0;
        const uint32_t j1 = d1 - a1;
        ns = (int)(a1 - a0);
        nt = (int)(j1 - j0);
        if (ns < 0 || nt < 0 || ns + nt != (int)(d1 - d0) || a1 > n || j1 > na) {
            if (t == 0) atomicOr(&ctl->err, 2u);
            ns = 0;
            nt = 0;
        }
    };
    uint32_t tile;
    if (L.dynamic) {
        if (t == 0) s_tile = __hip_atomic_fetch_add(&ctl->counter[L.slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        tile = s_tile;
    } else {
        tile = blockIdx.x;
    }
    if (tile >= ntiles) return;
    uint32_t i0, j0;
    int ns, nt;
    {
        const bool first = tile == pt && !L.dynamic;
        i0 = first ? pre0 : bf.splits[tile];
        const uint32_t i1 = first ? pre1 : bf.splits[tile + 1];
        bounds(tile, i0, i1, j0, ns, nt);
    }
    MergeRegs<W, TILE, REC> F;
    merge_fetch<W, TILE, REC>(bf, keys, coeffs, stride, na, i0, j0, ns, F);
    const uint32_t G = gridDim.x;
    uint64_t incl = 0;
    uint32_t s1a = 0, s1b = 0;
    if (!L.dynamic) {
        const bool ok1 = tile + G < ntiles;
        s1a = bf.splits[ok1 ? tile + G : 0u];
        s1b = bf.splits[ok1 ? tile + G + 1 : 0u];
    }
    // the tile whose outputs wait in o_key / o_val, and its packed counts
    bool waiting = false;
    uint32_t wtile = 0;
    uint64_t waggp = 0;
    // the waiting tile's prefix (its window is almost always complete by now)
    auto resolve = [&]() {
        const uint64_t ex = tile_resolve<EMIT ? 2 : 1>(ctl, L, bf, wtile, waggp, incl, s_wacc);
        const uint32_t nkept = (uint32_t)(waggp & 0xffffffffull), ntouch = (uint32_t)(waggp >> 32);
        const uint64_t ek = ex & 0xffffffffull, et = ex >> 32;
        if (t == 0) {
            s_excl = ek;
            s_excl_t = et;
            if (wtile == ntiles - 1) {
                const uint32_t nout = (uint32_t)(ek + nkept);
                set_state(ctl, p ^ 1, nout, cur ^ 1);
                if (EMIT) {
                    ctl->st[p ^ 1].na = (uint32_t)(et + ntouch);
#pragma unroll
                    for (int l = 0; l < W; ++l) ctl->tail_mask[l] = nmask.l[l];
                    ctl->tail_seq = L.seq + 1;
                    ctl->tail_ready = 1u;
                }
                ctl->ssamp_for = L.seq + 2u;
                if (L.glog) L.glog[2] = nout;
            }
        }
        __syncthreads();
    };
    // stores the waiting tile's outputs at that prefix
    auto store = [&]() {
        const uint32_t nkept = (uint32_t)(waggp & 0xffffffffull), ntouch = (uint32_t)(waggp >> 32);
        const uint32_t obase = (uint32_t)s_excl, tbase = EMIT ? (uint32_t)s_excl_t : 0u;
        // every 64th output is also the next gate's partition sample
#pragma unroll 1
        for (uint32_t j = t; j < nkept; j += WG) {
            const uint32_t o = obase + j;
#pragma unroll
            for (int l = 0; l < W; ++l) out_keys[l * stride + o] = o_key[l][j];
            out_coeffs[o] = o_val[j];
            if ((o & 63u) == 0) {
#pragma unroll
                for (int l = 0; l < W; ++l) bf.ssamp[(size_t)(o >> 6) * W + l] = o_key[l][j];
            }
        }
        if (EMIT) {
            // the next rotation's list: positions, and the terms themselves
#pragma unroll 1
            for (uint32_t j = t; j < ntouch; j += WG) {
                const uint32_t e = o_tlp[j];
                const uint32_t lpj = e & 0x7fffu;
                next_tail[tbase + j] = (obase + lpj) | ((e & 0x8000u) ? SRC_NEG : 0u);
#pragma unroll
                for (int l = 0; l < W; ++l) bf.tl_keys[l * stride + tbase + j] = o_key[l][lpj];
                bf.tl_coef[tbase + j] = o_val[lpj];
            }
        }
    };
    // each pass merges a tile (until done) and then stores the tile before it
    bool done = false;
    while (true) {
        int slot[ITEMS];
        double oc[ITEMS];
        uint32_t keep = 0, touch = 0, neg = 0, pre = 0, nxt = 0, s2a = 0, s2b = 0, ni0 = 0, ni1 = 0, ci0 = 0, cj0 = 0;
        int cns = 0, cnt = 0;
        uint64_t aggp = 0;
        bool more = false;
        if (!done) {
            __syncthreads();  // the previous tile's staged inputs are free again
            if (L.dynamic) {
                if (t == 0)
                    s_tile = __hip_atomic_fetch_add(&ctl->counter[L.slot], 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const bool ok2 = tile + 2 * G < ntiles;
                s2a = bf.splits[ok2 ? tile + 2 * G : 0u];
                s2b = bf.splits[ok2 ? tile + 2 * G + 1 : 0u];
            }
            merge_put<W, TILE, REC>(F, na, i0, j0, ns, nt, mask, rp.cosv, rp.damp_q, rp.fxy, rp.comm, s_key, s_val,
                                    nullptr);
            __syncthreads();
            nxt = L.dynamic ? s_tile : tile + G;
            more = nxt < ntiles;
            ni0 = s1a;
            ni1 = s1b;
            if (L.dynamic) {
                ni0 = bf.splits[more ? nxt : 0u];
                ni1 = bf.splits[more ? nxt + 1 : 0u];
            }
            // this tile's bounds; the next tile's loads go out at once (always,
            // from clamped addresses after the last tile), so they are in
            // flight through this tile's merge, publish and stores
            ci0 = i0;
            cj0 = j0;
            cns = ns;
            cnt = nt;
            if (more) {
                i0 = ni0;
                bounds(nxt, ni0, ni1, j0, ns, nt);
            } else {
                i0 = 0;
                j0 = 0;
                ns = 0;
                nt = 0;
            }
            merge_fetch<W, TILE, REC>(bf, keys, coeffs, stride, na, i0, j0, ns, F);
            auto S = [&](int x) {
                Key<W> k;
#pragma unroll
                for (int l = 0; l < W; ++l) k.l[l] = s_key[l][1 + x];
                return k;
            };
            auto T = [&](int y) {
                Key<W> k;
#pragma unroll
                for (int l = 0; l < W; ++l) k.l[l] = s_key[l][cns + 1 + y];
                return k;
            };
            const int dl = t * ITEMS;
            const int dtot = cns + cnt;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                slot[k] = 0;
                oc[k] = 0.0;
            }
            if (dl < dtot) {
                int lo = dl > cnt ? dl - cnt : 0, hi = min(dl, cns);
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (key_le<W>(S(mid), T(dl - 1 - mid))) lo = mid + 1;
                    else hi = mid;
                }
                int x = lo, y = dl - lo;
#pragma unroll
                for (int k = 0; k < ITEMS; ++k) {
                    if (x + y >= dtot) continue;
                    const bool hs = x < cns, ht = y < cnt;
                    const Key<W> ks = S(hs ? x : 0), kt = T(y);
                    const bool take_s = hs && (!ht || key_le<W>(ks, kt));
                    bool live;
                    Key<W> ko;
                    if (take_s) {
                        double cv = s_val[1 + x];
                        const bool merged = cj0 + y < na && key_eq<W>(kt, ks);
                        if (merged) cv = cv + s_val[cns + 1 + y];
                        slot[k] = 1 + x;
                        oc[k] = cv;
                        ko = ks;
                        live = true;
                        ++x;
                    } else {
                        live = !((ci0 + x > 0) && key_eq<W>(S(x - 1), kt));
                        slot[k] = cns + 1 + y;
                        oc[k] = s_val[cns + 1 + y];
                        ko = kt;
                        ++y;
                    }
                    if (rp.norm) oc[k] = oc[k] * inv;
                    if (live &&
                        !((rp.max_weight >= 0 && weight<W>(ko) > rp.max_weight) || fabs(oc[k]) < rp.min_abs)) {
                        keep |= 1u << k;
                        if (EMIT && anticommutes<W>(nmask, ko)) {
                            touch |= 1u << k;
                            if ((sign_exponent<W>(nmask, ko) & 2) == 0) neg |= 1u << k;
                        }
                    }
                }
            }
            uint32_t agg;
            const uint32_t mine = (uint32_t)__popc(keep) | ((uint32_t)__popc(touch) << 16);
            static_assert(ITEMS < 8, "per-thread counts fit 3 bits");
            pre = block_excl_sum_small<3>(mine, s_w, agg);
            aggp = (uint64_t)(agg & 0xffffu) | ((uint64_t)(agg >> 16) << 32);
            tile_publish<EMIT ? 2 : 1>(L, bf, tile, aggp);
        }
        if (waiting) resolve();
        if (waiting) store();
        if (done) break;
        __syncthreads();  // the waiting tile's buffer is free again
        {
            uint32_t lp = pre & 0xffffu, lt = pre >> 16;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                if ((keep >> k) & 1u) {
#pragma unroll
                    for (int l = 0; l < W; ++l) o_key[l][lp] = s_key[l][slot[k]];
                    o_val[lp] = oc[k];
                    if (EMIT && ((touch >> k) & 1u)) {
                        o_tlp[lt] = (uint16_t)(lp | (((neg >> k) & 1u) ? 0x8000u : 0u));
                        ++lt;
                    }
                    ++lp;
                }
            }
        }
        waiting = true;
        wtile = tile;
        waggp = aggp;
        if (more) {
            tile = nxt;
            s1a = s2a;
            s1b = s2b;
        } else {
            done = true;
            __syncthreads();  // the last tile's outputs are in o_key / o_val
        }
    }
}

// ============================================================ k_merge (B coefficient columns)
//
// The same merge paths, one entry per thread; then every wave computes whole
// coefficient rows (lane b = column b), stages them in LDS and writes the kept
// rows after the look-back. trig holds cos_b then sin_b of this gate.

template <int W, int TB, bool NARROW = false>
__global__ __launch_bounds__(WG, (W <= 2 ? 4 : 3)) void k_mergeB(Bufs bf, Ctl* ctl, Launch L, Key<W> mask, int B,
                                               const double* __restrict__ trig, int max_weight, double min_abs,
                                               int damp_q, double fxy, int comm, const double* __restrict__ norm,
                                               int part, int emit, Key<W> nmask) {
    const KTimer ktimer(L.ktime);
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (!lb_start(ctl, L, c)) return;
    const int p = L.seq & 1;
    const uint32_t na = c.s.na;
    if (!(na || c.s.run)) return;
    const uint32_t n = c.s.n, cur = c.s.cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const uint32_t* __restrict__ tail = bf.tail[1];
    const size_t stride = bf.cap;
    const double* __restrict__ cosb = trig;
    const double* __restrict__ sinb = trig + B;
    const uint32_t ntot = n + na;
    const uint32_t ntiles = (ntot + TB - 1) / TB;
    extern __shared__ double s_rows[];  // TB x B
    __shared__ uint64_t s_key[W][TB + 2];
    __shared__ uint32_t s_src[TB + 1];
    __shared__ int s_head[TB];          // head index, -(index+2) when scaled by cos, -1 none, INT_MIN dead
    // per tile (two, by tile parity: a tile's rows are stored after the next
    // tile's merge paths have written its own): tail src or ~0u, reused for
    // the output offset; whether the row is kept
    __shared__ uint32_t s_tlb[2][TB];
    __shared__ uint32_t s_keepb[2][TB];
    __shared__ uint32_t s_pj[TB + 1];    // recording: parent-order index of tail slot y
    __shared__ uint32_t s_rhb[2][TB];    // recording: head index | REC_FLAG if touched, or ~0u
    __shared__ uint32_t s_rtb[2][TB];    // recording: parent-order index of the consumed tail, or ~0u
    __shared__ uint8_t s_hxy[TB];        // amplitude damping: head has X or Y on the qubit
    __shared__ uint32_t s_split[2];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl, s_excl_t;
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    uint32_t* __restrict__ next_tail = bf.tail[0];
    // with `part`, tiles take their splits from the partition kernel (launched
    // before the merge, tiles of 64 rows or more); otherwise each tile
    // searches its own (better on small sums, where the extra launch costs
    // more than the searches)
    const bool PART = (TB % 64 == 0) && part != 0;
    bool have_next = false;
    uint32_t nx0 = 0, nx1 = 0;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    constexpr int RPW = TB / NWAVES;  // rows of a wave in a tile
    // (NARROW: B <= 32, several rows per wave; otherwise one row per wave)
    const int rps = !NARROW ? 1 : (B <= 16 ? 4 : 2);
    const int sub = lane / (64 / rps), bcol = lane - sub * (64 / rps);
    const bool col = bcol < B;
    // Tiles are software-pipelined by one: a workgroup merges tile k's paths
    // (A), then stores tile k-1, whose counts it published before (C), then
    // forms tile k's rows and publishes its counts (B). The wait for tile
    // k-1's prefix thereby overlaps tile k's merge paths. With static tiles
    // and partition splits, tile k+1's keys are fetched at the start of B, so
    // A starts from registers. Each pass runs the phases once, so each is
    // inlined once.
    const bool PREF = PART && !L.dynamic;
    uint32_t tile = 0, ptile = 0;  // tile k, tile k-1
    bool cur_ok = false, prev_ok = false, fin = false, started = false;
    bool pf = false;               // tile k's keys were fetched during tile k-1
    uint32_t pi0 = 0, pj0 = 0;     // and its bounds
    int pns = 0, pnt = 0;
    LoadRegs<W, TB> KR;
    auto bounds = [&](uint32_t tl, uint32_t a0, uint32_t a1, uint32_t& j0, int& ns, int& nt) {
        const uint32_t d0 = tl * (uint32_t)TB;
        const uint32_t d1 = min(d0 + (uint32_t)TB, ntot);
        j0 = d0 - a0;
        const uint32_t j1 = d1 - a1;
        ns = (int)(a1 - a0);
        nt = (int)(j1 - j0);
        if (ns < 0 || nt < 0 || ns + nt != (int)(d1 - d0) || a1 > n || j1 > na) {
            if (t == 0) atomicOr(&ctl->err, 2u);
            ns = 0;
            nt = 0;
        }
    };
    int pb = 0;                    // tile k's metadata buffer
    Key<W> okey, pokey;            // the output term of this thread's row, tiles k and k-1
    uint32_t pmine = 0, ppre = 0, ppre_t = 0;
    bool ptouch = false;
    uint64_t paggp = 0;
    while (true) {
        // ---- A: tile k's merge paths
        if (!fin) {
            if (PREF) {
                __syncthreads();  // the previous tile's shared memory is free again
                tile = started ? tile + gridDim.x : blockIdx.x;
                started = true;
            } else {
                tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
            }
            cur_ok = tile < ntiles;
            fin = !cur_ok;
        }
        if (cur_ok) {
            uint32_t i0, j0;
            int ns, nt;
            if (pf) {
                i0 = pi0;
                j0 = pj0;
                ns = pns;
                nt = pnt;
            } else {
                const uint32_t d0 = tile * (uint32_t)TB;
                const uint32_t d1 = min(d0 + (uint32_t)TB, ntot);
                uint32_t i1;
                if (PART) {
                    // splits from the partition kernel (static tiles: loaded during the previous tile)
                    i0 = (L.dynamic || !have_next) ? bf.splits[tile] : nx0;
                    i1 = (L.dynamic || !have_next) ? bf.splits[tile + 1] : nx1;
                } else {
                    if (wave < 2) {
                        const uint32_t s = merge_split<W>(keys, stride, n, bf.tkeys, na, wave == 0 ? d0 : d1);
                        if (lane == 0) s_split[wave] = s;
                    }
                    __syncthreads();
                    i0 = s_split[0];
                    i1 = s_split[1];
                }
                bounds(tile, i0, i1, j0, ns, nt);
                merge_load_fetch<W, TB>(keys, stride, tail, bf.tkeys, na, i0, j0, ns, bf.tpj, KR);
            }
            if (PART) {
                // the next static tile's splits go out now
                const uint32_t nt2 = tile + gridDim.x;
                have_next = !L.dynamic && nt2 < ntiles;
                if (have_next) {
                    nx0 = bf.splits[nt2];
                    nx1 = bf.splits[nt2 + 1];
                }
            }
            merge_load_put<W, TB>(KR, na, i0, j0, ns, nt, s_key, s_src, bf.tpj, s_pj);
            __syncthreads();
            auto S = [&](int x) {
                Key<W> k;
#pragma unroll
                for (int l = 0; l < W; ++l) k.l[l] = s_key[l][1 + x];
                return k;
            };
            auto T = [&](int y) {
                Key<W> k;
#pragma unroll
                for (int l = 0; l < W; ++l) k.l[l] = s_key[l][ns + 1 + y];
                return k;
            };
            const int dtot = ns + nt;
            if (t < dtot) {
                int lo = t > nt ? t - nt : 0, hi = min(t, ns);
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (key_le<W>(S(mid), T(t - 1 - mid))) lo = mid + 1;
                    else hi = mid;
                }
                const int x = lo, y = t - lo;
                bool take_s;
                if (x >= ns) take_s = false;
                else if (y >= nt) take_s = true;
                else take_s = key_le<W>(S(x), T(y));
                int head = -1;
                uint32_t tl = ~0u;
                bool live = false;
                uint32_t rh = ~0u, rt = ~0u;
                bool xy = false;
                if (take_s) {
                    okey = S(x);
                    head = (int)(i0 + x);
                    rh = (uint32_t)head;
                    if (j0 + y < na && key_eq<W>(T(y), okey)) {
                        tl = s_src[y];
                        rt = s_pj[y];
                    }
                    live = true;
                    if (damp_q >= 0) {
                        const uint32_t code = code_at<W>(okey, damp_q);
                        if (code == 3u) head = -(head + 2);
                        xy = code == 1u || code == 2u;
                    } else if (anticommutes<W>(mask, okey) != (comm != 0)) {
                        head = -(head + 2);
                        rh |= REC_FLAG;
                    }
                } else {
                    okey = T(y);
                    const bool dup = (i0 + x > 0) && key_eq<W>(S(x - 1), okey);
                    if (!dup) {
                        tl = s_src[y];
                        rt = s_pj[y];
                        live = true;
                    }
                }
                s_rhb[pb][t] = rh;
                s_rtb[pb][t] = rt;
                s_hxy[t] = xy ? 1 : 0;
                if (live && max_weight >= 0 && weight<W>(okey) > max_weight) live = false;
                s_head[t] = live ? head : INT_MIN;
                s_tlb[pb][t] = live ? tl : ~0u;
                s_keepb[pb][t] = live ? 1u : 0u;
            } else if (t < TB) {
                s_keepb[pb][t] = 0u;
                s_head[t] = INT_MIN;
                s_rhb[pb][t] = ~0u;
                s_rtb[pb][t] = ~0u;
            }
            __syncthreads();
        }
        // ---- C: tile k-1's prefix and stores
        if (prev_ok) {
            const int qb = pb ^ 1;
            {
                const uint64_t ex = emit ? tile_resolve<2>(ctl, L, bf, ptile, paggp, lb_incl, s_wacc)
                                         : tile_resolve<1>(ctl, L, bf, ptile, paggp, lb_incl, s_wacc);
                const uint64_t excl = ex & 0xffffffffull, excl_t = ex >> 32;
                if (t == 0) {
                    s_excl = excl;
                    s_excl_t = excl_t;
                    if (ptile == ntiles - 1) {
                        const uint32_t nout = (uint32_t)(excl + (paggp & 0xffffffffull));
                        set_state(ctl, p ^ 1, nout, cur ^ 1);
                        if (emit) {
                            ctl->st[p ^ 1].na = (uint32_t)(excl_t + (paggp >> 32));
#pragma unroll
                            for (int l = 0; l < W; ++l) ctl->tail_mask[l] = nmask.l[l];
                            ctl->tail_seq = L.seq + 1;
                            ctl->tail_ready = 1u;
                        }
                        if (L.glog) L.glog[2] = nout;
                    }
                }
            }
            __syncthreads();
            if (t < TB && pmine) {
                store_key<W>(out_keys, stride, s_excl + ppre, pokey);
                s_tlb[qb][t] = ppre;
                if (ptouch) {
                    const uint64_t j = s_excl_t + ppre_t;
                    const bool neg = (sign_exponent<W>(nmask, pokey) & 2) == 0;
                    next_tail[j] = (uint32_t)(s_excl + ppre) | (neg ? SRC_NEG : 0u);
#pragma unroll
                    for (int l = 0; l < W; ++l) bf.tl_keys[l * stride + j] = pokey.l[l];
                }
            }
            if (bf.rec && t < TB) {
                const uint32_t at = pmine ? (uint32_t)(s_excl + ppre) : REC_NONE;
                const uint32_t roff = c.s.roff;
                if (s_rhb[qb][t] != ~0u) bf.rec[(size_t)roff + (s_rhb[qb][t] & SRC_MASK)] = (s_rhb[qb][t] & REC_FLAG) | at;
                if (s_rtb[qb][t] != ~0u) bf.rec[(size_t)roff + n + na + s_rtb[qb][t]] = at;
            }
            __syncthreads();
            if (NARROW) {
                for (int k0 = 0; k0 < RPW; k0 += rps) {
                    const int k = k0 + sub;
                    const int e = wave + k * NWAVES;
                    if (k < RPW && col && s_keepb[qb][e])
                        out_coeffs[(s_excl + s_tlb[qb][e]) * B + bcol] = s_rows[(size_t)e * B + bcol];
                }
            } else {
                for (int e = wave; e < TB; e += NWAVES) {
                    if (!s_keepb[qb][e]) continue;
                    const uint64_t pos = s_excl + s_tlb[qb][e];
                    for (int b = lane; b < B; b += 64) out_coeffs[pos * B + b] = s_rows[(size_t)e * B + b];
                }
            }
            __syncthreads();  // tile k-1's rows are out: the row buffer is free
        }
        if (!cur_ok) break;
        // ---- B: tile k's rows, counts and publication
        // (static tiles with partition splits: the next tile's keys go out first)
        pf = PREF && have_next;
        if (pf) {
            pi0 = nx0;
            bounds(tile + gridDim.x, nx0, nx1, pj0, pns, pnt);
            merge_load_fetch<W, TB>(keys, stride, tail, bf.tkeys, na, pi0, pj0, pns, bf.tpj, KR);
        }
        // B <= 64: one column per lane, rps rows of the batch side by side in a
        // wave (segments of 64 / rps lanes); a wave loads the coefficients of
        // RD slots of its rows at once and then finishes them (same values
        // and operations as the loop for larger B below)
        if (!NARROW && B <= 64) {
            // one column per lane; a wave loads the coefficients of RD of its
            // rows at once and then finishes them, so RD rows' loads are in
            // flight together (same values and operations as the loop below)
            constexpr int RD = TB / NWAVES < 8 ? TB / NWAVES : 8;
            const int b = lane;
            const bool col = b < B;
            for (int e0 = wave; e0 < TB; e0 += NWAVES * RD) {
                double hv[RD], tv[RD];
#pragma unroll
                for (int d = 0; d < RD; ++d) {
                    const int e = e0 + d * NWAVES;
                    hv[d] = 0.0;
                    tv[d] = 0.0;
                    if (e < TB && s_keepb[pb][e] && col) {
                        const int head = s_head[e];
                        const uint32_t tl = s_tlb[pb][e];
                        if (head != -1) hv[d] = coeffs[(size_t)(head >= 0 ? head : -head - 2) * B + b];
                        if (tl != ~0u) tv[d] = coeffs[(size_t)(tl & SRC_MASK) * B + b];
                    }
                }
#pragma unroll
                for (int d = 0; d < RD; ++d) {
                    const int e = e0 + d * NWAVES;
                    if (e >= TB || !s_keepb[pb][e]) continue;
                    const int head = s_head[e];
                    const uint32_t tl = s_tlb[pb][e];
                    double m = 0.0;
                    if (col) {
                        double c = 0.0;
                        bool have = false;
                        if (head != -1) {
                            if (head >= 0) {
                                c = hv[d];
                                if (s_hxy[e]) c = c * fxy;
                            } else {
                                c = hv[d] * cosb[b];
                            }
                            have = true;
                        }
                        if (tl != ~0u) {
                            double tc = tv[d] * sinb[b];
                            if (tl & SRC_NEG) tc = -tc;
                            c = have ? c + tc : tc;
                        }
                        if (norm) c = c * norm[b];
                        s_rows[(size_t)e * B + b] = c;
                        m = fabs(c);
                    }
                    m = wave_max_nonneg_f64(m);
                    if (lane == 0 && m < min_abs) s_keepb[pb][e] = 0u;
                }
            }
        } else if (NARROW) {
            constexpr int RD = RPW < 8 ? RPW : 8;
            const int nslots = (RPW + rps - 1) / rps;
            for (int q0 = 0; q0 < nslots; q0 += RD) {
                double hv[RD], tv[RD];
#pragma unroll
                for (int d = 0; d < RD; ++d) {
                    const int k = (q0 + d) * rps + sub;
                    const int e = wave + k * NWAVES;
                    hv[d] = 0.0;
                    tv[d] = 0.0;
                    if (k < RPW && s_keepb[pb][e] && col) {
                        const int head = s_head[e];
                        const uint32_t tl = s_tlb[pb][e];
                        if (head != -1) hv[d] = coeffs[(size_t)(head >= 0 ? head : -head - 2) * B + bcol];
                        if (tl != ~0u) tv[d] = coeffs[(size_t)(tl & SRC_MASK) * B + bcol];
                    }
                }
#pragma unroll
                for (int d = 0; d < RD; ++d) {
                    if (q0 + d >= nslots) break;  // uniform: every lane's slot count is the same
                    const int k = (q0 + d) * rps + sub;
                    const int e = wave + k * NWAVES;
                    const bool live = k < RPW && s_keepb[pb][e];
                    double m = 0.0;
                    if (live && col) {
                        const int head = s_head[e];
                        const uint32_t tl = s_tlb[pb][e];
                        double c = 0.0;
                        bool have = false;
                        if (head != -1) {
                            if (head >= 0) {
                                c = hv[d];
                                if (s_hxy[e]) c = c * fxy;
                            } else {
                                c = hv[d] * cosb[bcol];
                            }
                            have = true;
                        }
                        if (tl != ~0u) {
                            double tc = tv[d] * sinb[bcol];
                            if (tl & SRC_NEG) tc = -tc;
                            c = have ? c + tc : tc;
                        }
                        if (norm) c = c * norm[bcol];
                        s_rows[(size_t)e * B + bcol] = c;
                        m = fabs(c);
                    }
                    m = wave_seg_max_nonneg_f64(m, rps);
                    if (bcol == 0 && live && m < min_abs) s_keepb[pb][e] = 0u;
                }
            }
        } else
        for (int e = wave; e < TB; e += NWAVES) {
            if (!s_keepb[pb][e]) continue;
            const int head = s_head[e];
            const uint32_t tl = s_tlb[pb][e];
            double m = 0.0;
            for (int b = lane; b < B; b += 64) {
                double c = 0.0;
                bool have = false;
                if (head != -1) {
                    if (head >= 0) {
                        c = coeffs[(size_t)head * B + b];
                        if (s_hxy[e]) c = c * fxy;
                    } else {
                        const int hi = -head - 2;
                        c = coeffs[(size_t)hi * B + b] * cosb[b];
                    }
                    have = true;
                }
                if (tl != ~0u) {
                    double tc = coeffs[(size_t)(tl & SRC_MASK) * B + b] * sinb[b];
                    if (tl & SRC_NEG) tc = -tc;
                    c = have ? c + tc : tc;
                }
                if (norm) c = c * norm[b];
                s_rows[(size_t)e * B + b] = c;
                m = fmax(m, fabs(c));
            }
            m = wave_max_nonneg_f64(m);
            if (lane == 0 && m < min_abs) s_keepb[pb][e] = 0u;
        }
        __syncthreads();
        // kept rows; with `emit`, the kept rows the next rotation (nmask)
        // touches are also listed for it, as its branch pass would list them
        // (second prefix chain, counts in the high half)
        const uint32_t mine = (t < TB) ? s_keepb[pb][t] : 0u;
        const bool touch = emit && mine && anticommutes<W>(nmask, okey);
        uint32_t agg;
        const uint32_t pre2 = block_excl_sum_small<1>(mine | (touch ? 1u << 16 : 0u), s_w, agg);
        const uint32_t pre = pre2 & 0xffffu, pre_t = pre2 >> 16;
        const uint64_t aggp = (uint64_t)(agg & 0xffffu) | ((uint64_t)(agg >> 16) << 32);
        if (emit) tile_publish<2>(L, bf, tile, aggp);
        else tile_publish<1>(L, bf, tile, aggp);
        ptile = tile;
        prev_ok = true;
        pokey = okey;
        pmine = mine;
        ptouch = touch;
        ppre = pre;
        ppre_t = pre_t;
        paggp = aggp;
        pb ^= 1;
    }
}

// ============================================================ Pauli noise

// entries per thread of the one-column noise pass: fewer for 256-bit keys so the
// loaded keys stay in registers
template <int W>
struct KNCfg {
    static constexpr int ITEMS = W == 4 ? 4 : 8;
    static constexpr int TILE = WG * ITEMS;
};

// c *= factor for the damped codes on `qubit`, then truncate every term and
// compact in order (src/Base/Primitives/mapandtruncate.jl:135-178)
template <int W>
__global__ __launch_bounds__(WG) void k_noise1(Bufs bf, Ctl* ctl, Launch L, int qubit, uint32_t damp_codes,
                                               double factor, int max_weight, double min_abs) {
    const KTimer ktimer(L.ktime);
    constexpr int ITEMS = KNCfg<W>::ITEMS, TILE = KNCfg<W>::TILE;
    const CtlView c = load_ctl(ctl, L.seq & 1);
    if (!lb_start(ctl, L, c)) return;
    const int p = L.seq & 1;
    const uint32_t n = c.s.n, cur = c.s.cur;
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) {
            set_state(ctl, p ^ 1, 0u, cur);
            if (L.glog) L.glog[0] = L.glog[1] = L.glog[2] = 0u;
            carry_records(ctl, L, p);
        }
        return;
    }
    const uint32_t roff = c.s.roff;
    if (bf.rec && (uint64_t)roff + n > bf.rec_words) {
        if (blockIdx.x == 0 && t == 0) record_overflow(ctl, L, (uint64_t)roff + n, c.s.coff);
        return;
    }
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const size_t stride = bf.cap;
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[ITEMS][NWAVES];
    __shared__ uint64_t s_excl;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t base = tile * (uint32_t)TILE;
        uint64_t ballots[ITEMS];
        uint32_t damped = 0;
        double cv[ITEMS];
        Key<W> kv[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = min(base + k * WG + t, n - 1);
            kv[k] = load_key<W>(keys, stride, i);
            cv[k] = coeffs[i];
        }
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = base + k * WG + t;
            const uint32_t pc = code_at<W>(kv[k], qubit);
            if ((damp_codes >> pc) & 1u) {
                cv[k] = cv[k] * factor;
                damped |= 1u << k;
            }
            const bool keep =
                (i < n) && !((max_weight >= 0 && weight<W>(kv[k]) > max_weight) || fabs(cv[k]) < min_abs);
            ballots[k] = __ballot(keep);
            if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(ballots[k]);
        }
        __syncthreads();
        uint32_t agg = 0, before[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            uint32_t rb = 0, row = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; ++w) {
                rb += (w < wave) ? s_cnt[k][w] : 0u;
                row += s_cnt[k][w];
            }
            before[k] = agg + rb;
            agg += row;
        }
        {
            const uint64_t excl = tile_prefix<1>(ctl, L, bf, tile, agg, lb_incl, s_wacc);
            if (t == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) {
                    const uint32_t nout = (uint32_t)(excl + agg);
                    set_state(ctl, p ^ 1, nout, cur ^ 1);
                    ctl->st[p ^ 1].roff = roff + (bf.rec ? n : 0u);
                    ctl->st[p ^ 1].coff = ctl->st[p].coff;
                    if (L.glog) {
                        L.glog[0] = n;
                        L.glog[1] = n;
                        L.glog[2] = nout;
                        L.glog[3] = roff;
                        L.glog[4] = ctl->st[p].coff;
                    }
                }
            }
        }
        __syncthreads();
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const bool kept = (ballots[k] >> lane) & 1ull;
            const uint64_t pos = s_excl + before[k] + (uint32_t)__popcll(ballots[k] & lt);
            if (kept) {
                store_key<W>(out_keys, stride, pos, kv[k]);
                out_coeffs[pos] = cv[k];
            }
            const uint32_t i = base + k * WG + t;
            if (bf.rec && i < n)
                bf.rec[(size_t)roff + i] = (((damped >> k) & 1u) ? REC_FLAG : 0u) | (kept ? (uint32_t)pos : REC_NONE);
        }
    }
}

// B columns: one wave per term row, 64 rows per tile
constexpr int KNB_ROWS = 64;

template <int W, int RPS = 1>
__global__ __launch_bounds__(WG) void k_noiseB(Bufs bf, Ctl* ctl, Launch L, int B, int qubit, uint32_t damp_codes,
                                               const double* __restrict__ factor, int max_weight, double min_abs) {
    const KTimer ktimer(L.ktime);
    if (!lb_start(ctl, L)) return;
    const int p = L.seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) {
            set_state(ctl, p ^ 1, 0u, cur);
            if (L.glog) L.glog[0] = L.glog[1] = L.glog[2] = 0u;
            carry_records(ctl, L, p);
        }
        return;
    }
    const uint32_t roff = ctl->st[p].roff;
    if (bf.rec && (uint64_t)roff + n > bf.rec_words) {
        if (blockIdx.x == 0 && t == 0) record_overflow(ctl, L, (uint64_t)roff + n, ctl->st[p].coff);
        return;
    }
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const size_t stride = bf.cap;
    const uint32_t ntiles = (n + KNB_ROWS - 1) / KNB_ROWS;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_keep[KNB_ROWS];
    __shared__ uint32_t s_pos[KNB_ROWS];
    __shared__ uint8_t s_dmp[KNB_ROWS];  // recording: the row's Pauli on the qubit is damped
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t base = tile * (uint32_t)KNB_ROWS;
        constexpr int RPW = KNB_ROWS / NWAVES;  // rows of a wave in a tile
        // B <= 64: one column per lane, RPS rows side by side in a wave
        // (segments of 64 / RPS lanes; RPS > 1 only for B <= 64 / RPS); a wave
        // loads the keys and coefficients of RG of its row slots at once,
        // decides them, and keeps their damped values (cvv, one per slot) for
        // the stores (same operations as the loop below)
        constexpr int NSL = RPW / RPS;           // row slots of a wave
        constexpr int RG = NSL < 8 ? NSL : 8;    // slots loaded at once
        const int sub = RPS == 1 ? 0 : lane / (64 / RPS), bcol = lane - sub * (64 / RPS);
        const bool col = bcol < B;
        double cvv[NSL];
        if (B <= 64) {
#pragma unroll
            for (int g0 = 0; g0 < NSL; g0 += RG) {
                Key<W> kk[RG];
                double cc[RG];
#pragma unroll
                for (int d = 0; d < RG; ++d) {
                    const int k = (g0 + d) * RPS + sub;
                    const uint32_t i = base + wave + (uint32_t)k * NWAVES;
                    const uint32_t ic = min(i, n - 1);
                    kk[d] = load_key<W>(keys, stride, ic);
                    cc[d] = col ? coeffs[(size_t)ic * B + bcol] : 0.0;
                }
#pragma unroll
                for (int d = 0; d < RG; ++d) {
                    const int k = (g0 + d) * RPS + sub;
                    const int r = wave + k * NWAVES;
                    const uint32_t i = base + r;
                    uint32_t keep = 0;
                    double c = cc[d];
                    double m = 0.0;
                    const bool damped = (damp_codes >> code_at<W>(kk[d], qubit)) & 1u;
                    if (i < n && col) {
                        if (damped) c = c * factor[bcol];
                        m = fabs(c);
                    }
                    m = RPS == 1 ? wave_max_nonneg_f64(m) : wave_seg_max_nonneg_f64(m, RPS);
                    if (i < n) keep = !((max_weight >= 0 && weight<W>(kk[d]) > max_weight) || m < min_abs);
                    cvv[g0 + d] = c;
                    if (bcol == 0) {
                        s_keep[r] = keep;
                        s_dmp[r] = damped ? 1 : 0;
                    }
                }
            }
        } else
        for (int r = wave; r < KNB_ROWS; r += NWAVES) {
            const uint32_t i = base + r;
            uint32_t keep = 0;
            if (i < n) {
                const Key<W> kk = load_key<W>(keys, stride, i);
                const bool damped = (damp_codes >> code_at<W>(kk, qubit)) & 1u;
                double m = 0.0;
                for (int b = lane; b < B; b += 64) {
                    double c = coeffs[(size_t)i * B + b];
                    if (damped) c = c * factor[b];
                    m = fmax(m, fabs(c));
                }
                m = wave_max_nonneg_f64(m);
                keep = !((max_weight >= 0 && weight<W>(kk) > max_weight) || m < min_abs);
            }
            if (lane == 0) s_keep[r] = keep;
        }
        __syncthreads();
        const uint32_t mine = (t < KNB_ROWS) ? s_keep[t] : 0u;
        uint32_t agg;
        const uint32_t pre = block_excl_sum_small<1>(mine, s_w, agg);
        if (t < KNB_ROWS) s_pos[t] = pre;
        {
            const uint64_t excl = tile_prefix<1>(ctl, L, bf, tile, agg, lb_incl, s_wacc);
            if (t == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) {
                    const uint32_t nout = (uint32_t)(excl + agg);
                    set_state(ctl, p ^ 1, nout, cur ^ 1);
                    ctl->st[p ^ 1].roff = roff + (bf.rec ? n : 0u);
                    ctl->st[p ^ 1].coff = ctl->st[p].coff;
                    if (L.glog) {
                        L.glog[0] = n;
                        L.glog[1] = n;
                        L.glog[2] = nout;
                        L.glog[3] = roff;
                        L.glog[4] = ctl->st[p].coff;
                    }
                }
            }
        }
        __syncthreads();
        if (B <= 64) {
            // the kept rows' values are in registers: keys copied, values
            // stored; recording, every row's output position (or none) and
            // whether it was damped
#pragma unroll
            for (int q = 0; q < NSL; ++q) {
                const int r = wave + (q * RPS + sub) * NWAVES;
                const uint32_t i = base + r;
                if (bf.rec && bcol == 0 && i < n)
                    bf.rec[(size_t)roff + i] = (s_dmp[r] ? REC_FLAG : 0u) |
                                               (s_keep[r] ? (uint32_t)(s_excl + s_pos[r]) : REC_NONE);
                if (!s_keep[r]) continue;
                const uint64_t pos = s_excl + s_pos[r];
                if (bcol == 0) store_key<W>(out_keys, stride, pos, load_key<W>(keys, stride, i));
                if (col) out_coeffs[pos * B + bcol] = cvv[q];
            }
            continue;
        }
        for (int r = wave; r < KNB_ROWS; r += NWAVES) {
            const uint32_t i = base + r;
            if (bf.rec && i < n && lane == 0) {
                const Key<W> kr = load_key<W>(keys, stride, i);
                const bool dmp = (damp_codes >> code_at<W>(kr, qubit)) & 1u;
                bf.rec[(size_t)roff + i] = (dmp ? REC_FLAG : 0u) | (s_keep[r] ? (uint32_t)(s_excl + s_pos[r]) : REC_NONE);
            }
            if (!s_keep[r]) continue;
            const uint64_t pos = s_excl + s_pos[r];
            const Key<W> kk = load_key<W>(keys, stride, i);
            if (lane == 0) store_key<W>(out_keys, stride, pos, kk);
            const bool damped = (damp_codes >> code_at<W>(kk, qubit)) & 1u;
            for (int b = lane; b < B; b += 64) {
                double c = coeffs[(size_t)i * B + b];
                if (damped) c = c * factor[b];
                out_coeffs[pos * B + b] = c;
            }
        }
    }
}

// ============================================================ imaginary rotation normalization

// The identity's coefficient after an imaginary rotation's merge, per column,
// computed as the merge computes it: the identity head (it commutes with every
// generator) times cosh, plus the term created from the generator's own
// string (P P = I): (c sinh), negated for a negative sign. out[b] = 1 / that,
// the reference's mult!(cache, 1 / getcoeff(sum, 0))
// (src/Propagation/specializations.jl:94-98). One workgroup; trig holds cosh_b
// then sinh_b for B > 1.
template <int W>
__global__ __launch_bounds__(WG) void k_imag_norm(Bufs bf, const Ctl* ctl, uint32_t seq, Key<W> mask, int B,
                                                  double cosh1, double sinh1, const double* __restrict__ trig,
                                                  double* __restrict__ out) {
    __shared__ uint32_t s_p;
    if (ctl->overflow) return;
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    if (threadIdx.x < 64) {
        const uint32_t x =
            wave_first_true(0u, n, [&](uint32_t i) { return key_le<W>(mask, load_key<W>(keys, bf.cap, i)); });
        if (threadIdx.x == 0) s_p = (x < n && key_eq<W>(load_key<W>(keys, bf.cap, x), mask)) ? x : 0xffffffffu;
    }
    __syncthreads();
    const uint32_t xp = s_p;
    // the identity is the smallest key: present when entry 0 is zero
    bool has_i = false;
    if (n > 0) {
        const Key<W> k0 = load_key<W>(keys, bf.cap, 0);
        uint64_t any = 0;
#pragma unroll
        for (int l = 0; l < W; ++l) any |= k0.l[l];
        has_i = any == 0;
    }
    const bool neg = (sign_exponent<W>(mask, mask) & 2) == 0;
    for (int b = threadIdx.x; b < B; b += WG) {
        const double ch = B > 1 ? trig[b] : cosh1, sh = B > 1 ? trig[B + b] : sinh1;
        double c = 0.0;
        bool have = false;
        if (has_i) {
            c = coeffs[b] * ch;
            have = true;
        }
        if (xp != 0xffffffffu) {
            double tc = coeffs[(size_t)xp * B + b] * sh;
            if (neg) tc = -tc;
            c = have ? c + tc : tc;
        }
        out[b] = 1.0 / c;
    }
}

// ============================================================ readouts
//
// Overlap with a product state (src/stateoverlap.jl:12-80): every term counts
// with its coefficient times its expectation in the state, 0 or +-1.
//   RD_ZERO   |0...0>: terms holding X or Y vanish (containsXorY), others count +1
//   RD_PLUS   |+...+>: terms holding Y or Z vanish (containsYorZ), others count +1
//   RD_COMP   a computational basis state: terms holding X or Y vanish, the
//             others count (-1)^(number of Z on the qubits in state 1); `ones`
//             has bit 2q set for every such qubit q (the key layout)
//   RD_MAXMIX the maximally mixed state: the identity counts 1, every other
//             string 0 (overlapwithmaxmixed, src/stateoverlap.jl:87-93)
constexpr int RD_ZERO = 0, RD_PLUS = 1, RD_COMP = 2, RD_MAXMIX = 3;

struct Readout {
    int kind;
    int pad;
    uint64_t ones[4];
};

template <int W>
__device__ __forceinline__ int readout_factor(const Readout& r, const Key<W>& k) {
    uint64_t orth = 0;
    int nz = 0;
    if (r.kind == RD_MAXMIX) {
#pragma unroll
        for (int l = 0; l < W; ++l) orth |= k.l[l];
        return orth ? 0 : 1;
    }
#pragma unroll
    for (int l = 0; l < W; ++l) {
        const uint64_t lo = k.l[l] & ALT, hi = (k.l[l] >> 1) & ALT;  // X 01, Y 10, Z 11
        orth |= r.kind == RD_PLUS ? hi : (lo ^ hi);
        nz += __popcll(lo & hi & r.ones[l]);
    }
    if (orth) return 0;
    return (r.kind == RD_COMP && (nz & 1)) ? -1 : 1;
}

// per-block double-double partial sums of the terms' contributions
template <int W>
__global__ __launch_bounds__(WG) void k_overlap1(Bufs bf, const Ctl* ctl, uint32_t seq, Readout rd,
                                                 double* __restrict__ partial) {
    __shared__ double s_hi[NWAVES], s_lo[NWAVES];
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    DD acc{0.0, 0.0};
    for (uint32_t i = blockIdx.x * WG + t; i < n; i += gridDim.x * WG) {
        const int f = readout_factor<W>(rd, load_key<W>(keys, bf.cap, i));
        if (f) acc = dd_add_d(acc, f > 0 ? coeffs[i] : -coeffs[i]);
    }
    acc = wave_sum_dd(acc);
    if (lane == 0) {
        s_hi[wave] = acc.hi;
        s_lo[wave] = acc.lo;
    }
    __syncthreads();
    if (t == 0) {
        DD r{0.0, 0.0};
        for (int w = 0; w < NWAVES; ++w) r = dd_add(r, DD{s_hi[w], s_lo[w]});
        partial[2 * blockIdx.x] = r.hi;
        partial[2 * blockIdx.x + 1] = r.lo;
    }
}

// B columns: a wave per row, lanes over columns; partial is gridDim x B x 2
template <int W>
__global__ __launch_bounds__(WG) void k_overlapB(Bufs bf, const Ctl* ctl, uint32_t seq, Readout rd, int B,
                                                 double* __restrict__ partial) {
    extern __shared__ double s_dd[];  // NWAVES x B x 2
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    for (int b0 = 0; b0 < B; b0 += 64) {
        DD acc{0.0, 0.0};
        const int b = b0 + lane;
        for (uint32_t i = blockIdx.x * NWAVES + wave; i < n; i += gridDim.x * NWAVES) {
            const int f = readout_factor<W>(rd, load_key<W>(keys, bf.cap, i));
            if (f && b < B) {
                const double c = coeffs[(size_t)i * B + b];
                acc = dd_add_d(acc, f > 0 ? c : -c);
            }
        }
        if (b < B) {
            s_dd[(wave * B + b) * 2] = acc.hi;
            s_dd[(wave * B + b) * 2 + 1] = acc.lo;
        }
    }
    __syncthreads();
    for (int b = t; b < B; b += WG) {
        DD r{0.0, 0.0};
        for (int w = 0; w < NWAVES; ++w) r = dd_add(r, DD{s_dd[(w * B + b) * 2], s_dd[(w * B + b) * 2 + 1]});
        partial[((size_t)blockIdx.x * B + b) * 2] = r.hi;
        partial[((size_t)blockIdx.x * B + b) * 2 + 1] = r.lo;
    }
}

// maxabscoeff (src/Base/Primitives/maxabscoeff.jl:7) per column: a wave per
// row, lanes over columns; partial is gridDim x B.
__global__ __launch_bounds__(WG) void k_maxabs(Bufs bf, const Ctl* ctl, uint32_t seq, int B,
                                               double* __restrict__ partial) {
    extern __shared__ double s_m[];  // NWAVES x B
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n;
    const double* __restrict__ coeffs = bf.coeffs[ctl->st[p].cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    for (int b0 = 0; b0 < B; b0 += 64) {
        const int b = b0 + lane;
        double m = 0.0;
        for (uint32_t i = blockIdx.x * NWAVES + wave; i < n; i += gridDim.x * NWAVES)
            if (b < B) m = fmax(m, fabs(coeffs[(size_t)i * B + b]));
        if (b < B) s_m[wave * B + b] = m;
    }
    __syncthreads();
    for (int b = t; b < B; b += WG) {
        double m = 0.0;
        for (int w = 0; w < NWAVES; ++w) m = fmax(m, s_m[w * B + b]);
        partial[(size_t)blockIdx.x * B + b] = m;
    }
}

// First position in the sorted sum whose term is not below `key` (n if none).
template <int W>
__device__ __forceinline__ uint32_t sum_lower_bound(const uint64_t* __restrict__ keys, size_t stride, uint32_t n,
                                                    const Key<W>& key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (key_le<W>(key, load_key<W>(keys, stride, mid))) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// getcoeff (src/Base/termsum.jl:104-124 on a sorted, duplicate-free sum): the
// B coefficients of `key`, or zeros when it is absent. One workgroup; the
// search runs 64-ary in one wave.
template <int W>
__global__ __launch_bounds__(WG) void k_getcoeff(Bufs bf, const Ctl* ctl, uint32_t seq, Key<W> key, int B,
                                                 double* __restrict__ out) {
    __shared__ uint32_t s_pos;
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    if (threadIdx.x < 64) {
        const uint32_t x = wave_first_true(0u, n, [&](uint32_t i) { return key_le<W>(key, load_key<W>(keys, bf.cap, i)); });
        if (threadIdx.x == 0) s_pos = (x < n && key_eq<W>(load_key<W>(keys, bf.cap, x), key)) ? x : 0xffffffffu;
    }
    __syncthreads();
    const uint32_t x = s_pos;
    for (int b = threadIdx.x; b < B; b += WG) out[b] = x == 0xffffffffu ? 0.0 : bf.coeffs[cur][(size_t)x * B + b];
}

// scalarproduct (src/stateoverlap.jl:120-135): for every term of sum A, its
// coefficients times those of the same term in sum B (binary search), summed
// per column in double-double; partial is gridDim x B x 2. A wave per term of
// A, lanes over columns.
template <int W>
__global__ __launch_bounds__(WG) void k_scalar(Bufs ba, const Ctl* ca, uint32_t sa, Bufs bb, const Ctl* cb,
                                               uint32_t sb, int B, double* __restrict__ partial) {
    extern __shared__ double s_dd[];  // NWAVES x B x 2
    const uint32_t na = ca->st[sa & 1].n, cura = ca->st[sa & 1].cur;
    const uint32_t nb = cb->st[sb & 1].n, curb = cb->st[sb & 1].cur;
    const uint64_t* __restrict__ ka = ba.keys[cura];
    const uint64_t* __restrict__ kb = bb.keys[curb];
    const double* __restrict__ cfa = ba.coeffs[cura];
    const double* __restrict__ cfb = bb.coeffs[curb];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    for (int b0 = 0; b0 < B; b0 += 64) {
        DD acc{0.0, 0.0};
        const int b = b0 + lane;
        for (uint32_t i = blockIdx.x * NWAVES + wave; i < na; i += gridDim.x * NWAVES) {
            const Key<W> k = load_key<W>(ka, ba.cap, i);
            const uint32_t j = sum_lower_bound<W>(kb, bb.cap, nb, k);
            if (j < nb && key_eq<W>(load_key<W>(kb, bb.cap, j), k) && b < B)
                acc = dd_add_d(acc, cfa[(size_t)i * B + b] * cfb[(size_t)j * B + b]);
        }
        if (b < B) {
            s_dd[(wave * B + b) * 2] = acc.hi;
            s_dd[(wave * B + b) * 2 + 1] = acc.lo;
        }
    }
    __syncthreads();
    for (int b = t; b < B; b += WG) {
        DD r{0.0, 0.0};
        for (int w = 0; w < NWAVES; ++w) r = dd_add(r, DD{s_dd[(w * B + b) * 2], s_dd[(w * B + b) * 2 + 1]});
        partial[((size_t)blockIdx.x * B + b) * 2] = r.hi;
        partial[((size_t)blockIdx.x * B + b) * 2 + 1] = r.lo;
    }
}

// final reduction of the partials: one block, one column after the other
__global__ __launch_bounds__(WG) void k_overlap_final(const double* __restrict__ partial, int nblocks, int B,
                                                      double* __restrict__ out) {
    __shared__ double s_hi[NWAVES], s_lo[NWAVES];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    for (int b = 0; b < B; ++b) {
        DD acc{0.0, 0.0};
        for (int i = t; i < nblocks; i += WG)
            acc = dd_add(acc, DD{partial[((size_t)i * B + b) * 2], partial[((size_t)i * B + b) * 2 + 1]});
        acc = wave_sum_dd(acc);
        if (lane == 0) {
            s_hi[wave] = acc.hi;
            s_lo[wave] = acc.lo;
        }
        __syncthreads();
        if (t == 0) {
            DD r{0.0, 0.0};
            for (int w = 0; w < NWAVES; ++w) r = dd_add(r, DD{s_hi[w], s_lo[w]});
            out[b] = r.hi + r.lo;
        }
        __syncthreads();
    }
}

// ============================================================ rewindgradient (reference gradient)
//
// The reference's gradient (src/gradient.jl:40-108) carries, beside the
// operator, a dual sum on the same strings whose coefficients are the loss's
// weight of each string, both undone gate by gate. Three steps of it are not
// gates: seeding the dual, the commutator dot of a rotation, and capping the
// dual to the operator's strings. All on B = 1 sums.

// The dual's coefficients: the readout factor of every string, in place
// (_dualcache, src/gradient.jl:116-124).
template <int W>
__global__ __launch_bounds__(WG) void k_dual_seed(Bufs bf, const Ctl* ctl, uint32_t seq, Readout rd) {
    const int p = seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    double* __restrict__ c = bf.coeffs[cur];
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG)
        c[i] = (double)readout_factor<W>(rd, load_key<W>(keys, bf.cap, i));
}

// _generatorcommutatordot (src/gradient.jl:137-241): over operator terms t
// that anticommute with the generator P, (s c_t) d(P xor t) with s = (e & 2) - 1
// (-0.5 times the imaginary part of commutator(P, t) = 2 i^e (P xor t),
// src/PauliAlgebra/paulioperations.jl:414-424) and d the dual's coefficient
// (0 when absent), found by binary search; per-block double-double partials
// as k_overlap1 writes them.
template <int W>
__global__ __launch_bounds__(WG) void k_comm_dot(Bufs ob, const Ctl* octl, uint32_t oseq, Bufs db, const Ctl* dctl,
                                                 uint32_t dseq, Key<W> mask, double* __restrict__ partial) {
    __shared__ double s_hi[NWAVES], s_lo[NWAVES];
    const int op = oseq & 1, dp = dseq & 1;
    const uint32_t n = octl->st[op].n, nd = dctl->st[dp].n;
    const uint64_t* __restrict__ keys = ob.keys[octl->st[op].cur];
    const double* __restrict__ coeffs = ob.coeffs[octl->st[op].cur];
    const uint64_t* __restrict__ dkeys = db.keys[dctl->st[dp].cur];
    const double* __restrict__ dcoef = db.coeffs[dctl->st[dp].cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    DD acc{0.0, 0.0};
    for (uint32_t i = blockIdx.x * WG + t; i < n; i += gridDim.x * WG) {
        const Key<W> k = load_key<W>(keys, ob.cap, i);
        if (!anticommutes<W>(mask, k)) continue;
        Key<W> pk;
#pragma unroll
        for (int l = 0; l < W; ++l) pk.l[l] = k.l[l] ^ mask.l[l];
        const uint32_t j = sum_lower_bound<W>(dkeys, db.cap, nd, pk);
        if (j >= nd || !key_eq<W>(load_key<W>(dkeys, db.cap, j), pk)) continue;
        const double sg = (sign_exponent<W>(mask, k) & 2) ? 1.0 : -1.0;
        acc = dd_add_d(acc, (sg * coeffs[i]) * dcoef[j]);
    }
    acc = wave_sum_dd(acc);
    if (lane == 0) {
        s_hi[wave] = acc.hi;
        s_lo[wave] = acc.lo;
    }
    __syncthreads();
    if (t == 0) {
        DD r{0.0, 0.0};
        for (int w = 0; w < NWAVES; ++w) r = dd_add(r, DD{s_hi[w], s_lo[w]});
        partial[2 * blockIdx.x] = r.hi;
        partial[2 * blockIdx.x + 1] = r.lo;
    }
}

// _intersectfilter! (src/gradient.jl:278-350): the dual (this launch's sum)
// keeps, in order, the terms whose strings the operator (ob at its state oseq)
// holds; both sorted. A look-back compaction into the dual's other buffer, as
// k_noise1 compacts.
constexpr int KI_ITEMS = 4;
constexpr int KI_TILE = WG * KI_ITEMS;

template <int W>
__global__ __launch_bounds__(WG) void k_intersect(Bufs bf, Ctl* ctl, Launch L, Bufs ob, const Ctl* octl,
                                                  uint32_t oseq) {
    if (!lb_start(ctl, L)) return;
    const int p = L.seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const uint32_t on = octl->st[oseq & 1].n;
    const uint64_t* __restrict__ okeys = ob.keys[octl->st[oseq & 1].cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) set_state(ctl, p ^ 1, 0u, cur);
        return;
    }
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const size_t stride = bf.cap;
    const uint32_t ntiles = (n + KI_TILE - 1) / KI_TILE;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[KI_ITEMS][NWAVES];
    __shared__ uint64_t s_excl;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t base = tile * (uint32_t)KI_TILE;
        uint64_t ballots[KI_ITEMS];
        Key<W> kv[KI_ITEMS];
        double cv[KI_ITEMS];
#pragma unroll
        for (int k = 0; k < KI_ITEMS; ++k) {
            const uint32_t i = min(base + k * WG + t, n - 1);
            kv[k] = load_key<W>(keys, stride, i);
            cv[k] = coeffs[i];
        }
#pragma unroll
        for (int k = 0; k < KI_ITEMS; ++k) {
            const uint32_t i = base + k * WG + t;
            bool keep = false;
            if (i < n) {
                const uint32_t j = sum_lower_bound<W>(okeys, ob.cap, on, kv[k]);
                keep = j < on && key_eq<W>(load_key<W>(okeys, ob.cap, j), kv[k]);
            }
            ballots[k] = __ballot(keep);
            if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(ballots[k]);
        }
        __syncthreads();
        uint32_t agg = 0, before[KI_ITEMS];
#pragma unroll
        for (int k = 0; k < KI_ITEMS; ++k) {
            uint32_t rb = 0, row = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; ++w) {
                rb += (w < wave) ? s_cnt[k][w] : 0u;
                row += s_cnt[k][w];
            }
            before[k] = agg + rb;
            agg += row;
        }
        {
            const uint64_t excl = tile_prefix<1>(ctl, L, bf, tile, agg, lb_incl, s_wacc);
            if (t == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) set_state(ctl, p ^ 1, (uint32_t)(excl + agg), cur ^ 1);
            }
        }
        __syncthreads();
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
        for (int k = 0; k < KI_ITEMS; ++k) {
            if (!((ballots[k] >> lane) & 1ull)) continue;
            const uint64_t pos = s_excl + before[k] + (uint32_t)__popcll(ballots[k] & lt);
            store_key<W>(out_keys, stride, pos, kv[k]);
            out_coeffs[pos] = cv[k];
        }
    }
}

// ============================================================ Pauli presence
//
// Which Paulis occur on each qubit anywhere in the sum: three bit planes in
// the key's layout (bit 2q of limb word l set when some term holds X, Y or Z
// on qubit q). A rotation whose generator anticommutes, qubit by qubit, with
// nothing present on its qubits touches no term, so the host can leave it
// out (it is the identity, xortailmerge.jl:121-123). partial is gridDim x 3W
// words; the sum's state is st[seq & 1] (nothing is read after an overflow).
template <int W>
__global__ __launch_bounds__(WG) void k_presence(Bufs bf, const Ctl* ctl, uint32_t seq,
                                                 uint64_t* __restrict__ partial) {
    __shared__ uint64_t s_w[NWAVES][3 * W];
    const int p = seq & 1;
    const uint32_t n = ctl->overflow ? 0u : ctl->st[p].n;
    const uint64_t* __restrict__ keys = bf.keys[ctl->st[p].cur];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    uint64_t acc[3 * W];
#pragma unroll
    for (int j = 0; j < 3 * W; ++j) acc[j] = 0;
    for (uint32_t i = blockIdx.x * WG + t; i < n; i += gridDim.x * WG) {
#pragma unroll
        for (int l = 0; l < W; ++l) {
            const uint64_t k = keys[(size_t)l * bf.cap + i];
            const uint64_t lo = k & ALT, hi = (k >> 1) & ALT;
            acc[l] |= lo & ~hi;       // X
            acc[W + l] |= hi & ~lo;   // Y
            acc[2 * W + l] |= lo & hi;  // Z
        }
    }
#pragma unroll
    for (int j = 0; j < 3 * W; ++j) {
        uint64_t v = acc[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, 64);
        if (lane == 0) s_w[wave][j] = v;
    }
    __syncthreads();
    if (t < 3 * W) {
        uint64_t v = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) v |= s_w[w][t];
        partial[(size_t)blockIdx.x * 3 * W + t] = v;
    }
}

// out[j] = OR of partial[b * 3W + j] over nblocks blocks; one workgroup, every
// block's words requested in one pass.
template <int W>
__global__ __launch_bounds__(WG) void k_presence_final(const uint64_t* __restrict__ partial, int nblocks,
                                                       uint64_t* __restrict__ out) {
    __shared__ uint64_t s_w[NWAVES][3 * W];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    uint64_t acc[3 * W];
#pragma unroll
    for (int j = 0; j < 3 * W; ++j) acc[j] = 0;
    for (int b = t; b < nblocks; b += WG) {
#pragma unroll
        for (int j = 0; j < 3 * W; ++j) acc[j] |= partial[(size_t)b * 3 * W + j];
    }
#pragma unroll
    for (int j = 0; j < 3 * W; ++j) {
        uint64_t v = acc[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, 64);
        if (lane == 0) s_w[wave][j] = v;
    }
    __syncthreads();
    if (t < 3 * W) {
        uint64_t v = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) v |= s_w[w][t];
        out[t] = v;
    }
}

// ============================================================ adjoint sweep
//
// The loss is overlapwithzero of the propagated sum. Its derivative with
// respect to every coefficient (the adjoint) runs backwards through the gates
// along the records of the forward run, with every truncation held as it
// happened. A rotation maps input term i to outputs ph (itself, times cos
// when it anticommutes) and pt (its created term, times sign * sin), so
//   a_in[i] = a[ph] cos + a[pt] sign sin   (a[ph] alone when it commutes)
//   dL/dtheta = sum over touched i of c_i (a[pt] sign cos - a[ph] sin),
// the derivative of the branch rule of src/Propagation/specializations.jl:34-44.
// Threads form rows of BP = 2^lgBP threads, thread b of a row handling
// column b < B; coefficient rows are B doubles.

// Adjoint of the final sum: the loss's weight of every string, the readout
// factor of `rd` (overlapwithzero: 1 without X or Y, else 0; src/stateoverlap.jl:12-93).
template <int W>
__global__ __launch_bounds__(WG) void k_adj_seed(const uint64_t* __restrict__ keys, size_t stride, uint32_t n, int B,
                                                 Readout rd, double* __restrict__ a) {
    const uint64_t total = (uint64_t)n * B;
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < total; e += (uint64_t)gridDim.x * WG) {
        const uint32_t i = (uint32_t)(e / B);
        a[e] = (double)readout_factor<W>(rd, load_key<W>(keys, stride, i));
    }
}

// Sum of a block's per-thread values of each column, in a fixed order; the
// block's columns are c0 + b for the BP lanes b of a row.
__device__ __forceinline__ void adj_block_columns(double g, int lgBP, int B, int c0, double* __restrict__ out) {
    __shared__ double s_g[WG];
    const int t = threadIdx.x, BP = 1 << lgBP, r = t >> lgBP, b = t & (BP - 1);
    s_g[t] = g;
    __syncthreads();
    for (int k = (WG >> lgBP) / 2; k >= 1; k >>= 1) {
        if (r < k) s_g[t] += s_g[t + k * BP];
        __syncthreads();
    }
    if (r == 0 && c0 + b < B) out[c0 + b] = s_g[t];
    __syncthreads();  // s_g is reused by the next column group
}

// Column groups a row of BP = 2^lgBP threads covers: one for B <= WG (BP the
// next power of two of B), else ceil(B / WG) groups of WG columns.
constexpr int ADJ_MAX_GROUPS = 4;  // B <= 1024

// One rotation backwards. rec points at the gate's records, rcoef at its
// coefficient rows, trig at cos_b then sin_b. Items [0, n) are the gate's input
// terms (those that commute are done here), items [n, n + na) its touched terms.
// Thread b of a row handles columns b, b + BP, ... below B.
__global__ __launch_bounds__(WG) void k_adj_rotation(const double* __restrict__ a_out, double* __restrict__ a_in,
                                                     const uint32_t* __restrict__ rec, uint32_t n, uint32_t na,
                                                     const double* __restrict__ rcoef, int B, int lgBP,
                                                     const double* __restrict__ trig, double* __restrict__ gpart) {
    const int t = threadIdx.x, BP = 1 << lgBP, rows = WG >> lgBP, r = t >> lgBP, b = t & (BP - 1);
    const int groups = (B + BP - 1) / BP;
    const uint32_t* __restrict__ rsrc = rec + n;
    const uint32_t* __restrict__ rpt = rec + n + na;
    const uint32_t items = n + na;
    double g[ADJ_MAX_GROUPS];
#pragma unroll
    for (int k = 0; k < ADJ_MAX_GROUPS; ++k) g[k] = 0.0;
    for (uint32_t it = blockIdx.x * rows + r; it < items; it += gridDim.x * rows) {
#pragma unroll
        for (int k = 0; k < ADJ_MAX_GROUPS; ++k) {
            const int col = b + k * BP;
            if (k >= groups || col >= B) continue;
            const double cb = trig[col], sb = trig[B + col];
            if (it < n) {
                const uint32_t ph = rec[it];
                if (ph & REC_FLAG) continue;  // touched: done with its created term
                const uint32_t at = ph & SRC_MASK;
                a_in[(size_t)it * B + col] = at != REC_NONE ? a_out[(size_t)at * B + col] : 0.0;
            } else {
                const uint32_t j = it - n;
                const uint32_t src = rsrc[j], i = src & SRC_MASK;
                const uint32_t hp = rec[i] & SRC_MASK, tp = rpt[j];
                const double ah = hp != REC_NONE ? a_out[(size_t)hp * B + col] : 0.0;
                double at = tp != REC_NONE ? a_out[(size_t)tp * B + col] : 0.0;
                if (src & SRC_NEG) at = -at;
                a_in[(size_t)i * B + col] = ah * cb + at * sb;
                const double c = rcoef[(size_t)j * B + col];
                g[k] += c * (at * cb) - c * (ah * sb);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < ADJ_MAX_GROUPS; ++k)
        if (k < groups) adj_block_columns(g[k], lgBP, B, k * BP, gpart + (size_t)blockIdx.x * B);
}

// Pauli noise backwards: a_in[i] = a[pos] times the factor when i was damped.
__global__ __launch_bounds__(WG) void k_adj_noise(const double* __restrict__ a_out, double* __restrict__ a_in,
                                                  const uint32_t* __restrict__ rec, uint32_t n, int B,
                                                  const double* __restrict__ factor) {
    const uint64_t total = (uint64_t)n * B;
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < total; e += (uint64_t)gridDim.x * WG) {
        const uint32_t i = (uint32_t)(e / B), b = (uint32_t)(e % B);
        const uint32_t w = rec[i], at = w & SRC_MASK;
        double v = at != REC_NONE ? a_out[(size_t)at * B + b] : 0.0;
        if (w & REC_FLAG) v = v * factor[b];
        a_in[e] = v;
    }
}

// The gate's gradient per column: the blocks' partial sums in block order.
__global__ __launch_bounds__(WG) void k_adj_final(const double* __restrict__ gpart, int nblocks, int B,
                                                  double* __restrict__ out) {
    for (int b = threadIdx.x; b < B; b += WG) {
        double s = 0.0;
        for (int k = 0; k < nblocks; ++k) s += gpart[(size_t)k * B + b];
        out[b] = s;
    }
}

// Every rotation of a segment of the sweep: slot s (one workgroup each) holds
// the partial gradients of nbs[s] blocks of the rotation at gate offset
// gidx[s]; they are summed per column in block order, as k_adj_final sums one
// gate's.
__global__ __launch_bounds__(WG) void k_adj_final_all(const double* __restrict__ gpart,
                                                      const uint32_t* __restrict__ nbs,
                                                      const uint32_t* __restrict__ gidx, uint32_t gblocks, int B,
                                                      double* __restrict__ out) {
    const uint32_t s = blockIdx.x;
    const int nb = (int)nbs[s];
    if (nb == 0) return;
    const double* __restrict__ p = gpart + (size_t)s * gblocks * B;
    double* __restrict__ o = out + (size_t)gidx[s] * B;
    for (int b = threadIdx.x; b < B; b += WG) {
        double acc = 0.0;
        for (int k = 0; k < nb; ++k) acc += p[(size_t)k * B + b];
        o[b] = acc;
    }
}

}  // namespace pp
