// pp_device.h -- device-side building blocks for the gfx950 Pauli-propagation kernels.
//
// Keys are W uint64 limbs (W = 1, 2, 4) stored as limb planes: limb l of term i
// lives at keys[l * stride + i]. Qubit q occupies bits 2q, 2q+1 of the
// 64W-bit integer (I=0, X=1, Y=2, Z=3; src/PauliAlgebra/bitoperations.jl:154).
// Order is unsigned integer order, highest limb first
// (src/Base/NTupleInteger/datatype.jl:280-309).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr uint64_t ALT = 0x5555555555555555ull;
constexpr int WG = 256;          // threads per workgroup: four wave64
constexpr int NWAVES = WG / 64;
constexpr int MAX_LEVELS = 16;   // bit levels one tail-permutation launch handles

template <int W>
struct Key {
    uint64_t l[W];
};

// ---------------------------------------------------------------- bit algebra

__device__ __forceinline__ uint64_t swap_pairs(uint64_t w) { return ((w >> 1) & ALT) | ((w & ALT) << 1); }

// anticommutation: parity of popcount(swap(mask) & key) over the limbs
// (src/PauliAlgebra/wide_bitoperations.jl:16-25)
template <int W>
__device__ __forceinline__ bool anticommutes(const Key<W>& m, const Key<W>& k) {
    uint64_t f = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) f ^= swap_pairs(m.l[l]) & k.l[l];
    return __popcll(f) & 1;
}

// exponent e of P1 * P2 = i^e (P1 xor P2), from the single-qubit table:
// with x = low ^ high and z = high, e = x1 z1 + x2 z2 + 2 z1 x2 - x3 z3 summed mod 4
template <int W>
__device__ __forceinline__ int sign_exponent(const Key<W>& p1, const Key<W>& p2) {
    int e = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) {
        uint64_t a = p1.l[l], b = p2.l[l];
        uint64_t z1 = (a >> 1) & ALT, x1 = (a ^ (a >> 1)) & ALT;
        uint64_t z2 = (b >> 1) & ALT, x2 = (b ^ (b >> 1)) & ALT;
        uint64_t x3 = x1 ^ x2, z3 = z1 ^ z2;
        e += __popcll(x1 & z1) + __popcll(x2 & z2) + 2 * __popcll(z1 & x2) - __popcll(x3 & z3);
    }
    return e & 3;
}

// one limb's share of sign_exponent (the shares add up mod 4)
__device__ __forceinline__ int limb_sign_exponent(uint64_t a, uint64_t b) {
    const uint64_t z1 = (a >> 1) & ALT, x1 = (a ^ (a >> 1)) & ALT;
    const uint64_t z2 = (b >> 1) & ALT, x2 = (b ^ (b >> 1)) & ALT;
    const uint64_t x3 = x1 ^ x2, z3 = z1 ^ z2;
    return __popcll(x1 & z1) + __popcll(x2 & z2) + 2 * __popcll(z1 & x2) - __popcll(x3 & z3);
}

template <int W>
__device__ __forceinline__ int weight(const Key<W>& k) {
    int c = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) c += __popcll((k.l[l] | (k.l[l] >> 1)) & ALT);
    return c;
}

template <int W>
__device__ __forceinline__ bool contains_xy(const Key<W>& k) {
    uint64_t f = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) f |= (k.l[l] ^ (k.l[l] >> 1)) & ALT;
    return f != 0;
}

template <int W>
__device__ __forceinline__ bool key_le(const Key<W>& a, const Key<W>& b) {
#pragma unroll
    for (int l = W - 1; l >= 0; --l)
        if (a.l[l] != b.l[l]) return a.l[l] < b.l[l];
    return true;
}

template <int W>
__device__ __forceinline__ bool key_eq(const Key<W>& a, const Key<W>& b) {
    uint64_t d = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) d |= a.l[l] ^ b.l[l];
    return d == 0;
}

// equal on every bit at position >= bit (bit may be 64W: always equal)
template <int W>
__device__ __forceinline__ bool eq_from(const Key<W>& a, const Key<W>& b, int bit) {
    bool eq = true;
#pragma unroll
    for (int l = 0; l < W; ++l) {
        uint64_t d = a.l[l] ^ b.l[l];
        int lo = l * 64;
        if (bit >= lo + 64) continue;
        if (bit > lo) d >>= (bit - lo);
        eq = eq && (d == 0);
    }
    return eq;
}

template <int W>
__device__ __forceinline__ Key<W> load_key(const uint64_t* __restrict__ keys, size_t stride, size_t i) {
    Key<W> k;
#pragma unroll
    for (int l = 0; l < W; ++l) k.l[l] = keys[l * stride + i];
    return k;
}

// only the limbs at or above limb `from` (the rest read as zero)
template <int W>
__device__ __forceinline__ Key<W> load_key_from(const uint64_t* __restrict__ keys, size_t stride, size_t i,
                                                int from) {
    Key<W> k;
#pragma unroll
    for (int l = 0; l < W; ++l) k.l[l] = (l >= from) ? keys[l * stride + i] : 0ull;
    return k;
}

template <int W>
__device__ __forceinline__ void store_key(uint64_t* __restrict__ keys, size_t stride, size_t i, const Key<W>& k) {
#pragma unroll
    for (int l = 0; l < W; ++l) keys[l * stride + i] = k.l[l];
}

template <int W>
__device__ __forceinline__ Key<W> key_xor(const Key<W>& a, const Key<W>& b) {
    Key<W> r;
#pragma unroll
    for (int l = 0; l < W; ++l) r.l[l] = a.l[l] ^ b.l[l];
    return r;
}

// Limb `l` of a key for a run-time l, selected with constant indices so that the
// key stays in registers (a variable index would move it to scratch memory).
// The limbs are combined with masks rather than selected: a select between
// limbs can be folded into one load from a variable address.
template <int W>
__device__ __forceinline__ uint64_t limb_at(const Key<W>& k, int l) {
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) r |= k.l[i] & (0ull - (uint64_t)(i == l));
    return r;
}

// Pauli code (0..3) of qubit q of a key.
template <int W>
__device__ __forceinline__ uint32_t code_at(const Key<W>& k, int q) {
    return (uint32_t)((limb_at<W>(k, q >> 5) >> (2 * (q & 31))) & 3ull);
}

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Sum over a full wave by data-parallel primitives: prefix sums inside each
// row of 16 lanes (row_shr steps, zeros shifted in at the row start), then
// the four row totals read as scalars. No LDS traffic, unlike shuffles.
__device__ __forceinline__ uint64_t wave_sum_u64_dpp(uint64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#define PP_ROW_STEP(CTRL)                                                                            \
    {                                                                                                \
        const uint32_t tl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, CTRL, 0xF, 0xF, true); \
        const uint32_t th = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, CTRL, 0xF, 0xF, true); \
        const uint64_t sum = (((uint64_t)hi << 32) | lo) + (((uint64_t)th << 32) | tl);             \
        lo = (uint32_t)sum;                                                                          \
        hi = (uint32_t)(sum >> 32);                                                                  \
    }
    PP_ROW_STEP(0x111)  // row_shr:1
    PP_ROW_STEP(0x112)  // row_shr:2
    PP_ROW_STEP(0x114)  // row_shr:4
    PP_ROW_STEP(0x118)  // row_shr:8
#undef PP_ROW_STEP
    uint64_t total = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t rl = (uint32_t)__builtin_amdgcn_readlane((int)lo, 16 * r + 15);
        const uint32_t rh = (uint32_t)__builtin_amdgcn_readlane((int)hi, 16 * r + 15);
        total += ((uint64_t)rh << 32) | rl;
    }
    return total;
}

// Maximum over a full wave of values >= 0 (zeros shifted in are neutral), by
// the same row steps and row reads as wave_sum_u64_dpp.
__device__ __forceinline__ double wave_max_nonneg_f64(double v) {
#define PP_ROW_STEP(CTRL)                                                                              \
    {                                                                                                  \
        const uint64_t u = (uint64_t)__double_as_longlong(v);                                          \
        const uint32_t tl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, true); \
        const uint32_t th = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true); \
        v = fmax(v, __longlong_as_double((long long)(((uint64_t)th << 32) | tl)));                     \
    }
    PP_ROW_STEP(0x111)  // row_shr:1
    PP_ROW_STEP(0x112)  // row_shr:2
    PP_ROW_STEP(0x114)  // row_shr:4
    PP_ROW_STEP(0x118)  // row_shr:8
#undef PP_ROW_STEP
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    double m = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t rl = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 16 * r + 15);
        const uint32_t rh = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 16 * r + 15);
        m = fmax(m, __longlong_as_double((long long)(((uint64_t)rh << 32) | rl)));
    }
    return m;
}

// Maximum of values >= 0 over segments of 64 / rps lanes (rps = 1, 2 or 4
// rows of a batch side by side in one wave): the same row steps, then each
// lane takes the maximum of its segment's 16-lane rows. Every lane calls it.
__device__ __forceinline__ double wave_seg_max_nonneg_f64(double v, int rps) {
#define PP_ROW_STEP(CTRL)                                                                              \
    {                                                                                                  \
        const uint64_t u = (uint64_t)__double_as_longlong(v);                                          \
        const uint32_t tl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, true); \
        const uint32_t th = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true); \
        v = fmax(v, __longlong_as_double((long long)(((uint64_t)th << 32) | tl)));                     \
    }
    PP_ROW_STEP(0x111)  // row_shr:1
    PP_ROW_STEP(0x112)  // row_shr:2
    PP_ROW_STEP(0x114)  // row_shr:4
    PP_ROW_STEP(0x118)  // row_shr:8
#undef PP_ROW_STEP
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t rl = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 16 * k + 15);
        const uint32_t rh = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 16 * k + 15);
        r[k] = __longlong_as_double((long long)(((uint64_t)rh << 32) | rl));
    }
    const int lane = lane_id();
    if (rps == 4) {
        const int q = lane >> 4;
        return q == 0 ? r[0] : (q == 1 ? r[1] : (q == 2 ? r[2] : r[3]));
    }
    if (rps == 2) return lane < 32 ? fmax(r[0], r[1]) : fmax(r[2], r[3]);
    return fmax(fmax(r[0], r[1]), fmax(r[2], r[3]));
}

__device__ __forceinline__ uint32_t wave_incl_sum_u32(uint32_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// ---------------------------------------------------------------- double-double sums

struct DD {
    double hi, lo;
};

// error-free sum of two doubles (Knuth's TwoSum)
__device__ __forceinline__ DD two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double err = (a - (s - bb)) + (b - bb);
    return {s, err};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    double lo = s.lo + (a.lo + b.lo);
    return two_sum(s.hi, lo);
}

__device__ __forceinline__ DD dd_add_d(DD a, double b) {
    DD s = two_sum(a.hi, b);
    return two_sum(s.hi, s.lo + a.lo);
}

__device__ __forceinline__ DD wave_sum_dd(DD v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        DD u{__shfl_xor(v.hi, o, 64), __shfl_xor(v.lo, o, 64)};
        v = dd_add(v, u);
    }
    return v;
}

// ---------------------------------------------------------------- decoupled look-back
//
// One 64-bit word per tile: [63:62] status (1 = tile aggregate, 2 = inclusive
// prefix), [61:40] launch epoch, [39:0] value. A word from an older launch
// carries another epoch and reads as "not ready", so the array is never
// cleared between launches. All accesses are relaxed agent-scope atomics:
// only the words themselves are communicated.

constexpr uint64_t LB_VAL_MASK = (1ull << 40) - 1;
constexpr uint32_t LB_GROUP_SHIFT = 6;  // tiles per group: 1 << LB_GROUP_SHIFT (at most 64)
constexpr uint32_t LB_GROUP = 1u << LB_GROUP_SHIFT;
constexpr uint32_t LB_EPOCH_MASK = (1u << 22) - 1;

__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroups take tile ids in the order they start, so a tile only waits on
// tiles that are already running. The workgroup that takes the last id resets
// the counter for the next launch on the stream.
__device__ __forceinline__ uint32_t lb_tile_id(uint32_t* counter, uint32_t* s_slot) {
    if (threadIdx.x == 0) {
        uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == gridDim.x - 1) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_slot = t;
    }
    __syncthreads();
    return *s_slot;
}

// Status and value of a look-back word read by this launch: words of older
// launches carry another epoch and read as "not ready" (status 0).
__device__ __forceinline__ uint32_t lb_status(uint64_t w, uint32_t ep32) {
    return (((w >> 40) & LB_EPOCH_MASK) == ep32) ? (uint32_t)(w >> 62) : 0u;
}

// One poll of a window of up to 64 predecessor words, nearest first, for NV
// independent prefix chains whose words lie `cstride` words apart: lane l reads
// word `base - l` of every chain when l < count. For each chain still active
// it finds the nearest inclusive prefix in the window; the poll fails while a
// word up to it (or up to the window's end, when it holds none) is not ready.
// On success sum[c] is the window's contribution to chain c's exclusive prefix
// and found[c] tells whether that chain ended at an inclusive prefix.
template <int NV>
__device__ __forceinline__ bool lb_window(const uint64_t* words, size_t cstride, int64_t base, int count,
                                          uint32_t ep32, const bool (&active)[NV], uint64_t (&sum)[NV],
                                          bool (&found)[NV]) {
    const int lane = lane_id();
    const bool valid = lane < count;
    const size_t idx = valid && base - lane >= 0 ? (size_t)(base - lane) : 0;
    uint64_t w[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) w[c] = lb_load(words + c * cstride + idx);
    uint64_t s[NV];
    bool f[NV];
    bool ok = true;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const uint32_t st = valid ? lb_status(w[c], ep32) : 0u;
        const uint64_t pm = __ballot(st == 2u);
        const int P = pm ? (__ffsll((unsigned long long)pm) - 1) : count;
        ok = ok && (!active[c] || __ballot(lane < P && st == 0u) == 0);
        s[c] = wave_sum_u64_dpp(lane <= P && valid ? (w[c] & LB_VAL_MASK) : 0ull);
        f[c] = P < count;
    }
    if (!ok) return false;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        sum[c] = active[c] ? s[c] : 0ull;
        found[c] = active[c] ? f[c] : true;
    }
    return true;
}

// First half of lb_exclusive: publish this tile's aggregate (tile 0: its
// inclusive prefix). A workgroup may do other work between this and
// lb_resolve (its next tile's loads): nothing waits on it until then.
template <int NV>
__device__ __forceinline__ void lb_publish(uint64_t* states, size_t tstride, uint32_t tile, uint32_t epoch,
                                           const uint64_t (&agg)[NV]) {
    if (lane_id() != 0) return;
    const uint64_t ep = (uint64_t)(epoch & LB_EPOCH_MASK) << 40;
    const uint64_t st = tile == 0 ? (2ull << 62) : (1ull << 62);
#pragma unroll
    for (int c = 0; c < NV; ++c) lb_store(states + c * tstride + tile, st | ep | agg[c]);
}

// Second half of lb_exclusive: the exclusive prefix of `tile`, whose aggregate
// lb_publish has stored.
template <int NV>
__device__ void lb_resolve(uint64_t* states, size_t tstride, uint64_t* gstates, size_t gstride, uint32_t tile,
                           uint32_t epoch, const uint64_t (&agg)[NV], uint32_t spin_limit, bool* aborted,
                           uint64_t (&excl)[NV]) {
    const int lane = lane_id();
    const uint32_t ep32 = epoch & LB_EPOCH_MASK;
    const uint64_t ep = (uint64_t)ep32 << 40;
    const uint32_t g = tile >> LB_GROUP_SHIFT;
    const int r = (int)(tile & (LB_GROUP - 1u));
    const bool last_of_group = r == (int)LB_GROUP - 1;
#pragma unroll
    for (int c = 0; c < NV; ++c) excl[c] = 0;
    if (tile == 0) return;
    uint32_t spins = 0;
    bool active[NV], found[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) active[c] = true;
    // this group's earlier tiles (tile 0 of the first group is always inclusive)
    while (!lb_window<NV>(states, tstride, (int64_t)tile - 1, r, ep32, active, excl, found)) {
        if (spin_limit && ++spins > spin_limit) {
            *aborted = true;
#pragma unroll
            for (int c = 0; c < NV; ++c) excl[c] = 0;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    bool any = false;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        active[c] = !found[c];
        any = any || active[c];
        if (active[c] && last_of_group && lane == 0)
            lb_store(gstates + c * gstride + g, (1ull << 62) | ep | (excl[c] + agg[c]));
    }
    int64_t gb = (int64_t)g - 1;
    while (any) {
        uint64_t part[NV];
        bool gfound[NV];
        const int cnt = gb + 1 < 64 ? (int)(gb + 1) : 64;
        if (lb_window<NV>(gstates, gstride, gb, cnt, ep32, active, part, gfound)) {
            any = false;
#pragma unroll
            for (int c = 0; c < NV; ++c) {
                excl[c] += part[c];
                // group 0's word is inclusive once ready
                active[c] = active[c] && !gfound[c] && gb >= 64;
                any = any || active[c];
            }
            gb -= 64;
            continue;
        }
        if (spin_limit && ++spins > spin_limit) {
            *aborted = true;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            lb_store(states + c * tstride + tile, (2ull << 62) | ep | (excl[c] + agg[c]));
            if (last_of_group) lb_store(gstates + c * gstride + g, (2ull << 62) | ep | (excl[c] + agg[c]));
        }
    }
}

// Two-level decoupled look-back over NV prefix chains at once (the chains
// share every poll). Called by all 64 lanes of one wave; publishes agg[c] for
// `tile` and returns in excl[c] the exclusive prefix of all earlier tiles.
// Tile words of chain c start at states + c * tstride, group words at
// gstates + c * gstride.
//
// Tiles form groups of 64. A tile first reads its own group's earlier words;
// when none of them holds an inclusive prefix it adds the group-level words
// (one per earlier group: the group's aggregate, or its inclusive prefix),
// 64 groups per poll. The last tile of a group publishes the group's word.
// A prefix therefore resolves in a few polls however many tiles run at once,
// where a flat look-back walks back through every running tile, 64 per poll.
//
// spin_limit bounds the polls of one wait (0 = unbounded). When it runs out,
// *aborted is set and the partial sums read so far are returned: they never
// exceed the true prefixes, so positions built from them stay in range.
template <int NV>
__device__ void lb_exclusive(uint64_t* states, size_t tstride, uint64_t* gstates, size_t gstride, uint32_t tile,
                             uint32_t epoch, const uint64_t (&agg)[NV], uint32_t spin_limit, bool* aborted,
                             uint64_t (&excl)[NV]) {
    lb_publish<NV>(states, tstride, tile, epoch, agg);
    lb_resolve<NV>(states, tstride, gstates, gstride, tile, epoch, agg, spin_limit, aborted, excl);
}

}  // namespace pp
