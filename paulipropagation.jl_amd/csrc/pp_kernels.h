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
#pragma once
#include "pp_device.h"

namespace pp {

constexpr uint32_t SRC_MASK = 0x7fffffffu;
constexpr uint32_t SRC_NEG = 0x80000000u;

// ============================================================ k_branch

constexpr int K1_ITEMS = 8;
constexpr int K1_TILE = WG * K1_ITEMS;

template <int W>
__global__ __launch_bounds__(WG) void k_branch(const uint64_t* __restrict__ keys, size_t stride, uint32_t n,
                                               Key<W> mask, uint32_t* __restrict__ tail, uint64_t* states,
                                               uint32_t* counter, uint32_t epoch, uint32_t* total) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[K1_ITEMS][NWAVES];
    __shared__ uint64_t s_excl;
    const uint32_t tile = lb_tile_id(counter, &s_tile);
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t base = tile * (uint32_t)K1_TILE;

    uint64_t ballots[K1_ITEMS];
    uint32_t negbits = 0;
#pragma unroll
    for (int k = 0; k < K1_ITEMS; ++k) {
        const uint32_t i = base + k * WG + t;
        bool f = false;
        if (i < n) {
            Key<W> kk;
#pragma unroll
            for (int l = 0; l < W; ++l) kk.l[l] = mask.l[l] ? keys[l * stride + i] : 0ull;
            f = anticommutes(mask, kk);
            if (f && (sign_exponent(mask, kk) & 2) == 0) negbits |= 1u << k;
        }
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
    if (wave == 0) {
        const uint64_t excl = lb_exclusive(states, tile, epoch, agg);
        if (lane == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1) *total = (uint32_t)(excl + agg);
        }
    }
    __syncthreads();
    const uint64_t excl = s_excl;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < K1_ITEMS; ++k) {
        if ((ballots[k] >> lane) & 1ull) {
            const uint32_t i = base + k * WG + t;
            const uint32_t pos = (uint32_t)excl + before[k] + (uint32_t)__popcll(ballots[k] & lt);
            tail[pos] = i | (((negbits >> k) & 1u) ? SRC_NEG : 0u);
        }
    }
}

// ============================================================ k_xorperm

constexpr int K2_ITEMS = 4;
constexpr int K2_TILE = WG * K2_ITEMS;
constexpr int INF_POS = 0x7fffffff;

struct Levels {
    int n;                   // number of levels
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
        const bool p = (xl >= hi) ? true : fn((uint32_t)xl);
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

template <int W>
__global__ __launch_bounds__(WG) void k_xorperm(const uint64_t* __restrict__ keys, size_t stride,
                                                const uint32_t* __restrict__ tail_in, uint32_t na, Levels lv,
                                                int64_t const_off, int from_limb,
                                                uint32_t* __restrict__ tail_out, uint32_t* err) {
    __shared__ Key<W> s_last[WG];
    __shared__ Key<W> s_kb, s_kl;
    __shared__ int s_edge_s[MAX_LEVELS], s_edge_e[MAX_LEVELS];
    __shared__ int s_wagg[NWAVES];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t base = blockIdx.x * (uint32_t)K2_TILE;
    const uint32_t cnt = min((uint32_t)K2_TILE, na - base);
    const uint32_t last = base + cnt - 1;

    uint32_t src[K2_ITEMS];
    Key<W> key[K2_ITEMS];
    const uint32_t j0 = base + t * K2_ITEMS;
    int nvalid = 0;
#pragma unroll
    for (int i = 0; i < K2_ITEMS; ++i) {
        const uint32_t j = j0 + i;
        if (j < base + cnt) {
            src[i] = tail_in[j];
            key[i] = load_key_from<W>(keys, stride, src[i] & SRC_MASK, from_limb);
            ++nvalid;
        } else {
            src[i] = 0;
#pragma unroll
            for (int l = 0; l < W; ++l) key[i].l[l] = 0;
        }
    }
    if (nvalid > 0) s_last[t] = key[nvalid - 1];

    auto key_at = [&](uint32_t x) -> Key<W> {
        return load_key_from<W>(keys, stride, tail_in[x] & SRC_MASK, from_limb);
    };

    // tile edges: the block around `base` and the block around `last`, per level
    if (t == 0) s_kb = key[0];
    if (nvalid > 0 && j0 + nvalid - 1 == last) s_kl = key[nvalid - 1];
    __syncthreads();
    const Key<W> kb = s_kb, kl = s_kl;
    for (int task = wave; task < 2 * lv.n; task += NWAVES) {
        const int v = task >> 1;
        const int bit = lv.bit[v];
        if ((task & 1) == 0) {
            // start of the block containing `base`
            uint32_t s = 0;
            if (base > 0) {
                const Key<W> kp = key_at(base - 1);
                if (!eq_from<W>(kp, kb, bit)) s = base;
                else s = wave_first_true(0u, base, [&](uint32_t x) { return eq_from<W>(key_at(x), kb, bit); });
            }
            if (lane == 0) s_edge_s[v] = (int)s;
        } else {
            uint32_t e = na;
            if (last + 1 < na) {
                const Key<W> kn = key_at(last + 1);
                if (!eq_from<W>(kn, kl, bit)) e = last + 1;
                else e = wave_first_true(last + 1, na, [&](uint32_t x) { return !eq_from<W>(key_at(x), kl, bit); });
            }
            if (lane == 0) s_edge_e[v] = (int)e;
        }
    }
    // previous entry of this thread's first item
    Key<W> prev;
    if (t > 0) prev = s_last[t - 1];
    else if (base > 0) prev = key_at(base - 1);
    else prev = key[0];
    __syncthreads();

    int64_t dest[K2_ITEMS];
#pragma unroll
    for (int i = 0; i < K2_ITEMS; ++i) dest[i] = (int64_t)(j0 + i) + const_off;

    for (int v = 0; v < lv.n; ++v) {
        const int bit = lv.bit[v];
        bool h[K2_ITEMS];
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) {
            const uint32_t j = j0 + i;
            const Key<W>& pk = (i == 0) ? prev : key[i - 1];
            h[i] = (i < nvalid) && (j == 0 || !eq_from<W>(key[i], pk, bit));
        }
        // starts: last head at or before the item
        int sl[K2_ITEMS];
        int run = -1;
#pragma unroll
        for (int i = 0; i < K2_ITEMS; ++i) {
            if (h[i]) run = (int)(j0 + i);
            sl[i] = run;
        }
        // ends: first head after the item
        int el[K2_ITEMS];
        int nxt = INF_POS;
#pragma unroll
        for (int i = K2_ITEMS - 1; i >= 0; --i) {
            el[i] = nxt;
            if (h[i]) nxt = (int)(j0 + i);
        }
        // exclusive max-scan of `run` over threads, exclusive suffix-min of `nxt`
        int incl = wave_incl_max(run);
        int wex = __shfl_up(incl, 1, 64);
        if (lane == 0) wex = -1;
        if (lane == 63) s_wagg[wave] = incl;
        __syncthreads();
        int pre = wex;
        for (int w = 0; w < wave; ++w) pre = max(pre, s_wagg[w]);
        __syncthreads();
        int sincl = wave_suffix_min(nxt);
        int sex = __shfl_down(sincl, 1, 64);
        if (lane == 63) sex = INF_POS;
        if (lane == 0) s_wagg[wave] = sincl;
        __syncthreads();
        int suf = sex;
        for (int w = wave + 1; w < NWAVES; ++w) suf = min(suf, s_wagg[w]);
        __syncthreads();
        const int es = s_edge_s[v], ee = s_edge_e[v];
        const int sg = lv.sgn[v];
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
            if (dest[i] >= 0 && dest[i] < (int64_t)na) tail_out[dest[i]] = src[i];
            else atomicOr(err, 1u);
        }
    }
}

// ============================================================ k_merge (one coefficient column)

constexpr int K3_ITEMS = 4;
constexpr int K3_TILE = WG * K3_ITEMS;

struct RotParams {
    double cosv, sinv;
    int max_weight;   // < 0: no weight truncation
    double min_abs;
};

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

// number of head entries among the first d merged entries; heads go first on ties
template <int W>
__device__ __forceinline__ uint32_t merge_split(const uint64_t* __restrict__ keys, size_t stride, uint32_t n,
                                                const uint32_t* __restrict__ tail, uint32_t na, const Key<W>& mask,
                                                uint32_t d) {
    const uint32_t xlo = d > na ? d - na : 0u;
    const uint32_t xhi = min(d, n);
    return wave_first_true(xlo, xhi, [&](uint32_t x) {
        const Key<W> a = load_key<W>(keys, stride, x);
        const Key<W> b = key_xor(load_key<W>(keys, stride, tail[d - 1 - x] & SRC_MASK), mask);
        return !key_le<W>(a, b);
    });
}

template <int W>
__global__ __launch_bounds__(WG) void k_merge1(const uint64_t* __restrict__ keys, size_t stride,
                                               const double* __restrict__ coeffs, uint32_t n,
                                               const uint32_t* __restrict__ tail, uint32_t na, Key<W> mask,
                                               RotParams rp, uint64_t* __restrict__ out_keys, size_t out_stride,
                                               double* __restrict__ out_coeffs, uint64_t* states,
                                               uint32_t* counter, uint32_t epoch, uint32_t* total, uint32_t* err) {
    __shared__ uint64_t s_key[W][K3_TILE + 2];
    __shared__ uint32_t s_src[K3_TILE + 1];
    __shared__ uint32_t s_split[2];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl;
    const uint32_t tile = lb_tile_id(counter, &s_tile);
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t ntot = n + na;
    const uint32_t d0 = tile * (uint32_t)K3_TILE;
    const uint32_t d1 = min(d0 + (uint32_t)K3_TILE, ntot);
    if (wave < 2) {
        const uint32_t s = merge_split<W>(keys, stride, n, tail, na, mask, wave == 0 ? d0 : d1);
        if (lane == 0) s_split[wave] = s;
    }
    __syncthreads();
    const uint32_t i0 = s_split[0], i1 = s_split[1];
    const uint32_t j0 = d0 - i0, j1 = d1 - i1;
    int ns = (int)(i1 - i0), nt = (int)(j1 - j0);
    if (ns < 0 || nt < 0 || ns + nt != (int)(d1 - d0) || i1 > n || j1 > na) {
        // inconsistent partition (a bug): report it, write nothing, still publish
        if (t == 0) atomicOr(err, 2u);
        ns = 0;
        nt = 0;
    }
    // slots: [0] = S[i0-1], [1 .. ns] = S[i0 .. i1), [ns+1 .. ns+1+nt] = T[j0 .. j1]
    for (int x = t; x < ns + 1; x += WG) {
        if (x > 0 || i0 > 0) {
            const size_t gi = (size_t)i0 + x - 1;
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][x] = keys[l * stride + gi];
        }
    }
    for (int y = t; y < nt + 1; y += WG) {
        const uint32_t gj = j0 + y;
        if (gj < na) {
            const uint32_t src = tail[gj];
            s_src[y] = src;
            const uint32_t si = src & SRC_MASK;
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][ns + 1 + y] = keys[l * stride + si] ^ mask.l[l];
        }
    }
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
    const int dl = t * K3_ITEMS;
    const int dtot = ns + nt;
    Key<W> okey[K3_ITEMS];
    double oc[K3_ITEMS];
    uint32_t keep = 0;
    if (dl < dtot) {
        // thread-level merge path inside the tile
        int lo = dl > nt ? dl - nt : 0, hi = min(dl, ns);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (key_le<W>(S(mid), T(dl - 1 - mid))) lo = mid + 1;
            else hi = mid;
        }
        int x = lo, y = dl - lo;
#pragma unroll
        for (int k = 0; k < K3_ITEMS; ++k) {
            if (x + y >= dtot) break;
            bool take_s;
            if (x >= ns) take_s = false;
            else if (y >= nt) take_s = true;
            else take_s = key_le<W>(S(x), T(y));
            if (take_s) {
                const Key<W> ks = S(x);
                const uint32_t gi = i0 + x;
                double c = coeffs[gi];
                if (anticommutes<W>(mask, ks)) c = c * rp.cosv;
                if (j0 + y < na && key_eq<W>(T(y), ks)) {
                    const uint32_t src = s_src[y];
                    double tc = coeffs[src & SRC_MASK] * rp.sinv;
                    if (src & SRC_NEG) tc = -tc;
                    c = c + tc;
                }
                okey[k] = ks;
                oc[k] = c;
                keep |= 1u << k;
                ++x;
            } else {
                const Key<W> kt = T(y);
                const bool dup = (i0 + x > 0) && key_eq<W>(S(x - 1), kt);
                if (!dup) {
                    const uint32_t src = s_src[y];
                    double tc = coeffs[src & SRC_MASK] * rp.sinv;
                    if (src & SRC_NEG) tc = -tc;
                    okey[k] = kt;
                    oc[k] = tc;
                    keep |= 1u << k;
                }
                ++y;
            }
            if ((keep >> k) & 1u) {
                const bool drop = (rp.max_weight >= 0 && weight<W>(okey[k]) > rp.max_weight) ||
                                  (fabs(oc[k]) < rp.min_abs);
                if (drop) keep &= ~(1u << k);
            }
        }
    }
    uint32_t agg;
    const uint32_t mine = (uint32_t)__popc(keep);
    const uint32_t pre = block_excl_sum(mine, s_w, agg);
    if (wave == 0) {
        const uint64_t excl = lb_exclusive(states, tile, epoch, agg);
        if (lane == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1) *total = (uint32_t)(excl + agg);
        }
    }
    __syncthreads();
    uint64_t pos = s_excl + pre;
#pragma unroll
    for (int k = 0; k < K3_ITEMS; ++k) {
        if ((keep >> k) & 1u) {
            store_key<W>(out_keys, out_stride, pos, okey[k]);
            out_coeffs[pos] = oc[k];
            ++pos;
        }
    }
}

// ============================================================ k_merge (B coefficient columns)
//
// The same merge paths, one entry per thread; then every wave computes whole
// coefficient rows (lane b = column b), stages them in LDS and writes the kept
// rows after the look-back.

template <int W, int TB>
__global__ __launch_bounds__(WG) void k_mergeB(const uint64_t* __restrict__ keys, size_t stride,
                                               const double* __restrict__ coeffs, uint32_t n, int B,
                                               const uint32_t* __restrict__ tail, uint32_t na, Key<W> mask,
                                               const double* __restrict__ cosb, const double* __restrict__ sinb,
                                               int max_weight, double min_abs, uint64_t* __restrict__ out_keys,
                                               size_t out_stride, double* __restrict__ out_coeffs, uint64_t* states,
                                               uint32_t* counter, uint32_t epoch, uint32_t* total, uint32_t* err) {
    extern __shared__ double s_rows[];  // TB x B
    __shared__ uint64_t s_key[W][TB + 2];
    __shared__ uint32_t s_src[TB + 1];
    __shared__ int s_head[TB];          // head index or -1
    __shared__ uint32_t s_tl[TB];       // tail src, or ~0u for none
    __shared__ uint32_t s_keep[TB];
    __shared__ uint32_t s_split[2];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl;
    const uint32_t tile = lb_tile_id(counter, &s_tile);
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t ntot = n + na;
    const uint32_t d0 = tile * (uint32_t)TB;
    const uint32_t d1 = min(d0 + (uint32_t)TB, ntot);
    if (wave < 2) {
        const uint32_t s = merge_split<W>(keys, stride, n, tail, na, mask, wave == 0 ? d0 : d1);
        if (lane == 0) s_split[wave] = s;
    }
    __syncthreads();
    const uint32_t i0 = s_split[0], i1 = s_split[1];
    const uint32_t j0 = d0 - i0, j1 = d1 - i1;
    int ns = (int)(i1 - i0), nt = (int)(j1 - j0);
    if (ns < 0 || nt < 0 || ns + nt != (int)(d1 - d0) || i1 > n || j1 > na) {
        if (t == 0) atomicOr(err, 2u);
        ns = 0;
        nt = 0;
    }
    for (int x = t; x < ns + 1; x += WG) {
        if (x > 0 || i0 > 0) {
            const size_t gi = (size_t)i0 + x - 1;
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][x] = keys[l * stride + gi];
        }
    }
    for (int y = t; y < nt + 1; y += WG) {
        const uint32_t gj = j0 + y;
        if (gj < na) {
            const uint32_t src = tail[gj];
            s_src[y] = src;
            const uint32_t si = src & SRC_MASK;
#pragma unroll
            for (int l = 0; l < W; ++l) s_key[l][ns + 1 + y] = keys[l * stride + si] ^ mask.l[l];
        }
    }
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
    Key<W> okey;
    bool live = false;
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
        if (take_s) {
            okey = S(x);
            head = (int)(i0 + x);
            if (j0 + y < na && key_eq<W>(T(y), okey)) tl = s_src[y];
            live = true;
            if (anticommutes<W>(mask, okey)) head = -(head + 2);  // encodes "scale by cos"
        } else {
            okey = T(y);
            const bool dup = (i0 + x > 0) && key_eq<W>(S(x - 1), okey);
            if (!dup) {
                tl = s_src[y];
                live = true;
            }
        }
        if (live && max_weight >= 0 && weight<W>(okey) > max_weight) live = false;
        s_head[t] = head;
        s_tl[t] = live ? tl : ~0u;
        s_keep[t] = live ? 1u : 0u;
        if (!live) s_head[t] = INT_MIN;
    } else if (t < TB) {
        s_keep[t] = 0u;
        s_head[t] = INT_MIN;
    }
    __syncthreads();
    // coefficient rows
    for (int e = wave; e < TB; e += NWAVES) {
        if (!s_keep[e]) continue;
        const int head = s_head[e];
        const uint32_t tl = s_tl[e];
        double m = 0.0;
        for (int b = lane; b < B; b += 64) {
            double c = 0.0;
            bool have = false;
            if (head != INT_MIN && head != -1) {
                if (head >= 0) {
                    c = coeffs[(size_t)head * B + b];
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
            s_rows[(size_t)e * B + b] = c;
            m = fmax(m, fabs(c));
        }
        m = wave_max_f64(m);
        if (lane == 0 && m < min_abs) s_keep[e] = 0u;
    }
    __syncthreads();
    const uint32_t mine = (t < TB) ? s_keep[t] : 0u;
    uint32_t agg;
    const uint32_t pre = block_excl_sum(mine, s_w, agg);
    if (wave == 0) {
        const uint64_t excl = lb_exclusive(states, tile, epoch, agg);
        if (lane == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1) *total = (uint32_t)(excl + agg);
        }
    }
    __syncthreads();
    // positions of kept rows: reuse s_tl as the output position
    if (t < TB) {
        if (mine) {
            store_key<W>(out_keys, out_stride, s_excl + pre, okey);
            s_tl[t] = pre;
        }
    }
    __syncthreads();
    for (int e = wave; e < TB; e += NWAVES) {
        if (!s_keep[e]) continue;
        const uint64_t pos = s_excl + s_tl[e];
        for (int b = lane; b < B; b += 64) out_coeffs[pos * B + b] = s_rows[(size_t)e * B + b];
    }
}

// ============================================================ Pauli noise

constexpr int KN_ITEMS = 8;
constexpr int KN_TILE = WG * KN_ITEMS;

// c *= factor for the damped codes on `qubit`, then truncate every term and
// compact in order (src/Base/Primitives/mapandtruncate.jl:135-178)
template <int W>
__global__ __launch_bounds__(WG) void k_noise1(const uint64_t* __restrict__ keys, size_t stride,
                                               const double* __restrict__ coeffs, uint32_t n, int qubit,
                                               uint32_t damp_codes, double factor, int max_weight, double min_abs,
                                               uint64_t* __restrict__ out_keys, size_t out_stride,
                                               double* __restrict__ out_coeffs, uint64_t* states, uint32_t* counter,
                                               uint32_t epoch, uint32_t* total) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[KN_ITEMS][NWAVES];
    __shared__ uint64_t s_excl;
    const uint32_t tile = lb_tile_id(counter, &s_tile);
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t base = tile * (uint32_t)KN_TILE;
    const int ql = qubit >> 5, qs = 2 * (qubit & 31);
    uint64_t ballots[KN_ITEMS];
    double cv[KN_ITEMS];
#pragma unroll
    for (int k = 0; k < KN_ITEMS; ++k) {
        const uint32_t i = base + k * WG + t;
        bool keep = false;
        cv[k] = 0.0;
        if (i < n) {
            const Key<W> kk = load_key<W>(keys, stride, i);
            double c = coeffs[i];
            const uint32_t p = (uint32_t)((kk.l[ql] >> qs) & 3ull);
            if ((damp_codes >> p) & 1u) c = c * factor;
            cv[k] = c;
            keep = !((max_weight >= 0 && weight<W>(kk) > max_weight) || fabs(c) < min_abs);
        }
        ballots[k] = __ballot(keep);
        if (lane == 0) s_cnt[k][wave] = (uint32_t)__popcll(ballots[k]);
    }
    __syncthreads();
    uint32_t agg = 0, before[KN_ITEMS];
#pragma unroll
    for (int k = 0; k < KN_ITEMS; ++k) {
        uint32_t rb = 0, row = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) {
            rb += (w < wave) ? s_cnt[k][w] : 0u;
            row += s_cnt[k][w];
        }
        before[k] = agg + rb;
        agg += row;
    }
    if (wave == 0) {
        const uint64_t excl = lb_exclusive(states, tile, epoch, agg);
        if (lane == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1) *total = (uint32_t)(excl + agg);
        }
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < KN_ITEMS; ++k) {
        if ((ballots[k] >> lane) & 1ull) {
            const uint32_t i = base + k * WG + t;
            const uint64_t pos = s_excl + before[k] + (uint32_t)__popcll(ballots[k] & lt);
#pragma unroll
            for (int l = 0; l < W; ++l) out_keys[l * out_stride + pos] = keys[l * stride + i];
            out_coeffs[pos] = cv[k];
        }
    }
}

// B columns: one wave per term row
template <int W>
__global__ __launch_bounds__(WG) void k_noiseB(const uint64_t* __restrict__ keys, size_t stride,
                                               const double* __restrict__ coeffs, uint32_t n, int B, int qubit,
                                               uint32_t damp_codes, const double* __restrict__ factor,
                                               int max_weight, double min_abs, uint64_t* __restrict__ out_keys,
                                               size_t out_stride, double* __restrict__ out_coeffs,
                                               uint64_t* states, uint32_t* counter, uint32_t epoch,
                                               uint32_t* total) {
    constexpr int ROWS = 64;  // rows per tile
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_keep[ROWS];
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl;
    const uint32_t tile = lb_tile_id(counter, &s_tile);
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint32_t base = tile * (uint32_t)ROWS;
    const int ql = qubit >> 5, qs = 2 * (qubit & 31);
    for (int r = wave; r < ROWS; r += NWAVES) {
        const uint32_t i = base + r;
        uint32_t keep = 0;
        if (i < n) {
            const Key<W> kk = load_key<W>(keys, stride, i);
            const bool damped = (damp_codes >> ((kk.l[ql] >> qs) & 3ull)) & 1u;
            double m = 0.0;
            for (int b = lane; b < B; b += 64) {
                double c = coeffs[(size_t)i * B + b];
                if (damped) c = c * factor[b];
                m = fmax(m, fabs(c));
            }
            m = wave_max_f64(m);
            keep = !((max_weight >= 0 && weight<W>(kk) > max_weight) || m < min_abs);
        }
        if (lane == 0) s_keep[r] = keep;
    }
    __syncthreads();
    const uint32_t mine = (t < ROWS) ? s_keep[t] : 0u;
    uint32_t agg;
    const uint32_t pre = block_excl_sum(mine, s_w, agg);
    __shared__ uint32_t s_pos[ROWS];
    if (t < ROWS) s_pos[t] = pre;
    if (wave == 0) {
        const uint64_t excl = lb_exclusive(states, tile, epoch, agg);
        if (lane == 0) {
            s_excl = excl;
            if (tile == gridDim.x - 1) *total = (uint32_t)(excl + agg);
        }
    }
    __syncthreads();
    for (int r = wave; r < ROWS; r += NWAVES) {
        if (!s_keep[r]) continue;
        const uint32_t i = base + r;
        const uint64_t pos = s_excl + s_pos[r];
        const Key<W> kk = load_key<W>(keys, stride, i);
        if (lane == 0) store_key<W>(out_keys, out_stride, pos, kk);
        const bool damped = (damp_codes >> ((kk.l[ql] >> qs) & 3ull)) & 1u;
        for (int b = lane; b < B; b += 64) {
            double c = coeffs[(size_t)i * B + b];
            if (damped) c = c * factor[b];
            out_coeffs[pos * B + b] = c;
        }
    }
}

// ============================================================ overlapwithzero

// per-block double-double partial sums of the coefficients of terms with no X or Y
template <int W>
__global__ __launch_bounds__(WG) void k_overlap1(const uint64_t* __restrict__ keys, size_t stride,
                                                 const double* __restrict__ coeffs, uint32_t n,
                                                 double* __restrict__ partial) {
    __shared__ double s_hi[NWAVES], s_lo[NWAVES];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    DD acc{0.0, 0.0};
    for (uint32_t i = blockIdx.x * WG + t; i < n; i += gridDim.x * WG) {
        const Key<W> kk = load_key<W>(keys, stride, i);
        if (!contains_xy<W>(kk)) acc = dd_add_d(acc, coeffs[i]);
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
__global__ __launch_bounds__(WG) void k_overlapB(const uint64_t* __restrict__ keys, size_t stride,
                                                 const double* __restrict__ coeffs, uint32_t n, int B,
                                                 double* __restrict__ partial) {
    extern __shared__ double s_dd[];  // NWAVES x B x 2
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    for (int b0 = 0; b0 < B; b0 += 64) {
        DD acc{0.0, 0.0};
        const int b = b0 + lane;
        for (uint32_t i = blockIdx.x * NWAVES + wave; i < n; i += gridDim.x * NWAVES) {
            const Key<W> kk = load_key<W>(keys, stride, i);
            if (!contains_xy<W>(kk) && b < B) acc = dd_add_d(acc, coeffs[(size_t)i * B + b]);
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

// final reduction of the partials: one block, one column per thread group
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

}  // namespace pp
