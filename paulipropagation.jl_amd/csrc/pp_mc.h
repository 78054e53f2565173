// pp_mc.h -- gfx950 kernels for Monte Carlo resampling of a B = 1 sum
// (src/Base/MonteCarlo/resample.jl).
//
// Every term lies on [0, total) on an interval as wide as its weight (|c|, or
// c^2 when resampling squared; zero for the terms a semideterministic comb
// keeps as they are), in sum order. A term gets the teeth of a comb, or the
// random draws, inside its interval; its new coefficient is teeth * spacing
// with its sign (the square root of that when squared), and the terms left at
// zero are dropped with the order of the others kept.
//
// The sum is cut into chunks of MC_CHUNK consecutive terms, one workgroup
// each: a thread adds its MC_ITEMS consecutive weights left to right, the
// chunk adds its threads' sums in thread order, and the host adds the chunks
// in order. A term's interval starts at its chunk's start plus the sums of the
// threads and items before it. (Weights with few significant bits, such as
// dyadic ones, sum exactly in any order; the positions then equal those of
// the reference's sequential walk, resample.jl:263-275.)
#pragma once
#include "pp_kernels.h"

namespace pp {

constexpr int MC_ITEMS = 16;
constexpr int MC_CHUNK = WG * MC_ITEMS;

// what k_mc_weigh counts and adds, per term of coefficient c and weight w
constexpr int MC_ABS = 0;    // add |c|
constexpr int MC_COMB = 1;   // count |c| > p (kept); add |c| of the others
constexpr int MC_W = 2;      // add w
constexpr int MC_HEAVY = 3;  // count w >= p (heavy); add w of the others

struct McParams {
    int method;          // 0 semideterministic systematic, 1 systematic, 2 multinomial
    int squared;         // weights c^2
    double thr;          // method 0: terms with |c| > thr are kept as they are
    double spacing;      // comb spacing / weight per draw
    double inv;          // teeth per weight: 1 / spacing (0 when spacing is 0)
    double offset;       // comb offset in [0, 1)
    const double* draws;  // method 2: the draws times the total weight, ascending
    uint32_t ndraws;
};

__device__ __forceinline__ double mc_weight(const McParams& m, double c) {
    if (m.method == 0) return fabs(c) > m.thr ? 0.0 : fabs(c);
    return m.squared ? c * c : fabs(c);
}

// teeth (draws) below a position: floor(position * inv - offset), or the
// number of draws below it (searchsortedfirst - 1, resample.jl:71-72, 183-194)
__device__ __forceinline__ double mc_below(const McParams& m, double pos) {
    if (m.method != 2) return floor(pos * m.inv - m.offset);
    uint32_t lo = 0, hi = m.ndraws;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (m.draws[mid] < pos) lo = mid + 1;
        else hi = mid;
    }
    return (double)lo;
}

// Per chunk: the count and the sum of `mode` (see above); p is the threshold
// or spacing the mode compares with, squared picks c^2 weights.
__global__ __launch_bounds__(WG) void k_mc_weigh(Bufs bf, const Ctl* ctl, uint32_t seq, int mode, double p,
                                                 int squared, double* __restrict__ sums, uint32_t* __restrict__ cnts) {
    __shared__ double s_sum[WG];
    __shared__ uint32_t s_cnt[WG];
    const int ps = seq & 1;
    const uint32_t n = ctl->st[ps].n;
    const double* __restrict__ coeffs = bf.coeffs[ctl->st[ps].cur];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * (uint32_t)MC_CHUNK + t * MC_ITEMS;
    double acc = 0.0;
    uint32_t cnt = 0;
    for (int k = 0; k < MC_ITEMS; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const double c = coeffs[i];
        const double w = squared ? c * c : fabs(c);
        if (mode == MC_ABS) {
            acc += fabs(c);
        } else if (mode == MC_COMB) {
            if (fabs(c) > p) ++cnt;
            else acc += fabs(c);
        } else if (mode == MC_W) {
            acc += w;
        } else {
            if (w >= p) ++cnt;
            else acc += w;
        }
    }
    s_sum[t] = acc;
    s_cnt[t] = cnt;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        uint32_t c = 0;
        for (int j = 0; j < WG; ++j) {
            s += s_sum[j];
            c += s_cnt[j];
        }
        sums[blockIdx.x] = s;
        cnts[blockIdx.x] = c;
    }
}

// Per chunk: every term's new coefficient into newc (zero when it is
// dropped), and the chunk's survivor count. starts[c] is where chunk c's
// intervals begin.
__global__ __launch_bounds__(WG) void k_mc_apply(Bufs bf, const Ctl* ctl, uint32_t seq, McParams m,
                                                 const double* __restrict__ starts, double* __restrict__ newc,
                                                 uint32_t* __restrict__ keeps) {
    __shared__ double s_pre[WG];
    __shared__ uint32_t s_cnt[WG];
    const int ps = seq & 1;
    const uint32_t n = ctl->st[ps].n;
    const double* __restrict__ coeffs = bf.coeffs[ctl->st[ps].cur];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * (uint32_t)MC_CHUNK + t * MC_ITEMS;
    double c[MC_ITEMS];
    double own = 0.0;
#pragma unroll
    for (int k = 0; k < MC_ITEMS; ++k) {
        c[k] = base + k < n ? coeffs[base + k] : 0.0;
        if (base + k < n) own += mc_weight(m, c[k]);
    }
    s_pre[t] = own;
    __syncthreads();
    if (t == 0) {
        double r = starts[blockIdx.x];
        for (int j = 0; j < WG; ++j) {
            const double w = s_pre[j];
            s_pre[j] = r;
            r += w;
        }
    }
    __syncthreads();
    double pos = s_pre[t];
    double below_end = mc_below(m, pos);
    uint32_t kept = 0;
#pragma unroll
    for (int k = 0; k < MC_ITEMS; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const bool keep_as_is = m.method == 0 && fabs(c[k]) > m.thr;
        const double below_start = below_end;
        pos += mc_weight(m, c[k]);
        below_end = mc_below(m, pos);
        const double teeth = below_end - below_start;
        double nc;
        if (keep_as_is) nc = c[k];
        else if (m.squared) nc = sqrt(teeth * m.spacing) * (c[k] == 0.0 ? 0.0 : 1.0);
        else nc = copysign(teeth * m.spacing, c[k]);
        newc[i] = nc;
        kept += nc != 0.0;
    }
    s_cnt[t] = kept;
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int j = 0; j < WG; ++j) s += s_cnt[j];
        keeps[blockIdx.x] = s;
    }
}

// Per chunk: the survivors (newc != 0) in order, keys and coefficients, into
// the other buffer from outs[c] on.
template <int W>
__global__ __launch_bounds__(WG) void k_mc_scatter(Bufs bf, const Ctl* ctl, uint32_t seq,
                                                   const double* __restrict__ newc, const uint32_t* __restrict__ outs) {
    __shared__ uint32_t s_pre[WG];
    const int ps = seq & 1;
    const uint32_t n = ctl->st[ps].n, cur = ctl->st[ps].cur;
    const uint64_t* __restrict__ keys = bf.keys[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * (uint32_t)MC_CHUNK + t * MC_ITEMS;
    uint32_t own = 0;
    for (int k = 0; k < MC_ITEMS; ++k)
        if (base + k < n && newc[base + k] != 0.0) ++own;
    s_pre[t] = own;
    __syncthreads();
    if (t == 0) {
        uint32_t r = outs[blockIdx.x];
        for (int j = 0; j < WG; ++j) {
            const uint32_t v = s_pre[j];
            s_pre[j] = r;
            r += v;
        }
    }
    __syncthreads();
    uint32_t o = s_pre[t];
    for (int k = 0; k < MC_ITEMS; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const double v = newc[i];
        if (v == 0.0) continue;
        store_key<W>(out_keys, bf.cap, o, load_key<W>(keys, bf.cap, i));
        out_coeffs[o] = v;
        ++o;
    }
}

}  // namespace pp
