// pp_lds.h -- a run of rotations and Pauli noise gates on a small sum held in
// shared memory by one workgroup, many gates per launch.
//
// While the sum is small (a few thousand terms) every kernel of the per-gate
// pipeline is latency-bound: each gate pays three launches, their control
// block reads and their look-back waits for a few tiles of work. Here one
// workgroup loads the sum into shared memory once and applies gate after gate
// to it, with barriers between the passes of a gate instead of launches:
//
//   rotation  A: anticommutation of every term, the touched terms' indices
//                listed in parent order (src/Base/xorbranch.jl:214-240);
//             B: the created terms term ^ m put in ascending order by the
//                rank formula of k_xorperm (block bounds per level from head
//                flags and two block-wide scans, no sort), with their values
//                coefficient x sin, negated for a negative sign
//                (src/Propagation/specializations.jl:40);
//             C: the merge of the sum (coefficients times cos where the term
//                anticommutes) with the created terms along merge paths, equal
//                keys summed head + tail, every output truncated, compacted
//                (src/Base/Merge/sortedtailmerge.jl:140-202);
//   noise     damped codes scaled, every term truncated, compacted
//             (src/Propagation/specializations.jl:218-238).
//
// Every value is formed by the same operations, in the same order, as the
// pipeline's kernels, so the result is bit-identical to it. A rotation that
// touches no term is the identity (nothing truncated). When a gate's output
// would not fit (more than `limit` terms), the run stops before it with the
// sum intact; the host goes on with the per-gate pipeline from that gate.
#pragma once
#include "pp_kernels.h"

namespace pp {

struct LdsGate {
    uint64_t mask[4];       // rotation: generator mask
    double a, b;            // rotation: cos, sin; noise: the damping factor (a)
    int32_t kind;           // 0 rotation, 1 Pauli noise
    int32_t qubit;          // noise: 0-based qubit
    uint32_t damp_codes;    // noise: bit c set when code c is damped
    int32_t nlev;           // rotation: levels of the rank formula (Levels)
    int32_t top_runs;
    uint32_t slot;          // the gate's log entry: glog + LOG_WORDS * slot
    int16_t bit[MAX_LEVELS];
    int16_t sgn[MAX_LEVELS];
};

// Terms a run holds (the two sum buffers, the sorted created terms and the
// list of touched indices must fit 160 KiB of shared memory) and the items a
// thread takes of one pass.
template <int W>
struct LdsCfg {
    static constexpr int NL = W == 1 ? 2944 : (W == 2 ? 1984 : 1152);
    static constexpr int KM = (NL + WG - 1) / WG;
};

constexpr int LDS_GCHUNK = 32;  // gate descriptors staged in shared memory at a time

template <int W>
__global__ __launch_bounds__(WG) void k_lds_run(Bufs bf, Ctl* ctl, uint32_t seq0, const LdsGate* __restrict__ gates,
                                                int ngates, uint32_t limit, int max_weight, double min_abs,
                                                uint32_t* __restrict__ glog) {
    constexpr int NL = LdsCfg<W>::NL, KM = LdsCfg<W>::KM;
    __shared__ uint64_t s_k[2][W][NL];
    __shared__ double s_v[2][NL];
    __shared__ uint64_t t_k[W][NL];
    __shared__ double t_v[NL];
    __shared__ uint16_t s_li[NL];
    __shared__ uint32_t s_w[NWAVES];
    __shared__ int s_lmax[MAX_LEVELS][NWAVES], s_lmin[MAX_LEVELS][NWAVES];
    __shared__ LdsGate s_g[LDS_GCHUNK];
    static_assert(sizeof(LdsGate) % 8 == 0, "descriptors are staged in 8-byte words");
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const int p0 = seq0 & 1;
    const CtlView c = load_ctl(ctl, p0);
    if (c.overflow) return;
    const uint32_t cur = c.s.cur;
    uint32_t n = c.s.n;
    if (n > limit || limit > (uint32_t)NL) {
        if (t == 0) {
            atomicOr(&ctl->err, 8u);
            ctl->lds_done = 0u;
        }
        return;
    }
    const size_t stride = bf.cap;
    uint64_t* __restrict__ gkeys = bf.keys[cur];
    double* __restrict__ gcoef = bf.coeffs[cur];
    for (uint32_t i = t; i < n; i += WG) {
#pragma unroll
        for (int l = 0; l < W; ++l) s_k[0][l][i] = gkeys[l * stride + i];
        s_v[0][i] = gcoef[i];
    }
    __syncthreads();
    int buf = 0;
    int gdone = 0;
    auto Skey = [&](int b, uint32_t x) {
        Key<W> k;
#pragma unroll
        for (int l = 0; l < W; ++l) k.l[l] = s_k[b][l][x];
        return k;
    };
    auto Tkey = [&](uint32_t y) {
        Key<W> k;
#pragma unroll
        for (int l = 0; l < W; ++l) k.l[l] = t_k[l][y];
        return k;
    };
    for (int g = 0; g < ngates; ++g) {
        if (g % LDS_GCHUNK == 0) {
            // the next descriptors, staged together (one round trip per chunk)
            const int cnt = min(LDS_GCHUNK, ngates - g);
            constexpr int WORDS = (int)(sizeof(LdsGate) / 8);
            const uint64_t* __restrict__ src = reinterpret_cast<const uint64_t*>(gates + g);
            uint64_t* dst = reinterpret_cast<uint64_t*>(s_g);
            __syncthreads();
            for (int w = t; w < cnt * WORDS; w += WG) dst[w] = src[w];
            __syncthreads();
        }
        const LdsGate& G = s_g[g % LDS_GCHUNK];
        const int kind = G.kind;
        uint32_t* lg = glog + (size_t)LOG_WORDS * G.slot;
        if (kind == 1) {
            // ---- Pauli noise: scale the damped codes, truncate, compact
            const int qubit = G.qubit;
            const uint32_t codes = G.damp_codes;
            const double f = G.a;
            const uint32_t kn = (n + WG - 1) / WG;
            const uint32_t i0 = min(n, t * kn), i1 = min(n, i0 + kn);
            uint32_t cnt = 0;
            for (uint32_t i = i0; i < i1; ++i) {
                const Key<W> k = Skey(buf, i);
                double cv = s_v[buf][i];
                if ((codes >> code_at<W>(k, qubit)) & 1u) cv = cv * f;
                cnt += !((max_weight >= 0 && weight<W>(k) > max_weight) || fabs(cv) < min_abs) ? 1u : 0u;
            }
            uint32_t tot;
            uint32_t pos = block_excl_sum(cnt, s_w, tot);
            for (uint32_t i = i0; i < i1; ++i) {
                const Key<W> k = Skey(buf, i);
                double cv = s_v[buf][i];
                if ((codes >> code_at<W>(k, qubit)) & 1u) cv = cv * f;
                if (!((max_weight >= 0 && weight<W>(k) > max_weight) || fabs(cv) < min_abs)) {
#pragma unroll
                    for (int l = 0; l < W; ++l) s_k[buf ^ 1][l][pos] = k.l[l];
                    s_v[buf ^ 1][pos] = cv;
                    ++pos;
                }
            }
            __syncthreads();
            if (t == 0) {
                lg[0] = n;
                lg[1] = n;
                lg[2] = tot;
                lg[3] = lg[4] = 0u;
            }
            n = tot;
            buf ^= 1;
            ++gdone;
            continue;
        }
        // ---- rotation
        Key<W> mask;
#pragma unroll
        for (int l = 0; l < W; ++l) mask.l[l] = G.mask[l];
        const double cosv = G.a, sinv = G.b;
        // A: touched terms, listed in parent order
        uint32_t na;
        {
            const uint32_t ka = (n + WG - 1) / WG;
            const uint32_t i0 = min(n, t * ka), i1 = min(n, i0 + ka);
            uint32_t cnt = 0;
            for (uint32_t i = i0; i < i1; ++i) cnt += anticommutes<W>(mask, Skey(buf, i)) ? 1u : 0u;
            uint32_t pos = block_excl_sum(cnt, s_w, na);
            for (uint32_t i = i0; i < i1; ++i)
                if (anticommutes<W>(mask, Skey(buf, i))) s_li[pos++] = (uint16_t)i;
            __syncthreads();
        }
        if (na == 0) {
            // the identity: nothing created, nothing truncated
            if (t == 0) {
                lg[0] = n;
                lg[1] = 0u;
                lg[2] = n;
                lg[3] = lg[4] = 0u;
            }
            ++gdone;
            continue;
        }
        // B: rank of every created term in ascending order of term ^ mask
        {
            const uint32_t kr = (na + WG - 1) / WG;
            const uint32_t j0 = min(na, t * kr);
            const int nv = (int)(min(na, j0 + kr) - j0);
            Key<W> key[KM];
            int64_t dest[KM];
#pragma unroll
            for (int k = 0; k < KM; ++k) {
                key[k] = Skey(buf, s_li[min(j0 + (uint32_t)k, na - 1)]);
                dest[k] = (int64_t)(j0 + k) + (int64_t)G.top_runs * na;
            }
            const Key<W> prevk = Skey(buf, s_li[j0 > 0 ? j0 - 1 : 0]);
            const int nlev = G.nlev;
            for (int v = 0; v < nlev; ++v) {
                const int bit = G.bit[v];
                const int sg = G.sgn[v];
                bool h[KM];
#pragma unroll
                for (int k = 0; k < KM; ++k) {
                    const uint32_t j = j0 + k;
                    const Key<W>& pk = (k == 0) ? prevk : key[k > 0 ? k - 1 : 0];
                    h[k] = k < nv && (j == 0 || !eq_from<W>(key[k], pk, bit));
                }
                int run[KM], nxt[KM];
                int r = -1;
#pragma unroll
                for (int k = 0; k < KM; ++k) {
                    if (h[k]) r = (int)(j0 + k);
                    run[k] = r;
                }
                int q = INF_POS;
#pragma unroll
                for (int k = KM - 1; k >= 0; --k) {
                    nxt[k] = q;
                    if (h[k]) q = (int)(j0 + k);
                }
                // exclusive max over earlier threads of their last head, and min
                // over later threads of their first head
                int mx = r, mn = q;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int u = __shfl_up(mx, o, 64);
                    const int d = __shfl_down(mn, o, 64);
                    if (lane >= o) mx = max(mx, u);
                    if (lane + o < 64) mn = min(mn, d);
                }
                int pre = __shfl_up(mx, 1, 64);
                int suf = __shfl_down(mn, 1, 64);
                if (lane == 0) pre = -1;
                if (lane == 63) suf = INF_POS;
                if (lane == 63) s_lmax[v][wave] = mx;
                if (lane == 0) s_lmin[v][wave] = mn;
                __syncthreads();
                for (int w = 0; w < wave; ++w) pre = max(pre, s_lmax[v][w]);
                for (int w = wave + 1; w < NWAVES; ++w) suf = min(suf, s_lmin[v][w]);
#pragma unroll
                for (int k = 0; k < KM; ++k) {
                    const int s = run[k] >= 0 ? run[k] : pre;
                    const int e = nxt[k] < INF_POS ? nxt[k] : (suf < INF_POS ? suf : (int)na);
                    dest[k] += (int64_t)sg * ((int64_t)s + (int64_t)e);
                }
            }
#pragma unroll
            for (int k = 0; k < KM; ++k) {
                if (k < nv) {
                    const int64_t d = dest[k];
                    if (d >= 0 && d < (int64_t)na) {
                        double val = s_v[buf][s_li[j0 + k]] * sinv;
                        if ((sign_exponent<W>(mask, key[k]) & 2) == 0) val = -val;
#pragma unroll
                        for (int l = 0; l < W; ++l) t_k[l][d] = key[k].l[l] ^ mask.l[l];
                        t_v[d] = val;
                    } else {
                        atomicOr(&ctl->err, 1u);
                    }
                }
            }
            __syncthreads();
        }
        // C: merge along merge paths, truncate, compact; counted first, then written
        const uint32_t M = n + na;
        const uint32_t km = (M + WG - 1) / WG;
        const uint32_t d0 = min(M, t * km), d1 = min(M, d0 + km);
        uint32_t x0;
        {
            uint32_t lo = d0 > na ? d0 - na : 0u, hi = min(d0, n);
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (key_le<W>(Skey(buf, mid), Tkey(d0 - 1 - mid))) lo = mid + 1;
                else hi = mid;
            }
            x0 = lo;
        }
        auto walk = [&](bool write, uint32_t pos) -> uint32_t {
            uint32_t x = x0, y = d0 - x0, cnt = 0;
            for (uint32_t d = d0; d < d1; ++d) {
                const bool hs = x < n, ht = y < na;
                const Key<W> ks = Skey(buf, hs ? x : 0u);
                const Key<W> kt = Tkey(ht ? y : 0u);
                const bool take_s = hs && (!ht || key_le<W>(ks, kt));
                bool live;
                double oc;
                Key<W> ko;
                if (take_s) {
                    const double cv = s_v[buf][x];
                    oc = anticommutes<W>(mask, ks) ? cv * cosv : cv;
                    if (ht && key_eq<W>(kt, ks)) oc = oc + t_v[y];
                    ko = ks;
                    live = true;
                    ++x;
                } else {
                    live = !(x > 0 && key_eq<W>(Skey(buf, x - 1), kt));
                    oc = t_v[y];
                    ko = kt;
                    ++y;
                }
                if (live && !((max_weight >= 0 && weight<W>(ko) > max_weight) || fabs(oc) < min_abs)) {
                    if (write) {
#pragma unroll
                        for (int l = 0; l < W; ++l) s_k[buf ^ 1][l][pos + cnt] = ko.l[l];
                        s_v[buf ^ 1][pos + cnt] = oc;
                    }
                    ++cnt;
                }
            }
            return cnt;
        };
        uint32_t tot;
        const uint32_t base = block_excl_sum(walk(false, 0u), s_w, tot);
        if (tot > limit) break;  // does not fit: stop before this gate, the sum intact
        walk(true, base);
        __syncthreads();
        if (t == 0) {
            lg[0] = n;
            lg[1] = na;
            lg[2] = tot;
            lg[3] = lg[4] = 0u;
        }
        n = tot;
        buf ^= 1;
        ++gdone;
    }
    // the sum back to its buffer, and the state of the gate after the last one run
    for (uint32_t i = t; i < n; i += WG) {
#pragma unroll
        for (int l = 0; l < W; ++l) gkeys[l * stride + i] = s_k[buf][l][i];
        gcoef[i] = s_v[buf][i];
    }
    if (t == 0) {
        const int pe = (int)((seq0 + (uint32_t)gdone) & 1u);
        ctl->st[pe].n = n;
        ctl->st[pe].cur = cur;
        ctl->st[pe].na = 0u;
        ctl->st[pe].run = 0u;
        ctl->st[pe].roff = c.s.roff;
        ctl->st[pe].coff = c.s.coff;
        ctl->tail_ready = 0u;
        ctl->ssamp_for = 0u;
        ctl->lds_done = (uint32_t)gdone;
    }
}

}  // namespace pp
