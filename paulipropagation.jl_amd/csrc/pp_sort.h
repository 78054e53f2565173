// pp_sort.h -- stable LSD radix sort of (u64 key, u32 value) pairs and the
// Clifford step built on it.
//
// A Clifford gate maps every term to another term (a bijection on the Pauli
// codes of its qubits, src/Propagation/specializations.jl:144-192), so the
// sum loses its order but gains no duplicates. Only the bits of the gate's
// qubits change; the sum was sorted, so its terms fall into segments of equal
// bits above the highest changed bit, already in order, and only the order
// inside a segment changes. The map kernel therefore writes, per kept term, the
// sort key (segment index << window bits) | (its new bits from the lowest
// changed bit up to the highest); a stable radix sort of those keys with the
// term's position as the value, and one gather, leave the sum sorted again.
// When the window is wider than 32 bits the key's limbs are sorted instead,
// least significant first (the bits below the window keep their order).
#pragma once
#include "pp_kernels.h"

namespace pp {

constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = WG * RS_ITEMS;
constexpr int RS_BINS = 256;

// Per tile, how many keys have each digit: hist[digit * max_tiles + tile].
__global__ __launch_bounds__(WG) void k_rs_upsweep(const Ctl* ctl, const uint32_t* n_ptr,
                                                   const uint64_t* __restrict__ keys, int shift, int bits,
                                                   uint32_t* __restrict__ hist, uint32_t max_tiles) {
    if (ctl->overflow) return;
    const uint32_t n = *n_ptr;
    const uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const uint64_t dmask = (1ull << bits) - 1;
    __shared__ uint32_t s_h[RS_BINS];
    const int t = threadIdx.x;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        s_h[t] = 0;
        __syncthreads();
        const uint32_t base = tile * (uint32_t)RS_TILE;
        uint64_t v[RS_ITEMS];
#pragma unroll
        for (int k = 0; k < RS_ITEMS; ++k) v[k] = keys[min(base + k * WG + t, n - 1)];
#pragma unroll
        for (int k = 0; k < RS_ITEMS; ++k)
            if (base + k * WG + t < n) atomicAdd(&s_h[(v[k] >> shift) & dmask], 1u);
        __syncthreads();
        hist[(size_t)t * max_tiles + tile] = s_h[t];
        __syncthreads();
    }
}

// Exclusive scan of the histogram in digit-major order, in place: the first
// output position of every (digit, tile). One workgroup; thread d owns digit d.
__global__ __launch_bounds__(WG) void k_rs_scan(const Ctl* ctl, const uint32_t* n_ptr, uint32_t* __restrict__ hist,
                                                uint32_t max_tiles) {
    if (ctl->overflow) return;
    const uint32_t n = *n_ptr;
    const uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const int t = threadIdx.x;
    uint32_t* row = hist + (size_t)t * max_tiles;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < ntiles; ++k) sum += row[k];
    __shared__ uint32_t s_w[NWAVES];
    uint32_t total;
    uint32_t run = block_excl_sum(sum, s_w, total);
    for (uint32_t k = 0; k < ntiles; ++k) {
        const uint32_t c = row[k];
        row[k] = run;
        run += c;
    }
}

// Scatter every key (and value) to its position: the first position of its
// (digit, tile) plus its rank among the tile's keys with that digit, in index
// order, so the sort is stable. Keys are taken in striped order (item k of
// thread t is index base + k * WG + t); each wave ranks equal digits with
// ballots.
__global__ __launch_bounds__(WG) void k_rs_downsweep(const Ctl* ctl, const uint32_t* n_ptr,
                                                     const uint64_t* __restrict__ kin,
                                                     const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
                                                     uint32_t* __restrict__ vout, int shift, int bits,
                                                     const uint32_t* __restrict__ hist, uint32_t max_tiles) {
    if (ctl->overflow) return;
    const uint32_t n = *n_ptr;
    const uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const uint64_t dmask = (1ull << bits) - 1;
    __shared__ uint32_t s_off[RS_BINS];
    __shared__ uint32_t s_wc[NWAVES][RS_BINS];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        s_off[t] = hist[(size_t)t * max_tiles + tile];
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) s_wc[w][t] = 0;
        __syncthreads();
        const uint32_t base = tile * (uint32_t)RS_TILE;
        for (int k = 0; k < RS_ITEMS; ++k) {
            const uint32_t i = base + k * WG + t;
            const bool valid = i < n;
            const uint32_t ic = valid ? i : n - 1;
            const uint64_t key = kin[ic];
            const uint32_t val = vin[ic];
            const uint32_t d = (uint32_t)((key >> shift) & dmask);
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < bits; ++b) {
                const uint64_t bb = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? bb : ~bb;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            if (valid && rank == 0) s_wc[wave][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = s_off[d] + rank;
                for (int w = 0; w < wave; ++w) pos += s_wc[w][d];
                kout[pos] = key;
                vout[pos] = val;
            }
            __syncthreads();
            // this row's counts move into the tile's running offsets
            uint32_t row = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; ++w) {
                row += s_wc[w][t];
                s_wc[w][t] = 0;
            }
            s_off[t] += row;
            __syncthreads();
        }
    }
}

// One round of a limb-by-limb sort: key = bits [lo, 64) of limb `limb` of term
// perm[i] (term i when perm is null), value = that term's index. The terms are
// the Clifford's mapped sum (the other buffer of the gate's input, n terms) or,
// with `plain`, buffer `plain_buf` with n at n_ptr.
__global__ __launch_bounds__(WG) void k_rs_limb_keys(const Ctl* ctl, int p, const Bufs bf, int plain, int plain_buf,
                                                     const uint32_t* n_ptr, int limb, int lo,
                                                     const uint32_t* __restrict__ perm, uint64_t* __restrict__ kout,
                                                     uint32_t* __restrict__ vout) {
    if (ctl->overflow) return;
    const uint32_t n = *n_ptr;
    const uint64_t* __restrict__ keys = plain ? bf.keys[plain_buf] : bf.keys[ctl->st[p].cur ^ 1];
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const uint32_t src = perm ? perm[i] : i;
        kout[i] = keys[(size_t)limb * bf.cap + src] >> lo;
        vout[i] = src;
    }
}

struct CliffArgs {
    const uint64_t* change;  // 16 x W limbs: key ^= change[idx]; idx = code(q0) | code(q1) << 2
    const double* sign;      // 16 signs
    int nq, q0, q1;
    int max_weight;   // < 0: nothing is truncated (a Clifford truncates only under a weight limit)
    double min_abs;
    int wlo, wbits;   // sort window: new bits [wlo, wlo + wbits)
    int composite;    // 1: keys (segment << wbits) | window; 0: wide window, limb rounds follow
};

// Map every term (new key, coefficient times the sign), truncate under a weight
// limit, compact into the other buffer in the input's order, and write each kept
// term's sort key and position. Two look-back chains: kept terms, and segment
// starts (a term whose bits above the window differ from its predecessor's).
constexpr int KC_ITEMS = 4;
constexpr int KC_TILE = WG * KC_ITEMS;

template <int W>
__global__ __launch_bounds__(WG) void k_clifford_map(Bufs bf, Ctl* ctl, Launch L, CliffArgs a, int B,
                                                     uint64_t* __restrict__ skey, uint32_t* __restrict__ sval) {
    if (!lb_start(ctl, L)) return;
    const int p = L.seq & 1;
    const uint32_t n = ctl->st[p].n, cur = ctl->st[p].cur;
    const int t = threadIdx.x;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) {
            set_state(ctl, p ^ 1, 0u, cur ^ 1);
            if (L.glog) L.glog[0] = L.glog[1] = L.glog[2] = 0u;
            carry_records(ctl, L, p);
        }
        return;
    }
    // recording (adjoint sweep): term i's output position and sign, one word
    // each (the positions are those of the mapped sum until the sort is done)
    const uint32_t roff = ctl->st[p].roff, coff = ctl->st[p].coff;
    if (bf.rec && (uint64_t)roff + n > bf.rec_words) {
        if (blockIdx.x == 0 && t == 0) record_overflow(ctl, L, (uint64_t)roff + n, coff);
        return;
    }
    const uint64_t* __restrict__ keys = bf.keys[cur];
    const double* __restrict__ coeffs = bf.coeffs[cur];
    uint64_t* __restrict__ out_keys = bf.keys[cur ^ 1];
    double* __restrict__ out_coeffs = bf.coeffs[cur ^ 1];
    const size_t stride = bf.cap;
    const uint32_t ntiles = (n + KC_TILE - 1) / KC_TILE;
    const int top = a.wlo + a.wbits;  // first bit above the window
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl, s_excl_s;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t i0 = tile * (uint32_t)KC_TILE + t * KC_ITEMS;  // blocked: thread t owns i0 .. i0 + 3
        Key<W> kv[KC_ITEMS];
        uint32_t keep = 0, segs = 0;
        int idxs[KC_ITEMS];
        Key<W> prev = load_key<W>(keys, stride, i0 > 0 ? min(i0 - 1, n - 1) : 0);
#pragma unroll
        for (int k = 0; k < KC_ITEMS; ++k) kv[k] = load_key<W>(keys, stride, min(i0 + k, n - 1));
#pragma unroll
        for (int k = 0; k < KC_ITEMS; ++k) {
            const uint32_t i = i0 + k;
            const bool in = i < n;
            // segment start: bits above the window differ from the previous term's
            if (in && i > 0 && !eq_from<W>(kv[k], prev, top)) segs |= 1u << k;
            prev = kv[k];
            int idx = (int)code_at<W>(kv[k], a.q0);
            if (a.nq == 2) idx |= (int)code_at<W>(kv[k], a.q1) << 2;
            idxs[k] = idx;
#pragma unroll
            for (int l = 0; l < W; ++l) kv[k].l[l] ^= a.change[idx * W + l];
            bool kept = in;
            if (in && a.max_weight >= 0) {
                double m = 0.0;
                for (int b = 0; b < B; ++b) m = fmax(m, fabs(coeffs[(size_t)i * B + b] * a.sign[idx]));
                kept = !(weight<W>(kv[k]) > a.max_weight || m < a.min_abs);
            }
            if (kept) keep |= 1u << k;
        }
        const uint32_t mine = (uint32_t)__popc(keep) | ((uint32_t)__popc(segs) << 16);
        uint32_t agg;
        static_assert(KC_ITEMS < 8, "per-thread counts fit 3 bits");
        const uint32_t pre = block_excl_sum_small<3>(mine, s_w, agg);
        {
            const uint64_t a2[2] = {agg & 0xffffu, agg >> 16};
            const uint64_t ex2 = tile_prefix<2>(ctl, L, bf, tile, a2[0] | (a2[1] << 32), lb_incl, s_wacc);
            const uint64_t e2[2] = {ex2 & 0xffffffffull, ex2 >> 32};
            if (t == 0) {
                s_excl = e2[0];
                s_excl_s = e2[1];
                if (tile == ntiles - 1) {
                    const uint32_t nout = (uint32_t)(e2[0] + a2[0]);
                    set_state(ctl, p ^ 1, nout, cur ^ 1);
                    if (bf.rec) {
                        ctl->st[p ^ 1].roff = roff + n;
                        ctl->st[p ^ 1].coff = coff;
                        if (L.glog) {
                            L.glog[3] = roff;
                            L.glog[4] = coff;
                        }
                    } else {
                        carry_records(ctl, L, p);
                    }
                    if (L.glog) {
                        L.glog[0] = n;
                        L.glog[1] = n;
                        L.glog[2] = nout;
                    }
                }
            }
        }
        __syncthreads();
        uint64_t pos = s_excl + (pre & 0xffffu);
        uint64_t seg = s_excl_s + (pre >> 16);
#pragma unroll
        for (int k = 0; k < KC_ITEMS; ++k) {
            seg += (segs >> k) & 1u;
            const uint32_t i = i0 + k;
            const bool kept = (keep >> k) & 1u;
            if (bf.rec && i < n)
                bf.rec[(size_t)roff + i] = (a.sign[idxs[k]] < 0.0 ? REC_FLAG : 0u) | (kept ? (uint32_t)pos : REC_NONE);
            if (!kept) continue;
            store_key<W>(out_keys, stride, pos, kv[k]);
            for (int b = 0; b < B; ++b) out_coeffs[pos * B + b] = coeffs[(size_t)i * B + b] * a.sign[idxs[k]];
            if (a.composite) {
                uint64_t win = 0;
#pragma unroll
                for (int l = 0; l < W; ++l) {
                    // bits [wlo, wlo + wbits) of the new key, wbits <= 32
                    const int lo = l * 64;
                    if (a.wlo >= lo + 64 || a.wlo + a.wbits <= lo) continue;
                    const int sh = a.wlo - lo;
                    win |= sh >= 0 ? (kv[k].l[l] >> sh) : (kv[k].l[l] << (-sh));
                }
                win &= (1ull << a.wbits) - 1;
                skey[pos] = (seg << a.wbits) | win;
                sval[pos] = (uint32_t)pos;
            }
            ++pos;
        }
    }
}

// Rows of the sorted sum: term i is the mapped term perm[i]. The mapped sum is
// in the other buffer of the gate's input; the sorted one goes back to the
// input's buffer.
template <int W>
__global__ __launch_bounds__(WG) void k_clifford_gather(Bufs bf, Ctl* ctl, int p, const uint32_t* __restrict__ perm,
                                                        int B) {
    if (ctl->overflow) return;
    const uint32_t n = ctl->st[p ^ 1].n;
    const uint32_t dst = ctl->st[p].cur, src = dst ^ 1;
    const uint64_t* __restrict__ skeys = bf.keys[src];
    const double* __restrict__ scoef = bf.coeffs[src];
    uint64_t* __restrict__ dkeys = bf.keys[dst];
    double* __restrict__ dcoef = bf.coeffs[dst];
    const size_t stride = bf.cap;
    const bool rec = bf.rec && bf.tpj;
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const uint32_t s = perm[i];
        store_key<W>(dkeys, stride, i, load_key<W>(skeys, stride, s));
        for (int b = 0; b < B; ++b) dcoef[(size_t)i * B + b] = scoef[(size_t)s * B + b];
        if (rec) bf.tpj[s] = i;  // mapped term s ends at i
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->st[p ^ 1].cur = dst;
}

// Recording only: the records of a Clifford's input terms point at the mapped
// sum; after the sort they are moved to the sorted positions (tpj, from the
// gather). The sign stays in REC_FLAG.
__global__ __launch_bounds__(WG) void k_clifford_rec_fix(Bufs bf, const Ctl* ctl, int p) {
    if (ctl->overflow) return;
    const uint32_t n = ctl->st[p].n, roff = ctl->st[p].roff;
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const uint32_t w = bf.rec[(size_t)roff + i], at = w & SRC_MASK;
        if (at != REC_NONE) bf.rec[(size_t)roff + i] = (w & REC_FLAG) | bf.tpj[at];
    }
}

// ============================================================ k_reduce_runs
//
// Terms of buffer 0 (n at n_ptr, any order, equal strings allowed), in the
// stable order perm: every run of equal strings becomes one term whose
// coefficients are the run's summed left to right (input order), truncated,
// and compacted in ascending order into buffer 1; the count goes to st[p ^ 1]
// (mergeandtruncate! on an unsorted vector, src/Base/Merge/mergeandtruncate.jl:104-120).
constexpr int KR_ITEMS = 4;
constexpr int KR_TILE = WG * KR_ITEMS;

template <int W>
__global__ __launch_bounds__(WG) void k_reduce_runs(Bufs bf, Ctl* ctl, Launch L, const uint32_t* __restrict__ perm,
                                                    int B, int max_weight, double min_abs) {
    if (!lb_start(ctl, L)) return;
    const int p = L.seq & 1;
    const uint32_t n = ctl->st[p].n;
    const uint64_t* __restrict__ keys = bf.keys[0];
    const double* __restrict__ coeffs = bf.coeffs[0];
    uint64_t* __restrict__ out_keys = bf.keys[1];
    double* __restrict__ out_coeffs = bf.coeffs[1];
    const size_t stride = bf.cap;
    const int t = threadIdx.x;
    if (n == 0) {
        if (blockIdx.x == 0 && t == 0) set_state(ctl, p ^ 1, 0u, 1u);
        return;
    }
    const uint32_t ntiles = (n + KR_TILE - 1) / KR_TILE;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_w[NWAVES];
    __shared__ uint64_t s_excl;
    TileSource ts;
    uint64_t lb_incl = 0;  // static tiles: this workgroup's inclusive prefix
    __shared__ uint64_t s_wacc[NWAVES];
    while (true) {
        const uint32_t tile = ts.take(L, &ctl->counter[L.slot], &s_tile);
        if (tile >= ntiles) break;
        const uint32_t i0 = tile * (uint32_t)KR_TILE + t * KR_ITEMS;
        uint32_t keep = 0;
        Key<W> prev = load_key<W>(keys, stride, perm[i0 > 0 ? min(i0 - 1, n - 1) : 0]);
        Key<W> kk[KR_ITEMS];
#pragma unroll
        for (int k = 0; k < KR_ITEMS; ++k) {
            const uint32_t i = i0 + k;
            kk[k] = load_key<W>(keys, stride, perm[min(i, n - 1)]);
            const bool head = i < n && (i == 0 || !key_eq<W>(kk[k], prev));
            prev = kk[k];
            if (!head) continue;
            // the run's sum, left to right, and its largest magnitude over the batch
            double m = 0.0;
            for (int b = 0; b < B; ++b) {
                double c = coeffs[(size_t)perm[i] * B + b];
                for (uint32_t j = i + 1; j < n; ++j) {
                    const uint32_t pj = perm[j];
                    if (!key_eq<W>(load_key<W>(keys, stride, pj), kk[k])) break;
                    c = c + coeffs[(size_t)pj * B + b];
                }
                m = fmax(m, fabs(c));
            }
            if (!((max_weight >= 0 && weight<W>(kk[k]) > max_weight) || m < min_abs)) keep |= 1u << k;
        }
        uint32_t agg;
        static_assert(KR_ITEMS < 8, "per-thread counts fit 3 bits");
        const uint32_t pre = block_excl_sum_small<3>((uint32_t)__popc(keep), s_w, agg);
        {
            const uint64_t excl = tile_prefix<1>(ctl, L, bf, tile, agg, lb_incl, s_wacc);
            if (t == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) set_state(ctl, p ^ 1, (uint32_t)(excl + agg), 1u);
            }
        }
        __syncthreads();
        uint64_t pos = s_excl + pre;
#pragma unroll
        for (int k = 0; k < KR_ITEMS; ++k) {
            if (!((keep >> k) & 1u)) continue;
            const uint32_t i = i0 + k;
            store_key<W>(out_keys, stride, pos, kk[k]);
            for (int b = 0; b < B; ++b) {
                double c = coeffs[(size_t)perm[i] * B + b];
                for (uint32_t j = i + 1; j < n; ++j) {
                    const uint32_t pj = perm[j];
                    if (!key_eq<W>(load_key<W>(keys, stride, pj), kk[k])) break;
                    c = c + coeffs[(size_t)pj * B + b];
                }
                out_coeffs[pos * B + b] = c;
            }
            ++pos;
        }
    }
}

}  // namespace pp
