/*
 * ppo.c -- CPU oracle: a plain-C restatement of the VectorPauliSum
 * propagate() path of PauliPropagation.jl v0.9.2.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, as the checker and as the timed host
 * baseline. The product library never links or calls it.
 *
 * What each function follows (paths relative to the reference checkout):
 *   anticommutation          src/PauliAlgebra/wide_bitoperations.jl:16-25
 *   sign exponent            src/PauliAlgebra/paulioperations.jl:518-564 (restated
 *                            here from the single-qubit product table, see below)
 *   rotation rule            src/Propagation/specializations.jl:34-44, 62-73
 *   branch (append in parent order)   src/Base/xorbranch.jl:214-240
 *   nothing touched => nothing merged or truncated   src/Base/Merge/xortailmerge.jl:121-123
 *   XOR block-reversal tail sort      src/Base/Merge/xortailmerge.jl:317-452
 *   head/tail merge, truncation of every output  src/Base/Merge/sortedtailmerge.jl:140-202
 *   full sort + run reduction fallback           src/Base/Merge/mergeandtruncate.jl:87-120,
 *                                                src/Base/Merge/deduplicate.jl:14-104
 *   truncation predicate     src/Propagation/generics.jl:223-250, src/truncations.jl:42-65
 *   PauliNoise               src/Propagation/specializations.jl:218-238,
 *                            src/Base/Primitives/mapandtruncate.jl:115-197
 *   AmplitudeDampingNoise    src/Propagation/specializations.jl:251-292
 *   CliffordGate             src/Propagation/specializations.jl:144-192
 *   overlapwithzero          src/stateoverlap.jl:12-35, src/Base/Primitives/mapreduce.jl:43-65
 *
 * Batched columns (B > 1) share one key set; a term is truncated when its
 * weight exceeds max_weight or when max_b |c_b| < min_abs. For B = 1 this is
 * exactly the reference predicate.
 *
 * Compile with -ffp-contract=off: every coefficient is one product or a sum of
 * two products, written out in the reference's order.
 */
#include "ppo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ALT 0x5555555555555555ULL
#define INLINE static inline __attribute__((always_inline))

struct ppo_sum {
    int W, B;
    size_t n, cap;
    uint64_t *keys;
    double *coeffs;
    int sorted; /* the whole active set is ascending and duplicate-free */
    /* working buffers kept between gates (a gate allocates nothing once they
     * are large enough): the merge output, swapped with keys/coeffs; the
     * created terms of a branching gate; the XOR sort's second buffer */
    uint64_t *akeys, *tkeys, *xkeys;
    double *acoeffs, *tcoeffs, *xcoeffs;
    size_t acap, tcap, xcap;
};

/* ------------------------------------------------------------------ */
/* bit algebra                                                        */
/* ------------------------------------------------------------------ */

INLINE uint64_t swap_pairs(uint64_t w) { return ((w >> 1) & ALT) | ((w & ALT) << 1); }

INLINE int anticommutes_w(const uint64_t *m, const uint64_t *t, int W) {
    uint64_t f = 0;
    for (int l = 0; l < W; ++l) f ^= swap_pairs(m[l]) & t[l];
    return __builtin_popcountll(f) & 1;
}

/* Phase exponent e of P1*P2 = i^e (P1 xor P2), from the per-qubit table.
 * With x = low ^ high, z = high for each 2-bit code, sigma_{x,z} = i^{xz} X^x Z^z,
 * so sigma_a sigma_b = i^{x1 z1 + x2 z2 + 2 z1 x2 - x3 z3} sigma_{x3,z3}. */
INLINE int sign_exponent_w(const uint64_t *p1, const uint64_t *p2, int W) {
    int e = 0;
    for (int l = 0; l < W; ++l) {
        uint64_t a = p1[l], b = p2[l];
        uint64_t z1 = (a >> 1) & ALT, x1 = (a ^ (a >> 1)) & ALT;
        uint64_t z2 = (b >> 1) & ALT, x2 = (b ^ (b >> 1)) & ALT;
        uint64_t x3 = x1 ^ x2, z3 = z1 ^ z2;
        e += __builtin_popcountll(x1 & z1) + __builtin_popcountll(x2 & z2) +
             2 * __builtin_popcountll(z1 & x2) - __builtin_popcountll(x3 & z3);
    }
    return e & 3;
}

INLINE int weight_w(const uint64_t *a, int W) {
    int c = 0;
    for (int l = 0; l < W; ++l) c += __builtin_popcountll((a[l] | (a[l] >> 1)) & ALT);
    return c;
}

INLINE int contains_xy_w(const uint64_t *a, int W) {
    for (int l = 0; l < W; ++l)
        if ((a[l] ^ (a[l] >> 1)) & ALT) return 1;
    return 0;
}

INLINE int key_cmp(const uint64_t *a, const uint64_t *b, int W) {
    for (int l = W - 1; l >= 0; --l)
        if (a[l] != b[l]) return a[l] < b[l] ? -1 : 1;
    return 0;
}

INLINE int key_eq(const uint64_t *a, const uint64_t *b, int W) {
    for (int l = 0; l < W; ++l)
        if (a[l] != b[l]) return 0;
    return 1;
}

/* equal on every bit at position >= bit */
INLINE int eq_from(const uint64_t *a, const uint64_t *b, int bit, int W) {
    int L = bit >> 6, off = bit & 63;
    if (L >= W) return 1;
    for (int l = W - 1; l > L; --l)
        if (a[l] != b[l]) return 0;
    return (a[L] >> off) == (b[L] >> off);
}

INLINE int get_pauli(const uint64_t *a, int q) { return (int)((a[q >> 5] >> (2 * (q & 31))) & 3); }

INLINE int truncated(const uint64_t *key, const double *c, int W, int B, int max_weight,
                     double min_abs) {
    if (max_weight >= 0 && weight_w(key, W) > max_weight) return 1;
    double m = fabs(c[0]);
    for (int b = 1; b < B; ++b) {
        double a = fabs(c[b]);
        if (a > m) m = a;
    }
    return m < min_abs;
}

int ppo_anticommutes(const uint64_t *a, const uint64_t *b, int W) { return anticommutes_w(a, b, W); }
int ppo_sign_exponent(const uint64_t *p1, const uint64_t *p2, int W) { return sign_exponent_w(p1, p2, W); }
int ppo_weight(const uint64_t *a, int W) { return weight_w(a, W); }
int ppo_contains_xy(const uint64_t *a, int W) { return contains_xy_w(a, W); }

/* ------------------------------------------------------------------ */
/* storage                                                            */
/* ------------------------------------------------------------------ */

static void *xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 8);
    if (!p) abort();
    return p;
}

static void reserve(ppo_sum *s, size_t cap) {
    if (cap <= s->cap) return;
    size_t nc = s->cap ? s->cap : 16;
    while (nc < cap) nc = nc + nc / 2 + 16;
    s->keys = realloc(s->keys, nc * s->W * sizeof(uint64_t));
    s->coeffs = realloc(s->coeffs, nc * s->B * sizeof(double));
    if (!s->keys || !s->coeffs) abort();
    s->cap = nc;
}

ppo_sum *ppo_create(int W, int B) {
    if (!(W == 1 || W == 2 || W == 4) || B < 1) return NULL;
    ppo_sum *s = calloc(1, sizeof(ppo_sum));
    s->W = W;
    s->B = B;
    s->sorted = 1;
    return s;
}

void ppo_destroy(ppo_sum *s) {
    if (!s) return;
    free(s->keys);
    free(s->coeffs);
    free(s->akeys);
    free(s->acoeffs);
    free(s->tkeys);
    free(s->tcoeffs);
    free(s->xkeys);
    free(s->xcoeffs);
    free(s);
}

/* room for `need` rows in a working buffer pair (grown by half again) */
static void ensure_buf(uint64_t **k, double **c, size_t *cap, size_t need, int W, int B) {
    if (need <= *cap) return;
    size_t nc = *cap ? *cap : 16;
    while (nc < need) nc = nc + nc / 2 + 16;
    free(*k);
    free(*c);
    *k = malloc(nc * W * sizeof(uint64_t));
    *c = malloc(nc * B * sizeof(double));
    if (!*k || !*c) abort();
    *cap = nc;
}

size_t ppo_size(const ppo_sum *s) { return s->n; }
int ppo_sorted(const ppo_sum *s) { return s->sorted; }

void ppo_get(const ppo_sum *s, uint64_t *keys, double *coeffs) {
    if (keys) memcpy(keys, s->keys, s->n * s->W * sizeof(uint64_t));
    if (coeffs) memcpy(coeffs, s->coeffs, s->n * s->B * sizeof(double));
}

/* index sort by key (ties by index, so equal keys keep their order) */
static const uint64_t *g_sort_keys;
static int g_sort_W;
static int idx_cmp(const void *pa, const void *pb) {
    size_t a = *(const size_t *)pa, b = *(const size_t *)pb;
    int c = key_cmp(g_sort_keys + a * g_sort_W, g_sort_keys + b * g_sort_W, g_sort_W);
    if (c) return c;
    return a < b ? -1 : (a > b);
}

static size_t *sorted_index(const uint64_t *keys, size_t n, int W) {
    size_t *idx = xmalloc(n * sizeof(size_t));
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    g_sort_keys = keys;
    g_sort_W = W;
    qsort(idx, n, sizeof(size_t), idx_cmp);
    return idx;
}

int ppo_set(ppo_sum *s, const uint64_t *keys, const double *coeffs, size_t n) {
    int W = s->W, B = s->B;
    size_t *idx = sorted_index(keys, n, W);
    for (size_t i = 1; i < n; ++i)
        if (key_eq(keys + idx[i] * W, keys + idx[i - 1] * W, W)) {
            free(idx);
            return -1;
        }
    reserve(s, n);
    for (size_t i = 0; i < n; ++i) {
        memcpy(s->keys + i * W, keys + idx[i] * W, W * sizeof(uint64_t));
        memcpy(s->coeffs + i * B, coeffs + idx[i] * B, B * sizeof(double));
    }
    free(idx);
    s->n = n;
    s->sorted = 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* XOR block-reversal tail sort                                        */
/* ------------------------------------------------------------------ */

/* maximal runs of adjacent set bits of the mask, highest first */
static int mask_groups(const uint64_t *mask, int W, int *lo, int *hi, int maxg) {
    int g = 0;
    int b = 64 * W - 1;
    while (b >= 0) {
        if ((mask[b >> 6] >> (b & 63)) & 1) {
            int top = b;
            while (b >= 0 && ((mask[b >> 6] >> (b & 63)) & 1)) --b;
            if (g < maxg) {
                hi[g] = top;
                lo[g] = b + 1;
            }
            ++g;
        } else {
            --b;
        }
    }
    return g;
}

/* One pass: within every run agreeing above `hi`, write the sub-blocks that
 * agree from `lo` up in reverse order; order inside a sub-block is kept. */
static void xor_pass(const uint64_t *ik, const double *ic, uint64_t *ok, double *oc, size_t n,
                     int W, int B, int lo, int hi) {
    size_t rs = 0;
    while (rs < n) {
        size_t re = rs + 1;
        while (re < n && eq_from(ik + re * W, ik + rs * W, hi + 1, W)) ++re;
        /* sub-blocks of [rs, re), written back to front */
        size_t w = rs;
        size_t be = re;
        while (be > rs) {
            size_t bs = be - 1;
            while (bs > rs && eq_from(ik + (bs - 1) * W, ik + (be - 1) * W, lo, W)) --bs;
            memcpy(ok + w * W, ik + bs * W, (be - bs) * W * sizeof(uint64_t));
            memcpy(oc + w * B, ic + bs * B, (be - bs) * B * sizeof(double));
            w += be - bs;
            be = bs;
        }
        rs = re;
    }
}

/* The XOR passes over (keys, coeffs), ping-ponging with (tk, tc); *rk, *rc
 * receive whichever pair holds the sorted tail. Returns the pass count, or
 * -1 for more than 256 groups. */
static int xor_sort_buf(uint64_t *keys, double *coeffs, uint64_t *tk, double *tc, size_t n, int W, int B,
                        const uint64_t *mask, uint64_t **rk, double **rc) {
    int lo[256], hi[256];
    int g = mask_groups(mask, W, lo, hi, 256);
    if (g > 256) return -1;
    uint64_t *ak = keys, *bk = tk;
    double *ac = coeffs, *bc = tc;
    for (int i = 0; i < g; ++i) {
        xor_pass(ak, ac, bk, bc, n, W, B, lo[i], hi[i]);
        uint64_t *t1 = ak; ak = bk; bk = t1;
        double *t2 = ac; ac = bc; bc = t2;
    }
    *rk = ak;
    *rc = ac;
    return g;
}

int ppo_xor_sort_tail(uint64_t *keys, double *coeffs, size_t n, int W, int B, const uint64_t *mask) {
    uint64_t *tk = xmalloc(n * W * sizeof(uint64_t));
    double *tc = xmalloc(n * B * sizeof(double));
    uint64_t *rk;
    double *rc;
    int g = xor_sort_buf(keys, coeffs, tk, tc, n, W, B, mask, &rk, &rc);
    if (g >= 0 && rk != keys) {
        memcpy(keys, rk, n * W * sizeof(uint64_t));
        memcpy(coeffs, rc, n * B * sizeof(double));
    }
    free(tk);
    free(tc);
    return g;
}

/* ------------------------------------------------------------------ */
/* merge with truncation                                               */
/* ------------------------------------------------------------------ */

/* Two-pointer merge of the sorted head (the sum itself) with a sorted,
 * duplicate-free tail. Equal keys merge as head + tail; every output term is
 * truncated. The result replaces the sum. */
/* One column, key width W known at compile time: the same merge without a
 * call per copied key or coefficient. */
/* a < b on keys of W limbs (W = 2 as one 128-bit compare) */
INLINE int key_lt(const uint64_t *a, const uint64_t *b, const int W) {
    if (W == 1) return a[0] < b[0];
    if (W == 2) {
        const unsigned __int128 x = ((unsigned __int128)a[1] << 64) | a[0];
        const unsigned __int128 y = ((unsigned __int128)b[1] << 64) | b[0];
        return x < y;
    }
    return key_cmp(a, b, W) < 0;
}

INLINE size_t merge_head_tail_1(const uint64_t *hk, const double *hc, size_t nh, const uint64_t *tk,
                                const double *tc, size_t nt, uint64_t *ok, double *oc, int max_weight,
                                double min_abs, const int W) {
    size_t i = 0, j = 0, o = 0;
    while (i < nh || j < nt) {
        const uint64_t *key;
        double v;
        if (j < nt) {
            /* the heads below the next tail entry go out as they are (each still truncated) */
            const uint64_t *tj = tk + j * W;
            while (i < nh && key_lt(hk + i * W, tj, W)) {
                const uint64_t *kk = hk + i * W;
                const double hv = hc[i];
                ++i;
                if ((max_weight >= 0 && weight_w(kk, W) > max_weight) || fabs(hv) < min_abs) continue;
                for (int l = 0; l < W; ++l) ok[o * W + l] = kk[l];
                oc[o] = hv;
                ++o;
            }
        }
        int c = (i < nh && j < nt) ? key_cmp(hk + i * W, tk + j * W, W) : (i < nh ? -1 : 1);
        if (c == 0) {
            key = hk + i * W;
            v = hc[i] + tc[j];
            ++i;
            ++j;
        } else if (c < 0) {
            key = hk + i * W;
            v = hc[i];
            ++i;
        } else {
            key = tk + j * W;
            v = tc[j];
            ++j;
        }
        if ((max_weight >= 0 && weight_w(key, W) > max_weight) || fabs(v) < min_abs) continue;
        for (int l = 0; l < W; ++l) ok[o * W + l] = key[l];
        oc[o] = v;
        ++o;
    }
    return o;
}

static void merge_head_tail(ppo_sum *s, const uint64_t *tk, const double *tc, size_t nt,
                            int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    if (B == 1) {
        const size_t nh = s->n;
        ensure_buf(&s->akeys, &s->acoeffs, &s->acap, nh + nt, W, B);
        size_t o;
        switch (W) {
        case 1: o = merge_head_tail_1(s->keys, s->coeffs, nh, tk, tc, nt, s->akeys, s->acoeffs, max_weight, min_abs, 1); break;
        case 2: o = merge_head_tail_1(s->keys, s->coeffs, nh, tk, tc, nt, s->akeys, s->acoeffs, max_weight, min_abs, 2); break;
        default: o = merge_head_tail_1(s->keys, s->coeffs, nh, tk, tc, nt, s->akeys, s->acoeffs, max_weight, min_abs, 4); break;
        }
        uint64_t *k = s->keys;
        double *c = s->coeffs;
        s->keys = s->akeys;
        s->coeffs = s->acoeffs;
        s->akeys = k;
        s->acoeffs = c;
        const size_t oldcap = s->cap;
        s->cap = s->acap;
        s->acap = oldcap;
        s->n = o;
        s->sorted = 1;
        return;
    }
    size_t nh = s->n;
    ensure_buf(&s->akeys, &s->acoeffs, &s->acap, nh + nt, W, B);
    uint64_t *ok = s->akeys;
    double *oc = s->acoeffs;
    double tmp_small[64];
    double *tmp = B <= 64 ? tmp_small : xmalloc(B * sizeof(double));
    size_t i = 0, j = 0, o = 0;
    while (i < nh || j < nt) {
        const uint64_t *key;
        int c = (i < nh && j < nt) ? key_cmp(s->keys + i * W, tk + j * W, W) : (i < nh ? -1 : 1);
        if (c == 0) {
            key = s->keys + i * W;
            for (int b = 0; b < B; ++b) tmp[b] = s->coeffs[i * B + b] + tc[j * B + b];
            ++i;
            ++j;
        } else if (c < 0) {
            key = s->keys + i * W;
            memcpy(tmp, s->coeffs + i * B, B * sizeof(double));
            ++i;
        } else {
            key = tk + j * W;
            memcpy(tmp, tc + j * B, B * sizeof(double));
            ++j;
        }
        if (truncated(key, tmp, W, B, max_weight, min_abs)) continue;
        memcpy(ok + o * W, key, W * sizeof(uint64_t));
        memcpy(oc + o * B, tmp, B * sizeof(double));
        ++o;
    }
    if (tmp != tmp_small) free(tmp);
    /* the output becomes the sum; the old arrays are the next merge's output */
    s->akeys = s->keys;
    s->acoeffs = s->coeffs;
    s->keys = ok;
    s->coeffs = oc;
    const size_t oldcap = s->cap;
    s->cap = s->acap;
    s->acap = oldcap;
    s->n = o;
    s->sorted = 1;
}

/* Fallback for an unsorted head: sort head and tail together, reduce runs of
 * equal keys left to right (head before tail), truncate every output. */
static void sort_reduce_truncate(ppo_sum *s, const uint64_t *tk, const double *tc, size_t nt,
                                 int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    size_t nh = s->n, n = nh + nt;
    uint64_t *ak = xmalloc(n * W * sizeof(uint64_t));
    double *ac = xmalloc(n * B * sizeof(double));
    memcpy(ak, s->keys, nh * W * sizeof(uint64_t));
    memcpy(ac, s->coeffs, nh * B * sizeof(double));
    if (nt) {
        memcpy(ak + nh * W, tk, nt * W * sizeof(uint64_t));
        memcpy(ac + nh * B, tc, nt * B * sizeof(double));
    }
    size_t *idx = sorted_index(ak, n, W);
    uint64_t *ok = xmalloc((n ? n : 1) * W * sizeof(uint64_t));
    double *oc = xmalloc((n ? n : 1) * B * sizeof(double));
    double *tmp = xmalloc(B * sizeof(double));
    size_t o = 0, r = 0;
    while (r < n) {
        const uint64_t *key = ak + idx[r] * W;
        memcpy(tmp, ac + idx[r] * B, B * sizeof(double));
        size_t q = r + 1;
        while (q < n && key_eq(ak + idx[q] * W, key, W)) {
            for (int b = 0; b < B; ++b) tmp[b] = tmp[b] + ac[idx[q] * B + b];
            ++q;
        }
        r = q;
        if (truncated(key, tmp, W, B, max_weight, min_abs)) continue;
        memcpy(ok + o * W, key, W * sizeof(uint64_t));
        memcpy(oc + o * B, tmp, B * sizeof(double));
        ++o;
    }
    free(tmp);
    free(idx);
    free(ak);
    free(ac);
    free(s->keys);
    free(s->coeffs);
    s->keys = ok;
    s->coeffs = oc;
    s->cap = n ? n : 1;
    s->n = o;
    s->sorted = 1;
}

/* Load unsorted terms that may repeat: sort, reduce equal keys in input
 * order and truncate (mergeandtruncate! on an unsorted array,
 * src/Base/Merge/mergeandtruncate.jl:104-120). */
int ppo_load_merge(ppo_sum *s, const uint64_t *keys, const double *coeffs, size_t n, int max_weight,
                   double min_abs) {
    s->n = 0;
    s->sorted = 0;
    sort_reduce_truncate(s, keys, coeffs, n, max_weight, min_abs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* gates                                                               */
/* ------------------------------------------------------------------ */

/* Shared tail of a branching gate: an empty touch leaves everything as is;
 * a sorted head merges a tail sorted by XOR passes; otherwise sort + reduce. */
static void finish_branch(ppo_sum *s, uint64_t *tk, double *tc, size_t nt, const uint64_t *mask,
                          int max_weight, double min_abs) {
    if (s->sorted) {
        uint64_t *rk;
        double *rc;
        ensure_buf(&s->xkeys, &s->xcoeffs, &s->xcap, nt, s->W, s->B);
        xor_sort_buf(tk, tc, s->xkeys, s->xcoeffs, nt, s->W, s->B, mask, &rk, &rc);
        merge_head_tail(s, rk, rc, nt, max_weight, min_abs);
    } else {
        sort_reduce_truncate(s, tk, tc, nt, max_weight, min_abs);
    }
}

/* The branch half of a rotation: anticommuting terms keep c * cos in place;
 * their created terms, (c * sin) * sign at term ^ mask, go in parent order to
 * the sum's tail buffer (src/Base/xorbranch.jl:214-240). Returns the count. */
INLINE size_t branch_into_tail(ppo_sum *s, const uint64_t *mask, const double *cv, const double *sv, const int W) {
    const int B = s->B;
    const size_t n = s->n;
    ensure_buf(&s->tkeys, &s->tcoeffs, &s->tcap, n, W, B);
    uint64_t *tk = s->tkeys;
    double *tc = s->tcoeffs;
    size_t j = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t *k = s->keys + i * W;
        if (!anticommutes_w(mask, k, W)) continue;
        int e = sign_exponent_w(mask, k, W);
        double sign = (double)((e & 2) - 1);
        for (int l = 0; l < W; ++l) tk[j * W + l] = k[l] ^ mask[l];
        double *c = s->coeffs + i * B;
        for (int b = 0; b < B; ++b) {
            tc[j * B + b] = (c[b] * sv[b]) * sign;
            c[b] = c[b] * cv[b];
        }
        ++j;
    }
    return j;
}

INLINE long rotation_impl(ppo_sum *s, const uint64_t *mask, const double *cv, const double *sv,
                          int max_weight, double min_abs, const int W) {
    const size_t na = branch_into_tail(s, mask, cv, sv, W);
    if (na == 0) return 0;
    finish_branch(s, s->tkeys, s->tcoeffs, na, mask, max_weight, min_abs);
    return (long)na;
}

long ppo_rotation(ppo_sum *s, const uint64_t *mask, const double *cosv, const double *sinv,
                  int max_weight, double min_abs) {
    switch (s->W) {
    case 1: return rotation_impl(s, mask, cosv, sinv, max_weight, min_abs, 1);
    case 2: return rotation_impl(s, mask, cosv, sinv, max_weight, min_abs, 2);
    default: return rotation_impl(s, mask, cosv, sinv, max_weight, min_abs, 4);
    }
}

/* ImaginaryPauliRotation (src/Propagation/specializations.jl:86-133): a term
 * that COMMUTES with the generator keeps c * cosh and branches into
 * (c * sinh) * sign at term ^ mask (_imaginaryrotationrule :120-131, the sign
 * of paulirotationproduct); the sum is merged without truncation (xormerge!);
 * with `normalize` every coefficient is multiplied by 1 / (the identity's
 * coefficient after the merge), mult!(cache, 1 / getcoeff(sum, 0)) (:94-98);
 * then every term is truncated (truncate!, :100). */
long ppo_imaginary_rotation(ppo_sum *s, const uint64_t *mask, const double *coshv, const double *sinhv,
                            int normalize, int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    size_t n = s->n, nc = 0;
    for (size_t i = 0; i < n; ++i) nc += !anticommutes_w(mask, s->keys + i * W, W);
    uint64_t *tk = xmalloc((nc ? nc : 1) * W * sizeof(uint64_t));
    double *tc = xmalloc((nc ? nc : 1) * B * sizeof(double));
    size_t j = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t *k = s->keys + i * W;
        if (anticommutes_w(mask, k, W)) continue;
        int e = sign_exponent_w(mask, k, W);
        double sign = (double)((e & 2) - 1);
        for (int l = 0; l < W; ++l) tk[j * W + l] = k[l] ^ mask[l];
        double *c = s->coeffs + i * B;
        for (int b = 0; b < B; ++b) {
            tc[j * B + b] = (c[b] * sinhv[b]) * sign;
            c[b] = c[b] * coshv[b];
        }
        ++j;
    }
    finish_branch(s, tk, tc, nc, mask, -1, 0.0);
    free(tk);
    free(tc);
    if (normalize) {
        uint64_t zero[4] = {0, 0, 0, 0};
        double *c0 = xmalloc(B * sizeof(double));
        ppo_get_coeff(s, zero, c0);
        for (int b = 0; b < B; ++b) c0[b] = 1 / c0[b];
        for (size_t i = 0; i < s->n; ++i)
            for (int b = 0; b < B; ++b) s->coeffs[i * B + b] = s->coeffs[i * B + b] * c0[b];
        free(c0);
    }
    size_t o = 0;
    for (size_t i = 0; i < s->n; ++i) {
        uint64_t *k = s->keys + i * W;
        double *c = s->coeffs + i * B;
        if (truncated(k, c, W, B, max_weight, min_abs)) continue;
        if (o != i) {
            memmove(s->keys + o * W, k, W * sizeof(uint64_t));
            memmove(s->coeffs + o * B, c, B * sizeof(double));
        }
        ++o;
    }
    s->n = o;
    return (long)nc;
}

void ppo_pauli_noise(ppo_sum *s, int qubit, unsigned damp_codes, const double *factor,
                     int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    size_t o = 0;
    for (size_t i = 0; i < s->n; ++i) {
        uint64_t *k = s->keys + i * W;
        double *c = s->coeffs + i * B;
        if ((damp_codes >> get_pauli(k, qubit)) & 1)
            for (int b = 0; b < B; ++b) c[b] = c[b] * factor[b];
        if (truncated(k, c, W, B, max_weight, min_abs)) continue;
        if (o != i) {
            memmove(s->keys + o * W, k, W * sizeof(uint64_t));
            memmove(s->coeffs + o * B, c, B * sizeof(double));
        }
        ++o;
    }
    s->n = o;
}

long ppo_amplitude_damping(ppo_sum *s, int qubit, double gamma, int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    uint64_t mask[4] = {0, 0, 0, 0};
    mask[qubit >> 5] = 3ULL << (2 * (qubit & 31));
    double damp_val = sqrt(1 - gamma);
    size_t nz = 0, touched = 0;
    for (size_t i = 0; i < s->n; ++i) {
        int p = get_pauli(s->keys + i * W, qubit);
        nz += (p == 3);
        touched += (p != 0);
    }
    if (touched == 0) return 0;
    uint64_t *tk = xmalloc((nz ? nz : 1) * W * sizeof(uint64_t));
    double *tc = xmalloc((nz ? nz : 1) * B * sizeof(double));
    size_t j = 0;
    for (size_t i = 0; i < s->n; ++i) {
        uint64_t *k = s->keys + i * W;
        double *c = s->coeffs + i * B;
        int p = get_pauli(k, qubit);
        if (p == 3) {
            for (int l = 0; l < W; ++l) tk[j * W + l] = k[l] ^ mask[l];
            for (int b = 0; b < B; ++b) {
                tc[j * B + b] = gamma * c[b];
                c[b] = (1 - gamma) * c[b];
            }
            ++j;
        } else if (p != 0) {
            for (int b = 0; b < B; ++b) c[b] = damp_val * c[b];
        }
    }
    finish_branch(s, tk, tc, nz, mask, max_weight, min_abs);
    free(tk);
    free(tc);
    return (long)touched;
}

void ppo_clifford(ppo_sum *s, int nq, const int *qubits, const uint64_t *changes,
                  const double *signs, int max_weight, double min_abs) {
    int W = s->W, B = s->B;
    size_t o = 0;
    for (size_t i = 0; i < s->n; ++i) {
        uint64_t *k = s->keys + i * W;
        double *c = s->coeffs + i * B;
        int idx = 0;
        for (int a = 0; a < nq; ++a) idx |= get_pauli(k, qubits[a]) << (2 * a);
        for (int l = 0; l < W; ++l) k[l] ^= changes[idx * W + l];
        for (int b = 0; b < B; ++b) c[b] = c[b] * signs[idx];
        if (max_weight >= 0 && truncated(k, c, W, B, max_weight, min_abs)) continue;
        if (o != i) {
            memmove(s->keys + o * W, k, W * sizeof(uint64_t));
            memmove(s->coeffs + o * B, c, B * sizeof(double));
        }
        ++o;
    }
    s->n = o;
    s->sorted = 0;
}

void ppo_overlap_zero(const ppo_sum *s, double *out) {
    int W = s->W, B = s->B;
    for (int b = 0; b < B; ++b) out[b] = 0.0;
    for (size_t i = 0; i < s->n; ++i)
        if (!contains_xy_w(s->keys + i * W, W))
            for (int b = 0; b < B; ++b) out[b] += s->coeffs[i * B + b];
}

/* Pauli code of qubit q (0-based): I 0, X 1, Y 2, Z 3 (src/PauliAlgebra/bitoperations.jl). */
static unsigned code_of(const uint64_t *key, int q) { return (unsigned)(key[q >> 5] >> (2 * (q & 31))) & 3u; }

/* The expectation of one Pauli string in a product state, qubit by qubit
 * (src/stateoverlap.jl:12-80): kind 0 |0..0> (X or Y anywhere -> 0), kind 1
 * |+..+> (Y or Z anywhere -> 0, src/stateoverlap.jl:44), kind 2 the basis
 * state with qubit q in |1> when bit 2q of `ones` is set (X or Y -> 0, else
 * (-1)^(number of Z on those qubits), _calcsignwithones :70-80); kind 3 the
 * maximally mixed state (1 for the identity, else 0; overlapwithmaxmixed :87-93). */
static int state_factor(const uint64_t *key, int W, int kind, const uint64_t *ones) {
    int f = 1;
    if (kind == 3) { /* the maximally mixed state: only the identity counts (src/stateoverlap.jl:87-93) */
        for (int l = 0; l < W; ++l)
            if (key[l]) return 0;
        return 1;
    }
    for (int q = 0; q < 32 * W; ++q) {
        unsigned c = code_of(key, q);
        if (kind == 1) {
            if (c == 2u || c == 3u) return 0;
        } else {
            if (c == 1u || c == 2u) return 0;
            if (kind == 2 && c == 3u && ((ones[q >> 5] >> (2 * (q & 31))) & 1u)) f = -f;
        }
    }
    return f;
}

void ppo_overlap_kind(const ppo_sum *s, int kind, const uint64_t *ones, double *out) {
    int W = s->W, B = s->B;
    for (int b = 0; b < B; ++b) out[b] = 0.0;
    for (size_t i = 0; i < s->n; ++i) {
        int f = state_factor(s->keys + i * W, W, kind, ones);
        if (f)
            for (int b = 0; b < B; ++b) out[b] += f * s->coeffs[i * B + b];
    }
}

/* getcoeff (src/Base/termsum.jl:104-124): every entry equal to `key` is
 * added (a linear scan also covers an unsorted sum). */
void ppo_get_coeff(const ppo_sum *s, const uint64_t *key, double *out) {
    int W = s->W, B = s->B;
    for (int b = 0; b < B; ++b) out[b] = 0.0;
    for (size_t i = 0; i < s->n; ++i)
        if (memcmp(s->keys + i * W, key, W * sizeof(uint64_t)) == 0)
            for (int b = 0; b < B; ++b) out[b] += s->coeffs[i * B + b];
}

/* scalarproduct (src/stateoverlap.jl:120-135): loop over b, look every term up in a. */
int ppo_scalar_product(const ppo_sum *a, const ppo_sum *b, double *out) {
    if (a->W != b->W || a->B != b->B) return -1;
    int W = a->W, B = a->B;
    double *c = xmalloc(B * sizeof(double));
    for (int k = 0; k < B; ++k) out[k] = 0.0;
    for (size_t i = 0; i < b->n; ++i) {
        ppo_get_coeff(a, b->keys + i * W, c);
        for (int k = 0; k < B; ++k) out[k] += c[k] * b->coeffs[i * B + k];
    }
    free(c);
    return 0;
}

/* maxabscoeff (src/Base/Primitives/maxabscoeff.jl:7): per column */
void ppo_max_abs(const ppo_sum *s, double *out) {
    int B = s->B;
    for (int b = 0; b < B; ++b) out[b] = 0.0;
    for (size_t i = 0; i < s->n; ++i)
        for (int b = 0; b < B; ++b) out[b] = fmax(out[b], fabs(s->coeffs[i * B + b]));
}

/* truncate! (src/Propagation/generics.jl:192-250): drop terms over max_weight
 * or under the coefficient threshold max(min_rel * maxabscoeff, min_abs) (a
 * negative min_rel: min_abs alone); B = 1 for a relative threshold. Returns
 * the threshold used. */
double ppo_truncate(ppo_sum *s, int max_weight, double min_abs, double min_rel) {
    int W = s->W, B = s->B;
    double thr = min_abs;
    if (min_rel >= 0) {
        double m;
        ppo_max_abs(s, &m);
        thr = fmax(min_rel * m, min_abs);
    }
    size_t o = 0;
    for (size_t i = 0; i < s->n; ++i) {
        uint64_t *k = s->keys + i * W;
        double *c = s->coeffs + i * B;
        if (truncated(k, c, W, B, max_weight, thr)) continue;
        if (o != i) {
            memmove(s->keys + o * W, k, W * sizeof(uint64_t));
            memmove(s->coeffs + o * B, c, B * sizeof(double));
        }
        ++o;
    }
    s->n = o;
    return thr;
}

/* ------------------------------------------------------------------ */
/* resampling (src/Base/MonteCarlo/resample.jl)                        */
/* ------------------------------------------------------------------ */

static double rs_weight(double c, int squared) { return squared ? c * c : fabs(c); }

/* The comb spacing that keeps `target` terms in expectation, Newton steps from
 * total / target (_calibrate_comb_spacing, resample.jl:118-130). */
static double rs_calibrate(const ppo_sum *s, double total, size_t target, int squared, double rtol, double atol) {
    double lowest = (1 - rtol) * (double)target - atol;
    double spacing = total / (double)target;
    for (int it = 0; it < 10; ++it) {
        size_t n_heavy = 0;
        double light = 0.0;
        for (size_t i = 0; i < s->n; ++i) {
            double w = rs_weight(s->coeffs[i], squared);
            if (w >= spacing) ++n_heavy;
            else light += w;
        }
        if ((double)n_heavy + light / spacing >= lowest || light == 0.0) return spacing;
        spacing = light / (double)(target - n_heavy);
    }
    return spacing;
}

/* Resampling of a B = 1 sum down to about `target` terms (resample.jl):
 * method 0 semideterministic_systematic_resample! (:139-170), 1
 * systematic_resample! (:96-112, calibrated when `calibrate`), 2
 * multinomial_resample! (:64-78) with `draws` = target uniforms in [0, 1),
 * ascending. `offset` is the comb's random offset in [0, 1). Every term lies
 * on [0, total) on an interval as wide as its weight; a term gets the teeth
 * (comb: floor(position / spacing - offset) counts; multinomial: draws
 * * total) in its interval, its new coefficient is copysign(teeth * spacing,
 * c) (squared: sqrt(teeth * spacing)), and terms left at zero are dropped
 * in order (_set_coeffs_and_drop_zeros!, :250-300). Returns the spacing used,
 * or -1 on bad arguments. */
double ppo_resample(ppo_sum *s, int method, size_t target, int squared, double offset, const double *draws,
                    int calibrate, double rtol, double atol) {
    if (s->B != 1 || target == 0) return -1.0;
    if (method == 0 && squared) return -1.0;
    if (method != 2 && target > s->n) return -1.0;
    int W = s->W;
    size_t n = s->n;
    double total = 0.0, thr = 0.0, spacing = 0.0;
    size_t n_kept = 0;
    if (method == 0) {
        double sabs = 0.0;
        for (size_t i = 0; i < n; ++i) sabs += fabs(s->coeffs[i]);
        thr = sabs / (double)target;
        for (size_t i = 0; i < n; ++i) {
            double c = s->coeffs[i];
            if (fabs(c) > thr) ++n_kept;
            else total += fabs(c);
        }
        spacing = target > n_kept ? total / (double)(target - n_kept) : 0.0;
    } else {
        for (size_t i = 0; i < n; ++i) total += rs_weight(s->coeffs[i], squared);
        if (method == 1)
            spacing = calibrate ? rs_calibrate(s, total, target, squared, rtol, atol) : total / (double)target;
        else
            spacing = total / (double)target;
    }
    double inv = spacing == 0.0 ? 0.0 : 1 / spacing;
    double *scaled = NULL;
    if (method == 2) {
        scaled = xmalloc((target ? target : 1) * sizeof(double));
        for (size_t k = 0; k < target; ++k) scaled[k] = draws[k] * total;
    }
    double pos = 0.0;
    size_t o = 0;
    double below_end = 0.0;
    if (method == 2) {
        size_t k = 0;
        while (k < target && scaled[k] < pos) ++k;
        below_end = (double)k;
    } else {
        below_end = floor(pos * inv - offset);
    }
    for (size_t i = 0; i < n; ++i) {
        double c = s->coeffs[i];
        int kept = method == 0 && fabs(c) > thr;
        double w = method == 0 ? (kept ? 0.0 : fabs(c)) : rs_weight(c, squared);
        double below_start = below_end;
        pos += w;
        if (method == 2) {
            /* number of draws below the position: searchsortedfirst(draws, pos) - 1 */
            size_t lo = 0, hi = target;
            while (lo < hi) {
                size_t mid = lo + (hi - lo) / 2;
                if (scaled[mid] < pos) lo = mid + 1;
                else hi = mid;
            }
            below_end = (double)lo;
        } else {
            below_end = floor(pos * inv - offset);
        }
        double teeth = below_end - below_start;
        double nc;
        if (kept) nc = c;
        else if (squared) nc = sqrt(teeth * spacing) * (c == 0.0 ? 0.0 : 1.0);
        else nc = copysign(teeth * spacing, c);
        if (nc == 0.0) continue;
        if (o != i) memmove(s->keys + o * W, s->keys + i * W, W * sizeof(uint64_t));
        s->coeffs[o] = nc;
        ++o;
    }
    s->n = o;
    free(scaled);
    return spacing;
}

/* ------------------------------------------------------------------ */
/* gate lists                                                          */
/* ------------------------------------------------------------------ */

static double gate_value(const ppo_gate *g, const double *theta, int B, int b) {
    return g->param < 0 ? g->value : theta[(size_t)g->param * B + b];
}

static int apply_gate(ppo_sum *s, const ppo_gate *g, const double *theta, const uint64_t *lut_changes,
                      const double *lut_signs, int max_weight, double min_abs, double *cv, double *sv) {
    int B = s->B, W = s->W;
    switch (g->kind) {
    case PPO_ROTATION:
        for (int b = 0; b < B; ++b) {
            double t = gate_value(g, theta, B, b);
            cv[b] = cos(t);
            sv[b] = sin(t);
        }
        ppo_rotation(s, g->mask, cv, sv, max_weight, min_abs);
        return 0;
    case PPO_IMAG_ROTATION:
        for (int b = 0; b < B; ++b) {
            double t = gate_value(g, theta, B, b);
            cv[b] = cosh(t);
            sv[b] = sinh(t);
        }
        ppo_imaginary_rotation(s, g->mask, cv, sv, g->flags & 1, max_weight, min_abs);
        return 0;
    case PPO_PAULI_NOISE:
        for (int b = 0; b < B; ++b) cv[b] = 1 - gate_value(g, theta, B, b);
        ppo_pauli_noise(s, g->qubit, g->damp_codes, cv, max_weight, min_abs);
        return 0;
    case PPO_AMP_DAMPING:
        if (g->param >= 0 && B > 1) return -1;
        ppo_amplitude_damping(s, g->qubit, gate_value(g, theta, B, 0), max_weight, min_abs);
        return 0;
    case PPO_CLIFFORD: {
        int nent = g->nq == 1 ? 4 : 16;
        size_t base = 0;
        /* tables are packed back to back, 16 entries per table slot */
        base = (size_t)g->lut * 16;
        ppo_clifford(s, g->nq, g->qubits, lut_changes + base * W, lut_signs + base, max_weight, min_abs);
        (void)nent;
        return 0;
    }
    default:
        return -1;
    }
}

int ppo_run(ppo_sum *s, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
            const uint64_t *lut_changes, const double *lut_signs, int max_weight, double min_abs,
            int64_t *counts) {
    (void)nparams;
    double *cv = xmalloc(s->B * sizeof(double));
    double *sv = xmalloc(s->B * sizeof(double));
    int rc = 0;
    for (size_t g = 0; g < ngates && rc == 0; ++g) {
        if (counts) counts[g] = (int64_t)s->n;
        rc = apply_gate(s, gates + g, theta, lut_changes, lut_signs, max_weight, min_abs, cv, sv);
    }
    free(cv);
    free(sv);
    return rc;
}

/* ------------------------------------------------------------------ */
/* reverse-mode gradient (every intermediate sum kept)                 */
/* ------------------------------------------------------------------ */

static long find_key(const uint64_t *keys, size_t n, const uint64_t *k, int W) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        int c = key_cmp(keys + mid * W, k, W);
        if (c == 0) return (long)mid;
        if (c < 0) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

static ppo_sum *clone_sum(const ppo_sum *a) {
    ppo_sum *s = ppo_create(a->W, a->B);
    reserve(s, a->n ? a->n : 1);
    memcpy(s->keys, a->keys, a->n * a->W * sizeof(uint64_t));
    memcpy(s->coeffs, a->coeffs, a->n * a->B * sizeof(double));
    s->n = a->n;
    s->sorted = a->sorted;
    return s;
}

static void merge_only(ppo_sum *s);

int ppo_adjoint(const ppo_sum *s0, const ppo_gate *gates, size_t ngates, const double *theta,
                size_t nparams, const uint64_t *lut_changes, const double *lut_signs, int max_weight, double min_abs,
                int kind, const uint64_t *ones, double *loss, double *grad) {
    int W = s0->W, B = s0->B;
    if (!s0->sorted || kind < 0 || kind > 3) return -1;
    for (size_t g = 0; g < ngates; ++g)
        if (gates[g].kind != PPO_ROTATION && gates[g].kind != PPO_PAULI_NOISE && gates[g].kind != PPO_CLIFFORD)
            return -1;
    const uint64_t zero_ones[4] = {0, 0, 0, 0};
    if (!ones) ones = zero_ones;
    double *cv = xmalloc(B * sizeof(double));
    double *sv = xmalloc(B * sizeof(double));
    /* The sum before every S-th gate is kept (S ~ sqrt(ngates)); a segment's
     * intermediate sums are computed again from its checkpoint when the sweep
     * reaches it, so memory holds about 2 sqrt(ngates) sums. A Clifford leaves
     * its output unsorted; it is sorted here for the lookups (the next
     * branching gate sorts it anyway, with the same sums). */
    size_t S = 1;
    while (S * S < ngates) ++S;
    const size_t nseg = (ngates + S - 1) / S;
    ppo_sum **ck = xmalloc((nseg ? nseg : 1) * sizeof(ppo_sum *));
    ppo_sum *fin = clone_sum(s0);
    for (size_t g = 0; g < ngates; ++g) {
        if (g % S == 0) ck[g / S] = clone_sum(fin);
        apply_gate(fin, gates + g, theta, lut_changes, lut_signs, max_weight, min_abs, cv, sv);
        merge_only(fin);
    }
    for (int b = 0; b < B; ++b) loss[b] = 0.0;
    for (size_t i = 0; i < fin->n; ++i) {
        const int f = state_factor(fin->keys + i * W, W, kind, ones);
        if (f)
            for (int b = 0; b < B; ++b) loss[b] += f * fin->coeffs[i * B + b];
    }
    memset(grad, 0, nparams * B * sizeof(double));
    /* adjoint of the final sum: d overlap / d c_t, the readout's weight of t */
    double *adj = xmalloc((fin->n ? fin->n : 1) * B * sizeof(double));
    for (size_t i = 0; i < fin->n; ++i)
        for (int b = 0; b < B; ++b) adj[i * B + b] = (double)state_factor(fin->keys + i * W, W, kind, ones);
    uint64_t pk[4];
    ppo_sum **st = xmalloc((S + 1) * sizeof(ppo_sum *));
    for (size_t seg = nseg; seg-- > 0;) {
      const size_t g0 = seg * S, g1 = g0 + S < ngates ? g0 + S : ngates;
      st[0] = clone_sum(ck[seg]);
      for (size_t g = g0; g < g1; ++g) {
          st[g - g0 + 1] = clone_sum(st[g - g0]);
          apply_gate(st[g - g0 + 1], gates + g, theta, lut_changes, lut_signs, max_weight, min_abs, cv, sv);
          merge_only(st[g - g0 + 1]);
      }
      for (size_t gg = g1; gg-- > g0;) {
        const ppo_gate *g = gates + gg;
        ppo_sum *in = st[gg - g0], *out = st[gg - g0 + 1];
        double *nadj = xmalloc((in->n ? in->n : 1) * B * sizeof(double));
        if (g->kind == PPO_ROTATION) {
            size_t na = 0;
            for (size_t i = 0; i < in->n; ++i) na += anticommutes_w(g->mask, in->keys + i * W, W);
            for (int b = 0; b < B; ++b) {
                double t = gate_value(g, theta, B, b);
                cv[b] = cos(t);
                sv[b] = sin(t);
            }
            for (size_t i = 0; i < in->n; ++i) {
                const uint64_t *k = in->keys + i * W;
                const double *c = in->coeffs + i * B;
                if (na == 0) { /* untouched gate: identity, no truncation */
                    for (int b = 0; b < B; ++b) nadj[i * B + b] = adj[i * B + b];
                    continue;
                }
                long ps = find_key(out->keys, out->n, k, W);
                if (!anticommutes_w(g->mask, k, W)) {
                    for (int b = 0; b < B; ++b) nadj[i * B + b] = ps >= 0 ? adj[ps * B + b] : 0.0;
                    continue;
                }
                for (int l = 0; l < W; ++l) pk[l] = k[l] ^ g->mask[l];
                long pp = find_key(out->keys, out->n, pk, W);
                double sign = (double)((sign_exponent_w(g->mask, k, W) & 2) - 1);
                for (int b = 0; b < B; ++b) {
                    double as = ps >= 0 ? adj[ps * B + b] : 0.0;
                    double ap = pp >= 0 ? adj[pp * B + b] : 0.0;
                    nadj[i * B + b] = as * cv[b] + ap * sv[b] * sign;
                    if (g->param >= 0)
                        grad[(size_t)g->param * B + b] += as * (-c[b] * sv[b]) + ap * (c[b] * cv[b] * sign);
                }
            }
        } else if (g->kind == PPO_CLIFFORD) { /* a signed permutation */
            const size_t base = (size_t)g->lut * 16;
            const uint64_t *ch = lut_changes + base * W;
            const double *sg = lut_signs + base;
            for (size_t i = 0; i < in->n; ++i) {
                const uint64_t *k = in->keys + i * W;
                int idx = 0;
                for (int a = 0; a < g->nq; ++a) idx |= get_pauli(k, g->qubits[a]) << (2 * a);
                for (int l = 0; l < W; ++l) pk[l] = k[l] ^ ch[idx * W + l];
                long ps = find_key(out->keys, out->n, pk, W);
                for (int b = 0; b < B; ++b) nadj[i * B + b] = ps >= 0 ? adj[ps * B + b] * sg[idx] : 0.0;
            }
        } else { /* PauliNoise: diagonal */
            for (size_t i = 0; i < in->n; ++i) {
                const uint64_t *k = in->keys + i * W;
                long ps = find_key(out->keys, out->n, k, W);
                int damped = (g->damp_codes >> get_pauli(k, g->qubit)) & 1;
                for (int b = 0; b < B; ++b) {
                    double f = damped ? 1 - gate_value(g, theta, B, b) : 1.0;
                    nadj[i * B + b] = ps >= 0 ? adj[ps * B + b] * f : 0.0;
                    if (g->param >= 0 && damped && ps >= 0)
                        grad[(size_t)g->param * B + b] += adj[ps * B + b] * (-in->coeffs[i * B + b]);
                }
            }
        }
        free(adj);
        adj = nadj;
      }
      for (size_t j = 0; j <= g1 - g0; ++j) ppo_destroy(st[j]);
    }
    free(adj);
    for (size_t k = 0; k < nseg; ++k) ppo_destroy(ck[k]);
    ppo_destroy(fin);
    free(ck);
    free(st);
    free(cv);
    free(sv);
    return 0;
}

/* ------------------------------------------------------------------ */
/* rewindgradient: the reference gradient (src/gradient.jl)            */
/* ------------------------------------------------------------------ */

/* merge! on a sum the last step may have left unsorted: sort, sum equal
 * strings, truncate nothing (src/Base/Merge/mergeandtruncate.jl:13-18). */
static void merge_only(ppo_sum *s) {
    if (!s->sorted) sort_reduce_truncate(s, NULL, NULL, 0, -1, 0.0);
}

/* _generatorcommutatordot (src/gradient.jl:137-153, 190-241): the sum over
 * operator terms t that anticommute with the generator P of
 * -0.5 * imag(comm) * c_t * dual(P xor t), where commutator(P, t) = 2 i^e
 * (P xor t) (src/PauliAlgebra/paulioperations.jl:414-424) with e odd, so that
 * -0.5 * imag(comm) = (e & 2) - 1. Both sums sorted; terms in operator order. */
static double commutator_dot(const ppo_sum *op, const ppo_sum *dual, const uint64_t *mask) {
    const int W = op->W;
    double total = 0.0;
    uint64_t pk[4];
    for (size_t i = 0; i < op->n; ++i) {
        const uint64_t *k = op->keys + i * W;
        if (!anticommutes_w(mask, k, W)) continue;
        for (int l = 0; l < W; ++l) pk[l] = k[l] ^ mask[l];
        long j = find_key(dual->keys, dual->n, pk, W);
        if (j < 0) continue;
        double sgn = (double)((sign_exponent_w(mask, k, W) & 2) - 1);
        total += (sgn * op->coeffs[i]) * dual->coeffs[j];
    }
    return total;
}

/* _intersectfilter! (src/gradient.jl:278-350): keep the dual terms whose
 * strings the (sorted) operator sum still holds, in order. */
static void intersect_filter(ppo_sum *dual, const ppo_sum *op) {
    const int W = dual->W, B = dual->B;
    size_t j = 0, o = 0;
    for (size_t i = 0; i < dual->n; ++i) {
        const uint64_t *dk = dual->keys + i * W;
        while (j < op->n && key_cmp(op->keys + j * W, dk, W) < 0) ++j;
        if (j < op->n && key_eq(op->keys + j * W, dk, W)) {
            if (o != i) {
                memmove(dual->keys + o * W, dk, W * sizeof(uint64_t));
                memmove(dual->coeffs + o * B, dual->coeffs + i * B, B * sizeof(double));
            }
            ++o;
        }
    }
    dual->n = o;
}

/* transposecliffordmap (src/Gates/cliffordgates.jl:94-98) on a packed table:
 * entry idx sends pattern idx to idx' = idx xor (the change on the gate's
 * qubits) with sign s; the transposed table sends idx' back with the same
 * change and sign. */
static void transpose_lut(int nq, const int *qubits, const uint64_t *change, const double *sign, int W,
                          uint64_t *tchange, double *tsign) {
    memset(tchange, 0, 16 * W * sizeof(uint64_t));
    for (int idx = 0; idx < 16; ++idx) tsign[idx] = 1.0;
    const int nent = 1 << (2 * nq);
    for (int idx = 0; idx < nent; ++idx) {
        int moved = 0;
        for (int a = 0; a < nq; ++a) moved |= get_pauli(change + idx * W, qubits[a]) << (2 * a);
        const int to = idx ^ moved;
        memcpy(tchange + to * W, change + idx * W, W * sizeof(uint64_t));
        tsign[to] = sign[idx];
    }
}

/* rewindgradient!(circuit, cache, params, overlapfunc; kwargs...)
 * (src/gradient.jl:40-108) on a B = 1 sum, for a gate list in Heisenberg
 * order (rotations, frozen rotations and Cliffords; anything else returns
 * -1, as the reference rejects noise, :42-43):
 *   forward  the list as propagate runs it, truncating (max_weight, min_abs);
 *   dual     the final operator's strings, each with overlapfunc of that
 *            string alone (_dualcache, :116-124): kinds 0..3 of state_factor,
 *            or kind 4, scale * (the string's coefficient in rho)
 *            (overlapwithpaulisum(rho, .) = 2^n scalarproduct, scale = 2^n);
 *   undo     the list backwards (the circuit's own order), each gate in its
 *            Schrodinger form (rotation angle negated, Clifford table
 *            transposed; src/Gates/heisenbergschrodinger.jl:134-161): a
 *            parametrized rotation first records grad[param] =
 *            commutator_dot(operator, dual); the operator takes the gate
 *            with truncation, the dual without, then the dual keeps only
 *            strings the operator holds (_undostep!, :84-108).
 * grad has one entry per Heisenberg parameter row; *loss is overlapfunc of
 * the forward result, in term order. */
int ppo_rewind(const ppo_sum *s0, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
               const uint64_t *lut_changes, const double *lut_signs, int max_weight, double min_abs, int kind,
               const uint64_t *ones, const ppo_sum *rho, double scale, double *loss, double *grad) {
    const int W = s0->W;
    if (s0->B != 1 || (kind == 4 && (!rho || rho->W != W || rho->B != 1)) || kind < 0 || kind > 4) return -1;
    for (size_t g = 0; g < ngates; ++g) {
        if (gates[g].kind != PPO_ROTATION && gates[g].kind != PPO_CLIFFORD) return -1;
        if (gates[g].param >= 0 && (size_t)gates[g].param >= nparams) return -1;
    }
    const uint64_t zero_ones[4] = {0, 0, 0, 0};
    if (!ones) ones = zero_ones;
    double cv, sv;
    ppo_sum *op = clone_sum(s0);
    for (size_t g = 0; g < ngates; ++g)
        apply_gate(op, gates + g, theta, lut_changes, lut_signs, max_weight, min_abs, &cv, &sv);
    /* the dual sum: the operator's strings with their own overlaps */
    ppo_sum *dual = clone_sum(op);
    double ex = 0.0;
    for (size_t i = 0; i < dual->n; ++i) {
        const uint64_t *k = dual->keys + i * W;
        double d;
        if (kind == 4) {
            ppo_get_coeff(rho, k, &d);
            d = scale * d;
        } else {
            d = (double)state_factor(k, W, kind, ones);
        }
        dual->coeffs[i] = d;
        ex += op->coeffs[i] * d;
    }
    *loss = ex;
    memset(grad, 0, nparams * sizeof(double));
    uint64_t tch[16 * 4];
    double tsg[16];
    for (size_t gg = ngates; gg-- > 0;) {
        const ppo_gate *g = gates + gg;
        merge_only(op);
        merge_only(dual);
        if (g->kind == PPO_ROTATION) {
            const double t = -gate_value(g, theta, 1, 0);
            const double c = cos(t), sn = sin(t);
            if (g->param >= 0) grad[g->param] = commutator_dot(op, dual, g->mask);
            ppo_rotation(op, g->mask, &c, &sn, max_weight, min_abs);
            ppo_rotation(dual, g->mask, &c, &sn, -1, 0.0);
        } else {
            const size_t base = (size_t)g->lut * 16;
            transpose_lut(g->nq, g->qubits, lut_changes + base * W, lut_signs + base, W, tch, tsg);
            ppo_clifford(op, g->nq, g->qubits, tch, tsg, max_weight, min_abs);
            ppo_clifford(dual, g->nq, g->qubits, tch, tsg, -1, 0.0);
            merge_only(op);
            merge_only(dual);
        }
        intersect_filter(dual, op);
    }
    ppo_destroy(op);
    ppo_destroy(dual);
    return 0;
}

/* ------------------------------------------------------------------ */
/* zones: the MultiPauliSum restatement (the all-cores CPU baseline)   */
/* ------------------------------------------------------------------ */

/* A term's zone: bit r is the parity of term & mask_r, with pseudo-random masks
 * from SplitMix64's mixer, seed 0x9e3779b97f4a7c15, one draw per 64-bit word
 * (src/Base/MultiSum/multisumstorage.jl:64-151). The map is linear over GF(2),
 * so a rotation moves the created terms of zone z to zone z ^ zone(mask). */
#define PPO_ZONE_SEED 0x9e3779b97f4a7c15ULL
#define PPO_MAX_ZBITS 8

struct ppo_multi {
    int W, B, zbits, Z;
    uint64_t masks[PPO_MAX_ZBITS][4];
    ppo_sum **zones;
};

static uint64_t mix64(uint64_t x) {
    uint64_t z = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static int zone_of(const ppo_multi *m, const uint64_t *key) {
    int z = 0;
    for (int r = 0; r < m->zbits; ++r) {
        uint64_t f = 0;
        for (int l = 0; l < m->W; ++l) f ^= key[l] & m->masks[r][l];
        z |= (__builtin_popcountll(f) & 1) << r;
    }
    return z;
}

ppo_multi *ppo_multi_create(int W, int B, int zbits) {
    if (zbits < 0 || zbits > PPO_MAX_ZBITS) return NULL;
    ppo_multi *m = xmalloc(sizeof(ppo_multi));
    m->W = W;
    m->B = B;
    m->zbits = zbits;
    m->Z = 1 << zbits;
    uint64_t state = PPO_ZONE_SEED;
    memset(m->masks, 0, sizeof(m->masks));
    for (int r = 0; r < zbits; ++r)
        for (int l = 0; l < W; ++l) {
            state += PPO_ZONE_SEED;
            m->masks[r][l] = mix64(state);
        }
    m->zones = xmalloc(m->Z * sizeof(ppo_sum *));
    for (int z = 0; z < m->Z; ++z) m->zones[z] = ppo_create(W, B);
    return m;
}

void ppo_multi_destroy(ppo_multi *m) {
    if (!m) return;
    for (int z = 0; z < m->Z; ++z) ppo_destroy(m->zones[z]);
    free(m->zones);
    free(m);
}

/* Distinct terms in any order: sorted once, then dealt to their zones in
 * order, so every zone is sorted. */
int ppo_multi_set(ppo_multi *m, const uint64_t *keys, const double *coeffs, size_t n) {
    ppo_sum *all = ppo_create(m->W, m->B);
    int rc = ppo_set(all, keys, coeffs, n);
    if (rc) {
        ppo_destroy(all);
        return rc;
    }
    int W = m->W, B = m->B;
    for (int z = 0; z < m->Z; ++z) m->zones[z]->n = 0;
    for (size_t i = 0; i < all->n; ++i) {
        ppo_sum *zs = m->zones[zone_of(m, all->keys + i * W)];
        reserve(zs, zs->n + 1);
        memcpy(zs->keys + zs->n * W, all->keys + i * W, W * sizeof(uint64_t));
        memcpy(zs->coeffs + zs->n * B, all->coeffs + i * B, B * sizeof(double));
        zs->n += 1;
        zs->sorted = 1;
    }
    ppo_destroy(all);
    return 0;
}

size_t ppo_multi_size(const ppo_multi *m) {
    size_t n = 0;
    for (int z = 0; z < m->Z; ++z) n += m->zones[z]->n;
    return n;
}

/* All terms, zone after zone (each zone ascending). */
void ppo_multi_get(const ppo_multi *m, uint64_t *keys, double *coeffs) {
    size_t o = 0;
    for (int z = 0; z < m->Z; ++z) {
        const ppo_sum *zs = m->zones[z];
        memcpy(keys + o * m->W, zs->keys, zs->n * m->W * sizeof(uint64_t));
        memcpy(coeffs + o * m->B, zs->coeffs, zs->n * m->B * sizeof(double));
        o += zs->n;
    }
}

/* Run rotations and Pauli noise over the zones on nthreads threads: a rotation
 * branches every zone in parallel, then every zone merges the tail of its
 * partner zone (zone ^ zone(mask)), also in parallel. Other gate kinds are
 * not part of this baseline (returns -1). counts receives n_in per gate. */
int ppo_multi_run(ppo_multi *m, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
                  int max_weight, double min_abs, int nthreads, int64_t *counts) {
    const int W = m->W, B = m->B, Z = m->Z;
    uint64_t **tk = xmalloc(Z * sizeof(uint64_t *));
    double **tc = xmalloc(Z * sizeof(double *));
    size_t *nt = xmalloc(Z * sizeof(size_t));
    double *cv = xmalloc(B * sizeof(double)), *sv = xmalloc(B * sizeof(double)), *fv = xmalloc(B * sizeof(double));
    int rc = 0;
    for (size_t g = 0; g < ngates && !rc; ++g) {
        const ppo_gate *gt = &gates[g];
        if (gt->param >= 0 && (size_t)gt->param >= nparams) {
            rc = -1;
            break;
        }
        if (counts) counts[g] = (int64_t)ppo_multi_size(m);
        if (gt->kind == PPO_ROTATION) {
            for (int b = 0; b < B; ++b) {
                const double th = gate_value(gt, theta, B, b);
                cv[b] = cos(th);
                sv[b] = sin(th);
            }
            const int zm = zone_of(m, gt->mask);
            size_t total = 0;
            /* every zone branches into its own tail buffer; a zone's tail is
             * sorted in its own XOR buffer and merged by its partner zone
             * z ^ zone(mask), which writes its own merge buffer */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : total)
            for (int z = 0; z < Z; ++z) {
                ppo_sum *zs = m->zones[z];
                switch (W) {
                case 1: nt[z] = branch_into_tail(zs, gt->mask, cv, sv, 1); break;
                case 2: nt[z] = branch_into_tail(zs, gt->mask, cv, sv, 2); break;
                default: nt[z] = branch_into_tail(zs, gt->mask, cv, sv, 4); break;
                }
                if (nt[z]) {
                    ensure_buf(&zs->xkeys, &zs->xcoeffs, &zs->xcap, nt[z], W, B);
                    xor_sort_buf(zs->tkeys, zs->tcoeffs, zs->xkeys, zs->xcoeffs, nt[z], W, B, gt->mask, &tk[z], &tc[z]);
                }
                total += nt[z];
            }
            if (total > 0) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
                for (int z = 0; z < Z; ++z) {
                    const int src = z ^ zm;
                    merge_head_tail(m->zones[z], tk[src], tc[src], nt[src], max_weight, min_abs);
                }
            }
        } else if (gt->kind == PPO_PAULI_NOISE) {
            for (int b = 0; b < B; ++b) fv[b] = 1.0 - gate_value(gt, theta, B, b);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
            for (int z = 0; z < Z; ++z) ppo_pauli_noise(m->zones[z], gt->qubit, gt->damp_codes, fv, max_weight, min_abs);
        } else {
            rc = -1;
        }
    }
    free(tk);
    free(tc);
    free(nt);
    free(cv);
    free(sv);
    free(fv);
    return rc;
}

