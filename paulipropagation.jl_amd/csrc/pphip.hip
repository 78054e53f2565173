// pphip.hip -- host side of libpphip.so: handle, buffers, launches, C ABI.
//
// Device layout of a handle (capacity `cap` terms, W key limbs, B columns):
//   keys[2]    W limb planes of cap uint64 each (limb l of term i at l*cap + i)
//   coeffs[2]  cap x B doubles, row-major
//   tail[2]    cap uint32 entries: index of an anticommuting term, sign in bit 31
//   lb_states  one look-back word per tile of the largest compaction launch
// Buffer `cur` holds the live sum; each merging gate writes the other buffer
// and flips `cur`.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pphip.h"
#include "pp_kernels.h"

using namespace pp;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(PP_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr size_t MAX_TERMS = 0x7fffffffu;  // indices carry a sign bit

}  // namespace

struct pp_handle {
    int nqubits = 0, W = 1, B = 1;
    size_t cap = 0, n = 0;
    int cur = 0;
    hipStream_t stream = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    double* coeffs[2] = {nullptr, nullptr};
    uint32_t* tail[2] = {nullptr, nullptr};
    uint64_t* lb_states = nullptr;
    size_t lb_cap = 0;
    uint32_t* counter = nullptr;   // look-back tile counter
    uint32_t* d_scalar = nullptr;  // [0] touched count, [1] output count, [2] error bits
    uint32_t* h_scalar = nullptr;  // pinned mirror
    double* d_param = nullptr;     // 3 x B: cos, sin / factors of the current gate
    double* h_param = nullptr;     // pinned staging for d_param
    double* d_partial = nullptr;   // overlap partials
    size_t partial_cap = 0;
    double* d_out = nullptr;       // B results
    uint32_t epoch = 0;
    // profile
    bool prof_on = false;
    hipEvent_t ev[8] = {};
    int64_t gates = 0, term_apps = 0;
    int64_t launches[PP_KERNEL_KINDS] = {};
    double ms[PP_KERNEL_KINDS] = {};
    double alg_bytes[PP_KERNEL_KINDS] = {};
};

namespace {

int free_buffers(pp_handle* h) {
    for (int b = 0; b < 2; ++b) {
        if (h->keys[b]) hipFree(h->keys[b]);
        if (h->coeffs[b]) hipFree(h->coeffs[b]);
        if (h->tail[b]) hipFree(h->tail[b]);
        h->keys[b] = nullptr;
        h->coeffs[b] = nullptr;
        h->tail[b] = nullptr;
    }
    if (h->lb_states) hipFree(h->lb_states);
    h->lb_states = nullptr;
    return PP_OK;
}

int tile_rows(int B) {
    if (B <= 16) return 256;
    if (B <= 32) return 128;
    if (B <= 64) return 64;
    if (B <= 128) return 32;
    return 16;
}

// look-back words for the largest compaction launch over up to 2*cap entries:
// the smallest tile is the batched merge's or the batched noise's 64 rows
size_t tiles_needed(size_t cap, int B) {
    const size_t tile = B > 1 ? (size_t)std::min(64, tile_rows(B)) : 1024;
    return (2 * cap) / tile + 64;
}

// (Re)allocate every term buffer for `newcap` terms, keeping the live sum.
int grow(pp_handle* h, size_t newcap) {
    if (newcap <= h->cap) return PP_OK;
    if (newcap > MAX_TERMS + 1) newcap = MAX_TERMS + 1;
    const int W = h->W, B = h->B;
    uint64_t* nk[2] = {nullptr, nullptr};
    double* nc[2] = {nullptr, nullptr};
    uint32_t* nt[2] = {nullptr, nullptr};
    uint64_t* ns = nullptr;
    const size_t nlb = tiles_needed(newcap, B);
    bool ok = true;
    for (int b = 0; b < 2 && ok; ++b) {
        ok = ok && hipMalloc(&nk[b], newcap * W * sizeof(uint64_t)) == hipSuccess;
        ok = ok && hipMalloc(&nc[b], newcap * B * sizeof(double)) == hipSuccess;
        ok = ok && hipMalloc(&nt[b], newcap * sizeof(uint32_t)) == hipSuccess;
    }
    ok = ok && hipMalloc(&ns, nlb * sizeof(uint64_t)) == hipSuccess;
    if (!ok) {
        for (int b = 0; b < 2; ++b) {
            if (nk[b]) hipFree(nk[b]);
            if (nc[b]) hipFree(nc[b]);
            if (nt[b]) hipFree(nt[b]);
        }
        if (ns) hipFree(ns);
        (void)hipGetLastError();
        return fail(PP_ERR_NOMEM, "device allocation of " + std::to_string(newcap) + " terms failed");
    }
    HIP_TRY(hipMemsetAsync(ns, 0, nlb * sizeof(uint64_t), h->stream));
    if (h->n > 0) {
        for (int l = 0; l < W; ++l)
            HIP_TRY(hipMemcpyAsync(nk[0] + l * newcap, h->keys[h->cur] + l * h->cap, h->n * sizeof(uint64_t),
                                   hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(nc[0], h->coeffs[h->cur], h->n * B * sizeof(double), hipMemcpyDeviceToDevice,
                               h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    free_buffers(h);
    for (int b = 0; b < 2; ++b) {
        h->keys[b] = nk[b];
        h->coeffs[b] = nc[b];
        h->tail[b] = nt[b];
    }
    h->lb_states = ns;
    h->lb_cap = nlb;
    h->cap = newcap;
    h->cur = 0;
    return PP_OK;
}

uint32_t next_epoch(pp_handle* h) {
    h->epoch = (h->epoch + 1) & LB_EPOCH_MASK;
    if (h->epoch == 0) {
        // wrapped: words of old launches could match again, so clear them
        hipMemsetAsync(h->lb_states, 0, h->lb_cap * sizeof(uint64_t), h->stream);
        h->epoch = 1;
    }
    return h->epoch;
}

// copy `count` device scalars from index idx and wait for the stream
int read_scalars(pp_handle* h, int idx, int count, uint32_t* out) {
    HIP_TRY(hipMemcpyAsync(h->h_scalar + idx, h->d_scalar + idx, count * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int i = 0; i < count; ++i) out[i] = h->h_scalar[idx + i];
    return PP_OK;
}

inline unsigned blocks(size_t n, size_t per) { return (unsigned)((n + per - 1) / per); }

template <int W>
Key<W> to_key(const uint64_t* mask) {
    Key<W> k;
    for (int l = 0; l < W; ++l) k.l[l] = mask[l];
    return k;
}

// Levels of the tail permutation for mask m: every maximal run [lo, hi] of set
// bits contributes +(s+e) at level hi+1 and -(s+e) at level lo; a level at or
// above 64W spans the whole tail, so s+e = na there.
int mask_levels(const uint64_t* m, int W, uint32_t na, Levels* lv, int64_t* const_off, int* from_limb) {
    lv->n = 0;
    *const_off = 0;
    int minbit = 64 * W;
    int b = 64 * W - 1;
    while (b >= 0) {
        if ((m[b >> 6] >> (b & 63)) & 1ull) {
            const int hi = b;
            while (b >= 0 && ((m[b >> 6] >> (b & 63)) & 1ull)) --b;
            const int lo = b + 1;
            if (hi + 1 >= 64 * W) {
                *const_off += (int64_t)na;
            } else {
                if (lv->n >= MAX_LEVELS) return -1;
                lv->bit[lv->n] = hi + 1;
                lv->sgn[lv->n] = +1;
                ++lv->n;
            }
            if (lv->n >= MAX_LEVELS) return -1;
            lv->bit[lv->n] = lo;
            lv->sgn[lv->n] = -1;
            ++lv->n;
            minbit = std::min(minbit, lo);
        } else {
            --b;
        }
    }
    *from_limb = minbit >> 6;
    return 0;
}

int check_mask(const pp_handle* h, const uint64_t* mask) {
    bool any = false;
    for (int l = 0; l < h->W; ++l) {
        uint64_t m = mask[l];
        any = any || m != 0;
        const int lo = l * 64, hi = lo + 64;
        const int qmax = 2 * h->nqubits;
        if (qmax < hi) {
            const int keep = qmax - lo;
            const uint64_t allowed = keep <= 0 ? 0ull : (keep >= 64 ? ~0ull : ((1ull << keep) - 1));
            if (m & ~allowed) return fail(PP_ERR_ARG, "rotation mask acts on a qubit beyond nqubits");
        }
    }
    for (int l = h->W; l < 4; ++l)
        if (mask[l]) return fail(PP_ERR_ARG, "rotation mask acts on a qubit beyond the key width");
    if (!any) return fail(PP_ERR_ARG, "rotation mask is the identity");
    return PP_OK;
}

// ---------------------------------------------------------------- profile

void prof_mark(pp_handle* h, int ev) {
    if (h->prof_on) hipEventRecord(h->ev[ev], h->stream);
}

// after a stream synchronization: the events of slot (2*slot, 2*slot+1) are complete
void prof_add(pp_handle* h, int slot, int kind, double bytes) {
    h->launches[kind] += 1;
    h->alg_bytes[kind] += bytes;
    if (h->prof_on) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->ev[2 * slot], h->ev[2 * slot + 1]) == hipSuccess) h->ms[kind] += ms;
    }
}

template <int W>
int rotation_impl(pp_handle* h, const uint64_t* mask_in, const double* cos_b, const double* sin_b, int max_weight,
                  double min_abs, pp_gate_stats* st) {
    const size_t n = h->n;
    if (st) *st = pp_gate_stats{(int64_t)n, 0, (int64_t)n};
    h->gates += 1;
    h->term_apps += (int64_t)n;
    if (n == 0) return PP_OK;
    const Key<W> mask = to_key<W>(mask_in);
    int mask_limbs = 0;
    for (int l = 0; l < W; ++l) mask_limbs += mask_in[l] != 0;
    const double row = 8.0 * W + 8.0 * h->B;
    // K1: anticommuting terms, in order
    prof_mark(h, 0);
    hipLaunchKernelGGL(k_branch<W>, dim3(blocks(n, K1_TILE)), dim3(WG), 0, h->stream, h->keys[h->cur], h->cap,
                       (uint32_t)n, mask, h->tail[0], h->lb_states, h->counter, next_epoch(h), h->d_scalar);
    HIP_TRY(hipGetLastError());
    prof_mark(h, 1);
    uint32_t na = 0;
    if (int rc = read_scalars(h, 0, 1, &na)) return rc;
    prof_add(h, 0, PP_KERNEL_BRANCH, 8.0 * mask_limbs * n + 4.0 * na);
    if (st) st->n_touched = na;
    if (na == 0) return PP_OK;  // nothing touched: nothing merged or truncated
    if (n + na > MAX_TERMS) return fail(PP_ERR_CAPACITY, "term count exceeds 2^31 - 1");
    if (n + na > h->cap) {
        // the new buffers do not carry the tail over: branch again into them
        if (int rc = grow(h, std::max(n + na, h->cap + h->cap / 2))) return rc;
        hipLaunchKernelGGL(k_branch<W>, dim3(blocks(n, K1_TILE)), dim3(WG), 0, h->stream, h->keys[h->cur],
                           h->cap, (uint32_t)n, mask, h->tail[0], h->lb_states, h->counter, next_epoch(h),
                           h->d_scalar);
        HIP_TRY(hipGetLastError());
    }
    const int c = h->cur;
    // K2: tail into ascending order of term ^ mask
    Levels lv;
    int64_t const_off;
    int from_limb;
    if (mask_levels(mask_in, W, na, &lv, &const_off, &from_limb))
        return fail(PP_ERR_UNSUPPORTED, "rotation generator with more than 8 runs of set bits (16 levels)");
    prof_mark(h, 2);
    hipLaunchKernelGGL(k_xorperm<W>, dim3(blocks(na, K2_TILE)), dim3(WG), 0, h->stream, h->keys[c], h->cap,
                       h->tail[0], na, lv, const_off, from_limb, h->tail[1], h->d_scalar + 2);
    HIP_TRY(hipGetLastError());
    prof_mark(h, 3);
    // K3: merge with truncation into the other buffer
    const size_t ntot = n + na;
    if (h->B == 1) {
        RotParams rp{cos_b[0], sin_b[0], max_weight, min_abs};
        prof_mark(h, 4);
        hipLaunchKernelGGL(k_merge1<W>, dim3(blocks(ntot, K3_TILE)), dim3(WG), 0, h->stream, h->keys[c], h->cap,
                           h->coeffs[c], (uint32_t)n, h->tail[1], na, mask, rp, h->keys[1 - c], h->cap,
                           h->coeffs[1 - c], h->lb_states, h->counter, next_epoch(h), h->d_scalar + 1,
                           h->d_scalar + 2);
    } else {
        const int B = h->B;
        std::memcpy(h->h_param, cos_b, B * sizeof(double));
        std::memcpy(h->h_param + B, sin_b, B * sizeof(double));
        HIP_TRY(hipMemcpyAsync(h->d_param, h->h_param, 2 * B * sizeof(double), hipMemcpyHostToDevice, h->stream));
        const int TB = tile_rows(B);
        const size_t shm = (size_t)TB * B * sizeof(double);
        const unsigned g = blocks(ntot, TB);
        prof_mark(h, 4);
#define LAUNCH_MB(TBV)                                                                                          \
    hipLaunchKernelGGL((k_mergeB<W, TBV>), dim3(g), dim3(WG), shm, h->stream, h->keys[c], h->cap, h->coeffs[c], \
                       (uint32_t)n, B, h->tail[1], na, mask, h->d_param, h->d_param + B, max_weight, min_abs,   \
                       h->keys[1 - c], h->cap, h->coeffs[1 - c], h->lb_states, h->counter, next_epoch(h),       \
                       h->d_scalar + 1, h->d_scalar + 2)
        switch (TB) {
        case 256: LAUNCH_MB(256); break;
        case 128: LAUNCH_MB(128); break;
        case 64: LAUNCH_MB(64); break;
        case 32: LAUNCH_MB(32); break;
        default: LAUNCH_MB(16); break;
        }
#undef LAUNCH_MB
    }
    HIP_TRY(hipGetLastError());
    prof_mark(h, 5);
    uint32_t res[2] = {0, 0};  // output count, error bits
    if (int rc = read_scalars(h, 1, 2, res)) return rc;
    if (res[1]) {
        HIP_TRY(hipMemsetAsync(h->d_scalar + 2, 0, sizeof(uint32_t), h->stream));
        return fail(PP_ERR_STATE, "internal consistency check failed in the rotation kernels (bits " +
                                      std::to_string(res[1]) + ")");
    }
    const uint32_t nout = res[0];
    prof_add(h, 1, PP_KERNEL_TAIL_PERMUTE, na * (8.0 + 8.0 * (W - from_limb)));
    prof_add(h, 2, PP_KERNEL_MERGE, row * ((double)n + (double)nout));
    h->cur = 1 - c;
    h->n = nout;
    if (st) st->n_out = nout;
    return PP_OK;
}

template <int W>
int noise_impl(pp_handle* h, int qubit, uint32_t damp_codes, const double* factor_b, int max_weight,
               double min_abs, pp_gate_stats* st) {
    const size_t n = h->n;
    if (st) *st = pp_gate_stats{(int64_t)n, (int64_t)n, (int64_t)n};
    h->gates += 1;
    h->term_apps += (int64_t)n;
    if (n == 0) return PP_OK;
    const int c = h->cur;
    if (h->B == 1) {
        prof_mark(h, 4);
        hipLaunchKernelGGL(k_noise1<W>, dim3(blocks(n, KN_TILE)), dim3(WG), 0, h->stream, h->keys[c], h->cap,
                           h->coeffs[c], (uint32_t)n, qubit, damp_codes, factor_b[0], max_weight, min_abs,
                           h->keys[1 - c], h->cap, h->coeffs[1 - c], h->lb_states, h->counter, next_epoch(h),
                           h->d_scalar + 1);
    } else {
        const int B = h->B;
        std::memcpy(h->h_param + 2 * B, factor_b, B * sizeof(double));
        HIP_TRY(hipMemcpyAsync(h->d_param + 2 * B, h->h_param + 2 * B, B * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
        prof_mark(h, 4);
        hipLaunchKernelGGL(k_noiseB<W>, dim3(blocks(n, 64)), dim3(WG), 0, h->stream, h->keys[c], h->cap,
                           h->coeffs[c], (uint32_t)n, B, qubit, damp_codes, h->d_param + 2 * B, max_weight,
                           min_abs, h->keys[1 - c], h->cap, h->coeffs[1 - c], h->lb_states, h->counter,
                           next_epoch(h), h->d_scalar + 1);
    }
    HIP_TRY(hipGetLastError());
    prof_mark(h, 5);
    uint32_t nout = 0;
    if (int rc = read_scalars(h, 1, 1, &nout)) return rc;
    prof_add(h, 2, PP_KERNEL_NOISE, (8.0 * h->W + 8.0 * h->B) * ((double)n + (double)nout));
    h->cur = 1 - c;
    h->n = nout;
    if (st) st->n_out = nout;
    return PP_OK;
}

template <int W>
int overlap_impl(pp_handle* h, double* out) {
    const int B = h->B;
    const size_t n = h->n;
    if (n == 0) {
        for (int b = 0; b < B; ++b) out[b] = 0.0;
        return PP_OK;
    }
    const int c = h->cur;
    const unsigned g = std::min<unsigned>(2048u, blocks(n, B == 1 ? WG : NWAVES));
    const size_t need = (size_t)g * B * 2;
    if (need > h->partial_cap) {
        if (h->d_partial) hipFree(h->d_partial);
        h->d_partial = nullptr;
        HIP_TRY(hipMalloc(&h->d_partial, need * sizeof(double)));
        h->partial_cap = need;
    }
    prof_mark(h, 6);
    if (B == 1) {
        hipLaunchKernelGGL(k_overlap1<W>, dim3(g), dim3(WG), 0, h->stream, h->keys[c], h->cap, h->coeffs[c],
                           (uint32_t)n, h->d_partial);
    } else {
        hipLaunchKernelGGL(k_overlapB<W>, dim3(g), dim3(WG), (size_t)NWAVES * B * 2 * sizeof(double), h->stream,
                           h->keys[c], h->cap, h->coeffs[c], (uint32_t)n, B, h->d_partial);
    }
    HIP_TRY(hipGetLastError());
    prof_mark(h, 7);
    hipLaunchKernelGGL(k_overlap_final, dim3(1), dim3(WG), 0, h->stream, h->d_partial, (int)g, B, h->d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, h->d_out, B * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    prof_add(h, 3, PP_KERNEL_OVERLAP, (8.0 * h->W + 8.0 * B) * (double)n);
    return PP_OK;
}

int key_cmp(const uint64_t* a, const uint64_t* b, int W) {
    for (int l = W - 1; l >= 0; --l)
        if (a[l] != b[l]) return a[l] < b[l] ? -1 : 1;
    return 0;
}

}  // namespace

// ============================================================ C ABI

extern "C" {

int pp_version(void) { return 100; }

const char* pp_last_error(void) { return g_err.c_str(); }

int pp_create(int nqubits, int key_words, int batch, size_t capacity, pp_handle** out) {
    if (!out) return fail(PP_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (nqubits < 1 || nqubits > 128) return fail(PP_ERR_ARG, "nqubits must be in 1..128");
    const int need = nqubits <= 32 ? 1 : (nqubits <= 64 ? 2 : 4);
    if (key_words == 0) key_words = need;
    if (!(key_words == 1 || key_words == 2 || key_words == 4) || key_words < need)
        return fail(PP_ERR_ARG, "key_words must be 1, 2 or 4 and hold 2*nqubits bits");
    if (batch < 1 || batch > 1024) return fail(PP_ERR_ARG, "batch must be in 1..1024");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(PP_ERR_HIP, "no HIP device available");
    }
    pp_handle* h = new pp_handle();
    h->nqubits = nqubits;
    h->W = key_words;
    h->B = batch;
    auto bail = [&](int rc) {
        pp_destroy(h);
        return rc;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PP_ERR_HIP, "hipStreamCreate failed"));
    if (hipMalloc(&h->counter, 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&h->d_scalar, 4 * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(&h->h_scalar, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&h->d_param, 3 * batch * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&h->h_param, 3 * batch * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&h->d_out, batch * sizeof(double)) != hipSuccess)
        return bail(fail(PP_ERR_NOMEM, "device allocation failed"));
    if (hipMemsetAsync(h->counter, 0, 4 * sizeof(uint32_t), h->stream) != hipSuccess ||
        hipMemsetAsync(h->d_scalar, 0, 4 * sizeof(uint32_t), h->stream) != hipSuccess)
        return bail(fail(PP_ERR_HIP, "hipMemset failed"));
    if (int rc = grow(h, std::max<size_t>(capacity, 1024))) return bail(rc);
    *out = h;
    return PP_OK;
}

int pp_destroy(pp_handle* h) {
    if (!h) return PP_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (int e = 0; e < 8; ++e)
        if (h->ev[e]) hipEventDestroy(h->ev[e]);
    free_buffers(h);
    if (h->counter) hipFree(h->counter);
    if (h->d_scalar) hipFree(h->d_scalar);
    if (h->h_scalar) hipHostFree(h->h_scalar);
    if (h->d_param) hipFree(h->d_param);
    if (h->h_param) hipHostFree(h->h_param);
    if (h->d_partial) hipFree(h->d_partial);
    if (h->d_out) hipFree(h->d_out);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return PP_OK;
}

int pp_info(const pp_handle* h, int* nqubits, int* key_words, int* batch, size_t* n, size_t* capacity) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (nqubits) *nqubits = h->nqubits;
    if (key_words) *key_words = h->W;
    if (batch) *batch = h->B;
    if (n) *n = h->n;
    if (capacity) *capacity = h->cap;
    return PP_OK;
}

int pp_upload(pp_handle* h, const uint64_t* keys, const double* coeffs, size_t n) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (n > 0 && (!keys || !coeffs)) return fail(PP_ERR_ARG, "keys or coeffs is NULL");
    if (n > MAX_TERMS) return fail(PP_ERR_CAPACITY, "term count exceeds 2^31 - 1");
    const int W = h->W, B = h->B;
    const int qbits = 2 * h->nqubits;
    for (size_t i = 0; i < n; ++i)
        for (int l = 0; l < W; ++l) {
            const int lo = 64 * l;
            if (qbits >= lo + 64) continue;
            const uint64_t allowed = qbits <= lo ? 0ull : ((1ull << (qbits - lo)) - 1);
            if (keys[i * W + l] & ~allowed) return fail(PP_ERR_ARG, "a term acts on a qubit beyond nqubits");
        }
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(),
              [&](size_t a, size_t b) { return key_cmp(keys + a * W, keys + b * W, W) < 0; });
    for (size_t i = 1; i < n; ++i)
        if (key_cmp(keys + idx[i] * W, keys + idx[i - 1] * W, W) == 0)
            return fail(PP_ERR_ARG, "duplicate Pauli string in upload");
    if (n > h->cap) {
        h->n = 0;
        if (int rc = grow(h, n)) return rc;
    }
    std::vector<uint64_t> planes((size_t)W * std::max<size_t>(n, 1));
    std::vector<double> rows((size_t)B * std::max<size_t>(n, 1));
    for (size_t i = 0; i < n; ++i) {
        for (int l = 0; l < W; ++l) planes[(size_t)l * n + i] = keys[idx[i] * W + l];
        std::memcpy(&rows[i * B], coeffs + idx[i] * B, B * sizeof(double));
    }
    h->cur = 0;
    for (int l = 0; l < W && n > 0; ++l)
        HIP_TRY(hipMemcpyAsync(h->keys[0] + (size_t)l * h->cap, &planes[(size_t)l * n], n * sizeof(uint64_t),
                               hipMemcpyHostToDevice, h->stream));
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(h->coeffs[0], rows.data(), n * B * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->n = n;
    return PP_OK;
}

int pp_download(pp_handle* h, uint64_t* keys, double* coeffs, size_t cap, size_t* n) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (n) *n = h->n;
    if (cap < h->n) return fail(PP_ERR_ARG, "output arrays too small");
    const size_t m = h->n;
    const int W = h->W, B = h->B;
    if (m == 0) return PP_OK;
    std::vector<uint64_t> planes((size_t)W * m);
    for (int l = 0; l < W; ++l)
        HIP_TRY(hipMemcpyAsync(&planes[(size_t)l * m], h->keys[h->cur] + (size_t)l * h->cap, m * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, h->stream));
    if (coeffs)
        HIP_TRY(hipMemcpyAsync(coeffs, h->coeffs[h->cur], m * B * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (keys)
        for (size_t i = 0; i < m; ++i)
            for (int l = 0; l < W; ++l) keys[i * W + l] = planes[(size_t)l * m + i];
    return PP_OK;
}

int pp_apply_rotation(pp_handle* h, const uint64_t* mask, const double* cos_b, const double* sin_b, int max_weight,
                      double min_abs_coeff, pp_gate_stats* stats) {
    if (!h || !mask || !cos_b || !sin_b) return fail(PP_ERR_ARG, "NULL argument");
    uint64_t m4[4] = {0, 0, 0, 0};
    for (int l = 0; l < h->W; ++l) m4[l] = mask[l];
    if (int rc = check_mask(h, m4)) return rc;
    switch (h->W) {
    case 1: return rotation_impl<1>(h, m4, cos_b, sin_b, max_weight, min_abs_coeff, stats);
    case 2: return rotation_impl<2>(h, m4, cos_b, sin_b, max_weight, min_abs_coeff, stats);
    default: return rotation_impl<4>(h, m4, cos_b, sin_b, max_weight, min_abs_coeff, stats);
    }
}

int pp_apply_pauli_noise(pp_handle* h, int qubit, uint32_t damp_codes, const double* factor_b, int max_weight,
                         double min_abs_coeff, pp_gate_stats* stats) {
    if (!h || !factor_b) return fail(PP_ERR_ARG, "NULL argument");
    if (qubit < 0 || qubit >= h->nqubits) return fail(PP_ERR_ARG, "noise qubit out of range");
    if (damp_codes & ~0xeu) return fail(PP_ERR_ARG, "damp_codes may only hold X, Y, Z (bits 1..3)");
    switch (h->W) {
    case 1: return noise_impl<1>(h, qubit, damp_codes, factor_b, max_weight, min_abs_coeff, stats);
    case 2: return noise_impl<2>(h, qubit, damp_codes, factor_b, max_weight, min_abs_coeff, stats);
    default: return noise_impl<4>(h, qubit, damp_codes, factor_b, max_weight, min_abs_coeff, stats);
    }
}

int pp_run_circuit(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
                   const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight,
                   double min_abs_coeff, int64_t* counts) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (ngates > 0 && !gates) return fail(PP_ERR_ARG, "gates is NULL");
    const int B = h->B;
    std::vector<double> a(B), s(B);
    for (size_t g = 0; g < ngates; ++g) {
        const pp_gate& gt = gates[g];
        if (counts) counts[g] = (int64_t)h->n;
        if (gt.param >= 0 && ((size_t)gt.param >= nparams || !theta))
            return fail(PP_ERR_ARG, "gate " + std::to_string(g) + " reads a parameter beyond nparams");
        auto val = [&](int b) { return gt.param < 0 ? gt.value : theta[(size_t)gt.param * B + b]; };
        int rc = PP_OK;
        switch (gt.kind) {
        case PP_GATE_ROTATION:
            for (int b = 0; b < B; ++b) {
                a[b] = std::cos(val(b));
                s[b] = std::sin(val(b));
            }
            rc = pp_apply_rotation(h, gt.mask, a.data(), s.data(), max_weight, min_abs_coeff, nullptr);
            break;
        case PP_GATE_PAULI_NOISE:
            for (int b = 0; b < B; ++b) {
                const double lam = val(b);
                if (!(lam >= 0.0 && lam <= 1.0)) return fail(PP_ERR_ARG, "noise strength must be in [0, 1]");
                a[b] = 1 - lam;
            }
            rc = pp_apply_pauli_noise(h, gt.qubit, gt.damp_codes, a.data(), max_weight, min_abs_coeff, nullptr);
            break;
        default:
            (void)lut_changes;
            (void)lut_signs;
            (void)nluts;
            return fail(PP_ERR_UNSUPPORTED, "gate kind " + std::to_string(gt.kind) + " is not implemented");
        }
        if (rc) return rc;
    }
    return PP_OK;
}

int pp_overlap_zero(pp_handle* h, double* out_b) {
    if (!h || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    switch (h->W) {
    case 1: return overlap_impl<1>(h, out_b);
    case 2: return overlap_impl<2>(h, out_b);
    default: return overlap_impl<4>(h, out_b);
    }
}

int pp_synchronize(pp_handle* h) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PP_OK;
}

int pp_profile_control(pp_handle* h, int enable_timing, int reset) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (enable_timing == 1 && !h->ev[0]) {
        for (int e = 0; e < 8; ++e) HIP_TRY(hipEventCreate(&h->ev[e]));
    }
    if (enable_timing >= 0) h->prof_on = enable_timing == 1;
    if (reset) {
        h->gates = h->term_apps = 0;
        for (int k = 0; k < PP_KERNEL_KINDS; ++k) {
            h->launches[k] = 0;
            h->ms[k] = 0.0;
            h->alg_bytes[k] = 0.0;
        }
    }
    return PP_OK;
}

int pp_profile_read(const pp_handle* h, pp_profile* out) {
    if (!h || !out) return fail(PP_ERR_ARG, "NULL argument");
    out->gates = h->gates;
    out->term_applications = h->term_apps;
    for (int k = 0; k < PP_KERNEL_KINDS; ++k) {
        out->launches[k] = h->launches[k];
        out->milliseconds[k] = h->ms[k];
        out->algorithmic_bytes[k] = h->alg_bytes[k];
    }
    return PP_OK;
}

}  // extern "C"
