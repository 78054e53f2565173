// pphip.hip -- host side of libpphip.so: handle, buffers, launches, C ABI.
//
// Device layout of a handle (capacity `cap` terms, W key limbs, B columns):
//   keys[2]    W limb planes of cap uint64 each (limb l of term i at l*cap + i)
//   coeffs[2]  cap x B doubles, row-major
//   tail[2]    cap uint32 entries: index of an anticommuting term, sign in bit 31
//   tkeys      W limb planes of cap words: permuted tail keys (term ^ mask);
//              tcoef (B = 1) their values, tail[1] (B > 1) their indices
//   tl_keys    W limb planes of cap words: the terms listed in tail[0], in list
//              order (tl_coef: their coefficients, B = 1)
//   states     one look-back word per tile of the largest compaction launch, two
//              chains, then one round-window word per tile (one-column merge)
//   gstates    one look-back word per group of 64 such tiles
//   splits     merge-path split of every merge tile boundary
//   ssamp      every 64th sum entry (W words each), written by the merge that
//              made the sum, else by the permutation
//   tsamp      permuted tail entries 63 modulo 64 (W words each)
//   ctl        the live size and buffer of the sum, double-buffered by gate
//              parity, plus the overflow record (pp_kernels.h)
// The host enqueues gates without reading sizes back; it synchronizes once
// per chunk of gates to catch an overflow, grows the buffers and resumes at
// the gate that overflowed.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pphip.h"
#include "pp_kernels.h"
#include "pp_sort.h"
#include "pp_mc.h"
#include "pp_lds.h"

using namespace pp;

// 1024 blocks x 64 columns x 2 (double-double partials) + 64 results
constexpr size_t SCRATCH_DOUBLES_DECL = 1024 * 64 * 2 + 64;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES_DECL];

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

// Scratch of pp_overlap_noalloc: the overlap kernels' per-block double-double
// partials (at most SCRATCH_BLOCKS blocks of SCRATCH_COLS columns) and the
// SCRATCH_COLS results, about 1 MiB per GPU, held until the process ends.
// Host code never passes the array's name to a kernel: each device's address
// comes from hipGetSymbolAddress, taken once per device by pp_create with that
// device current (scratch_on).
constexpr int SCRATCH_COLS = 64;
constexpr unsigned SCRATCH_BLOCKS = 1024;
constexpr size_t SCRATCH_DOUBLES = (size_t)SCRATCH_BLOCKS * SCRATCH_COLS * 2 + SCRATCH_COLS;
static_assert(SCRATCH_DOUBLES == SCRATCH_DOUBLES_DECL, "scratch layout and declaration agree");
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[MAX_DEVICES] = {};
std::string g_scratch_err[MAX_DEVICES];

// The device address of `scratch` on device `dev` (current), taken once.
int scratch_setup(int dev, double** out);
// Launch shapes the tuning table may set (pp_tune, pp_load_tuning); every
// choice gives bit-identical results.
struct Tuning {
    int merge_items = 0;   // one-column merge tile = 256 x this (2, 3, 4); 0: 3 for 256-bit keys, else 4
    int merge_defer = 0;   // one-column merge (not recording) stores each tile one tile behind (k_merge1d)
    int64_t split_below = 1 << 16;  // one-column merge: sums below this many terms search their own splits
    int perm_items = 2;    // tail permutation: list entries per thread (1, 2, 4)
    int batch_rows = 0;    // batched merge: term rows per tile (16 .. 256); 0: by B
    int radix_bits = 8;    // Clifford / load-merge radix sort: digit width (4 .. 8)
    int grid_per_cu = 4;   // persistent grids: at most this many workgroups per compute unit (1 .. 8)
    int run_chunk = 128;   // gates enqueued between overflow checks (16 .. 1024)
};
constexpr int MAX_GRID_PER_CU = 8;
constexpr unsigned ADJ_BLOCKS = 1024;      // blocks of the adjoint sweep's rotation kernels
constexpr int PRES_WORDS = 12;             // X, Y, Z presence planes of up to 4 limbs

}  // namespace

struct pp_handle {
    int nqubits = 0, W = 1, B = 1;
    int device = 0;                // HIP device of the handle's stream and buffers
    double* scratch = nullptr;     // that device's address of the file-scope scratch array
    size_t cap = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;        // false: a work handle on its parent's stream
    Bufs bufs{};
    size_t lb_cap = 0;
    unsigned gmax = 2048;          // persistent grid size (tun.grid_per_cu x compute units)
    int cus = 256;                 // compute units
    Tuning tun;                    // launch shapes
    std::unordered_map<const void*, unsigned> resident;  // co-resident workgroups per kernel
    uint32_t spin_limit = 1u << 16;  // bound of one look-back wait with static tiles
    bool force_dynamic = false;      // take tiles from the atomic counter always
    bool force_rerun = false;        // static-tile waits report running out (tests the rerun)
    bool chain_tails = true;         // merges list the next rotation's touched terms
    bool prev_lists = false;         // the last gate enqueued lists the next rotation's touched terms
    bool skip_identity = true;       // leave out rotations the Pauli presence shows touch nothing
    int64_t identity_gates = 0;      // rotations left out that way
    int64_t lds_gates = 0;           // gates applied by shared-memory runs (k_lds_run)
    bool lds_runs = true;            // run small sums' rotations and noise in shared memory
    bool use_graphs = false;         // capture chunks of gates in hipGraphs and replay them
    struct ChunkGraph {
        hipGraphExec_t exec = nullptr;
        uint32_t seq_delta = 0, lb_delta = 0;
        bool prev_lists = false;
    };
    std::unordered_map<uint64_t, ChunkGraph> graphs;
    int64_t graph_launches = 0, graph_captures = 0;
    LdsGate* d_ldsg = nullptr;       // their descriptors, and a pinned staging copy
    LdsGate* h_ldsg = nullptr;
    size_t ldsg_cap = 0;
    uint32_t safe_until = 0;         // gates before this sequence number use dynamic tiles
    int64_t fallbacks = 0;           // gates run again with dynamic tiles
    size_t adj_defer_bytes = (size_t)2 << 30;  // largest store of per-rotation partial gradients (adjoint)
    bool last_adj_deferred = false;  // the last adjoint sweep kept its partials and summed them per segment
    Ctl* d_ctl = nullptr;
    Ctl* h_ctl = nullptr;          // pinned mirror
    uint32_t seq = 0;              // next gate's sequence number; the sum's state is st[seq & 1]
    uint32_t lb_count = 0;         // look-back launches so far (counter slot parity)
    uint32_t epoch = 0;
    // per-run device arrays
    uint32_t* d_log = nullptr;     // LOG_WORDS per gate: n_in, n_touched, n_out, roff, coff
    uint32_t* h_log = nullptr;
    size_t log_cap = 0;            // gates
    double* d_trig = nullptr;      // per gate: B cos, B sin (rotation) or B factors (noise)
    double* h_trig = nullptr;
    size_t trig_cap = 0;           // doubles
    // recording for the adjoint sweep (allocated on first use)
    bool recording = false;
    uint32_t* d_tpj = nullptr;
    uint32_t* d_rec = nullptr;
    double* d_rcoef = nullptr;
    size_t tpj_cap = 0, rec_words = 0, rec_rows = 0;
    // Clifford: radix-sort buffers (sized for cap) and the gate tables of a run
    uint64_t* d_rs_k[2] = {nullptr, nullptr};
    uint32_t* d_rs_v[2] = {nullptr, nullptr};
    uint32_t* d_rs_hist = nullptr;
    size_t rs_cap = 0, rs_tiles = 0;
    uint64_t* d_lut_c = nullptr;
    double* d_lut_s = nullptr;
    size_t lut_cap = 0;            // tables
    // overlap
    double* d_partial = nullptr;
    size_t partial_cap = 0;
    double* d_out = nullptr;
    double* d_norm = nullptr;         // B doubles: 1 / the identity's coefficient (imaginary rotations)
    // resampling: per chunk sums, counts, interval starts and output offsets; the draws
    double* d_mc_sum = nullptr;
    uint32_t* d_mc_cnt = nullptr;
    double* d_mc_start = nullptr;
    uint32_t* d_mc_out = nullptr;
    size_t mc_chunks = 0;
    double* d_mc_draws = nullptr;
    size_t mc_draws_cap = 0;
    uint64_t* d_pres_part = nullptr;  // presence partials, gmax x 12 words, then the 12 result words
    uint64_t* h_pres = nullptr;       // pinned copy of the result
    // profile
    int prof_mode = 0;             // 0 off, 1 merge/noise kernels, 2 every kernel
    std::vector<hipEvent_t> ev;    // 6 per gate slot
    uint64_t* d_ktime = nullptr;   // profile mode 1: device-clock start/end of each gate's merge or noise kernel
    std::vector<uint64_t> h_ktime;
    size_t ktime_cap = 0;
    uint64_t* kt = nullptr;        // the enqueued gate's d_ktime slot (mode 1), else null
    uint64_t* d_wgstamp = nullptr; // mode 1: start and end of every workgroup of the timed launch
    unsigned last_grid = 0;        // grid of the last launch()
    double clock_khz = 100000.0;   // rate of that clock
    int64_t gates = 0, term_apps = 0;
    int64_t launches[PP_KERNEL_KINDS] = {};
    double ms[PP_KERNEL_KINDS] = {};
    double alg_bytes[PP_KERNEL_KINDS] = {};
};

namespace {

void free_terms(pp_handle* h) {
    for (int b = 0; b < 2; ++b) {
        if (h->bufs.keys[b]) hipFree(h->bufs.keys[b]);
        if (h->bufs.coeffs[b]) hipFree(h->bufs.coeffs[b]);
        if (h->bufs.tail[b]) hipFree(h->bufs.tail[b]);
        h->bufs.keys[b] = nullptr;
        h->bufs.coeffs[b] = nullptr;
        h->bufs.tail[b] = nullptr;
    }
    if (h->bufs.states) hipFree(h->bufs.states);
    if (h->bufs.tkeys) hipFree(h->bufs.tkeys);
    if (h->bufs.tcoef) hipFree(h->bufs.tcoef);
    h->bufs.tcoef = nullptr;
    if (h->bufs.tl_keys) hipFree(h->bufs.tl_keys);
    if (h->bufs.tl_coef) hipFree(h->bufs.tl_coef);
    h->bufs.tl_keys = nullptr;
    h->bufs.tl_coef = nullptr;
    if (h->bufs.splits) hipFree(h->bufs.splits);
    if (h->bufs.gstates) hipFree(h->bufs.gstates);
    if (h->bufs.ssamp) hipFree(h->bufs.ssamp);
    if (h->bufs.tsamp) hipFree(h->bufs.tsamp);
    h->bufs.gstates = nullptr;
    h->bufs.ssamp = nullptr;
    h->bufs.tsamp = nullptr;
    h->bufs.states = nullptr;
    h->bufs.tkeys = nullptr;
    h->bufs.splits = nullptr;
}

void drop_graphs(pp_handle* h) {
    for (auto& kv : h->graphs)
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
    h->graphs.clear();
}

int scratch_setup(int dev, double** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(PP_ERR_HIP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev] && g_scratch_err[dev].empty()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            g_scratch_err[dev] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        } else {
            g_scratch[dev] = static_cast<double*>(p);
        }
    }
    *out = g_scratch[dev];
    if (!*out) return fail(PP_ERR_HIP, g_scratch_err[dev]);
    return PP_OK;
}

int tile_rows(int B) {
    if (B <= 16) return 256;
    if (B <= 64) return 64;  // 128 rows for B = 32 measured 7 % slower (a noisy forward run), 32 rows for B = 64 20 % slower
    if (B <= 128) return 32;
    return 16;
}

// the batched merge's rows per tile: the tuning table's, or by B
int batch_rows(const pp_handle* h) { return h->tun.batch_rows ? h->tun.batch_rows : tile_rows(h->B); }

// the batched merge's dynamic shared memory limit (rows x B doubles)
constexpr size_t MAX_ROWS_LDS = 64 * 1024;

// look-back words for the largest compaction launch over up to cap entries
// (every look-back kernel's tiles are at least this wide: 512 terms for one
// column, 16 rows for batches, whatever the tuning table picks)
static_assert(K1_TILE >= WG * MIN_MERGE_ITEMS, "branch tiles narrower than the merge's would outnumber the look-back words");
size_t tiles_needed(size_t cap, int B) {
    const size_t tile = B > 1 ? (size_t)16 : (size_t)WG * MIN_MERGE_ITEMS;
    return cap / tile + 64;
}

int sync_ctl(pp_handle* h) {
    HIP_TRY(hipMemcpyAsync(h->h_ctl, h->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PP_OK;
}

int put_ctl(pp_handle* h) {
    HIP_TRY(hipMemcpyAsync(h->d_ctl, h->h_ctl, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PP_OK;
}

// live size and buffer of the sum
int live(pp_handle* h, uint32_t* n, uint32_t* cur) {
    if (int rc = sync_ctl(h)) return rc;
    const GateState& s = h->h_ctl->st[h->seq & 1];
    *n = s.n;
    *cur = s.cur;
    return PP_OK;
}

// Set the sum's state (its size, buffer and record offsets) and clear the
// overflow record and counters.
int reset_ctl(pp_handle* h, const GateState& st) {
    std::memset(h->h_ctl, 0, sizeof(Ctl));
    h->prev_lists = false;
    GateState& s = h->h_ctl->st[h->seq & 1];
    s = GateState{};
    s.n = st.n;
    s.cur = st.cur;
    s.roff = st.roff;
    s.coff = st.coff;
    return put_ctl(h);
}

int reset_ctl(pp_handle* h, uint32_t n, uint32_t cur) {
    GateState st{};
    st.n = n;
    st.cur = cur;
    return reset_ctl(h, st);
}

// (Re)allocate every term buffer for `newcap` terms, keeping the live sum.
int grow(pp_handle* h, size_t newcap) {
    if (newcap <= h->cap) return PP_OK;
    drop_graphs(h);  // they hold the old buffers' addresses
    if (newcap > MAX_TERMS + 1) newcap = MAX_TERMS + 1;
    const int W = h->W, B = h->B;
    uint32_t n = 0, cur = 0;
    GateState keep{};
    if (h->cap) {
        if (int rc = live(h, &n, &cur)) return rc;
        keep = h->h_ctl->st[h->seq & 1];
    }
    uint64_t* nk[2] = {nullptr, nullptr};
    double* nc[2] = {nullptr, nullptr};
    uint32_t* nt[2] = {nullptr, nullptr};
    uint64_t* ns = nullptr;
    uint64_t* ntk = nullptr;
    double* ntc = nullptr;
    uint64_t* nlk = nullptr;
    double* nlc = nullptr;
    uint32_t* nsp = nullptr;
    uint64_t* ngs = nullptr;
    uint64_t* nss = nullptr;
    uint64_t* nts = nullptr;
    const size_t nsamp = (newcap / 64 + 2) * W;
    const size_t nlb = tiles_needed(newcap, B);
    const size_t ngw = nlb / LB_GROUP + 2;
    // two look-back chains (merge offsets and the next gate's touched list)
    bool ok = true;
    for (int b = 0; b < 2 && ok; ++b) {
        ok = ok && hipMalloc(&nk[b], newcap * W * sizeof(uint64_t)) == hipSuccess;
        ok = ok && hipMalloc(&nc[b], newcap * B * sizeof(double)) == hipSuccess;
        ok = ok && hipMalloc(&nt[b], newcap * sizeof(uint32_t)) == hipSuccess;
    }
    ok = ok && hipMalloc(&ns, 3 * nlb * sizeof(uint64_t)) == hipSuccess;
    ok = ok && hipMalloc(&ntk, newcap * W * sizeof(uint64_t)) == hipSuccess;
    if (B == 1) ok = ok && hipMalloc(&ntc, newcap * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&nlk, newcap * W * sizeof(uint64_t)) == hipSuccess;
    if (B == 1) ok = ok && hipMalloc(&nlc, newcap * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&nsp, nlb * sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMalloc(&ngs, 2 * ngw * sizeof(uint64_t)) == hipSuccess;
    ok = ok && hipMalloc(&nss, nsamp * sizeof(uint64_t)) == hipSuccess;
    ok = ok && hipMalloc(&nts, nsamp * sizeof(uint64_t)) == hipSuccess;
    if (!ok) {
        for (int b = 0; b < 2; ++b) {
            if (nk[b]) hipFree(nk[b]);
            if (nc[b]) hipFree(nc[b]);
            if (nt[b]) hipFree(nt[b]);
        }
        if (ns) hipFree(ns);
        if (ntk) hipFree(ntk);
        if (ntc) hipFree(ntc);
        if (nlk) hipFree(nlk);
        if (nlc) hipFree(nlc);
        if (nsp) hipFree(nsp);
        if (ngs) hipFree(ngs);
        if (nss) hipFree(nss);
        if (nts) hipFree(nts);
        (void)hipGetLastError();
        return fail(PP_ERR_NOMEM, "device allocation of " + std::to_string(newcap) + " terms failed");
    }
    HIP_TRY(hipMemsetAsync(ns, 0, 3 * nlb * sizeof(uint64_t), h->stream));
    HIP_TRY(hipMemsetAsync(ngs, 0, 2 * ngw * sizeof(uint64_t), h->stream));
    if (n > 0) {
        for (int l = 0; l < W; ++l)
            HIP_TRY(hipMemcpyAsync(nk[0] + (size_t)l * newcap, h->bufs.keys[cur] + (size_t)l * h->cap,
                                   (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(nc[0], h->bufs.coeffs[cur], (size_t)n * B * sizeof(double), hipMemcpyDeviceToDevice,
                               h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    free_terms(h);
    for (int b = 0; b < 2; ++b) {
        h->bufs.keys[b] = nk[b];
        h->bufs.coeffs[b] = nc[b];
        h->bufs.tail[b] = nt[b];
    }
    h->bufs.states = ns;
    h->bufs.tkeys = ntk;
    h->bufs.tcoef = ntc;
    h->bufs.tl_keys = nlk;
    h->bufs.tl_coef = nlc;
    h->bufs.splits = nsp;
    h->bufs.gstates = ngs;
    h->bufs.lb_tiles = (uint32_t)nlb;
    h->bufs.lb_groups = (uint32_t)ngw;
    h->bufs.ssamp = nss;
    h->bufs.tsamp = nts;
    h->bufs.cap = (uint32_t)newcap;
    h->bufs.batch = (uint32_t)h->B;
    h->lb_cap = nlb;
    h->cap = newcap;
    h->epoch = 0;
    if (h->recording) {
        if (h->d_tpj) hipFree(h->d_tpj);
        h->d_tpj = nullptr;
        h->tpj_cap = 0;
        HIP_TRY(hipMalloc(&h->d_tpj, newcap * sizeof(uint32_t)));
        h->tpj_cap = newcap;
        h->bufs.tpj = h->d_tpj;
    }
    keep.n = n;
    keep.cur = 0;
    return reset_ctl(h, keep);
}

uint32_t next_epoch(pp_handle* h) {
    h->epoch = (h->epoch + 1) & LB_EPOCH_MASK;
    if (h->epoch == 0) {
        // wrapped: words of old launches could match again, so clear them
        hipMemsetAsync(h->bufs.states, 0, 3 * (size_t)h->bufs.lb_tiles * sizeof(uint64_t), h->stream);
        hipMemsetAsync(h->bufs.gstates, 0, 2 * (size_t)h->bufs.lb_groups * sizeof(uint64_t), h->stream);
        h->epoch = 1;
    }
    return h->epoch;
}

Launch lb_launch(pp_handle* h, uint32_t* glog) {
    Launch L{};
    L.seq = h->seq;
    L.epoch = next_epoch(h);
    L.slot = (h->lb_count++) & 1u;
    L.dynamic = (h->force_dynamic || h->seq < h->safe_until) ? 1u : 0u;
    L.spin_limit = h->spin_limit;
    L.flags = (h->force_rerun && !L.dynamic) ? 1u : 0u;
    L.glog = glog;
    L.ktime = nullptr;
    return L;
}

// lb_launch of a gate's merge or noise kernel: in profile mode 1 it stamps
// its device-clock start and end into the gate's slot
Launch kt_launch(pp_handle* h, uint32_t* glog) {
    Launch L = lb_launch(h, glog);
    L.ktime = h->kt ? h->d_wgstamp : nullptr;
    return L;
}

// after a kt_launch kernel: fold its workgroups' stamps into the gate's slot
void ktime_finish(pp_handle* h) {
    if (h->kt) hipLaunchKernelGGL(k_ktime_reduce, dim3(1), dim3(WG), 0, h->stream, h->d_wgstamp, h->last_grid, h->kt);
}

// Workgroups of `fn` the whole chip holds at once (occupancy x compute units).
unsigned resident_wgs(pp_handle* h, const void* fn, size_t shm) {
    auto it = h->resident.find(fn);
    if (it != h->resident.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, WG, shm) != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    const unsigned r = (unsigned)nb * (unsigned)h->cus;
    h->resident[fn] = r;
    return r;
}

// Grid of a persistent kernel: enough workgroups for `expect` entries, at most
// what the chip holds at once. Gates run back to back with the live size on
// the device only, so the forward pipeline passes the capacity: a grid sized
// from an older, smaller count makes every workgroup walk several tiles in
// series on the gates where the sum has grown, which costs far more than the
// idle workgroups a full grid dispatches on the gates where it is small.
unsigned grid_for(pp_handle* h, size_t expect, size_t per_tile) {
    const size_t tiles = (expect + per_tile - 1) / per_tile;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(tiles, h->gmax));
}

// Grid of a look-back kernel: as grid_for, and never more workgroups than are
// resident together, which static tiles need.
unsigned lb_grid(pp_handle* h, const void* fn, size_t shm, size_t expect, size_t per_tile) {
    return std::min(grid_for(h, expect, per_tile), resident_wgs(h, fn, shm));
}

int ensure_run(pp_handle* h, size_t ngates) {
    if (ngates > h->log_cap) {
        if (h->d_log) hipFree(h->d_log);
        if (h->h_log) hipHostFree(h->h_log);
        h->d_log = nullptr;
        h->h_log = nullptr;
        const size_t c = std::max<size_t>(ngates, 1024);
        HIP_TRY(hipMalloc(&h->d_log, LOG_WORDS * c * sizeof(uint32_t)));
        HIP_TRY(hipHostMalloc(&h->h_log, LOG_WORDS * c * sizeof(uint32_t), hipHostMallocDefault));
        h->log_cap = c;
    }
    const size_t need = ngates * 2 * (size_t)h->B;
    if (h->B > 1 && need > h->trig_cap) {
        if (h->d_trig) hipFree(h->d_trig);
        if (h->h_trig) hipHostFree(h->h_trig);
        h->d_trig = nullptr;
        h->h_trig = nullptr;
        const size_t c = std::max<size_t>(need, 4096);
        HIP_TRY(hipMalloc(&h->d_trig, c * sizeof(double)));
        HIP_TRY(hipHostMalloc(&h->h_trig, c * sizeof(double), hipHostMallocDefault));
        h->trig_cap = c;
    }
    if (h->prof_mode == 1 && h->ktime_cap < ngates) {
        if (h->d_ktime) hipFree(h->d_ktime);
        h->d_ktime = nullptr;
        h->ktime_cap = 0;
        HIP_TRY(hipMalloc(&h->d_ktime, 2 * ngates * sizeof(uint64_t)));
        if (!h->d_wgstamp)
            HIP_TRY(hipMalloc(&h->d_wgstamp, 2 * (size_t)std::max(64, MAX_GRID_PER_CU * h->cus) * sizeof(uint64_t)));
        h->h_ktime.resize(2 * ngates);
        h->ktime_cap = ngates;
    }
    if (h->prof_mode && h->ev.size() < 6 * ngates) {
        const size_t old = h->ev.size();
        h->ev.resize(6 * ngates, nullptr);
        for (size_t e = old; e < h->ev.size(); ++e) HIP_TRY(hipEventCreate(&h->ev[e]));
    }
    return PP_OK;
}

template <int W>
Key<W> to_key(const uint64_t* mask) {
    Key<W> k;
    for (int l = 0; l < W; ++l) k.l[l] = mask[l];
    return k;
}

// Levels of the tail permutation for mask m: every maximal run [lo, hi] of set
// bits contributes +(s+e) at level hi+1 and -(s+e) at level lo; a level at or
// above 64W spans the whole tail, so s+e = na there.
int mask_levels(const uint64_t* m, int W, Levels* lv) {
    lv->n = 0;
    lv->top_runs = 0;
    int minbit = 64 * W;
    int b = 64 * W - 1;
    while (b >= 0) {
        if ((m[b >> 6] >> (b & 63)) & 1ull) {
            const int hi = b;
            while (b >= 0 && ((m[b >> 6] >> (b & 63)) & 1ull)) --b;
            const int lo = b + 1;
            if (hi + 1 >= 64 * W) {
                lv->top_runs += 1;
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
    lv->from_limb = minbit >> 6;
    return 0;
}

int check_mask(const pp_handle* h, const uint64_t* mask) {
    bool any = false;
    const int qmax = 2 * h->nqubits;
    for (int l = 0; l < 4; ++l) {
        const uint64_t m = mask[l];
        any = any || m != 0;
        if (l >= h->W) {
            if (m) return fail(PP_ERR_ARG, "rotation mask acts on a qubit beyond the key width");
            continue;
        }
        const int lo = l * 64;
        if (qmax < lo + 64) {
            const int keep = qmax - lo;
            const uint64_t allowed = keep <= 0 ? 0ull : ((1ull << keep) - 1);
            if (m & ~allowed) return fail(PP_ERR_ARG, "rotation mask acts on a qubit beyond nqubits");
        }
    }
    if (!any) return fail(PP_ERR_ARG, "rotation mask is the identity");
    Levels lv;
    if (mask_levels(mask, h->W, &lv))
        return fail(PP_ERR_UNSUPPORTED, "rotation generator with more than 8 runs of set bits");
    return PP_OK;
}

// ---------------------------------------------------------------- one gate, enqueued

// Rotation: cos/sin by value for B = 1, or at d_trig (B cos then B sin) for B > 1.
// Launch on the handle's stream; with events, the dispatch itself records
// them at its start and end (no marker packets between kernels).
template <typename F, typename... Args>
void launch(pp_handle* h, hipEvent_t e0, hipEvent_t e1, F kernel, unsigned grid, size_t shm, Args... args) {
    h->last_grid = grid;
    if (e0 || e1)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(WG), (uint32_t)shm, h->stream, e0, e1, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(WG), shm, h->stream, args...);
}

// The tail permutation with the tuning table's list entries per thread; on a
// look-back launch (lbgrid) the grid is capped at the kernel's residency.
template <int W, bool BR, int PI>
void enq_perm_pi(pp_handle* h, hipEvent_t e0, hipEvent_t e1, const Launch& L, bool lbgrid, const Levels& lv,
                 const Key<W>& mask, const Key<W>& nmask, uint32_t has_next, uint32_t check, double tscale) {
    constexpr int T2 = WG * PI;
    const unsigned grid = lbgrid ? lb_grid(h, (const void*)k_xorperm<W, BR, PI>, 0, h->cap, T2)
                                 : grid_for(h, h->cap, T2);
    launch(h, e0, e1, k_xorperm<W, BR, PI>, grid, 0, h->bufs, h->d_ctl, L, lv, mask, nmask, has_next, check, tscale);
}

template <int W, bool BR>
void enq_perm(pp_handle* h, hipEvent_t e0, hipEvent_t e1, const Launch& L, bool lbgrid, const Levels& lv,
              const Key<W>& mask, const Key<W>& nmask, uint32_t has_next, uint32_t check, double tscale) {
    switch (h->tun.perm_items) {
    case 1: enq_perm_pi<W, BR, 1>(h, e0, e1, L, lbgrid, lv, mask, nmask, has_next, check, tscale); break;
    case 4: enq_perm_pi<W, BR, 4>(h, e0, e1, L, lbgrid, lv, mask, nmask, has_next, check, tscale); break;
    default: enq_perm_pi<W, BR, 2>(h, e0, e1, L, lbgrid, lv, mask, nmask, has_next, check, tscale); break;
    }
}

// The batched merge, with several rows per wave for B <= 32 (NARROW).
template <int W, int TB, typename... Args>
void launch_mb(pp_handle* h, hipEvent_t e0, hipEvent_t e1, unsigned grid, size_t shm, Args... args) {
    if (h->B <= 32)
        launch(h, e0, e1, k_mergeB<W, TB, true>, grid, shm, args...);
    else
        launch(h, e0, e1, k_mergeB<W, TB, false>, grid, shm, args...);
}

// The batched merge's partition (tiles of 64 rows or more; narrower tiles
// search their own splits).
// Sums of at least this many terms at the last synchronization take the
// batched merge's splits from the partition kernel.
constexpr uint32_t PART_B_MIN_TERMS = 1u << 18;

int part_b(const pp_handle* h, int tb) {
    return (tb % 64 == 0 && h->h_ctl->st[h->seq & 1].n >= PART_B_MIN_TERMS) ? 1 : 0;
}

template <int W, int TB>
void enq_partition_b(pp_handle* h) {
    if constexpr (TB % 64 == 0) {
        if (!part_b(h, TB)) return;
        Launch LP{};
        LP.seq = h->seq;
        launch(h, nullptr, nullptr, k_partition<W, TB>, grid_for(h, h->cap + h->cap / 2, (size_t)TB * NWAVES), 0,
               h->bufs, h->d_ctl, LP);
    }
}

// One one-column merge launch, with or without its own split searches.
template <int W, bool EMIT, bool REC, int MI>
void launch_merge1(pp_handle* h, hipEvent_t e4, hipEvent_t e5, bool self, const Launch& L, const Key<W>& mask,
                   const RotParams& rp, const Key<W>& nmask) {
    constexpr int TILE = WG * MI;
    const size_t expect = h->cap + h->cap / 2;
    if (self)
        launch(h, e4, e5, k_merge1<W, EMIT, REC, MI, true>,
               lb_grid(h, (const void*)k_merge1<W, EMIT, REC, MI, true>, 0, expect, TILE), 0, h->bufs, h->d_ctl, L,
               mask, rp, nmask);
    else
        launch(h, e4, e5, k_merge1<W, EMIT, REC, MI, false>,
               lb_grid(h, (const void*)k_merge1<W, EMIT, REC, MI, false>, 0, expect, TILE), 0, h->bufs, h->d_ctl, L,
               mask, rp, nmask);
}

// The one-column merge with the tuning table's tile (256 x MI diagonals): its
// partition (or, small sums, its own split searches), then the merge; variant
// 0 plain, 1 listing the next rotation's touched terms (nmask), 2 recording
// for the adjoint sweep, 3 both.
template <int W, int MI>
void enq_merge1_mi(pp_handle* h, hipEvent_t e3, hipEvent_t e4, hipEvent_t e5, int variant, const Key<W>& mask,
                   const RotParams& rp, const Key<W>& nmask, uint32_t* glog) {
    constexpr int TILE = WG * MI;
    const size_t expect = h->cap + h->cap / 2;
    const Launch LM = kt_launch(h, glog);
    // (event timing, profile mode 2, times the partition kernel: it keeps it)
    // sums of fewer terms than tun.split_below at the last synchronization
    // search their tiles' splits in the merge (static tiles), so no partition
    // kernel is launched: on a sum of a few tiles per workgroup the launch
    // costs more than the searches
    const bool self = !LM.dynamic && !h->tun.merge_defer && h->prof_mode != 2 &&
                      (int64_t)h->h_ctl->st[h->seq & 1].n < h->tun.split_below;
    if (!self) {
        Launch LP{};
        LP.seq = h->seq;
        launch(h, nullptr, e3, k_partition<W, TILE>, grid_for(h, expect, (size_t)TILE * NWAVES), 0, h->bufs,
               h->d_ctl, LP);
    }
    if (variant == 3)
        launch_merge1<W, true, true, MI>(h, e4, e5, self, LM, mask, rp, nmask);
    else if (variant == 2)
        launch_merge1<W, false, true, MI>(h, e4, e5, self, LM, mask, rp, mask);
    else if (h->tun.merge_defer && variant == 1)
        launch(h, e4, e5, k_merge1d<W, true, MI>, lb_grid(h, (const void*)k_merge1d<W, true, MI>, 0, expect, TILE), 0,
               h->bufs, h->d_ctl, LM, mask, rp, nmask);
    else if (h->tun.merge_defer)
        launch(h, e4, e5, k_merge1d<W, false, MI>, lb_grid(h, (const void*)k_merge1d<W, false, MI>, 0, expect, TILE), 0,
               h->bufs, h->d_ctl, LM, mask, rp, mask);
    else if (variant == 1)
        launch_merge1<W, true, false, MI>(h, e4, e5, self, LM, mask, rp, nmask);
    else
        launch_merge1<W, false, false, MI>(h, e4, e5, self, LM, mask, rp, mask);
    ktime_finish(h);
}

template <int W>
void enq_merge1(pp_handle* h, hipEvent_t e3, hipEvent_t e4, hipEvent_t e5, int variant, const Key<W>& mask,
                const RotParams& rp, const Key<W>& nmask, uint32_t* glog) {
    switch (h->tun.merge_items ? h->tun.merge_items : K3Cfg<W>::ITEMS) {
    case 2: enq_merge1_mi<W, 2>(h, e3, e4, e5, variant, mask, rp, nmask, glog); break;
    case 3: enq_merge1_mi<W, 3>(h, e3, e4, e5, variant, mask, rp, nmask, glog); break;
    default: enq_merge1_mi<W, 4>(h, e3, e4, e5, variant, mask, rp, nmask, glog); break;
    }
}

// Whether a rotation's merge may list the next rotation's touched terms (the
// merges that record list them too).
bool chain_ok(const pp_handle* h) { return h->chain_tails; }

// next_mask: the mask of the rotation enqueued right after this one, if any;
// the merge then lists that gate's touched terms.
// Events (6 per gate): branch [0, 1], permutation + partition [2, 3] (every
// kernel timed, mode 2); merge [4, 5] (mode 1 and 2).
template <int W>
void enq_rotation(pp_handle* h, const uint64_t* mask_in, double c1, double s1, const double* d_trig, int max_weight,
                  double min_abs, uint32_t* glog, hipEvent_t* ev, const uint64_t* next_mask) {
    const Key<W> mask = to_key<W>(mask_in);
    Levels lv;
    mask_levels(mask_in, W, &lv);
    const bool all = h->prof_mode == 2 && ev;
    const bool mrg = h->prof_mode == 2 && ev;
    auto E = [&](int k, bool on) -> hipEvent_t { return on ? ev[k] : nullptr; };
    // Chained rotations (recording too): every gate lists the next one's
    // touched terms (its merge, or its permutation kernel when it touches
    // nothing), so after the first gate of a chain k_branch is not launched and
    // the permutation kernel ends the branch step.
    const bool chained = chain_ok(h) && h->prof_mode != 2;
    const bool listed = chained && h->prev_lists;
    h->prev_lists = false;
    if (!listed)
        launch(h, E(0, all), E(1, all), k_branch<W, false>,
               lb_grid(h, (const void*)k_branch<W, false>, 0, h->cap, K1_TILE), 0, h->bufs, h->d_ctl,
               lb_launch(h, glog), mask);
    Launch L2{};
    L2.seq = h->seq;
    L2.glog = glog;
    const bool part = h->B == 1;
    if (chained)
        enq_perm<W, true>(h, E(2, all), nullptr, lb_launch(h, glog), true, lv, mask,
                          next_mask ? to_key<W>(next_mask) : mask, next_mask ? 1u : 0u, listed ? 1u : 0u, s1);
    else
        enq_perm<W, false>(h, E(2, all), part ? nullptr : E(3, all), L2, false, lv, mask, mask, 0u, 0u, s1);
    const size_t expect = h->cap + h->cap / 2;
    if (part) {
        RotParams rp{c1, s1, max_weight, min_abs};
        enq_merge1<W>(h, E(3, all), E(4, mrg), E(5, mrg), h->recording ? (next_mask ? 3 : 2) : (next_mask ? 1 : 0), mask, rp,
                      next_mask ? to_key<W>(next_mask) : mask, glog);
    } else {
        const int B = h->B;
        const int TB = batch_rows(h);
        const size_t shm = (size_t)TB * B * sizeof(double);
        const Launch L3 = kt_launch(h, glog);
        const bool emitB = chained && next_mask;
#define LAUNCH_MB(TBV)                                                                                         \
    enq_partition_b<W, TBV>(h);                                                                                \
    launch_mb<W, TBV>(h, E(4, mrg), E(5, mrg), lb_grid(h, B <= 32 ? (const void*)k_mergeB<W, TBV, true> : (const void*)k_mergeB<W, TBV>, shm, expect, TBV), \
           shm, h->bufs, h->d_ctl, L3, mask, B, d_trig, max_weight, min_abs, -1, 1.0, 0, nullptr, part_b(h, TBV), emitB ? 1 : 0, emitB ? to_key<W>(next_mask) : mask)
        switch (TB) {
        case 256: LAUNCH_MB(256); break;
        case 128: LAUNCH_MB(128); break;
        case 64: LAUNCH_MB(64); break;
        case 32: LAUNCH_MB(32); break;
        default: LAUNCH_MB(16); break;
        }
#undef LAUNCH_MB
        ktime_finish(h);
    }
    h->prev_lists = chained && next_mask;
    h->seq += 1;
}

// ImaginaryPauliRotation (src/Propagation/specializations.jl:86-133): the
// rotation's kernels with the commuting terms listed, cosh / sinh values,
// and (normalize) every output divided by the identity's coefficient before it
// is truncated. Never chained: the next rotation lists its own terms.
template <int W>
void enq_imag_rotation(pp_handle* h, const uint64_t* mask_in, double ch, double sh, const double* d_trig,
                       bool normalize, int max_weight, double min_abs, uint32_t* glog, hipEvent_t* ev) {
    const Key<W> mask = to_key<W>(mask_in);
    Levels lv;
    mask_levels(mask_in, W, &lv);
    const bool all = h->prof_mode == 2 && ev;
    const bool mrg = h->prof_mode == 2 && ev;
    auto E = [&](int k, bool on) -> hipEvent_t { return on ? ev[k] : nullptr; };
    launch(h, E(0, all), E(1, all), k_branch<W, false, true>,
           lb_grid(h, (const void*)k_branch<W, false, true>, 0, h->cap, K1_TILE), 0, h->bufs, h->d_ctl,
           lb_launch(h, glog), mask);
    Launch L2{};
    L2.seq = h->seq;
    L2.glog = glog;
    const bool part = h->B == 1;
    enq_perm<W, false>(h, E(2, all), part ? nullptr : E(3, all), L2, false, lv, mask, mask, 0u, 0u, sh);
    if (normalize)
        hipLaunchKernelGGL(k_imag_norm<W>, dim3(1), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, mask, h->B,
                           ch, sh, d_trig, h->d_norm);
    const double* norm = normalize ? h->d_norm : nullptr;
    const size_t expect = h->cap + h->cap / 2;
    if (part) {
        RotParams rp{ch, sh, max_weight, min_abs};
        rp.comm = 1;
        rp.norm = norm;
        enq_merge1<W>(h, E(3, all), E(4, mrg), E(5, mrg), 0, mask, rp, mask, glog);
    } else {
        const int B = h->B;
        const int TB = batch_rows(h);
        const size_t shm = (size_t)TB * B * sizeof(double);
        const Launch L3 = kt_launch(h, glog);
#define LAUNCH_MB(TBV)                                                                                         \
    enq_partition_b<W, TBV>(h);                                                                                \
    launch_mb<W, TBV>(h, E(4, mrg), E(5, mrg), lb_grid(h, B <= 32 ? (const void*)k_mergeB<W, TBV, true> : (const void*)k_mergeB<W, TBV>, shm, expect, TBV), \
           shm, h->bufs, h->d_ctl, L3, mask, B, d_trig, max_weight, min_abs, -1, 1.0, 1, norm, part_b(h, TBV), 0, mask)
        switch (TB) {
        case 256: LAUNCH_MB(256); break;
        case 128: LAUNCH_MB(128); break;
        case 64: LAUNCH_MB(64); break;
        case 32: LAUNCH_MB(32); break;
        default: LAUNCH_MB(16); break;
        }
#undef LAUNCH_MB
        ktime_finish(h);
    }
    h->seq += 1;
}

// Amplitude damping (gamma) on `qubit`: the rotation's kernels with the
// damping rule (src/Propagation/specializations.jl:258-276); d_trig holds
// 1 - gamma_b then gamma_b for B > 1 (all columns share fxy = sqrt(1 - gamma)).
template <int W>
void enq_damping(pp_handle* h, int qubit, double gamma, double fxy, const double* d_trig, int max_weight,
                 double min_abs, uint32_t* glog, hipEvent_t* ev) {
    uint64_t m[4] = {0, 0, 0, 0};
    m[qubit >> 5] = 3ull << (2 * (qubit & 31));
    const Key<W> mask = to_key<W>(m);
    Levels lv;
    mask_levels(m, W, &lv);
    const bool all = h->prof_mode == 2 && ev;
    const bool mrg = h->prof_mode == 2 && ev;
    auto E = [&](int k, bool on) -> hipEvent_t { return on ? ev[k] : nullptr; };
    launch(h, E(0, all), E(1, all), k_branch<W, true>,
           lb_grid(h, (const void*)k_branch<W, true>, 0, h->cap, K1_TILE), 0, h->bufs, h->d_ctl,
           lb_launch(h, glog), mask);
    Launch L2{};
    L2.seq = h->seq;
    L2.glog = glog;
    const bool part = h->B == 1;
    enq_perm<W, false>(h, E(2, all), part ? nullptr : E(3, all), L2, false, lv, mask, mask, 0u, 0u, gamma);
    const size_t expect = h->cap + h->cap / 2;
    if (part) {
        RotParams rp{1 - gamma, gamma, max_weight, min_abs, qubit, fxy};
        enq_merge1<W>(h, E(3, all), E(4, mrg), E(5, mrg), 0, mask, rp, mask, glog);
    } else {
        const int B = h->B;
        const int TB = batch_rows(h);
        const size_t shm = (size_t)TB * B * sizeof(double);
        const Launch L3 = kt_launch(h, glog);
#define LAUNCH_MB(TBV)                                                                                         \
    enq_partition_b<W, TBV>(h);                                                                                \
    launch_mb<W, TBV>(h, E(4, mrg), E(5, mrg), lb_grid(h, B <= 32 ? (const void*)k_mergeB<W, TBV, true> : (const void*)k_mergeB<W, TBV>, shm, expect, TBV), \
           shm, h->bufs, h->d_ctl, L3, mask, B, d_trig, max_weight, min_abs, qubit, fxy, 0, nullptr, part_b(h, TBV), 0, mask)
        switch (TB) {
        case 256: LAUNCH_MB(256); break;
        case 128: LAUNCH_MB(128); break;
        case 64: LAUNCH_MB(64); break;
        case 32: LAUNCH_MB(32); break;
        default: LAUNCH_MB(16); break;
        }
#undef LAUNCH_MB
        ktime_finish(h);
    }
    h->seq += 1;
}

template <int W>
void enq_noise(pp_handle* h, int qubit, uint32_t damp_codes, double f1, const double* d_factor, int max_weight,
               double min_abs, uint32_t* glog, hipEvent_t* ev) {
    const bool mrg = h->prof_mode == 2 && ev;
    hipEvent_t e0 = mrg ? ev[4] : nullptr, e1 = mrg ? ev[5] : nullptr;
    if (h->B == 1) {
        launch(h, e0, e1, k_noise1<W>, lb_grid(h, (const void*)k_noise1<W>, 0, h->cap, KNCfg<W>::TILE), 0, h->bufs,
               h->d_ctl, kt_launch(h, glog), qubit, damp_codes, f1, max_weight, min_abs);
    } else {
        // two or four rows side by side in a wave for narrow batches
#define LAUNCH_NB(R)                                                                                          \
    launch(h, e0, e1, k_noiseB<W, R>, lb_grid(h, (const void*)k_noiseB<W, R>, 0, h->cap, KNB_ROWS), 0, h->bufs, \
           h->d_ctl, kt_launch(h, glog), h->B, qubit, damp_codes, d_factor, max_weight, min_abs)
        if (h->B <= 16) LAUNCH_NB(4);
        else if (h->B <= 32) LAUNCH_NB(2);
        else LAUNCH_NB(1);
#undef LAUNCH_NB
    }
    ktime_finish(h);
    h->seq += 1;
}

// A gate of a run, resolved on the host: what enq_* needs.
struct HostGate {
    int kind;              // enum pp_gate_kind
    const uint64_t* mask;
    int qubit;
    uint32_t damp_codes;
    double a, b;           // B = 1: cos, sin (rotation), factor (noise); gamma, sqrt(1 - gamma) (damping)
    int nq = 0, q1 = 0, lut = 0;  // Clifford: qubits (qubit, q1) and table slot
    int flags = 0;                 // imaginary rotation: bit 0 = normalize by the identity's coefficient
};

// Radix-sort buffers for `cap` terms.
int ensure_sort(pp_handle* h) {
    if (h->rs_cap >= h->cap && h->d_rs_hist) return PP_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int b = 0; b < 2; ++b) {
        if (h->d_rs_k[b]) hipFree(h->d_rs_k[b]);
        if (h->d_rs_v[b]) hipFree(h->d_rs_v[b]);
        h->d_rs_k[b] = nullptr;
        h->d_rs_v[b] = nullptr;
    }
    if (h->d_rs_hist) hipFree(h->d_rs_hist);
    h->d_rs_hist = nullptr;
    h->rs_cap = 0;
    const size_t cap = std::max<size_t>(h->cap, 1);
    const size_t tiles = cap / RS_TILE + 2;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMalloc(&h->d_rs_k[b], cap * sizeof(uint64_t)));
        HIP_TRY(hipMalloc(&h->d_rs_v[b], cap * sizeof(uint32_t)));
    }
    HIP_TRY(hipMalloc(&h->d_rs_hist, RS_BINS * tiles * sizeof(uint32_t)));
    h->rs_cap = cap;
    h->rs_tiles = tiles;
    return PP_OK;
}

// One stable LSD sort of (rs_k[c], rs_v[c]) over key bits [0, nbits); returns
// the buffer index holding the result. n is read on the device at n_ptr.
int enq_radix_sort(pp_handle* h, const uint32_t* n_ptr, int c, int nbits) {
    const unsigned grid = grid_for(h, h->cap, RS_TILE);
    const int step = h->tun.radix_bits;
    for (int shift = 0; shift < nbits; shift += step) {
        const int bits = std::min(step, nbits - shift);
        hipLaunchKernelGGL(k_rs_upsweep, dim3(grid), dim3(WG), 0, h->stream, h->d_ctl, n_ptr, h->d_rs_k[c], shift,
                           bits, h->d_rs_hist, (uint32_t)h->rs_tiles);
        hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(WG), 0, h->stream, h->d_ctl, n_ptr, h->d_rs_hist,
                           (uint32_t)h->rs_tiles);
        hipLaunchKernelGGL(k_rs_downsweep, dim3(grid), dim3(WG), 0, h->stream, h->d_ctl, n_ptr, h->d_rs_k[c],
                           h->d_rs_v[c], h->d_rs_k[c ^ 1], h->d_rs_v[c ^ 1], shift, bits, h->d_rs_hist,
                           (uint32_t)h->rs_tiles);
        c ^= 1;
    }
    return c;
}

// Clifford on up to two qubits with table slot `lut` (uploaded to d_lut_*):
// map + truncate + compact, then re-sort inside segments (see pp_sort.h).
template <int W>
void enq_clifford(pp_handle* h, const HostGate& g, int max_weight, double min_abs, uint32_t* glog, hipEvent_t* ev) {
    const bool mrg = h->prof_mode == 2 && ev;
    const int q0 = g.qubit, q1 = g.nq == 2 ? g.q1 : g.qubit;
    const int qmin = std::min(q0, q1), qmax = std::max(q0, q1);
    CliffArgs a{};
    a.change = h->d_lut_c + (size_t)g.lut * 16 * W;
    a.sign = h->d_lut_s + (size_t)g.lut * 16;
    a.nq = g.nq;
    a.q0 = q0;
    a.q1 = q1;
    a.max_weight = max_weight;
    a.min_abs = min_abs;
    a.wlo = 2 * qmin;
    a.wbits = 2 * (qmax - qmin) + 2;
    a.composite = a.wbits <= 32 ? 1 : 0;
    const int p = h->seq & 1;
    const uint32_t* n_ptr = &h->d_ctl->st[p ^ 1].n;
    launch(h, mrg ? ev[4] : nullptr, mrg ? ev[5] : nullptr, k_clifford_map<W>,
           lb_grid(h, (const void*)k_clifford_map<W>, 0, h->cap, KC_TILE), 0, h->bufs, h->d_ctl,
           lb_launch(h, glog), a, h->B, h->d_rs_k[0], h->d_rs_v[0]);
    int c = 0;
    if (a.composite) {
        int segbits = 1;
        while (segbits < 31 && (1ull << segbits) <= h->cap) ++segbits;
        c = enq_radix_sort(h, n_ptr, 0, a.wbits + segbits);
    } else {
        // wide window: the new key's limbs from the window's lowest, least significant first
        const unsigned grid = grid_for(h, h->cap, WG * 4);
        bool first = true;
        for (int l = a.wlo / 64; l < W; ++l) {
            const int lo = l == a.wlo / 64 ? a.wlo % 64 : 0;
            hipLaunchKernelGGL(k_rs_limb_keys, dim3(grid), dim3(WG), 0, h->stream, h->d_ctl, p, h->bufs, 0, 0, n_ptr,
                               l, lo, first ? nullptr : h->d_rs_v[c], h->d_rs_k[c ^ 1], h->d_rs_v[c ^ 1]);
            c ^= 1;
            first = false;
            c = enq_radix_sort(h, n_ptr, c, 64 - lo);
        }
    }
    hipLaunchKernelGGL(k_clifford_gather<W>, dim3(grid_for(h, h->cap, WG * 4)), dim3(WG), 0, h->stream, h->bufs,
                       h->d_ctl, p, h->d_rs_v[c], h->B);
    if (h->recording)
        hipLaunchKernelGGL(k_clifford_rec_fix, dim3(grid_for(h, h->cap, WG * 4)), dim3(WG), 0, h->stream, h->bufs,
                           h->d_ctl, p);
    h->seq += 1;
}


int enqueue(pp_handle* h, const HostGate& g, size_t slot, int max_weight, double min_abs,
            const uint64_t* next_mask) {
    uint32_t* glog = h->d_log + LOG_WORDS * slot;
    hipEvent_t* ev = h->prof_mode ? &h->ev[6 * slot] : nullptr;
    h->kt = (h->prof_mode == 1 && h->d_ktime && slot < h->ktime_cap) ? h->d_ktime + 2 * slot : nullptr;
    const double* dt = h->B > 1 ? h->d_trig + slot * 2 * (size_t)h->B : nullptr;
    if (g.kind != PP_GATE_ROTATION) h->prev_lists = false;
    if (g.kind == PP_GATE_CLIFFORD) {
        if (int rc = ensure_sort(h)) return rc;
        switch (h->W) {
        case 1: enq_clifford<1>(h, g, max_weight, min_abs, glog, ev); break;
        case 2: enq_clifford<2>(h, g, max_weight, min_abs, glog, ev); break;
        default: enq_clifford<4>(h, g, max_weight, min_abs, glog, ev); break;
        }
    } else if (g.kind == PP_GATE_AMP_DAMPING) {
        switch (h->W) {
        case 1: enq_damping<1>(h, g.qubit, g.a, g.b, dt, max_weight, min_abs, glog, ev); break;
        case 2: enq_damping<2>(h, g.qubit, g.a, g.b, dt, max_weight, min_abs, glog, ev); break;
        default: enq_damping<4>(h, g.qubit, g.a, g.b, dt, max_weight, min_abs, glog, ev); break;
        }
    } else if (g.kind == PP_GATE_IMAG_ROTATION) {
        const bool nrm = (g.flags & 1) != 0;
        switch (h->W) {
        case 1: enq_imag_rotation<1>(h, g.mask, g.a, g.b, dt, nrm, max_weight, min_abs, glog, ev); break;
        case 2: enq_imag_rotation<2>(h, g.mask, g.a, g.b, dt, nrm, max_weight, min_abs, glog, ev); break;
        default: enq_imag_rotation<4>(h, g.mask, g.a, g.b, dt, nrm, max_weight, min_abs, glog, ev); break;
        }
    } else if (g.kind == PP_GATE_ROTATION) {
        switch (h->W) {
        case 1: enq_rotation<1>(h, g.mask, g.a, g.b, dt, max_weight, min_abs, glog, ev, next_mask); break;
        case 2: enq_rotation<2>(h, g.mask, g.a, g.b, dt, max_weight, min_abs, glog, ev, next_mask); break;
        default: enq_rotation<4>(h, g.mask, g.a, g.b, dt, max_weight, min_abs, glog, ev, next_mask); break;
        }
    } else {
        switch (h->W) {
        case 1: enq_noise<1>(h, g.qubit, g.damp_codes, g.a, dt, max_weight, min_abs, glog, ev); break;
        case 2: enq_noise<2>(h, g.qubit, g.damp_codes, g.a, dt, max_weight, min_abs, glog, ev); break;
        default: enq_noise<4>(h, g.qubit, g.damp_codes, g.a, dt, max_weight, min_abs, glog, ev); break;
        }
    }
    h->kt = nullptr;
    HIP_TRY(hipGetLastError());
    return PP_OK;
}

// Run resolved gates (trig already staged at d_trig for B > 1), resuming
// after an overflow; then account the gate log into the profile.
// Grow the records to hold `words` words and `rows` coefficient rows, keeping
// the first `keep_words` and `keep_rows` already written.
int grow_records(pp_handle* h, size_t words, size_t rows, size_t keep_words, size_t keep_rows) {
    const size_t B = (size_t)h->B;
    if (words > 0xffffffffull || rows > 0xffffffffull)
        return fail(PP_ERR_CAPACITY, "adjoint records exceed 2^32 words; use checkpoint segments");
    if (words > h->rec_words) {
        uint32_t* nr = nullptr;
        HIP_TRY(hipMalloc(&nr, words * sizeof(uint32_t)));
        if (keep_words)
            HIP_TRY(hipMemcpyAsync(nr, h->d_rec, keep_words * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (h->d_rec) hipFree(h->d_rec);
        h->d_rec = nr;
        h->rec_words = words;
    }
    if (rows > h->rec_rows) {
        double* nc = nullptr;
        HIP_TRY(hipMalloc(&nc, std::max<size_t>(rows, 1) * B * sizeof(double)));
        if (keep_rows)
            HIP_TRY(hipMemcpyAsync(nc, h->d_rcoef, keep_rows * B * sizeof(double), hipMemcpyDeviceToDevice,
                                   h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (h->d_rcoef) hipFree(h->d_rcoef);
        h->d_rcoef = nc;
        h->rec_rows = rows;
    }
    if (h->recording) {
        h->bufs.rec = h->d_rec;
        h->bufs.rcoef = h->d_rcoef;
        h->bufs.rec_words = (uint32_t)h->rec_words;
        h->bufs.rec_rows = (uint32_t)h->rec_rows;
    }
    return PP_OK;
}

// Which Paulis may occur on each qubit: X, Y and Z bit planes in the key's
// layout (bit 2q of limb q / 32), exact when read from the device and widened
// on the host gate by gate, so that they stay a superset until the next read.
struct Presence {
    uint64_t x[4], y[4], z[4];
};

// The X, Y and Z positions of a mask's qubits.
void mask_planes(const uint64_t* m, int l, uint64_t* mx, uint64_t* my, uint64_t* mz) {
    const uint64_t lo = m[l] & 0x5555555555555555ull, hi = (m[l] >> 1) & 0x5555555555555555ull;
    *mx = lo & ~hi;
    *my = hi & ~lo;
    *mz = lo & hi;
}

// Whether some term may anticommute with the mask: only if some qubit of the
// mask may hold one of the two Paulis that anticommute with the mask's there.
// (A term anticommutes when an odd number of its qubits do; none means none.)
bool may_touch(const Presence& P, const uint64_t* m, int W) {
    uint64_t any = 0;
    for (int l = 0; l < W; ++l) {
        uint64_t mx, my, mz;
        mask_planes(m, l, &mx, &my, &mz);
        any |= ((P.y[l] | P.z[l]) & mx) | ((P.x[l] | P.z[l]) & my) | ((P.x[l] | P.y[l]) & mz);
    }
    return any != 0;
}

// After a rotation that may touch terms: a created term holds t * m on every
// qubit of the mask (t the term's Pauli there, I included), so m and every
// product with a present Pauli may occur.
void widen_rotation(Presence& P, const uint64_t* m, int W) {
    for (int l = 0; l < W; ++l) {
        uint64_t mx, my, mz;
        mask_planes(m, l, &mx, &my, &mz);
        const uint64_t x = P.x[l], y = P.y[l], z = P.z[l];
        P.x[l] = x | mx | (z & my) | (y & mz);  // Z.Y ~ X, Y.Z ~ X
        P.y[l] = y | my | (z & mx) | (x & mz);  // Z.X ~ Y, X.Z ~ Y
        P.z[l] = z | mz | (y & mx) | (x & my);  // Y.X ~ Z, X.Y ~ Z
    }
}

// After a Clifford: anything may occur on its qubits.
void widen_qubit(Presence& P, int q) {
    const uint64_t b = 1ull << (2 * (q & 31));
    P.x[q >> 5] |= b;
    P.y[q >> 5] |= b;
    P.z[q >> 5] |= b;
}

// Enqueue the presence of the sum's current state into h_pres (valid after
// the next stream synchronization).
int enq_presence(pp_handle* h) {
    const unsigned grid = grid_for(h, h->cap, (size_t)WG * 16);
    uint64_t* res = h->d_pres_part + (size_t)std::max(64, MAX_GRID_PER_CU * h->cus) * PRES_WORDS;
    switch (h->W) {
    case 1:
        hipLaunchKernelGGL(k_presence<1>, dim3(grid), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq,
                           h->d_pres_part);
        hipLaunchKernelGGL(k_presence_final<1>, dim3(1), dim3(WG), 0, h->stream, h->d_pres_part, (int)grid, res);
        break;
    case 2:
        hipLaunchKernelGGL(k_presence<2>, dim3(grid), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq,
                           h->d_pres_part);
        hipLaunchKernelGGL(k_presence_final<2>, dim3(1), dim3(WG), 0, h->stream, h->d_pres_part, (int)grid, res);
        break;
    default:
        hipLaunchKernelGGL(k_presence<4>, dim3(grid), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq,
                           h->d_pres_part);
        hipLaunchKernelGGL(k_presence_final<4>, dim3(1), dim3(WG), 0, h->stream, h->d_pres_part, (int)grid, res);
        break;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h->h_pres, res, 3 * h->W * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream));
    return PP_OK;
}

Presence read_presence(const pp_handle* h) {
    Presence P{};
    for (int l = 0; l < h->W; ++l) {
        P.x[l] = h->h_pres[l];
        P.y[l] = h->h_pres[h->W + l];
        P.z[l] = h->h_pres[2 * h->W + l];
    }
    return P;
}

// ---------------------------------------------------------------- chunks as graphs
//
// With PP_TUNE_GRAPHS a chunk of rotations and Pauli noise gates is captured
// in a hipGraph the first time it is enqueued and replayed when the same chunk
// comes again (same gates, angles, truncation, state parity, buffers and
// launch shapes: a repeated propagation). The look-back words carry the epochs
// of the capture, so the graph starts by clearing them (a fill kernel node);
// the host's epoch counter has moved past them, so no later launch reads them
// as its own.

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
    template <typename T>
    void val(const T& v) { add(&v, sizeof(T)); }
};

bool graph_ok(const pp_handle* h, const std::vector<HostGate>& gs, const std::vector<size_t>& order) {
    if (!h->use_graphs || h->prof_mode != 0 || h->recording || order.empty()) return false;
    if (h->force_dynamic || h->seq < h->safe_until) return false;
    if ((size_t)h->epoch + 8 * order.size() + 64 >= (size_t)LB_EPOCH_MASK) return false;  // no wrap inside
    for (size_t i : order)
        if (gs[i].kind != PP_GATE_ROTATION && gs[i].kind != PP_GATE_PAULI_NOISE) return false;
    return true;
}

int enqueue(pp_handle* h, const HostGate& g, size_t slot, int max_weight, double min_abs,
            const uint64_t* next_mask);

// The chunk's gates (order; chaining as in run_resolved), from the cache or
// captured now; either way launched on the handle's stream.
int enq_chunk_graph(pp_handle* h, const std::vector<HostGate>& gs, const std::vector<size_t>& order, size_t ng,
                    int max_weight, double min_abs) {
    Fnv f;
    f.val(h->seq);  // baked into the launches (state parity, list and sample tags, overflow reports)
    f.val(h->lb_count & 1u);
    f.val(h->cap);
    f.val(h->W);
    f.val(h->B);
    f.val(max_weight);
    f.val(min_abs);
    f.val(h->tun);
    f.val(h->chain_tails);
    f.val(h->spin_limit);
    f.val(h->force_rerun);
    for (size_t k = 0; k < order.size(); ++k) {
        const size_t i = order[k];
        const HostGate& q = gs[i];
        f.val(i);
        f.val(q.kind);
        if (q.kind == PP_GATE_ROTATION) f.add(q.mask, sizeof(uint64_t) * h->W);
        f.val(q.qubit);
        f.val(q.damp_codes);
        f.val(q.a);
        f.val(q.b);
        f.val(q.flags);
    }
    const size_t nl = order.back() + 1;
    f.val(nl < ng ? gs[nl].kind : -1);  // the gate after the chunk (chaining of the last one)
    if (nl < ng && gs[nl].kind == PP_GATE_ROTATION) f.add(gs[nl].mask, sizeof(uint64_t) * h->W);
    const uint64_t key = f.h;
    auto it = h->graphs.find(key);
    if (it != h->graphs.end()) {
        HIP_TRY(hipGraphLaunch(it->second.exec, h->stream));
        h->seq += it->second.seq_delta;
        h->lb_count += it->second.lb_delta;
        h->prev_lists = it->second.prev_lists;
        h->graph_launches += 1;
        return PP_OK;
    }
    const uint32_t seq0 = h->seq, lb0 = h->lb_count;
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = PP_OK;
    const size_t nw = 3 * (size_t)h->bufs.lb_tiles, ngw = 2 * (size_t)h->bufs.lb_groups;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)std::min<size_t>((nw + WG - 1) / WG, 1024)), dim3(WG), 0, h->stream,
                       h->bufs.states, nw, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)std::min<size_t>((ngw + WG - 1) / WG, 1024)), dim3(WG), 0,
                       h->stream, h->bufs.gstates, ngw, 0ull);
    for (size_t k = 0; k < order.size() && rc == PP_OK; ++k) {
        const size_t i = order[k];
        const size_t nx = k + 1 < order.size() ? order[k + 1] : i + 1;
        const bool chain = chain_ok(h) && nx < ng && gs[nx].kind == PP_GATE_ROTATION;
        rc = enqueue(h, gs[i], i, max_weight, min_abs, chain ? gs[nx].mask : nullptr);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
    if (rc != PP_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess || !graph) return fail(PP_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(PP_ERR_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
    if (h->graphs.size() >= 4096) drop_graphs(h);  // bounded cache
    pp_handle::ChunkGraph cg;
    cg.exec = exec;
    cg.seq_delta = h->seq - seq0;
    cg.lb_delta = h->lb_count - lb0;
    cg.prev_lists = h->prev_lists;
    h->graphs[key] = cg;
    h->graph_captures += 1;
    HIP_TRY(hipGraphLaunch(exec, h->stream));
    h->graph_launches += 1;
    return PP_OK;
}

// Terms a shared-memory run may hold on this handle.
uint32_t lds_limit(const pp_handle* h) {
    const size_t nl = h->W == 1 ? LdsCfg<1>::NL : (h->W == 2 ? LdsCfg<2>::NL : LdsCfg<4>::NL);
    return (uint32_t)std::min(nl, h->cap);
}

// Whether gate i may go to a shared-memory run: a rotation or Pauli noise on
// one coefficient column, not recorded.
bool lds_gate(const pp_handle* h, const HostGate& q) {
    return h->lds_runs && h->B == 1 && !h->recording &&
           (q.kind == PP_GATE_ROTATION || q.kind == PP_GATE_PAULI_NOISE);
}

// Gates sel[0 .. cnt) (all lds_gate, gate indices = log slots) in one
// k_lds_run launch; *done receives how many it applied before the sum outgrew
// shared memory. Reads the state back.
int lds_run(pp_handle* h, const std::vector<HostGate>& gs, const std::vector<size_t>& sel, int max_weight,
            double min_abs, size_t* done) {
    *done = 0;
    const size_t cnt = sel.size();
    if (cnt > h->ldsg_cap) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (h->d_ldsg) hipFree(h->d_ldsg);
        if (h->h_ldsg) hipHostFree(h->h_ldsg);
        h->d_ldsg = nullptr;
        h->h_ldsg = nullptr;
        h->ldsg_cap = 0;
        const size_t c = std::max<size_t>(cnt, 1024);
        HIP_TRY(hipMalloc(&h->d_ldsg, c * sizeof(LdsGate)));
        HIP_TRY(hipHostMalloc(&h->h_ldsg, c * sizeof(LdsGate), hipHostMallocDefault));
        h->ldsg_cap = c;
    }
    for (size_t k = 0; k < cnt; ++k) {
        const HostGate& q = gs[sel[k]];
        LdsGate d{};
        d.slot = (uint32_t)sel[k];
        if (q.kind == PP_GATE_ROTATION) {
            d.kind = 0;
            uint64_t m4[4] = {0, 0, 0, 0};
            for (int l = 0; l < h->W; ++l) m4[l] = q.mask[l];
            for (int l = 0; l < 4; ++l) d.mask[l] = m4[l];
            d.a = q.a;
            d.b = q.b;
            Levels lv;
            if (mask_levels(m4, h->W, &lv)) return fail(PP_ERR_UNSUPPORTED, "rotation generator with too many runs");
            d.nlev = lv.n;
            d.top_runs = lv.top_runs;
            for (int v = 0; v < lv.n; ++v) {
                d.bit[v] = (int16_t)lv.bit[v];
                d.sgn[v] = (int16_t)lv.sgn[v];
            }
        } else {
            d.kind = 1;
            d.qubit = q.qubit;
            d.damp_codes = q.damp_codes;
            d.a = q.a;
        }
        h->h_ldsg[k] = d;
    }
    HIP_TRY(hipMemcpyAsync(h->d_ldsg, h->h_ldsg, cnt * sizeof(LdsGate), hipMemcpyHostToDevice, h->stream));
    const uint32_t lim = lds_limit(h);
    uint32_t* glog = h->d_log;
    switch (h->W) {
    case 1:
        hipLaunchKernelGGL(k_lds_run<1>, dim3(1), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, h->d_ldsg,
                           (int)cnt, lim, max_weight, min_abs, glog);
        break;
    case 2:
        hipLaunchKernelGGL(k_lds_run<2>, dim3(1), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, h->d_ldsg,
                           (int)cnt, lim, max_weight, min_abs, glog);
        break;
    default:
        hipLaunchKernelGGL(k_lds_run<4>, dim3(1), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, h->d_ldsg,
                           (int)cnt, lim, max_weight, min_abs, glog);
        break;
    }
    HIP_TRY(hipGetLastError());
    if (int rc = sync_ctl(h)) return rc;
    if (h->h_ctl->err)
        return fail(PP_ERR_STATE, "internal consistency check failed in the shared-memory run (bits " +
                                      std::to_string(h->h_ctl->err) + ")");
    const size_t d = h->h_ctl->lds_done;
    if (d > cnt) return fail(PP_ERR_STATE, "shared-memory run reports more gates than it was given");
    h->seq += (uint32_t)d;
    h->prev_lists = false;
    *done = d;
    return PP_OK;
}

// Run resolved gates first .. last-1 (trig staged at d_trig for B > 1, slots
// are gate indices), resuming after an overflow; then account the gate log
// into the profile. counts[i - first] receives gate i's input size.
//
// Rotations that touch no term are the identity (xortailmerge.jl:121-123:
// nothing branches, nothing is truncated). The host reads which Paulis occur
// on each qubit at every synchronization and leaves out the rotations of the
// next chunk that this presence, widened gate by gate, shows to touch nothing;
// their log entries are filled in afterwards (no touched terms, so no records:
// the adjoint sweep passes them as it passes an identity that ran).
int run_resolved(pp_handle* h, const std::vector<HostGate>& gs, size_t first, size_t last, int max_weight,
                 double min_abs, int64_t* counts) {
    const size_t ng = last;
    size_t g = first;
    const bool skip = h->skip_identity && last - first > 1;
    std::vector<uint8_t> skipped(last - first, 0);
    std::vector<size_t> order;  // gates of the chunk that are enqueued
    if (h->prof_mode == 1 && h->d_ktime && last > first && last <= h->ktime_cap) {
        const size_t nk = last - first;
        hipLaunchKernelGGL(k_ktime_init, dim3((unsigned)((nk + WG - 1) / WG)), dim3(WG), 0, h->stream,
                           h->d_ktime + 2 * first, nk);
    }
    order.reserve(h->tun.run_chunk);
    std::vector<uint8_t> in_lds(last - first, 0);  // applied by a shared-memory run
    bool pres_ok = false;
    // h_ctl holds the current state when fresh (after every synchronization
    // that no host-side repair followed)
    bool fresh = false;
    if (skip && first < last) {
        if (int rc = enq_presence(h)) return rc;
        if (int rc = sync_ctl(h)) return rc;
        pres_ok = h->h_ctl->overflow == 0;
        fresh = true;
    } else if (last - first > 1 && lds_gate(h, gs[first])) {
        if (int rc = sync_ctl(h)) return rc;
        fresh = true;
    }
    while (g < ng) {
        // a small sum: the rotations and noise from here run in shared memory,
        // one launch for as many gates as the sum stays small (with the
        // presence, rotations that touch nothing are left out as below)
        if (fresh && h->h_ctl->overflow == 0 && lds_gate(h, gs[g]) &&
            h->h_ctl->st[h->seq & 1].n <= lds_limit(h)) {
            Presence P{};
            if (pres_ok) P = read_presence(h);
            order.clear();
            size_t e = g;
            while (e < ng && e - g < (1u << 16) && lds_gate(h, gs[e])) {
                const HostGate& q = gs[e];
                if (pres_ok && q.kind == PP_GATE_ROTATION) {
                    if (!may_touch(P, q.mask, h->W)) {
                        skipped[e - first] = 1;
                        ++e;
                        continue;
                    }
                    widen_rotation(P, q.mask, h->W);
                }
                order.push_back(e);
                ++e;
            }
            size_t done = 0;
            if (!order.empty())
                if (int rc = lds_run(h, gs, order, max_weight, min_abs, &done)) return rc;
            if (done > 0 || order.empty()) {
                for (size_t k = 0; k < done; ++k) in_lds[order[k] - first] = 1;
                // the rest, from the gate the run stopped at, goes to the pipeline
                const size_t stop = done < order.size() ? order[done] : e;
                for (size_t i = stop; i < e; ++i) skipped[i - first] = 0;
                g = stop;
                pres_ok = false;
                if (skip && g < ng) {
                    if (int rc = enq_presence(h)) return rc;
                    if (int rc = sync_ctl(h)) return rc;
                    pres_ok = h->h_ctl->overflow == 0;
                }
                continue;
            }
            for (size_t i = g; i < e; ++i) skipped[i - first] = 0;
        }
        const size_t end = std::min(ng, g + (size_t)h->tun.run_chunk);
        const uint32_t seq0 = h->seq;
        order.clear();
        if (pres_ok) {
            Presence P = read_presence(h);
            for (size_t i = g; i < end; ++i) {
                const HostGate& q = gs[i];
                if (q.kind == PP_GATE_ROTATION) {
                    if (!may_touch(P, q.mask, h->W)) {
                        skipped[i - first] = 1;
                        continue;
                    }
                    widen_rotation(P, q.mask, h->W);
                } else if (q.kind == PP_GATE_IMAG_ROTATION) {
                    widen_rotation(P, q.mask, h->W);
                } else if (q.kind == PP_GATE_CLIFFORD) {
                    widen_qubit(P, q.qubit);
                    if (q.nq == 2) widen_qubit(P, q.q1);
                }
                order.push_back(i);
            }
        } else {
            for (size_t i = g; i < end; ++i) order.push_back(i);
        }
        // a chunk starts a new chain: the last gate of the previous one listed
        // the terms of the next gate of the circuit, which may have been left out
        h->prev_lists = false;
        if (graph_ok(h, gs, order)) {
            if (int rc = enq_chunk_graph(h, gs, order, ng, max_weight, min_abs)) return rc;
        } else {
            for (size_t k = 0; k < order.size(); ++k) {
                const size_t i = order[k];
                const size_t nx = k + 1 < order.size() ? order[k + 1] : i + 1;
                const bool chain = chain_ok(h) && nx < ng && gs[nx].kind == PP_GATE_ROTATION;
                if (int rc = enqueue(h, gs[i], i, max_weight, min_abs, chain ? gs[nx].mask : nullptr)) return rc;
            }
        }
        if (skip && end < ng)
            if (int rc = enq_presence(h)) return rc;
        if (int rc = sync_ctl(h)) return rc;
        if (h->h_ctl->err)
            return fail(PP_ERR_STATE, "internal consistency check failed in the rotation kernels (bits " +
                                          std::to_string(h->h_ctl->err) + ")");
        pres_ok = skip && h->h_ctl->overflow == 0;
        fresh = h->h_ctl->overflow == 0;
        if (h->h_ctl->overflow) {
            const uint32_t fseq = h->h_ctl->overflow_seq;
            const size_t gi = order[fseq - seq0];
            for (size_t i = gi; i < end; ++i) skipped[i - first] = 0;
            h->seq = fseq;  // the state before the stopped gate is still in st[fseq & 1]
            if (h->h_ctl->overflow == 1) {
                const size_t need = h->h_ctl->overflow_need;
                if (need > MAX_TERMS) return fail(PP_ERR_CAPACITY, "term count exceeds 2^31 - 1");
                if (int rc = grow(h, std::max(need + need / 4, h->cap + h->cap / 2))) return rc;
            } else if (h->h_ctl->overflow == 3) {
                // the gate's records did not fit: grow them and run it again
                const GateState st = h->h_ctl->st[fseq & 1];
                const size_t words = std::max<size_t>((size_t)h->h_ctl->overflow_need * 3 / 2, 2 * h->rec_words);
                const size_t rows = std::max<size_t>((size_t)h->h_ctl->overflow_rows * 3 / 2, 2 * h->rec_rows);
                if (int rc = grow_records(h, std::min<size_t>(words, 0xffffffffull),
                                          std::min<size_t>(rows, 0xffffffffull), st.roff, st.coff))
                    return rc;
                if (int rc = reset_ctl(h, st)) return rc;
            } else {
                // a look-back wait ran out: run that gate again with dynamic tiles
                const GateState st = h->h_ctl->st[fseq & 1];
                if (int rc = reset_ctl(h, st)) return rc;
                h->safe_until = fseq + 1;
                h->fallbacks += 1;
            }
            g = gi;
            continue;
        }
        g = end;
    }
    if (first >= last) return PP_OK;
    HIP_TRY(hipMemcpyAsync(h->h_log + LOG_WORDS * first, h->d_log + LOG_WORDS * first,
                           LOG_WORDS * (last - first) * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    const bool ktimed = h->prof_mode == 1 && h->d_ktime && last <= h->ktime_cap;
    if (ktimed)
        HIP_TRY(hipMemcpyAsync(h->h_ktime.data() + 2 * first, h->d_ktime + 2 * first,
                               2 * (last - first) * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // a gate left out passes the sum on: its size is the next run gate's input
    // size, or the final size
    uint32_t n_after = h->h_ctl->st[h->seq & 1].n;
    for (size_t i = last; i-- > first;) {
        uint32_t* lg = h->h_log + LOG_WORDS * i;
        if (skipped[i - first]) {
            lg[0] = n_after;
            lg[1] = 0;
            lg[2] = n_after;
            lg[3] = lg[4] = 0;
        } else {
            n_after = lg[0];
        }
    }
    const double row = 8.0 * h->W + 8.0 * h->B;
    for (size_t i = first; i < last; ++i) {
        const uint32_t nin = h->h_log[LOG_WORDS * i], na = h->h_log[LOG_WORDS * i + 1],
                       nout = h->h_log[LOG_WORDS * i + 2];
        if (counts) counts[i - first] = nin;
        h->gates += 1;
        h->term_apps += nin;
        if (skipped[i - first]) {
            h->identity_gates += 1;
            continue;
        }
        if (in_lds[i - first]) {
            h->lds_gates += 1;
            continue;
        }
        hipEvent_t* ev = h->prof_mode ? &h->ev[6 * i] : nullptr;
        auto span = [&](int a) -> double {
            if (a == 4 && ktimed) {
                // the gate's merge or noise kernel on the device clock (0 when it did not run)
                const uint64_t t0 = h->h_ktime[2 * i], t1 = h->h_ktime[2 * i + 1];
                return t1 > t0 ? (double)(t1 - t0) / h->clock_khz : 0.0;
            }
            float ms = 0.f;
            if (ev && hipEventElapsedTime(&ms, ev[a], ev[a + 1]) == hipSuccess) return ms;
            return 0.0;
        };
        if (gs[i].kind == PP_GATE_ROTATION || gs[i].kind == PP_GATE_AMP_DAMPING) {
            uint64_t m[4] = {0, 0, 0, 0};
            if (gs[i].kind == PP_GATE_ROTATION)
                for (int l = 0; l < h->W; ++l) m[l] = gs[i].mask[l];
            else
                m[gs[i].qubit >> 5] = 3ull << (2 * (gs[i].qubit & 31));
            int mask_limbs = 0;
            for (int l = 0; l < h->W; ++l) mask_limbs += m[l] != 0;
            // every rotation launches all its kernels (those of a gate that touches
            // nothing exit at once), so launches count dispatches as a profiler does
            h->launches[PP_KERNEL_BRANCH] += 1;
            h->alg_bytes[PP_KERNEL_BRANCH] += 8.0 * mask_limbs * nin + 4.0 * na;
            if (h->prof_mode == 2) h->ms[PP_KERNEL_BRANCH] += span(0);
            h->launches[PP_KERNEL_TAIL_PERMUTE] += 1;
            h->launches[PP_KERNEL_MERGE] += 1;
            if (h->prof_mode == 2) h->ms[PP_KERNEL_TAIL_PERMUTE] += span(2);
            if (h->prof_mode >= 1) h->ms[PP_KERNEL_MERGE] += span(4);
            if (na > 0) {
                Levels lv;
                mask_levels(m, h->W, &lv);
                h->alg_bytes[PP_KERNEL_TAIL_PERMUTE] += na * (8.0 + 8.0 * (h->W - lv.from_limb));
                h->alg_bytes[PP_KERNEL_MERGE] += row * ((double)nin + (double)nout);
            }
        } else {
            h->launches[PP_KERNEL_NOISE] += 1;
            h->alg_bytes[PP_KERNEL_NOISE] += row * ((double)nin + (double)nout);
            if (h->prof_mode >= 1) h->ms[PP_KERNEL_NOISE] += span(4);
        }
    }
    return PP_OK;
}

int ensure_partial(pp_handle* h, size_t need) {
    if (need <= h->partial_cap) return PP_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->d_partial) hipFree(h->d_partial);
    h->d_partial = nullptr;
    h->partial_cap = 0;
    HIP_TRY(hipMalloc(&h->d_partial, need * sizeof(double)));
    h->partial_cap = need;
    return PP_OK;
}

// Double-double partials of `rd`'s readout in `partial` (max_blocks x B x 2
// doubles; nullptr: the handle's own, grown as needed), finished into B doubles
// at d_res on the device and copied to out.
template <int W>
int overlap_impl(pp_handle* h, double* out, const Readout& rd = Readout{}, double* partial = nullptr,
                 double* d_res = nullptr, unsigned max_blocks = 2048u) {
    const int B = h->B;
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    if (n == 0) {
        for (int b = 0; b < B; ++b) out[b] = 0.0;
        return PP_OK;
    }
    const unsigned per = B == 1 ? WG : NWAVES;
    const unsigned g = std::min<unsigned>(max_blocks, (n + per - 1) / per);
    if (!partial) {
        if (int rc = ensure_partial(h, (size_t)g * B * 2)) return rc;
        partial = h->d_partial;
        d_res = h->d_out;
    }
    if (B == 1) {
        hipLaunchKernelGGL(k_overlap1<W>, dim3(g), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, rd, partial);
    } else {
        hipLaunchKernelGGL(k_overlapB<W>, dim3(g), dim3(WG), (size_t)NWAVES * B * 2 * sizeof(double), h->stream,
                           h->bufs, h->d_ctl, h->seq, rd, B, partial);
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_overlap_final, dim3(1), dim3(WG), 0, h->stream, partial, (int)g, B, d_res);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d_res, B * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->launches[PP_KERNEL_OVERLAP] += 1;
    h->alg_bytes[PP_KERNEL_OVERLAP] += (8.0 * h->W + 8.0 * B) * (double)n;
    return PP_OK;
}

// Per-column largest |c| into out (B doubles).
int maxabs_impl(pp_handle* h, double* out) {
    const int B = h->B;
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    for (int b = 0; b < B; ++b) out[b] = 0.0;
    if (n == 0) return PP_OK;
    const unsigned g = std::min<unsigned>(1024u, (n + NWAVES - 1) / NWAVES);
    if (int rc = ensure_partial(h, (size_t)g * B)) return rc;
    hipLaunchKernelGGL(k_maxabs, dim3(g), dim3(WG), (size_t)NWAVES * B * sizeof(double), h->stream, h->bufs,
                       h->d_ctl, h->seq, B, h->d_partial);
    HIP_TRY(hipGetLastError());
    std::vector<double> part((size_t)g * B);
    HIP_TRY(hipMemcpyAsync(part.data(), h->d_partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (unsigned k = 0; k < g; ++k)
        for (int b = 0; b < B; ++b) out[b] = std::fmax(out[b], part[(size_t)k * B + b]);
    return PP_OK;
}

template <int W>
int getcoeff_impl(pp_handle* h, const uint64_t* key, double* out) {
    if (int rc = sync_ctl(h)) return rc;
    hipLaunchKernelGGL(k_getcoeff<W>, dim3(1), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, to_key<W>(key),
                       h->B, h->d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, h->d_out, h->B * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PP_OK;
}

// Terms of a are looked up in b; a's stream runs the kernels after b's is idle.
template <int W>
int scalar_impl(pp_handle* a, pp_handle* b, double* out) {
    const int B = a->B;
    if (int rc = sync_ctl(b)) return rc;
    uint32_t n = 0, cur = 0;
    if (int rc = live(a, &n, &cur)) return rc;
    if (n == 0) {
        for (int c = 0; c < B; ++c) out[c] = 0.0;
        return PP_OK;
    }
    const unsigned g = std::min<unsigned>(2048u, (n + NWAVES - 1) / NWAVES);
    if (int rc = ensure_partial(a, (size_t)g * B * 2)) return rc;
    hipLaunchKernelGGL(k_scalar<W>, dim3(g), dim3(WG), (size_t)NWAVES * B * 2 * sizeof(double), a->stream, a->bufs,
                       a->d_ctl, a->seq, b->bufs, b->d_ctl, b->seq, B, a->d_partial);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_overlap_final, dim3(1), dim3(WG), 0, a->stream, a->d_partial, (int)g, B, a->d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, a->d_out, B * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    return PP_OK;
}

int key_cmp(const uint64_t* a, const uint64_t* b, int W) {
    for (int l = W - 1; l >= 0; --l)
        if (a[l] != b[l]) return a[l] < b[l] ? -1 : 1;
    return 0;
}

// one gate with host values, synchronously, with its statistics
int single_gate(pp_handle* h, const HostGate& g, const double* hv1, const double* hv2, int max_weight,
                double min_abs, pp_gate_stats* st) {
    if (int rc = ensure_run(h, 1)) return rc;
    if (h->B > 1 && hv1) {  // per-column values (a Clifford has none)
        std::memcpy(h->h_trig, hv1, h->B * sizeof(double));
        if (hv2) std::memcpy(h->h_trig + h->B, hv2, h->B * sizeof(double));
        HIP_TRY(hipMemcpyAsync(h->d_trig, h->h_trig, 2 * h->B * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    std::vector<HostGate> gs(1, g);
    if (int rc = run_resolved(h, gs, 0, 1, max_weight, min_abs, nullptr)) return rc;
    if (st) {
        st->n_in = h->h_log[0];
        st->n_touched = h->h_log[1];
        st->n_out = h->h_log[2];
    }
    return PP_OK;
}

// Validate and resolve gates for a run: cos/sin or noise factors, staged at
// d_trig for B > 1. With atrig, every gate's B cos then B sin (rotation) or B
// factors (noise) also go to atrig (2B doubles per gate).
int resolve_gates(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
                  const uint64_t* lut_changes, const double* lut_signs, size_t nluts, std::vector<HostGate>& gs,
                  std::vector<double>* atrig) {
    if (ngates > 0 && !gates) return fail(PP_ERR_ARG, "gates is NULL");
    if (nluts > 0 && (!lut_changes || !lut_signs)) return fail(PP_ERR_ARG, "Clifford tables are NULL");
    const int B = h->B;
    gs.assign(ngates, HostGate{});
    if (atrig) atrig->assign(ngates * 2 * (size_t)B, 0.0);
    if (int rc = ensure_run(h, ngates)) return rc;
    for (size_t g = 0; g < ngates; ++g) {
        const pp_gate& gt = gates[g];
        if (gt.param >= 0 && ((size_t)gt.param >= nparams || !theta))
            return fail(PP_ERR_ARG, "gate " + std::to_string(g) + " reads a parameter beyond nparams");
        auto val = [&](int b) { return gt.param < 0 ? gt.value : theta[(size_t)gt.param * B + b]; };
        HostGate& hg = gs[g];
        hg.kind = gt.kind;
        if (gt.kind == PP_GATE_ROTATION) {
            if (int rc = check_mask(h, gt.mask)) return rc;
            hg.mask = gt.mask;
            // cos and sin as two separate libm calls, as the reference evaluates them
            // (src/Propagation/specializations.jl:30)
            hg.a = std::cos(val(0));
            hg.b = std::sin(val(0));
            for (int b = 0; b < B; ++b) {
                const double c = std::cos(val(b)), sn = std::sin(val(b));
                if (B > 1) {
                    h->h_trig[g * 2 * B + b] = c;
                    h->h_trig[g * 2 * B + B + b] = sn;
                }
                if (atrig) {
                    (*atrig)[g * 2 * B + b] = c;
                    (*atrig)[g * 2 * B + B + b] = sn;
                }
            }
        } else if (gt.kind == PP_GATE_IMAG_ROTATION) {
            if (int rc = check_mask(h, gt.mask)) return rc;
            if (atrig) return fail(PP_ERR_UNSUPPORTED, "the adjoint sweep does not take imaginary rotations");
            hg.mask = gt.mask;
            hg.flags = gt.flags;
            hg.a = std::cosh(val(0));
            hg.b = std::sinh(val(0));
            for (int b = 0; b < B && B > 1; ++b) {
                h->h_trig[g * 2 * B + b] = std::cosh(val(b));
                h->h_trig[g * 2 * B + B + b] = std::sinh(val(b));
            }
        } else if (gt.kind == PP_GATE_PAULI_NOISE) {
            if (gt.qubit < 0 || gt.qubit >= h->nqubits) return fail(PP_ERR_ARG, "noise qubit out of range");
            if (gt.damp_codes & ~0xeu) return fail(PP_ERR_ARG, "damp_codes may only hold X, Y, Z (bits 1..3)");
            hg.qubit = gt.qubit;
            hg.damp_codes = gt.damp_codes;
            for (int b = 0; b < B; ++b) {
                const double lam = val(b);
                if (!(lam >= 0.0 && lam <= 1.0)) return fail(PP_ERR_ARG, "noise strength must be in [0, 1]");
                if (B > 1) h->h_trig[g * 2 * B + b] = 1 - lam;
                if (atrig) (*atrig)[g * 2 * B + b] = 1 - lam;
            }
            hg.a = 1 - val(0);
        } else if (gt.kind == PP_GATE_AMP_DAMPING) {
            if (gt.qubit < 0 || gt.qubit >= h->nqubits) return fail(PP_ERR_ARG, "damping qubit out of range");
            const double g0 = val(0);
            for (int b = 0; b < B; ++b) {
                const double gam = val(b);
                if (!(gam >= 0.0 && gam <= 1.0)) return fail(PP_ERR_ARG, "noise strength must be in [0, 1]");
                if (gam != g0)
                    return fail(PP_ERR_UNSUPPORTED, "amplitude damping with a different gamma per batch column");
                if (B > 1) {
                    h->h_trig[g * 2 * B + b] = 1 - gam;
                    h->h_trig[g * 2 * B + B + b] = gam;
                }
            }
            hg.qubit = gt.qubit;
            hg.a = g0;
            hg.b = std::sqrt(1 - g0);
            if (atrig) return fail(PP_ERR_UNSUPPORTED, "the adjoint sweep does not take amplitude damping yet");
        } else if (gt.kind == PP_GATE_CLIFFORD) {
            if (gt.nq < 1 || gt.nq > 2) return fail(PP_ERR_ARG, "a Clifford gate acts on 1 or 2 qubits");
            for (int a = 0; a < gt.nq; ++a)
                if (gt.qubits[a] < 0 || gt.qubits[a] >= h->nqubits)
                    return fail(PP_ERR_ARG, "Clifford qubit out of range");
            if (gt.nq == 2 && gt.qubits[0] == gt.qubits[1]) return fail(PP_ERR_ARG, "Clifford qubits must differ");
            if (gt.lut < 0 || (size_t)gt.lut >= nluts) return fail(PP_ERR_ARG, "Clifford table slot out of range");
            hg.qubit = gt.qubits[0];
            hg.q1 = gt.nq == 2 ? gt.qubits[1] : gt.qubits[0];
            hg.nq = gt.nq;
            hg.lut = gt.lut;
            // the sweep multiplies the adjoint of a term whose sign flips by -1
            if (atrig)
                for (int b = 0; b < B; ++b) (*atrig)[g * 2 * B + b] = -1.0;
        } else {
            return fail(PP_ERR_UNSUPPORTED, "gate kind " + std::to_string(gt.kind) + " is not implemented");
        }
    }
    if (B > 1 && ngates > 0)
        HIP_TRY(hipMemcpyAsync(h->d_trig, h->h_trig, ngates * 2 * B * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
    if (nluts > 0) {
        if (nluts > h->lut_cap) {
            HIP_TRY(hipStreamSynchronize(h->stream));
            if (h->d_lut_c) hipFree(h->d_lut_c);
            if (h->d_lut_s) hipFree(h->d_lut_s);
            h->d_lut_c = nullptr;
            h->d_lut_s = nullptr;
            h->lut_cap = 0;
            HIP_TRY(hipMalloc(&h->d_lut_c, nluts * 16 * h->W * sizeof(uint64_t)));
            HIP_TRY(hipMalloc(&h->d_lut_s, nluts * 16 * sizeof(double)));
            h->lut_cap = nluts;
        }
        // the tables of a run go to the device before any of its gates
        HIP_TRY(hipMemcpyAsync(h->d_lut_c, lut_changes, nluts * 16 * h->W * sizeof(uint64_t),
                               hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->d_lut_s, lut_signs, nluts * 16 * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return PP_OK;
}

// Switch recording on (allocating its buffers) or off.
int set_recording(pp_handle* h, bool on) {
    h->recording = on;
    if (!on) {
        h->bufs.tpj = nullptr;
        h->bufs.rec = nullptr;
        h->bufs.rcoef = nullptr;
        h->bufs.rec_words = h->bufs.rec_rows = 0;
        return PP_OK;
    }
    if (h->tpj_cap < h->cap) {
        if (h->d_tpj) hipFree(h->d_tpj);
        h->d_tpj = nullptr;
        h->tpj_cap = 0;
        HIP_TRY(hipMalloc(&h->d_tpj, h->cap * sizeof(uint32_t)));
        h->tpj_cap = h->cap;
    }
    if (!h->d_rec || !h->d_rcoef)
        if (int rc = grow_records(h, std::max<size_t>(h->rec_words, std::max<size_t>(1 << 20, 4 * h->cap)),
                                  std::max<size_t>(h->rec_rows, std::max<size_t>(1 << 16, h->cap)), 0, 0))
            return rc;
    h->bufs.tpj = h->d_tpj;
    h->bufs.rec = h->d_rec;
    h->bufs.rcoef = h->d_rcoef;
    h->bufs.rec_words = (uint32_t)h->rec_words;
    h->bufs.rec_rows = (uint32_t)h->rec_rows;
    return PP_OK;
}

// A device copy of the live sum (keys and coefficients).
struct Snapshot {
    uint64_t* keys = nullptr;
    double* coeffs = nullptr;
    uint32_t n = 0;
};

void free_snapshot(Snapshot& sn) {
    if (sn.keys) hipFree(sn.keys);
    if (sn.coeffs) hipFree(sn.coeffs);
    sn = Snapshot{};
}

int take_snapshot(pp_handle* h, Snapshot& sn) {
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    const size_t W = h->W, B = h->B, m = std::max<uint32_t>(n, 1);
    HIP_TRY(hipMalloc(&sn.keys, m * W * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&sn.coeffs, m * B * sizeof(double)));
    sn.n = n;
    for (size_t l = 0; l < W && n; ++l)
        HIP_TRY(hipMemcpyAsync(sn.keys + l * n, h->bufs.keys[cur] + l * h->cap, n * sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, h->stream));
    if (n)
        HIP_TRY(hipMemcpyAsync(sn.coeffs, h->bufs.coeffs[cur], (size_t)n * B * sizeof(double),
                               hipMemcpyDeviceToDevice, h->stream));
    return PP_OK;
}

int restore_snapshot(pp_handle* h, const Snapshot& sn) {
    if (sn.n > h->cap)
        if (int rc = grow(h, sn.n)) return rc;
    const size_t W = h->W, B = h->B, n = sn.n;
    for (size_t l = 0; l < W && n; ++l)
        HIP_TRY(hipMemcpyAsync(h->bufs.keys[0] + l * h->cap, sn.keys + l * n, n * sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, h->stream));
    if (n)
        HIP_TRY(hipMemcpyAsync(h->bufs.coeffs[0], sn.coeffs, n * B * sizeof(double), hipMemcpyDeviceToDevice,
                               h->stream));
    return reset_ctl(h, (uint32_t)n, 0);
}

template <int W>
int seed_adjoint(pp_handle* h, const Readout& rd, double* a) {
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    if (n == 0) return PP_OK;
    hipLaunchKernelGGL(k_adj_seed<W>, dim3(grid_for(h, (size_t)n * h->B, WG * 4)), dim3(WG), 0, h->stream,
                       h->bufs.keys[cur], (size_t)h->cap, n, h->B, rd, a);
    HIP_TRY(hipGetLastError());
    return PP_OK;
}

// The readout of an overlap kind (enum pp_overlap_kind), checking `ones`.
int make_readout(const pp_handle* h, int kind, const uint64_t* ones, Readout* rd) {
    *rd = Readout{};
    if (kind != PP_OVERLAP_ZERO && kind != PP_OVERLAP_PLUS && kind != PP_OVERLAP_COMPUTATIONAL &&
        kind != PP_OVERLAP_MAXMIXED)
        return fail(PP_ERR_ARG, "unknown overlap kind");
    rd->kind = kind == PP_OVERLAP_PLUS ? RD_PLUS
               : kind == PP_OVERLAP_COMPUTATIONAL ? RD_COMP
               : kind == PP_OVERLAP_MAXMIXED ? RD_MAXMIX
                                             : RD_ZERO;
    if (kind == PP_OVERLAP_COMPUTATIONAL) {
        if (!ones) return fail(PP_ERR_ARG, "computational overlap needs the one-bit mask");
        for (int l = 0; l < h->W; ++l) {
            if (ones[l] & ~0x5555555555555555ull)
                return fail(PP_ERR_ARG, "one-bit mask: only bit 2q of a qubit q may be set");
            rd->ones[l] = ones[l];
        }
        for (int q = h->nqubits; q < 32 * h->W; ++q)
            if ((rd->ones[q >> 5] >> (2 * (q & 31))) & 1ull) return fail(PP_ERR_ARG, "one-bit index out of range");
    }
    return PP_OK;
}

// Loss = the readout `rd` of the propagated sum, and its gradient, by a
// forward run that records every gate (per segment of checkpoint_gates gates,
// from a snapshot of the segment's input sum) and a reverse sweep over the
// records.
int adjoint_impl(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
                 const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight, double min_abs,
                 size_t checkpoint_gates, const Readout& rd, double* loss, double* grad) {
    const int B = h->B;
    std::vector<HostGate> gs;
    std::vector<double> atrig;
    if (int rc = resolve_gates(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, gs, &atrig))
        return rc;
    const size_t seg = checkpoint_gates ? std::min(checkpoint_gates, std::max<size_t>(ngates, 1))
                                        : std::max<size_t>(ngates, 1);
    const size_t nseg = ngates ? (ngates + seg - 1) / seg : 1;
    // device arrays of the sweep
    double* d_atrig = nullptr;
    double* d_a[2] = {nullptr, nullptr};
    double* d_gpart = nullptr;
    double* d_dgrad = nullptr;
    uint32_t* d_slots = nullptr;  // per rotation slot of a segment: its block count, then its gate offset
    std::vector<Snapshot> snaps;
    Snapshot final_sum;
    int rc = PP_OK;
    auto cleanup = [&]() {
        if (d_atrig) hipFree(d_atrig);
        if (d_a[0]) hipFree(d_a[0]);
        if (d_a[1]) hipFree(d_a[1]);
        if (d_gpart) hipFree(d_gpart);
        if (d_dgrad) hipFree(d_dgrad);
        if (d_slots) hipFree(d_slots);
        for (auto& sn : snaps) free_snapshot(sn);
        free_snapshot(final_sum);
    };
    auto run_segment = [&](size_t k) -> int {
        return run_resolved(h, gs, k * seg, std::min(ngates, (k + 1) * seg), max_weight, min_abs, nullptr);
    };
    // rows of BP = 2^lgBP threads, one column per thread (B <= WG), or rows
    // of WG threads covering B in groups of WG columns
    const int lgBP = B <= 1 ? 0 : std::min(8, 32 - __builtin_clz((unsigned)(B - 1)));
    static_assert(ADJ_MAX_GROUPS * WG >= 1024, "the column groups cover the largest batch");
    // a fixed block count (not the tuned persistent grid), so the per-block
    // partial gradients, and their sums, do not depend on a launch shape
    const unsigned gblocks = ADJ_BLOCKS;
    // The blocks' partial gradients of every rotation of a segment are kept and
    // summed by one launch after the segment's sweep (one workgroup per
    // rotation, blocks in order: the sums a launch per gate would give), unless
    // that takes more than adj_defer_bytes or the allocation fails; then each
    // rotation sums its own.
    size_t maxrot = 0;
    for (size_t k = 0; k < nseg; ++k) {
        size_t r = 0;
        for (size_t g = k * seg; g < std::min(ngates, (k + 1) * seg); ++g) r += gs[g].kind == PP_GATE_ROTATION;
        maxrot = std::max(maxrot, r);
    }
    bool deferred = maxrot > 0 && maxrot * (size_t)gblocks * B * sizeof(double) <= h->adj_defer_bytes;
    std::vector<uint32_t> slots(deferred ? 2 * maxrot : 0, 0u);
    do {
        if (ngates) {
            if (hipMalloc(&d_atrig, atrig.size() * sizeof(double)) != hipSuccess ||
                hipMalloc(&d_dgrad, ngates * (size_t)B * sizeof(double)) != hipSuccess) {
                rc = fail(PP_ERR_NOMEM, "adjoint: device allocation failed");
                break;
            }
            if (deferred && (hipMalloc(&d_gpart, maxrot * gblocks * B * sizeof(double)) != hipSuccess ||
                             hipMalloc(&d_slots, 2 * maxrot * sizeof(uint32_t)) != hipSuccess)) {
                // no room for every rotation's partials: sum them gate by gate
                (void)hipGetLastError();
                if (d_gpart) hipFree(d_gpart);
                d_gpart = nullptr;
                deferred = false;
                slots.clear();
            }
            if (!d_gpart && hipMalloc(&d_gpart, (size_t)gblocks * B * sizeof(double)) != hipSuccess) {
                rc = fail(PP_ERR_NOMEM, "adjoint: device allocation failed");
                break;
            }
            if (hipMemcpyAsync(d_atrig, atrig.data(), atrig.size() * sizeof(double), hipMemcpyHostToDevice,
                               h->stream) != hipSuccess ||
                hipMemsetAsync(d_dgrad, 0, ngates * (size_t)B * sizeof(double), h->stream) != hipSuccess) {
                rc = fail(PP_ERR_HIP, "adjoint: staging failed");
                break;
            }
        }
        // pass 1 (segments only): snapshots of every segment's input sum
        if (nseg > 1) {
            snaps.resize(nseg);
            for (size_t k = 0; k < nseg && !rc; ++k) {
                if ((rc = take_snapshot(h, snaps[k]))) break;
                rc = run_segment(k);
            }
            if (rc) break;
            if ((rc = take_snapshot(h, final_sum))) break;
        }
        size_t acap = 0;
        int acur = 0;  // d_a[acur] holds the adjoint of the current end of the sweep
        for (size_t kk = nseg; kk-- > 0;) {
            if (nseg > 1) {
                if ((rc = restore_snapshot(h, snaps[kk]))) break;
            } else {
                uint32_t n = 0, cur = 0;
                if ((rc = live(h, &n, &cur))) break;
                if ((rc = reset_ctl(h, n, cur))) break;
            }
            if ((rc = set_recording(h, true))) break;
            rc = run_segment(kk);
            set_recording(h, false);
            if (rc) break;
            if (kk == nseg - 1) {
                // the sweep starts at the final sum: loss and seed
                acap = h->cap;
                if (hipMalloc(&d_a[0], acap * B * sizeof(double)) != hipSuccess ||
                    hipMalloc(&d_a[1], acap * B * sizeof(double)) != hipSuccess) {
                    rc = fail(PP_ERR_NOMEM, "adjoint: device allocation failed");
                    break;
                }
                switch (h->W) {
                case 1: rc = overlap_impl<1>(h, loss, rd); break;
                case 2: rc = overlap_impl<2>(h, loss, rd); break;
                default: rc = overlap_impl<4>(h, loss, rd); break;
                }
                if (rc) break;
                switch (h->W) {
                case 1: rc = seed_adjoint<1>(h, rd, d_a[0]); break;
                case 2: rc = seed_adjoint<2>(h, rd, d_a[0]); break;
                default: rc = seed_adjoint<4>(h, rd, d_a[0]); break;
                }
                if (rc) break;
            }
            // the sweep through this segment, last gate first
            const size_t g0 = kk * seg, g1 = std::min(ngates, (kk + 1) * seg);
            uint32_t nslot = 0;
            for (size_t g = g1; g-- > g0;) {
                const uint32_t* lg = h->h_log + LOG_WORDS * g;
                const uint32_t nin = lg[0], na = lg[1], nout = lg[2], roff = lg[3], coff = lg[4];
                if (std::max(nin, nout) > acap) {
                    rc = fail(PP_ERR_STATE, "adjoint: a gate's sum outgrew the adjoint buffers");
                    break;
                }
                const double* tr = d_atrig + g * 2 * (size_t)B;
                if (gs[g].kind == PP_GATE_ROTATION) {
                    if (na == 0 || nin == 0) continue;  // identity
                    const unsigned rows = WG >> lgBP;
                    const unsigned nb = (unsigned)std::min<size_t>(gblocks, ((size_t)nin + na + rows - 1) / rows);
                    double* part = d_gpart + (deferred ? (size_t)nslot * gblocks * B : 0);
                    hipLaunchKernelGGL(k_adj_rotation, dim3(nb), dim3(WG), 0, h->stream, d_a[acur], d_a[acur ^ 1],
                                       h->d_rec + roff, nin, na, h->d_rcoef + (size_t)coff * B, B, lgBP, tr,
                                       part);
                    if (deferred) {
                        slots[nslot] = nb;
                        slots[maxrot + nslot] = (uint32_t)(g - g0);
                        ++nslot;
                    } else {
                        hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(WG), 0, h->stream, d_gpart, (int)nb, B,
                                           d_dgrad + g * (size_t)B);
                    }
                } else {
                    // Pauli noise (factor on damped terms) or a Clifford (factor -1 on
                    // terms whose sign flips): a_in[i] = a_out[its position] x factor
                    if (nin == 0) continue;
                    hipLaunchKernelGGL(k_adj_noise, dim3(grid_for(h, (size_t)nin * B, WG * 4)), dim3(WG), 0,
                                       h->stream, d_a[acur], d_a[acur ^ 1], h->d_rec + roff, nin, B, tr);
                }
                if (hipGetLastError() != hipSuccess) {
                    rc = fail(PP_ERR_HIP, "adjoint: kernel launch failed");
                    break;
                }
                acur ^= 1;
            }
            if (rc) break;
            if (deferred && nslot > 0) {
                if (hipMemcpyAsync(d_slots, slots.data(), 2 * maxrot * sizeof(uint32_t), hipMemcpyHostToDevice,
                                   h->stream) != hipSuccess) {
                    rc = fail(PP_ERR_HIP, "adjoint: staging failed");
                    break;
                }
                hipLaunchKernelGGL(k_adj_final_all, dim3(nslot), dim3(WG), 0, h->stream, d_gpart, d_slots,
                                   d_slots + maxrot, gblocks, B, d_dgrad + g0 * (size_t)B);
                // slots are refilled for the next segment only once the copy has been read
                if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
                    rc = fail(PP_ERR_HIP, "adjoint: gradient sums failed");
                    break;
                }
            }
        }
        if (rc) break;
        // per-gate gradients, summed into parameters in gate order
        std::vector<double> dg(ngates * (size_t)B);
        if (ngates && hipMemcpyAsync(dg.data(), d_dgrad, dg.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     h->stream) != hipSuccess) {
            rc = fail(PP_ERR_HIP, "adjoint: download failed");
            break;
        }
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail(PP_ERR_HIP, "adjoint: synchronization failed");
            break;
        }
        for (size_t q = 0; q < nparams * (size_t)B; ++q) grad[q] = 0.0;
        for (size_t g = 0; g < ngates; ++g)
            if (gs[g].kind == PP_GATE_ROTATION && gates[g].param >= 0)
                for (int b = 0; b < B; ++b) grad[(size_t)gates[g].param * B + b] += dg[g * B + b];
        // leave the propagated sum in the handle
        if (nseg > 1) rc = restore_snapshot(h, final_sum);
    } while (false);
    cleanup();
    h->last_adj_deferred = deferred;
    return rc;
}

// Chunk buffers of the resampling kernels for `nch` chunks.
int ensure_mc(pp_handle* h, size_t nch) {
    if (nch <= h->mc_chunks) return PP_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->d_mc_sum) hipFree(h->d_mc_sum);
    if (h->d_mc_cnt) hipFree(h->d_mc_cnt);
    if (h->d_mc_start) hipFree(h->d_mc_start);
    if (h->d_mc_out) hipFree(h->d_mc_out);
    h->d_mc_sum = nullptr;
    h->d_mc_cnt = nullptr;
    h->d_mc_start = nullptr;
    h->d_mc_out = nullptr;
    h->mc_chunks = 0;
    HIP_TRY(hipMalloc(&h->d_mc_sum, nch * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_mc_cnt, nch * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&h->d_mc_start, nch * sizeof(double)));
    HIP_TRY(hipMalloc(&h->d_mc_out, nch * sizeof(uint32_t)));
    h->mc_chunks = nch;
    return PP_OK;
}

// One k_mc_weigh pass: per-chunk sums (kept in `sums`) and their total and
// count, added chunk by chunk in order.
int mc_weigh(pp_handle* h, size_t nch, int mode, double p, int squared, std::vector<double>& sums, double* total,
             uint64_t* count) {
    hipLaunchKernelGGL(k_mc_weigh, dim3((unsigned)nch), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, mode, p,
                       squared, h->d_mc_sum, h->d_mc_cnt);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> cnts(nch);
    sums.resize(nch);
    HIP_TRY(hipMemcpyAsync(sums.data(), h->d_mc_sum, nch * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(cnts.data(), h->d_mc_cnt, nch * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    double t = 0.0;
    uint64_t c = 0;
    for (size_t k = 0; k < nch; ++k) {
        t += sums[k];
        c += cnts[k];
    }
    *total = t;
    *count = c;
    return PP_OK;
}

// Resampling of a B = 1 sum (src/Base/MonteCarlo/resample.jl): the totals and
// the comb spacing on the host from per-chunk sums, then the new coefficients
// and the survivors' compaction on the device (pp_mc.h).
template <int W>
int resample_impl(pp_handle* h, int method, size_t target, int squared, double offset, const double* draws,
                  int calibrate, double rtol, double atol, double* spacing_out) {
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    if (method != PP_RESAMPLE_MULTINOMIAL && target > n)
        return fail(PP_ERR_ARG, "this resampler cannot grow " + std::to_string(n) + " terms to target_size " +
                                    std::to_string(target) + "; use the multinomial one");
    if (n == 0) {
        if (spacing_out) *spacing_out = 0.0;
        return PP_OK;
    }
    const size_t nch = ((size_t)n + MC_CHUNK - 1) / MC_CHUNK;
    if (int rc = ensure_mc(h, nch)) return rc;
    std::vector<double> sums;
    double total = 0.0, thr = 0.0, spacing = 0.0;
    uint64_t cnt = 0;
    if (method == PP_RESAMPLE_SEMIDETERMINISTIC) {
        // keep the terms heavier than the mean weight per target term; a comb over the others
        // (semideterministic_systematic_resample!, resample.jl:139-170)
        double sabs = 0.0;
        if (int rc = mc_weigh(h, nch, MC_ABS, 0.0, 0, sums, &sabs, &cnt)) return rc;
        thr = sabs / (double)target;
        uint64_t n_kept = 0;
        if (int rc = mc_weigh(h, nch, MC_COMB, thr, 0, sums, &total, &n_kept)) return rc;
        spacing = target > n_kept ? total / (double)(target - n_kept) : 0.0;
    } else {
        if (int rc = mc_weigh(h, nch, MC_W, 0.0, squared, sums, &total, &cnt)) return rc;
        spacing = total / (double)target;
        if (method == PP_RESAMPLE_SYSTEMATIC && calibrate) {
            // Newton steps from total / target toward the expected survivor count
            // (_calibrate_comb_spacing, resample.jl:118-130)
            const double lowest = (1 - rtol) * (double)target - atol;
            std::vector<double> hs;
            for (int it = 0; it < 10; ++it) {
                double light = 0.0;
                uint64_t n_heavy = 0;
                if (int rc = mc_weigh(h, nch, MC_HEAVY, spacing, squared, hs, &light, &n_heavy)) return rc;
                if ((double)n_heavy + light / spacing >= lowest || light == 0.0) break;
                spacing = light / (double)(target - n_heavy);
            }
        }
    }
    // chunk starts: the chunks' weights added in order
    std::vector<double> starts(nch);
    double r = 0.0;
    for (size_t k = 0; k < nch; ++k) {
        starts[k] = r;
        r += sums[k];
    }
    HIP_TRY(hipMemcpyAsync(h->d_mc_start, starts.data(), nch * sizeof(double), hipMemcpyHostToDevice, h->stream));
    McParams m{};
    m.method = method;
    m.squared = squared;
    m.thr = thr;
    m.spacing = spacing;
    m.inv = spacing == 0.0 ? 0.0 : 1 / spacing;
    m.offset = offset;
    if (method == PP_RESAMPLE_MULTINOMIAL) {
        if (target > h->mc_draws_cap) {
            HIP_TRY(hipStreamSynchronize(h->stream));
            if (h->d_mc_draws) hipFree(h->d_mc_draws);
            h->d_mc_draws = nullptr;
            h->mc_draws_cap = 0;
            HIP_TRY(hipMalloc(&h->d_mc_draws, target * sizeof(double)));
            h->mc_draws_cap = target;
        }
        std::vector<double> scaled(target);
        for (size_t k = 0; k < target; ++k) scaled[k] = draws[k] * total;
        HIP_TRY(hipMemcpyAsync(h->d_mc_draws, scaled.data(), target * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        m.draws = h->d_mc_draws;
        m.ndraws = (uint32_t)target;
    }
    hipLaunchKernelGGL(k_mc_apply, dim3((unsigned)nch), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq, m,
                       h->d_mc_start, h->bufs.tcoef, h->d_mc_cnt);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> keeps(nch), outs(nch);
    HIP_TRY(hipMemcpyAsync(keeps.data(), h->d_mc_cnt, nch * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    uint64_t kept = 0;
    for (size_t k = 0; k < nch; ++k) {
        outs[k] = (uint32_t)kept;
        kept += keeps[k];
    }
    HIP_TRY(hipMemcpyAsync(h->d_mc_out, outs.data(), nch * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_mc_scatter<W>, dim3((unsigned)nch), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq,
                       h->bufs.tcoef, h->d_mc_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (spacing_out) *spacing_out = spacing;
    return reset_ctl(h, (uint32_t)kept, cur ^ 1);
}

// Terms in any order with repeats: upload as given, sort on the device (stable,
// limb by limb), sum every run of equal strings in input order, truncate.
template <int W>
int load_merge_impl(pp_handle* h, int max_weight, double min_abs) {
    if (int rc = ensure_sort(h)) return rc;
    const int p = h->seq & 1;
    const uint32_t* n_ptr = &h->d_ctl->st[p].n;
    const unsigned grid = grid_for(h, h->cap, WG * 4);
    int c = 0;
    for (int l = 0; l < W; ++l) {
        hipLaunchKernelGGL(k_rs_limb_keys, dim3(grid), dim3(WG), 0, h->stream, h->d_ctl, p, h->bufs, 1, 0, n_ptr, l,
                           0, l == 0 ? nullptr : h->d_rs_v[c], h->d_rs_k[c ^ 1], h->d_rs_v[c ^ 1]);
        c ^= 1;
        c = enq_radix_sort(h, n_ptr, c, 64);
    }
    hipLaunchKernelGGL(k_reduce_runs<W>, dim3(lb_grid(h, (const void*)k_reduce_runs<W>, 0, h->cap, KR_TILE)),
                       dim3(WG), 0, h->stream, h->bufs, h->d_ctl, lb_launch(h, nullptr), h->d_rs_v[c], h->B,
                       max_weight, min_abs);
    HIP_TRY(hipGetLastError());
    if (int rc = sync_ctl(h)) return rc;
    if (h->h_ctl->overflow) return fail(PP_ERR_STATE, "load_merge: a look-back wait ran out");
    return reset_ctl(h, h->h_ctl->st[p ^ 1].n, 1);
}

}  // namespace

// ============================================================ C ABI

namespace {
int create_handle(int nqubits, int key_words, int batch, size_t capacity, hipStream_t shared, pp_handle** out);

// The transposed table of a packed Clifford table (transposecliffordmap,
// src/Gates/cliffordgates.jl:94-98): entry idx moves the pattern idx to
// idx ^ (the change on the gate's qubits) with a sign; the transpose sends that
// pattern back with the same change and sign.
void transpose_table(int nq, const int* qubits, const uint64_t* change, const double* sign, int W, uint64_t* tch,
                     double* tsg) {
    std::fill(tch, tch + 16 * W, 0ull);
    std::fill(tsg, tsg + 16, 1.0);
    const int nent = 1 << (2 * nq);
    for (int idx = 0; idx < nent; ++idx) {
        int moved = 0;
        for (int a = 0; a < nq; ++a)
            moved |= (int)((change[idx * W + (qubits[a] >> 5)] >> (2 * (qubits[a] & 31))) & 3u) << (2 * a);
        const int to = idx ^ moved;
        for (int l = 0; l < W; ++l) tch[to * W + l] = change[idx * W + l];
        tsg[to] = sign[idx];
    }
}

// Cap the dual (hd) to the strings of the operator (h), both current; on a
// look-back wait that ran out the compaction runs again with atomic tile claims.
template <int W>
int intersect(pp_handle* hd, pp_handle* h) {
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (int rc = sync_ctl(hd)) return rc;
        const GateState in = hd->h_ctl->st[hd->seq & 1];
        Launch L = lb_launch(hd, nullptr);
        if (attempt) L.dynamic = 1;
        hipLaunchKernelGGL(k_intersect<W>, dim3(lb_grid(hd, (const void*)k_intersect<W>, 0, hd->cap, KI_TILE)),
                           dim3(WG), 0, hd->stream, hd->bufs, hd->d_ctl, L, h->bufs, h->d_ctl, h->seq);
        HIP_TRY(hipGetLastError());
        if (int rc = sync_ctl(hd)) return rc;
        if (hd->h_ctl->overflow == 0) {
            hd->seq += 1;
            return PP_OK;
        }
        hd->fallbacks += 1;
        if (int rc = reset_ctl(hd, in)) return rc;
    }
    return fail(PP_ERR_STATE, "rewind: the dual's compaction did not complete");
}

// rewindgradient (src/gradient.jl:40-108) on a B = 1 handle: forward run,
// loss, the dual sum on a work handle that shares the stream, then the undo
// sweep (the gate list backwards, each gate's Schrodinger form) with the
// commutator dot of every parametrized rotation taken before it is undone.
template <int W>
int rewind_impl(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
                const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight, double min_abs,
                const Readout& rd, double* loss, double* grad) {
    if (h->B != 1) return fail(PP_ERR_UNSUPPORTED, "rewindgradient takes one coefficient column");
    for (size_t g = 0; g < ngates; ++g)
        if (gates[g].kind != PP_GATE_ROTATION && gates[g].kind != PP_GATE_CLIFFORD)
            return fail(PP_ERR_UNSUPPORTED, "rewindgradient: gate " + std::to_string(g) +
                                                " is not a rotation or a Clifford (noise channels cannot be undone)");
    std::vector<HostGate> gs;
    if (int rc = resolve_gates(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, gs, nullptr))
        return rc;
    // the undo list: the circuit's own order, rotations at -theta, transposed tables
    std::vector<pp_gate> ug(ngates);
    for (size_t k = 0; k < ngates; ++k) {
        ug[k] = gates[ngates - 1 - k];
        if (ug[k].kind == PP_GATE_ROTATION) {
            ug[k].value = -(ug[k].param >= 0 ? theta[(size_t)ug[k].param] : ug[k].value);
            ug[k].param = -1;
        }
    }
    std::vector<uint64_t> tch(std::max<size_t>(nluts, 1) * 16 * W, 0ull);
    std::vector<double> tsg(std::max<size_t>(nluts, 1) * 16, 1.0);
    std::vector<uint8_t> done(nluts, 0);
    for (size_t g = 0; g < ngates; ++g) {
        if (gates[g].kind != PP_GATE_CLIFFORD || done[gates[g].lut]) continue;
        const size_t sl = gates[g].lut;
        transpose_table(gates[g].nq, gates[g].qubits, lut_changes + sl * 16 * W, lut_signs + sl * 16, W,
                        &tch[sl * 16 * W], &tsg[sl * 16]);
        done[sl] = 1;
    }
    if (int rc = run_resolved(h, gs, 0, ngates, max_weight, min_abs, nullptr)) return rc;
    if (int rc = overlap_impl<W>(h, loss, rd)) return rc;
    for (size_t q = 0; q < nparams; ++q) grad[q] = 0.0;
    if (ngates == 0) return PP_OK;
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    pp_handle* hd = nullptr;
    double* d_g = nullptr;
    int rc = PP_OK;
    do {
        if ((rc = create_handle(h->nqubits, W, 1, std::max<size_t>(h->cap, 1024), h->stream, &hd))) break;
        hd->skip_identity = false;
        hd->tun = h->tun;
        hd->gmax = h->gmax;
        for (int l = 0; l < W && n; ++l)
            if (hipMemcpyAsync(hd->bufs.keys[0] + (size_t)l * hd->cap, h->bufs.keys[cur] + (size_t)l * h->cap,
                               n * sizeof(uint64_t), hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
                rc = fail(PP_ERR_HIP, "rewind: dual copy failed");
                break;
            }
        if (rc) break;
        if ((rc = reset_ctl(hd, n, 0))) break;
        hipLaunchKernelGGL(k_dual_seed<W>, dim3(grid_for(hd, n, WG * 4)), dim3(WG), 0, h->stream, hd->bufs,
                           hd->d_ctl, hd->seq, rd);
        std::vector<HostGate> gu, gd;
        if ((rc = resolve_gates(h, ug.data(), ngates, nullptr, 0, tch.data(), tsg.data(), nluts, gu, nullptr))) break;
        if ((rc = resolve_gates(hd, ug.data(), ngates, nullptr, 0, tch.data(), tsg.data(), nluts, gd, nullptr))) break;
        if (hipMalloc(&d_g, ngates * sizeof(double)) != hipSuccess ||
            hipMemsetAsync(d_g, 0, ngates * sizeof(double), h->stream) != hipSuccess) {
            rc = fail(PP_ERR_NOMEM, "rewind: device allocation failed");
            break;
        }
        for (size_t k = 0; k < ngates && !rc; ++k) {
            const pp_gate& og = gates[ngates - 1 - k];
            if (og.kind == PP_GATE_ROTATION && og.param >= 0) {
                // _generatorcommutatordot before the gate is undone
                const unsigned gb = std::min<unsigned>(2048u, (unsigned)std::max<size_t>(1, (h->cap + WG - 1) / WG));
                if ((rc = ensure_partial(h, (size_t)gb * 2))) break;
                hipLaunchKernelGGL(k_comm_dot<W>, dim3(gb), dim3(WG), 0, h->stream, h->bufs, h->d_ctl, h->seq,
                                   hd->bufs, hd->d_ctl, hd->seq, to_key<W>(og.mask), h->d_partial);
                hipLaunchKernelGGL(k_overlap_final, dim3(1), dim3(WG), 0, h->stream, h->d_partial, (int)gb, 1,
                                   d_g + k);
                if (hipGetLastError() != hipSuccess) {
                    rc = fail(PP_ERR_HIP, "rewind: kernel launch failed");
                    break;
                }
            }
            // the operator truncates; the dual does not (_undostep!, :84-108)
            if ((rc = run_resolved(h, gu, k, k + 1, max_weight, min_abs, nullptr))) break;
            if ((rc = run_resolved(hd, gd, k, k + 1, -1, 0.0, nullptr))) break;
            rc = intersect<W>(hd, h);
        }
        if (rc) break;
        std::vector<double> hg(ngates);
        if (hipMemcpyAsync(hg.data(), d_g, ngates * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail(PP_ERR_HIP, "rewind: download failed");
            break;
        }
        for (size_t k = 0; k < ngates; ++k) {
            const pp_gate& og = gates[ngates - 1 - k];
            if (og.kind == PP_GATE_ROTATION && og.param >= 0) grad[og.param] = hg[k];
        }
    } while (false);
    if (d_g) hipFree(d_g);
    if (hd) pp_destroy(hd);
    return rc;
}
}  // namespace

extern "C" {

int pp_version(void) { return 200; }

const char* pp_last_error(void) { return g_err.c_str(); }

int pp_create(int nqubits, int key_words, int batch, size_t capacity, pp_handle** out) {
    return create_handle(nqubits, key_words, batch, capacity, nullptr, out);
}

}  // extern "C"

namespace {

// A handle, on its own new stream or (work handles) on `shared`.
int create_handle(int nqubits, int key_words, int batch, size_t capacity, hipStream_t shared, pp_handle** out) {
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
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    h->device = dev;
    // a device without the scratch address still gets its handle; only
    // pp_overlap_noalloc reports the failure
    (void)scratch_setup(dev, &h->scratch);
    h->cus = std::max(1, cus);
    h->gmax = (unsigned)std::max(64, h->tun.grid_per_cu * h->cus);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
        h->clock_khz = (double)khz;
    if (shared) {
        h->stream = shared;
        h->own_stream = false;
    } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        return bail(fail(PP_ERR_HIP, "hipStreamCreate failed"));
    }
    if (hipMalloc(&h->d_ctl, sizeof(Ctl)) != hipSuccess ||
        hipHostMalloc(&h->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&h->d_out, batch * sizeof(double)) != hipSuccess ||
        hipMalloc(&h->d_norm, batch * sizeof(double)) != hipSuccess ||
        hipMalloc(&h->d_pres_part, ((size_t)std::max(64, MAX_GRID_PER_CU * h->cus) + 1) * PRES_WORDS *
                                       sizeof(uint64_t)) != hipSuccess ||
        hipHostMalloc(&h->h_pres, PRES_WORDS * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
        return bail(fail(PP_ERR_NOMEM, "device allocation failed"));
    if (int rc = reset_ctl(h, 0, 0)) return bail(rc);
    if (int rc = grow(h, std::max<size_t>(capacity, 1024))) return bail(rc);
    // a work handle takes its parent's launch shapes (the caller copies them)
    if (!shared)
        if (int rc = pp_load_tuning(h, nullptr)) return bail(rc);
    *out = h;
    return PP_OK;
}

}  // namespace

extern "C" {

int pp_destroy(pp_handle* h) {
    if (!h) return PP_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (hipEvent_t e : h->ev)
        if (e) hipEventDestroy(e);
    free_terms(h);
    if (h->d_ctl) hipFree(h->d_ctl);
    if (h->h_ctl) hipHostFree(h->h_ctl);
    if (h->d_log) hipFree(h->d_log);
    if (h->h_log) hipHostFree(h->h_log);
    if (h->d_ktime) hipFree(h->d_ktime);
    if (h->d_wgstamp) hipFree(h->d_wgstamp);
    if (h->d_ldsg) hipFree(h->d_ldsg);
    if (h->h_ldsg) hipHostFree(h->h_ldsg);
    drop_graphs(h);
    if (h->d_trig) hipFree(h->d_trig);
    if (h->h_trig) hipHostFree(h->h_trig);
    if (h->d_partial) hipFree(h->d_partial);
    if (h->d_out) hipFree(h->d_out);
    if (h->d_norm) hipFree(h->d_norm);
    if (h->d_mc_sum) hipFree(h->d_mc_sum);
    if (h->d_mc_cnt) hipFree(h->d_mc_cnt);
    if (h->d_mc_start) hipFree(h->d_mc_start);
    if (h->d_mc_out) hipFree(h->d_mc_out);
    if (h->d_mc_draws) hipFree(h->d_mc_draws);
    if (h->d_pres_part) hipFree(h->d_pres_part);
    if (h->h_pres) hipHostFree(h->h_pres);
    if (h->d_tpj) hipFree(h->d_tpj);
    if (h->d_rec) hipFree(h->d_rec);
    if (h->d_rcoef) hipFree(h->d_rcoef);
    for (int b = 0; b < 2; ++b) {
        if (h->d_rs_k[b]) hipFree(h->d_rs_k[b]);
        if (h->d_rs_v[b]) hipFree(h->d_rs_v[b]);
    }
    if (h->d_rs_hist) hipFree(h->d_rs_hist);
    if (h->d_lut_c) hipFree(h->d_lut_c);
    if (h->d_lut_s) hipFree(h->d_lut_s);
    if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
    delete h;
    return PP_OK;
}

int pp_clone(pp_handle* h, pp_handle** out) {
    if (!h || !out) return fail(PP_ERR_ARG, "NULL argument");
    *out = nullptr;
    uint32_t n = 0, cur = 0;
    if (int rc = live(h, &n, &cur)) return rc;
    pp_handle* c = nullptr;
    if (int rc = create_handle(h->nqubits, h->W, h->B, std::max<size_t>(n, 1024), nullptr, &c)) return rc;
    c->spin_limit = h->spin_limit;
    c->force_dynamic = h->force_dynamic;
    c->force_rerun = h->force_rerun;
    c->lds_runs = h->lds_runs;
    c->use_graphs = h->use_graphs;
    c->chain_tails = h->chain_tails;
    c->skip_identity = h->skip_identity;
    c->adj_defer_bytes = h->adj_defer_bytes;
    c->tun = h->tun;
    c->gmax = h->gmax;
    // the copies run on the source's stream, which is idle once they are done
    bool ok = true;
    for (int l = 0; l < h->W && n; ++l)
        ok = ok && hipMemcpyAsync(c->bufs.keys[0] + (size_t)l * c->cap, h->bufs.keys[cur] + (size_t)l * h->cap,
                                  (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
    if (n)
        ok = ok && hipMemcpyAsync(c->bufs.coeffs[0], h->bufs.coeffs[cur], (size_t)n * h->B * sizeof(double),
                                  hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
    if (!ok) {
        pp_destroy(c);
        return fail(PP_ERR_HIP, "pp_clone: device copy failed");
    }
    if (int rc = reset_ctl(c, n, 0)) {
        pp_destroy(c);
        return rc;
    }
    *out = c;
    return PP_OK;
}

int pp_info(pp_handle* h, int* nqubits, int* key_words, int* batch, size_t* n, size_t* capacity) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (nqubits) *nqubits = h->nqubits;
    if (key_words) *key_words = h->W;
    if (batch) *batch = h->B;
    if (capacity) *capacity = h->cap;
    if (n) {
        uint32_t nn = 0, cur = 0;
        if (int rc = live(h, &nn, &cur)) return rc;
        *n = nn;
    }
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
    // a new sum: gate numbers (which the kernels compare on the device) start
    // again, so a repeated propagation repeats them and can replay its graphs
    h->seq = 0;
    h->safe_until = 0;
    h->lb_count = 0;
    if (int rc = reset_ctl(h, 0, 0)) return rc;
    if (n > h->cap)
        if (int rc = grow(h, n)) return rc;
    std::vector<uint64_t> planes((size_t)W * std::max<size_t>(n, 1));
    std::vector<double> rows((size_t)B * std::max<size_t>(n, 1));
    for (size_t i = 0; i < n; ++i) {
        for (int l = 0; l < W; ++l) planes[(size_t)l * n + i] = keys[idx[i] * W + l];
        std::memcpy(&rows[i * B], coeffs + idx[i] * B, B * sizeof(double));
    }
    for (int l = 0; l < W && n > 0; ++l)
        HIP_TRY(hipMemcpyAsync(h->bufs.keys[0] + (size_t)l * h->cap, &planes[(size_t)l * n], n * sizeof(uint64_t),
                               hipMemcpyHostToDevice, h->stream));
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(h->bufs.coeffs[0], rows.data(), n * B * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return reset_ctl(h, (uint32_t)n, 0);
}

int pp_resample(pp_handle* h, int method, size_t target_size, int squared, double offset, const double* draws,
                int calibrate, double rtol, double atol, double* spacing_out) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (h->B != 1) return fail(PP_ERR_UNSUPPORTED, "resampling takes one coefficient column");
    if (method != PP_RESAMPLE_SEMIDETERMINISTIC && method != PP_RESAMPLE_SYSTEMATIC &&
        method != PP_RESAMPLE_MULTINOMIAL)
        return fail(PP_ERR_ARG, "unknown resampling method");
    if (target_size == 0 || target_size > MAX_TERMS) return fail(PP_ERR_ARG, "target_size must be positive");
    if (method == PP_RESAMPLE_SEMIDETERMINISTIC && squared)
        return fail(PP_ERR_ARG, "semideterministic_systematic_resample! does not support squared=true");
    if (method != PP_RESAMPLE_MULTINOMIAL && !(offset >= 0.0 && offset < 1.0))
        return fail(PP_ERR_ARG, "the comb offset must lie in [0, 1)");
    if (method == PP_RESAMPLE_MULTINOMIAL) {
        if (!draws) return fail(PP_ERR_ARG, "multinomial resampling needs target_size draws");
        for (size_t k = 0; k < target_size; ++k)
            if (!(draws[k] >= 0.0 && draws[k] < 1.0) || (k && draws[k] < draws[k - 1]))
                return fail(PP_ERR_ARG, "draws must lie in [0, 1) in ascending order");
    }
    switch (h->W) {
    case 1: return resample_impl<1>(h, method, target_size, squared, offset, draws, calibrate, rtol, atol, spacing_out);
    case 2: return resample_impl<2>(h, method, target_size, squared, offset, draws, calibrate, rtol, atol, spacing_out);
    default: return resample_impl<4>(h, method, target_size, squared, offset, draws, calibrate, rtol, atol, spacing_out);
    }
}

int pp_load_merge(pp_handle* h, const uint64_t* keys, const double* coeffs, size_t n, int max_weight,
                  double min_abs_coeff) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (n > 0 && (!keys || !coeffs)) return fail(PP_ERR_ARG, "keys or coeffs is NULL");
    if (n > MAX_TERMS) return fail(PP_ERR_CAPACITY, "term count exceeds 2^31 - 1");
    const int W = h->W, B = h->B, qbits = 2 * h->nqubits;
    for (size_t i = 0; i < n; ++i)
        for (int l = 0; l < W; ++l) {
            const int lo = l * 64;
            if (qbits >= lo + 64) continue;
            const uint64_t allowed = qbits <= lo ? 0ull : ((1ull << (qbits - lo)) - 1);
            if (keys[i * W + l] & ~allowed) return fail(PP_ERR_ARG, "a term acts on a qubit beyond nqubits");
        }
    if (int rc = reset_ctl(h, 0, 0)) return rc;
    if (n > h->cap)
        if (int rc = grow(h, n)) return rc;
    if (n == 0) return PP_OK;
    std::vector<uint64_t> planes((size_t)W * n);
    for (size_t i = 0; i < n; ++i)
        for (int l = 0; l < W; ++l) planes[(size_t)l * n + i] = keys[i * W + l];
    for (int l = 0; l < W; ++l)
        HIP_TRY(hipMemcpyAsync(h->bufs.keys[0] + (size_t)l * h->cap, &planes[(size_t)l * n], n * sizeof(uint64_t),
                               hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->bufs.coeffs[0], coeffs, n * B * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (int rc = reset_ctl(h, (uint32_t)n, 0)) return rc;
    switch (W) {
    case 1: return load_merge_impl<1>(h, max_weight, min_abs_coeff);
    case 2: return load_merge_impl<2>(h, max_weight, min_abs_coeff);
    default: return load_merge_impl<4>(h, max_weight, min_abs_coeff);
    }
}

int pp_download(pp_handle* h, uint64_t* keys, double* coeffs, size_t cap, size_t* n) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    uint32_t m = 0, cur = 0;
    if (int rc = live(h, &m, &cur)) return rc;
    if (n) *n = m;
    if (cap < m) return fail(PP_ERR_ARG, "output arrays too small");
    const int W = h->W, B = h->B;
    if (m == 0) return PP_OK;
    std::vector<uint64_t> planes((size_t)W * m);
    for (int l = 0; l < W; ++l)
        HIP_TRY(hipMemcpyAsync(&planes[(size_t)l * m], h->bufs.keys[cur] + (size_t)l * h->cap, m * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, h->stream));
    if (coeffs)
        HIP_TRY(hipMemcpyAsync(coeffs, h->bufs.coeffs[cur], (size_t)m * B * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
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
    HostGate g{PP_GATE_ROTATION, m4, 0, 0, cos_b[0], sin_b[0]};
    return single_gate(h, g, cos_b, sin_b, max_weight, min_abs_coeff, stats);
}

int pp_apply_imaginary_rotation(pp_handle* h, const uint64_t* mask, const double* cosh_b, const double* sinh_b,
                                int normalize, int max_weight, double min_abs_coeff, pp_gate_stats* stats) {
    if (!h || !mask || !cosh_b || !sinh_b) return fail(PP_ERR_ARG, "NULL argument");
    uint64_t m4[4] = {0, 0, 0, 0};
    for (int l = 0; l < h->W; ++l) m4[l] = mask[l];
    if (int rc = check_mask(h, m4)) return rc;
    HostGate g{PP_GATE_IMAG_ROTATION, m4, 0, 0, cosh_b[0], sinh_b[0]};
    g.flags = normalize ? 1 : 0;
    return single_gate(h, g, cosh_b, sinh_b, max_weight, min_abs_coeff, stats);
}

int pp_apply_pauli_noise(pp_handle* h, int qubit, uint32_t damp_codes, const double* factor_b, int max_weight,
                         double min_abs_coeff, pp_gate_stats* stats) {
    if (!h || !factor_b) return fail(PP_ERR_ARG, "NULL argument");
    if (qubit < 0 || qubit >= h->nqubits) return fail(PP_ERR_ARG, "noise qubit out of range");
    if (damp_codes & ~0xeu) return fail(PP_ERR_ARG, "damp_codes may only hold X, Y, Z (bits 1..3)");
    HostGate g{PP_GATE_PAULI_NOISE, nullptr, qubit, damp_codes, factor_b[0], 0.0};
    return single_gate(h, g, factor_b, nullptr, max_weight, min_abs_coeff, stats);
}

int pp_apply_amplitude_damping(pp_handle* h, int qubit, const double* gamma_b, int max_weight, double min_abs_coeff,
                               pp_gate_stats* stats) {
    if (!h || !gamma_b) return fail(PP_ERR_ARG, "NULL argument");
    if (qubit < 0 || qubit >= h->nqubits) return fail(PP_ERR_ARG, "damping qubit out of range");
    for (int b = 0; b < h->B; ++b) {
        if (!(gamma_b[b] >= 0.0 && gamma_b[b] <= 1.0)) return fail(PP_ERR_ARG, "noise strength must be in [0, 1]");
        if (gamma_b[b] != gamma_b[0])
            return fail(PP_ERR_UNSUPPORTED, "amplitude damping with a different gamma per batch column");
    }
    std::vector<double> keep(h->B), created(h->B);
    for (int b = 0; b < h->B; ++b) {
        keep[b] = 1 - gamma_b[b];
        created[b] = gamma_b[b];
    }
    HostGate g{PP_GATE_AMP_DAMPING, nullptr, qubit, 0, gamma_b[0], std::sqrt(1 - gamma_b[0])};
    return single_gate(h, g, keep.data(), created.data(), max_weight, min_abs_coeff, stats);
}

int pp_apply_clifford(pp_handle* h, int nq, const int* qubits, const uint64_t* changes, const double* signs,
                      int max_weight, double min_abs_coeff, pp_gate_stats* stats) {
    if (!h || !qubits || !changes || !signs) return fail(PP_ERR_ARG, "NULL argument");
    pp_gate gt{};
    gt.kind = PP_GATE_CLIFFORD;
    gt.param = -1;
    gt.nq = nq;
    for (int a = 0; a < nq && a < 2; ++a) gt.qubits[a] = qubits[a];
    gt.lut = 0;
    std::vector<HostGate> gs;
    if (int rc = resolve_gates(h, &gt, 1, nullptr, 0, changes, signs, 1, gs, nullptr)) return rc;
    return single_gate(h, gs[0], nullptr, nullptr, max_weight, min_abs_coeff, stats);
}

int pp_run_circuit(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
                   const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight,
                   double min_abs_coeff, int64_t* counts) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    std::vector<HostGate> gs;
    if (int rc = resolve_gates(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, gs, nullptr))
        return rc;
    return run_resolved(h, gs, 0, ngates, max_weight, min_abs_coeff, counts);
}

int pp_adjoint(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
               const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight,
               double min_abs_coeff, size_t checkpoint_gates, int overlap_kind, const uint64_t* ones, double* loss,
               double* grad) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (!loss || (nparams > 0 && !grad)) return fail(PP_ERR_ARG, "loss or grad is NULL");
    Readout rd;
    if (int rc = make_readout(h, overlap_kind, ones, &rd)) return rc;
    const int rc = adjoint_impl(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, max_weight,
                                min_abs_coeff, checkpoint_gates, rd, loss, grad);
    set_recording(h, false);
    return rc;
}

int pp_rewind(pp_handle* h, const pp_gate* gates, size_t ngates, const double* theta, size_t nparams,
              const uint64_t* lut_changes, const double* lut_signs, size_t nluts, int max_weight,
              double min_abs_coeff, int overlap_kind, const uint64_t* ones, double* loss, double* grad) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (!loss || (nparams > 0 && !grad)) return fail(PP_ERR_ARG, "loss or grad is NULL");
    Readout rd;
    if (int rc = make_readout(h, overlap_kind, ones, &rd)) return rc;
    switch (h->W) {
    case 1:
        return rewind_impl<1>(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, max_weight,
                              min_abs_coeff, rd, loss, grad);
    case 2:
        return rewind_impl<2>(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, max_weight,
                              min_abs_coeff, rd, loss, grad);
    default:
        return rewind_impl<4>(h, gates, ngates, theta, nparams, lut_changes, lut_signs, nluts, max_weight,
                              min_abs_coeff, rd, loss, grad);
    }
}

int pp_overlap_zero(pp_handle* h, double* out_b) {
    if (!h || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    switch (h->W) {
    case 1: return overlap_impl<1>(h, out_b);
    case 2: return overlap_impl<2>(h, out_b);
    default: return overlap_impl<4>(h, out_b);
    }
}

int pp_overlap(pp_handle* h, int kind, const uint64_t* ones, double* out_b) {
    if (!h || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    Readout rd;
    if (int rc = make_readout(h, kind, ones, &rd)) return rc;
    switch (h->W) {
    case 1: return overlap_impl<1>(h, out_b, rd);
    case 2: return overlap_impl<2>(h, out_b, rd);
    default: return overlap_impl<4>(h, out_b, rd);
    }
}

int pp_overlap_noalloc(pp_handle* h, int kind, const uint64_t* ones, double* out_b) {
    if (!h || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    if (h->B > SCRATCH_COLS)
        return fail(PP_ERR_ARG, "pp_overlap_noalloc takes at most " + std::to_string(SCRATCH_COLS) +
                                    " coefficient columns (its static scratch); this sum has " +
                                    std::to_string(h->B));
    if (!h->scratch) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        return fail(PP_ERR_HIP, "no scratch address on the handle's device: " +
                                    g_scratch_err[h->device < MAX_DEVICES && h->device >= 0 ? h->device : 0]);
    }
    Readout rd;
    if (int rc = make_readout(h, kind, ones, &rd)) return rc;
    double* part = h->scratch;
    double* res = h->scratch + (size_t)SCRATCH_BLOCKS * SCRATCH_COLS * 2;
    switch (h->W) {
    case 1: return overlap_impl<1>(h, out_b, rd, part, res, SCRATCH_BLOCKS);
    case 2: return overlap_impl<2>(h, out_b, rd, part, res, SCRATCH_BLOCKS);
    default: return overlap_impl<4>(h, out_b, rd, part, res, SCRATCH_BLOCKS);
    }
}

int pp_max_abs_coeff(pp_handle* h, double* out_b) {
    if (!h || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    return maxabs_impl(h, out_b);
}

int pp_truncate(pp_handle* h, int max_weight, double min_abs_coeff, double min_rel_coeff, double* threshold_out) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    double thr = min_abs_coeff;
    if (min_rel_coeff >= 0) {
        if (h->B != 1) return fail(PP_ERR_UNSUPPORTED, "a relative threshold takes one coefficient column");
        double m = 0.0;
        if (int rc = maxabs_impl(h, &m)) return rc;
        thr = std::fmax(min_rel_coeff * m, min_abs_coeff);
    }
    if (threshold_out) *threshold_out = thr;
    // Pauli noise that damps nothing: every term truncated, the rest compacted
    std::vector<double> one((size_t)h->B, 1.0);
    HostGate g{PP_GATE_PAULI_NOISE, nullptr, 0, 0u, 1.0, 0.0};
    return single_gate(h, g, one.data(), nullptr, max_weight, thr, nullptr);
}

int pp_get_coeff(pp_handle* h, const uint64_t* key, double* out_b) {
    if (!h || !key || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    switch (h->W) {
    case 1: return getcoeff_impl<1>(h, key, out_b);
    case 2: return getcoeff_impl<2>(h, key, out_b);
    default: return getcoeff_impl<4>(h, key, out_b);
    }
}

int pp_scalar_product(pp_handle* h1, pp_handle* h2, double* out_b) {
    if (!h1 || !h2 || !out_b) return fail(PP_ERR_ARG, "NULL argument");
    if (h1->nqubits != h2->nqubits || h1->W != h2->W)
        return fail(PP_ERR_ARG, "scalar product of sums on different numbers of qubits");
    if (h1->B != h2->B) return fail(PP_ERR_ARG, "scalar product of sums with different batch sizes");
    switch (h1->W) {
    case 1: return scalar_impl<1>(h1, h2, out_b);
    case 2: return scalar_impl<2>(h1, h2, out_b);
    default: return scalar_impl<4>(h1, h2, out_b);
    }
}

int pp_synchronize(pp_handle* h) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PP_OK;
}

int pp_profile_control(pp_handle* h, int enable_timing, int reset) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    if (enable_timing >= 0) h->prof_mode = std::min(enable_timing, 2);
    if (reset) {
        h->gates = h->term_apps = h->fallbacks = h->identity_gates = h->lds_gates = 0;
        h->graph_launches = h->graph_captures = 0;
        for (int k = 0; k < PP_KERNEL_KINDS; ++k) {
            h->launches[k] = 0;
            h->ms[k] = 0.0;
            h->alg_bytes[k] = 0.0;
        }
    }
    return PP_OK;
}

int pp_tune(pp_handle* h, int key, int64_t value) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    switch (key) {
    case PP_TUNE_SPIN_LIMIT:
        if (value < 1 || value > 0x7fffffff) return fail(PP_ERR_ARG, "spin limit must be in 1..2^31-1");
        h->spin_limit = (uint32_t)value;
        return PP_OK;
    case PP_TUNE_DYNAMIC_TILES:
        h->force_dynamic = value != 0;
        return PP_OK;
    case PP_TUNE_FORCE_RERUN:
        h->force_rerun = value != 0;
        return PP_OK;
    case PP_TUNE_LDS_RUNS:
        h->lds_runs = value != 0;
        return PP_OK;
    case PP_TUNE_GRAPHS:
        h->use_graphs = value != 0;
        if (!h->use_graphs) drop_graphs(h);
        return PP_OK;
    case PP_TUNE_CHAIN_TAILS:
        h->chain_tails = value != 0;
        return PP_OK;
    case PP_TUNE_SKIP_IDENTITY:
        h->skip_identity = value != 0;
        return PP_OK;
    case PP_TUNE_ADJ_DEFER_BYTES:
        if (value < 0) return fail(PP_ERR_ARG, "the partial-gradient store size must be non-negative");
        h->adj_defer_bytes = (size_t)value;
        return PP_OK;
    case PP_TUNE_MERGE_ITEMS:
        if (value != 0 && value != 2 && value != 3 && value != 4)
            return fail(PP_ERR_ARG, "merge items must be 2, 3 or 4 (0: the key width's default)");
        h->tun.merge_items = (int)value;
        return PP_OK;
    case PP_TUNE_MERGE_DEFER:
        if (value != 0 && value != 1) return fail(PP_ERR_ARG, "merge defer must be 0 or 1");
        h->tun.merge_defer = (int)value;
        return PP_OK;
    case PP_TUNE_SPLIT_BELOW:
        if (value < 0 || value > (int64_t)1 << 31) return fail(PP_ERR_ARG, "split_below must be 0..2^31 terms");
        h->tun.split_below = value;
        return PP_OK;
    case PP_TUNE_PERM_ITEMS:
        if (value != 1 && value != 2 && value != 4) return fail(PP_ERR_ARG, "permutation items must be 1, 2 or 4");
        h->tun.perm_items = (int)value;
        return PP_OK;
    case PP_TUNE_BATCH_ROWS:
        if (value != 0 && !(value >= 16 && value <= 256 && (value & (value - 1)) == 0))
            return fail(PP_ERR_ARG, "batch rows must be a power of two in 16..256 (0: by the batch)");
        if ((size_t)value * h->B * sizeof(double) > MAX_ROWS_LDS)
            return fail(PP_ERR_ARG, "batch rows x columns exceed the merge's shared memory");
        h->tun.batch_rows = (int)value;
        return PP_OK;
    case PP_TUNE_RADIX_BITS:
        if (value < 4 || value > 8) return fail(PP_ERR_ARG, "radix digits must be 4..8 bits");
        h->tun.radix_bits = (int)value;
        return PP_OK;
    case PP_TUNE_GRID_PER_CU:
        if (value < 1 || value > MAX_GRID_PER_CU) return fail(PP_ERR_ARG, "grid per compute unit must be 1..8");
        h->tun.grid_per_cu = (int)value;
        h->gmax = (unsigned)std::max(64, h->tun.grid_per_cu * h->cus);
        return PP_OK;
    case PP_TUNE_RUN_CHUNK:
        if (value < 16 || value > 1024) return fail(PP_ERR_ARG, "run chunk must be 16..1024 gates");
        h->tun.run_chunk = (int)value;
        return PP_OK;
    default:
        return fail(PP_ERR_ARG, "unknown tuning key");
    }
}

int pp_load_tuning(pp_handle* h, const char* path) {
    if (!h) return fail(PP_ERR_ARG, "handle is NULL");
    std::string file;
    if (path) {
        file = path;
    } else if (const char* env = std::getenv("PPHIP_TUNE_FILE")) {
        file = env;
    } else {
        // pphip_tune.txt beside the library, if there is one
        Dl_info info{};
        if (dladdr((const void*)&pp_version, &info) && info.dli_fname) {
            file = info.dli_fname;
            const size_t slash = file.rfind('/');
            file = (slash == std::string::npos ? std::string(".") : file.substr(0, slash)) + "/pphip_tune.txt";
        }
        FILE* probe = file.empty() ? nullptr : std::fopen(file.c_str(), "r");
        if (!probe) return PP_OK;
        std::fclose(probe);
    }
    FILE* f = std::fopen(file.c_str(), "r");
    if (!f) return fail(PP_ERR_ARG, "cannot open the tuning table " + file);
    struct Entry {
        int w, bmax;
        std::string name;
        long long value;
    };
    std::vector<Entry> es;
    char line[512];
    int lineno = 0;
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        const char* p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '#' || *p == '\n' || *p == '\0') continue;
        int w = 0, bmax = 0;
        char name[64] = {0};
        long long value = 0;
        if (std::sscanf(p, "%d %d %63s %lld", &w, &bmax, name, &value) != 4) {
            std::fclose(f);
            return fail(PP_ERR_ARG, file + ":" + std::to_string(lineno) + ": expected <key_words> <max_batch> <name> <value>");
        }
        es.push_back({w, bmax, name, value});
    }
    std::fclose(f);
    // the class of this handle: its key width and the smallest batch bound that covers B
    int best = -1;
    for (const Entry& e : es)
        if (e.w == h->W && e.bmax >= h->B && (best < 0 || e.bmax < best)) best = e.bmax;
    if (best < 0) return PP_OK;
    static const struct {
        const char* name;
        int key;
    } names[] = {{"merge_items", PP_TUNE_MERGE_ITEMS}, {"merge_defer", PP_TUNE_MERGE_DEFER},
                 {"split_below", PP_TUNE_SPLIT_BELOW},
                 {"perm_items", PP_TUNE_PERM_ITEMS},
                 {"batch_rows", PP_TUNE_BATCH_ROWS},   {"radix_bits", PP_TUNE_RADIX_BITS},
                 {"grid_per_cu", PP_TUNE_GRID_PER_CU}, {"run_chunk", PP_TUNE_RUN_CHUNK}};
    for (const Entry& e : es) {
        if (e.w != h->W || e.bmax != best) continue;
        int key = 0;
        for (const auto& n : names)
            if (e.name == n.name) key = n.key;
        if (!key) return fail(PP_ERR_ARG, file + ": unknown tuning name " + e.name);
        if (int rc = pp_tune(h, key, e.value)) return rc;
    }
    return PP_OK;
}

int pp_profile_read(const pp_handle* h, pp_profile* out) {
    if (!h || !out) return fail(PP_ERR_ARG, "NULL argument");
    out->gates = h->gates;
    out->fallbacks = h->fallbacks;
    out->term_applications = h->term_apps;
    out->identity_gates = h->identity_gates;
    out->persistent_gates = h->lds_gates;
    out->graph_launches = h->graph_launches;
    out->graph_captures = h->graph_captures;
    for (int k = 0; k < PP_KERNEL_KINDS; ++k) {
        out->launches[k] = h->launches[k];
        out->milliseconds[k] = h->ms[k];
        out->algorithmic_bytes[k] = h->alg_bytes[k];
    }
    return PP_OK;
}

}  // extern "C"
