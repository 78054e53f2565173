/*
 * pphip.h -- C ABI of libpphip.so, the MI355X (gfx950) backend for the
 * propagate() hot path of PauliPropagation.jl v0.9.2.
 *
 * The reference has no FFI for this path: it dispatches in Julia on the
 * propagation cache type. Each entry point below replaces one Julia method
 * of that dispatch (paths relative to the reference checkout); a Julia host
 * package binds them with `ccall`.
 *
 * Conventions
 *   - A Pauli string on n qubits is W = 1, 2 or 4 little-endian uint64 limbs
 *     (n <= 32, 64, 128), qubit q (0-based) at bits 2q, 2q+1, with I=0, X=1,
 *     Y=2, Z=3 (src/PauliAlgebra/bitoperations.jl:154, utils.jl:2). Host key
 *     arrays are row-major: term i occupies keys[i*W .. i*W+W-1].
 *   - Coefficients are FP64, row-major n x B (B = batch columns sharing one
 *     key set; B = 1 is the reference's scalar case).
 *   - max_weight < 0 means no weight truncation (the reference's Inf);
 *     min_abs_coeff is compared against max_b |c_b| (|c| for B = 1).
 *   - Every function returns PP_OK (0) or a pp_status code; pp_last_error()
 *     returns the calling thread's message for the last failure.
 *   - A handle owns all device memory and one HIP stream. Host pointers are
 *     only borrowed for the duration of a call. Calls on one handle must not
 *     overlap (as the reference's owner-only workers,
 *     src/Base/threading_utils.jl:111-122); different handles are independent.
 */
#ifndef PPHIP_H
#define PPHIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pp_handle pp_handle;

enum pp_status {
    PP_OK = 0,
    PP_ERR_ARG = 1,          /* invalid argument (qubit out of range, bad noise strength, ...) */
    PP_ERR_NOMEM = 2,        /* device allocation failed */
    PP_ERR_HIP = 3,          /* HIP runtime error */
    PP_ERR_UNSUPPORTED = 4,  /* gate or option this backend does not implement */
    PP_ERR_CAPACITY = 5,     /* more terms than the backend's index range (2^31 - 1) */
    PP_ERR_STATE = 6         /* call not valid in the handle's current state */
};

enum pp_gate_kind {
    PP_GATE_ROTATION = 0,    /* PauliRotation: mask = generator */
    PP_GATE_PAULI_NOISE = 1, /* PauliNoise family: damp_codes on qubit */
    PP_GATE_AMP_DAMPING = 2, /* AmplitudeDampingNoise on qubit */
    PP_GATE_CLIFFORD = 3,    /* CliffordGate: table `lut` on qubits[0..nq) */
    PP_GATE_IMAG_ROTATION = 4 /* ImaginaryPauliRotation exp(-tau/2 P): mask, parameter tau, flags bit 0 */
};

/* One gate of a Heisenberg-ordered gate list. Parametrized gates read
 * theta[param * B + b]; param < 0 uses `value` (a FrozenGate). */
typedef struct pp_gate {
    uint64_t mask[4];     /* rotation generator, limb 0 first (unused limbs 0) */
    int32_t kind;         /* enum pp_gate_kind */
    int32_t param;        /* parameter row, or -1 */
    double value;         /* frozen parameter */
    int32_t qubit;        /* noise / damping qubit, 0-based */
    uint32_t damp_codes;  /* PauliNoise: bit v set <=> Pauli code v is damped */
    int32_t nq;           /* Clifford: number of qubits (1 or 2) */
    int32_t qubits[2];    /* Clifford qubits, 0-based; qubits[0] is the table's low code */
    int32_t lut;          /* Clifford: table slot */
    int32_t flags;        /* ImaginaryPauliRotation: bit 0 = normalize by the identity's coefficient */
} pp_gate;

typedef struct pp_gate_stats {
    int64_t n_in;         /* live terms before the gate */
    int64_t n_touched;    /* anticommuting (rotation) / touched terms */
    int64_t n_out;        /* live terms after the gate */
} pp_gate_stats;

/* Library version (major * 10000 + minor * 100 + patch). */
int pp_version(void);

/* Message for the calling thread's last failed call ("" if none). */
const char *pp_last_error(void);

/* Create a device Pauli sum for `nqubits` qubits with key_words limbs
 * (0 = smallest that fits), `batch` coefficient columns and room for
 * `capacity` terms (it grows on demand).
 * Replaces: PropagationCache(::VectorPauliSum) -> VectorPauliPropagationCache
 *           (src/Base/propagationcache.jl:9-10, src/Propagation/propagationcache.jl:92-101). */
int pp_create(int nqubits, int key_words, int batch, size_t capacity, pp_handle **out);

/* Free every device buffer of the handle. */
int pp_destroy(pp_handle *h);

/* A new handle (its own stream and buffers) holding a copy of h's sum, with
 * the same qubits, key width and batch, room for its live terms, and h's
 * tuning settings. The two are independent afterwards.
 * Replaces: deepcopy of the sum that propagate and rewindgradient take
 *           (src/Base/propagate.jl:3, src/gradient.jl:27). */
int pp_clone(pp_handle *h, pp_handle **out);

/* nqubits, key_words, batch, live terms, capacity (any pointer may be NULL). */
int pp_info(pp_handle *h, int *nqubits, int *key_words, int *batch, size_t *n, size_t *capacity);

/* Replace the sum with n distinct terms (any order; duplicates are an error).
 * Replaces: VectorPauliSum(::PauliSum) + cache construction
 *           (src/PauliDataTypes/vectorpaulisum.jl:21-49). */
int pp_upload(pp_handle *h, const uint64_t *keys, const double *coeffs, size_t n);

/* Replace the sum with n terms in any order, equal strings allowed: sorted on
 * the device (stable LSD radix sort, limb by limb), every run of equal strings
 * summed in input order, then truncated.
 * Replaces: mergeandtruncate!(cache) on an unsorted vector, sortterms! +
 *           _deduplicateandtruncatecpu! (src/Base/Merge/mergeandtruncate.jl:87-120,
 *           src/Base/Merge/deduplicate.jl:14-104). */
int pp_load_merge(pp_handle *h, const uint64_t *keys, const double *coeffs, size_t n, int max_weight,
                  double min_abs_coeff);

/* Copy the live terms, ascending by key, into caller memory with room for
 * `cap` terms; *n receives the term count (also when cap is too small, in
 * which case nothing is copied and PP_ERR_ARG is returned).
 * Replaces: extractsum!(cache, sum) (src/Base/propagationcache.jl:293-328). */
int pp_download(pp_handle *h, uint64_t *keys, double *coeffs, size_t cap, size_t *n);

/* PauliRotation step: branch, merge and truncate in one call.
 * cos_b / sin_b are B host values (computed by the caller, as the reference
 * does once per gate, src/Propagation/specializations.jl:30).
 * Replaces: applymergetruncate!(::PauliRotation, cache, theta; max_weight, min_abs_coeff)
 *           (src/Propagation/specializations.jl:53, 365-388). */
int pp_apply_rotation(pp_handle *h, const uint64_t *mask, const double *cos_b, const double *sin_b,
                      int max_weight, double min_abs_coeff, pp_gate_stats *stats);

/* ImaginaryPauliRotation step (Schrodinger picture): terms that commute with
 * the generator keep c cosh and branch into (c sinh) sign at term ^ mask; the
 * sum is merged; with `normalize` every coefficient is divided by (multiplied
 * by 1 / ) the identity's coefficient; every term is then truncated.
 * cosh_b / sinh_b are B host values of cosh(tau), sinh(tau). n_touched counts
 * the commuting terms.
 * Replaces: applymergetruncate!(::ImaginaryPauliRotation, cache, tau;
 *           normalize_coeffs) (src/Propagation/specializations.jl:86-133). */
int pp_apply_imaginary_rotation(pp_handle *h, const uint64_t *mask, const double *cosh_b, const double *sinh_b,
                                int normalize, int max_weight, double min_abs_coeff, pp_gate_stats *stats);

/* PauliNoise step: c *= factor_b[b] for the Pauli codes set in damp_codes
 * (bit v <=> code v) on `qubit`, then truncate every term.
 * Replaces: applymergetruncate!(::PauliNoise, cache, lambda; ...) with factor = 1 - lambda
 *           (src/Propagation/specializations.jl:218-238). */
int pp_apply_pauli_noise(pp_handle *h, int qubit, uint32_t damp_codes, const double *factor_b,
                         int max_weight, double min_abs_coeff, pp_gate_stats *stats);

/* AmplitudeDampingNoise step on `qubit` (gamma_b: B values, equal across
 * columns): X and Y scale by sqrt(1 - gamma); Z keeps 1 - gamma and branches
 * into I with gamma; then every output is truncated (no truncation when no
 * term has X, Y or Z there).
 * Replaces: applymergetruncate!(::AmplitudeDampingNoise, cache, gamma; ...)
 *           (src/Propagation/specializations.jl:250-285). */
int pp_apply_amplitude_damping(pp_handle *h, int qubit, const double *gamma_b, int max_weight,
                               double min_abs_coeff, pp_gate_stats *stats);

/* CliffordGate step on nq = 1 or 2 qubits: a term whose Pauli codes there form
 * idx = code(qubits[0]) | code(qubits[1]) << 2 becomes term ^ changes[idx]
 * (16 entries of W limbs) with its coefficient times signs[idx]; terms are
 * truncated only when max_weight >= 0 (then by weight and min_abs_coeff); the
 * sum is sorted again on the device.
 * Replaces: applymergetruncate!(::CliffordGate, cache; ...)
 *           (src/Propagation/specializations.jl:144-192, cliffordgates.jl:176-182). */
int pp_apply_clifford(pp_handle *h, int nq, const int *qubits, const uint64_t *changes, const double *signs,
                      int max_weight, double min_abs_coeff, pp_gate_stats *stats);

/* Run a Heisenberg-ordered gate list (the reference's _propagate! loop over
 * applymergetruncate!, src/Base/propagate.jl:44-71). theta is nparams x B
 * row-major. Clifford tables are nluts slots of 16 entries: lut_changes
 * (16 x W limbs per slot) and lut_signs (16 per slot). counts, if not NULL,
 * receives n_in of every gate (ngates entries; @countpaulis,
 * src/Base/countterms.jl:41-66). */
int pp_run_circuit(pp_handle *h, const pp_gate *gates, size_t ngates, const double *theta, size_t nparams,
                   const uint64_t *lut_changes, const double *lut_signs, size_t nluts, int max_weight,
                   double min_abs_coeff, int64_t *counts);

/* Loss and gradient by the adjoint sweep: loss_b = the overlap `overlap_kind`
 * (enum pp_overlap_kind; `ones` as for pp_overlap) of column b after the
 * Heisenberg-ordered gate list (as pp_run_circuit), and grad[p * B + b] =
 * d loss_b / d theta[p * B + b]. The forward run records, per gate, where
 * every input term went and the touched terms' coefficients; the sweep then
 * applies the transpose of each gate's linear map to the coefficient adjoints,
 * last gate first, with every truncation held as it happened: a rotation's
 * 1 / cos / +-sin branch, a Pauli noise gate's damping factors, a Clifford's
 * signed permutation. It is the exact derivative of the truncated forward map
 * (wherever no coefficient crosses the threshold), which is what central
 * differences of pp_run_circuit measure. checkpoint_gates = 0 records the
 * whole run; k > 0 records k gates at a time, each segment run again from a
 * snapshot of its input sum (one extra forward run), so records need room
 * for k gates only. On return the handle holds the propagated sum.
 * Amplitude damping and imaginary rotations are not taken (PP_ERR_UNSUPPORTED).
 * Counterpart: rewindgradient(circuit, psum, params, overlapfunc; kwargs...)
 *   (src/gradient.jl:26-108) computes the gradient of the same loss, equal to
 *   this one when nothing is truncated; under truncation it differs, since it
 *   truncates the operator again while undoing the gates. pp_rewind is that
 *   algorithm. The reference rejects noise; here noise is allowed. */
int pp_adjoint(pp_handle *h, const pp_gate *gates, size_t ngates, const double *theta, size_t nparams,
               const uint64_t *lut_changes, const double *lut_signs, size_t nluts, int max_weight,
               double min_abs_coeff, size_t checkpoint_gates, int overlap_kind, const uint64_t *ones, double *loss,
               double *grad);

/* The reference gradient, rewindgradient!(circuit, psum, params, overlapfunc;
 * max_weight, min_abs_coeff) (src/gradient.jl:40-108), on a B = 1 handle: the
 * Heisenberg-ordered gate list (rotations, frozen rotations, Cliffords; noise
 * is rejected with PP_ERR_UNSUPPORTED, as the reference rejects it) runs
 * forward with truncation and *loss receives the overlap `overlap_kind` of the
 * result; a dual sum starts on the result's strings with each string's own
 * overlap; then the gates are undone last to first in their Schrodinger form
 * (rotation angle negated, Clifford table transposed), the operator truncated
 * as in the forward run and the dual not, the dual capped to the operator's
 * strings after every gate, and for every parametrized rotation, before it is
 * undone, grad[param] = sum over operator terms t anticommuting with its
 * generator P of ((e & 2) - 1) c_t d(P xor t), P t = i^e (P xor t). Under
 * truncation this differs from pp_adjoint's gradient (which holds the forward
 * run's truncations fixed); without truncation both are exact. On return the
 * handle holds the operator after the undo sweep, as the reference's cache. */
int pp_rewind(pp_handle *h, const pp_gate *gates, size_t ngates, const double *theta, size_t nparams,
              const uint64_t *lut_changes, const double *lut_signs, size_t nluts, int max_weight,
              double min_abs_coeff, int overlap_kind, const uint64_t *ones, double *loss, double *grad);

/* overlapwithzero for every column: sum of the coefficients of terms with no
 * X or Y, accumulated in double-double and rounded once.
 * Replaces: overlapwithzero(::VectorPauliSum) (src/stateoverlap.jl:12-35). */
int pp_overlap_zero(pp_handle *h, double *out_b);

/* Product states of pp_overlap. */
enum pp_overlap_kind {
    PP_OVERLAP_ZERO = 0,          /* |0...0>: overlapwithzero (src/stateoverlap.jl:35) */
    PP_OVERLAP_PLUS = 1,          /* |+...+>: overlapwithplus (src/stateoverlap.jl:44) */
    PP_OVERLAP_COMPUTATIONAL = 2, /* a computational basis state: overlapwithcomputational (src/stateoverlap.jl:59-80) */
    PP_OVERLAP_MAXMIXED = 3       /* the maximally mixed state: overlapwithmaxmixed (src/stateoverlap.jl:87-93) */
};

/* Overlap of the sum with a product state, for every column: the sum over
 * terms of c times the term's expectation in the state (0 or +-1), in
 * double-double. PP_OVERLAP_COMPUTATIONAL reads `ones` (key_words limbs, the
 * key layout): bit 2q set for every qubit q (0-based) in state |1>; terms with
 * X or Y vanish and the others count (-1)^(number of Z on those qubits).
 * `ones` may be NULL for the other kinds.
 * PP_OVERLAP_MAXMIXED reads the identity's coefficient.
 * Replaces: overlapwithplus / overlapwithcomputational (src/stateoverlap.jl:44,59-80),
 * overlapwithmaxmixed (:87-93) and overlapwithzero (:35) on a VectorPauliSum. */
int pp_overlap(pp_handle *h, int kind, const uint64_t *ones, double *out_b);

/* pp_overlap without any device allocation: the same kernels, with their
 * per-block partial sums and results in a static device array of the library
 * (about 1 MiB on each GPU it runs on, held until the process ends) instead
 * of memory the handle allocates on demand. At most 64 coefficient columns
 * (B > 64 is rejected with PP_ERR_ARG). Every call on one GPU shares that
 * array, which keeps the last call's contents: calls that overlap in time on
 * one GPU race (calls on one handle never overlap, as for every entry point).
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). */
int pp_overlap_noalloc(pp_handle *h, int kind, const uint64_t *ones, double *out_b);

/* The largest |c| of every column (0 for an empty sum).
 * Replaces: maxabscoeff(cache) (src/Base/Primitives/maxabscoeff.jl:7). */
int pp_max_abs_coeff(pp_handle *h, double *out_b);

/* Truncate every term: weight above max_weight (-1: none), or coefficient
 * below the threshold -- min_abs_coeff, or with min_rel_coeff >= 0 (B = 1)
 * max(min_rel_coeff * maxabscoeff, min_abs_coeff); batched sums drop a term
 * when max_b |c_b| is below it. *threshold_out (may be NULL) receives it.
 * Replaces: truncate!(cache; min_abs_coeff, max_weight, min_rel_coeff)
 *           (src/Propagation/generics.jl:192-250). */
int pp_truncate(pp_handle *h, int max_weight, double min_abs_coeff, double min_rel_coeff, double *threshold_out);

/* The B coefficients of one Pauli string (key_words limbs), zeros when it is
 * absent. overlapwithmaxmixed is pp_get_coeff of the identity (key 0).
 * Replaces: getcoeff(::VectorPauliSum, trm) (src/Base/termsum.jl:104-124) and
 * overlapwithmaxmixed (src/stateoverlap.jl:87-93). */
int pp_get_coeff(pp_handle *h, const uint64_t *key, double *out_b);

/* scalarproduct for every column: the sum, over Pauli strings present in both
 * sums, of the products of their coefficients (Tr[psum1 psum2] / 2^n).
 * Both handles must have the same qubit count and batch. The terms of h1 are
 * looked up in h2; the call waits for both handles' work.
 * Replaces: scalarproduct(psum1, psum2) (src/stateoverlap.jl:120-135);
 * overlapwithpaulisum(rho, psum) is 2^n times it (:102-104). */
int pp_scalar_product(pp_handle *h1, pp_handle *h2, double *out_b);

/* Resampling methods of pp_resample (src/Base/MonteCarlo/resample.jl). */
enum pp_resample_method {
    PP_RESAMPLE_SEMIDETERMINISTIC = 0, /* semideterministic_systematic_resample! (:139-170), the default */
    PP_RESAMPLE_SYSTEMATIC = 1,        /* systematic_resample! (:96-112), calibrated when `calibrate` */
    PP_RESAMPLE_MULTINOMIAL = 2        /* multinomial_resample! (:64-78), the default with squared */
};

/* Monte Carlo resampling of a B = 1 sum down to about target_size terms.
 * Every term lies on [0, total) on an interval as wide as its weight (|c|, or
 * c^2 with `squared`; zero for the terms the semideterministic method keeps as
 * they are, those with |c| > sum|c| / target_size), in key order. A term gets
 * the comb's teeth in its interval -- floor(position / spacing - offset)
 * counts them, `offset` the caller's uniform draw in [0, 1) -- or, for the
 * multinomial method, the draws (target_size uniforms in [0, 1), ascending,
 * times the total weight); its new coefficient is teeth * spacing with its
 * sign (the square root with `squared`); terms left at zero are dropped and
 * the others keep their order. The systematic method's spacing is total /
 * target_size, or with `calibrate` the Newton iteration toward (1 - rtol)
 * target_size - atol expected survivors. *spacing_out (may be NULL) receives
 * the spacing (the weight per draw). Random numbers come from the caller.
 * Replaces: resample!(cache, target_size; resample_func, squared)
 *           (src/Base/MonteCarlo/resample.jl:24-60), the step of mcpropagate!
 *           (src/Base/MonteCarlo/mcpropagate.jl:42-52). */
int pp_resample(pp_handle *h, int method, size_t target_size, int squared, double offset, const double *draws,
                int calibrate, double rtol, double atol, double *spacing_out);

/* Wait for all work on the handle's stream. */
int pp_synchronize(pp_handle *h);

/* Kernel kinds of the per-handle profile. */
enum pp_kernel_kind {
    PP_KERNEL_BRANCH = 0,        /* anticommuting terms, compacted in order */
    PP_KERNEL_TAIL_PERMUTE = 1,  /* tail into ascending order of term ^ mask, and the merge partition */
    PP_KERNEL_MERGE = 2,         /* merge with truncation */
    PP_KERNEL_NOISE = 3,         /* Pauli noise with truncation */
    PP_KERNEL_OVERLAP = 4,       /* overlapwithzero partial sums */
    PP_KERNEL_KINDS = 5
};

/* Counters since the last reset: gates applied, term-gate applications
 * (sum of n_in over those gates), and per kernel kind the launches, the time
 * between HIP events recorded around them on the handle's stream (only while
 * timing is enabled), and their algorithmic HBM bytes: one read and one write
 * of the live set for a merge, (K + 8B)(n_in + n_out). */
typedef struct pp_profile {
    int64_t gates;
    int64_t term_applications;
    int64_t fallbacks;           /* gates run again with atomic tile claims (see pp_tune) */
    int64_t launches[PP_KERNEL_KINDS];
    double milliseconds[PP_KERNEL_KINDS];
    double algorithmic_bytes[PP_KERNEL_KINDS];
    int64_t identity_gates;      /* rotations left out as touching no term (see pp_tune) */
    int64_t persistent_gates;    /* gates applied by shared-memory runs of many gates (see pp_tune) */
    int64_t graph_launches;      /* chunks launched as hipGraphs (see pp_tune), of which captured: */
    int64_t graph_captures;
} pp_profile;

/* Kernel timing: 0 off; 1 the merge and noise kernels of pp_run_circuit, each
 * stamping its first workgroup's start and last workgroup's end on the
 * device's constant-rate clock (no events, so the timed run is not slowed);
 * 2 HIP events around every kernel (the events' completion adds to each timed
 * dispatch); -1 leave as is; reset != 0 zeroes the counters. No counterpart in the reference; its analogue is the per-gate
 * term counter of @countpaulis (src/Base/countterms.jl:41-66). */
int pp_profile_control(pp_handle *h, int enable_timing, int reset);
int pp_profile_read(const pp_handle *h, pp_profile *out);

/* Tuning keys of pp_tune. The compaction kernels hand out tiles statically,
 * which needs every workgroup resident at once; the host sizes grids from the
 * kernels' occupancy. Every look-back wait is bounded by PP_TUNE_SPIN_LIMIT
 * polls; a gate whose wait runs out is run again with tiles claimed through an
 * atomic counter, which needs no residency (counted in pp_profile.fallbacks).
 * PP_TUNE_DYNAMIC_TILES != 0 claims tiles through the counter always.
 * PP_TUNE_LDS_RUNS != 0 (the default) lets pp_run_circuit apply the
 * rotations and Pauli noise gates of one coefficient column to a sum of at
 * most a few thousand terms (2944, 1984, 1152 for 64-, 128-, 256-bit keys, and
 * the capacity) in one launch of one workgroup that holds the sum in shared
 * memory, gate after gate, until the sum outgrows it (counted in
 * pp_profile.persistent_gates); 0 turns this off (results are identical).
 * PP_TUNE_GRAPHS != 0 captures every chunk of rotations and Pauli noise gates
 * pp_run_circuit enqueues in a hipGraph and replays it when the same chunk
 * comes again (same gates, angles, truncation, buffers and launch shapes: a
 * repeated propagation); growing the buffers drops the graphs (counted in
 * pp_profile.graph_launches; results are identical). Default 0.
 * PP_TUNE_FORCE_RERUN != 0 makes every look-back kernel with static tiles
 * report its wait as run out once it has finished, so that every gate is run
 * again with atomic tile claims (a test of that path; results are identical).
 * PP_TUNE_CHAIN_TAILS != 0 (the default) lets the merge of a rotation in
 * pp_run_circuit list the terms the next rotation touches, so that rotation
 * skips its own pass over the sum; 0 turns this off (results are identical).
 * PP_TUNE_SKIP_IDENTITY != 0 (the default) lets pp_run_circuit leave out the
 * rotations that touch no term, found from which Paulis occur on each qubit
 * (read from the device every 128 gates and widened gate by gate); they count
 * in pp_profile.identity_gates. 0 turns this off (results are identical).
 * PP_TUNE_ADJ_DEFER_BYTES bounds the device memory pp_adjoint may take to keep
 * every rotation's per-block partial gradients of a segment and sum them in
 * one launch (default 2 GiB); above it, or when that allocation fails, each
 * rotation sums its own (results are identical).
 * Launch shapes (any value gives identical results): PP_TUNE_MERGE_ITEMS the
 * one-column merge's tile of 256 x 2, 3 or 4 diagonals (0: 3 for 256-bit
 * keys, 4 otherwise); PP_TUNE_MERGE_DEFER != 0 makes that merge (outside
 * recorded runs) keep each tile's outputs in shared memory until it has merged
 * its next tile, and only then wait for the tile's prefix and store it (0 or
 * 1, default 0); PP_TUNE_SPLIT_BELOW the size (terms, at the last
 * synchronization) below which that merge searches its tiles' splits itself
 * instead of after a partition kernel (0..2^31, default 2^16); PP_TUNE_PERM_ITEMS the tail permutation's list entries
 * per thread (1, 2, 4; default 2); PP_TUNE_BATCH_ROWS the batched merge's term
 * rows per tile (a power of two 16..256 with rows x B x 8 bytes <= 64 KiB; 0:
 * by B); PP_TUNE_RADIX_BITS the radix sort's digit width (4..8, default 8);
 * PP_TUNE_GRID_PER_CU the persistent grids' workgroups per compute unit
 * (1..8, default 4); PP_TUNE_RUN_CHUNK the gates pp_run_circuit enqueues
 * between overflow checks (16..1024, default 128). */
enum pp_tune_key {
    PP_TUNE_SPIN_LIMIT = 1,
    PP_TUNE_DYNAMIC_TILES = 2,
    PP_TUNE_CHAIN_TAILS = 3,
    PP_TUNE_SKIP_IDENTITY = 4,
    PP_TUNE_ADJ_DEFER_BYTES = 5,
    PP_TUNE_MERGE_ITEMS = 6,
    PP_TUNE_PERM_ITEMS = 7,
    PP_TUNE_BATCH_ROWS = 8,
    PP_TUNE_RADIX_BITS = 9,
    PP_TUNE_GRID_PER_CU = 10,
    PP_TUNE_RUN_CHUNK = 11,
    PP_TUNE_FORCE_RERUN = 12,
    PP_TUNE_LDS_RUNS = 13,
    PP_TUNE_GRAPHS = 14,
    PP_TUNE_MERGE_DEFER = 15,
    PP_TUNE_SPLIT_BELOW = 16
};

/* Set a tuning value. No counterpart in the reference. */
int pp_tune(pp_handle *h, int key, int64_t value);

/* Apply a tuning table: lines "<key_words> <max_batch> <name> <value>" ('#'
 * starts a comment), name one of merge_items, merge_defer, split_below, perm_items, batch_rows,
 * radix_bits, grid_per_cu, run_chunk. The handle takes the lines of its key
 * width with the smallest max_batch not below its batch. path NULL: the file
 * named by the environment variable PPHIP_TUNE_FILE, else pphip_tune.txt in
 * the library's directory if it exists (pp_create applies that one). The
 * table is written by the autotuner (python -m pauliprop_hip.autotune). No
 * counterpart in the reference. */
int pp_load_tuning(pp_handle *h, const char *path);

#ifdef __cplusplus
}
#endif
#endif
