/*
 * ppo.h -- CPU oracle for the propagate() hot path of PauliPropagation.jl.
 *
 * TEST INFRASTRUCTURE. This is a plain-C restatement of the reference's
 * VectorPauliSum algorithm. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only as the checker or as the
 * timed host baseline; the product (libpphip.so) never links it.
 *
 * Key layout: a Pauli string on n qubits is W little-endian uint64 limbs
 * (W = 1, 2 or 4), qubit q (0-based) at bits 2q, 2q+1 with I=0, X=1, Y=2, Z=3
 * (src/PauliAlgebra/bitoperations.jl:154, utils.jl:2). Keys are stored
 * row-major (term i at keys[i*W .. i*W+W-1]); coefficients row-major n x B.
 */
#ifndef PPO_H
#define PPO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ppo_sum ppo_sum;

/* gate kinds of a gate list (same numbering as include/pphip.h) */
enum { PPO_ROTATION = 0, PPO_PAULI_NOISE = 1, PPO_AMP_DAMPING = 2, PPO_CLIFFORD = 3, PPO_IMAG_ROTATION = 4 };

typedef struct ppo_gate {
    uint64_t mask[4];     /* rotation generator / unused */
    int32_t kind;
    int32_t param;        /* index into theta rows, -1 = frozen */
    double value;         /* frozen parameter */
    int32_t qubit;        /* noise / damping qubit (0-based) */
    uint32_t damp_codes;  /* PauliNoise: bit v set <=> Pauli code v is damped */
    int32_t nq;           /* Clifford: 1 or 2 qubits */
    int32_t qubits[2];    /* Clifford qubits (0-based) */
    int32_t lut;          /* Clifford: index of its table */
    int32_t flags;        /* ImaginaryPauliRotation: bit 0 = normalize by the identity's coefficient */
} ppo_gate;

ppo_sum *ppo_create(int W, int B);
void ppo_destroy(ppo_sum *s);
/* Load n terms; they are sorted here. Returns -1 on a duplicate key. */
int ppo_set(ppo_sum *s, const uint64_t *keys, const double *coeffs, size_t n);
/* Load terms that may repeat: sort, sum equal keys in input order, truncate. */
int ppo_load_merge(ppo_sum *s, const uint64_t *keys, const double *coeffs, size_t n, int max_weight,
                   double min_abs);
size_t ppo_size(const ppo_sum *s);
int ppo_sorted(const ppo_sum *s);
void ppo_get(const ppo_sum *s, uint64_t *keys, double *coeffs);

/* PauliRotation with host cos/sin per batch column. Returns the number of
 * anticommuting terms (0 = nothing touched, nothing truncated). */
long ppo_rotation(ppo_sum *s, const uint64_t *mask, const double *cosv, const double *sinv,
                  int max_weight, double min_abs);
/* ImaginaryPauliRotation exp(-tau/2 P) . exp(-tau/2 P): terms commuting with P
 * branch into cosh and sinh parts; merge; optionally divide by the identity's
 * coefficient; truncate all. Returns the number of branched terms. */
long ppo_imaginary_rotation(ppo_sum *s, const uint64_t *mask, const double *coshv, const double *sinhv,
                            int normalize, int max_weight, double min_abs);
/* PauliNoise: damp the codes in damp_codes on `qubit` by factor[b]; truncate all. */
void ppo_pauli_noise(ppo_sum *s, int qubit, unsigned damp_codes, const double *factor,
                     int max_weight, double min_abs);
/* AmplitudeDampingNoise(gamma) on `qubit`. Returns the number of touched terms. */
long ppo_amplitude_damping(ppo_sum *s, int qubit, double gamma, int max_weight, double min_abs);
/* CliffordGate via a lookup of 4^nq entries: changes[idx*W..] XORed into the key,
 * coefficient multiplied by signs[idx]. Truncates only when max_weight >= 0. */
void ppo_clifford(ppo_sum *s, int nq, const int *qubits, const uint64_t *changes,
                  const double *signs, int max_weight, double min_abs);
/* sum of coefficients over terms without X or Y, in array order, per column */
void ppo_overlap_zero(const ppo_sum *s, double *out);
/* overlap with a product state, per column: kind 0 |0..0>, 1 |+..+>, 2 the
 * computational basis state with ones where bit 2q of `ones` is set, 3 the
 * maximally mixed state (the identity's coefficient) */
void ppo_overlap_kind(const ppo_sum *s, int kind, const uint64_t *ones, double *out);
/* Monte Carlo resampling of a B = 1 sum to about `target` terms: method 0
 * semideterministic systematic, 1 systematic (calibrated when `calibrate`),
 * 2 multinomial with `draws` (target uniforms, ascending); `offset` is the
 * comb offset in [0, 1). Returns the comb spacing (weight per draw), or -1. */
double ppo_resample(ppo_sum *s, int method, size_t target, int squared, double offset, const double *draws,
                    int calibrate, double rtol, double atol);
/* per-column largest |c| */
void ppo_max_abs(const ppo_sum *s, double *out);
/* truncate! with max(min_rel * maxabscoeff, min_abs) (min_rel < 0: min_abs); returns the threshold */
double ppo_truncate(ppo_sum *s, int max_weight, double min_abs, double min_rel);
/* the coefficients of `key` (zeros when absent) */
void ppo_get_coeff(const ppo_sum *s, const uint64_t *key, double *out);
/* sum over common strings of coefficient products; -1 on mismatched W or B */
int ppo_scalar_product(const ppo_sum *a, const ppo_sum *b, double *out);

/* Run a Heisenberg-ordered gate list. theta is P x B row-major. counts, if not
 * NULL, receives n_in for every gate (ngates entries). */
int ppo_run(ppo_sum *s, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
            const uint64_t *lut_changes, const double *lut_signs, int max_weight, double min_abs,
            int64_t *counts);

/* Reverse-mode gradient of overlap(propagate(gates, s0, theta)) for every
 * column, every truncation held as the forward run made it; every
 * intermediate sum is kept. The overlap is `kind` of ppo_overlap_kind (0..3).
 * grad is P x B row-major and is overwritten; loss (B entries) receives the
 * forward overlap. Rotations, PauliNoise and Cliffords (tables as ppo_run).
 * Returns 0, or -1 on an unsupported gate. */
int ppo_adjoint(const ppo_sum *s0, const ppo_gate *gates, size_t ngates, const double *theta,
                size_t nparams, const uint64_t *lut_changes, const double *lut_signs, int max_weight,
                double min_abs, int kind, const uint64_t *ones, double *loss, double *grad);

/* rewindgradient (src/gradient.jl:40-108), the reference gradient: forward
 * propagation with truncation, then an undo sweep that re-applies each gate's
 * Schrodinger form to the truncated operator and carries a dual sum. B = 1;
 * gates are rotations (parametrized or frozen) and Cliffords, in Heisenberg
 * order; grad (nparams entries, Heisenberg rows) is overwritten. kind: 0
 * overlapwithzero, 1 overlapwithplus, 2 overlapwithcomputational (`ones`), 3
 * overlapwithmaxmixed, 4 scale * scalarproduct with rho (overlapwithpaulisum
 * with scale 2^n). Returns 0, or -1 on an unsupported gate or argument. */
int ppo_rewind(const ppo_sum *s0, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
               const uint64_t *lut_changes, const double *lut_signs, int max_weight, double min_abs, int kind,
               const uint64_t *ones, const ppo_sum *rho, double scale, double *loss, double *grad);

/* bit-algebra helpers, exposed for the brute-force tests */
int ppo_anticommutes(const uint64_t *a, const uint64_t *b, int W);
int ppo_sign_exponent(const uint64_t *p1, const uint64_t *p2, int W);
int ppo_weight(const uint64_t *a, int W);
int ppo_contains_xy(const uint64_t *a, int W);
/* sort a tail of `n` keys (with coefficients) produced as term^mask of an
 * ascending run, by XOR block-reversal passes; returns the pass count. */
int ppo_xor_sort_tail(uint64_t *keys, double *coeffs, size_t n, int W, int B, const uint64_t *mask);

/* The zone-parallel restatement of MultiPauliSum (src/Base/MultiSum/): 2^zbits
 * zones by a GF(2)-linear hash, rotations branched and merged zone by zone on
 * nthreads OpenMP threads. It is the all-cores CPU baseline. */
typedef struct ppo_multi ppo_multi;
ppo_multi *ppo_multi_create(int W, int B, int zbits);
void ppo_multi_destroy(ppo_multi *m);
int ppo_multi_set(ppo_multi *m, const uint64_t *keys, const double *coeffs, size_t n);
size_t ppo_multi_size(const ppo_multi *m);
void ppo_multi_get(const ppo_multi *m, uint64_t *keys, double *coeffs);
int ppo_multi_run(ppo_multi *m, const ppo_gate *gates, size_t ngates, const double *theta, size_t nparams,
                  int max_weight, double min_abs, int nthreads, int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
