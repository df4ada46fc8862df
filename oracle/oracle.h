/*
 * oracle.h — CPU restatement of the Stwo CpuBackend semantics for the
 * backend hot path (fields, circle domains, twiddle tree, circle FFT,
 * Blake2s Merkle trees, FRI folds, DEEP quotients, accumulation,
 * barycentric evaluation, proof-of-work grinding, Blake2s channel).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP backend is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * Conventions (all words canonical M31 values in [0, P)):
 *   - QM31 values are 4 words [a.0, a.1, b.0, b.1]
 *     (crates/stwo/src/core/fields/qm31.rs:42-44).
 *   - A QM31 circle point is 8 words: x[4] then y[4].
 *   - A column of QM31 values is 4 coordinate arrays ("by coords",
 *     crates/stwo/src/prover/secure_column.rs:13-63).
 *   - Cosets are (initial_index, log_size) with step = 2^(31 - log_size)
 *     (crates/stwo/src/core/circle.rs:260-272).
 *   - Evaluations are stored in bit-reversed order of the circle domain
 *     (crates/stwo/src/core/utils.rs:104-111).
 */
#ifndef STWO_ORACLE_H
#define STWO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 0x7fffffffu

/* ---- fields (crates/stwo/src/core/fields/{m31,cm31,qm31}.rs) ---- */
uint32_t or_m31_mul(uint32_t a, uint32_t b);
uint32_t or_m31_add(uint32_t a, uint32_t b);
uint32_t or_m31_reduce(uint64_t v);
uint32_t or_m31_inv(uint32_t a);
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]);

/* ---- circle (crates/stwo/src/core/circle.rs) ---- */
void or_index_to_point(uint32_t index, uint32_t out[2]);
void or_circle_domain_at(uint32_t half_initial_index, uint32_t half_log_size, uint64_t i,
                         uint32_t out[2]);

/* ---- twiddle tree (crates/stwo/src/prover/backend/cpu/circle.rs:240-285) ---- */
/* tw and itw receive 2^log_size words each. */
void or_precompute_twiddles(uint32_t initial_index, uint32_t log_size, uint32_t *tw,
                            uint32_t *itw);
/* crates/stwo/src/prover/poly/twiddles.rs:40-76; out has 2^(sub_log-1) words */
void or_extract_subdomain_twiddles(const uint32_t *buf, uint64_t buf_len, uint32_t domain_log,
                                   uint32_t sub_log, uint32_t *out);

/* ---- circle FFT (crates/stwo/src/prover/backend/cpu/circle.rs:26-79, 178-238) ----
 * The domain is CircleDomain(Coset(half_initial, half_log)), of log size half_log + 1.
 * twiddle buffers: tw/itw of tw_len words (the tree of the root coset). */
void or_interpolate(const uint32_t *itw, uint64_t tw_len, uint32_t half_initial,
                    uint32_t half_log, uint32_t *values);
void or_evaluate(const uint32_t *tw, uint64_t tw_len, uint32_t half_initial, uint32_t half_log,
                 const uint32_t *coeffs, uint32_t coeff_log, uint32_t *out);
/* crates/stwo/src/prover/backend/cpu/circle.rs:81-98 */
void or_eval_at_point(const uint32_t *coeffs, uint32_t log_size, const uint32_t point[8],
                      uint32_t out[4]);

/* ---- barycentric (crates/stwo/src/prover/backend/cpu/circle.rs:100-140) ---- */
void or_barycentric_weights(uint32_t log_size, const uint32_t point[8], uint32_t *w0,
                            uint32_t *w1, uint32_t *w2, uint32_t *w3);
void or_barycentric_eval(const uint32_t *evals, uint32_t log_size, const uint32_t *w0,
                         const uint32_t *w1, const uint32_t *w2, const uint32_t *w3,
                         uint32_t out[4]);

/* ---- Blake2s (RFC 7693; crates/stwo/src/core/vcs/blake2_hash.rs) ---- */
void or_blake2s(const uint8_t *data, uint64_t len, uint8_t out[32]);

/* ---- Merkle trees ---- */
/* crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:46-90. columns sorted ascending. */
void or_merkle_build_leaves(const uint32_t *const *cols, const uint32_t *log_sizes, uint64_t n,
                            uint32_t lifting_log_size, uint8_t *out);
/* crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:92-97 */
void or_merkle_next_layer(const uint8_t *prev, uint32_t prev_log_size, uint8_t *out);
/* crates/stwo/src/prover/backend/cpu/blake2s.rs:11-24 (prev may be NULL) */
void or_commit_on_layer(uint32_t log_size, const uint8_t *prev, const uint32_t *const *cols,
                        uint64_t n, uint8_t *out);

/* ---- FRI folds (crates/stwo/src/prover/backend/cpu/fri.rs:66-116) ---- */
/* line domain = Coset(initial_index, log_size) x-coordinates; src/dst by coords. */
void or_fold_line(const uint32_t *const src[4], uint32_t line_initial, uint32_t log_size,
                  const uint32_t alpha[4], uint32_t *const dst[4]);
void or_fold_circle_into_line(const uint32_t *const src[4], uint32_t half_initial,
                              uint32_t half_log, const uint32_t alpha[4], uint32_t *const dst[4]);

/* ---- DEEP quotients (crates/stwo/src/prover/backend/cpu/quotients.rs) ----
 * One sample batch = one point; for its k-th column term: col_idx[k], a/b/c line coefficients
 * already multiplied by the random power (crates/stwo/src/core/constraints.rs:119-134).
 * line_coeffs: for each term 12 words (a[4], b[4], c[4]).
 * out: 4 coordinate arrays of length 2^(col_log - log_blowup). first_linear: sum of a. */
void or_line_coeffs(const uint32_t point[8], const uint32_t value[4], const uint32_t alpha[4],
                    uint32_t out[12]);
void or_accumulate_numerators(const uint32_t *const *cols, uint32_t col_log, uint32_t log_blowup,
                              const uint32_t *col_idx, const uint32_t *line_coeffs,
                              uint64_t n_terms, uint32_t *const out[4],
                              uint32_t first_linear[4]);
/* partials: per accumulation 4 coord arrays (flattened: partials[4*k + c]); logs[k] their log
 * sizes; points 8 words each; first_linear 4 words each.  tw/itw: tree of the root coset.
 * out: 4 coordinate arrays of 2^lifting_log_size. */
void or_compute_quotients_and_combine(const uint32_t *const *partials, const uint32_t *logs,
                                      const uint32_t *points, const uint32_t *first_linear,
                                      uint64_t n_acc, uint32_t lifting_log_size,
                                      uint32_t log_blowup, const uint32_t *tw,
                                      const uint32_t *itw, uint64_t tw_len,
                                      uint32_t *const out[4]);

/* ---- accumulation (crates/stwo/src/prover/backend/cpu/accumulation.rs:8-47) ---- */
/* cols: n QM31 columns (4 coord arrays each, flattened), strictly ascending log sizes. */
void or_lift_and_accumulate(const uint32_t *const *cols, const uint32_t *logs, uint64_t n,
                            uint32_t *const out[4]);

/* ---- proof of work (crates/stwo/src/prover/backend/cpu/grind.rs:5-15) ---- */
int or_verify_pow_nonce(const uint8_t digest[32], uint32_t n_bits, uint64_t nonce);
uint64_t or_grind(const uint8_t digest[32], uint32_t n_bits);

/* ---- Blake2s channel (crates/stwo/src/core/channel/blake2s.rs) ---- */
typedef struct {
  uint8_t digest[32];
  uint32_t n_draws;
} or_channel;
void or_channel_mix_u32s(or_channel *ch, const uint32_t *words, uint64_t n);
void or_channel_mix_u64(or_channel *ch, uint64_t v);
void or_channel_mix_felts(or_channel *ch, const uint32_t *felts, uint64_t n_felts);
void or_channel_mix_hash(or_channel *ch, const uint8_t root[32]);
void or_channel_draw_u32s(or_channel *ch, uint32_t out[8]);
void or_channel_draw_secure_felt(or_channel *ch, uint32_t out[4]);
void or_channel_random_point(or_channel *ch, uint32_t out[8]);

#ifdef __cplusplus
}
#endif
#endif
