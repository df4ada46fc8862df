/*
 * stwo_hip.h — C-ABI of the MI355X (gfx950) backend for the Stwo Circle-STARK
 * prover.  Each entry point replaces one method of the reference's backend trait
 * family (crates/stwo/src/prover/backend/mod.rs:18-66); the comment above it
 * cites the trait method and the SimdBackend implementation it stands in for.
 *
 * Conventions
 *  - Every function returns SH_OK (0) or a negative SH_ERR_* code; the message
 *    of the last failure is available from sh_last_error().  The reference
 *    panics (assert!) where these return SH_ERR_INVALID.
 *  - Device buffers are plain device pointers (from sh_alloc or any HIP
 *    allocation); "host arrays of device pointers" are ordinary host arrays
 *    whose elements are device addresses.
 *  - M31 words are canonical, in [0, 2^31 - 1).  A QM31 column is 4 device
 *    arrays, one per coordinate (SecureColumnByCoords,
 *    crates/stwo/src/prover/secure_column.rs:13-63).  QM31 scalars are 4 words
 *    [a.0, a.1, b.0, b.1]; QM31 circle points are 8 words (x then y).
 *  - Evaluations are in bit-reversed order of the circle domain; coefficients
 *    are in the FFT basis order of CpuBackend (natural order).
 *  - Hashes are 32-byte Blake2s-256 digests, stored contiguously.
 *  - Work is enqueued on the context's stream; downloads and sh_sync block.
 *  - Contexts are independent; use one per host thread.
 */
#ifndef STWO_HIP_H
#define STWO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SH_OK 0
#define SH_ERR_INVALID (-1) /* bad sizes/arguments (a reference assert!) */
#define SH_ERR_OOM (-2)
#define SH_ERR_HIP (-3)
#define SH_ERR_NOT_FOUND (-4)

typedef struct sh_ctx sh_ctx;
typedef struct sh_twiddles sh_twiddles;

/* ---------------------------------------------------------------- context */
/* The reference backend is a zero-sized marker type
 * (crates/stwo/src/prover/backend/simd/mod.rs:38-45); a context carries the
 * device state the GPU needs instead (stream, staging ring, scratch). */
int sh_ctx_create(int device, sh_ctx **out);
int sh_ctx_destroy(sh_ctx *ctx);
const char *sh_last_error(const sh_ctx *ctx);
int sh_sync(sh_ctx *ctx);
/* Span timing by class, mirroring the reference's tracing classes
 * (crates/stwo/src/tracing/mod.rs:17-104): hipEvents around every launch. */
int sh_profile_enable(sh_ctx *ctx, int on);
/* Writes a JSON object {"class": {"n": launches, "ms": total}, ...} and resets. */
int sh_profile_report(sh_ctx *ctx, char *buf, uint64_t cap);
/* Stream-ordered wall clock of the context stream (hipEvent pair). */
int sh_timer_start(sh_ctx *ctx);
int sh_timer_stop(sh_ctx *ctx, float *ms);

/* ---------------------------------------------------------------- columns */
/* Column::zeros / uninitialized / to_cpu / from_iter
 * (crates/stwo/src/prover/backend/mod.rs:45-66). */
int sh_alloc(sh_ctx *ctx, uint64_t bytes, void **d_out);
int sh_free(sh_ctx *ctx, void *d);
/* sh_alloc/sh_free go through a per-context pool (BaseColumnPool,
 * crates/stwo/src/prover/mempool.rs:14-58): freed blocks are reused by later
 * allocations of the same rounded size, stream-ordered, without a device sync.
 * A buffer still read by work enqueued on this context may be freed; one that
 * another context or the host is using asynchronously must not be.  sh_free of an address the
 * context did not hand out (or handed out and took back already) releases nothing.
 * sh_pool_trim releases every cached block (synchronising the context first). */
int sh_pool_trim(sh_ctx *ctx);
int sh_upload(sh_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes);
int sh_download(sh_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes);
int sh_memset(sh_ctx *ctx, void *d_dst, int value, uint64_t bytes);
/* Column clone (device to device, stream-ordered). */
int sh_copy(sh_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes);
/* ColumnOps::bit_reverse_column (crates/stwo/src/prover/backend/mod.rs:37-43;
 * SIMD crates/stwo/src/prover/backend/simd/bit_reverse.rs:25-107). */
int sh_bit_reverse_column(sh_ctx *ctx, uint32_t *d_col, uint32_t log_size);

/* Tuning of the FFT pass shape for transforms of log size log_size (no
 * counterpart in the reference; the SIMD backend's split is fixed at compile
 * time): contiguous block pass of 2^block_log words (1..13), strided passes of
 * stride_layers layers (1..8), and the rows per lane of the block and the
 * strided passes, 2^r_block and 2^r_stride quads of 4 positions (1..4: more
 * rows per lane run more layers per LDS round trip, fewer rows put more lanes
 * on a small transform).  Defaults are the autotuner's picks on MI355X for the
 * benchmark's sizes.  Results never depend on it. */
int sh_set_fft_tuning(sh_ctx *ctx, uint32_t log_size, uint32_t block_log, uint32_t stride_layers,
                      uint32_t r_block, uint32_t r_stride);
int sh_get_fft_tuning(sh_ctx *ctx, uint32_t log_size, uint32_t out[4]);
/* Interpolation followed by extension of the same columns (sh_commit_trace;
 * no counterpart in the reference, whose SIMD backend reads the coefficients
 * back for each subdomain): mode 0 runs them as separate transforms; mode 1
 * (the default) runs the interpolation's last strided pass and the
 * extension's first as one pass when their pass plans give them the same
 * layers, so the coefficients are written once and never read back; mode 2
 * also plans the extension with the interpolation's layer split so that they
 * always do.  Results never depend on it. */
int sh_set_fft_fusion(sh_ctx *ctx, uint32_t mode);
/* Autotuner (no counterpart in the reference): for the column sizes of a
 * workload, times candidate pass shapes of every interpolation and
 * evaluation size it uses, on batches of the workload's width, and installs
 * the fastest in the context. */
int sh_autotune_fft(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *log_sizes, uint64_t n,
                    uint32_t log_blowup);

/* ---------------------------------------------------------------- PolyOps */
/* PolyOps::precompute_twiddles (crates/stwo/src/prover/poly/circle/ops.rs:143;
 * SIMD crates/stwo/src/prover/backend/simd/circle.rs:431-462).  Root coset is
 * Coset(root_initial_index, root_log_size); buffers hold 2^root_log_size words
 * in CpuBackend's layout (crates/stwo/src/prover/backend/cpu/circle.rs:274-285). */
int sh_twiddles_create(sh_ctx *ctx, uint32_t root_initial_index, uint32_t root_log_size,
                       sh_twiddles **out);
int sh_twiddles_destroy(sh_ctx *ctx, sh_twiddles *tw);
int sh_twiddles_download(sh_ctx *ctx, const sh_twiddles *tw, uint32_t *h_tw, uint32_t *h_itw);
/* PolyOps::interpolate (crates/stwo/src/prover/poly/circle/ops.rs:25-28;
 * SIMD crates/stwo/src/prover/backend/simd/circle.rs:135-163).  In place:
 * bit-reversed evaluations on CanonicCoset(log_size) -> coefficients. */
int sh_interpolate(sh_ctx *ctx, const sh_twiddles *tw, uint32_t *d_values, uint32_t log_size);
/* PolyOps::interpolate_columns (crates/stwo/src/prover/poly/circle/ops.rs:30-40):
 * many columns of uneven sizes in one call (batched launches per size class). */
int sh_interpolate_columns(sh_ctx *ctx, const sh_twiddles *tw, uint32_t *const *d_cols,
                           const uint32_t *log_sizes, uint64_t n);
/* PolyOps::evaluate / evaluate_into (crates/stwo/src/prover/poly/circle/ops.rs:94-107;
 * SIMD crates/stwo/src/prover/backend/simd/circle.rs:389-425): coefficients of
 * length 2^coeff_log -> bit-reversed evaluations on CanonicCoset(domain_log). */
int sh_evaluate(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *d_coeffs, uint32_t coeff_log,
                uint32_t *d_out, uint32_t domain_log);
/* PolyOps::evaluate_polynomials (crates/stwo/src/prover/poly/circle/ops.rs:109-140). */
int sh_evaluate_polynomials(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_coeffs,
                            const uint32_t *coeff_logs, uint64_t n, uint32_t log_blowup,
                            uint32_t *const *d_outs);

/* PolyOps::split_at_mid (crates/stwo/src/prover/poly/circle/ops.rs:145-165;
 * CpuBackend crates/stwo/src/prover/backend/cpu/circle.rs:266-271): the
 * 2^log_size coefficients split into halves, p = p_left + pi^(log_size-2)(x) * p_right;
 * d_left and d_right receive 2^(log_size-1) words each.  d_left may equal d_coeffs. */
int sh_split_at_mid(sh_ctx *ctx, const uint32_t *d_coeffs, uint32_t log_size, uint32_t *d_left,
                    uint32_t *d_right);

/* ---------------------------------------------------------------- Merkle */
/* MerkleOpsLifted::build_leaves (crates/stwo/src/prover/vcs_lifted/ops.rs:14-15;
 * SIMD crates/stwo/src/prover/backend/simd/blake2s_lifted.rs:46-229).  Columns
 * must be sorted ascending by size (as the trait requires); d_out receives
 * 2^lifting_log_size hashes (one hash when n == 0). */
int sh_merkle_build_leaves(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *log_sizes,
                           uint64_t n, uint32_t lifting_log_size, uint8_t *d_out);
/* MerkleOpsLifted::build_next_layer (crates/stwo/src/prover/vcs_lifted/ops.rs:19;
 * SIMD crates/stwo/src/prover/backend/simd/blake2s_lifted.rs:232-263). */
int sh_merkle_build_next_layer(sh_ctx *ctx, const uint8_t *d_prev, uint32_t prev_log_size,
                               uint8_t *d_out);
/* MerkleProverLifted::commit (crates/stwo/src/prover/vcs_lifted/prover.rs:35-80):
 * stable-sorts columns ascending, builds the leaves and every layer.
 * d_layers[k] (k = 0..lifting_log_size) receives the 2^k hashes of layer k
 * (layer 0 is the root); root (host) receives the root. */
int sh_merkle_commit_lifted(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *log_sizes,
                            uint64_t n, uint32_t lifting_log_size, uint8_t *const *d_layers,
                            uint8_t root[32]);
/* MerkleOps::commit_on_layer (crates/stwo/src/prover/vcs/ops.rs:24-29;
 * SIMD crates/stwo/src/prover/backend/simd/blake2s.rs:60-143): mixed-degree
 * node hash Blake2s(prev[2i] || prev[2i+1] || LE32(col[i]) ...); d_prev may be NULL. */
int sh_merkle_commit_on_layer(sh_ctx *ctx, uint32_t log_size, const uint8_t *d_prev,
                              const uint32_t *const *d_cols, uint64_t n, uint8_t *d_out);
/* MerkleProver::commit (crates/stwo/src/prover/vcs/prover.rs:39-59): columns
 * stable-sorted descending, commit_on_layer for log_size = max..0.
 * d_layers[k] receives layer k (2^k hashes). */
int sh_merkle_commit_mixed(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *log_sizes,
                           uint64_t n, uint8_t *const *d_layers, uint8_t root[32]);

/* ---------------------------------------------------------------- trace commit */
/* TreeBuilder::extend_evals + CommitmentTreeProver::new
 * (crates/stwo/src/prover/pcs/mod.rs:325-334, 360-386): interpolate every
 * column of bit-reversed evaluations d_evals[i] into coefficients d_coeffs[i]
 * (may alias d_evals[i]), evaluate each on CanonicCoset(log_size + log_blowup)
 * into d_ldes[i], and commit the lifted Merkle tree of the LDEs with lifting
 * size max(log_size) + log_blowup (d_layers[k] receives layer k, 2^k hashes).
 * Batched: one launch chain per size class for the FFTs, grouped leaf hashing
 * with hash-state sharing, fused inner layers.  root may be NULL (no sync). */
int sh_commit_trace(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_evals,
                    uint32_t *const *d_coeffs, const uint32_t *log_sizes, uint64_t n,
                    uint32_t log_blowup, uint32_t *const *d_ldes, uint8_t *const *d_layers,
                    uint8_t root[32]);
/* The same with the mixed-degree tree: TreeBuilder-style interpolate + LDE as above, then
 * MerkleProver::commit (crates/stwo/src/prover/vcs/prover.rs:39-59 over MerkleOps::
 * commit_on_layer, crates/stwo/src/prover/vcs/ops.rs:24-29), which injects each column at the
 * layer of its LDE size (d_layers[k] receives layer k, k = 0..max(log_size) + log_blowup).
 * Batched: the size classes' FFTs run largest first on a helper stream, and each column level of
 * the tree starts when its class is ready (one launch per column level, the column-free levels
 * above it reduced in the same launch).  root may be NULL (no sync). */
int sh_commit_trace_mixed(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_evals,
                          uint32_t *const *d_coeffs, const uint32_t *log_sizes, uint64_t n,
                          uint32_t log_blowup, uint32_t *const *d_ldes, uint8_t *const *d_layers,
                          uint8_t root[32]);

/* ---------------------------------------------------------------- OOD sampling */
/* PolyOps::eval_at_point (crates/stwo/src/prover/poly/circle/ops.rs:44-48;
 * CpuBackend crates/stwo/src/prover/backend/cpu/circle.rs:81-98): evaluates the
 * polynomial with 2^log_size coefficients at a QM31 circle point (8 words). */
int sh_eval_at_point(sh_ctx *ctx, const uint32_t *d_coeffs, uint32_t log_size, const uint32_t point[8],
                     uint32_t out[4]);
/* eval_at_point for many polynomials at many points in one call (compute_samples,
 * crates/stwo/src/prover/pcs/mod.rs:176-222, evaluates every column at its sampled points):
 * column i (2^log_sizes[i] coefficients) at point j (points[8j..8j+8]) goes to
 * out[4 * (i * n_points + j) ..] (host).  Each column is read once per 4 points. */
int sh_eval_at_points(sh_ctx *ctx, const uint32_t *const *d_coeffs, const uint32_t *log_sizes, uint64_t n_cols,
                      const uint32_t *points, uint32_t n_points, uint32_t *out);
/* compute_samples (crates/stwo/src/prover/pcs/mod.rs:176-222) at one point held in device memory
 * (d_point, 8 words a device channel drew): column i (2^log_sizes[i] coefficients) evaluated at
 * the point doubled doublings[i] times goes to d_values[4i..4i+4] (device).  No host wait. */
int sh_eval_at_dpoint(sh_ctx *ctx, const uint32_t *const *d_coeffs, const uint32_t *log_sizes,
                      const uint32_t *doublings, uint64_t n_cols, const uint32_t *d_point, uint32_t *d_values);
/* PolyOps::barycentric_weights_into (crates/stwo/src/prover/poly/circle/ops.rs:50-62;
 * SIMD crates/stwo/src/prover/backend/simd/circle.rs:165-300): weights for
 * CanonicCoset(log_size) and the point, written to 4 coordinate arrays. */
int sh_barycentric_weights(sh_ctx *ctx, uint32_t log_size, const uint32_t point[8], uint32_t *const d_w[4]);
/* PolyOps::barycentric_eval_at_point (crates/stwo/src/prover/poly/circle/ops.rs:64-70;
 * SIMD crates/stwo/src/prover/backend/simd/circle.rs:302-353). */
int sh_barycentric_eval_at_point(sh_ctx *ctx, const uint32_t *d_evals, uint32_t log_size,
                                 const uint32_t *const d_w[4], uint32_t out[4]);

/* ---------------------------------------------------------------- FriOps */
/* FriOps::fold_line (crates/stwo/src/prover/fri.rs:41-45; SIMD
 * crates/stwo/src/prover/backend/simd/fri.rs:32-109): src is a QM31 evaluation on
 * LineDomain(Coset(line_initial_index, log_size)) (must be in the twiddle tower);
 * alphas holds fold_step QM31 values (alpha, alpha^2, ...); dst receives
 * 2^(log_size - fold_step) values.  fold_step in [1, 4]. */
int sh_fold_line(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const src[4],
                 uint32_t line_initial_index, uint32_t log_size, const uint32_t *alphas,
                 uint32_t fold_step, uint32_t *const dst[4]);
/* FriOps::fold_circle_into_line (crates/stwo/src/prover/fri.rs:62-66; SIMD
 * crates/stwo/src/prover/backend/simd/fri.rs:111-166): src on CanonicCoset(log_size)
 * -> dst (2^(log_size-1) values) on LineDomain(half_odds(log_size - 1)). */
int sh_fold_circle_into_line(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const src[4],
                             uint32_t log_size, const uint32_t alpha[4], uint32_t *const dst[4]);
/* FriOps::decompose (crates/stwo/src/prover/fri.rs:68-75; CpuBackend
 * crates/stwo/src/prover/backend/cpu/fri.rs:38-62, 118-146). */
int sh_fri_decompose(sh_ctx *ctx, const uint32_t *const src[4], uint32_t log_size, uint32_t *const dst[4],
                     uint32_t lambda[4]);

/* ---------------------------------------------------------------- QuotientOps */
/* QuotientOps::accumulate_numerators (crates/stwo/src/prover/pcs/quotient_ops.rs:29-34;
 * SIMD crates/stwo/src/prover/backend/simd/quotients.rs:35-78): n_cols columns of
 * log col_log; sample batch b has point batch_points[8b..8b+8] and terms
 * [batch_term_offsets[b], batch_term_offsets[b+1]) with column index, sampled value
 * (4 words) and random coefficient (4 words) each, as ColumnSampleBatch holds them
 * (crates/stwo/src/core/pcs/quotients.rs:40-76).  Batch b's partial numerator goes to
 * d_outs[4b..4b+4] (2^(col_log - log_blowup) rows) and its summed a-coefficients
 * to first_linear[4b..4b+4] (host). */
int sh_accumulate_numerators(sh_ctx *ctx, const uint32_t *const *d_cols, uint64_t n_cols, uint32_t col_log,
                             uint32_t log_blowup, const uint32_t *batch_points,
                             const uint32_t *batch_term_offsets, const uint32_t *term_col_idx,
                             const uint32_t *term_values, const uint32_t *term_random_coeffs,
                             uint64_t n_batches, uint32_t *const *d_outs, uint32_t *first_linear);
/* QuotientOps::compute_quotients_and_combine (crates/stwo/src/prover/pcs/quotient_ops.rs:47-53;
 * SIMD crates/stwo/src/prover/backend/simd/quotients.rs:82-200): accumulation k has
 * partial numerator d_partials[4k..4k+4] of log partial_logs[k], sample point
 * points[8k..8k+8] and first linear term first_linear[4k..4k+4]; d_out receives the
 * combined quotient evaluation on CanonicCoset(lifting_log_size) (4 arrays). */
int sh_compute_quotients_and_combine(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_partials,
                                     const uint32_t *partial_logs, const uint32_t *points,
                                     const uint32_t *first_linear, uint64_t n_acc,
                                     uint32_t lifting_log_size, uint32_t log_blowup, uint32_t *const *d_out);
/* sh_compute_quotients_and_combine with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The subdomain quotient lives in a static device array
 * of 32 MiB per device (4 coordinate arrays of 2^21 words), so lifting_log_size - log_blowup must
 * be at most 21 (SH_ERR_INVALID otherwise).  Every call on one device shares that array:
 * calls that overlap on one device (other contexts or streams) race.  Same arguments and
 * results as sh_compute_quotients_and_combine. */
int sh_compute_quotients_and_combine_noalloc(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_partials,
                                             const uint32_t *partial_logs, const uint32_t *points,
                                             const uint32_t *first_linear, uint64_t n_acc,
                                             uint32_t lifting_log_size, uint32_t log_blowup,
                                             uint32_t *const *d_out);

/* ---------------------------------------------------------------- AccumulationOps */
/* AccumulationOps::lift_and_accumulate (crates/stwo/src/prover/air/accumulation.rs:198-200;
 * SIMD crates/stwo/src/prover/backend/simd/accumulation.rs:12-78): QM31 columns
 * d_cols[4k..4k+4] of strictly ascending log sizes -> d_out (size of the largest). */
int sh_lift_and_accumulate(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *log_sizes, uint64_t n,
                           uint32_t *const *d_out);
/* AccumulationOps::accumulate (crates/stwo/src/prover/air/accumulation.rs:184-186):
 * column += other. */
int sh_accumulate(sh_ctx *ctx, uint32_t *const *d_col, const uint32_t *const *d_other, uint32_t log_size);

/* ---------------------------------------------------------------- GrindOps */
/* GrindOps::grind (crates/stwo/src/core/proof_of_work.rs:10-14; SIMD
 * crates/stwo/src/prover/backend/simd/grind.rs:20-122) for the Blake2s channel:
 * the smallest nonce whose proof-of-work hash over the channel digest has at
 * least pow_bits trailing zero bits. */
int sh_grind(sh_ctx *ctx, const uint8_t digest[32], uint32_t pow_bits, uint64_t *nonce);
/* GrindOps::grind on a device channel (d_ch: 16 device words, stwo_prover.h) followed by
 * mix_u64(nonce) (crates/stwo/src/prover/mod.rs:74-81): the nonce lands as a little-endian u64 in
 * the first 8 bytes of d_ctl (16 device bytes); the host does not wait.  pow_bits <= 32; one
 * search over 2^(pow_bits + 12) nonces (a nonce reading as all ones means none was found). */
int sh_grind_dchannel(sh_ctx *ctx, uint32_t *d_ch, uint32_t pow_bits, uint32_t *d_ctl);

/* ---------------------------------------------------------------- composition */
/* ComponentProver::evaluate_constraint_quotients_on_domain
 * (crates/constraint-framework/src/prover/component_prover.rs:118-193) for the wide-Fibonacci
 * AIR (WideFibonacciEval, crates/examples/src/wide_fibonacci/mod.rs:77-98): n_cols trace
 * columns d_cols[i] (committed evaluations, at least 2^eval_log words each), constraints
 * c_{j+2} - (c_j^2 + c_{j+1}^2) for j = 0..n_cols-3 weighted by powers[4j..4j+4] (the
 * accumulator's random-coefficient powers in constraint order), divided by the trace coset's
 * vanishing polynomial, added into d_accum (4 arrays of 2^eval_log) on the evaluation domain
 * CanonicCoset(eval_log + log_expansion).circle_domain().split(log_expansion).0. */
int sh_wide_fibonacci_composition(sh_ctx *ctx, const uint32_t *const *d_cols, uint64_t n_cols,
                                  uint32_t trace_log, uint32_t eval_log, uint32_t log_expansion,
                                  const uint32_t *powers, uint32_t *const d_accum[4]);
/* sh_wide_fibonacci_composition with the powers in device memory (d_powers, 4 words each, as
 * sh_secure_powers_dev writes them). */
int sh_wide_fibonacci_composition_dev(sh_ctx *ctx, const uint32_t *const *d_cols, uint64_t n_cols,
                                      uint32_t trace_log, uint32_t eval_log, uint32_t log_expansion,
                                      const uint32_t *d_powers, uint32_t *const d_accum[4]);

/* The sanity check closing prove_ex (crates/stwo/src/prover/mod.rs:118-132) for the
 * wide-Fibonacci AIR, on device data: StarkProof::extract_composition_oods_eval
 * (crates/stwo/src/core/proof.rs:27-53) of the 8 composition samples against the constraints
 * evaluated on the n_cols trace samples at the OODS point d_point (8 words), the constraints
 * weighted by d_powers_rev[j] = alpha^(n_cols - 3 - j) and divided by the vanishing polynomial
 * of CanonicCoset(log_n).  d_values: the n_cols trace values then the 8 composition values.
 * d_out (9 device words): 1 if they differ (ProvingError::ConstraintsNotSatisfied) else 0, then
 * the two values. */
int sh_wide_fibonacci_oods_check(sh_ctx *ctx, const uint32_t *d_values, uint64_t n_cols,
                                 const uint32_t *d_powers_rev, const uint32_t *d_point, uint32_t log_n,
                                 uint32_t *d_out);

/* generate_trace of the wide-Fibonacci AIR (crates/examples/src/wide_fibonacci/mod.rs:27-48)
 * on the device from its two input columns d_c0, d_c1 (2^log_size rows): d_cols[0..n_cols)
 * receive c0, c1 and c_k = c_{k-2}^2 + c_{k-1}^2 (bit-reversed evaluations on
 * CanonicCoset(log_size), as the reference's generator lays them out). */
int sh_wide_fibonacci_trace(sh_ctx *ctx, const uint32_t *d_c0, const uint32_t *d_c1, uint32_t log_size,
                            uint64_t n_cols, uint32_t *const *d_cols);

/* ---------------------------------------------------------------- graph capture */
/* hipGraph of sh_commit_trace with these exact arguments (no counterpart in the
 * reference): one graph launch replays interpolate + LDE + Merkle on the same
 * buffers.  The graph owns its staging tables and scratch; the buffers passed
 * here must outlive it.  Contents of d_evals may change between launches. */
typedef struct sh_graph sh_graph;
int sh_commit_trace_graph_create(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_evals,
                                 uint32_t *const *d_coeffs, const uint32_t *log_sizes, uint64_t n,
                                 uint32_t log_blowup, uint32_t *const *d_ldes, uint8_t *const *d_layers,
                                 sh_graph **out);
/* hipGraph of sh_commit_trace_mixed with these exact arguments (no counterpart in the
 * reference); the same rules as sh_commit_trace_graph_create. */
int sh_commit_trace_mixed_graph_create(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_evals,
                                       uint32_t *const *d_coeffs, const uint32_t *log_sizes, uint64_t n,
                                       uint32_t log_blowup, uint32_t *const *d_ldes, uint8_t *const *d_layers,
                                       sh_graph **out);
/* Capture of any sequence of entry points that only enqueue device work (no counterpart in the
 * reference; the prover captures a proof's device work with it).  sh_graph_create makes an empty
 * graph that owns a staging ring, scratch buffers and an arena of device blocks.  Between
 * sh_graph_warmup_begin and sh_graph_warmup_end the calls run eagerly with those resources (every
 * pooled allocation comes from the arena, and a block freed at any later time goes back to it);
 * between sh_graph_capture_begin and sh_graph_capture_end the same calls, with the same arguments
 * and the warm-up's allocations freed again, are recorded instead of run.  A call that waits for
 * the device or stages a table the warm-up did not stage fails the capture.  sh_graph_launch then
 * replays the recorded work; sh_graph_destroy frees the arena's blocks, so buffers taken from it
 * must not be used afterwards.  One graph at a time per context is open; all on the main stream. */
int sh_graph_create(sh_ctx *ctx, sh_graph **out);
int sh_graph_warmup_begin(sh_ctx *ctx, sh_graph *g);
int sh_graph_warmup_end(sh_ctx *ctx, sh_graph *g);
int sh_graph_capture_begin(sh_ctx *ctx, sh_graph *g);
int sh_graph_capture_end(sh_ctx *ctx, sh_graph *g);
int sh_graph_launch(sh_ctx *ctx, sh_graph *g);
int sh_graph_destroy(sh_ctx *ctx, sh_graph *g);

#ifdef __cplusplus
}
#endif
#endif
