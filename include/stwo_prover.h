/*
 * stwo_prover.h — host-side protocol drivers above the backend C-ABI
 * (stwo_hip.h).  The reference implements these generically over its Backend
 * trait; Rust cannot be built here, so they are restated in C++ and drive the
 * HIP backend directly.  Each entry point cites the reference function it
 * restates.  Conventions as in stwo_hip.h.
 */
#ifndef STWO_PROVER_H
#define STWO_PROVER_H

#include <stdint.h>

#include "stwo_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- channel */
/* Blake2sChannel (crates/stwo/src/core/channel/blake2s.rs:16-117): plain state. */
typedef struct {
  uint8_t digest[32];
  uint32_t n_draws;
} sh_channel;
void sh_channel_mix_u32s(sh_channel *ch, const uint32_t *words, uint64_t n);
void sh_channel_mix_u64(sh_channel *ch, uint64_t value);
/* mix_felts: n QM31 values, 4 words each. */
void sh_channel_mix_felts(sh_channel *ch, const uint32_t *felts, uint64_t n);
/* MerkleChannel::mix_hash (crates/stwo/src/core/vcs_lifted/blake2_merkle.rs:56-58). */
void sh_channel_mix_root(sh_channel *ch, const uint8_t root[32]);
void sh_channel_draw_u32s(sh_channel *ch, uint32_t out[8]);
void sh_channel_draw_secure_felt(sh_channel *ch, uint32_t out[4]);
/* CirclePoint::get_random_point (crates/stwo/src/core/circle.rs:150-160). */
void sh_channel_draw_random_point(sh_channel *ch, uint32_t out[8]);
/* draw_queries (crates/stwo/src/core/queries.rs:11-28): n unsorted positions. */
void sh_channel_draw_queries(sh_channel *ch, uint32_t log_domain_size, uint32_t n, uint32_t *out);
/* AccumulationOps::generate_secure_powers (crates/stwo/src/prover/air/accumulation.rs:188-196):
 * out[4k..4k+4] = felt^k for k < n (host arithmetic). */
void sh_generate_secure_powers(const uint32_t felt[4], uint64_t n, uint32_t *out);

/* The same channel kept in device memory (16 words: the digest as 8 little-endian words, then
 * n_draws), so that a proof mixes its roots, nonces and sampled values and draws its randomness
 * without the host waiting for any of them.  Each call is stream-ordered on the context stream. */
/* The device channel set to a host state. */
int sh_dchannel_load(sh_ctx *ctx, uint32_t *d_ch, const uint8_t digest[32], uint32_t n_draws);
/* mix_felts / mix_u32s / mix_u64 (crates/stwo/src/core/channel/blake2s.rs:42-67) and mix_root over
 * n device words. */
int sh_dchannel_mix(sh_ctx *ctx, uint32_t *d_ch, const uint32_t *d_words, uint64_t n);
/* draw_secure_felts (crates/stwo/src/core/channel/blake2s.rs:69-78): n felts into d_out (4 words
 * each). */
int sh_dchannel_draw_secure_felts(sh_ctx *ctx, uint32_t *d_ch, uint32_t n, uint32_t *d_out);
/* CirclePoint::get_random_point (crates/stwo/src/core/circle.rs:150-160) into d_out (8 words). */
int sh_dchannel_draw_random_point(sh_ctx *ctx, uint32_t *d_ch, uint32_t *d_out);
/* generate_secure_powers (crates/stwo/src/prover/air/accumulation.rs:188-196) of a device felt:
 * d_out[4k..4k+4] = felt^k for k < n, or felt^(n-1-k) when reversed. */
int sh_secure_powers_dev(sh_ctx *ctx, const uint32_t *d_felt, uint64_t n, int reversed, uint32_t *d_out);

/* ---------------------------------------------------------------- Merkle decommitment */
/* MerkleProverLifted::decommit hash witness (crates/stwo/src/prover/vcs_lifted/prover.rs:97-162)
 * for a lifted tree of height L whose layers are d_layers[0..L] (layer k: 2^k hashes).
 * positions: sorted, deduplicated leaf indices.  Writes up to cap hashes (32 bytes each) to
 * witness and their count to n_witness. */
int sh_merkle_decommit(sh_ctx *ctx, const uint8_t *const *d_layers, uint32_t L, const uint32_t *positions,
                       uint64_t n_positions, uint8_t *witness, uint64_t cap, uint64_t *n_witness);
/* CommitmentTreeProver::decommit (crates/stwo/src/prover/pcs/mod.rs:392-398) ->
 * MerkleProverLifted::decommit (crates/stwo/src/prover/vcs_lifted/prover.rs:97-162) in one round
 * trip: values[c * n_pos + k] = column c (log size log_sizes[c], in [1, L]) at the lifted row of
 * positions[k] (positions in the given order, duplicates kept), and the hash witness of the
 * tree of height L (layers d_layers[0..L]) for the sorted, deduplicated positions: up to cap
 * hashes (32 bytes each) in witness, their count in n_witness. */
int sh_tree_decommit(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *log_sizes, uint64_t n_cols,
                     const uint8_t *const *d_layers, uint32_t L, const uint32_t *positions, uint64_t n_pos,
                     uint32_t *values, uint8_t *witness, uint64_t cap, uint64_t *n_witness);
/* Column::at for many (column, row) pairs: out[i] = d_cols[i][rows[i]] (one gather launch). */
int sh_gather(sh_ctx *ctx, const uint32_t *const *d_cols, const uint32_t *rows, uint64_t n, uint32_t *out);

/* ---------------------------------------------------------------- FRI */
typedef struct {
  uint32_t pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries, fold_step;
} sh_fri_params; /* FriParams (crates/stwo/src/core/fri.rs:31-43) */

/* FriProver::commit + FriProver::decommit (crates/stwo/src/prover/fri.rs:113-284) on a QM31
 * evaluation on CanonicCoset(log_size) (4 device coordinate arrays).  Advances the channel.
 * Writes the FriProof in CairoSerde felt encoding (crates/cairo-serialize/src/serialize.rs:66-90)
 * to felts (each felt < 2^64) and the sorted, deduplicated query positions to
 * query_positions (at most n_queries).  fold_step 1-4 is supported: with fold_step > 1 a layer
 * of at least 4 values is committed with packed leaves (crates/stwo/src/prover/fri.rs:295-297),
 * each leaf holding 4 consecutive QM31 values (LOG_PACKED_LEAF_SIZE = 2,
 * crates/stwo/src/core/vcs_lifted/verifier.rs:16), and each layer is folded fold_step times
 * (the last fold may be shorter). */
int sh_fri_prove(sh_ctx *ctx, const sh_twiddles *tw, sh_channel *ch, const sh_fri_params *params,
                 const uint32_t *const column[4], uint32_t log_size, uint64_t *felts, uint64_t cap,
                 uint64_t *n_felts, uint32_t *query_positions, uint64_t *n_query_positions);
/* A committed tree opened at the proof's queries (CommitmentTreeProver::decommit,
 * crates/stwo/src/prover/pcs/mod.rs:392-398, as sh_tree_decommit): inputs as sh_tree_decommit's;
 * preprocessed: the positions are prepare_preprocessed_query_positions of the queries
 * (crates/stwo/src/core/pcs/utils.rs:185-203), else the queries themselves.  Outputs: values
 * (n_cols * n_positions words, room for n_cols * n_queries), witness (up to witness_cap
 * hashes), n_positions, n_witness. */
typedef struct {
  const uint32_t *const *d_cols;
  const uint32_t *log_sizes;
  uint64_t n_cols;
  const uint8_t *const *d_layers;
  uint32_t L;
  uint32_t preprocessed;
  uint32_t *values;
  uint8_t *witness;
  uint64_t witness_cap;
  uint64_t n_positions;
  uint64_t n_witness;
} sh_tree_opening;

/* sh_fri_prove on a proof whose channel is the device channel d_ch (the rest of prove_values,
 * crates/stwo/src/prover/pcs/mod.rs:265-315): the driver starts from it without the host reading
 * it, and ch_out receives the channel's state at the end of the proof (as sh_fri_prove leaves
 * ch).  n_extra words at d_extra come back to h_extra with the commit phase's own download (the
 * proof's device-drawn values, read in the same round trip), and the n_trees trees are opened at
 * the queries in the FRI decommitment's gather (one more round trip in all). */
int sh_fri_prove_dch(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *d_ch, sh_channel *ch_out,
                     const uint32_t *d_extra, uint64_t n_extra, uint32_t *h_extra, sh_tree_opening *trees,
                     uint32_t n_trees, const sh_fri_params *params, const uint32_t *const column[4],
                     uint32_t log_size, uint64_t *felts, uint64_t cap, uint64_t *n_felts, uint32_t *query_positions,
                     uint64_t *n_query_positions);
/* sh_fri_prove_dch in two halves, so that the first can be recorded into a graph
 * (sh_graph_capture_begin) with the rest of a proof's device work.  sh_fri_commit_dch enqueues
 * FriProver::commit (crates/stwo/src/prover/fri.rs:113-226) on the device channel d_ch without
 * waiting for the device; *out receives the run (every layer's evaluation and tree).
 * sh_fri_decommit_dch waits for that work, then does what sh_fri_prove_dch does after it (the host
 * replay of the channel, checked against the device's alphas; the last layer; the proof of work;
 * the queries; one gather for the FRI witnesses and the trees' openings; the felts).  It leaves
 * the run's buffers in place: after a graph replay of the commit half it may be called again on
 * the same run.  sh_fri_run_free releases the run. */
typedef struct sh_fri_run sh_fri_run;
int sh_fri_commit_dch(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *d_ch, const uint32_t *d_extra,
                      uint64_t n_extra, const sh_fri_params *params, const uint32_t *const column[4],
                      uint32_t log_size, sh_fri_run **out);
int sh_fri_decommit_dch(sh_ctx *ctx, sh_fri_run *run, sh_channel *ch_out, uint32_t *h_extra, sh_tree_opening *trees,
                        uint32_t n_trees, uint64_t *felts, uint64_t cap, uint64_t *n_felts, uint32_t *query_positions,
                        uint64_t *n_query_positions);
int sh_fri_run_free(sh_ctx *ctx, sh_fri_run *run);

/* ---------------------------------------------------------------- DEEP quotients */
/* compute_fri_quotients (crates/stwo/src/prover/pcs/quotient_ops.rs:86-160): columns are the
 * committed evaluations of every tree, flattened in tree/column order; column i has samples
 * [sample_offsets[i], sample_offsets[i+1]) with points (8 words) and values (4 words).  Builds
 * the samples with random powers and periodicity samples
 * (crates/stwo/src/core/pcs/quotients.rs:274-319), accumulates numerators per column size,
 * lifts and accumulates per point, and combines into d_out (4 arrays, 2^lifting_log_size). */
int sh_fri_quotients(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_cols,
                     const uint32_t *col_log_sizes, uint64_t n_cols, const uint32_t *sample_offsets,
                     const uint32_t *sample_points, const uint32_t *sample_values,
                     const uint32_t random_coeff[4], uint32_t lifting_log_size, uint32_t log_blowup,
                     uint32_t *const d_out[4]);
/* compute_fri_quotients when every column has one sample and all share one point held in device
 * memory with the values and the random coefficient (a device channel's draws): column i is
 * sampled at d_point (8 words) with value d_values[4i..4i+4]; the line coefficients, per-size
 * sums and combine constants are computed on the device.  Same output as sh_fri_quotients. */
int sh_fri_quotients_dpoint(sh_ctx *ctx, const sh_twiddles *tw, const uint32_t *const *d_cols,
                            const uint32_t *col_log_sizes, uint64_t n_cols, const uint32_t *d_point,
                            const uint32_t *d_values, const uint32_t *d_random_coeff, uint32_t lifting_log_size,
                            uint32_t log_blowup, uint32_t *const d_out[4]);

#ifdef __cplusplus
}
#endif
#endif
