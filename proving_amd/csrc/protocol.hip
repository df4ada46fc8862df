// Host protocol drivers: Blake2s channel, lifted-Merkle decommitment, FRI prover,
// DEEP-quotient driver.  Restated from the reference (citations per function);
// all heavy work goes to the HIP kernels through the backend entry points.
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "../../include/stwo_prover.h"
#include "blake2s.h"
#include "blake2s_quad.h"
#include "common.h"

int sh_build_leaves_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* log_sizes, uint64_t n,
                         uint32_t L, uint32_t* d_out, const hipEvent_t* group_ready);
int sh_build_layers_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers);
int sh_accumulate_numerators_fl(sh_ctx* ctx, const uint32_t* const* d_cols, uint64_t n_cols, uint32_t col_log,
                                uint32_t log_blowup, const uint32_t* batch_points, const uint32_t* batch_term_offsets,
                                const uint32_t* term_col_idx, const uint32_t* term_values,
                                const uint32_t* term_random_coeffs, uint64_t n_batches, uint32_t* const* d_outs,
                                uint32_t* first_linear, const uint8_t* apply_fl);
int sh_fold_line_dev(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4], uint32_t line_initial_index,
                     uint32_t log_size, const uint32_t* d_alphas, uint32_t fold_step, uint32_t* const dst[4],
                     uint32_t* leaves = nullptr);
int sh_fold_circle_dev(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const src[4], uint32_t log_size,
                       const uint32_t* d_alpha, uint32_t* const dst[4], uint32_t* leaves = nullptr);
int sh_dch_mix_impl(sh_ctx* ctx, uint32_t* d_ch, const uint32_t* d_words, uint64_t n);

namespace {

typedef std::array<uint8_t, 32> Hash;

// ------------------------------------------------------------------ channel
// crates/stwo/src/core/channel/blake2s.rs
void ch_update(sh_channel* ch, const uint8_t d[32]) {
  memcpy(ch->digest, d, 32);
  ch->n_draws = 0;
}

sh::qm31 draw_felt(sh_channel* ch) {
  uint32_t u[8];
  sh_channel_draw_u32s(ch, u);
  uint32_t f[8];
  // M31::reduce(u) for u < 2^32 (m31.rs:58-60) = u mod P
  for (int i = 0; i < 8; i++) f[i] = (uint32_t)((uint64_t)u[i] % sh::P);
  uint32_t w[4];
  for (int i = 0; i < 4; i++) w[i] = sh::madd(sh::madd(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
  return sh::qload(w);
}

// ------------------------------------------------------------------ gather
__global__ void k_gather(const uint32_t* const* cols, const uint32_t* rows, uint32_t n, uint32_t* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cols[i][rows[i]];
}

// 32-byte nodes: out[2i], out[2i + 1] = the two halves of *nodes[i].
__global__ void k_gather_nodes(const uint4* const* nodes, uint32_t n, uint4* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * n) out[i] = nodes[i >> 1][i & 1];
}

// Queried values at lifted rows: out[c * n_pos + k] = cols[c][lift(pos[k])] with
// lift(p) = ((p >> (L - log_c + 1)) << 1) + (p & 1) (crates/stwo/src/prover/vcs_lifted/prover.rs:97-116).
__global__ void k_gather_lifted(const uint32_t* const* cols, const uint32_t* logs, uint32_t n_cols, uint32_t L,
                                const uint32_t* pos, uint32_t n_pos, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cols * n_pos) return;
  const uint32_t c = i / n_pos, k = i - c * n_pos;
  const uint32_t p = pos[k];
  out[i] = cols[c][((p >> (L - logs[c] + 1)) << 1) + (p & 1)];
}

// Up to 6 device regions copied into one buffer (dst + dst_off[r]), for a single download.
struct PackArgs {
  const uint32_t* src[6];
  uint32_t len[6];
  uint32_t dst_off[6];
  uint32_t n;
};
__global__ void k_pack(PackArgs a, uint32_t* dst) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t r = 0; r < a.n; r++)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.len[r]; i += stride)
      dst[a.dst_off[r] + i] = a.src[r][i];
}

// The queried values of one committed tree at its positions (k_gather_lifted).
struct LiftedJob {
  const uint32_t* const* cols;  // host array of n_cols device column addresses
  const uint32_t* logs;
  uint64_t n_cols;
  uint32_t L;
  std::vector<uint32_t> pos;
  uint64_t n_values() const { return pos.empty() ? 0 : n_cols * pos.size(); }
};

// Reads words (cols[i][rows[i]]), 32-byte nodes (*nodes[j]) and the lifted values of every job
// into host memory: one upload of the address tables (from the context's pinned buffer), one
// launch per kind and job, one download, one wait.  values receives the jobs' values in order.
int gather_all(sh_ctx* ctx, const std::vector<const uint32_t*>& cols, const std::vector<uint32_t>& rows,
               const std::vector<const uint8_t*>& nodes, const std::vector<LiftedJob>& jobs,
               std::vector<uint32_t>& words, std::vector<uint32_t>& node_words, std::vector<uint32_t>& values) {
  const uint64_t nw = cols.size(), nn = nodes.size();
  uint64_t nv = 0;
  for (auto& j : jobs) nv += j.n_values();
  words.assign(nw, 0);
  node_words.assign(8 * nn, 0);
  values.assign(nv, 0);
  if (nw + nn + nv == 0) return SH_OK;
  auto pad16 = [](uint64_t x) { return (x + 15) & ~15ull; };
  // tables: [node addresses][column addresses][rows] then per job [column addresses][logs][positions];
  // outputs: [node words][words][job values]
  const uint64_t o_cols = 8 * nn, o_rows = o_cols + 8 * nw;
  uint64_t tab_bytes = pad16(o_rows + 4 * nw);
  std::vector<uint64_t> o_job(jobs.size());
  for (size_t k = 0; k < jobs.size(); k++) {
    o_job[k] = tab_bytes;
    if (jobs[k].n_values()) tab_bytes = pad16(tab_bytes + 12 * jobs[k].n_cols + 4 * jobs[k].pos.size());
  }
  const uint64_t out_bytes = 32 * nn + 4 * nw + 4 * nv;
  uint8_t* h;
  SH_TRY(sh_host_xfer(ctx, tab_bytes + out_bytes, &h));
  if (nn) memcpy(h, nodes.data(), 8 * nn);
  if (nw) {
    memcpy(h + o_cols, cols.data(), 8 * nw);
    memcpy(h + o_rows, rows.data(), 4 * nw);
  }
  for (size_t k = 0; k < jobs.size(); k++) {
    const LiftedJob& j = jobs[k];
    if (!j.n_values()) continue;
    memcpy(h + o_job[k], j.cols, 8 * j.n_cols);
    memcpy(h + o_job[k] + 8 * j.n_cols, j.logs, 4 * j.n_cols);
    memcpy(h + o_job[k] + 12 * j.n_cols, j.pos.data(), 4 * j.pos.size());
  }
  void* scr;
  SH_TRY(sh_scratch_get(ctx, SCR_MISC, tab_bytes + out_bytes, &scr));
  uint8_t* base = (uint8_t*)scr;
  uint8_t* d_out = base + tab_bytes;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(base, h, tab_bytes, hipMemcpyHostToDevice, ctx->stream));
  if (nn) {
    k_gather_nodes<<<(unsigned)((2 * nn + 255) / 256), 256, 0, ctx->stream>>>((const uint4* const*)base,
                                                                               (uint32_t)nn, (uint4*)d_out);
    SH_TRY(sh_launch_check(ctx, "gather_nodes"));
  }
  if (nw) {
    k_gather<<<(unsigned)((nw + 255) / 256), 256, 0, ctx->stream>>>(
        (const uint32_t* const*)(base + o_cols), (const uint32_t*)(base + o_rows), (uint32_t)nw,
        (uint32_t*)(d_out + 32 * nn));
    SH_TRY(sh_launch_check(ctx, "gather"));
  }
  uint64_t vo = 0;
  for (size_t k = 0; k < jobs.size(); k++) {
    const LiftedJob& j = jobs[k];
    const uint64_t n = j.n_values();
    if (!n) continue;
    const uint8_t* t = base + o_job[k];
    k_gather_lifted<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
        (const uint32_t* const*)t, (const uint32_t*)(t + 8 * j.n_cols), (uint32_t)j.n_cols, j.L,
        (const uint32_t*)(t + 12 * j.n_cols), (uint32_t)j.pos.size(), (uint32_t*)(d_out + 32 * nn + 4 * nw) + vo);
    SH_TRY(sh_launch_check(ctx, "gather_lifted"));
    vo += n;
  }
  uint8_t* ho = h + tab_bytes;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(ho, d_out, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nn) memcpy(node_words.data(), ho, 32 * nn);
  if (nw) memcpy(words.data(), ho + 32 * nn, 4 * nw);
  if (nv) memcpy(values.data(), ho + 32 * nn + 4 * nw, 4 * nv);
  return SH_OK;
}

int gather_words_nodes(sh_ctx* ctx, const std::vector<const uint32_t*>& cols, const std::vector<uint32_t>& rows,
                       const std::vector<const uint8_t*>& nodes, std::vector<uint32_t>& words,
                       std::vector<uint32_t>& node_words) {
  std::vector<uint32_t> none;
  return gather_all(ctx, cols, rows, nodes, {}, words, node_words, none);
}

// Reads n words (d_cols[i][rows[i]]) into host memory with one launch.
int gather_words(sh_ctx* ctx, const std::vector<const uint32_t*>& cols, const std::vector<uint32_t>& rows,
                 std::vector<uint32_t>& out) {
  std::vector<uint32_t> none;
  return gather_words_nodes(ctx, cols, rows, {}, out, none);
}

// The nodes (layer, index) of MerkleProverLifted::decommit's hash witness, in witness order
// (crates/stwo/src/prover/vcs_lifted/prover.rs:118-152): one per node whose sibling is not known.
std::vector<std::pair<uint32_t, uint32_t>> merkle_witness_nodes(uint32_t L, std::vector<uint32_t> queries) {
  std::sort(queries.begin(), queries.end());
  queries.erase(std::unique(queries.begin(), queries.end()), queries.end());
  std::vector<std::pair<uint32_t, uint32_t>> need;
  need.reserve(queries.size() * L);
  std::vector<uint32_t> next;
  next.reserve(queries.size());
  for (int layer = (int)L - 1; layer >= 0; layer--) {
    next.clear();
    for (size_t i = 0; i < queries.size();) {
      uint32_t first = queries[i];
      size_t len = (i + 1 < queries.size() && (first ^ 1u) == queries[i + 1]) ? 2 : 1;
      if (len == 1) need.push_back({(uint32_t)layer + 1, first ^ 1u});
      next.push_back(first >> 1);
      i += len;
    }
    queries.swap(next);
  }
  return need;
}

// MerkleProverLifted::decommit hash witness (crates/stwo/src/prover/vcs_lifted/prover.rs:118-152).
int merkle_decommit(sh_ctx* ctx, const std::vector<const uint8_t*>& layers, uint32_t L,
                    std::vector<uint32_t> queries, std::vector<Hash>& witness) {
  std::sort(queries.begin(), queries.end());
  queries.erase(std::unique(queries.begin(), queries.end()), queries.end());
  // which nodes are needed: (layer, index), in witness order
  std::vector<std::pair<uint32_t, uint32_t>> need;
  for (int layer = (int)L - 1; layer >= 0; layer--) {
    std::vector<uint32_t> next;
    for (size_t i = 0; i < queries.size();) {
      uint32_t first = queries[i];
      size_t len = (i + 1 < queries.size() && (first ^ 1u) == queries[i + 1]) ? 2 : 1;
      if (len == 1) need.push_back({(uint32_t)layer + 1, first ^ 1u});
      next.push_back(first >> 1);
      i += len;
    }
    queries = next;
  }
  std::vector<const uint8_t*> nodes(need.size());
  for (size_t k = 0; k < need.size(); k++) nodes[k] = layers[need[k].first] + 32ull * need[k].second;
  std::vector<uint32_t> none, words;
  SH_TRY(gather_words_nodes(ctx, {}, {}, nodes, none, words));
  witness.resize(need.size());
  for (size_t k = 0; k < need.size(); k++) memcpy(witness[k].data(), &words[8 * k], 32);
  return SH_OK;
}

// A lifted tree committed on the device, with its layer buffers.
struct DeviceTree {
  uint32_t L = 0;
  sh_ctx* owner = nullptr;
  uint8_t* buf = nullptr;  // all layers, contiguous (pooled)
  std::vector<uint8_t*> layers;
  Hash root{};
  void free_all() {
    if (buf) sh_dev_free(owner, buf);
    buf = nullptr;
  }
};

// The layer buffers of a tree of height log_size (pooled, all layers in one allocation).
int alloc_tree(sh_ctx* ctx, uint32_t log_size, DeviceTree& t) {
  t.L = log_size;
  uint64_t total = 32ull * ((2ull << log_size) - 1);
  t.owner = ctx;
  SH_TRY(sh_dev_alloc(ctx, total, (void**)&t.buf));
  t.layers.resize(log_size + 1);
  uint64_t off = 0;
  for (uint32_t k = 0; k <= log_size; k++) {
    t.layers[k] = t.buf + off;
    off += 32ull << k;
  }
  return SH_OK;
}

// leaves_ready: t is allocated already and its leaf layer written (by the fold that made the
// columns); only the layers above it are built.
// st: the FRI channel step, run by the launch that computes the root (k_fri_channel for a tree of
// one leaf).
__global__ void k_fri_channel(const uint32_t* root, uint32_t* ch, uint32_t k, uint32_t* alphas, uint32_t* root_out);
int commit_columns(sh_ctx* ctx, const std::vector<const uint32_t*>& cols, uint32_t log_size, DeviceTree& t,
                   bool fetch_root = true, bool leaves_ready = false, const sh_root_step* st = nullptr) {
  if (!leaves_ready) {
    SH_TRY(alloc_tree(ctx, log_size, t));
    std::vector<uint32_t> logs(cols.size(), log_size);
    SH_TRY(sh_build_leaves_impl(ctx, cols.data(), logs.data(), cols.size(), log_size,
                                (uint32_t*)t.layers[log_size], nullptr));
  }
  SH_TRY(sh_build_layers_step_impl(ctx, log_size, (uint32_t* const*)t.layers.data(), st));
  if (st && log_size == 0) {
    k_fri_channel<<<1, 64, 0, ctx->stream>>>((const uint32_t*)t.layers[0], st->ch, st->k, st->alphas, st->root_out);
    SH_TRY(sh_launch_check(ctx, "fri_channel"));
  }
  if (!fetch_root) return SH_OK;
  SH_CHECK_HIP(ctx, hipMemcpyAsync(t.root.data(), t.layers[0], 32, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SH_OK;
}

// PackLeavesOps::pack_leaves_input (crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:101-124):
// output column coord + 4*offset, row i, holds coordinate coord of value 4i + offset.
__global__ void k_pack_leaves(const uint32_t* c0, const uint32_t* c1, const uint32_t* c2, const uint32_t* c3,
                              uint32_t n_out, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  const uint32_t* cols[4] = {c0, c1, c2, c3};
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int off = 0; off < 4; off++) out[(uint64_t)(c + 4 * off) * n_out + i] = cols[c][4ull * i + off];
}

// FriFirstLayerProver::new / FriInnerLayerProver::new (crates/stwo/src/prover/fri.rs:294-306,
// 357-368): a tree over the 4 coordinate columns, one leaf per value, or with packed leaves one
// leaf per 4 consecutive values (16 words: one Blake2s block) and a tree 2 levels shorter.
int commit_fri_layer(sh_ctx* ctx, const uint32_t* const cols[4], uint32_t log_size, bool pack, DeviceTree& t,
                     bool fetch_root = true, const sh_root_step* st = nullptr) {
  if (!pack) return commit_columns(ctx, {cols[0], cols[1], cols[2], cols[3]}, log_size, t, fetch_root, false, st);
  const uint32_t n_out = 1u << (log_size - 2);
  uint32_t* packed;
  SH_TRY(sh_dev_alloc(ctx, 64ull * n_out, (void**)&packed));
  k_pack_leaves<<<(n_out + 255) / 256, 256, 0, ctx->stream>>>(cols[0], cols[1], cols[2], cols[3], n_out, packed);
  int rc = sh_launch_check(ctx, "pack_leaves");
  if (rc == SH_OK) {
    std::vector<const uint32_t*> pc(16);
    for (int k = 0; k < 16; k++) pc[k] = packed + (uint64_t)k * n_out;
    rc = commit_columns(ctx, pc, log_size - 2, t, fetch_root, false, st);
  }
  sh_dev_free(ctx, packed);  // stream-ordered: a later allocation reuses it after the leaves read it
  return rc;
}

// Blake2sChannel::mix_root followed by draw_secure_felt (crates/stwo/src/core/channel/blake2s.rs:
// 33-117) on the device, in one thread.  ch holds the digest (8 LE words) and n_draws; the root is
// layer 0 of a tree committed on the same stream.  Writes squared_alpha_powers(alpha, k)
// (crates/stwo/src/prover/fri.rs:79-87) to alphas (4 words each) and a copy of the root to root_out.
__global__ void k_fri_channel(const uint32_t* root, uint32_t* ch, uint32_t k, uint32_t* alphas, uint32_t* root_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t m[16], h[8], u[8];
  for (int i = 0; i < 8; i++) {
    m[i] = ch[i];
    m[8 + i] = root[i];
    root_out[i] = root[i];
  }
  sh::b2_init(h);
  sh::b2_compress(h, m, 64, true);  // digest = Blake2s(digest || root), one block
  for (int i = 0; i < 8; i++) m[i] = h[i];
  for (int i = 8; i < 16; i++) m[i] = 0;  // n_draws = 0 (LE), then the 0x00 byte
  sh::b2_init(u);
  sh::b2_compress(u, m, 37, true);  // Blake2s(digest || LE32(n_draws) || 0x00)
  for (int i = 0; i < 8; i++) ch[i] = h[i];
  ch[8] = 1;
  uint32_t f[8], w[4];
  for (int i = 0; i < 8; i++) f[i] = sh::reduce_once((u[i] & sh::P) + (u[i] >> 31));  // M31::reduce(u32)
  for (int i = 0; i < 4; i++) w[i] = sh::madd(sh::madd(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
  sh::qm31 a = sh::qload(w);
  for (uint32_t j = 0; j < k; j++) {
    sh::qstore(a, alphas + 4 * j);
    a = sh::qmul(a, a);
  }
}

// ---- the small FRI layers in one launch (fold_step 1)
// At 2^9 values and below a layer's leaves, tree, channel step and fold are each too small to fill
// the device, so each launch costs its latency (a layer of 2^8..2^13 values as separate launches
// takes about 60 us).  One workgroup of 1024 lanes runs every remaining layer: the leaves
// (Blake2s of the 4 coordinates, 16 bytes) and the tree with four lanes per hash through LDS, the
// channel step (mix the root, draw alpha: FriInnerLayerProver::new + draw_secure_felt,
// crates/stwo/src/prover/fri.rs:175-194) and the fold (fold_line, crates/stwo/src/prover/fri.rs:
// 41-45) into the next layer's evaluation.  Every layer's evaluation and tree stay in device memory
// for the decommitment, exactly as the per-layer launches leave them.
// (Measured on the FRI + DEEP phase of 200 x 2^21: layers up to 2^7 with 512 lanes 2.47 ms, up
// to 2^9 with 1024 lanes 2.37 ms, up to 2^10 with 1024 lanes 2.46 ms: one CU is the bound there.)
constexpr uint32_t kFriTailMaxLog = 9;
constexpr uint32_t kFriTailMaxLayers = 12;
constexpr uint32_t kFriTailThreads = 1024;
// the tree's next level (4 << log words at most) fits in the leaf-message area
static_assert((kFriTailThreads / 4) * 16 >= (4u << kFriTailMaxLog), "FRI tail work area too small");

struct FriTailLayer {
  const uint32_t* eval;  // 4 coordinate columns of 2^log words, contiguous
  uint8_t* tree;         // Merkle layer k (2^k hashes) at 32 * (2^k - 1)
  const uint32_t* itw;   // inverse twiddles of the line domain (one per fold output)
  uint32_t log, slot;    // log size; index into alphas (16 words each) and roots (8 words each)
};
struct FriTailArgs {
  FriTailLayer layer[kFriTailMaxLayers];
  uint32_t* next_eval[kFriTailMaxLayers];  // fold output of layer j (the evaluation of layer j + 1)
  uint32_t n_layers;
  uint32_t* ch;  // device channel: digest (8 words), n_draws
  uint32_t* alphas;
  uint32_t* roots;
};

__global__ __launch_bounds__(kFriTailThreads) void k_fri_tail(FriTailArgs a) {
  __shared__ uint32_t hashes[8u << kFriTailMaxLog];  // leaf hashes, then tree levels
  __shared__ uint32_t work[(kFriTailThreads / 4) * 16];  // leaf messages, then tree levels
  __shared__ uint32_t msg[16];
  __shared__ uint32_t digest[8];
  __shared__ uint32_t draw[8];
  __shared__ uint32_t alpha_w[4];
  const uint32_t tid = threadIdx.x, q = tid & 3, quad = tid >> 2;
  constexpr uint32_t kQuads = kFriTailThreads / 4;
  uint32_t sidx[40];
  quad_sigma(q, sidx);
  uint32_t h_lo, h_hi, d_leaf, d_node, d_draw;
  quad_init(q, 16, h_lo, h_hi, d_leaf);  // a leaf: the 4 coordinates, 16 bytes
  quad_init(q, 64, h_lo, h_hi, d_node);  // a node, or digest || root
  quad_init(q, 37, h_lo, h_hi, d_draw);  // digest || LE32(n_draws) || 0x00
  if (tid < 8) digest[tid] = a.ch[tid];
  for (uint32_t j = 0; j < a.n_layers; j++) {
    const FriTailLayer L = a.layer[j];
    const uint32_t N = 1u << L.log;
    // leaves, one per quad: lane q stores coordinate q into the quad's zero-padded message
    for (uint32_t base = 0; base < N; base += kQuads) {
      const uint32_t g = base + quad;
      for (uint32_t w = 4 + q; w < 16; w += 4) work[16 * quad + w] = 0;
      if (g < N) work[16 * quad + q] = L.eval[(uint64_t)q * N + g];
      __syncthreads();
      if (g < N) {
        uint32_t lo, hi;
        quad_node_hash(work + 16 * quad, sidx, q, h_lo, h_hi, d_leaf, lo, hi);
        hashes[8 * g + q] = lo;
        hashes[8 * g + 4 + q] = hi;
        uint32_t* o = (uint32_t*)(L.tree + 32ull * (N - 1)) + 8 * g;
        o[q] = lo;
        o[4 + q] = hi;
      }
      __syncthreads();
    }
    // the tree: level l (2^l nodes) from level l + 1
    uint32_t* cur = hashes;
    uint32_t* nxt = work;
    for (int l = (int)L.log - 1; l >= 0; l--) {
      const uint32_t cnt = 1u << l;
      for (uint32_t g = quad; g < cnt; g += kQuads) {
        uint32_t lo, hi;
        quad_node_hash(cur + 16 * g, sidx, q, h_lo, h_hi, d_node, lo, hi);
        nxt[8 * g + q] = lo;
        nxt[8 * g + 4 + q] = hi;
        uint32_t* o = (uint32_t*)(L.tree + 32ull * (cnt - 1)) + 8 * g;
        o[q] = lo;
        o[4 + q] = hi;
      }
      __syncthreads();
      uint32_t* t = cur;
      cur = nxt;
      nxt = t;
    }
    // channel: digest = Blake2s(digest || root); draw = Blake2s(digest || LE32(0) || 0x00)
    if (tid < 16) msg[tid] = tid < 8 ? digest[tid] : cur[tid - 8];
    if (tid < 8) a.roots[8 * L.slot + tid] = cur[tid];
    __syncthreads();
    if (tid < 4) {
      uint32_t lo, hi;
      quad_node_hash(msg, sidx, q, h_lo, h_hi, d_node, lo, hi);
      digest[q] = lo;
      digest[4 + q] = hi;
    }
    __syncthreads();
    if (tid < 16) msg[tid] = tid < 8 ? digest[tid] : 0;
    __syncthreads();
    if (tid < 4) {
      uint32_t lo, hi;
      quad_node_hash(msg, sidx, q, h_lo, h_hi, d_draw, lo, hi);
      draw[q] = lo;
      draw[4 + q] = hi;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t f[8], w[4];
      for (int i = 0; i < 8; i++) f[i] = sh::reduce_once((draw[i] & sh::P) + (draw[i] >> 31));  // M31::reduce(u32)
      for (int i = 0; i < 4; i++) w[i] = sh::madd(sh::madd(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
      for (int i = 0; i < 4; i++) {
        alpha_w[i] = w[i];
        a.alphas[16 * L.slot + i] = w[i];
      }
      for (int i = 0; i < 8; i++) a.ch[i] = digest[i];
      a.ch[8] = 1;
    }
    __syncthreads();
    // fold: out[i] = (v[2i] + v[2i+1]) + alpha * ((v[2i] - v[2i+1]) * x_i^-1)
    const sh::qm31 alpha = sh::qload(alpha_w);
    const uint32_t M = N / 2;
    uint32_t* dst = a.next_eval[j];
    for (uint32_t i = tid; i < M; i += kFriTailThreads) {
      sh::qm31 x = {{L.eval[2 * i], L.eval[N + 2 * i]}, {L.eval[2 * N + 2 * i], L.eval[3 * N + 2 * i]}};
      sh::qm31 y = {{L.eval[2 * i + 1], L.eval[N + 2 * i + 1]}, {L.eval[2 * N + 2 * i + 1], L.eval[3 * N + 2 * i + 1]}};
      sh::qm31 f0 = sh::qadd(x, y);
      sh::qm31 f1 = sh::qmulm(sh::qsub(x, y), L.itw[i]);
      sh::qm31 r = sh::qadd(f0, sh::qmul(alpha, f1));
      dst[i] = r.a.a;
      dst[M + i] = r.a.b;
      dst[2 * M + i] = r.b.a;
      dst[3 * M + i] = r.b.b;
    }
    __syncthreads();  // the next layer reads dst
  }
}

// Serialisation helpers (crates/cairo-serialize/src/serialize.rs).
struct Felts {
  std::vector<uint64_t> v;
  void u(uint64_t x) { v.push_back(x); }
  void q(const sh::qm31& x) {
    u(x.a.a);
    u(x.a.b);
    u(x.b.a);
    u(x.b.b);
  }
  void h(const Hash& x) {
    for (int i = 0; i < 8; i++) {
      uint32_t w;
      memcpy(&w, x.data() + 4 * i, 4);
      u(w);
    }
  }
};

// A committed FRI layer: its evaluation (4 device arrays) and Merkle tree.
struct FriLayer {
  uint32_t log_size = 0;
  sh_ctx* owner = nullptr;
  uint32_t* buf = nullptr;  // 4 * 2^log_size words (pooled)
  const uint32_t* cols[4] = {nullptr, nullptr, nullptr, nullptr};
  DeviceTree tree;
  uint32_t fold_step = 1;  // folds from this layer to the next
  bool pack = false;       // leaves of 4 consecutive values
  bool leaves_done = false;  // tree allocated and its leaves written by the fold that made this layer
  void free_all() {
    if (buf) sh_dev_free(owner, buf);
    buf = nullptr;
    tree.free_all();
  }
};

// line_ifft (crates/stwo/src/core/poly/line.rs:202-219) on natural-order values.
void line_ifft(std::vector<sh::qm31>& v, uint32_t init, uint32_t log) {
  uint32_t step = sh::subgroup_gen(log);
  while (log > 0) {
    uint64_t size = 1ull << log;
    for (uint64_t c = 0; c < v.size(); c += size)
      for (uint64_t i = 0; i < size / 2; i++) {
        uint32_t x = sh::index_to_point_host(sh::idx_reduce((uint64_t)init + (uint64_t)step * i)).x;
        uint32_t xi = sh::minv(x);
        sh::qm31 a = v[c + i], b = v[c + size / 2 + i];
        v[c + i] = sh::qadd(a, b);
        v[c + size / 2 + i] = sh::qmulm(sh::qsub(a, b), xi);
      }
    init = sh::idx_reduce((uint64_t)init * 2);
    step = sh::idx_reduce((uint64_t)step * 2);
    log--;
  }
}

void bit_reverse_q(std::vector<sh::qm31>& v) {
  uint32_t log = 0;
  while ((1ull << log) < v.size()) log++;
  for (uint32_t i = 0; i < v.size(); i++) {
    uint32_t j = sh::bitrev(i, log);
    if (j > i) std::swap(v[i], v[j]);
  }
}

}  // namespace

// ------------------------------------------------------------------ channel C-ABI
extern "C" void sh_channel_mix_u32s(sh_channel* ch, const uint32_t* words, uint64_t n) {
  sh::B2Stream s;
  sh::b2s_init(&s);
  sh::b2s_update(&s, ch->digest, 32);
  for (uint64_t i = 0; i < n; i++) sh::b2s_update(&s, &words[i], 4);
  uint8_t d[32];
  sh::b2s_final(&s, d);
  ch_update(ch, d);
}
extern "C" void sh_channel_mix_u64(sh_channel* ch, uint64_t v) {
  uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
  sh_channel_mix_u32s(ch, w, 2);
}
extern "C" void sh_channel_mix_felts(sh_channel* ch, const uint32_t* felts, uint64_t n) {
  sh_channel_mix_u32s(ch, felts, 4 * n);
}
extern "C" void sh_channel_mix_root(sh_channel* ch, const uint8_t root[32]) {
  uint8_t buf[64], d[32];
  memcpy(buf, ch->digest, 32);
  memcpy(buf + 32, root, 32);
  sh::b2s_hash(buf, 64, d);
  ch_update(ch, d);
}
extern "C" void sh_channel_draw_u32s(sh_channel* ch, uint32_t out[8]) {
  uint8_t buf[37], d[32];
  memcpy(buf, ch->digest, 32);
  memcpy(buf + 32, &ch->n_draws, 4);
  buf[36] = 0;
  ch->n_draws++;
  sh::b2s_hash(buf, 37, d);
  memcpy(out, d, 32);
}
extern "C" void sh_channel_draw_secure_felt(sh_channel* ch, uint32_t out[4]) { sh::qstore(draw_felt(ch), out); }
extern "C" void sh_channel_draw_random_point(sh_channel* ch, uint32_t out[8]) {
  sh::qm31 t = draw_felt(ch);
  sh::qm31 t2 = sh::qmul(t, t);
  sh::qm31 inv = sh::qinv(sh::qadd(t2, sh::qone()));
  sh::qstore(sh::qmul(sh::qsub(sh::qone(), t2), inv), out);
  sh::qstore(sh::qmul(sh::qadd(t, t), inv), out + 4);
}
extern "C" void sh_channel_draw_queries(sh_channel* ch, uint32_t log_domain_size, uint32_t n, uint32_t* out) {
  uint32_t mask = log_domain_size >= 32 ? 0xffffffffu : ((1u << log_domain_size) - 1);
  uint32_t k = 0;
  while (k < n) {
    uint32_t w[8];
    sh_channel_draw_u32s(ch, w);
    for (int i = 0; i < 8 && k < n; i++) out[k++] = w[i] & mask;
  }
}

extern "C" void sh_generate_secure_powers(const uint32_t felt[4], uint64_t n, uint32_t* out) {
  const sh::qm31 f = sh::qload(felt);
  sh::qm31 acc = sh::qone();
  for (uint64_t i = 0; i < n; i++) {
    sh::qstore(acc, out + 4 * i);
    acc = sh::qmul(acc, f);
  }
}

extern "C" int sh_gather(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* rows, uint64_t n, uint32_t* out) {
  if (!ctx || (n && (!d_cols || !rows || !out))) return SH_ERR_INVALID;
  std::vector<const uint32_t*> c(d_cols, d_cols + n);
  std::vector<uint32_t> r(rows, rows + n), o;
  SH_TRY(gather_words(ctx, c, r, o));
  if (n) memcpy(out, o.data(), 4 * n);
  return SH_OK;
}

extern "C" int sh_merkle_decommit(sh_ctx* ctx, const uint8_t* const* d_layers, uint32_t L, const uint32_t* positions,
                                  uint64_t n_positions, uint8_t* witness, uint64_t cap, uint64_t* n_witness) {
  if (!ctx || !d_layers || !n_witness || (n_positions && !positions)) return SH_ERR_INVALID;
  std::vector<const uint8_t*> layers(d_layers, d_layers + L + 1);
  std::vector<uint32_t> q(positions, positions + n_positions);
  for (auto p : q) SH_REQUIRE(ctx, L < 32 && p < (1u << L), "query position %u out of range", p);
  std::vector<Hash> w;
  SH_TRY(merkle_decommit(ctx, layers, L, q, w));
  *n_witness = w.size();
  SH_REQUIRE(ctx, w.size() <= cap, "witness of %zu hashes exceeds the buffer (%llu)", w.size(),
             (unsigned long long)cap);
  for (size_t k = 0; k < w.size(); k++) memcpy(witness + 32 * k, w[k].data(), 32);
  return SH_OK;
}

// One tree's opening (MerkleProverLifted::decommit, crates/stwo/src/prover/vcs_lifted/prover.rs:
// 97-162) added to a gather: its witness nodes (sorted, deduplicated positions) and its values at
// the positions as given.  *n_nodes receives the witness length.
static int plan_tree_opening(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes, uint64_t n_cols,
                             const uint8_t* const* d_layers, uint32_t L, std::vector<uint32_t> pos,
                             std::vector<const uint8_t*>& nodes, std::vector<LiftedJob>& jobs, uint64_t* n_nodes) {
  *n_nodes = 0;
  const uint64_t n_pos = pos.size();
  if ((n_cols && n_pos && (!d_cols || !log_sizes)) || (L > 0 && n_pos && !d_layers)) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, L < 31, "tree height %u out of range", L);
  SH_REQUIRE(ctx, n_cols * n_pos < (1ull << 31), "too many queried values");
  for (uint64_t k = 0; k < n_pos; k++) SH_REQUIRE(ctx, pos[k] < (1u << L), "query position %u out of range", pos[k]);
  for (uint64_t c = 0; c < n_cols && n_pos; c++)
    SH_REQUIRE(ctx, log_sizes[c] >= 1 && log_sizes[c] <= L, "column log size %u not in [1, %u]", log_sizes[c], L);
  if (L > 0 && n_pos) {
    auto need = merkle_witness_nodes(L, pos);
    for (auto& nd : need) nodes.push_back(d_layers[nd.first] + 32ull * nd.second);
    *n_nodes = need.size();
  }
  if (n_cols && n_pos) jobs.push_back({d_cols, log_sizes, n_cols, L, std::move(pos)});
  return SH_OK;
}

extern "C" int sh_tree_decommit(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes,
                                uint64_t n_cols, const uint8_t* const* d_layers, uint32_t L,
                                const uint32_t* positions, uint64_t n_pos, uint32_t* values, uint8_t* witness,
                                uint64_t cap, uint64_t* n_witness) {
  if (!ctx || !n_witness || (n_pos && !positions) || (n_cols && n_pos && !values)) return SH_ERR_INVALID;
  *n_witness = 0;
  std::vector<const uint8_t*> nodes;
  std::vector<LiftedJob> jobs;
  uint64_t nn = 0;
  SH_TRY(plan_tree_opening(ctx, d_cols, log_sizes, n_cols, d_layers, L,
                           std::vector<uint32_t>(positions, positions + n_pos), nodes, jobs, &nn));
  SH_REQUIRE(ctx, nn <= cap, "witness of %llu hashes exceeds the buffer (%llu)", (unsigned long long)nn,
             (unsigned long long)cap);
  std::vector<uint32_t> none, nw, vals;
  SH_TRY(gather_all(ctx, {}, {}, nodes, jobs, none, nw, vals));
  if (nn) memcpy(witness, nw.data(), 32 * nn);
  if (!vals.empty()) memcpy(values, vals.data(), 4 * vals.size());
  *n_witness = nn;
  return SH_OK;
}

// ------------------------------------------------------------------ FRI prover
namespace {
// LineEvaluation::interpolate with commit_last_layer's degree check and coefficient order
// (crates/stwo/src/prover/line.rs:54-65, crates/stwo/src/prover/fri.rs:208-226) in one workgroup:
// e[c] holds coordinate c of the n = 2^log values in bit-reversed order; itw holds the inverse
// x-coordinates of line_ifft's layers (crates/stwo/src/core/poly/line.rs:202-219), layer j
// (2^(log-1-j) values) after layer j - 1.  The ordered coefficient k is the scaled transform's
// value at bitrev(k); out[4m..4m+4] = coefficient bitrev(m, bound_log) (LinePoly keeps its
// coefficients bit-reversed), and flag[0] = 1 when a coefficient at or above 2^bound_log is
// nonzero (the reference's "invalid degree").
constexpr uint32_t kLastLayerMaxLog = 10;
__global__ __launch_bounds__(256) void k_fri_last_layer(const uint32_t* e0, const uint32_t* e1, const uint32_t* e2,
                                                        const uint32_t* e3, const uint32_t* __restrict__ itw,
                                                        uint32_t log, uint32_t bound_log, uint32_t inv_n,
                                                        uint32_t* __restrict__ out, uint32_t* __restrict__ flag) {
  __shared__ uint32_t s[4][1u << kLastLayerMaxLog];
  __shared__ uint32_t bad;
  const uint32_t n = 1u << log, tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (uint32_t i = tid; i < n; i += blockDim.x) {
    const uint32_t j = sh::bitrev(i, log);  // natural order
    s[0][j] = e0[i];
    s[1][j] = e1[i];
    s[2][j] = e2[i];
    s[3][j] = e3[i];
  }
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t l = log; l > 0; l--) {
    const uint32_t half = 1u << (l - 1);
    for (uint32_t t = tid; t < n / 2; t += blockDim.x) {
      const uint32_t c = t >> (l - 1), i = t & (half - 1);
      const uint32_t ia = (c << l) + i, ib = ia + half;
      const sh::qm31 a = {{s[0][ia], s[1][ia]}, {s[2][ia], s[3][ia]}};
      const sh::qm31 b = {{s[0][ib], s[1][ib]}, {s[2][ib], s[3][ib]}};
      const sh::qm31 x = sh::qadd(a, b), y = sh::qmulm(sh::qsub(a, b), itw[off + i]);
      s[0][ia] = x.a.a;
      s[1][ia] = x.a.b;
      s[2][ia] = x.b.a;
      s[3][ia] = x.b.b;
      s[0][ib] = y.a.a;
      s[1][ib] = y.a.b;
      s[2][ib] = y.b.a;
      s[3][ib] = y.b.b;
    }
    off += half;
    __syncthreads();
  }
  for (uint32_t i = tid; i < n; i += blockDim.x) {
    const uint32_t k = sh::bitrev(i, log);
    const sh::qm31 v = sh::qmulm({{s[0][i], s[1][i]}, {s[2][i], s[3][i]}}, inv_n);
    if (k >= (1u << bound_log)) {
      if (!sh::qeq(v, sh::qzero())) bad = 1;
    } else {
      sh::qstore(v, out + 4 * sh::bitrev(k, bound_log));
    }
  }
  __syncthreads();
  if (tid == 0) flag[0] = bad;
}

// Inverse twiddles of line_ifft over the line domain (init, log), layer after layer, as
// k_fri_last_layer reads them; computed once per domain.
const std::vector<uint32_t>& last_layer_itw(uint32_t init, uint32_t log) {
  static std::mutex mu;
  static std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto& t = cache[{init, log}];
  if (t.empty()) {
    uint32_t step = sh::subgroup_gen(log);
    for (uint32_t l = log; l > 0; l--) {
      for (uint64_t i = 0; i < (1ull << (l - 1)); i++)
        t.push_back(sh::minv(sh::index_to_point_host(sh::idx_reduce((uint64_t)init + (uint64_t)step * i)).x));
      init = sh::idx_reduce((uint64_t)init * 2);
      step = sh::idx_reduce((uint64_t)step * 2);
    }
    if (t.empty()) t.push_back(0);  // log 0: no layers (one word keeps the staged table non-empty)
  }
  return t;
}

// verify_pow_nonce (crates/stwo/src/core/proof_of_work.rs): Blake2s(prefixed digest || LE64(nonce))
// has pow_bits trailing zero bits, prefixed = Blake2s(LE32(0x12345678) || digest || LE32(pow_bits)).
bool pow_ok(const uint8_t digest[32], uint32_t pow_bits, uint64_t nonce) {
  uint8_t buf[40], pre[32], h[32];
  const uint32_t prefix = 0x12345678u;
  memcpy(buf, &prefix, 4);
  memcpy(buf + 4, digest, 32);
  memcpy(buf + 36, &pow_bits, 4);
  sh::b2s_hash(buf, 40, pre);
  memcpy(buf, pre, 32);
  memcpy(buf + 32, &nonce, 8);
  sh::b2s_hash(buf, 40, h);
  for (uint32_t b = 0; b < pow_bits; b++)
    if ((h[b >> 3] >> (b & 7)) & 1) return false;
  return true;
}
}  // namespace

// The state of one FRI proof between its commit phase (enqueued only) and its decommitment: every
// layer's evaluation and tree, and the device words of the commit phase's one download.
struct sh_fri_run {
  sh_ctx* owner = nullptr;
  sh_fri_params p{};
  uint32_t log_size = 0, fs = 0, last_log = 0, max_layers = 0, n_layers = 0;
  uint32_t last_line_init = 0;  // the last layer's line domain
  bool from_dch = false;        // the proof's channel is a device channel (copied to CH)
  bool device_tail = false;     // the last layer's coefficients, their mix and the proof of work ran on the device
  sh_channel ch_in{};           // the host channel the commit started from (host-channel proofs)
  uint64_t n_extra = 0;
  uint32_t* d_aux = nullptr;
  // word offsets in d_aux: the channel, the alphas (16 words per layer), the roots (8 per layer),
  // the initial channel, the device's last-layer coefficients, the degree flag, the nonce
  // control words; then the last layer's evaluation and the extra words (the download)
  uint64_t ALPHAS = 0, ROOTS = 0, INIT = 0, LASTC = 0, FLAG = 0, CTL = 0, EVAL = 0, EXTRA = 0, dl_words = 0;
  FriLayer first, cur;
  std::vector<FriLayer> inner;
  void free_all() {
    first.tree.free_all();
    for (auto& l : inner) l.free_all();
    inner.clear();
    cur.free_all();
    if (d_aux) sh_dev_free(owner, d_aux);
    d_aux = nullptr;
  }
};

// FriProver::commit (crates/stwo/src/prover/fri.rs:113-226), enqueued without waiting for the
// device: each layer's root is mixed into a device copy of the channel and the folding alphas are
// drawn there, by the launch that computes the root (k_next_layers_quad's channel step), so the
// folds read them from device memory.  When the proof of work
// fits the device search (pow_bits <= 32) and the last layer has at most 2^10 values, the last
// layer's coefficients are computed and mixed and the proof of work is ground on the device too
// (FriProver::decommit's first steps, fri.rs:229-244).  The host replays the same channel
// operations afterwards, from the downloaded roots, and checks that the device agrees.
// d_ch_in: when set, the proof's channel is this device channel (16 words, as dchannel.hip keeps
// it); else ch is staged.  d_extra: n_extra words of the caller's device memory that the commit
// phase's download also brings back.
static int fri_commit_impl(sh_ctx* ctx, const sh_twiddles* tw, const sh_channel* ch, const uint32_t* d_ch_in,
                           const uint32_t* d_extra, uint64_t n_extra, const sh_fri_params* p,
                           const uint32_t* const column[4], uint32_t log_size, sh_fri_run** out) {
  if (!ctx || !tw || !p || !column || !out || (!ch && !d_ch_in) || (n_extra && !d_extra)) return SH_ERR_INVALID;
  *out = nullptr;
  SH_REQUIRE(ctx, n_extra < (1ull << 28), "too many extra words (%llu)", (unsigned long long)n_extra);
  SH_REQUIRE(ctx, p->fold_step >= 1 && p->fold_step <= 4, "fold_step %u out of range [1, 4]", p->fold_step);
  SH_REQUIRE(ctx, p->log_blowup_factor >= 1 && p->log_blowup_factor <= 16 && p->log_last_layer_degree_bound <= 10,
             "invalid FRI parameters");
  const uint32_t fs = p->fold_step;
  const uint32_t last_log = p->log_last_layer_degree_bound + p->log_blowup_factor;
  SH_REQUIRE(ctx, log_size >= 2 && log_size <= 30 && log_size >= fs + last_log,
             "The circle-to-line fold results in a smaller line domain than the last layer.");
  sh_fri_run* run = new sh_fri_run();
  run->owner = ctx;
  run->p = *p;
  run->log_size = log_size;
  run->fs = fs;
  run->last_log = last_log;
  run->from_dch = d_ch_in != nullptr;
  if (ch) run->ch_in = *ch;
  run->n_extra = n_extra;
  run->device_tail = p->pow_bits <= 32 && last_log <= kLastLayerMaxLog;
  const uint32_t max_layers = run->max_layers = log_size + 1;
  run->ALPHAS = 16;
  run->ROOTS = run->ALPHAS + 16ull * max_layers;
  run->INIT = run->ROOTS + 8ull * max_layers;
  run->LASTC = run->INIT + 16;
  run->FLAG = run->LASTC + (4ull << last_log);
  run->CTL = run->FLAG + 4;
  run->EVAL = run->CTL + 4;
  run->EXTRA = run->EVAL + (4ull << last_log);
  run->dl_words = run->EXTRA + n_extra;
  int rc = SH_OK;
#define TRY_RUN(expr)     \
  do {                    \
    rc = (expr);          \
    if (rc != SH_OK) {    \
      run->free_all();    \
      delete run;         \
      return rc;          \
    }                     \
  } while (0)
#define CHECK_RUN(expr)                                                                         \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess)                                                                       \
      TRY_RUN(sh_fail(ctx, SH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)));        \
  } while (0)
  TRY_RUN(sh_dev_alloc(ctx, 4 * run->dl_words, (void**)&run->d_aux));
  uint32_t* const d_ch = run->d_aux;
  uint32_t* const d_alphas = run->d_aux + run->ALPHAS;
  uint32_t* const d_roots = run->d_aux + run->ROOTS;
  if (d_ch_in) {
    CHECK_RUN(hipMemcpyAsync(d_ch, d_ch_in, 64, hipMemcpyDeviceToDevice, ctx->stream));
    CHECK_RUN(hipMemcpyAsync(run->d_aux + run->INIT, d_ch_in, 64, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    std::vector<uint32_t> h_ch(16, 0);
    memcpy(h_ch.data(), ch->digest, 32);
    h_ch[8] = ch->n_draws;
    void* staged;
    TRY_RUN(sh_stage(ctx, h_ch.data(), 64, &staged));
    CHECK_RUN(hipMemcpyAsync(d_ch, staged, 64, hipMemcpyDeviceToDevice, ctx->stream));
  }
  FriLayer& first = run->first;
  FriLayer& cur = run->cur;  // the evaluation being folded (not committed yet)
  FriLayer tmp;              // ping-pong buffer of a multi-step fold
  auto fail_tmp = [&](int r) {  // on an error, tmp (not yet owned by the run) is released too
    if (r != SH_OK) tmp.free_all();
    return r;
  };
  // the channel step of layer `layer` (mix its root, draw alpha, k squared powers), run by the
  // launch that computes the layer's root
  auto channel_step = [&](uint32_t layer, uint32_t k) {
    return sh_root_step{d_ch, d_alphas + 16 * layer, d_roots + 8 * layer, k};
  };
  auto alloc_eval = [&](FriLayer& l, uint32_t lg) -> int {
    l.log_size = lg;
    l.owner = ctx;
    int r = sh_dev_alloc(ctx, 16ull << lg, (void**)&l.buf);
    if (r != SH_OK) return r;
    for (int c = 0; c < 4; c++) l.cols[c] = l.buf + ((uint64_t)c << lg);
    return SH_OK;
  };
  auto out_cols = [](FriLayer& l) {  // the 4 coordinate arrays of an owned evaluation, writable
    std::array<uint32_t*, 4> d;
    for (int c = 0; c < 4; c++) d[c] = l.buf + ((uint64_t)c << l.log_size);
    return d;
  };
  // commit_first_layer (fri.rs:130-140): packed leaves when fold_step > 1
  for (int c = 0; c < 4; c++) first.cols[c] = column[c];
  first.log_size = log_size;
  first.fold_step = fs;
  first.pack = log_size >= 2 && fs > 1;
  {
    const sh_root_step st = channel_step(0, fs);
    TRY_RUN(commit_fri_layer(ctx, column, log_size, first.pack, first.tree, false, &st));
  }
  // commit_inner_layers (fri.rs:145-196): the circle-to-line fold with alpha, then fold_step - 1
  // line folds with alpha^2, alpha^4, ... (one multi-step launch)
  uint32_t line_log = log_size - 1;
  uint32_t line_init = sh::subgroup_gen(line_log + 2);  // half_odds(line_log)
  // with fold_step 1 a fold also hashes the leaves of the layer it makes, when that layer is
  // committed by its own launches (the tail workgroup hashes its layers itself)
  auto leaves_by_fold = [&](uint32_t lg) {
    return fs == 1 && lg > last_log && !(lg <= kFriTailMaxLog && lg - last_log <= kFriTailMaxLayers);
  };
  auto prepare_leaves = [&](FriLayer& l, uint32_t lg, uint32_t** leaves) -> int {
    *leaves = nullptr;
    if (!leaves_by_fold(lg)) return SH_OK;
    SH_TRY(alloc_tree(ctx, lg, l.tree));
    l.leaves_done = true;
    *leaves = (uint32_t*)l.tree.layers[lg];
    return SH_OK;
  };
  TRY_RUN(alloc_eval(cur, log_size - 1));
  {
    uint32_t* leaves;
    TRY_RUN(prepare_leaves(cur, log_size - 1, &leaves));
    TRY_RUN(sh_fold_circle_dev(ctx, tw, column, log_size, d_alphas, out_cols(cur).data(), leaves));
  }
  if (fs > 1) {
    TRY_RUN(fail_tmp(alloc_eval(tmp, line_log - (fs - 1))));
    TRY_RUN(fail_tmp(
        sh_fold_line_dev(ctx, tw, cur.cols, line_init, line_log, d_alphas + 4, fs - 1, out_cols(tmp).data())));
    cur.free_all();  // stream-ordered release: the fold that reads it is already queued
    std::swap(cur, tmp);
    for (uint32_t j = 0; j < fs - 1; j++) line_init = sh::idx_reduce((uint64_t)line_init * 2);
    line_log -= fs - 1;
  }
  uint32_t n_layers = 1;
  while (line_log > last_log) {
    if (fs == 1 && line_log <= kFriTailMaxLog && line_log - last_log <= kFriTailMaxLayers) {
      // every remaining layer in one launch (k_fri_tail)
      if (!sh::is_doubling_of(line_init, line_log, tw->root_initial_index, tw->root_log_size))
        TRY_RUN(sh_fail(ctx, SH_ERR_INVALID, "line domain (initial %u, log %u) is not in the twiddle tower", line_init,
                        line_log));
      FriTailArgs ta = {};
      ta.n_layers = line_log - last_log;
      ta.ch = d_ch;
      ta.alphas = d_alphas;
      ta.roots = d_roots;
      for (uint32_t j = 0; j < ta.n_layers; j++) {
        const uint32_t lg = line_log - j;
        cur.fold_step = 1;
        cur.pack = false;
        DeviceTree& t = cur.tree;
        t.L = lg;
        t.owner = ctx;
        TRY_RUN(sh_dev_alloc(ctx, 32ull * ((2ull << lg) - 1), (void**)&t.buf));
        t.layers.resize(lg + 1);
        for (uint32_t k = 0; k <= lg; k++) t.layers[k] = t.buf + 32ull * ((1ull << k) - 1);
        TRY_RUN(fail_tmp(alloc_eval(tmp, lg - 1)));
        ta.layer[j] = {cur.buf, t.buf, tw->itw + (1ull << tw->root_log_size) - (1ull << lg), lg, n_layers + j};
        ta.next_eval[j] = tmp.buf;
        run->inner.push_back(cur);  // keeps this layer's evaluation and tree for the decommitment
        cur = FriLayer();
        std::swap(cur, tmp);
        line_init = sh::idx_reduce((uint64_t)line_init * 2);
      }
      {
        SpanGuard span(ctx, "fri_tail");
        k_fri_tail<<<1, kFriTailThreads, 0, ctx->stream>>>(ta);
        TRY_RUN(sh_launch_check(ctx, "fri_tail"));
      }
      n_layers += ta.n_layers;
      line_log = last_log;
      break;
    }
    // FriInnerLayerProver::new + mix root + draw alpha + fold_line; the last inner layer folds
    // exactly down to the last layer's domain (fri.rs:175-194)
    const uint32_t k = line_log > last_log + fs ? fs : line_log - last_log;
    cur.fold_step = k;
    cur.pack = line_log >= 2 && k > 1;
    const sh_root_step st = channel_step(n_layers, k);
    if (cur.leaves_done)
      TRY_RUN(commit_columns(ctx, {cur.cols[0], cur.cols[1], cur.cols[2], cur.cols[3]}, line_log, cur.tree, false,
                             true, &st));
    else
      TRY_RUN(commit_fri_layer(ctx, cur.cols, line_log, cur.pack, cur.tree, false, &st));
    TRY_RUN(fail_tmp(alloc_eval(tmp, line_log - k)));
    uint32_t* next_leaves = nullptr;
    if (k == 1) TRY_RUN(fail_tmp(prepare_leaves(tmp, line_log - 1, &next_leaves)));
    rc = sh_fold_line_dev(ctx, tw, cur.cols, line_init, line_log, d_alphas + 16 * n_layers, k, out_cols(tmp).data(),
                          next_leaves);
    run->inner.push_back(cur);  // keeps this layer's evaluation and tree for the decommitment
    cur = FriLayer();
    std::swap(cur, tmp);
    TRY_RUN(rc);
    for (uint32_t j = 0; j < k; j++) line_init = sh::idx_reduce((uint64_t)line_init * 2);
    line_log -= k;
    n_layers++;
  }
  run->n_layers = n_layers;
  run->last_line_init = line_init;
  if (run->device_tail) {
    // commit_last_layer (fri.rs:208-226) and the proof of work (fri.rs:229-244) on the device
    const auto& itw = last_layer_itw(line_init, last_log);
    void* d_itw;
    TRY_RUN(sh_stage(ctx, itw.data(), 4 * itw.size(), &d_itw));
    k_fri_last_layer<<<1, 256, 0, ctx->stream>>>(cur.cols[0], cur.cols[1], cur.cols[2], cur.cols[3],
                                                 (const uint32_t*)d_itw, last_log, p->log_last_layer_degree_bound,
                                                 sh::minv(1u << last_log), run->d_aux + run->LASTC,
                                                 run->d_aux + run->FLAG);
    TRY_RUN(sh_launch_check(ctx, "fri_last_layer"));
    TRY_RUN(sh_dch_mix_impl(ctx, d_ch, run->d_aux + run->LASTC, 4ull << p->log_last_layer_degree_bound));
    TRY_RUN(sh_grind_dchannel(ctx, d_ch, p->pow_bits, run->d_aux + run->CTL));
  }
  // one download: the device channel, the alphas and roots of every layer, the device tail's
  // words, the last layer, the extra words (packed here; the decommitment copies them out)
  {
    PackArgs pa = {};
    for (int c = 0; c < 4; c++) {
      pa.src[c] = cur.cols[c];
      pa.len[c] = 1u << last_log;
      pa.dst_off[c] = (uint32_t)(run->EVAL + ((uint64_t)c << last_log));
    }
    pa.n = 4;
    if (n_extra) {
      pa.src[4] = d_extra;
      pa.len[4] = (uint32_t)n_extra;
      pa.dst_off[4] = (uint32_t)run->EXTRA;
      pa.n = 5;
    }
    const uint64_t longest = std::max<uint64_t>(1ull << last_log, n_extra);
    k_pack<<<(unsigned)std::min<uint64_t>(256, (longest + 255) / 256), 256, 0, ctx->stream>>>(pa, run->d_aux);
    TRY_RUN(sh_launch_check(ctx, "pack"));
  }
#undef TRY_RUN
#undef CHECK_RUN
  *out = run;
  return SH_OK;
}

// FriProver::decommit (crates/stwo/src/prover/fri.rs:229-284) after fri_commit_impl: waits for the
// commit phase's download, replays the channel on the host (ch: the final state), interpolates
// the last layer and grinds the proof of work (or checks the device's), draws the queries and
// reads every layer's witness values and Merkle witness nodes, and the committed trees'
// openings, with one gather.  Leaves the run's buffers in place (a replayed graph writes them
// again).
static int fri_decommit_impl(sh_ctx* ctx, sh_fri_run* run, sh_channel* ch, uint32_t* h_extra, sh_tree_opening* trees,
                             uint32_t n_trees, uint64_t* felts, uint64_t cap, uint64_t* n_felts,
                             uint32_t* query_positions, uint64_t* n_query_positions) {
  if (!ctx || !run || !ch || !n_felts || !n_query_positions || (run->n_extra && !h_extra) || (n_trees && !trees))
    return SH_ERR_INVALID;
  SH_REQUIRE(ctx, run->owner == ctx && run->d_aux, "FRI run of another context, or freed");
  const sh_fri_params* p = &run->p;
  const uint32_t fs = run->fs, last_log = run->last_log, log_size = run->log_size, max_layers = run->max_layers;
  uint8_t* h_dl;
  SH_TRY(sh_host_xfer(ctx, 4 * run->dl_words, &h_dl));
  SH_CHECK_HIP(ctx, hipMemcpyAsync(h_dl, run->d_aux, 4 * run->dl_words, hipMemcpyDeviceToHost, ctx->stream));
  SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t* aux = (const uint32_t*)h_dl;
  if (run->n_extra) memcpy(h_extra, aux + run->EXTRA, 4 * run->n_extra);
  if (run->from_dch) {
    memcpy(ch->digest, aux + run->INIT, 32);
    ch->n_draws = aux[run->INIT + 8];
  } else {
    *ch = run->ch_in;
  }
  std::vector<const FriLayer*> layers_all;
  layers_all.push_back(&run->first);
  for (auto& l : run->inner) layers_all.push_back(&l);
  SH_REQUIRE(ctx, layers_all.size() == run->n_layers, "FRI run holds %zu layers, expected %u", layers_all.size(),
             run->n_layers);
  std::vector<Hash> roots(run->n_layers);
  for (uint32_t l = 0; l < run->n_layers; l++) {
    memcpy(roots[l].data(), aux + run->ROOTS + 8 * l, 32);
    sh_channel_mix_root(ch, roots[l].data());
    uint32_t a[4];
    sh_channel_draw_secure_felt(ch, a);
    if (memcmp(a, aux + run->ALPHAS + 16 * l, 16) != 0)
      return sh_fail(ctx, SH_ERR_INVALID, "device-side FRI channel diverged from the host channel at layer %u", l);
  }
  // commit_last_layer (fri.rs:208-226)
  std::vector<sh::qm31> last_coeffs;
  {
    const uint32_t* words = aux + run->EVAL;
    std::vector<sh::qm31> v(1ull << last_log);
    for (uint64_t i = 0; i < v.size(); i++)
      v[i] = {{words[i], words[v.size() + i]}, {words[2 * v.size() + i], words[3 * v.size() + i]}};
    bit_reverse_q(v);  // LineEvaluation::interpolate (crates/stwo/src/prover/line.rs:54-65)
    line_ifft(v, run->last_line_init, last_log);
    uint32_t inv = sh::minv((uint32_t)v.size());
    for (auto& x : v) x = sh::qmulm(x, inv);
    bit_reverse_q(v);  // into_ordered_coefficients
    const uint64_t bound = 1ull << p->log_last_layer_degree_bound;
    for (uint64_t i = bound; i < v.size(); i++)
      if (!sh::qeq(v[i], sh::qzero())) return sh_fail(ctx, SH_ERR_INVALID, "invalid degree");
    v.resize(bound);
    bit_reverse_q(v);  // LinePoly::from_ordered_coefficients stores bit-reversed coefficients
    last_coeffs = v;
    std::vector<uint32_t> fw(4 * v.size());
    for (size_t i = 0; i < v.size(); i++) sh::qstore(v[i], &fw[4 * i]);
    if (run->device_tail &&
        (aux[run->FLAG] != 0 || memcmp(fw.data(), aux + run->LASTC, 4 * fw.size()) != 0))
      return sh_fail(ctx, SH_ERR_INVALID, "device-side last FRI layer diverged from the host's");
    sh_channel_mix_felts(ch, fw.data(), v.size());
  }
  // decommit (fri.rs:229-251): the proof of work
  uint64_t nonce = 0;
  if (run->device_tail) {
    nonce = (uint64_t)aux[run->CTL] | ((uint64_t)aux[run->CTL + 1] << 32);
    if (!pow_ok(ch->digest, p->pow_bits, nonce))
      return sh_fail(ctx, SH_ERR_NOT_FOUND, "device proof of work returned no valid nonce");
  } else {
    SH_TRY(sh_grind(ctx, ch->digest, p->pow_bits, &nonce));
  }
  sh_channel_mix_u64(ch, nonce);
  if (run->device_tail && memcmp(ch->digest, aux, 32) != 0)
    return sh_fail(ctx, SH_ERR_INVALID, "device-side FRI channel diverged from the host channel after the nonce");
  std::vector<uint32_t> q(p->n_queries);
  sh_channel_draw_queries(ch, log_size, p->n_queries, q.data());
  std::sort(q.begin(), q.end());
  q.erase(std::unique(q.begin(), q.end()), q.end());
  // decommit_on_queries (fri.rs:253-284): queries fold by each layer's own step.  Every layer's
  // witness values and Merkle witness nodes are read with one gather.
  struct Plan {
    std::vector<uint32_t> witness_pos;
    std::vector<std::pair<uint32_t, uint32_t>> nodes;
  };
  std::vector<Plan> plans(layers_all.size());
  std::vector<const uint32_t*> gc;
  std::vector<uint32_t> gr;
  std::vector<const uint8_t*> gn;
  gc.reserve(4ull * q.size() * layers_all.size() << fs);
  gr.reserve(4ull * q.size() * layers_all.size() << fs);
  gn.reserve(32ull * q.size() * layers_all.size());
  std::vector<uint32_t> lq = q;
  for (size_t li = 0; li < layers_all.size(); li++) {
    const FriLayer& L = *layers_all[li];
    const uint32_t k = L.fold_step;
    // compute_decommitment_positions_and_witness_evals (fri.rs:408-438)
    std::vector<uint32_t> mpos;
    for (size_t i = 0; i < lq.size();) {
      const uint32_t start = (lq[i] >> k) << k;
      size_t j = i;
      while (j < lq.size() && (lq[j] >> k) == (lq[i] >> k)) j++;
      size_t qi = i;
      for (uint32_t pos = start; pos < start + (1u << k); pos++) {
        if (qi < j && lq[qi] == pos)
          qi++;
        else
          plans[li].witness_pos.push_back(pos);
        const uint32_t m = L.pack ? pos >> 2 : pos;
        if (mpos.empty() || mpos.back() != m) mpos.push_back(m);
      }
      i = j;
    }
    plans[li].nodes = merkle_witness_nodes(L.tree.L, mpos);
    for (uint32_t pos : plans[li].witness_pos)
      for (int c = 0; c < 4; c++) {
        gc.push_back(L.cols[c]);
        gr.push_back(pos);
      }
    for (auto& nd : plans[li].nodes) gn.push_back(L.tree.layers[nd.first] + 32ull * nd.second);
    for (auto& x : lq) x >>= k;  // Queries::fold (crates/stwo/src/core/queries.rs:48-55)
    lq.erase(std::unique(lq.begin(), lq.end()), lq.end());
  }
  // the committed trees' openings (CommitmentTreeProver::decommit, crates/stwo/src/prover/pcs/
  // mod.rs:286-302) join the same gather
  std::vector<LiftedJob> jobs;
  std::vector<uint64_t> tree_node0(n_trees), tree_nodes(n_trees), tree_npos(n_trees);
  for (uint32_t t = 0; t < n_trees; t++) {
    sh_tree_opening& T = trees[t];
    std::vector<uint32_t> pos;
    if (!T.preprocessed) {
      pos = q;
    } else if (T.L > 0) {
      // prepare_preprocessed_query_positions (crates/stwo/src/core/pcs/utils.rs:185-203)
      for (uint32_t x : q)
        pos.push_back(log_size < T.L ? ((x >> 1) << (T.L - log_size + 1)) + (x & 1)
                                     : ((x >> (log_size - T.L + 1)) << 1) + (x & 1));
    }
    tree_node0[t] = gn.size();
    tree_npos[t] = pos.size();
    const size_t jobs_before = jobs.size();
    SH_TRY(plan_tree_opening(ctx, T.d_cols, T.log_sizes, T.n_cols, T.d_layers, T.L, std::move(pos), gn, jobs,
                             &tree_nodes[t]));
    if (tree_nodes[t] > T.witness_cap || (jobs.size() > jobs_before && !T.values))
      return sh_fail(ctx, SH_ERR_INVALID, "tree %u: witness of %llu hashes exceeds the buffer or no value buffer", t,
                     (unsigned long long)tree_nodes[t]);
  }
  std::vector<uint32_t> gw, gnw, gv;
  SH_TRY(gather_all(ctx, gc, gr, gn, jobs, gw, gnw, gv));
  {
    uint64_t vo = 0;
    for (uint32_t t = 0; t < n_trees; t++) {
      sh_tree_opening& T = trees[t];
      T.n_positions = tree_npos[t];
      T.n_witness = tree_nodes[t];
      if (tree_nodes[t]) memcpy(T.witness, gnw.data() + 8 * tree_node0[t], 32 * tree_nodes[t]);
      const uint64_t nv = tree_npos[t] ? T.n_cols * tree_npos[t] : 0;
      if (nv) memcpy(T.values, gv.data() + vo, 4 * nv);
      vo += nv;
    }
  }
  // FriProof (crates/cairo-serialize/src/serialize.rs:79-90): the first layer, the inner layers
  // (length-prefixed), each FriLayerProof (serialize.rs:66-77) = fri_witness, decommitment,
  // commitment, written straight from the gathered words
  Felts f;
  f.v.reserve(gw.size() + gnw.size() + 16 * layers_all.size() + 8 * last_coeffs.size() + 16);
  size_t off = 0, noff = 0;
  for (size_t li = 0; li < layers_all.size(); li++) {
    if (li == 1) f.u(layers_all.size() - 1);
    const size_t nwit = plans[li].witness_pos.size(), nnodes = plans[li].nodes.size();
    f.u(nwit);
    f.v.insert(f.v.end(), gw.begin() + off, gw.begin() + off + 4 * nwit);
    off += 4 * nwit;
    f.u(nnodes);
    f.v.insert(f.v.end(), gnw.begin() + noff, gnw.begin() + noff + 8 * nnodes);
    noff += 8 * nnodes;
    f.h(roots[li]);
  }
  if (layers_all.size() == 1) f.u(0);
  f.u(last_coeffs.size());
  for (auto& x : last_coeffs) f.q(x);
  f.u(p->log_last_layer_degree_bound);
  f.u(nonce);
  (void)max_layers;
  *n_felts = f.v.size();
  *n_query_positions = q.size();
  SH_REQUIRE(ctx, f.v.size() <= cap, "FRI proof of %zu felts exceeds the buffer (%llu)", f.v.size(),
             (unsigned long long)cap);
  memcpy(felts, f.v.data(), 8 * f.v.size());
  if (query_positions) memcpy(query_positions, q.data(), 4 * q.size());
  return SH_OK;
}

static int fri_prove_impl(sh_ctx* ctx, const sh_twiddles* tw, sh_channel* ch, const uint32_t* d_ch_in,
                          const uint32_t* d_extra, uint64_t n_extra, uint32_t* h_extra, sh_tree_opening* trees,
                          uint32_t n_trees, const sh_fri_params* p, const uint32_t* const column[4],
                          uint32_t log_size, uint64_t* felts, uint64_t cap, uint64_t* n_felts,
                          uint32_t* query_positions, uint64_t* n_query_positions) {
  if (!ctx || !ch || !n_felts || !n_query_positions || (n_extra && !h_extra)) return SH_ERR_INVALID;
  sh_fri_run* run = nullptr;
  SH_TRY(fri_commit_impl(ctx, tw, d_ch_in ? nullptr : ch, d_ch_in, d_extra, n_extra, p, column, log_size, &run));
  const int rc = fri_decommit_impl(ctx, run, ch, h_extra, trees, n_trees, felts, cap, n_felts, query_positions,
                                   n_query_positions);
  run->free_all();
  delete run;
  return rc;
}

extern "C" int sh_fri_prove(sh_ctx* ctx, const sh_twiddles* tw, sh_channel* ch, const sh_fri_params* p,
                            const uint32_t* const column[4], uint32_t log_size, uint64_t* felts, uint64_t cap,
                            uint64_t* n_felts, uint32_t* query_positions, uint64_t* n_query_positions) {
  return fri_prove_impl(ctx, tw, ch, nullptr, nullptr, 0, nullptr, nullptr, 0, p, column, log_size, felts, cap,
                        n_felts, query_positions, n_query_positions);
}

extern "C" int sh_fri_prove_dch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* d_ch, sh_channel* ch_out,
                                const uint32_t* d_extra, uint64_t n_extra, uint32_t* h_extra,
                                sh_tree_opening* trees, uint32_t n_trees, const sh_fri_params* p,
                                const uint32_t* const column[4], uint32_t log_size, uint64_t* felts, uint64_t cap,
                                uint64_t* n_felts, uint32_t* query_positions, uint64_t* n_query_positions) {
  if (!d_ch) return SH_ERR_INVALID;
  return fri_prove_impl(ctx, tw, ch_out, d_ch, d_extra, n_extra, h_extra, trees, n_trees, p, column, log_size, felts,
                        cap, n_felts, query_positions, n_query_positions);
}

extern "C" int sh_fri_commit_dch(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* d_ch, const uint32_t* d_extra,
                                 uint64_t n_extra, const sh_fri_params* p, const uint32_t* const column[4],
                                 uint32_t log_size, sh_fri_run** out) {
  if (!d_ch) return SH_ERR_INVALID;
  return fri_commit_impl(ctx, tw, nullptr, d_ch, d_extra, n_extra, p, column, log_size, out);
}

extern "C" int sh_fri_decommit_dch(sh_ctx* ctx, sh_fri_run* run, sh_channel* ch_out, uint32_t* h_extra,
                                   sh_tree_opening* trees, uint32_t n_trees, uint64_t* felts, uint64_t cap,
                                   uint64_t* n_felts, uint32_t* query_positions, uint64_t* n_query_positions) {
  if (run && !run->from_dch) return SH_ERR_INVALID;
  return fri_decommit_impl(ctx, run, ch_out, h_extra, trees, n_trees, felts, cap, n_felts, query_positions,
                           n_query_positions);
}

extern "C" int sh_fri_run_free(sh_ctx* ctx, sh_fri_run* run) {
  if (!ctx) return SH_ERR_INVALID;
  if (!run) return SH_OK;
  run->free_all();
  delete run;
  return SH_OK;
}

// ------------------------------------------------------------------ DEEP quotient driver
extern "C" int sh_fri_quotients(sh_ctx* ctx, const sh_twiddles* tw, const uint32_t* const* d_cols,
                                const uint32_t* col_log_sizes, uint64_t n_cols, const uint32_t* sample_offsets,
                                const uint32_t* sample_points, const uint32_t* sample_values,
                                const uint32_t random_coeff[4], uint32_t lifting_log_size, uint32_t log_blowup,
                                uint32_t* const d_out[4]) {
  if (!ctx || !tw || !d_out || !random_coeff || (n_cols && (!d_cols || !col_log_sizes || !sample_offsets)))
    return SH_ERR_INVALID;
  struct S {
    sh::qpt point;
    sh::qm31 value, rc;
  };
  const uint32_t L = lifting_log_size;
  // build_samples_with_randomness_and_periodicity (crates/stwo/src/core/pcs/quotients.rs:274-319)
  sh::qm31 pw = sh::qone(), alpha = sh::qload(random_coeff);
  auto next_pow = [&]() {
    sh::qm31 r = pw;
    pw = sh::qmul(pw, alpha);
    return r;
  };
  std::vector<std::vector<S>> samples(n_cols);
  for (uint64_t i = 0; i < n_cols; i++) {
    uint32_t b = sample_offsets[i], e = sample_offsets[i + 1];
    if (e == b) continue;
    SH_REQUIRE(ctx, col_log_sizes[i] <= L, "column log size above the lifting size");
    if (e - b == 2) {
      // period generator: CanonicCoset(L).step() doubled log_size times
      uint32_t idx = sh::idx_reduce((uint64_t)sh::subgroup_gen(L) << col_log_sizes[i]);
      sh::mpt g = sh::index_to_point_host(idx);
      sh::qpt pt = {sh::qload(sample_points + 8 * (b + 1)), sh::qload(sample_points + 8 * (b + 1) + 4)};
      sh::qpt shifted = sh::qpadd(pt, {sh::qfromm(g.x), sh::qfromm(g.y)});
      samples[i].push_back({shifted, sh::qload(sample_values + 4 * (b + 1)), next_pow()});
    }
    for (uint32_t k = b; k < e; k++)
      samples[i].push_back({{sh::qload(sample_points + 8 * k), sh::qload(sample_points + 8 * k + 4)},
                            sh::qload(sample_values + 4 * k), next_pow()});
  }
  // accumulate_numerators per column log size, ascending (quotient_ops.rs:108-122)
  struct Acc {
    sh::qpt point;
    uint32_t log;
    uint32_t* buf;
    sh::qm31 fl;
    bool fl_applied;  // the numerator took its first-linear term already
  };
  auto key = [](const sh::qpt& p) {
    return std::array<uint32_t, 8>{p.x.a.a, p.x.a.b, p.x.b.a, p.x.b.b, p.y.a.a, p.y.a.b, p.y.b.a, p.y.b.b};
  };
  // a point sampled by columns of one size only, the lifting size, has its numerator at the
  // combine's size: the numerator kernel subtracts the first-linear term and the combine skips it
  std::map<std::array<uint32_t, 8>, std::set<uint32_t>> point_sizes;
  for (uint64_t i = 0; i < n_cols; i++)
    for (auto& smp : samples[i]) point_sizes[key(smp.point)].insert(col_log_sizes[i]);
  std::vector<Acc> accs;
  std::vector<uint32_t> sizes(col_log_sizes, col_log_sizes + n_cols);
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  int rc = SH_OK;
  auto free_accs = [&]() {
    for (auto& a : accs)
      if (a.buf) sh_dev_free(ctx, a.buf);
  };
  for (uint32_t lg : sizes) {
    std::vector<const uint32_t*> gcols;
    std::vector<uint64_t> gidx;
    for (uint64_t i = 0; i < n_cols; i++)
      if (col_log_sizes[i] == lg) {
        gidx.push_back(i);
        gcols.push_back(d_cols[i]);
      }
    // ColumnSampleBatch::new_vec: group by point in first-occurrence order
    std::vector<sh::qpt> pts;
    std::vector<std::vector<std::tuple<uint32_t, sh::qm31, sh::qm31>>> terms;
    for (uint32_t ci = 0; ci < gidx.size(); ci++)
      for (auto& smp : samples[gidx[ci]]) {
        size_t k = 0;
        while (k < pts.size() && !(sh::qeq(pts[k].x, smp.point.x) && sh::qeq(pts[k].y, smp.point.y))) k++;
        if (k == pts.size()) {
          pts.push_back(smp.point);
          terms.emplace_back();
        }
        terms[k].push_back({ci, smp.value, smp.rc});
      }
    if (pts.empty()) continue;
    std::vector<uint32_t> bp, off{0}, tidx, tval, trc;
    for (size_t k = 0; k < pts.size(); k++) {
      uint32_t w[8];
      sh::qstore(pts[k].x, w);
      sh::qstore(pts[k].y, w + 4);
      bp.insert(bp.end(), w, w + 8);
      for (auto& t : terms[k]) {
        tidx.push_back(std::get<0>(t));
        uint32_t v[4], r[4];
        sh::qstore(std::get<1>(t), v);
        sh::qstore(std::get<2>(t), r);
        tval.insert(tval.end(), v, v + 4);
        trc.insert(trc.end(), r, r + 4);
      }
      off.push_back((uint32_t)tidx.size());
    }
    SH_REQUIRE(ctx, lg >= log_blowup + 1, "column too small for the blowup");
    const uint32_t plog = lg - log_blowup;
    std::vector<uint32_t*> outs(4 * pts.size());
    for (size_t k = 0; k < pts.size(); k++) {
      uint32_t* buf = nullptr;
      if ((rc = sh_dev_alloc(ctx, 16ull << plog, (void**)&buf)) != SH_OK) {
        free_accs();
        return rc;
      }
      const auto& ps = point_sizes[key(pts[k])];
      accs.push_back({pts[k], plog, buf, sh::qzero(), ps.size() == 1 && lg == L});
      for (int c = 0; c < 4; c++) outs[4 * k + c] = buf + ((uint64_t)c << plog);
    }
    std::vector<uint32_t> fl(4 * pts.size());
    std::vector<uint8_t> apply(pts.size());
    for (size_t k = 0; k < pts.size(); k++) apply[k] = accs[accs.size() - pts.size() + k].fl_applied;
    rc = sh_accumulate_numerators_fl(ctx, gcols.data(), gcols.size(), lg, log_blowup, bp.data(), off.data(),
                                     tidx.data(), tval.data(), trc.data(), pts.size(), outs.data(), fl.data(),
                                     apply.data());
    if (rc != SH_OK) {
      free_accs();
      return rc;
    }
    for (size_t k = 0; k < pts.size(); k++) accs[accs.size() - pts.size() + k].fl = sh::qload(&fl[4 * k]);
  }
  // group per point, sorted by (x, y) (quotient_ops.rs:127-152); stable
  std::vector<size_t> order(accs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key(accs[a].point) < key(accs[b].point); });
  std::vector<uint32_t*> lifted;
  std::vector<const uint32_t*> parts;
  std::vector<uint32_t> plogs, ppts, pfls;
  for (size_t i = 0; i < order.size() && rc == SH_OK;) {
    size_t j = i;
    while (j < order.size() && key(accs[order[j]].point) == key(accs[order[i]].point)) j++;
    sh::qm31 fl = sh::qzero();
    std::vector<const uint32_t*> cols;
    std::vector<uint32_t> logs;
    for (size_t k = i; k < j; k++) {
      const Acc& a = accs[order[k]];
      if (!a.fl_applied) fl = sh::qadd(fl, a.fl);
      for (int c = 0; c < 4; c++) cols.push_back(a.buf + ((uint64_t)c << a.log));
      logs.push_back(a.log);
    }
    const Acc& top = accs[order[j - 1]];
    const uint32_t* pcols[4];
    uint32_t plog = top.log;
    if (j - i == 1) {
      for (int c = 0; c < 4; c++) pcols[c] = top.buf + ((uint64_t)c << top.log);
    } else {
      uint32_t* buf = nullptr;
      if ((rc = sh_dev_alloc(ctx, 16ull << plog, (void**)&buf)) != SH_OK) break;
      lifted.push_back(buf);
      uint32_t* o[4];
      for (int c = 0; c < 4; c++) o[c] = buf + ((uint64_t)c << plog);
      rc = sh_lift_and_accumulate(ctx, cols.data(), logs.data(), logs.size(), o);
      for (int c = 0; c < 4; c++) pcols[c] = o[c];
    }
    for (int c = 0; c < 4; c++) parts.push_back(pcols[c]);
    plogs.push_back(plog);
    auto kk = key(top.point);
    ppts.insert(ppts.end(), kk.begin(), kk.end());
    uint32_t fw[4];
    sh::qstore(fl, fw);
    pfls.insert(pfls.end(), fw, fw + 4);
    i = j;
  }
  if (rc == SH_OK)
    rc = sh_compute_quotients_and_combine(ctx, tw, parts.data(), plogs.data(), ppts.data(), pfls.data(), plogs.size(),
                                          L, log_blowup, d_out);
  // the temporaries go back to the context pool: later users are ordered after this work
  for (auto b : lifted) sh_dev_free(ctx, b);
  free_accs();
  return rc;
}
