// Blake2s Merkle trees: lifted leaves (MerkleOpsLifted::build_leaves), inner
// layers (build_next_layer) and the mixed-degree tree (MerkleOps::commit_on_layer).
//
// Lifted leaves (crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:46-90):
// with columns sorted ascending by size, leaf i of a tree of height L is
// Blake2s over LE32(col_c[((i >> (L - n_c + 1)) << 1) + (i & 1)]) for every
// column c in order.  Rows of a smaller column group share their hash state
// with every row they lift to, so each size group is absorbed once at its own
// size (the SIMD backend does the same, crates/stwo/src/prover/backend/simd/
// blake2s_lifted.rs:95-145): one launch per size group, one lane per row, the
// partial state (chaining value + pending words of the unfinished block) kept
// in HBM between groups, structure-of-arrays for coalescing.
//
// Streaming rule (the `blake2` crate): a full block is compressed only when
// more data follows; the last block, full or not, is compressed with the final
// flag, except that a group followed by another compresses its own last full block: between
// groups 0..15 words are pending.
#include "../../include/stwo_hip.h"
#include "blake2s.h"
#include "blake2s_quad.h"
#include "common.h"

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kMsgStride = 17;  // odd stride: conflict-free per-lane message staging in LDS

struct LeafGroupArgs {
  const uint32_t* const* cols;  // device table of this group's columns
  const uint32_t* base;         // with stride != 0: column k is base + k * stride (no table reads)
  uint64_t stride;              // words between consecutive columns, 0 when they are not evenly spaced
  uint32_t cnt;                 // columns in the group
  uint32_t log_n;               // rows = 2^log_n
  const uint32_t* state_in;     // SoA state of the previous group (nullptr for the first)
  uint32_t in_log;              // previous group's row log
  uint32_t pend_in;             // pending words carried in (uniform)
  uint32_t t_in;                // bytes compressed before this group (uniform)
  uint32_t* state_out;          // SoA state for the next group (nullptr when final)
  uint32_t* hash_out;           // final hashes (when state_out == nullptr), 8 words per row
};

__device__ __forceinline__ void load_msg(const uint32_t* lm, uint32_t m[16]) {
#pragma unroll
  for (int j = 0; j < 16; j++) m[j] = lm[j];
}

// Column k of a leaf group: with STRIDED the columns lie at a constant stride in one allocation
// (the address is uniform arithmetic, so the load takes a scalar base and no table read waits).
template <bool STRIDED>
__device__ __forceinline__ const uint32_t* group_col(const LeafGroupArgs& a, uint32_t k) {
  if constexpr (STRIDED) return a.base + (uint64_t)k * a.stride;
  return a.cols[k];
}

// One lane per row.  Whole 16-column blocks are loaded straight into registers (16 independent
// loads in flight) while the previous block is compressed; only a partial block at either end of
// the group goes through the lane's LDS slot.
template <bool STRIDED>
__global__ __launch_bounds__(kThreads) void k_leaves_group(LeafGroupArgs a) {
  __shared__ uint32_t msg_lds[kThreads * kMsgStride];
  const uint32_t rows = 1u << a.log_n;
  const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= rows) return;  // no barriers below: each lane only touches its own LDS slot
  uint32_t* lm = msg_lds + threadIdx.x * kMsgStride;
  uint32_t h[8];
  uint32_t pend = a.pend_in, t = a.t_in;
  if (a.state_in) {
    const uint32_t in_rows = 1u << a.in_log;
    const uint32_t ratio = a.log_n - a.in_log;
    const uint32_t parent = ((r >> (ratio + 1)) << 1) + (r & 1);
#pragma unroll
    for (int w = 0; w < 8; w++) h[w] = a.state_in[w * in_rows + parent];
    for (uint32_t j = 0; j < pend; j++) lm[j] = a.state_in[(8 + j) * in_rows + parent];
  } else {
    sh::b2_init(h);
  }
  const uint32_t cnt = a.cnt;
  uint32_t k = 0;
  uint32_t m[16];
  bool full = false;  // m holds a full block that is compressed once more data is known to follow
  if (pend) {         // complete the block carried in from the previous group
    while (pend < 16 && k < cnt) lm[pend++] = group_col<STRIDED>(a, k++)[r];
    if (pend == 16) {
      load_msg(lm, m);
      full = true;
      pend = 0;
    }
  }
  while (k + 16 <= cnt) {  // (pend is 0 whenever this runs)
    uint32_t n[16];
#pragma unroll
    for (int j = 0; j < 16; j++) n[j] = group_col<STRIDED>(a, k + j)[r];
    if (full) {
      t += 64;
      sh::b2_compress(h, m, t, false);
    }
#pragma unroll
    for (int j = 0; j < 16; j++) m[j] = n[j];
    full = true;
    k += 16;
  }
  if (k < cnt) {  // a tail of fewer than 16 words
    if (full) {
      t += 64;
      sh::b2_compress(h, m, t, false);
      full = false;
    }
    while (k < cnt) lm[pend++] = group_col<STRIDED>(a, k++)[r];
  }
  if (a.state_out) {
    if (full) {  // more data follows in a later group: compress here, at the smaller size
      t += 64;
      sh::b2_compress(h, m, t, false);
    }
#pragma unroll
    for (int w = 0; w < 8; w++) a.state_out[w * rows + r] = h[w];
    for (uint32_t j = 0; j < pend; j++) a.state_out[(8 + j) * rows + r] = lm[j];
  } else {
    if (full) {
      t += 64;
    } else {
      t += 4 * pend;
      for (uint32_t j = pend; j < 16; j++) lm[j] = 0;
      load_msg(lm, m);
    }
    sh::b2_compress(h, m, t, true);
    uint4* o = (uint4*)(a.hash_out + 8 * r);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
  }
}

// Leaves of a tree taller than the largest column: leaf i copies the hash of
// its lifted row ((i >> (ratio + 1)) << 1) + (i & 1).
__global__ void k_lift_hashes(const uint32_t* src, uint32_t ratio, uint32_t* dst, uint32_t log_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << log_out)) return;
  uint32_t p = ((i >> (ratio + 1)) << 1) + (i & 1);
  const uint4* s = (const uint4*)(src + 8 * p);
  uint4* d = (uint4*)(dst + 8 * i);
  d[0] = s[0];
  d[1] = s[1];
}

// Hash of the empty stream (the single leaf of an empty tree).
__global__ void k_empty_hash(uint32_t* out) {
  uint32_t h[8], m[16];
  sh::b2_init(h);
#pragma unroll
  for (int j = 0; j < 16; j++) m[j] = 0;
  sh::b2_compress(h, m, 0, true);
#pragma unroll
  for (int w = 0; w < 8; w++) out[w] = h[w];
}

// Up to 8 inner layers per launch: level 0 hashes pairs of the input layer
// from HBM, each further level hashes pairs held in LDS.  outs[l] receives
// level l (2^(in_log - 1 - l) nodes).
struct LayerOuts {
  uint32_t* p[8];  // up to 8 output levels per launch
};

__global__ __launch_bounds__(kThreads) void k_next_layers(const uint32_t* prev, uint32_t in_log,
                                                          uint32_t levels, LayerOuts outs) {
  __shared__ uint32_t node[kThreads * 8];
  const uint32_t n0 = 1u << (in_log - 1);
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  uint32_t h[8];
  if (i < n0) {
    const uint4* s = (const uint4*)(prev + 16 * i);
    uint4 q0 = s[0], q1 = s[1], q2 = s[2], q3 = s[3];
    uint32_t m[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                      q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
    sh::b2_init(h);
    sh::b2_compress(h, m, 64, true);
    uint4* o = (uint4*)(outs.p[0] + 8 * i);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    *(uint4*)&node[threadIdx.x * 8] = make_uint4(h[0], h[1], h[2], h[3]);
    *(uint4*)&node[threadIdx.x * 8 + 4] = make_uint4(h[4], h[5], h[6], h[7]);
  }
  uint32_t width = kThreads;  // nodes of the current level held by this workgroup
  for (uint32_t l = 1; l < levels; l++) {
    __syncthreads();
    width >>= 1;
    uint32_t mine = 0;
    const uint32_t tid = threadIdx.x;
    const uint32_t gi = blockIdx.x * width + tid;  // node index in level l
    const uint32_t nl = n0 >> l;
    if (tid < width && gi < nl) {
      uint32_t m[16];
#pragma unroll
      for (int w = 0; w < 16; w += 4) {
        uint4 q = *(const uint4*)&node[tid * 16 + w];
        m[w] = q.x;
        m[w + 1] = q.y;
        m[w + 2] = q.z;
        m[w + 3] = q.w;
      }
      sh::b2_init(h);
      sh::b2_compress(h, m, 64, true);
      mine = 1;
    }
    __syncthreads();
    if (mine) {
      *(uint4*)&node[tid * 8] = make_uint4(h[0], h[1], h[2], h[3]);
      *(uint4*)&node[tid * 8 + 4] = make_uint4(h[4], h[5], h[6], h[7]);
      uint4* o = (uint4*)(outs.p[l] + 8 * gi);
      o[0] = make_uint4(h[0], h[1], h[2], h[3]);
      o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    }
  }
}

// ---- tree tops: four lanes per node (blake2s_quad.h)
constexpr uint32_t kQuadMaxLog = 13;  // levels of up to 2^13 nodes use four lanes per node
constexpr uint64_t kQuadMaxNodes = 1ull << kQuadMaxLog;

// Up to 7 levels below an input layer of 2^in_log hashes, four lanes per node: a workgroup of
// 256 lanes (64 quads) reduces 128 input hashes to 64 nodes, then its own 64 to 1 through LDS;
// level l (2^(in_log-1-l) nodes) goes to outs.p[l].  For fewer than 64 nodes one workgroup
// takes them all.  When st.ch is set (only for the one-workgroup launch that computes the root),
// the workgroup then runs the FRI channel step on the root: Blake2sChannel::mix_root and
// draw_secure_felt (crates/stwo/src/core/channel/blake2s.rs:33-117) with four lanes per hash,
// and the k squared powers of alpha (crates/stwo/src/prover/fri.rs:79-87).
constexpr uint32_t kQuadGroupNodes = 64;
__global__ __launch_bounds__(256) void k_next_layers_quad(const uint32_t* __restrict__ prev, uint32_t in_log,
                                                          uint32_t levels, LayerOuts outs, sh_root_step st) {
  __shared__ uint32_t buf_a[2 * kQuadGroupNodes * 8];
  __shared__ uint32_t buf_b[kQuadGroupNodes * 8];
  const uint32_t tid = threadIdx.x, q = tid & 3, quad = tid >> 2;
  uint32_t sidx[40];
  quad_sigma(q, sidx);
  uint32_t h_lo, h_hi, d0;
  quad_init(q, 64, h_lo, h_hi, d0);  // a node: one 64-byte block
  const uint32_t n_out0 = 1u << (in_log - 1);
  const uint32_t w0 = n_out0 < kQuadGroupNodes ? n_out0 : kQuadGroupNodes;  // this group's level-0 nodes
  // its 2 * w0 input hashes
  const uint4* src = (const uint4*)(prev + 16ull * w0 * blockIdx.x);
  for (uint32_t i = tid; i < 4 * w0; i += 256) ((uint4*)buf_a)[i] = src[i];
  __syncthreads();
  uint32_t* cur = buf_a;
  uint32_t* nxt = buf_b;
  for (uint32_t l = 0; l < levels; l++) {
    const uint32_t w = w0 >> l;
    if (quad < w) {
      uint32_t lo, hi;
      quad_node_hash(cur + 16 * quad, sidx, q, h_lo, h_hi, d0, lo, hi);
      nxt[8 * quad + q] = lo;
      nxt[8 * quad + 4 + q] = hi;
      uint32_t* o = outs.p[l] + 8ull * ((uint64_t)blockIdx.x * w + quad);
      o[q] = lo;
      o[4 + q] = hi;
    }
    __syncthreads();
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (st.ch) {  // uniform: the barriers below are reached by the whole workgroup
    __shared__ uint32_t msg[16], digest[8], draw[8];
    uint32_t dh_lo, dh_hi, d_draw;
    quad_init(q, 37, dh_lo, dh_hi, d_draw);  // digest || LE32(n_draws) || 0x00
    if (tid < 16) msg[tid] = tid < 8 ? st.ch[tid] : cur[tid - 8];
    if (tid < 8) st.root_out[tid] = cur[tid];
    __syncthreads();
    if (tid < 4) {  // digest = Blake2s(digest || root)
      uint32_t lo, hi;
      quad_node_hash(msg, sidx, q, h_lo, h_hi, d0, lo, hi);
      digest[q] = lo;
      digest[4 + q] = hi;
    }
    __syncthreads();
    if (tid < 16) msg[tid] = tid < 8 ? digest[tid] : 0u;  // n_draws = 0 after the mix
    __syncthreads();
    if (tid < 4) {  // the draw: Blake2s(digest || LE32(0) || 0x00)
      uint32_t lo, hi;
      quad_node_hash(msg, sidx, q, dh_lo, dh_hi, d_draw, lo, hi);
      draw[q] = lo;
      draw[4 + q] = hi;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t f[8], w[4];
      for (int i = 0; i < 8; i++) f[i] = sh::reduce_once((draw[i] & sh::P) + (draw[i] >> 31));  // M31::reduce(u32)
      for (int i = 0; i < 4; i++) w[i] = sh::madd(sh::madd(f[2 * i], f[2 * i + 1]), f[2 * i + 1]);
      sh::qm31 a = sh::qload(w);
      for (uint32_t j = 0; j < st.k; j++) {
        sh::qstore(a, st.alphas + 4 * j);
        a = sh::qmul(a, a);
      }
      for (int i = 0; i < 8; i++) st.ch[i] = digest[i];
      st.ch[8] = 1;
    }
  }
}

// Mixed-degree levels in one launch: level `log_size` (one lane per node) is
// Blake2s(prev[2i] || prev[2i+1] || LE32(col_0[i]) || ...) (crates/stwo/src/core/vcs/
// blake2_merkle.rs:13-29); each 16-word block of column values is loaded straight into registers
// (16 independent loads, coalesced across lanes) while the previous block is compressed.  Then
// up to kMixedExtra column-free levels above it are reduced inside the workgroup through LDS.
constexpr uint32_t kMixedExtra = 3;
struct MixedArgs {
  const uint32_t* prev;         // level log_size + 1, or nullptr at the bottom of the tree
  const uint32_t* const* cols;  // this level's columns (device table), cnt of them
  uint32_t cnt;
  uint32_t log_size;
  uint32_t extra;                   // column-free levels computed above this one
  uint32_t* out[kMixedExtra + 1];   // out[l]: level log_size - l
};

__global__ __launch_bounds__(kThreads) void k_mixed_levels(MixedArgs a) {
  __shared__ uint32_t node[kThreads * 8];
  const uint32_t n0 = 1u << a.log_size;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n0) {
    uint32_t h[8], m[16];
    sh::b2_init(h);
    bool full = false;  // m holds a full block, compressed once more data is known to follow
    uint32_t t = 0;
    if (a.prev) {
      const uint4* s = (const uint4*)(a.prev + 16ull * i);
      const uint4 q0 = s[0], q1 = s[1], q2 = s[2], q3 = s[3];
      m[0] = q0.x; m[1] = q0.y; m[2] = q0.z; m[3] = q0.w;
      m[4] = q1.x; m[5] = q1.y; m[6] = q1.z; m[7] = q1.w;
      m[8] = q2.x; m[9] = q2.y; m[10] = q2.z; m[11] = q2.w;
      m[12] = q3.x; m[13] = q3.y; m[14] = q3.z; m[15] = q3.w;
      full = true;
    }
    uint32_t k = 0;
    for (; k + 16 <= a.cnt; k += 16) {
      uint32_t nx[16];
#pragma unroll
      for (int j = 0; j < 16; j++) nx[j] = a.cols[k + j][i];
      if (full) {
        t += 64;
        sh::b2_compress(h, m, t, false);
      }
#pragma unroll
      for (int j = 0; j < 16; j++) m[j] = nx[j];
      full = true;
    }
    const uint32_t rem = a.cnt - k;
    if (rem) {
      uint32_t nx[16];
#pragma unroll
      for (int j = 0; j < 16; j++) nx[j] = (uint32_t)j < rem ? a.cols[k + j][i] : 0u;
      if (full) {
        t += 64;
        sh::b2_compress(h, m, t, false);
      }
#pragma unroll
      for (int j = 0; j < 16; j++) m[j] = nx[j];
      t += 4 * rem;
    } else if (full) {
      t += 64;
    } else {  // no children and no columns: the hash of nothing
#pragma unroll
      for (int j = 0; j < 16; j++) m[j] = 0;
    }
    sh::b2_compress(h, m, t, true);
    uint4* o = (uint4*)(a.out[0] + 8ull * i);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    *(uint4*)&node[threadIdx.x * 8] = make_uint4(h[0], h[1], h[2], h[3]);
    *(uint4*)&node[threadIdx.x * 8 + 4] = make_uint4(h[4], h[5], h[6], h[7]);
  }
  uint32_t width = kThreads;
  for (uint32_t l = 1; l <= a.extra; l++) {
    __syncthreads();
    width >>= 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t gi = blockIdx.x * width + tid;
    uint32_t h[8];
    const bool mine = tid < width && gi < (n0 >> l);
    if (mine) {
      uint32_t m[16];
#pragma unroll
      for (int w = 0; w < 16; w += 4) {
        const uint4 q = *(const uint4*)&node[tid * 16 + w];
        m[w] = q.x;
        m[w + 1] = q.y;
        m[w + 2] = q.z;
        m[w + 3] = q.w;
      }
      sh::b2_init(h);
      sh::b2_compress(h, m, 64, true);
    }
    __syncthreads();
    if (mine) {
      *(uint4*)&node[tid * 8] = make_uint4(h[0], h[1], h[2], h[3]);
      *(uint4*)&node[tid * 8 + 4] = make_uint4(h[4], h[5], h[6], h[7]);
      uint4* o = (uint4*)(a.out[l] + 8ull * gi);
      o[0] = make_uint4(h[0], h[1], h[2], h[3]);
      o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    }
  }
}

}  // namespace

// Leaves for columns already sorted ascending (host table h_cols).
int sh_build_leaves_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* log_sizes, uint64_t n,
                         uint32_t L, uint32_t* d_out, const hipEvent_t* group_ready) {
  if (n == 0) {
    SpanGuard span(ctx, "merkle_empty");
    k_empty_hash<<<1, 1, 0, ctx->stream>>>(d_out);
    return sh_launch_check(ctx, "empty_hash");
  }
  for (uint64_t i = 0; i < n; i++) {
    SH_REQUIRE(ctx, log_sizes[i] >= 1, "A column must be of length >= 2.");
    SH_REQUIRE(ctx, log_sizes[i] <= 30, "column log size %u too large", log_sizes[i]);
    if (i) SH_REQUIRE(ctx, log_sizes[i] >= log_sizes[i - 1], "columns must be sorted ascending by size");
  }
  SH_REQUIRE(ctx, L >= log_sizes[n - 1], "lifting_log_size %u < %u", L, log_sizes[n - 1]);
  void* d_tab;
  SH_TRY(sh_stage(ctx, h_cols, n * sizeof(void*), &d_tab));
  const uint32_t* const* d_cols = (const uint32_t* const*)d_tab;
  // group boundaries
  std::vector<std::pair<uint64_t, uint64_t>> groups;
  for (uint64_t s = 0; s < n;) {
    uint64_t e = s;
    while (e < n && log_sizes[e] == log_sizes[s]) e++;
    groups.push_back({s, e});
    s = e;
  }
  const uint32_t* state_in = nullptr;
  uint32_t in_log = 0;
  uint32_t pend = 0, t = 0;  // uniform stream position: pending words, bytes compressed
  int slot = SCR_MERKLE_STATE_A;
  uint32_t n_last = log_sizes[n - 1];
  for (size_t g = 0; g < groups.size(); g++) {
    LeafGroupArgs a;
    a.cols = d_cols + groups[g].first;
    a.cnt = (uint32_t)(groups[g].second - groups[g].first);
    // evenly spaced columns (one allocation, equal sizes) are addressed as base + k * stride
    a.base = h_cols[groups[g].first];
    a.stride = 0;
    if (a.cnt > 1 && h_cols[groups[g].first + 1] > a.base) {
      const uint64_t st = (uint64_t)(h_cols[groups[g].first + 1] - a.base);
      bool even = st >= (1ull << log_sizes[groups[g].first]);
      for (uint64_t i = groups[g].first + 2; i < groups[g].second && even; i++)
        even = h_cols[i] - h_cols[i - 1] == (ptrdiff_t)st;
      if (even) a.stride = st;
    }
    a.log_n = log_sizes[groups[g].first];
    a.state_in = state_in;
    a.in_log = in_log;
    a.pend_in = pend;
    a.t_in = t;
    bool last = g + 1 == groups.size();
    // replay the kernel's uniform stream bookkeeping
    for (uint32_t k = 0; k < a.cnt; k++) {
      if (pend == 16) {
        t += 64;
        pend = 0;
      }
      pend++;
    }
    if (!last && pend == 16) {
      t += 64;
      pend = 0;
    }
    uint64_t rows = 1ull << a.log_n;
    if (!last) {
      uint32_t pend_after = pend;
      void* st;
      SH_TRY(sh_scratch_get(ctx, slot, 4ull * (8 + pend_after) * rows, &st));
      a.state_out = (uint32_t*)st;
      a.hash_out = nullptr;
    } else {
      a.state_out = nullptr;
      if (L == n_last) {
        a.hash_out = d_out;
      } else {
        void* tmp;
        SH_TRY(sh_scratch_get(ctx, SCR_MERKLE_TMP, 32ull * rows, &tmp));
        a.hash_out = (uint32_t*)tmp;
      }
    }
    if (group_ready) SH_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream, group_ready[g], 0));
    SpanGuard span(ctx, "merkle_leaves_group");
    const unsigned grid = (unsigned)((rows + kThreads - 1) / kThreads);
    if (a.stride)
      k_leaves_group<true><<<grid, kThreads, 0, ctx->stream>>>(a);
    else
      k_leaves_group<false><<<grid, kThreads, 0, ctx->stream>>>(a);
    SH_TRY(sh_launch_check(ctx, "leaves_group"));
    state_in = a.state_out;
    in_log = a.log_n;
    slot = slot == SCR_MERKLE_STATE_A ? SCR_MERKLE_STATE_B : SCR_MERKLE_STATE_A;
  }
  if (L > n_last) {
    void* tmp;
    SH_TRY(sh_scratch_get(ctx, SCR_MERKLE_TMP, 32ull << n_last, &tmp));
    uint64_t out_n = 1ull << L;
    SpanGuard span(ctx, "merkle_lift_hashes");
    k_lift_hashes<<<(unsigned)((out_n + 255) / 256), 256, 0, ctx->stream>>>((const uint32_t*)tmp, L - n_last,
                                                                            d_out, L);
    SH_TRY(sh_launch_check(ctx, "lift_hashes"));
  }
  return SH_OK;
}

// Layers below the leaves: layer k-1 from layer k for k = L..1; layers[k] has 2^k hashes.
int sh_build_layers_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers) {
  return sh_build_layers_step_impl(ctx, L, layers, nullptr);
}

int sh_build_layers_step_impl(sh_ctx* ctx, uint32_t L, uint32_t* const* layers, const sh_root_step* st) {
  const sh_root_step none = {nullptr, nullptr, nullptr, 0};
  uint32_t k = L;
  while (k > 0) {
    const uint64_t n0 = 1ull << (k - 1);  // nodes of the first level of this launch
    LayerOuts outs;
    SpanGuard span(ctx, "merkle_next_layers");
    if (n0 > kQuadMaxNodes) {
      // wide levels: one lane per node, three levels per launch (level l keeps 256 >> l lanes of a
      // workgroup busy), down to kQuadMaxNodes nodes
      const uint32_t levels = std::min<uint32_t>(3, (k - 1) - kQuadMaxLog);
      for (uint32_t l = 0; l < levels; l++) outs.p[l] = layers[k - 1 - l];
      k_next_layers<<<(unsigned)((n0 + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(layers[k], k,
                                                                                              levels, outs);
      SH_TRY(sh_launch_check(ctx, "next_layers"));
      k -= levels;
    } else {
      // the top: four lanes per node, 64 nodes per workgroup reduced through LDS, up to 7 levels
      const uint32_t levels = std::min<uint32_t>(7, k);
      for (uint32_t l = 0; l < levels; l++) outs.p[l] = layers[k - 1 - l];
      const unsigned grid = (unsigned)((n0 + kQuadGroupNodes - 1) / kQuadGroupNodes);
      // the launch that computes the root (levels == k) is a single workgroup: it runs the step
      k_next_layers_quad<<<grid, 256, 0, ctx->stream>>>(layers[k], k, levels, outs,
                                                         st && levels == k ? *st : none);
      SH_TRY(sh_launch_check(ctx, "next_layers_quad"));
      k -= levels;
    }
  }
  return SH_OK;
}

extern "C" int sh_merkle_build_leaves(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes,
                                      uint64_t n, uint32_t lifting_log_size, uint8_t* d_out) {
  if (!ctx || !d_out || (n && (!d_cols || !log_sizes))) return SH_ERR_INVALID;
  return sh_build_leaves_impl(ctx, d_cols, log_sizes, n, lifting_log_size, (uint32_t*)d_out, nullptr);
}

extern "C" int sh_merkle_build_next_layer(sh_ctx* ctx, const uint8_t* d_prev, uint32_t prev_log_size,
                                          uint8_t* d_out) {
  if (!ctx || !d_prev || !d_out) return SH_ERR_INVALID;
  SH_REQUIRE(ctx, prev_log_size >= 1 && prev_log_size <= 31, "bad layer log size %u", prev_log_size);
  SpanGuard span(ctx, "merkle_next_layers");
  LayerOuts outs;
  outs.p[0] = (uint32_t*)d_out;
  uint64_t n0 = 1ull << (prev_log_size - 1);
  k_next_layers<<<(unsigned)((n0 + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(
      (const uint32_t*)d_prev, prev_log_size, 1, outs);
  return sh_launch_check(ctx, "next_layer");
}

extern "C" int sh_merkle_commit_lifted(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes,
                                       uint64_t n, uint32_t lifting_log_size, uint8_t* const* d_layers,
                                       uint8_t root[32]) {
  if (!ctx || !d_layers || (n && (!d_cols || !log_sizes))) return SH_ERR_INVALID;
  if (n == 0) SH_REQUIRE(ctx, lifting_log_size == 0, "lifting_log_size must be 0 when no columns are committed");
  // stable sort ascending by size (crates/stwo/src/prover/vcs_lifted/prover.rs:68)
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t x, uint64_t y) { return log_sizes[x] < log_sizes[y]; });
  std::vector<const uint32_t*> cols(n);
  std::vector<uint32_t> logs(n);
  for (uint64_t i = 0; i < n; i++) {
    cols[i] = d_cols[order[i]];
    logs[i] = log_sizes[order[i]];
  }
  SH_TRY(sh_build_leaves_impl(ctx, cols.data(), logs.data(), n, lifting_log_size,
                              (uint32_t*)d_layers[lifting_log_size], nullptr));
  SH_TRY(sh_build_layers_impl(ctx, lifting_log_size, (uint32_t* const*)d_layers));
  if (root) {
    SH_CHECK_HIP(ctx, hipMemcpyAsync(root, d_layers[0], 32, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return SH_OK;
}

int sh_commit_on_layer_impl(sh_ctx* ctx, uint32_t log_size, const uint32_t* d_prev,
                            const uint32_t* const* h_cols, uint64_t n, uint32_t* d_out) {
  SH_REQUIRE(ctx, log_size <= 30, "layer log size %u too large", log_size);
  MixedArgs a = {};
  a.prev = d_prev;
  a.cnt = (uint32_t)n;
  a.log_size = log_size;
  a.extra = 0;
  a.out[0] = d_out;
  if (n) {
    void* d_tab;
    SH_TRY(sh_stage(ctx, h_cols, n * sizeof(void*), &d_tab));
    a.cols = (const uint32_t* const*)d_tab;
  }
  const uint64_t rows = 1ull << log_size;
  SpanGuard span(ctx, "merkle_commit_on_layer");
  k_mixed_levels<<<(unsigned)((rows + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(a);
  return sh_launch_check(ctx, "commit_on_layer");
}

extern "C" int sh_merkle_commit_on_layer(sh_ctx* ctx, uint32_t log_size, const uint8_t* d_prev,
                                         const uint32_t* const* d_cols, uint64_t n, uint8_t* d_out) {
  if (!ctx || !d_out || (n && !d_cols)) return SH_ERR_INVALID;
  return sh_commit_on_layer_impl(ctx, log_size, (const uint32_t*)d_prev, d_cols, n, (uint32_t*)d_out);
}

// The mixed-degree tree (MerkleProver::commit, crates/stwo/src/prover/vcs/prover.rs:39-59) over
// columns already sorted descending by size: one k_mixed_levels launch per level that carries
// columns (plus the bottom level), each also reducing up to kMixedExtra column-free levels above
// it, and the column-free top of the tree through sh_build_layers_impl.  level_ready, when given,
// holds one event per distinct column size (descending): a level waits for its columns' event.
int sh_merkle_mixed_impl(sh_ctx* ctx, const uint32_t* const* h_cols, const uint32_t* logs, uint64_t n,
                         uint8_t* const* d_layers, const hipEvent_t* level_ready) {
  for (uint64_t i = 0; i < n; i++) {
    SH_REQUIRE(ctx, logs[i] <= 30, "column log size %u too large", logs[i]);
    if (i) SH_REQUIRE(ctx, logs[i] <= logs[i - 1], "columns must be sorted descending by size");
  }
  const uint32_t max_log = n ? logs[0] : 0;
  uint64_t pos = 0;
  size_t cls = 0;
  int lv = (int)max_log;
  const uint32_t* prev = nullptr;
  while (lv >= 0) {
    const uint64_t first = pos;
    while (pos < n && logs[pos] == (uint32_t)lv) pos++;
    const uint64_t cnt = pos - first;
    if (cnt == 0 && prev && pos == n) {
      // every remaining level is column-free: the plain inner layers
      return sh_build_layers_impl(ctx, (uint32_t)lv + 1, (uint32_t* const*)d_layers);
    }
    // column-free levels directly above this one, up to the next level with columns
    uint32_t free_above = 0;
    const uint32_t next_log = pos < n ? logs[pos] : 0;
    const uint32_t limit = pos < n ? (uint32_t)lv - next_log - 1 : 0;  // all free when no columns remain
    free_above = std::min<uint32_t>(std::min<uint32_t>(limit, kMixedExtra), (uint32_t)lv);
    MixedArgs a = {};
    a.prev = prev;
    a.cnt = (uint32_t)cnt;
    a.log_size = (uint32_t)lv;
    a.extra = free_above;
    if (cnt) {
      void* d_tab;
      SH_TRY(sh_stage(ctx, h_cols + first, cnt * sizeof(void*), &d_tab));
      a.cols = (const uint32_t* const*)d_tab;
    }
    for (uint32_t l = 0; l <= free_above; l++) a.out[l] = (uint32_t*)d_layers[lv - (int)l];
    if (cnt && level_ready) SH_CHECK_HIP(ctx, hipStreamWaitEvent(ctx->stream, level_ready[cls], 0));
    if (cnt) cls++;
    {
      SpanGuard span(ctx, "merkle_mixed_levels");
      const uint64_t nodes = 1ull << lv;
      k_mixed_levels<<<(unsigned)((nodes + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(a);
      SH_TRY(sh_launch_check(ctx, "mixed_levels"));
    }
    prev = (const uint32_t*)d_layers[lv - (int)free_above];
    lv -= (int)free_above + 1;
  }
  return SH_OK;
}

extern "C" int sh_merkle_commit_mixed(sh_ctx* ctx, const uint32_t* const* d_cols, const uint32_t* log_sizes,
                                      uint64_t n, uint8_t* const* d_layers, uint8_t root[32]) {
  if (!ctx || !d_layers || (n && (!d_cols || !log_sizes))) return SH_ERR_INVALID;
  // stable sort descending by size (crates/stwo/src/prover/vcs/prover.rs:45)
  std::vector<uint64_t> order(n);
  for (uint64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t x, uint64_t y) { return log_sizes[x] > log_sizes[y]; });
  std::vector<const uint32_t*> cols(n);
  std::vector<uint32_t> logs(n);
  for (uint64_t i = 0; i < n; i++) {
    cols[i] = d_cols[order[i]];
    logs[i] = log_sizes[order[i]];
  }
  SH_TRY(sh_merkle_mixed_impl(ctx, cols.data(), logs.data(), n, d_layers, nullptr));
  if (root) {
    SH_CHECK_HIP(ctx, hipMemcpyAsync(root, d_layers[0], 32, hipMemcpyDeviceToHost, ctx->stream));
    SH_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return SH_OK;
}
