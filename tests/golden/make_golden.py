"""Writes tests/golden/golden.json: known-answer vectors for the oracle and the HIP backend.

Two kinds of data, both plain inputs/expected outputs:
  * "reference_kats": vectors that the reference's own tests assert (copied as data, with the
    reference test that holds them);
  * "hashlib": vectors computed here from first principles with Python's hashlib.blake2s (an
    independent RFC 7693 implementation) and integer arithmetic, following the definitions the
    reference documents (leaf/node contents, channel transcript rules, proof-of-work predicate).
Nothing from the reference is imported or executed.
Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import struct

P = (1 << 31) - 1
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")


def b2(data: bytes) -> str:
    return hashlib.blake2s(data).hexdigest()


def le32s(words):
    return b"".join(struct.pack("<I", w & 0xFFFFFFFF) for w in words)


def lift_index(i, n, L):
    return ((i >> (L - n + 1)) << 1) + (i & 1)


def lifted_leaves(cols, L):
    """Leaf i = Blake2s(LE32 of every column's lifted value), columns sorted ascending
    (crates/stwo/src/prover/backend/cpu/merkle_lifted.rs:14-45 docs)."""
    cols = sorted(cols, key=len)
    out = []
    for i in range(1 << L):
        words = [c[lift_index(i, len(c).bit_length() - 1, L)] for c in cols]
        out.append(hashlib.blake2s(le32s(words)).digest())
    return out


def tree_root(leaves):
    layer = leaves
    while len(layer) > 1:
        layer = [hashlib.blake2s(layer[2 * i] + layer[2 * i + 1]).digest() for i in range(len(layer) // 2)]
    return layer[0]


def mixed_root(cols):
    """MerkleProver::commit with hash_node (crates/stwo/src/core/vcs/blake2_merkle.rs:13-29)."""
    cols = sorted(cols, key=lambda c: -len(c))
    max_log = len(cols[0]).bit_length() - 1
    prev = None
    for ls in range(max_log, -1, -1):
        layer_cols = [c for c in cols if len(c) == 1 << ls]
        cur = []
        for i in range(1 << ls):
            data = b""
            if prev is not None:
                data += prev[2 * i] + prev[2 * i + 1]
            data += le32s([c[i] for c in layer_cols])
            cur.append(hashlib.blake2s(data).digest())
        prev = cur
    return prev[0]


def lcg_cols(seed, logs):
    x = seed
    cols = []
    for lg in logs:
        c = []
        for _ in range(1 << lg):
            x = (x * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
            c.append((x >> 33) % P)
        cols.append(c)
    return cols


class Channel:
    """Blake2sChannel transcript rules (crates/stwo/src/core/channel/blake2s.rs:16-100)."""

    def __init__(self):
        self.digest = bytes(32)
        self.n_draws = 0

    def mix_u32s(self, ws):
        self.digest = hashlib.blake2s(self.digest + le32s(ws)).digest()
        self.n_draws = 0

    def draw_u32s(self):
        h = hashlib.blake2s(self.digest + struct.pack("<I", self.n_draws) + b"\x00").digest()
        self.n_draws += 1
        return list(struct.unpack("<8I", h))

    def draw_secure_felt(self):
        f = [((u >> 31) + (u & P)) % P for u in self.draw_u32s()]  # M31::reduce of a u32
        return [(f[2 * i] + 2 * f[2 * i + 1]) % P for i in range(4)]


def pow_ok(digest, bits, nonce):
    pre = hashlib.blake2s(struct.pack("<I", 0x12345678) + digest + struct.pack("<I", bits)).digest()
    res = hashlib.blake2s(pre + struct.pack("<Q", nonce)).digest()
    v = int.from_bytes(res[:16], "little")
    tz = 128 if v == 0 else (v & -v).bit_length() - 1
    return tz >= bits


def main():
    g = {}
    g["reference_kats"] = {
        # crates/stwo/src/core/vcs/blake2_hash.rs:125-131
        "blake2s_a": {"input_hex": b"a".hex(),
                      "hash": "4a0d129873403037c2cd9b9048203687f6233fb6738956e0349bd4320fec3e90"},
        # crates/stwo/src/core/vcs_lifted/hasher_test.rs:10-19, 54-63
        "hash_u32s": {"words": [0, 1, 2, 3, 0x12345678, 0xFFFFFFFF, 7, 8, 9],
                      "hash": "efee1538e0216d3a09ce742ce768c44a7db004d801c40f66040994353f55fe85",
                      "followed_by_digest":
                          "18a60268d189dccf5065ddd361d280474fbc2b08d75a7e4e1c20e57a67f2f7fa"},
        # crates/stwo/src/core/channel/blake2s.rs:186-205
        "mix_u64": {"value": 0x1111222233334444,
                    "digest": bytes([0xbc, 0x9e, 0x3f, 0xc1, 0xd2, 0x4e, 0x88, 0x97, 0x95, 0x6d, 0x33,
                                     0x59, 0x32, 0x73, 0x97, 0x24, 0x9d, 0x6b, 0xca, 0xcd, 0x22, 0x4d,
                                     0x92, 0x74, 0x04, 0xe7, 0xba, 0x4a, 0x77, 0xdc, 0x6e, 0xce]).hex()},
        # crates/stwo/src/core/channel/blake2s.rs:207-219
        "mix_u32s": {"words": [1, 2, 3, 4, 5, 6, 7, 8, 9],
                     "digest": bytes([0x70, 0x91, 0x76, 0x83, 0x57, 0xbb, 0x1b, 0xb3, 0x34, 0x6f, 0xda,
                                      0xb6, 0xb3, 0x57, 0xd7, 0xfa, 0x46, 0xb8, 0xfb, 0xe3, 0x2c, 0x2e,
                                      0x43, 0x24, 0xa0, 0xff, 0xc2, 0x94, 0xcb, 0xf9, 0xa1,
                                      0xc7]).hex()},
        # crates/stwo/src/core/channel/blake2s.rs:221-240
        "verify_pow_nonce": {"digest_words": [968886948, 725376924, 836084817, 484428276, 1805658819,
                                              300032261, 172116750, 994058243],
                             "true": [[10, 649]], "false": [[11, 649], [9, 649], [10, 648], [10, 650]]},
        # crates/stwo/src/core/fields/m31.rs:38-60 doc tests
        "m31_partial_reduce": [[2 * P - 19, P - 19]],
        "m31_reduce": [[P * P - 19, P - 19]],
        # crates/stwo/src/core/fields/cm31.rs:88-108
        "cm31_mul": [[[1, 2], [4, 5], [P - 6, 13]]],
        # crates/stwo/src/core/fields/qm31.rs:151-167
        "qm31_mul": [[[1, 2, 3, 4], [4, 5, 6, 7], [P - 71, 93, P - 16, 50]]],
        # crates/stwo/src/core/fields/mod.rs:171 complex conjugate doc test
        "qm31_conj": [[[1, 2, 3, 4], [1, 2, P - 3, P - 4]]],
        # crates/stwo/src/core/circle.rs:163-185 / 188-191
        "m31_circle_gen": [2, 1268011823],
        "secure_field_circle_gen": [[1, 0, 478637715, 513582971],
                                    [992285211, 649143431, 740191619, 1186584352]],
        # crates/stwo/src/core/poly/utils.rs:137-145 (line twiddle layout of CanonicCoset(4))
        "domain_line_twiddles_log4": [[0, 1, 2, 3], [4, 5], [6]],
        # crates/stwo/src/prover/backend/cpu/circle.rs:355-369: the test asserts
        # eval == c0 + c1*y + c2*x + c3*x*y for coeffs [1, 3, 2, 4] at (x, y) = (5, 8)
        "eval_at_point_4coeffs": {"coeffs": [1, 3, 2, 4], "x": 5, "y": 8,
                                  "value": (1 + 3 * 8 + 2 * 5 + 4 * 5 * 8) % P},
        # crates/stwo/src/prover/backend/cpu/circle.rs:395-423 (test_cpu_eval_at_point_by_folding)
        "eval_by_folding_poly": [691, 805673, 5, 435684, 4832, 23876431, 197, 897346068],
        "eval_by_folding_point_indices": [348, 9736524, 13, 346752],
    }
    # crates/stwo/src/prover/vcs_lifted/prover.rs:194-236: columns 2^2, 2^3, 2^4 valued 0..n-1
    cols = [list(range(1 << i)) for i in (2, 3, 4)]
    leaves = lifted_leaves(cols, 4)
    g["reference_kats"]["lifted_leaves_small"] = {
        "logs": [2, 3, 4],
        "leaf0": hashlib.blake2s(bytes(12)).hexdigest(),
        "leaf4": hashlib.blake2s(le32s([0, 2, 4])).hexdigest(),
        "leaf15": hashlib.blake2s(le32s([3, 7, 15])).hexdigest(),
        "all_leaves": [l.hex() for l in leaves],
        "root": tree_root(leaves).hex(),
    }
    hl = {}
    hl["blake2s"] = [{"len": n, "hash": b2(bytes((i * 7 + 3) & 0xFF for i in range(n)))}
                     for n in (0, 1, 3, 4, 31, 32, 63, 64, 65, 127, 128, 129, 200)]
    # lifted and mixed trees over LCG columns (seeded), brute force by definition
    tree_cases = []
    for seed, logs, L in [(1, [1], 1), (2, [3, 3], 3), (3, [2, 5, 3, 5, 4], 6),
                          (4, [4] * 17 + [6] * 3, 6), (5, [1, 2, 3, 4, 5, 6], 6),
                          (6, [5] * 33, 5), (7, [3] * 15 + [5] * 16 + [7] * 2, 8)]:
        cs = lcg_cols(seed, logs)
        lv = lifted_leaves(cs, L)
        tree_cases.append({"seed": seed, "logs": logs, "lifting_log": L,
                           "leaves_sample": {str(i): lv[i].hex() for i in (0, 1, (1 << L) - 1)},
                           "lifted_root": tree_root(lv).hex(), "mixed_root": mixed_root(cs).hex()})
    hl["trees"] = tree_cases
    # channel transcript from the default channel
    ch = Channel()
    trace = {"draw_u32s_0": ch.draw_u32s()}
    trace["draw_secure_felt_1"] = ch.draw_secure_felt()
    ch.mix_u32s([7, 8, 9])
    trace["after_mix_digest"] = ch.digest.hex()
    trace["draw_secure_felt_after_mix"] = ch.draw_secure_felt()
    hl["channel"] = trace
    # smallest nonce for a few (digest, bits)
    grinds = []
    for k, bits in [(0, 4), (1, 8), (2, 10)]:
        d = hashlib.blake2s(bytes([k])).digest()
        nonce = 0
        while not pow_ok(d, bits, nonce):
            nonce += 1
        grinds.append({"digest": d.hex(), "bits": bits, "nonce": nonce})
    hl["grind"] = grinds
    g["hashlib"] = hl
    with open(OUT, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
