"""Writes tests/golden/cairo_proof_all_opcodes.npz from the reference's own test fixture
test_data/test_prove_verify_all_opcode_components/proof.json.

That file is the CairoSerde felt list of a Cairo proof made by the reference prover
(crates/prover/tests/slow_tests.rs:420-463, test_prove_all_opcode_components_proof_regression:
Blake2s channel, FriParams::new(26, 0, 1, 70, 3), channel salt 0).  Every felt of it is below 2^32,
so the fixture keeps it as one uint32 array (plus the SHA-256 of the source text), which is data
only: the file is read as JSON and nothing from the reference is imported or executed.
Run: python tests/golden/make_cairo_fixture.py
"""
import hashlib
import json
import os

import numpy as np

SRC = "/root/reference/test_data/test_prove_verify_all_opcode_components/proof.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cairo_proof_all_opcodes.npz")


def main():
    with open(SRC, "rb") as f:
        raw = f.read()
    felts = [int(x, 16) for x in json.loads(raw)]
    assert all(0 <= v < 1 << 32 for v in felts), "a felt does not fit 32 bits"
    np.savez_compressed(OUT, felts=np.array(felts, np.uint32),
                        source_sha256=np.frombuffer(hashlib.sha256(raw).digest(), np.uint8))
    print(f"wrote {OUT}: {len(felts)} felts")


if __name__ == "__main__":
    main()
