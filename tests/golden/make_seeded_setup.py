#!/usr/bin/env python3
"""Write tests/golden/seeded_setup.bin: an insecure setup with a known tau for
tests and benchmarks of verification and commitments.

Layout: 4096 x 48 bytes [tau^i]G1 (compressed, i < 4096), then 65 x 96 bytes
[tau^i]Q (compressed G2, i <= 64), Q the oracle's G2 generator.  Made by the
CPU oracle from the tau the tests use (tests/conftest.py SEEDED_TAU);
tests/test_verify_host.py checks the file against the oracle.

Usage: python3 tests/golden/make_seeded_setup.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402
from tests.conftest import SEEDED_TAU  # noqa: E402

OUT = os.path.join(HERE, "seeded_setup.bin")


def build():
    g1 = O.setup_g1_from_tau(SEEDED_TAU)
    raw = O.g2_powers_from_tau(SEEDED_TAU, 65)
    g2 = b"".join(O.g2_compress(raw[192 * i:192 * (i + 1)]) for i in range(65))
    return g1 + g2


if __name__ == "__main__":
    data = build()
    open(OUT, "wb").write(data)
    print("wrote", OUT, len(data))
