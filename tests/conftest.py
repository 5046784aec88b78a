"""Shared fixtures: consensus-vector fixtures, the CPU oracle, GPU provers."""
import json
import os
import random
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")
MAINNET_SETUP = os.path.join(REPO, "vkzg_amd", "data", "g1_monomial_mainnet.bin")
SEEDED_TAU = 0x2A5C1D0E7F3B9A4C6D8E0F1A2B3C4D5E6F708192A3B4C5D6E7F8091A2B3C4D5  # < r


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvkzg.so)")


class Vectors:
    """The consensus-spec vectors of the reference's tests/vectors, as fixtures."""

    def __init__(self):
        self.index = json.load(open(os.path.join(GOLDEN, "vectors.json")))
        self.data = open(os.path.join(GOLDEN, "compute_vectors.bin"), "rb").read()
        self.compute = {r["name"]: r for r in self.index["compute"]}

    def blob(self, name):
        r = self.compute[name]
        return self.data[r["offset"]:r["offset"] + r["blob_len"]]

    def proofs(self, name):
        r = self.compute[name]
        return self.data[r["proofs_offset"]:r["proofs_offset"] + 6144]

    def cells_sha256(self, name):
        return self.compute[name]["cells_sha256"]

    def valid_compute(self):
        return [r["name"] for r in self.index["compute"] if r["valid"]]

    def invalid_compute_131072(self):
        return [r["name"] for r in self.index["compute"]
                if not r["valid"] and r.get("offset") is not None]

    def recover_cases(self):
        return self.index["recover"]


@pytest.fixture(scope="session")
def vectors():
    return Vectors()


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def mainnet_setup_bytes():
    return open(MAINNET_SETUP, "rb").read()


@pytest.fixture(scope="session")
def oracle_mainnet(oracle_mod, mainnet_setup_bytes):
    return oracle_mod.Setup(mainnet_setup_bytes)


@pytest.fixture(scope="session")
def seeded_setup_bytes(oracle_mod):
    return oracle_mod.setup_g1_from_tau(SEEDED_TAU)


@pytest.fixture(scope="session")
def seeded_verify_setup(oracle_mod, seeded_setup_bytes):
    """Commitment and batch cell-proof verification for the seeded-tau setup
    (G1 powers plus 65 G2 powers [tau^i]Q)."""
    return oracle_mod.VerifySetup(seeded_setup_bytes, oracle_mod.g2_powers_from_tau(SEEDED_TAU, 65))


@pytest.fixture(scope="session")
def cells_of(vectors, oracle_mod):
    """Oracle cells of a compute vector (their sha256 is checked by test_oracle)."""
    cache = {}

    def get(name):
        if name not in cache:
            rc, cells = oracle_mod.blob_to_cells(vectors.blob(name))
            assert rc == 0
            cache[name] = cells
        return cache[name]

    return get


def random_blob(rng, R=0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001):
    return b"".join(rng.randrange(R).to_bytes(32, "big") for _ in range(4096))


@pytest.fixture(scope="session")
def rng():
    return random.Random(0x5EED)


# ---------------------------------------------------------------- GPU side

@pytest.fixture(scope="session")
def gpu_prover():
    import vkzg_amd
    p = vkzg_amd.Prover()
    yield p
    p.close()


@pytest.fixture(scope="session")
def gpu_prover_small_batch():
    """max_batch_size = 4 (chunking) with a 10-bit phase-A window, so a
    second table geometry is covered too."""
    import vkzg_amd
    old = os.environ.get("VKZG_AMD_WINDOW")
    os.environ["VKZG_AMD_WINDOW"] = "10"
    try:
        p = vkzg_amd.Prover(max_batch_size=4)
    finally:
        if old is None:
            del os.environ["VKZG_AMD_WINDOW"]
        else:
            os.environ["VKZG_AMD_WINDOW"] = old
    assert p.info()["window_bits"] == 10
    yield p
    p.close()


@pytest.fixture(scope="session")
def gpu_seeded_prover(seeded_setup_bytes):
    import vkzg_amd
    p = vkzg_amd.Prover(setup_g1_monomial=seeded_setup_bytes)
    yield p
    p.close()
