"""ctypes wrapper around oracle/_build/liboracle.so (test infrastructure only).

Each function mirrors a C entry point of oracle/oracle.h; byte buffers in and
out use the reference's public formats (include/vkzg.h:32-38).
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

BYTES_PER_BLOB = 131072
BYTES_PER_CELL = 2048
CELLS_PER_BLOB = 128
BYTES_PER_PROOF = 48
PROOF_BYTES_PER_BLOB = CELLS_PER_BLOB * BYTES_PER_PROOF
CELL_BYTES_PER_BLOB = CELLS_PER_BLOB * BYTES_PER_CELL


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes.c_char_p
        u8p = ctypes.c_char_p
        L.or_setup_new.restype = ctypes.c_void_p
        L.or_setup_new.argtypes = [u8p]
        L.or_setup_free.argtypes = [ctypes.c_void_p]
        L.or_compute_proofs.argtypes = [ctypes.c_void_p, c, u8p, ctypes.c_size_t]
        L.or_compute_proofs_direct.argtypes = [ctypes.c_void_p, c, u8p]
        L.or_setup_xext.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, c]
        L.or_recover_cells.argtypes = [c, u8p, u8p, ctypes.c_size_t]
        L.or_blob_to_cells.argtypes = [c, u8p]
        L.or_blob_to_poly.argtypes = [c, u8p]
        L.or_fr_ntt.argtypes = [c, ctypes.c_uint32, ctypes.c_int]
        L.or_setup_g1_from_tau.argtypes = [c, u8p]
        L.or_proofs_from_tau.argtypes = [c, u8p, u8p]
        L.or_g1_roundtrip.argtypes = [c, u8p]
        L.or_g1_mul_gen.argtypes = [c, u8p]
        L.or_g1_mul.argtypes = [c, u8p, u8p]
        L.or_g1_add.argtypes = [c, u8p, u8p]
        L.or_fr_mul_be.argtypes = [c, u8p, u8p]
        L.or_fp_mul_be.argtypes = [c, u8p, u8p]
        L.or_fr_inv_be.argtypes = [c, u8p]
        L.or_root_8192_be.argtypes = [c, ctypes.c_uint32]
        L.or_g2_generator.argtypes = [c]
        L.or_g2_raw_point.argtypes = [c, ctypes.c_uint64]
        L.or_g2_on_curve.argtypes = [u8p]
        L.or_g2_mul.argtypes = [c, u8p, u8p, ctypes.c_uint32]
        L.or_g2_add.argtypes = [c, u8p, u8p]
        L.or_g2_powers_from_tau.argtypes = [c, u8p, ctypes.c_uint32]
        L.or_g2_cofactor.argtypes = [c]
        L.or_final_exp_hard.argtypes = [c]
        L.or_pairing.argtypes = [c, u8p, u8p]
        L.or_pairing_check.argtypes = [u8p, u8p, ctypes.c_size_t]
        L.or_vsetup_new.restype = ctypes.c_void_p
        L.or_vsetup_new.argtypes = [u8p, u8p]
        L.or_vsetup_free.argtypes = [ctypes.c_void_p]
        L.or_commit.argtypes = [ctypes.c_void_p, c, u8p]
        L.or_commit_batch.argtypes = [ctypes.c_void_p, c, u8p, ctypes.c_size_t]
        L.or_verify_cell_proofs.argtypes = [ctypes.c_void_p, u8p, ctypes.POINTER(ctypes.c_uint32), u8p, u8p,
                                            ctypes.c_size_t, ctypes.c_uint64]
        L.or_verify_cell_proofs_r.argtypes = [ctypes.c_void_p, u8p, ctypes.POINTER(ctypes.c_uint32), u8p, u8p,
                                              ctypes.c_size_t, u8p, c, c]
        L.or_g2_compress.argtypes = [c, u8p]
        L.or_g2_decompress.argtypes = [c, u8p]
        L.or_g1_msm.argtypes = [c, u8p, u8p, ctypes.c_size_t]
        _lib = L
    return _lib


R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


class Setup:
    """FK20 setup built from 4096 compressed G1 monomial points."""

    def __init__(self, g1_monomial: bytes):
        assert len(g1_monomial) == 4096 * 48
        self._h = lib().or_setup_new(g1_monomial)
        if not self._h:
            raise ValueError("invalid setup point")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_setup_free(self._h)
            self._h = None

    def compute_proofs(self, blobs: bytes, n: int):
        """Returns (rc, proofs); rc 0 = ok, 6 = invalid blob."""
        out = ctypes.create_string_buffer(PROOF_BYTES_PER_BLOB * n)
        rc = lib().or_compute_proofs(self._h, out, blobs, n)
        return rc, out.raw

    def compute_proofs_direct(self, blob: bytes):
        out = ctypes.create_string_buffer(PROOF_BYTES_PER_BLOB)
        rc = lib().or_compute_proofs_direct(self._h, out, blob)
        return rc, out.raw

    def xext(self, j: int, i: int) -> bytes:
        out = ctypes.create_string_buffer(48)
        lib().or_setup_xext(self._h, j, i, out)
        return out.raw


def recover_cells(cells: bytes, present: bytes, n: int):
    out = ctypes.create_string_buffer(CELL_BYTES_PER_BLOB * n)
    rc = lib().or_recover_cells(out, cells, present, n)
    return rc, out.raw


def blob_to_cells(blob: bytes):
    out = ctypes.create_string_buffer(CELL_BYTES_PER_BLOB)
    rc = lib().or_blob_to_cells(out, blob)
    return rc, out.raw


def blob_to_poly(blob: bytes):
    out = ctypes.create_string_buffer(4096 * 32)
    rc = lib().or_blob_to_poly(out, blob)
    return rc, out.raw


def fr_ntt(data: bytes, n: int, inverse: bool) -> bytes:
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    lib().or_fr_ntt(buf, n, 1 if inverse else 0)
    return buf.raw[: len(data)]


def setup_g1_from_tau(tau: int) -> bytes:
    out = ctypes.create_string_buffer(4096 * 48)
    lib().or_setup_g1_from_tau(out, tau.to_bytes(32, "big"))
    return out.raw


def proofs_from_tau(blob: bytes, tau: int):
    out = ctypes.create_string_buffer(PROOF_BYTES_PER_BLOB)
    rc = lib().or_proofs_from_tau(out, blob, tau.to_bytes(32, "big"))
    return rc, out.raw


def g1_roundtrip(p: bytes):
    out = ctypes.create_string_buffer(48)
    rc = lib().or_g1_roundtrip(out, p)
    return rc, out.raw


def g1_mul_gen(k: int) -> bytes:
    out = ctypes.create_string_buffer(48)
    lib().or_g1_mul_gen(out, (k % (1 << 256)).to_bytes(32, "big"))
    return out.raw


def g1_mul(p: bytes, k: int) -> bytes:
    out = ctypes.create_string_buffer(48)
    if lib().or_g1_mul(out, p, (k % (1 << 256)).to_bytes(32, "big")):
        raise ValueError("bad point")
    return out.raw


def g1_add(p: bytes, q: bytes) -> bytes:
    out = ctypes.create_string_buffer(48)
    if lib().or_g1_add(out, p, q):
        raise ValueError("bad point")
    return out.raw


def fr_mul(a: int, b: int) -> int:
    out = ctypes.create_string_buffer(32)
    lib().or_fr_mul_be(out, a.to_bytes(32, "big"), b.to_bytes(32, "big"))
    return int.from_bytes(out.raw, "big")


def fp_mul(a: int, b: int) -> int:
    out = ctypes.create_string_buffer(48)
    lib().or_fp_mul_be(out, a.to_bytes(48, "big"), b.to_bytes(48, "big"))
    return int.from_bytes(out.raw, "big")


def fr_inv(a: int) -> int:
    out = ctypes.create_string_buffer(32)
    lib().or_fr_inv_be(out, a.to_bytes(32, "big"))
    return int.from_bytes(out.raw, "big")


def root_8192(k: int) -> int:
    out = ctypes.create_string_buffer(32)
    lib().or_root_8192_be(out, k)
    return int.from_bytes(out.raw, "big")


# ---- pairing and batch cell-proof verification ----

def g2_generator() -> bytes:
    out = ctypes.create_string_buffer(192)
    lib().or_g2_generator(out)
    return out.raw


def g2_raw_point(seed: int) -> bytes:
    out = ctypes.create_string_buffer(192)
    lib().or_g2_raw_point(out, seed)
    return out.raw


def g2_on_curve(p: bytes) -> bool:
    return bool(lib().or_g2_on_curve(p))


def g2_mul(p: bytes, k: int) -> bytes:
    kb = k.to_bytes(max(1, (k.bit_length() + 7) // 8), "big")
    out = ctypes.create_string_buffer(192)
    if lib().or_g2_mul(out, p, kb, len(kb)):
        raise ValueError("bad G2 point")
    return out.raw


def g2_add(p: bytes, q: bytes) -> bytes:
    out = ctypes.create_string_buffer(192)
    if lib().or_g2_add(out, p, q):
        raise ValueError("bad G2 point")
    return out.raw


def g2_powers_from_tau(tau: int, n: int = 65) -> bytes:
    out = ctypes.create_string_buffer(192 * n)
    lib().or_g2_powers_from_tau(out, (tau % R).to_bytes(32, "big"), n)
    return out.raw


def g2_cofactor() -> int:
    out = ctypes.create_string_buffer(64)
    lib().or_g2_cofactor(out)
    return int.from_bytes(out.raw, "big")


def final_exp_hard() -> int:
    out = ctypes.create_string_buffer(256)
    lib().or_final_exp_hard(out)
    return int.from_bytes(out.raw, "big")


def pairing(p: bytes, q: bytes) -> bytes:
    out = ctypes.create_string_buffer(576)
    if lib().or_pairing(out, p, q):
        raise ValueError("bad point")
    return out.raw


def pairing_check(pairs) -> bool:
    """True when the product of e(P, Q) over (P, Q) in pairs is 1."""
    g1 = b"".join(p for p, _ in pairs)
    g2 = b"".join(q for _, q in pairs)
    rc = lib().or_pairing_check(g1, g2, len(pairs))
    if rc < 0:
        raise ValueError("bad point")
    return rc == 1


class VerifySetup:
    """Monomial G1 setup plus G2 powers for commitments and cell-proof checks."""

    def __init__(self, g1_monomial: bytes, g2_powers: bytes):
        assert len(g1_monomial) == 4096 * 48 and len(g2_powers) >= 65 * 192
        self._h = lib().or_vsetup_new(g1_monomial, g2_powers)
        if not self._h:
            raise ValueError("invalid setup point")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_vsetup_free(self._h)
            self._h = None

    def commit(self, blob: bytes) -> bytes:
        out = ctypes.create_string_buffer(48)
        rc = lib().or_commit(self._h, out, blob)
        if rc:
            raise ValueError("non-canonical blob")
        return out.raw

    def commit_batch(self, blobs: bytes, n: int) -> bytes:
        out = ctypes.create_string_buffer(48 * max(1, n))
        if lib().or_commit_batch(self._h, out, blobs, n):
            raise ValueError("non-canonical blob")
        return out.raw[:48 * n]

    def verify(self, commitments: bytes, cell_indices, cells: bytes, proofs: bytes, seed: int = 1) -> bool:
        """Batch check of n items (commitment, cell index, cell, proof).
        Returns True/False; raises ValueError on malformed input."""
        n = len(cell_indices)
        assert len(commitments) == 48 * n and len(cells) == 2048 * n and len(proofs) == 48 * n
        idx = (ctypes.c_uint32 * max(1, n))(*cell_indices)
        rc = lib().or_verify_cell_proofs(self._h, commitments, idx, cells, proofs, n, seed)
        if rc < 0:
            raise ValueError("malformed verification input")
        return rc == 1

    def verify_r(self, commitments: bytes, cell_indices, cells: bytes, proofs: bytes, r: bytes):
        """The specs' check with weights r^k for a given challenge r (32 bytes big-endian):
        (valid, A, RL) with A = sum r^k pi_k and RL the other G1 input, compressed."""
        n = len(cell_indices)
        assert len(commitments) == 48 * n and len(cells) == 2048 * n and len(proofs) == 48 * n
        idx = (ctypes.c_uint32 * max(1, n))(*cell_indices)
        a = ctypes.create_string_buffer(48)
        b = ctypes.create_string_buffer(48)
        rc = lib().or_verify_cell_proofs_r(self._h, commitments, idx, cells, proofs, n, r, a, b)
        if rc < 0:
            raise ValueError("malformed verification input")
        return rc == 1, a.raw, b.raw

    def verify_blobs(self, blobs: bytes, proofs: bytes, n: int, cells: bytes = None, seed: int = 1) -> bool:
        """Every proof of n blobs against the blobs' commitments and cells."""
        one = self.commit_batch(blobs, n)
        coms = b"".join(one[48 * b:48 * (b + 1)] * CELLS_PER_BLOB for b in range(n))
        if cells is None:
            cells = b"".join(blob_to_cells(blobs[b * BYTES_PER_BLOB:(b + 1) * BYTES_PER_BLOB])[1]
                             for b in range(n))
        return self.verify(coms, list(range(CELLS_PER_BLOB)) * n, cells, proofs, seed)


def g2_compress(q192: bytes) -> bytes:
    """ZCash 96-byte compression of a point in this oracle's 192-byte G2 encoding."""
    out = ctypes.create_string_buffer(96)
    if lib().or_g2_compress(out, q192):
        raise ValueError("bad point")
    return out.raw


def g2_decompress(q96: bytes) -> bytes:
    out = ctypes.create_string_buffer(192)
    if lib().or_g2_decompress(out, q96):
        raise ValueError("bad encoding")
    return out.raw


def g1_msm(points: bytes, scalars: bytes) -> bytes:
    n = len(points) // 48
    out = ctypes.create_string_buffer(48)
    if lib().or_g1_msm(out, points, scalars, n):
        raise ValueError("bad point or scalar")
    return out.raw


def verify_challenge(commitments: bytes, cell_indices, cells: bytes, proofs: bytes) -> bytes:
    """compute_verify_cell_kzg_proof_batch_challenge of the consensus specs
    (polynomial-commitments-sampling), with the deduplication of
    verify_cell_kzg_proof_batch: 32 bytes big-endian, below r."""
    import hashlib
    n = len(cell_indices)
    coms = [commitments[48 * k:48 * (k + 1)] for k in range(n)]
    uniq = []
    for c in coms:
        if c not in uniq:
            uniq.append(c)
    h = b"RCKZGCBATCH__V1_"
    h += (4096).to_bytes(8, "big") + (64).to_bytes(8, "big")
    h += len(uniq).to_bytes(8, "big") + n.to_bytes(8, "big")
    h += b"".join(uniq)
    for k in range(n):
        h += uniq.index(coms[k]).to_bytes(8, "big") + int(cell_indices[k]).to_bytes(8, "big")
        h += cells[2048 * k:2048 * (k + 1)] + proofs[48 * k:48 * (k + 1)]
    r = int.from_bytes(hashlib.sha256(h).digest(), "big") % R
    return r.to_bytes(32, "big")
