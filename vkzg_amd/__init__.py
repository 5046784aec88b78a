"""vkzg for AMD Instinct MI355X: EIP-7594 FK20 cell proofs and cell recovery.

The computation runs in libvkzg.so (hand-written gfx950 HIP kernels behind the
reference's C ABI, include/vkzg.h).  This module is the Python host mirror of
the reference's Java binding (bindings/java/src/main/java/vkzg/Vkzg.java):

  * module-level ``init`` / ``deinit`` / ``compute_cell_kzg_proofs_batch`` /
    ``recover_cells_batch`` / ``device_name`` / ``gpu_core_count`` with the
    same argument checks and errors as Vkzg.java:164-305, and
  * a ``Prover`` class for callers that want several provers (custom setups,
    different batch sizes).
"""
import ctypes
import threading

from ._lib import (LIB_PATH, RequestInfo, ServerOptions, ServerStats, Tuning, VkzgAmdInfo, VkzgOptions,
                   load)

FIELD_ELEMENTS_PER_BLOB = 4096
BYTES_PER_FIELD_ELEMENT = 32
BYTES_PER_BLOB = FIELD_ELEMENTS_PER_BLOB * BYTES_PER_FIELD_ELEMENT
CELL_PROOFS_PER_BLOB = 128
BYTES_PER_PROOF = 48
FIELD_ELEMENTS_PER_CELL = 64
BYTES_PER_CELL = FIELD_ELEMENTS_PER_CELL * BYTES_PER_FIELD_ELEMENT

OK, ERR_BADARGS, ERR_MALLOC, ERR_IO, ERR_SETUP, ERR_GPU, ERR_INVALID_BLOB = range(7)


class VkzgError(Exception):
    """A nonzero vkzg_result (mirrors VkzgException: code + vkzg_error_string)."""

    def __init__(self, code):
        self.code = code
        super().__init__("vkzg error %d: %s" % (code, error_string(code)))


def error_string(code):
    return load().vkzg_error_string(code).decode()


def _check(rc):
    if rc != OK:
        raise VkzgError(rc)


class Prover:
    """One vkzg_prover on the current HIP device."""

    def __init__(self, max_batch_size=0, setup_g1_monomial=None):
        L = load()
        opts = VkzgOptions()
        L.vkzg_options_default(ctypes.byref(opts))
        if max_batch_size:
            opts.max_batch_size = max_batch_size
        h = ctypes.c_void_p()
        if setup_g1_monomial is None:
            rc = L.vkzg_prover_new(ctypes.byref(h), ctypes.byref(opts))
        else:
            rc = L.vkzg_amd_prover_new_with_setup(ctypes.byref(h), ctypes.byref(opts),
                                                  bytes(setup_g1_monomial),
                                                  len(setup_g1_monomial))
        _check(rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().vkzg_prover_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def device_name(self):
        return load().vkzg_prover_device_name(self._h).decode()

    def gpu_core_count(self):
        return int(load().vkzg_prover_gpu_core_count(self._h))

    def info(self):
        out = VkzgAmdInfo()
        _check(load().vkzg_amd_prover_info(self._h, ctypes.byref(out)))
        return {f: getattr(out, f) for f, _ in VkzgAmdInfo._fields_}

    def compute_cell_kzg_proofs_batch(self, blobs, blob_count):
        if len(blobs) != blob_count * BYTES_PER_BLOB:
            raise ValueError("blobs.length must be blobCount * %d" % BYTES_PER_BLOB)
        nbytes = blob_count * CELL_PROOFS_PER_BLOB * BYTES_PER_PROOF
        out = ctypes.create_string_buffer(max(1, nbytes))
        src = ctypes.create_string_buffer(bytes(blobs), max(1, len(blobs)))
        _check(load().vkzg_compute_proofs_batch(self._h, out, src, blob_count))
        return out.raw[:nbytes]

    def recover_cells_batch(self, cells, cell_present, blob_count):
        per_blob = CELL_PROOFS_PER_BLOB * BYTES_PER_CELL
        if len(cells) != blob_count * per_blob:
            raise ValueError("cells.length must be blobCount * %d" % per_blob)
        if len(cell_present) != blob_count * CELL_PROOFS_PER_BLOB:
            raise ValueError("cellPresent.length must be blobCount * %d" % CELL_PROOFS_PER_BLOB)
        out = ctypes.create_string_buffer(max(1, len(cells)))
        c = ctypes.create_string_buffer(bytes(cells), max(1, len(cells)))
        pm = ctypes.create_string_buffer(bytes(cell_present), max(1, len(cell_present)))
        _check(load().vkzg_recover_cells_batch(self._h, out, c, pm, blob_count))
        return out.raw[: len(cells)]

    def compute_cells_and_kzg_proofs_batch(self, blobs, blob_count):
        """(cells, proofs) of each blob: the 2x extension and the 128 proofs."""
        if len(blobs) != blob_count * BYTES_PER_BLOB:
            raise ValueError("blobs.length must be blobCount * %d" % BYTES_PER_BLOB)
        ncells = blob_count * CELL_PROOFS_PER_BLOB * BYTES_PER_CELL
        nproofs = blob_count * CELL_PROOFS_PER_BLOB * BYTES_PER_PROOF
        cells = ctypes.create_string_buffer(max(1, ncells))
        proofs = ctypes.create_string_buffer(max(1, nproofs))
        src = ctypes.create_string_buffer(bytes(blobs), max(1, len(blobs)))
        _check(load().vkzg_amd_compute_cells_and_proofs(self._h, cells, proofs, src, blob_count))
        return cells.raw[:ncells], proofs.raw[:nproofs]

    def recover_cells_and_kzg_proofs_batch(self, cells, cell_present, blob_count):
        """(all 128 cells, 128 proofs) of each blob from any >= 64 of its cells."""
        per_blob = CELL_PROOFS_PER_BLOB * BYTES_PER_CELL
        if len(cells) != blob_count * per_blob:
            raise ValueError("cells.length must be blobCount * %d" % per_blob)
        if len(cell_present) != blob_count * CELL_PROOFS_PER_BLOB:
            raise ValueError("cellPresent.length must be blobCount * %d" % CELL_PROOFS_PER_BLOB)
        nproofs = blob_count * CELL_PROOFS_PER_BLOB * BYTES_PER_PROOF
        out = ctypes.create_string_buffer(max(1, len(cells)))
        proofs = ctypes.create_string_buffer(max(1, nproofs))
        c = ctypes.create_string_buffer(bytes(cells), max(1, len(cells)))
        pm = ctypes.create_string_buffer(bytes(cell_present), max(1, len(cell_present)))
        _check(load().vkzg_amd_recover_cells_and_proofs(self._h, out, proofs, c, pm, blob_count))
        return out.raw[:len(cells)], proofs.raw[:nproofs]

    def load_g2_setup(self, g2_monomial):
        """The 65 compressed G2 setup points [tau^i]G2 (65 x 96 bytes) for verification."""
        g2 = bytes(g2_monomial)
        _check(load().vkzg_amd_load_g2_setup(self._h, g2, len(g2)))

    def verify_cell_kzg_proof_batch(self, commitments, cell_indices, cells, proofs):
        """verify_cell_kzg_proof_batch: commitments, cells and proofs as flat bytes
        (48, 2048 and 48 bytes per cell), cell_indices a sequence of ints."""
        n = len(cell_indices)
        if len(commitments) != 48 * n or len(proofs) != 48 * n or len(cells) != BYTES_PER_CELL * n:
            raise ValueError("commitments, cells and proofs must have one entry per cell index")
        ok = ctypes.c_int(0)
        idx = (ctypes.c_uint64 * max(1, n))(*cell_indices)
        _check(load().vkzg_amd_verify_cell_kzg_proof_batch(self._h, ctypes.byref(ok), bytes(commitments), idx,
                                                           bytes(cells), bytes(proofs), n))
        return bool(ok.value)

    def write_v1_tables(self, path):
        """The reference's tables file (version 1) for this prover's setup."""
        _check(load().vkzg_amd_write_v1_tables(self._h, str(path).encode()))

    def debug_fp_mul_rate(self, waves_per_simd, iters=2000):
        """(chip-wide Fp products per second at that occupancy, cycles per product of one wave)."""
        rate, cyc = ctypes.c_double(), ctypes.c_double()
        _check(load().vkzg_amd_debug_fp_mul_rate(self._h, waves_per_simd, iters, ctypes.byref(rate),
                                                 ctypes.byref(cyc)))
        return rate.value, cyc.value

    def debug_verify_terms(self, commitments, cell_indices, cells, proofs):
        """(verdict, A, RL): the verification and the two G1 inputs of its pairing check."""
        n = len(cell_indices)
        ok = ctypes.c_int(0)
        terms = ctypes.create_string_buffer(96)
        idx = (ctypes.c_uint64 * max(1, n))(*cell_indices)
        _check(load().vkzg_amd_debug_verify_terms(self._h, ctypes.byref(ok), terms, bytes(commitments), idx,
                                                  bytes(cells), bytes(proofs), n))
        return bool(ok.value), terms.raw[:48], terms.raw[48:]

    def blob_to_kzg_commitments(self, blobs, blob_count):
        """Commitments (blob_count x 48 bytes) of blob_count blobs."""
        if len(blobs) != blob_count * BYTES_PER_BLOB:
            raise ValueError("blobs.length must be blobCount * %d" % BYTES_PER_BLOB)
        out = ctypes.create_string_buffer(max(1, 48 * blob_count))
        src = ctypes.create_string_buffer(bytes(blobs), max(1, len(blobs)))
        _check(load().vkzg_amd_blob_to_commitments(self._h, out, src, blob_count))
        return out.raw[: 48 * blob_count]

    def debug_pairing(self, g1_points, g2_points):
        """Device e(P_i, Q_i)^3 for compressed P_i (48 B) and Q_i (96 B): 576 bytes each."""
        n = len(g1_points) // 48
        if len(g1_points) != 48 * n or len(g2_points) != 96 * n:
            raise ValueError("need as many 48-byte G1 as 96-byte G2 points")
        out = ctypes.create_string_buffer(max(1, 576 * n))
        _check(load().vkzg_amd_debug_pairing(self._h, out, bytes(g1_points), bytes(g2_points), n))
        return out.raw[: 576 * n]

    def debug_g1_msm(self, points, scalars):
        """sum [s_i] P_i on the device: 48-byte points, 32-byte big-endian scalars below r."""
        n = len(points) // 48
        if len(points) != 48 * n or len(scalars) != 32 * n:
            raise ValueError("need as many 48-byte points as 32-byte scalars")
        out = ctypes.create_string_buffer(48)
        _check(load().vkzg_amd_debug_g1_msm(self._h, out, bytes(points), bytes(scalars), n))
        return out.raw

    def get_tuning(self, blob_count):
        t = Tuning()
        _check(load().vkzg_amd_get_tuning(self._h, blob_count, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def set_tuning(self, blob_count, **fields):
        """Set fields of the tuning of blob_count's size class (others unchanged)."""
        t = Tuning()
        _check(load().vkzg_amd_get_tuning(self._h, blob_count, ctypes.byref(t)))
        for k, v in fields.items():
            setattr(t, k, v)
        _check(load().vkzg_amd_set_tuning(self._h, blob_count, ctypes.byref(t)))

    def set_graphs(self, enable=True):
        _check(load().vkzg_amd_set_graphs(self._h, 1 if enable else 0))

    def load_window_table(self, window_bits):
        """Keep a smaller phase-A table (window 10-13) resident for tunings to name."""
        _check(load().vkzg_amd_load_window_table(self._h, window_bits))

    def debug_mad_rate(self, waves_per_simd, iters=2000):
        """(v_mad_u64_u32 lane-ops per second over the chip, nominal clock in MHz)."""
        rate, mhz = ctypes.c_double(), ctypes.c_double()
        _check(load().vkzg_amd_debug_mad_rate(self._h, waves_per_simd, iters, ctypes.byref(rate),
                                              ctypes.byref(mhz)))
        return rate.value, mhz.value

    def debug_force_checked(self, enable=True):
        """Send phase A, phase B and the bucket reduction of later proof calls
        through their checked recompute (a test switch; proofs are unchanged)."""
        _check(load().vkzg_amd_debug_force_checked(self._h, 1 if enable else 0))

    def autotune(self):
        """Run the autotuner; returns its wall time in ms."""
        ms = ctypes.c_double()
        _check(load().vkzg_amd_autotune(self._h, ctypes.byref(ms)))
        return ms.value

    def set_stage_timing(self, enable=True):
        _check(load().vkzg_amd_set_stage_timing(self._h, 1 if enable else 0))

    def last_stage_times(self):
        """{stage name: ms} of the last call (HIP events on the kernels' stream)."""
        ms = (ctypes.c_float * 16)()
        names = (ctypes.c_char_p * 16)()
        n = load().vkzg_amd_last_stage_times(self._h, ms, names, 16)
        return {names[i].decode(): float(ms[i]) for i in range(n)}

    # Device-pointer entry points (device addresses as integers, e.g. from vkzg_amd.hip).
    def compute_proofs_device(self, d_proofs, d_blobs, blob_count, stream=None):
        _check(load().vkzg_amd_compute_proofs_device(self._h, d_proofs, d_blobs, blob_count,
                                                     stream))

    def recover_cells_device(self, d_out, d_cells, d_present, h_present, blob_count, stream=None):
        pm = ctypes.create_string_buffer(bytes(h_present), max(1, len(h_present)))
        _check(load().vkzg_amd_recover_cells_device(self._h, d_out, d_cells, d_present, pm,
                                                    blob_count, stream))


class Server:
    """Continuous-batching request server on a Prover (vkzg_amd_server_*).

    submit_* return a ticket immediately; the request's input and output
    buffers are held by the ticket until wait() returns."""

    def __init__(self, prover, batch_blobs=0, slots=0, hbm_budget_bytes=0):
        L = load()
        o = ServerOptions()
        L.vkzg_amd_server_options_default(ctypes.byref(o))
        o.batch_blobs, o.slots, o.hbm_budget_bytes = batch_blobs, slots, hbm_budget_bytes
        h = ctypes.c_void_p()
        _check(L.vkzg_amd_server_new(ctypes.byref(h), prover.handle, ctypes.byref(o)))
        self._h = h
        self._prover = prover
        self._held = {}

    def close(self):
        if getattr(self, "_h", None):
            load().vkzg_amd_server_free(self._h)
            self._h = None
            self._held.clear()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_proofs(self, blobs, n):
        if len(blobs) != n * BYTES_PER_BLOB:
            raise ValueError("blobs.length must be blobCount * %d" % BYTES_PER_BLOB)
        src = ctypes.create_string_buffer(bytes(blobs), max(1, len(blobs)))
        out = ctypes.create_string_buffer(max(1, n * CELL_PROOFS_PER_BLOB * BYTES_PER_PROOF))
        t = ctypes.c_uint64()
        _check(load().vkzg_amd_server_submit_proofs(self._h, out, src, n, ctypes.byref(t)))
        self._held[t.value] = (src, out, n * CELL_PROOFS_PER_BLOB * BYTES_PER_PROOF)
        return t.value

    def submit_recovery(self, cells, present, n):
        per_blob = CELL_PROOFS_PER_BLOB * BYTES_PER_CELL
        if len(cells) != n * per_blob or len(present) != n * CELL_PROOFS_PER_BLOB:
            raise ValueError("cells/present length mismatch")
        src = ctypes.create_string_buffer(bytes(cells), max(1, len(cells)))
        pm = ctypes.create_string_buffer(bytes(present), max(1, len(present)))
        out = ctypes.create_string_buffer(max(1, len(cells)))
        t = ctypes.c_uint64()
        _check(load().vkzg_amd_server_submit_recovery(self._h, out, src, pm, n, ctypes.byref(t)))
        self._held[t.value] = (src, pm, out, len(cells))
        return t.value

    # Raw-pointer submissions (addresses of caller-owned host buffers that
    # must outlive the request); used by benchmarks to avoid copies.
    def submit_proofs_raw(self, out_addr, in_addr, n):
        t = ctypes.c_uint64()
        _check(load().vkzg_amd_server_submit_proofs(self._h, out_addr, in_addr, n, ctypes.byref(t)))
        return t.value

    def submit_recovery_raw(self, out_addr, in_addr, present_addr, n):
        t = ctypes.c_uint64()
        _check(load().vkzg_amd_server_submit_recovery(self._h, out_addr, in_addr, present_addr, n,
                                                      ctypes.byref(t)))
        return t.value

    def wait_raw(self, ticket):
        """(result code, info dict) without raising."""
        info = RequestInfo()
        rc = load().vkzg_amd_server_wait(self._h, ticket, ctypes.byref(info))
        return rc, {f: getattr(info, f) for f, _ in RequestInfo._fields_}

    def poll(self, ticket):
        d = ctypes.c_int()
        _check(load().vkzg_amd_server_poll(self._h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        """Returns (result bytes, info dict); raises VkzgError on a failed request."""
        info = RequestInfo()
        rc = load().vkzg_amd_server_wait(self._h, ticket, ctypes.byref(info))
        held = self._held.pop(ticket)
        out, nbytes = held[-2], held[-1]
        meta = {f: getattr(info, f) for f, _ in RequestInfo._fields_}
        if rc != OK:
            raise VkzgError(rc)
        return out.raw[:nbytes], meta

    def stats(self):
        st = ServerStats()
        _check(load().vkzg_amd_server_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}


# ---- module-level singleton, as in Vkzg.java ----
_lock = threading.Lock()
_prover = None


def init(max_batch_size=0):
    global _prover
    with _lock:
        if _prover is not None:
            raise RuntimeError("already initialized")
        _prover = Prover(max_batch_size)


def deinit():
    global _prover
    with _lock:
        _loaded().close()
        _prover = None


def _loaded():
    if _prover is None:
        raise RuntimeError("not initialized; call init() first")
    return _prover


def compute_cell_kzg_proofs_batch(blobs, blob_count):
    with _lock:
        return _loaded().compute_cell_kzg_proofs_batch(blobs, blob_count)


def recover_cells_batch(cells, cell_present, blob_count):
    with _lock:
        return _loaded().recover_cells_batch(cells, cell_present, blob_count)


def device_name():
    with _lock:
        return _loaded().device_name()


def gpu_core_count():
    with _lock:
        return _loaded().gpu_core_count()


def debug_fp_mul(a_limbs, b_limbs, prover=None):
    """Products of canonical Fp values given as lists of 12-limb tuples (test hook)."""
    return _debug_mul(a_limbs, b_limbs, 12, prover, load().vkzg_amd_debug_fp_mul)


def debug_fr_mul(a_limbs, b_limbs, prover=None):
    """Products mod r of 256-bit values given as lists of 8-limb tuples (test hook)."""
    return _debug_mul(a_limbs, b_limbs, 8, prover, load().vkzg_amd_debug_fr_mul)


def debug_fp_inv(a_limbs, prover=None):
    """Inverses mod p (0 -> 0) of canonical Fp values as 12-limb tuples (test hook);
    an all-ones tuple marks a disagreement with the Fermat inverse."""
    n = len(a_limbs)
    A = (ctypes.c_uint32 * (n * 12))(*[x for v in a_limbs for x in v])
    C = (ctypes.c_uint32 * (n * 12))()
    _check(load().vkzg_amd_debug_fp_inv(prover.handle if prover else None, C, A, n))
    return [tuple(C[i * 12:(i + 1) * 12]) for i in range(n)]


def _debug_mul(a, b, limbs, prover, fn):
    n = len(a)
    A = (ctypes.c_uint32 * (n * limbs))(*[x for v in a for x in v])
    B = (ctypes.c_uint32 * (n * limbs))(*[x for v in b for x in v])
    C = (ctypes.c_uint32 * (n * limbs))()
    _check(fn(prover.handle if prover else None, C, A, B, n))
    return [tuple(C[i * limbs:(i + 1) * limbs]) for i in range(n)]


__all__ = [
    "Prover", "Server", "VkzgError", "init", "deinit", "compute_cell_kzg_proofs_batch",
    "recover_cells_batch", "device_name", "gpu_core_count", "error_string", "LIB_PATH",
]


def verify_challenge(commitments, cell_indices, cells, proofs, portable_sha=False):
    """The Fiat-Shamir challenge r of a verification call (32 bytes big-endian), host only."""
    n = len(cell_indices)
    out = ctypes.create_string_buffer(32)
    idx = (ctypes.c_uint64 * max(1, n))(*cell_indices)
    _check(load().vkzg_amd_verify_challenge(out, bytes(commitments), idx, bytes(cells), bytes(proofs), n,
                                            1 if portable_sha else 0))
    return out.raw


def sha256(data, portable=False):
    """SHA-256 of the library's host implementation (SHA extensions unless portable)."""
    out = ctypes.create_string_buffer(32)
    _check(load().vkzg_amd_sha256(out, bytes(data), len(data), 1 if portable else 0))
    return out.raw


def sha256_hw():
    return bool(load().vkzg_amd_sha256_hw())
