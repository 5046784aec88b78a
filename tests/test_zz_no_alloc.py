"""vkzg_amd_recover_cells_device_noalloc against vkzg_amd_recover_cells_device,
through the C ABI: bit-identical cells on two calls in a row (the second
smaller), the blob limit, and no device memory taken by the calls."""
import ctypes
import random

import pytest

import vkzg_amd
from vkzg_amd import hip

pytestmark = pytest.mark.gpu

CELL = 2048
CELLS = 128 * CELL
MAX_BLOBS = 16


def _inputs(rng, cells_of, names, n):
    ins, pms, srcs = [], [], []
    for b in range(n):
        src = cells_of(names[b % len(names)])
        keep = rng.sample(range(128), rng.randint(64, 127))
        c = bytearray(rng.randbytes(CELLS))
        pm = bytearray(128)
        for k in keep:
            c[CELL * k:CELL * (k + 1)] = src[CELL * k:CELL * (k + 1)]
            pm[k] = 1
        ins.append(bytes(c))
        pms.append(bytes(pm))
        srcs.append(src)
    return b"".join(ins), b"".join(pms), b"".join(srcs)


def _free_bytes():
    H = hip.lib()
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    H.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    hip.check(H.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)))
    return free.value


def test_noalloc_recovery_matches_device_entry(gpu_prover, vectors, cells_of):
    L = vkzg_amd._lib.load()
    rng = random.Random(0x7A11)
    names = vectors.valid_compute()
    for n in (MAX_BLOBS, 5):  # two calls in a row, the second smaller
        cells, present, want = _inputs(rng, cells_of, names, n)
        d_cells = hip.DeviceBuffer.from_bytes(cells)
        d_present = hip.DeviceBuffer.from_bytes(present)
        d_ref = hip.DeviceBuffer(n * CELLS)
        d_out = hip.DeviceBuffer(n * CELLS)
        d_out.zero()
        hpm = ctypes.create_string_buffer(present, len(present))
        assert L.vkzg_amd_recover_cells_device(gpu_prover.handle, d_ref.ptr, d_cells.ptr, d_present.ptr,
                                               hpm, n, None) == vkzg_amd.OK
        before = _free_bytes()
        assert L.vkzg_amd_recover_cells_device_noalloc(gpu_prover.handle, d_out.ptr, d_cells.ptr,
                                                       d_present.ptr, hpm, n, None) == vkzg_amd.OK
        assert _free_bytes() == before
        ref = d_ref.to_bytes()
        assert ref == want
        assert d_out.to_bytes() == ref
        for b in (d_cells, d_present, d_ref, d_out):
            b.free()


def test_noalloc_recovery_limits(gpu_prover):
    L = vkzg_amd._lib.load()
    n = MAX_BLOBS + 1
    d = hip.DeviceBuffer(n * CELLS)
    hpm = ctypes.create_string_buffer(bytes([1]) * (128 * n), 128 * n)
    assert L.vkzg_amd_recover_cells_device_noalloc(gpu_prover.handle, d.ptr, d.ptr, d.ptr, hpm, n,
                                                   None) == vkzg_amd.ERR_BADARGS
    assert L.vkzg_amd_recover_cells_device_noalloc(gpu_prover.handle, d.ptr, d.ptr, d.ptr, hpm, 0,
                                                   None) == vkzg_amd.OK
    assert L.vkzg_amd_recover_cells_device_noalloc(None, d.ptr, d.ptr, d.ptr, hpm, 1,
                                                   None) == vkzg_amd.ERR_BADARGS
    bad = ctypes.create_string_buffer(bytes([1] * 63 + [0] * 65), 128)  # 65 missing
    assert L.vkzg_amd_recover_cells_device_noalloc(gpu_prover.handle, d.ptr, d.ptr, d.ptr, bad, 1,
                                                   None) == vkzg_amd.ERR_BADARGS
    d.free()
