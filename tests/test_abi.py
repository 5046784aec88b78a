"""CPU tests of the drop-in boundary: libvkzg.so loads, exports every symbol
the headers declare, and honours the argument contract that needs no GPU
(reference include/vkzg.h:42-121, src/vkzg.cpp:170-194,1183-1218)."""
import ctypes
import os
import re
import subprocess

import pytest

import vkzg_amd
from vkzg_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(vkzg_\w+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = declared_functions("vkzg.h") | declared_functions("vkzg_amd.h")
    assert set(_lib.ABI_SYMBOLS) <= declared
    assert set(_lib.EXT_SYMBOLS) <= declared
    for name in sorted(declared):
        assert hasattr(L, name), name


def test_library_exports_only_the_declared_c_abi():
    """libvkzg.so exports the functions the two headers declare and nothing
    else: no C++ runtime symbols, no embedded data (a version script keeps
    them local)."""
    r = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exported = {line.split()[-1] for line in r.stdout.splitlines() if line.strip()}
    declared = declared_functions("vkzg.h") | declared_functions("vkzg_amd.h")
    assert exported == declared, (sorted(exported - declared), sorted(declared - exported))


def test_reference_abi_is_exactly_eight_functions():
    assert declared_functions("vkzg.h") == set(_lib.ABI_SYMBOLS)


def test_size_macros_and_codes_match_reference_values():
    text = open(os.path.join(REPO, "include", "vkzg.h")).read()
    for macro, value in [("VKZG_FIELD_ELEMENTS_PER_BLOB", "4096"), ("VKZG_NUM_CELL_PROOFS", "128"),
                         ("VKZG_BYTES_PER_FIELD_ELEMENT", "32"), ("VKZG_BYTES_PER_PROOF", "48"),
                         ("VKZG_FIELD_ELEMENTS_PER_CELL", "64")]:
        assert re.search(r"#define %s %s\b" % (macro, value), text), macro
    for name, code in [("VKZG_OK", 0), ("VKZG_ERR_BADARGS", 1), ("VKZG_ERR_MALLOC", 2),
                       ("VKZG_ERR_IO", 3), ("VKZG_ERR_SETUP", 4), ("VKZG_ERR_GPU", 5),
                       ("VKZG_ERR_INVALID_BLOB", 6)]:
        assert re.search(r"%s = %d\b" % (name, code), text), name


def test_error_strings():
    L = _lib.load()
    want = ["ok", "invalid argument", "allocation failed", "i/o error", "malformed trusted setup",
            "gpu error", "blob contains a non-canonical field element"]
    for code, s in enumerate(want):
        assert L.vkzg_error_string(code).decode() == s
    assert L.vkzg_error_string(99).decode() == "unknown error"


def test_options_are_four_bytes_and_default_to_zero():
    L = _lib.load()
    assert ctypes.sizeof(_lib.VkzgOptions) == 4
    o = _lib.VkzgOptions(1234)
    L.vkzg_options_default(ctypes.byref(o))
    assert o.max_batch_size == 0
    L.vkzg_options_default(None)  # NULL-safe


def test_null_arguments_without_gpu():
    L = _lib.load()
    buf = ctypes.create_string_buffer(262144)
    assert L.vkzg_compute_proofs_batch(None, buf, buf, 1) == vkzg_amd.ERR_BADARGS
    assert L.vkzg_recover_cells_batch(None, buf, buf, buf, 1) == vkzg_amd.ERR_BADARGS
    assert L.vkzg_prover_new(None, None) == vkzg_amd.ERR_BADARGS
    L.vkzg_prover_free(None)
    assert L.vkzg_prover_device_name(None) == b""
    assert L.vkzg_prover_gpu_core_count(None) == 0
    assert L.vkzg_amd_compute_proofs_device(None, buf, buf, 1, None) == vkzg_amd.ERR_BADARGS
    bad = ctypes.c_void_p()
    assert L.vkzg_amd_prover_new_with_setup(ctypes.byref(bad), None, b"x", 1) == \
        vkzg_amd.ERR_BADARGS


def test_python_mirror_argument_checks():
    p = object.__new__(vkzg_amd.Prover)
    p._h = None
    with pytest.raises(ValueError):
        p.compute_cell_kzg_proofs_batch(b"\0" * 100, 1)
    with pytest.raises(ValueError):
        p.recover_cells_batch(b"\0" * 262144, b"\0" * 127, 1)
    with pytest.raises(RuntimeError):
        vkzg_amd.compute_cell_kzg_proofs_batch(b"", 0)


def test_java_style_harness_builds_and_checks_symbols(tmp_path):
    """tests/c/java_like.c opens the library with dlopen/dlsym exactly like the
    Java FFM binding; without a GPU it must get as far as prover_new."""
    exe = tmp_path / "java_like"
    subprocess.run(["gcc", "-O1", "-o", str(exe), os.path.join(REPO, "tests", "c", "java_like.c"),
                    "-ldl"], check=True)
    r = subprocess.run([str(exe), _lib.LIB_PATH, "--symbols-only"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "8 symbols ok" in r.stdout


def test_verification_entry_points_reject_bad_arguments_without_a_gpu():
    """The additive verification, commitment and tables entry points check
    their arguments before any device call (NULL prover or buffers, an
    out-of-range cell index, a wrong G2 setup length, rate-hook bounds)."""
    L = _lib.load()
    ok = ctypes.c_int(7)
    idx = (ctypes.c_uint64 * 2)(1, 2)
    buf = ctypes.create_string_buffer(4096)
    BAD = 1
    assert L.vkzg_amd_verify_cell_kzg_proof_batch(None, ctypes.byref(ok), buf, idx, buf, buf, 2) == BAD
    fake = ctypes.c_void_p(1)  # never dereferenced: the argument checks come first
    assert L.vkzg_amd_verify_cell_kzg_proof_batch(fake, None, buf, idx, buf, buf, 2) == BAD
    assert L.vkzg_amd_verify_cell_kzg_proof_batch(fake, ctypes.byref(ok), None, idx, buf, buf, 2) == BAD
    assert ok.value == 0  # *ok is cleared before the other checks
    big = (ctypes.c_uint64 * 2)(1, 128)
    assert L.vkzg_amd_verify_cell_kzg_proof_batch(fake, ctypes.byref(ok), buf, big, buf, buf, 2) == BAD
    assert L.vkzg_amd_load_g2_setup(None, buf, 65 * 96) == BAD
    assert L.vkzg_amd_load_g2_setup(fake, buf, 64 * 96) == BAD
    assert L.vkzg_amd_load_g2_setup(fake, None, 65 * 96) == BAD
    assert L.vkzg_amd_blob_to_commitments(None, buf, buf, 1) == BAD
    assert L.vkzg_amd_blob_to_commitments(fake, None, buf, 1) == BAD
    assert L.vkzg_amd_write_v1_tables(None, b"/tmp/x") == BAD
    assert L.vkzg_amd_write_v1_tables(fake, None) == BAD
    r, c = ctypes.c_double(), ctypes.c_double()
    assert L.vkzg_amd_debug_fp_mul_rate(fake, 0, 10, ctypes.byref(r), ctypes.byref(c)) == BAD
    assert L.vkzg_amd_debug_fp_mul_rate(fake, 9, 10, ctypes.byref(r), ctypes.byref(c)) == BAD
    assert L.vkzg_amd_debug_g1_msm(fake, buf, buf, buf, 0) == BAD
    assert L.vkzg_amd_sha256(None, buf, 1, 0) == BAD
    assert L.vkzg_amd_verify_challenge(None, buf, idx, buf, buf, 2, 0) == BAD
