"""CPU test of the device's binary-GCD Fp inversion (vkzg_amd/csrc/inverse.h),
compiled for the host with g++: y * inverse(y) = 1 mod p, checked with Python
integers on random values, values near 0, p and 2p, powers of two and values
with long runs of equal bits."""
import os
import random
import subprocess

P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
HERE = os.path.dirname(os.path.abspath(__file__))


def test_binary_gcd_inverse(tmp_path):
    exe = str(tmp_path / "inverse_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "c", "inverse_host.cpp")],
                   check=True)
    rng = random.Random(2024)
    ys = [1, 2, 3, 4, P - 1, P + 1, 2 * P - 1, P - 2, (P - 1) // 2, (P + 1) // 2, 2 * P - 2]
    ys += [1 << k for k in range(0, 382)]
    ys += [(1 << k) - 1 for k in range(1, 382)]
    ys += [P - (1 << k) for k in range(0, 380)]
    ys += [rng.randrange(1, 2 * P) for _ in range(20000)]
    ys += [rng.randrange(1, 1 << 60) for _ in range(2000)]
    ys = [y for y in ys if y % P and y < 1 << 382]
    out = subprocess.run([exe], input="\n".join("%x" % y for y in ys) + "\n", capture_output=True,
                         text=True, check=True).stdout.split("\n")
    assert len([l for l in out if l]) == len(ys)
    for y, line in zip(ys, out):
        f = [int(t) for t in line.split()]
        assert f[0] == 1, hex(y)
        v = sum(f[1 + i] << (28 * i) for i in range(14))  # limb 13 carries the sign
        assert abs(v) < 31 * P, hex(y)
        assert v * y % P == 1, hex(y)
