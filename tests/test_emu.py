"""CPU tests of the device tile logic: tests/emu/emu.cpp compiles the same
headers as the HIP kernels (gpu-resident-library_amd/csrc/*.hpp) with g++ and
runs each workgroup sequentially; results must equal the oracle bit for bit
for every ring size, register radix and twiddle layout."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
u64p = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def emu():
    if shutil.which("g++") is None:
        pytest.skip("g++ absent")
    d = tempfile.mkdtemp()
    so = os.path.join(d, "libemu.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
                           "-I" + os.path.join(ROOT, "gpu-resident-library_amd", "csrc"),
                           "-o", so, os.path.join(HERE, "emu", "emu.cpp")])
    lib = C.CDLL(so)
    lib.emu_ntt.argtypes = [u64p, C.c_uint32, C.c_uint64, u64p, u64p, C.c_int, C.c_uint64, C.c_uint64,
                            C.c_int, C.c_uint32]
    for f in ("emu_mulmod", "emu_shoup", "emu_mulhi"):
        getattr(lib, f).restype = C.c_uint64
        getattr(lib, f).argtypes = [C.c_uint64] * 3 if f != "emu_mulhi" else [C.c_uint64] * 2
    lib.emu_redc.restype = C.c_uint64
    lib.emu_redc.argtypes = [C.c_uint64] * 3
    yield lib
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("logn", list(range(1, 18)))
def test_ntt_engine_matches_oracle(emu, logn):
    n = 1 << logn
    for bits in (50, 60):
        q = int(O.native_primes(1, bits if logn > 3 else 30, n)[0])
        psi = int(O.lib.or_native_root(n, q))
        fr, fp, ir, ip, ni, nip = O.root_tables(n, q, psi)
        x = np.random.default_rng(logn + bits).integers(0, q, n, dtype=np.uint64)
        x[:2] = [q - 1, q - 1]
        want = O.ntt(x, q, fr, fp)
        for rmax in (2, 3, 4):
            for paired in (0, 1):
                y = x.copy()
                emu.emu_ntt(O.P(y), n, q, O.P(fr), O.P(fp), 1, 0, 0, paired, rmax)
                assert np.array_equal(y, want), (bits, rmax, paired)
                z = want.copy()
                emu.emu_ntt(O.P(z), n, q, O.P(ir), O.P(ip), 0, ni, nip, paired, rmax)
                assert np.array_equal(z, x), (bits, rmax, paired)


def test_modarith_edges(emu):
    rng = np.random.default_rng(1)
    for bits in (30, 50, 59, 60):
        q = int(O.native_primes(1, bits, 1 << 10)[0])
        vals = [0, 1, q - 1, q - 2, q // 2] + [int(v) for v in rng.integers(0, q, 200, dtype=np.uint64)]
        for a in vals[:40]:
            for b in vals[:40]:
                assert emu.emu_mulmod(a, b, q) == (a * b) % q
                assert emu.emu_shoup(a, b, q) == (a * b) % q
        for _ in range(2000):
            a, b = (int(v) for v in rng.integers(0, 1 << 63, 2, dtype=np.uint64))
            a, b = a * 2 + 1, b
            assert emu.emu_mulhi(a, b) == (a * b) >> 64
        R = 1 << 64
        for _ in range(500):
            T = int(rng.integers(0, q, dtype=np.uint64)) * int(rng.integers(0, 1 << 63, dtype=np.uint64))
            T %= q * R
            assert emu.emu_redc(T % R, T >> 64, q) == (T * pow(R, -1, q)) % q
