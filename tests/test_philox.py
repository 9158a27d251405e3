"""The production generator, on the host (no GPU): Philox4x32-10 as
philox.hpp builds it for the host, pinned by the Random123 known answers, and a
numpy restatement of the generator and of the per-attempt draw bundle that the
GPU tests hold the device draws to.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of
  (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,
  c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0), k += (W0, W1).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
KAT = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
       ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
       ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]


def philox(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on numpy arrays of 32-bit values (uint64 holds each product)."""
    m = np.uint64(0xFFFFFFFF)
    c = [np.asarray(x, np.uint64) & m for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0) & m, np.uint64(k1) & m
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & m, p1 & m, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & m, p0 & m]
        k0, k1 = (k0 + np.uint64(W0)) & m, (k1 + np.uint64(W1)) & m
    return c


def draw_bundle(seed, spin0, n_spins, n_att):
    """The production draws of spins [spin0, spin0 + n_spins), attempts [0, n_att):
    out[..., 0:3] the displacement normals (float64), out[..., 3] the permeability
    uniform.  Counter (attempt_lo, attempt_hi, spin, 0), key (seed_lo, seed_hi)."""
    att = np.arange(n_att, dtype=np.uint64)[None, :].repeat(n_spins, 0)
    spin = (np.uint64(spin0) + np.arange(n_spins, dtype=np.uint64))[:, None].repeat(n_att, 1)
    r = philox(att & np.uint64(0xFFFFFFFF), att >> np.uint64(32), spin, np.zeros_like(att),
               int(seed) & 0xFFFFFFFF, int(seed) >> 32)
    top = [1.0 + (x >> np.uint64(9)).astype(np.float64) / 2.0 ** 23 for x in r]      # [1, 2)
    rho1 = np.sqrt(-2.0 * np.log(2.0 - top[0]))
    rho2 = np.sqrt(-2.0 * np.log(2.0 - top[2]))
    u = ((r[0] & np.uint64(255)) | ((r[1] & np.uint64(255)) << np.uint64(8)) |
         ((r[2] & np.uint64(255)) << np.uint64(16))).astype(np.float64) / 2.0 ** 24
    return np.stack([rho1 * np.cos(2 * np.pi * top[1]), rho1 * np.sin(2 * np.pi * top[1]),
                     rho2 * np.cos(2 * np.pi * top[3]), u], axis=-1)


def test_numpy_restatement_matches_random123_known_answers():
    for c, k, want in KAT:
        got = [int(x) for x in philox(*[np.uint64(v) for v in c], *k)]
        assert got == list(want)


def _host_program(body):
    code = '#include "%s"\n#include <cstdio>\n#include <cstdint>\nint main(){%s}' % (
        os.path.join(ROOT, "spinwalk_amd", "csrc", "philox.hpp"), body)
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "k.cpp"), "w") as fh:
        fh.write(code)
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", os.path.join(d, "k"), os.path.join(d, "k.cpp")], check=True)
    return subprocess.run([os.path.join(d, "k")], capture_output=True, text=True, check=True).stdout.split()


def test_host_build_matches_random123_known_answers():
    body = ''.join('{auto r=sw::philox4x32_10(%du,%du,%du,%du,%du,%du);printf("%%u %%u %%u %%u\\n",'
                   'r.v[0],r.v[1],r.v[2],r.v[3]);}' % (c + k) for c, k, _ in KAT)
    out = [int(x) for x in _host_program(body + "return 0;")]
    assert out == [w for _, _, want in KAT for w in want]


def test_precomputed_first_product_gives_the_same_block():
    """philox4x32_10_pre with (M1 * c2) given is philox4x32_10, for many counters."""
    rng = np.random.default_rng(5)
    cases = rng.integers(0, 2 ** 32, size=(64, 6), dtype=np.uint64)
    body = ""
    for c0, c1, c2, c3, k0, k1 in cases.tolist():
        m = M1 * c2
        body += ('{auto a=sw::philox4x32_10(%du,%du,%du,%du,%du,%du);'
                 'auto b=sw::philox4x32_10_pre(%du,%du,%du,%du,%du,%du,%du);'
                 'printf("%%d\\n",(int)(a.v[0]==b.v[0]&&a.v[1]==b.v[1]&&a.v[2]==b.v[2]&&a.v[3]==b.v[3]));}'
                 % (c0, c1, c2, c3, k0, k1, c0, c1, m >> 32, m & 0xFFFFFFFF, c3, k0, k1))
    out = _host_program(body + "return 0;")
    assert out == ["1"] * len(cases)
    want = philox(cases[:, 0], cases[:, 1], cases[:, 2], cases[:, 3], 0, 0)     # and numpy agrees with a zero key
    body = ''.join('{auto r=sw::philox4x32_10(%du,%du,%du,%du,0u,0u);printf("%%u %%u %%u %%u\\n",r.v[0],r.v[1],'
                   'r.v[2],r.v[3]);}' % tuple(c[:4]) for c in cases.tolist())
    got = np.array([int(x) for x in _host_program(body + "return 0;")], np.uint64).reshape(-1, 4)
    np.testing.assert_array_equal(got, np.stack(want, axis=-1))


def test_bundle_bit_helpers():
    """unit12 is 1.0 | r >> 9 as a float in [1, 2); low_bytes3 packs the low bytes
    of three words; u24_to_unit divides by 2^24 exactly."""
    vals = [0, 1, 511, 512, 0x12345678, 0xDEADBEEF, 0xFFFFFFFF]
    body = ''.join('printf("%%a %%u %%a\\n",(double)sw::unit12(%du),sw::low_bytes3(%du,%du,%du),'
                   '(double)sw::u24_to_unit(sw::low_bytes3(%du,%du,%du)));' % (v, v, v ^ 0x55, v >> 3, v, v ^ 0x55, v >> 3)
                   for v in vals)
    out = _host_program(body + "return 0;")
    for i, v in enumerate(vals):
        f, b, u = float.fromhex(out[3 * i]), int(out[3 * i + 1]), float.fromhex(out[3 * i + 2])
        assert f == 1.0 + (v >> 9) / 2.0 ** 23
        want = (v & 255) | (((v ^ 0x55) & 255) << 8) | (((v >> 3) & 255) << 16)
        assert b == want and u == want / 2.0 ** 24


def test_bundle_statistics():
    d = draw_bundle(0x1234567890ABCDEF, 7, 512, 512).reshape(-1, 4)
    n = d.shape[0]
    for k in range(3):
        assert abs(d[:, k].mean()) < 5 / np.sqrt(n)
        assert abs(d[:, k].var() - 1) < 5 * np.sqrt(2 / n)
    assert 0 <= d[:, 3].min() and d[:, 3].max() < 1
    assert abs(d[:, 3].mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
