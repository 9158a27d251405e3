"""fmod360 (spinwalk_amd/csrc/fmod360.hpp), the remainder the Kahan phase mode
rotates by at events, equals the C library's fmodf(v, 360) bit for bit: on the
host build of the same header, over every 61st float bit pattern below 2^15,
their negatives, multiples of 360 and their neighbours, and random bit patterns
(NaN, infinities, denormals and the 2^24 fallback boundary included)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

CODE = r'''
#include "%s"
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <random>
int main() {
    uint64_t bad = 0, n = 0;
    auto chk = [&](float v) {
        const float a = sw::fmod360(v), b = fmodf(v, 360.f);
        uint32_t ia, ib; std::memcpy(&ia, &a, 4); std::memcpy(&ib, &b, 4);
        ++n;
        if (ia != ib && !(a != a && b != b)) { if (bad < 5) std::printf("v=%%.9g got %%.9g want %%.9g\n", v, a, b); ++bad; }
    };
    for (uint32_t u = 0; u < 0x47000000u; u += 61) { float v; std::memcpy(&v, &u, 4); chk(v); chk(-v); }
    for (int k = -50000; k <= 50000; k++) {
        const float v = 360.f * k;
        chk(v); chk(nextafterf(v, 1e30f)); chk(nextafterf(v, -1e30f));
    }
    for (float v : {16777215.f, 16777216.f, 16777218.f, -16777215.f, -16777216.f, 1e30f, -1e30f}) chk(v);
    std::mt19937_64 g(1);
    for (int i = 0; i < 5000000; i++) { const uint32_t u = (uint32_t)g(); float v; std::memcpy(&v, &u, 4); chk(v); }
    std::printf("%%llu %%llu\n", (unsigned long long)n, (unsigned long long)bad);
    return bad != 0;
}
'''


def test_fmod360_equals_fmodf():
    src = os.path.join(HERE, "..", "spinwalk_amd", "csrc", "fmod360.hpp")
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "f.cpp"), "w") as fh:
        fh.write(CODE % os.path.abspath(src))
    exe = os.path.join(d, "f")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(d, "f.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    n, bad = map(int, r.stdout.split()[-2:])
    assert n > 30_000_000 and bad == 0
