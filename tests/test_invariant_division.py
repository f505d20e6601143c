"""ray-tracer_amd/csrc/fastdiv.hpp (division by a launch-invariant divisor
with a multiplier) compiled on the host and checked against plain unsigned
division: every divisor up to 4096, divisors near powers of two and near
2^32, and for each the dividends 0, 1, d - 1, d, d + 1, multiples of d +-1,
2^32 - 1 and random ones."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include "fastdiv.hpp"
int main() {
    unsigned long long bad = 0, n_checked = 0;
    unsigned d, n;
    while (std::scanf("%u %u", &d, &n) == 2) {
        const rsh::FastDiv f = rsh::fast_div(d);
        if (rsh::fdiv(n, f) != n / d) { if (bad < 5) std::printf("bad %u %u\n", d, n); bad++; }
        n_checked++;
    }
    std::printf("%llu %llu\n", n_checked, bad);
    return 0;
}
"""


def cases():
    rng = np.random.default_rng(11)
    ds = list(range(1, 4097))
    for b in range(1, 32):
        ds += [(1 << b) - 1, 1 << b, (1 << b) + 1]
    ds += [0xFFFFFFFF, 0xFFFFFFFE, 0x80000001, 1920, 1080, 1920 * 1080, 23_543_814, 1920 * 3]
    ds += [int(x) for x in rng.integers(2, 1 << 32, 2000, dtype=np.uint64)]
    out = []
    for d in ds:
        d &= 0xFFFFFFFF
        if d == 0:
            continue
        ns = {0, 1, d - 1, d, d + 1, 0xFFFFFFFF, 0xFFFFFFFE}
        k = 0xFFFFFFFF // d
        for q in (1, 2, k // 2, k - 1, k):
            for e in (-1, 0, 1):
                ns.add(q * d + e)
        ns |= {int(x) for x in rng.integers(0, 1 << 32, 8, dtype=np.uint64)}
        out += [(d, n) for n in ns if 0 <= n <= 0xFFFFFFFF]
    return out


def test_fastdiv_matches_division(tmp_path):
    src = tmp_path / "fd.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "fd"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "ray-tracer_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    data = "\n".join(f"{d} {n}" for d, n in cases())
    r = subprocess.run([str(exe)], input=data, capture_output=True, text=True, check=True)
    checked, bad = map(int, r.stdout.strip().splitlines()[-1].split())
    assert checked > 100_000
    assert bad == 0, r.stdout
