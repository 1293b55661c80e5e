"""The kernels' division by a launch-time divisor (dtfe_amd/csrc/fastdiv.h)
checked against the hardware '/' on the CPU: every divisor up to 4096 and
random divisors up to 2^32 - 1, each against numerators at the quotient
boundaries (q d - 1, q d, q d + 1), random numerators and 2^32 - 1."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdint>
#include <cstdio>
#include <random>
#include "fastdiv.h"

static long bad = 0, checked = 0;
static void check(const dtfe::FastDiv &f, uint32_t n) {
    ++checked;
    if (f.div(n) != n / f.d) {
        if (bad < 10) std::printf("d=%u n=%u got %u want %u\n", f.d, n, f.div(n), n / f.d);
        ++bad;
    }
}
static void divisor(uint32_t d, std::mt19937_64 &rng) {
    dtfe::FastDiv f(d);
    const uint32_t top = 0xffffffffu;
    for (uint32_t n : {0u, 1u, 2u, top, top - 1, d - 1, d, d + 1}) check(f, n);
    // quotient boundaries spread over the whole range
    for (int i = 0; i < 64; ++i) {
        uint64_t q = rng() % ((uint64_t)top / d + 1);
        uint64_t n = q * d;
        if (n > 0) check(f, (uint32_t)(n - 1));
        if (n <= top) check(f, (uint32_t)n);
        if (n + 1 <= top) check(f, (uint32_t)(n + 1));
        check(f, (uint32_t)rng());
    }
}
int main() {
    std::mt19937_64 rng(12345);
    for (uint32_t d = 1; d <= 4096; ++d) divisor(d, rng);
    for (int k = 0; k < 32; ++k) {  // powers of two and their neighbours
        uint64_t p = 1ull << k;
        for (int64_t e = -1; e <= 1; ++e) {
            uint64_t d = p + e;
            if (d >= 1 && d <= 0xffffffffull) divisor((uint32_t)d, rng);
        }
    }
    divisor(0xffffffffu, rng);
    for (int i = 0; i < 20000; ++i) {
        uint32_t d = (uint32_t)(rng() >> (32 + rng() % 32));
        if (d) divisor(d, rng);
    }
    std::printf("checked %ld bad %ld\n", checked, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fastdiv_matches_integer_division(tmp_path):
    src = tmp_path / "fastdiv_check.cpp"
    exe = tmp_path / "fastdiv_check"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "dtfe_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "bad 0" in r.stdout
