"""The persistent launches' size guard (cgspline_amd/csrc/cgs_limits.hpp),
compiled with g++ on the CPU: a launch is refused exactly when its 32-bit
work counter could wrap."""
from __future__ import annotations

import os
import subprocess

import conftest

SRC = r"""
#include <cstdio>
#include <cstdint>
#include "cgs_limits.hpp"
int main()
{
    using namespace cgs;
    int bad = 0;
    auto expect = [&](bool got, bool want, const char* what) {
        if (got != want) { std::printf("FAIL %s\n", what); bad++; }
    };
    // 256 CUs x 3 workgroups x 4 waves, 64 items per fetch
    const uint64_t g = 768;
    const uint64_t lim = (1ull << 32) - g * 4 * 64;
    expect(counterFits(lim - 1, g, 64), true, "just below");
    expect(counterFits(lim, g, 64), false, "at the limit");
    expect(counterFits(0xFFFFFFFFull, g, 64), false, "2^32-1 items");
    expect(counterFits(0, g, 64), true, "empty");
    // the batched kernel's chunk multiplies the overshoot
    expect(counterFits(lim - 1, g, 4096), false, "chunk 4096");
    expect(counterFits((1ull << 32) - g * 4 * 4096 - 1, g, 4096), true, "chunk 4096, below");
    // a 65535 x 65535 frame: 8192^2 tiles x 64 = 2^32 items
    expect(renderItems(65535, 65535) == (1ull << 32), true, "render items");
    expect(counterFits(renderItems(65535, 65535), g, 64), false, "65535^2 frame");
    expect(counterFits(renderItems(2048, 2048), g, 64), true, "2048^2 frame");
    expect(renderItems(9, 1) == 128, true, "edge tiles whole");
    std::printf("%s\n", bad ? "failed" : "ok");
    return bad;
}
"""


def test_counter_guard(tmp_path):
    src = tmp_path / "limits.cpp"
    src.write_text(SRC)
    exe = tmp_path / "limits"
    inc = os.path.join(conftest.ROOT, "cgspline_amd", "csrc")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", inc, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
