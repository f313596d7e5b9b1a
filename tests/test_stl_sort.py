"""The device restatement of libstdc++'s std::sort (indialign_amd/csrc/stl_sort.h)
must produce the same permutation as the host std::sort for the comparator
the anchor-seed builder uses (count descending, seeds.h:237-238), including
tie-heavy inputs and inputs long enough to reach the heapsort fallback."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include "stl_sort.h"
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
struct CI { int count, a; };
struct Desc { bool operator()(const CI &x, const CI &y) const { return x.count > y.count; } };
int main() {
    std::mt19937_64 rng(12345);
    long checked = 0;
    for (int trial = 0; trial < 20000; trial++) {
        int n = (int)(rng() % (trial < 15000 ? 200 : 3000));
        int range = 1 + (int)(rng() % (trial % 3 == 0 ? 3 : (trial % 3 == 1 ? 20 : 100000)));
        std::vector<CI> v(n);
        for (int i = 0; i < n; i++) v[i] = CI{(int)(rng() % range), i};
        if (trial % 7 == 0) std::sort(v.begin(), v.end(), [](const CI &x, const CI &y) { return x.count < y.count; });
        if (trial % 11 == 0) for (int i = 0; i < n; i++) v[i].count = (i % 2) ? i : n - i;  // organ pipe
        std::vector<CI> a = v, b = v;
        std::sort(a.begin(), a.end(), Desc());
        if (!b.empty()) ia_sort::sort(b.data(), b.data() + b.size(), Desc());
        for (int i = 0; i < n; i++)
            if (a[i].a != b[i].a || a[i].count != b[i].count) { std::printf("MISMATCH trial %d n %d at %d\n", trial, n, i); return 1; }
        checked += n;
    }
    // heapsort fallback: median-of-3 killer style sequences force depth exhaustion
    for (int n : {64, 256, 1024, 4096}) {
        std::vector<CI> v(n);
        for (int i = 0; i < n; i++) v[i] = CI{i % 2 ? n + i : i, i};
        for (int rep = 0; rep < 50; rep++) {
            std::shuffle(v.begin(), v.end(), rng);
            std::vector<CI> a = v, b = v;
            std::sort(a.begin(), a.end(), Desc());
            ia_sort::sort(b.data(), b.data() + n, Desc());
            for (int i = 0; i < n; i++) if (a[i].a != b[i].a) { std::printf("MISMATCH heap n %d\n", n); return 1; }
        }
    }
    // direct check of the heap path against std::partial_sort(first, last, last)
    for (int trial = 0; trial < 2000; trial++) {
        int n = 1 + (int)(rng() % 300);
        std::vector<CI> v(n);
        for (int i = 0; i < n; i++) v[i] = CI{(int)(rng() % 7), i};
        std::vector<CI> a = v, b = v;
        std::partial_sort(a.begin(), a.end(), a.end(), Desc());
        ia_sort::heap_sort_(b.data(), (long)n, Desc());
        for (int i = 0; i < n; i++) if (a[i].a != b[i].a) { std::printf("MISMATCH partial_sort n %d\n", n); return 1; }
    }
    std::printf("OK %ld\n", checked);
    return 0;
}
"""


def test_stl_sort_matches_libstdcxx():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        with open(src, "w") as f:
            f.write(PROG)
        r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "indialign_amd", "csrc"),
                            src, "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.fail(r.stderr[-3000:])
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout
