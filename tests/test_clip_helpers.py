"""CPU check of the clip kernel's closed-form vertex move.

The reference moves the vertices in conflict with a new plane to the tail
of the vertex list with a serial swap loop (convex_cell.cu:706-721): i scans
from the front, end from the back, a conflicting vertex at i is swapped with
the one at end - 1.  k_clip computes every vertex's destination from the
conflict mask instead (popcounts and k-th set-bit selections on a 128-bit
mask).  This test takes that code and its mask helpers verbatim from
rpd_kernels.hip, compiles them for the host with g++ (one lane per cell,
up to 128 vertices, both mask words) and compares the permutation with the
serial loop on random masks.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "libmat_amd", "csrc", "device",
                   "rpd_kernels.hip")

HEAD = r"""
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
using std::max;
using std::min;
static inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
static inline int __ffsll(long long x) { return __builtin_ffsll(x); }
static inline int __clzll(long long x) {
  return __builtin_clzll((unsigned long long)x);
}
// one lane per cell: the wavefront vote is the lane's own predicate
static inline unsigned long long ballot(bool p) { return p ? 1ull : 0ull; }
"""

MAIN = r"""
// the reference's serial swap loop: perm[pos] = old index at pos
static void serial(int nb, const int* conf, int* perm) {
  for (int p = 0; p < nb; ++p) perm[p] = -1;
  int i = 0, end = nb, occ = 0;
  while (i < end) {
    int pos, nxt;
    if (!conf[occ]) {
      pos = i;
      nxt = ++i;
    } else {
      pos = --end;
      nxt = end;
    }
    perm[pos] = occ;
    occ = nxt;
  }
  for (int p = 0; p < nb; ++p)
    if (perm[p] < 0) perm[p] = p;  // positions the loop never reached
}

int main() {
  srand(11);
  constexpr int L = 1, VS = 128;  // one lane holds all 128 vertex slots
  const int j = 0;
  long bad = 0, n = 0;
  for (int trial = 0; trial < 40000; ++trial) {
    const int nb_v = 2 + rand() % 127;
    const int dens = rand() % 100;
    int conf[128] = {0}, nr = 0;
    uint64_t conf_lo = 0, conf_hi = 0;
    for (int v = 0; v < nb_v; ++v) {
      conf[v] = (rand() % 100) < dens;
      nr += conf[v];
      if (conf[v]) {
        if (v < 64) conf_lo |= 1ull << v; else conf_hi |= 1ull << (v - 64);
      }
    }
    if (nr == 0 || nr == nb_v) continue;  // the kernel never moves these
    uint32_t s_ver_row[128];
    for (int v = 0; v < 128; ++v) s_ver_row[v] = (uint32_t)v;
    uint32_t val[VS];
    int dst[VS];
    {
BODY
    }
    int perm[128], got[128];
    serial(nb_v, conf, perm);
    for (int p = 0; p < nb_v; ++p) got[p] = -1;
    bool ok = true;
    for (int q = 0; q < VS && ok; ++q)
      if (dst[q] >= 0) {
        if (dst[q] >= nb_v || got[dst[q]] != -1) ok = false;
        else got[dst[q]] = (int)val[q];
      }
    for (int p = 0; p < nb_v && ok; ++p) ok = got[p] == perm[p];
    bad += !ok;
    ++n;
  }
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


def _extract():
    s = open(SRC).read()
    a = s.index("// 128-bit vertex masks as (lo, hi) words")
    b = s.index("template <int L>\n__device__ __forceinline__ uint64_t "
                "group_bits")
    helpers = s[a:b].replace("__device__ __forceinline__", "static inline")
    c = s.index("      const int nk = nb_v - nr;")
    d = s.index("      wsync();\n#pragma unroll\n      for (int q = 0; q < VS; "
                "++q)\n        if (dst[q] >= 0)", c)
    body = s[c:d]
    # the kernel declares val/dst in the block; the harness reads them after
    body = body.replace("      uint32_t val[VS];\n      int dst[VS];\n", "")
    assert "uint32_t val[VS]" not in body
    body = body.replace("s_ver[g][v]", "s_ver_row[v]")
    return helpers, body


def test_closed_form_vertex_move_matches_serial_swap_loop(tmp_path):
    helpers, body = _extract()
    prog = HEAD + helpers + MAIN.replace("BODY", body)
    cpp = tmp_path / "swap_check.cpp"
    exe = tmp_path / "swap_check"
    cpp.write_text(prog)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(cpp)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("harness does not compile:\n" + r.stderr[-2000:])
    out = subprocess.run([str(exe)], capture_output=True, text=True,
                         check=True).stdout.split()
    n, bad = int(out[0]), int(out[1])
    assert n > 30000 and bad == 0, (n, bad)
