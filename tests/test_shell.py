"""The ghost fill's thread -> cell enumeration (particleclem_amd/csrc/
pc_shell.h) on the CPU: for boxes of several shapes and positions, the
first ghost_shell_count(bx) indices map one to one onto the cells of
grow(bx,4) outside bx, and every later index is rejected.  Compiled with the
host compiler from the header itself (it has no HIP dependency)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdio>
#include <set>
#include <tuple>
#include "pc_shell.h"
int main() {
  const int boxes[][6] = {{0, 0, 0, 31, 31, 31}, {5, -3, 7, 12, 9, 7},
                          {0, 0, 0, 0, 0, 0},    {-4, 2, 1, 20, 2, 30},
                          {3, 3, 3, 10, 40, 5},  {-9, -9, -9, -2, -2, -2}};
  int bad = 0;
  for (const auto& c : boxes) {
    const int L[3] = {c[0], c[1], c[2]}, H[3] = {c[3], c[4], c[5]};
    const long long n = ghost_shell_count(L, H);
    long long expect = 1, inner = 1;
    for (int d = 0; d < 3; ++d) {
      expect *= H[d] - L[d] + 9;
      inner *= H[d] - L[d] + 1;
    }
    if (n != expect - inner) ++bad;
    std::set<std::tuple<int, int, int>> seen;
    for (long long t = 0; t < n + 1000; ++t) {
      int i, j, k;
      const bool ok = ghost_shell_cell(L, H, t, i, j, k);
      if (t >= n) {
        if (ok) ++bad;
        continue;
      }
      if (!ok) { ++bad; continue; }
      const bool valid = i >= L[0] && i <= H[0] && j >= L[1] && j <= H[1] &&
                         k >= L[2] && k <= H[2];
      const bool grown = i >= L[0] - 4 && i <= H[0] + 4 && j >= L[1] - 4 &&
                         j <= H[1] + 4 && k >= L[2] - 4 && k <= H[2] + 4;
      if (valid || !grown) ++bad;
      seen.insert(std::make_tuple(i, j, k));
    }
    if ((long long)seen.size() != n) ++bad;
  }
  std::printf("bad=%d\n", bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_ghost_shell_enumeration_is_a_bijection(tmp_path):
    src = tmp_path / "shell.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "shell"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I",
                    os.path.join(ROOT, "particleclem_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "bad=0" in out.stdout
