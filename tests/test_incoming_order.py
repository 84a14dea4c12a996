"""The incoming-contact walk of the force kernels (next_incoming in
pacific_amd/csrc/gdem_sweep.h) visits every list entry exactly once, in
(partner slot, code, position) order, also when an entry appears twice, and
then reports the end: it never yields an index outside the list; with a mask
it visits the masked entries only, in the same order.  Built for the
host with hipcc (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pacific_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

DRIVER = r"""
#include <cstdio>
#include "gdem_sweep.h"
int main() {
  const int n = 7;
  int2 il[n] = {make_int2(7, 5), make_int2(3, 1), make_int2(7, 5), make_int2(3, 9),
                make_int2(12, 0), make_int2(7, 2), make_int2(0, 127)};
  long long last = -1, key = 0;
  int m = 0;
  for (int e = 0; e < n + 2; ++e) {
    const int f = gdem::next_incoming(il, n, 1, last, key);
    if (f < 0) break;
    if (f >= n) { printf("out of range %d\n", f); return 1; }
    printf("%d ", f);
    last = key;
    ++m;
  }
  printf("| %d ", m);
  // only the entries of a mask (here 0, 3, 5, 6), in the same order
  const unsigned mask = (1u << 0) | (1u << 3) | (1u << 5) | (1u << 6);
  last = -1;
  m = 0;
  for (int e = 0; e < n + 2; ++e) {
    const int f = gdem::next_incoming(il, n, 1, last, key, mask);
    if (f < 0) break;
    if (f >= n || !((mask >> f) & 1u)) { printf("bad %d\n", f); return 1; }
    printf("%d ", f);
    last = key;
    ++m;
  }
  printf("| %d\n", m);
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_incoming_walk_visits_each_entry_once(tmp_path):
    src = tmp_path / "walk.hip"
    src.write_text(DRIVER)
    exe = tmp_path / "walk"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", CSRC,
                           "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)]).decode().split()
    # order: (0,127)#6, (3,1)#1, (3,9)#3, (7,2)#5, (7,5)#0, (7,5)#2, (12,0)#4
    assert out == ["6", "1", "3", "5", "0", "2", "4", "|", "7",
                   "6", "3", "5", "0", "|", "4"]
