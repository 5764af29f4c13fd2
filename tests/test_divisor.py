"""The cell-index split's division by invariant integers (make_divisor/divide in lbm_step.hpp),
compiled for the host and checked against exact integer division: every divisor up to 4096,
powers of two and their neighbours, large divisors up to 2^31 - 1, and numerators at 0, the
2^31/2^32 edges, around multiples of the divisor and pseudo-random."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROGRAM = r'''
#include <cstdio>
#include <cstdint>
#include <vector>
#include "lbm_step.hpp"
int main() {
  std::vector<uint32_t> ds;
  for (uint32_t d = 1; d <= 4096; ++d) ds.push_back(d);
  for (int k = 12; k < 31; ++k) {
    ds.push_back((1u << k) - 1); ds.push_back(1u << k); ds.push_back((1u << k) + 1);
  }
  ds.push_back(0x7FFFFFFFu); ds.push_back(0x7FFFFFFEu); ds.push_back(16776960u);
  uint64_t state = 0x9E3779B97F4A7C15ull, checked = 0, bad = 0;
  for (uint32_t d : ds) {
    const tlbm::Divisor dv = tlbm::make_divisor(d);
    std::vector<uint32_t> ns = {0u, 1u, 2u, 0x7FFFFFFEu, 0x7FFFFFFFu, 0x80000000u, 0xFFFFFFFEu,
                                0xFFFFFFFFu};
    for (int i = 0; i < 2000; ++i) {
      state = state * 6364136223846793005ull + 1442695040888963407ull;
      const uint32_t r = static_cast<uint32_t>(state >> 32);
      ns.push_back(r);
      const uint64_t m = (static_cast<uint64_t>(r) % (0xFFFFFFFFull / d + 1)) * d;
      for (int64_t o = -2; o <= 2; ++o) {
        const int64_t v = static_cast<int64_t>(m) + o;
        if (v >= 0 && v <= 0xFFFFFFFFll) ns.push_back(static_cast<uint32_t>(v));
      }
    }
    for (uint32_t n : ns) {
      ++checked;
      if (tlbm::divide(n, dv) != n / d) ++bad;
    }
  }
  std::printf("%llu %llu\n", static_cast<unsigned long long>(checked),
              static_cast<unsigned long long>(bad));
  return 0;
}
'''


def test_invariant_integer_division_is_exact(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    src = tmp_path / "divisor_check.hip"
    src.write_text(PROGRAM)
    exe = tmp_path / "divisor_check"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "tack-lbm_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120)
    checked, bad = (int(v) for v in out.stdout.split())
    assert checked > 10_000_000 and bad == 0
