"""The device's expf (turbo-cat_amd/csrc/tc_expf.h) against the host's glibc
expf, the function every std::exp(float) of the reference resolves to
(loss.hpp:99-101, loss.cpp:157-162, :213-215, :273-285).

The header compiles under g++ as well as hipcc; here it is built into a small
host program that compares it bit for bit with glibc on a strided sample of
all 2^32 float bit patterns (every 61st), the two inputs where glibc's FMA
build and an unfused evaluation round differently, and the special values
(+-0, +-inf, NaN, the overflow and underflow bounds).  Every float was
compared once when the restatement was written: no difference.  The device's
own results are compared with the oracle's (glibc) gradients in
tests/test_gpu_primitives.py."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "turbo-cat_amd", "csrc", "tc_expf.h")

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include "tc_expf.h"
static int check(uint32_t u) {
  float x; memcpy(&x, &u, 4);
  const float a = expf(x), c = tc::ref_expf(x);
  uint32_t ua, uc; memcpy(&ua, &a, 4); memcpy(&uc, &c, 4);
  if (ua == uc || (a != a && c != c)) return 0;
  printf("0x%08x %a %a\n", u, a, c);
  return 1;
}
int main() {
  unsigned long long n = 0, bad = 0;
  for (unsigned long long b = 0; b < (1ull << 32); b += 61) { bad += check((uint32_t)b); ++n; }
  const uint32_t edge[] = {0x4202422fu, 0xc27c65d9u, 0x00000000u, 0x80000000u, 0x7f800000u,
                           0xff800000u, 0x7fc00000u, 0x42b00000u, 0xc2b00000u, 0x42b17217u,
                           0x42b17218u, 0xc2cff1b4u, 0xc2cff1b5u, 0x41700000u, 0xc1700000u};
  for (uint32_t u : edge) { bad += check(u); ++n; }
  printf("checked %llu mismatches %llu\n", n, bad);
  return bad != 0;
}
"""


def _has_fma():
    try:
        with open("/proc/cpuinfo") as fh:
            return " fma " in fh.read()
    except OSError:
        return False


@pytest.mark.skipif(not _has_fma(), reason="glibc runs its non-FMA expf on a CPU without FMA")
def test_ref_expf_equals_glibc(tmp_path):
    src = tmp_path / "expf_check.cpp"
    exe = tmp_path / "expf_check"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    os.path.dirname(HEADER), str(src), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:]
    assert "mismatches 0" in out.stdout
