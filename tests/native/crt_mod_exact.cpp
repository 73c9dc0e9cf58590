// Host-only check of crt_mod_exact (csrc/qs_crt.h), the INT8 CRT GEMM
// epilogue's reduction of an INT32 sum mod m, against the integer remainder
// for every modulus 2..256: at x = q m + d (d in -2..2, the values next to a
// multiple of m, where a quotient off by one would show) for the 8192
// quotients nearest zero, the 8192 at each end of the range |q m| < 2^31 and
// 2^14 random ones; at every |x| < 2^16; and at 2^22 random (x, m) pairs.
// (Run once with every quotient of every modulus, 1.1e11 values: no failure.)
// Built and run by tests/test_tiles_host.py.
#include <cstdint>
#include <cstdio>

#include "../../quasistrided.jl_amd/csrc/qs_crt.h"

static long long checked = 0, failures = 0;

static void check(long long x, int m) {
  if (x <= -2147483648LL || x >= 2147483648LL) return;
  const unsigned got = qs::crt_mod_exact((int)x, (double)m, 1.0 / (double)m);
  long long want = x % m;
  if (want < 0) want += m;
  ++checked;
  if ((long long)got != want && failures++ < 10) std::printf("FAIL m=%d x=%lld got %u want %lld\n", m, x, got, want);
}

int main() {
  uint64_t s = 0x9e3779b97f4a7c15ull;
  for (int m = 2; m <= 256; ++m) {
    const long long qmax = (1LL << 31) / m + 1;
    auto near = [&](long long q) {
      for (int d = -2; d <= 2; ++d) check(q * m + d, m);
    };
    for (long long q = -4096; q < 4096; ++q) near(q);
    for (long long q = 0; q < 8192; ++q) {
      near(qmax - q);
      near(-qmax + q);
    }
    for (int i = 0; i < (1 << 14); ++i) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      near((long long)(s % (uint64_t)(2 * qmax + 1)) - qmax);
    }
    for (long long x = -(1LL << 16); x < (1LL << 16); ++x) check(x, m);
  }
  for (int i = 0; i < (1 << 22); ++i) {   // random moduli too
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    check((long long)(int32_t)(uint32_t)(s >> 20), 2 + (int)(s % 255));
  }
  std::printf("checked %lld failures %lld\n", checked, failures);
  return failures != 0;
}
