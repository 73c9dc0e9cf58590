// The exact reduction of an INT32 sum mod m that the INT8 CRT GEMM's
// epilogue stores, shared with the host-only test that checks it
// (tests/native/crt_mod_exact.cpp).  Plain C++ apart from the HIP
// qualifiers, so a host compiler can include it on its own.
#pragma once

#include <math.h>

#if defined(__HIP__)
#define QS_HOST_DEVICE __host__ __device__
#else
#define QS_HOST_DEVICE
#endif

namespace qs {

// x mod m in [0, m), for |x| < 2^31 and 2 <= m <= 256, with im = 1 / m
// rounded to double.  im has a relative error of at most 2^-53, |x / m| <
// 2^30, and the FMA rounds once, so fma(x, im, 2^-20) is within 2^-22 of
// x / m + 2^-20.  When x / m is an integer q, that lies in (q, q + 1); when
// it is not, its fraction is at least 1 / m >= 2^-8, so the 2^-20 offset
// never carries it past the next integer.  The floor is therefore
// floor(x / m), and x - q m, an integer in [0, m), comes out of the second
// FMA exactly.
QS_HOST_DEVICE inline unsigned int crt_mod_exact(int x, double m, double im) {
  const double xd = (double)x;
  const double q = floor(__builtin_fma(xd, im, 0x1p-20));
  return (unsigned int)__builtin_fma(-q, m, xd);
}

}  // namespace qs
