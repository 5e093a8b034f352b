// Elementary functions that must give the same bits on the host (oracle, CPU
// baseline) and on gfx950.  Only IEEE-exact operations are used (+, -, *, /,
// comparisons, bit manipulation); with floating-point contraction disabled on
// both compilers the host and device results are identical.
//
// vx_log: natural logarithm by the classical reduction
//   x = 2^k * m, m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f / (2 + f),
//   log(1+f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R)),
//   R = 2 (s^2/3 + s^4/5 + ... + s^20/21)   (Taylor series of atanh),
//   log(x) = k ln2_hi + (f - (hfsq - (s (hfsq + R) + k ln2_lo))),
// with ln2 split so that k * ln2_hi is exact.  Accuracy is within 1 ulp of the
// correctly rounded result for positive normal arguments (checked against the
// C library in the tests).  The reference evaluates these logarithms with
// MKL's vdLn (Rectangle.cpp:227-229).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define VX_HD __host__ __device__ inline
#else
#define VX_HD inline
#endif

VX_HD double vx_bits_to_double(uint64_t b) {
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}
VX_HD uint64_t vx_double_to_bits(double d) {
    uint64_t b;
    memcpy(&b, &d, sizeof b);
    return b;
}

VX_HD double vx_log(double x) {
    const double ln2_hi = 0.6931471803691238;      // 0x3fe62e42fee00000
    const double ln2_lo = 1.9082149292705877e-10;  // ln 2 - ln2_hi
    uint64_t b = vx_double_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff);
    int kbias = 1023;
    if (!(x > 0.0) || e == 0x7ff) {
        // zero, negative, inf, nan: not produced by the method; give the
        // IEEE-style answers without claiming accuracy.
        if (x == 0.0) return -vx_bits_to_double(0x7ff0000000000000ULL);
        if (x < 0.0 || x != x) return vx_bits_to_double(0x7ff8000000000000ULL);
        return x;
    }
    if (e == 0) {  // subnormal: scale by 2^54 (exact) and correct the exponent
        b = vx_double_to_bits(x * 18014398509481984.0);
        e = (int)((b >> 52) & 0x7ff);
        kbias = 1023 + 54;
    }
    int k = e - kbias;
    // mantissa in [1, 2)
    double m = vx_bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) {  // move to [sqrt(1/2), sqrt(2))
        m = m * 0.5;
        k = k + 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = z * (0.6666666666666666 +
               z * (0.4 +
               z * (0.2857142857142857 +
               z * (0.2222222222222222 +
               z * (0.18181818181818182 +
               z * (0.15384615384615385 +
               z * (0.13333333333333333 +
               z * (0.11764705882352941 +
               z * (0.10526315789473684 +
               z * 0.09523809523809523)))))))));
    double hfsq = 0.5 * f * f;
    double dk = (double)k;
    return dk * ln2_hi + (f - (hfsq - (s * (hfsq + R) + dk * ln2_lo)));
}
