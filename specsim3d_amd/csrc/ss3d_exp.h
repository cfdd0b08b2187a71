// ss3d_exp.h -- FP64 exp with the results of the reference build's libm, bit for bit.
//
// The reference evaluates its exponential and Gaussian covariances with the C library's exp
// (src/specsim3d.f90:153-178).  With glibc (2.28 and later; the image ships 2.35) that is the
// table-driven method of ARM's optimized-routines (Szabolcs Nagy, 2018, MIT licence), in the build
// x86-64 selects on CPUs with FMA.  It is not correctly rounded (about 1 result in 1300 differs
// from the correctly rounded one), so a covariance grid equal to the reference's needs that method
// itself: this file restates it, with the same constants, the same fused multiply-adds and the
// same special cases, so that device results equal glibc's (checked against glibc on 4e7
// arguments over every range, subnormal results included, by tests/test_libm_exp.py).
//
// Method: x = k ln2/128 + r, |r| <= ln2/256; 2^(k/128) = H_j (1 + t_j) from a 128-entry table
// (j = k mod 128, exponent k/128 added to H's bits); exp(r) - 1 ~= r + C2 r^2 + ... + C5 r^5;
// exp(x) = s + s (t_j + exp(r) - 1), s the scaled H_j.  |x| >= 512 goes through a rescaled path
// that rounds once into the subnormal range.
//
// Compiled with -ffp-contract=off: every fused multiply-add below is written as fma().
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SS3D_EXP_FN __device__ __forceinline__
#define SS3D_EXP_TAB __constant__
#else  // host build of the same header (tests/test_libm_exp.py compares it with glibc)
#include <math.h>
#define SS3D_EXP_FN static inline
#define SS3D_EXP_TAB static
#endif

namespace ss3d {
namespace libm {

SS3D_EXP_TAB const uint64_t kExpTab[256] = {
#include "ss3d_exp_table.inc"
};

SS3D_EXP_FN uint64_t bits_of(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
SS3D_EXP_FN double from_bits(uint64_t u) {
    double x;
    memcpy(&x, &u, 8);
    return x;
}

// |x| >= 512 (k the table index scaled by 128, kbits its integer form): the scale's exponent may
// leave the normal range, so it is built 1022 (k < 0) or 1009 (k > 0) binades in and applied last.
// Results below 2^-1022 are first rounded to their final precision by adding 1 (hi + lo), so that
// the scaling into the subnormal range rounds only once.  The k > 0 branch fuses s + s tmp, the
// k < 0 branch fuses nothing (as in the library's build; both found against its results).
SS3D_EXP_FN double exp_rescaled(double tmp, uint64_t sbits, uint64_t kbits) {
    if ((kbits & 0x80000000u) == 0) {
        const double s = from_bits(sbits - (1009ull << 52));
        return 0x1p1009 * fma(s, tmp, s);
    }
    const double s = from_bits(sbits + (1022ull << 52));
    double y = s + s * tmp;
    if (y < 1.0) {
        double lo = s - y + s * tmp;
        const double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) y = 0.0;  // no negative zero
    }
    return 0x1p-1022 * y;
}

SS3D_EXP_FN double exp(double x) {
    const double kInvLn2N = 0x1.71547652b82fep0 * 128;  // 128/ln2
    const double kNegLn2HiN = -0x1.62e42fefa0000p-8;   // -ln2/128, high part (exact k*hi)
    const double kNegLn2LoN = -0x1.cf79abc9e3b3ap-47;  // -ln2/128, low part
    const double kShift = 0x1.8p52;                     // round-to-integer shift
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3;
    const double C4 = 0x1.55555cf172b91p-5, C5 = 0x1.1111167a4d017p-7;
    // biased exponent of |x|: tiny (< 2^-54), huge (>= 1024), large (>= 512) or ordinary
    uint32_t etop = (uint32_t)(bits_of(x) >> 52) & 0x7ff;
    const uint32_t e_tiny = 0x3c9, e_512 = 0x408, e_1024 = 0x409, e_inf = 0x7ff;
    if (etop - e_tiny >= e_512 - e_tiny) {
        if (etop < e_tiny) return 1.0 + x;
        if (etop >= e_1024) {
            if (bits_of(x) == 0xfff0000000000000ull) return 0.0;  // -inf
            if (etop >= e_inf) return 1.0 + x;                    // nan, +inf
            return (bits_of(x) >> 63) ? 0.0 : from_bits(0x7ff0000000000000ull);
        }
        etop = 0;  // 512 <= |x| < 1024: rescaled path below
    }
    // k = nearest integer to 128 x / ln2 (the shift rounds it into the low bits)
    const double z = kInvLn2N * x;
    double kd = z + kShift;
    const uint64_t kbits = bits_of(kd);
    kd -= kShift;
    const double r = fma(kd, kNegLn2LoN, fma(kd, kNegLn2HiN, x));
    const uint32_t j = (uint32_t)(kbits % 128);
    const double tail = from_bits(kExpTab[2 * j]);
    const uint64_t sbits = kExpTab[2 * j + 1] + (kbits << 45);  // exponent += k/128
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
    if (etop == 0) return exp_rescaled(tmp, sbits, kbits);
    const double s = from_bits(sbits);
    return fma(s, tmp, s);
}

}  // namespace libm
}  // namespace ss3d
