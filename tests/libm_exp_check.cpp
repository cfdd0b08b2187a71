// Host build of specsim3d_amd/csrc/ss3d_exp.h compared with the C library's exp (glibc), bit for
// bit.  Usage: libm_exp_check n lo hi  -> prints "n <n> mismatches <m>" for n arguments drawn
// uniformly from [lo, hi] (xorshift64), then the special values.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ss3d_exp.h"

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const long n = atol(argv[1]);
    const double lo = atof(argv[2]), hi = atof(argv[3]);
    long bad = 0;
    uint64_t s = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
    auto differs = [](double x) {
        const double a = std::exp(x), b = ss3d::libm::exp(x);
        return ss3d::libm::bits_of(a) != ss3d::libm::bits_of(b) && !(std::isnan(a) && std::isnan(b));
    };
    for (long i = 0; i < n; ++i) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        const double x = lo + (hi - lo) * ((double)(s >> 11) * 0x1p-53);
        if (differs(x)) {
            if (bad < 3) printf("x %a glibc %a here %a\n", x, std::exp(x), ss3d::libm::exp(x));
            ++bad;
        }
    }
    const double special[] = {0.0, -0.0, 1e-300, -1e-300, 0x1p-54, -0x1p-54, 512.0, -512.0, 1024.0, -1024.0,
                              -745.1332191019412, -745.1332191019411, -708.3964185322641, 709.782712893384,
                              709.7827128933841, -INFINITY, INFINITY, NAN};
    long sbad = 0;
    for (double x : special) sbad += differs(x);
    printf("n %ld mismatches %ld special %ld\n", n, bad, sbad);
    return 0;
}
