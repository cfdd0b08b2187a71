// ss3d_fast.h -- FP32 realization passes (the throughput path).
//
// Unlike the FP64 passes (ss3d_pass.h), which repeat the reference's radix-2 butterflies for
// parity, these passes are free to pick the algorithm: a mixed-radix Stockham inverse FFT with
// radices up to 32, every length and tile a compile-time constant so that all addressing folds
// into immediates.
//   - pass 0 reads its operands straight from HBM (or generates them: Hermitian coefficient
//     fill) into registers, the last pass writes straight from registers; LDS is used only for
//     the exchange between consecutive radix passes;
//   - pass 0 has compile-time twiddles only; a later pass applies its table twiddles in tangent
//     form, as radix-2 stages of six fused multiply-adds per butterfly (strided passes; the
//     table is built in LDS from FP64 values rounded to FP32), or multiplies by W_n table
//     entries (C2R passes, axes longer than 2^11);
//   - the x pass is a real-output (C2R) transform computed as an n/2-point complex FFT of
//     Z[k] = E[k] + i O[k], E = X[k] + conj X[n/2-k], O = (X[k] - conj X[n/2-k]) W_n^k, with
//     x[2m] = Re z[m], x[2m+1] = Im z[m].  The imaginary parts of X[0] and X[n/2] are dropped,
//     which is exactly what taking the real part of a complex-to-complex transform does;
//   - z-run layout (z and y axes of 2^9): the z fill holds whole column pairs per wave, so its
//     exchanges need no block barrier, and writes runs of 8 z per kx pair; the y pass reads and
//     writes those runs through its strides, and a C2R variant moves them to rows in LDS.  Every
//     column and row is transformed by the same arithmetic as in the row layouts.
//
// Stockham pass j (radix R = R_j, Ns = R_0 ... R_{j-1}) on butterfly b in [0, n/R):
//     v[r] = in[b + r n/R] * W_{Ns R}^{r (b mod Ns)},  v = DFT_R(v),
//     out[(b / Ns) Ns R + (b mod Ns) + r Ns] = v[r]
// (natural order in and out; inverse transform, W_m = exp(+2 pi i/m), no scaling).
#pragma once

#include <utility>

#include "ss3d_internal.h"

namespace ss3d {
namespace fast {

using f2 = float2;
using f2v = float __attribute__((ext_vector_type(2)));  // for non-temporal 8-byte accesses

__device__ __forceinline__ f2 mk(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ f2 add(f2 a, f2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ f2 sub(f2 a, f2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ f2 mul(f2 w, f2 x) { return mk(w.x * x.x - w.y * x.y, w.x * x.y + w.y * x.x); }

// ------------------------------------------------------------ compile-time plan
// A length 2^L is split into npass(L) radix passes of at most 2^RM each, as evenly as possible,
// largest first.  RM (5: radix <= 32, 4: <= 16, 3: <= 8) trades LDS exchanges for registers.
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
constexpr int npass(int L, int RM = 5) { return cdiv(L, RM); }
constexpr int rlog(int L, int j, int RM = 5) {
    int rem = L, m = npass(L, RM);
    for (int i = 0; i < j; ++i) rem -= cdiv(rem, m - i);
    return cdiv(rem, m - j);
}
constexpr int nslog(int L, int j, int RM = 5) {
    int s = 0;
    for (int i = 0; i < j; ++i) s += rlog(L, i, RM);
    return s;
}
constexpr int eptlog(int L, int RM = 5) { return L ? rlog(L, 0, RM) : 0; }

// cos / sin of 2 pi j / 32, rounded to float
__device__ constexpr float kC32[32] = {
    1.0f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
    0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f, 0.19509032201612826785f,
    0.0f, -0.19509032201612826785f, -0.38268343236508977173f, -0.55557023301960222474f,
    -0.70710678118654752440f, -0.83146961230254523708f, -0.92387953251128675613f, -0.98078528040323044913f,
    -1.0f, -0.98078528040323044913f, -0.92387953251128675613f, -0.83146961230254523708f,
    -0.70710678118654752440f, -0.55557023301960222474f, -0.38268343236508977173f, -0.19509032201612826785f,
    0.0f, 0.19509032201612826785f, 0.38268343236508977173f, 0.55557023301960222474f,
    0.70710678118654752440f, 0.83146961230254523708f, 0.92387953251128675613f, 0.98078528040323044913f};

// x * exp(+2 pi i j/32), with the exact forms for multiples of pi/2 and pi/4
template <int J>
__device__ __forceinline__ f2 rot32(f2 x) {
    constexpr int j = J & 31;
    constexpr float s = 0.70710678118654752440f;
    if constexpr (j == 0) return x;
    else if constexpr (j == 8) return mk(-x.y, x.x);
    else if constexpr (j == 16) return mk(-x.x, -x.y);
    else if constexpr (j == 24) return mk(x.y, -x.x);
    else if constexpr (j == 4) return mk(s * (x.x - x.y), s * (x.x + x.y));
    else if constexpr (j == 12) return mk(-s * (x.x + x.y), s * (x.x - x.y));
    else if constexpr (j == 20) return mk(s * (x.y - x.x), -s * (x.x + x.y));
    else if constexpr (j == 28) return mk(s * (x.x + x.y), s * (x.y - x.x));
    else {
        constexpr float c = kC32[j], sn = kC32[(j + 24) & 31];  // sin(t) = cos(t - pi/2)
        return mk(c * x.x - sn * x.y, c * x.y + sn * x.x);
    }
}

constexpr int brev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

// in-place inverse DFT of length 2^LR on v[0 .. 2^LR), radix-2 network with constant twiddles
template <int LR, int LEN = 2>
__device__ __forceinline__ void idft_stages(f2* t) {
    if constexpr (LEN <= (1 << LR)) {
        constexpr int H = LEN / 2;
#pragma unroll
        for (int i = 0; i < (1 << LR); i += LEN) {
            [&]<int... K>(std::integer_sequence<int, K...>) {
                ((void)[&] {
                    f2 u = t[i + K];
                    f2 w = rot32<K * (32 / LEN)>(t[i + K + H]);
                    t[i + K] = add(u, w);
                    t[i + K + H] = sub(u, w);
                }(), ...);
            }(std::make_integer_sequence<int, H>{});
        }
        idft_stages<LR, LEN * 2>(t);
    }
}

template <int LR>
__device__ __forceinline__ void idft(f2* v) {
    if constexpr (LR > 0) {
        f2 t[1 << LR];
        [&]<int... I>(std::integer_sequence<int, I...>) { ((t[brev(I, LR)] = v[I]), ...); }(
            std::make_integer_sequence<int, (1 << LR)>{});
        idft_stages<LR>(t);
#pragma unroll
        for (int i = 0; i < (1 << LR); ++i) v[i] = t[i];
    }
}

// ------------------------------------------------------------ random stream, FP32 coefficients
__device__ __forceinline__ uint32_t pm_mul(uint32_t a, uint32_t b) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    uint32_t s = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
    return s >= kPmM ? s - kPmM : s;
}

// Coefficient from the first draw `st` of its site: igauss=0 amp*exp(i 2 pi u) (self-conjugate:
// +-amp), igauss=1 amp*(g1 + i g2)/sqrt2 (self-conjugate: amp*g1), Box-Muller on (u1, u2);
// conj for the partner of a drawing site.  Branch-free (selects only).  Hardware sin/cos take
// revolutions: v_sin_f32(x) = sin(2 pi x); v_log_f32 is log2.
template <int IG>
__device__ __forceinline__ f2 coef_t(float amps, uint32_t st, bool conj, bool selfc) {
    // amps = amplitude * sqrt(ln 2) (the FP32 half-spectrum table)
    const float inv_m = 4.6566128752457969e-10f;  // 1/(2^31-1)
    float re, im;
    if constexpr (IG == 0) {
        const float amp = amps * 1.2011224087864498f;  // 1/sqrt(ln 2)
        const float u = (float)st * inv_m;
        const float c = __builtin_amdgcn_cosf(u), s = __builtin_amdgcn_sinf(u);
        const float sgn = (st <= 1073741823u) ? amp : -amp;
        re = selfc ? sgn : amp * c;
        im = selfc ? 0.0f : amp * s;
    } else {
        const float u1 = (float)st * inv_m;
        // u2 = (st a mod M)/M = frac(st a/M) (to 4e-12): only its sin/cos (period 1) are used
        const float u2 = (float)__builtin_amdgcn_fract((double)st * (16807.0 / 2147483647.0));
        // amp sqrt(-2 ln u1) / sqrt(2) = amps sqrt(-log2 u1); self-conjugate sites: times sqrt(2)
        const float a = amps * __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1)) * (selfc ? 1.41421356237309505f : 1.0f);
        re = a * __builtin_amdgcn_cosf(u2);
        im = selfc ? 0.0f : a * __builtin_amdgcn_sinf(u2);
    }
    return mk(re, conj ? -im : im);
}
// a * b mod 2^31-1 without the final reduction: the state s or s + (2^31-1), below 2^32 (a, b
// are nonzero states, so the product is no multiple of 2^31-1)
__device__ __forceinline__ uint32_t pm_mul_lazy(uint32_t a, uint32_t b) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    return (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
}

// The same for a site that is not self-conjugate, with the conjugation carried by the sign of
// the amplitude (negative: conjugate): the real part takes |amplitude|, the imaginary part the
// signed one -- the values of coef_t(|amps|, st, amps < 0, false), without a select.  `st` is
// pm_mul_lazy's state (s or s + M): only fractions of st/M are used, and a whole revolution more
// does not change a sine or cosine:
//   igauss=0: the phase st/M in [0, 2) revolutions;
//   igauss=1: u1 = frac(st/M) = s/M (absolute error < 2^-52 before its rounding to FP32, so the
//             logarithm's error stays below 5e-7 even at s = 1) and u2 = frac(st a/M).
template <int IG>
__device__ __forceinline__ f2 coef_signed(float amps, uint32_t st) {
    if constexpr (IG == 0) {
        const float inv_m = 4.6566128752457969e-10f;  // 1/(2^31-1)
        const float amp = amps * 1.2011224087864498f;
        const float u = (float)st * inv_m;
        return mk(fabsf(amp) * __builtin_amdgcn_cosf(u), amp * __builtin_amdgcn_sinf(u));
    } else {
        const double d = (double)st;
        const float u1 = (float)__builtin_amdgcn_fract(d * (1.0 / 2147483647.0));
        const float u2 = (float)__builtin_amdgcn_fract(d * (16807.0 / 2147483647.0));
        const float a = amps * __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1));
        return mk(fabsf(a) * __builtin_amdgcn_cosf(u2), a * __builtin_amdgcn_sinf(u2));
    }
}
__device__ __forceinline__ f2 coef(float amp, uint32_t st, int igauss, bool conj, bool selfc) {
    return igauss ? coef_t<1>(amp, st, conj, selfc) : coef_t<0>(amp, st, conj, selfc);
}
// coef_signed's value as its parts: the coefficient is (|R| c, R s)
template <int IG>
__device__ __forceinline__ void coef_parts(float amps, uint32_t st, float& R, float& c, float& s) {
    if constexpr (IG == 0) {
        const float inv_m = 4.6566128752457969e-10f;  // 1/(2^31-1)
        const float u = (float)st * inv_m;
        R = amps * 1.2011224087864498f;
        c = __builtin_amdgcn_cosf(u);
        s = __builtin_amdgcn_sinf(u);
    } else {
        const double d = (double)st;
        const float u1 = (float)__builtin_amdgcn_fract(d * (1.0 / 2147483647.0));
        const float u2 = (float)__builtin_amdgcn_fract(d * (16807.0 / 2147483647.0));
        R = amps * __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1));
        c = __builtin_amdgcn_cosf(u2);
        s = __builtin_amdgcn_sinf(u2);
    }
}
// a radix-2 stage's sum and difference of the coefficients (|Ra| ca, Ra sa) and (|Rb| cb, Rb sb), the
// second one's products fused into the sum and difference
__device__ __forceinline__ void stage1_parts(f2& lo, f2& hi, float Ra, float ca, float sa, float Rb, float cb, float sb) {
    const float ax = fabsf(Ra) * ca, ay = Ra * sa, bx = fabsf(Rb);
    lo = mk(__builtin_fmaf(bx, cb, ax), __builtin_fmaf(Rb, sb, ay));
    hi = mk(__builtin_fmaf(-bx, cb, ax), __builtin_fmaf(-Rb, sb, ay));
}

// Fill of a half-spectrum column kx (kx <= npx/2), following the walk of
// src/specsim3d.f90:1305-1333: the site of (kx, ky, kz) is the point itself when its row
// R = ky + npy kz precedes the partner row Rp (or R == Rp and kx < npx - kx), the partner
// otherwise (conjugated); R == Rp and kx == npx - kx is self-conjugate.  The realization base
// is folded into the x factors of the column and of its partner column.
struct Column {
    int kx;
    bool xgt, xeq;
    uint32_t town, tpar;
};
__device__ __forceinline__ Column column_of(const PassArgs& a, uint32_t rb, int kx) {
    Column c;
    const int kxp = (a.npx - kx) & (a.npx - 1);
    c.kx = kx;
    c.xgt = kx > kxp;
    c.xeq = kx == kxp;
    c.town = pm_mul(rb, __ldg(a.txp + kx));
    c.tpar = pm_mul(rb, __ldg(a.txp + kxp));
    return c;
}
// Row part of the classification (shared by every column): the point's row, the drawing row
// of its site, and how the row compares with its partner row.
struct Site {
    uint32_t row, site;
    bool gt, eq;
};
__device__ __forceinline__ Site site_of(const PassArgs& a, int ky, int kz) {
    const int kyp = (a.npy - ky) & (a.npy - 1), kzp = (a.npz - kz) & (a.npz - 1);
    const uint32_t row = ky + a.npy * kz, rowp = kyp + a.npy * kzp;
    Site s;
    s.row = row;
    s.gt = row > rowp;
    s.eq = row == rowp;
    s.site = s.gt ? rowp : row;
    return s;
}
__device__ __forceinline__ bool site_conj(const Site& s, const Column& c) { return s.gt || (s.eq && c.xgt); }
__device__ __forceinline__ bool site_self(const Site& s, const Column& c) { return s.eq && c.xeq; }

// one coefficient (used where a single value is needed: C2R 1D fill, one-point x axis)
__device__ __forceinline__ f2 fill_at(const PassArgs& a, const Column& c, int ky, int kz) {
    const Site s = site_of(a, ky, kz);
    const bool cj = site_conj(s, c);
    const uint32_t st = pm_mul(__ldg(a.rowf + s.site), cj ? c.tpar : c.town);
    return coef(__ldg(a.amp32h + c.kx + a.hxa * s.row), st, a.igauss, cj, site_self(s, c));
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

extern __shared__ __attribute__((aligned(16))) unsigned char fast_smem[];

// ------------------------------------------------------------ one Stockham pass in registers
// Thread owns butterflies b = g + G*i, i < EPT/R; v[i*R + r] is element r of butterfly i.
template <int L, int J, int G, int RM, typename TW>
__device__ __forceinline__ void radix_pass(f2* v, int g, TW tw) {
    constexpr int LR = rlog(L, J, RM), R = 1 << LR, NS = 1 << nslog(L, J, RM);
    constexpr int NB = (1 << eptlog(L, RM)) / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if constexpr (NS > 1) {
            const int b = g + G * i;
            const int k = b & (NS - 1);
            // W_{NS R}^{r k} = W_n^{r k n/(NS R)}
            constexpr int SH = L - nslog(L, J, RM) - LR;
#pragma unroll
            for (int r = 1; r < R; ++r) v[i * R + r] = mul(tw((r * k) << SH), v[i * R + r]);
        }
        idft<LR>(v + i * R);
    }
}

// ------------------------------------------------------------ twiddled passes in tangent form
// A Stockham pass J > 0 of radix R = 2^LR is LR radix-2 Stockham stages done in registers: stage s
// pairs registers j and j + R/2^(s+1) of a butterfly with the twiddle W_{2^(s+1) NS}^(k + m NS),
// k = b mod NS, m = brev_s(j >> (LR - s)) (s-bit reversal); the sum stays in j, the difference goes
// to j + R/2^(s+1), so register j ends holding output brev_LR(j).  A twiddle w = c (1 + i t),
// t = tan(arg w), makes each stage butterfly six fused multiply-adds,
//     b' = (b.x - t b.y, b.y + t b.x),  a + w b = a + c b',  a - w b = a - c b',
// against the table twiddle's complex product and the DFT_R's own constant rotations (72 against 80
// operations per radix-8 butterfly).  cos(arg w) is never exactly 0 in the table (the FP32 values
// of cos(pi/2) and cos(3 pi/2) are ~1e-16), and a large t is harmless: c b' then equals the product
// to the table's own rounding.
//
// The table, in LDS: pass J from entry tan_base(J) on, one row of R - 1 float2 (c, t) per k < NS,
// entry 2^s - 1 + m of a row for stage s: a butterfly reads its row from one address with immediate
// offsets.  The passes J > 0 hold N - R_0 entries in all, no more than the W_n table.
template <int L, int RM>
constexpr int tan_base(int J) {
    int e = 0;
    for (int j = 1; j < J; ++j) e += (1 << nslog(L, j, RM)) * ((1 << rlog(L, j, RM)) - 1);
    return e;
}
template <int L, int RM>
constexpr int tan_entries() { return tan_base<L, RM>(npass(L, RM)); }

// entries of the tangent table from the W_n table (wn: W_n^m, m < n, FP32), by the block's threads
template <int L, int RM, int NT>
__device__ __forceinline__ void tan_table_build(f2* tt, const f2* __restrict__ wn, const int tid) {
    [&]<int... J>(std::integer_sequence<int, J...>) {
        (
            [&] {
                constexpr int Jc = J + 1, LR = rlog(L, Jc, RM), R = 1 << LR, NSL = nslog(L, Jc, RM);
                constexpr int E = (1 << NSL) * (R - 1), B = tan_base<L, RM>(Jc);
                for (int e = tid; e < E; e += NT) {
                    const int k = e / (R - 1), idx = e - k * (R - 1);
                    const int s = 31 - __builtin_clz(idx + 1), m = idx + 1 - (1 << s);
                    // W_{2^(s+1) NS}^(k + m NS) = W_n^((k + m NS) n / (2^(s+1) NS))
                    const f2 x = __ldg(wn + ((k + (m << NSL)) << (L - s - 1 - NSL)));
                    tt[B + e] = mk(x.x, x.y / x.x);
                }
            }(),
            ...);
    }(std::make_integer_sequence<int, npass(L, RM) - 1>{});
}

// a + w b and a - w b for w = c (1 + i t)
__device__ __forceinline__ void bfly_tan(f2& a, f2& b, const float c, const float t) {
    const float bx = __builtin_fmaf(-t, b.y, b.x), by = __builtin_fmaf(t, b.x, b.y);
    const f2 p = mk(__builtin_fmaf(c, bx, a.x), __builtin_fmaf(c, by, a.y));
    b = mk(__builtin_fmaf(-c, bx, a.x), __builtin_fmaf(-c, by, a.y));
    a = p;
}

// Stockham pass J > 0 on the thread's butterflies (as radix_pass), twiddles from the tangent table
template <int L, int J, int G, int RM>
__device__ __forceinline__ void radix_pass_tan(f2* v, int g, const f2* tt) {
    constexpr int LR = rlog(L, J, RM), R = 1 << LR, NS = 1 << nslog(L, J, RM);
    constexpr int NB = (1 << eptlog(L, RM)) / R;
    static_assert(J > 0 && NS > 1, "pass 0 has no twiddles");
    const f2* tj = tt + tan_base<L, RM>(J);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int k = (g + G * i) & (NS - 1);
        const f2* row = tj + k * (R - 1);
        f2 ent[R - 1];
#pragma unroll
        for (int h = 0; h < R - 1; ++h) ent[h] = row[h];
        f2* u = v + i * R;
#pragma unroll
        for (int s = 0; s < LR; ++s) {
            const int d = R >> (s + 1);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (j & d) continue;
                const int m = s == 0 ? 0 : brev(j >> (LR - s), s);
                const f2 e = ent[(1 << s) - 1 + m];
                bfly_tan(u[j], u[j + d], e.x, e.y);
            }
        }
        f2 t[R];
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = u[brev(r, LR)];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = t[r];
    }
}

// twiddle table W_n^m read from HBM (cached) or from an LDS copy
struct TwGlobal {
    const f2* __restrict__ p;
    __device__ __forceinline__ f2 operator()(int m) const { return __ldg(p + m); }
};
struct TwLds {
    const f2* p;
    __device__ __forceinline__ f2 operator()(int m) const { return p[m]; }
};

// output position of element r of butterfly b in pass J
template <int L, int J, int RM>
__device__ __forceinline__ int out_pos(int b, int r) {
    constexpr int LR = rlog(L, J, RM), NSL = nslog(L, J, RM);
    return ((b >> NSL) << (NSL + LR)) + (b & ((1 << NSL) - 1)) + (r << NSL);
}

// ------------------------------------------------------------ strided pass (y or z axis)
// Tile: P consecutive columns i0 (contiguous in HBM, P even) x the whole axis (2^L).  Thread
// tid handles the column pair (i0base + 2q, +1), q = tid % (P/2), and butterfly slot
// g = tid / (P/2): every HBM and LDS access moves 16 bytes (two complex values).  Element
// offsets inside the tile are 32-bit and relative to a wave-uniform base; strides are scalars.
// Buffers have an even row width, so the pair partner of the last active column is padding.
//
// Half-spectrum column 0 carries the Nyquist column kx = npx/2 in its imaginary part:
// c'(0, ky, kz) = c(0, ky, kz) + i c(npx/2, ky, kz).  Both planes are Hermitian in (ky, kz),
// so after the y and z transforms column 0 holds g0 + i gN with g0, gN real.
// PAD: one padding slot after every EPT positions in the exchange buffer (by default for tiles
// narrower than 16 columns, whose 16-byte lanes would otherwise meet in few LDS banks); a tile
// variant of the autotuner flips it.
template <int L, int P, int RM, bool PAD = (P < 16)>
struct StridedCfg {
    static_assert(P % 2 == 0, "column pairs");
    static constexpr int N = 1 << L, EL = eptlog(L, RM), EPT = 1 << EL, G = N / EPT, NT = G * (P / 2);
    static constexpr int M = npass(L, RM);
    static constexpr bool PADPOS = PAD;
    static constexpr int ppos(int pos) { return PADPOS ? pos + (pos >> EL) : pos; }
    static constexpr size_t lds_elems() { return (size_t)(N + (PADPOS ? (N >> EL) : 0)) * P; }
    // lengths up to 2^11: the passes J > 0 take their twiddles from a tangent table built in LDS
    // behind the exchange buffer (N - R_0 entries), so that no pass waits on vector-memory loads
    // (which complete in order with the stores); longer axes multiply by W_n table entries read
    // from memory
    static constexpr bool TWL = L <= 11;
    static constexpr size_t lds_bytes() {
        return lds_elems() * sizeof(f2) + (TWL ? (size_t)tan_entries<L, RM>() * sizeof(f2) : 0);
    }
};

__device__ __forceinline__ float4 pack2(f2 a, f2 b) { return make_float4(a.x, a.y, b.x, b.y); }

// Raw buffer access (offsets in bytes from a uniform base).  An offset >= kOob is out of range:
// a load returns zeros and a store is dropped, without touching memory -- so masked lanes and
// members past the end need no branch, and every wave issues the same memory instructions
// (which keeps the compiler's vector-memory counter waits exact across the member loop).
using u32x4 = unsigned int __attribute__((ext_vector_type(4)));
constexpr uint32_t kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, int bytes = 0x7fffffff) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
// loads non-temporal (aux 2): each operand is read once (2-4% faster y pass than cached loads)
__device__ __forceinline__ float4 bload16(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 2);
    return make_float4(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z), __uint_as_float(w.w));
}
__device__ __forceinline__ void bstore16(float4 v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const u32x4 w = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
    __builtin_amdgcn_raw_buffer_store_b128(w, rs, off, 0, 2);  // aux 2: non-temporal (streamed out)
}
// non-temporal 16-byte store: pass outputs are read back by the next pass, long after L2 has
// cycled, so they are not kept in cache
__device__ __forceinline__ void st_nt16(float4 v, float4* p) {
    using f4 = float __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store((f4){v.x, v.y, v.z, v.w}, reinterpret_cast<f4*>(p));
}

// Fill mode (IN_FILL32H): a block serves realizations b0 .. b1-1 of its tile (a.bpb per block).
// What depends on the site and not on the realization -- the amplitude, the product of the
// site's row factor and column factor, the conjugation and self-conjugate flags -- is read and
// computed once and kept in registers; per realization only its base factor is multiplied in
// (the Park-Miller state is a product mod 2^31-1, so the order of the factors does not matter).
// One tile of a strided pass (the work of one block): `tile` in XCD order, member group `grp`.
// HALF (passes that keep at most half the axis, n_out <= N/2): the last radix pass computes
// only the outputs r < R/2 of each butterfly.  Output r of a last-pass butterfly sits at position
// pos0 + r N/R with pos0 < N/R, so the outputs r >= R/2 lie at N/2 or beyond and are never stored.
//
// ZR (z fill only): wave-local z-run fill.  Thread tid holds column pair q = tid / G and butterfly
// slot g = tid % G, so a wave holds whole column pairs: both LDS exchanges of a member stay inside
// the wave, in the pairs' own padded slots of the exchange buffer, ordered by wave-scope fences
// (the LDS executes a wave's instructions in order), and the member loop has no block barrier.
// The output goes to the z-run layout: element z of column pair p in row `third` at float4
// (z / 8) sZ + p sP + third sK + z mod 8 (sZ, sP, sK = out_sA, out_sC, out_sT in float4 units), so
// that the 64 lanes of a store write whole runs of 8 z (128 bytes).  The site values are computed
// in the standard mapping (coalesced amplitude and row-factor reads) and moved to the wave-local
// one through LDS once per block.  Every column's arithmetic is that of the standard mapping.
template <int L, int P, int IN, int AXIS, int RM, int IG, bool HALF = false, bool PAD = (P < 16), bool ZR = false>
__device__ __forceinline__ void strided_tile(const PassArgs& a, const int tile, const int grp) {
    using C = StridedCfg<L, P, RM, PAD || ZR>;
    constexpr int N = C::N, EPT = C::EPT, G = C::G, M = C::M, PH = P / 2;
    static_assert(M > 0, "strided passes have length >= 2");
    static_assert(!ZR || (IN == IN_FILL32H && AXIS == 2 && 64 % G == 0 && C::TWL), "z-run: wave-local z fill");
    // slot of one column pair in the exchange buffer (z-run: padded positions, pair-major)
    constexpr int ZSLOT = N + (N >> C::EL);
    static_assert(!ZR || (size_t)ZSLOT * PH * 2 == C::lds_elems(), "z-run slots fill the exchange buffer");
    float4* lds4 = reinterpret_cast<float4*>(fast_smem);
    const f2* __restrict__ tw = reinterpret_cast<const f2*>(a.tw);
    f2* twl = reinterpret_cast<f2*>(fast_smem + C::lds_elems() * sizeof(f2));
    const int tid = threadIdx.x;
    const int q = ZR ? tid / G : tid % PH, g = ZR ? tid % G : tid / PH;
    if constexpr (C::TWL && M > 1) {
        // first read after the barrier of the first exchange (z-run: after the site transposition)
        tan_table_build<L, RM, C::NT>(twl, tw, tid);
    }
    // exchange slot of position pos of pair qq, and the order between an exchange's writes and reads
    auto xslot = [](const int pos, const int qq) __attribute__((always_inline)) {
        if constexpr (ZR) return qq * ZSLOT + C::ppos(pos);
        else return C::ppos(pos) * PH + qq;
    };
    auto xsync = [] __attribute__((always_inline)) {
        if constexpr (ZR) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        else __syncthreads();
    };

    const int i0base = (tile % a.ntiles0) * P;
    // element offsets of the tile's first column in the input and output (column-tile strides
    // sC: P for x-fastest buffers, the size of a column tile's slab for column-tiled ones)
    const int64_t cin = (int64_t)(tile % a.ntiles0) * (a.in_sC ? a.in_sC : P);
    const int64_t cout = (int64_t)(tile % a.ntiles0) * (a.out_sC ? a.out_sC : P);
    const int third = tile / a.ntiles0;
    const int c0 = i0base + 2 * q;
    const bool active = c0 < a.n0;
    f2 v0[EPT], v1[EPT];

    // radix passes (LDS exchange between them), then the store of batch member b:
    // butterfly g + G*i, element r at out_pos; positions >= n_out pruned
    auto transform_store = [&](int b, const int g, const int q) __attribute__((always_inline)) {
        // exchange read of position g + m of the thread's pair (m constant after unrolling): with
        // z-run slots, a multiple m of EPT adds ppos(m) to ppos(g) (the padding counts whole
        // groups of EPT positions), so the reads of a pass share one base register
        const int xbase = ZR ? q * ZSLOT + C::ppos(g) : 0;
        auto xread = [&](const int m) __attribute__((always_inline)) {
            if constexpr (ZR) {
                if (m % EPT == 0) return xbase + C::ppos(m);
            }
            return xslot(g + m, q);
        };
        [&]<int... J>(std::integer_sequence<int, J...>) {
            (
                [&] {
                    constexpr int Jc = J;
                    if constexpr (Jc > 0) {
                        constexpr int LRp = rlog(L, Jc - 1, RM), NBp = EPT >> LRp;
                        constexpr int LR = rlog(L, Jc, RM), NB = EPT >> LR;
                        if constexpr (Jc > 1) xsync();
#pragma unroll
                        for (int i = 0; i < NBp; ++i)
#pragma unroll
                            for (int r = 0; r < (1 << LRp); ++r) {
                                const int e = i * (1 << LRp) + r;
                                lds4[xslot(out_pos<L, Jc - 1, RM>(g + G * i, r), q)] = pack2(v0[e], v1[e]);
                            }
                        xsync();
#pragma unroll
                        for (int i = 0; i < NB; ++i)
#pragma unroll
                            for (int r = 0; r < (1 << LR); ++r) {
                                const int e = i * (1 << LR) + r;
                                const float4 w = lds4[xread(i * G + r * (N >> LR))];
                                v0[e] = mk(w.x, w.y);
                                v1[e] = mk(w.z, w.w);
                            }
                    }
                    if constexpr (Jc == 0 && IN == IN_FILL32H) {
                        // the member loop has done the first radix-2 stage of the single pass-0
                        // butterfly (radix EPT, no twiddles), in the order of idft_stages
                        idft_stages<C::EL, 4>(v0);
                        idft_stages<C::EL, 4>(v1);
                    } else if constexpr (Jc > 0 && C::TWL) {
                        radix_pass_tan<L, Jc, G, RM>(v0, g, twl);
                        radix_pass_tan<L, Jc, G, RM>(v1, g, twl);
                    } else {
                        radix_pass<L, Jc, G, RM>(v0, g, TwGlobal{tw});
                        radix_pass<L, Jc, G, RM>(v1, g, TwGlobal{tw});
                    }
                }(),
                ...);
        }(std::make_integer_sequence<int, M>{});
        constexpr int LRl = rlog(L, M - 1, RM), NBl = EPT >> LRl;
        const uint32_t sA = (uint32_t)a.out_sA >> 1;  // in float4 units
        const int nout = a.n_out;
        const f2* obase = reinterpret_cast<f2*>(a.out) + (int64_t)b * a.out_bstride + (int64_t)third * a.out_sT + cout;
        if constexpr (IN == IN_LOAD) {  // branch-free: pruned and inactive lanes store out of range
            const __amdgpu_buffer_rsrc_t rs = rsrc(obase);
#pragma unroll
            for (int i = 0; i < NBl; ++i)
#pragma unroll
                for (int r = 0; r < (1 << LRl); ++r) {
                    const int pos = out_pos<L, M - 1, RM>(g + G * i, r);
                    const int e = i * (1 << LRl) + r;
                    const uint32_t off = (active && pos < nout) ? (q + (uint32_t)pos * sA) * 16u : kOob;
                    if (!HALF || r < (1 << LRl) / 2) bstore16(pack2(v0[e], v1[e]), rs, off);
                }
        } else if constexpr (ZR) {
            // z-run stores through a resource spanning the member's buffer; pruned positions and
            // inactive lanes use an out-of-range offset, and a store whose positions are pruned for
            // the whole wave is not issued
            constexpr int NSLl = nslog(L, M - 1, RM);
            const uint32_t sZ = (uint32_t)(a.out_sA >> 1), sP = (uint32_t)(a.out_sC >> 1), sK = (uint32_t)(a.out_sT >> 1);
            const __amdgpu_buffer_rsrc_t rs =
                rsrc(reinterpret_cast<f2*>(a.out) + (int64_t)b * a.out_bstride, (int)(a.out_bstride * 8));
            static_assert(NSLl >= 3, "the elements of a last-pass butterfly lie whole runs apart");
            const uint32_t o0 = (uint32_t)(i0base / 2 + q) * sP + (uint32_t)third * sK;
            const uint32_t rstep = (sZ << (NSLl - 3)) * 16u;  // bytes from element r to r + 1 (uniform)
#pragma unroll
            for (int i = 0; i < NBl; ++i) {
                const int pos0 = out_pos<L, M - 1, RM>(g + G * i, 0);
                const int wpos0 = __builtin_amdgcn_readfirstlane(pos0);
                const uint32_t b0 = (o0 + (uint32_t)(pos0 >> 3) * sZ + (uint32_t)(pos0 & 7)) * 16u;
#pragma unroll
                for (int r = 0; r < (1 << LRl); ++r) {
                    const int e = i * (1 << LRl) + r;
                    const uint32_t off = active && pos0 + (r << NSLl) < nout ? b0 + (uint32_t)r * rstep : kOob;
                    if ((!HALF || r < (1 << LRl) / 2) && wpos0 + (r << NSLl) < nout) bstore16(pack2(v0[e], v1[e]), rs, off);
                }
            }
        } else {
            // fill passes: element r of butterfly i sits at position pos(i, 0) + r 2^NSL.  When the
            // axis spans less than 2 GiB the stores go through a buffer resource that ends at
            // position nout (pruned positions fall outside it; inactive lanes use an out-of-range
            // offset); a store whose positions are pruned for the whole wave is not issued (the
            // positions grow with the lane, so lane 0 holds the wave's smallest)
            constexpr int NSLl = nslog(L, M - 1, RM);
            const uint32_t step = sA << NSLl;  // float4 units from element r to r + 1
            if ((int64_t)N * a.out_sA * 8 <= 0x7fffffffLL) {
                const __amdgpu_buffer_rsrc_t rs = rsrc(obase, (int)((int64_t)nout * a.out_sA * 8));
#pragma unroll
                for (int i = 0; i < NBl; ++i) {
                    const int pos0 = out_pos<L, M - 1, RM>(g + G * i, 0);
                    const int wpos0 = __builtin_amdgcn_readfirstlane(pos0);
                    const uint32_t o0 = active ? (q + (uint32_t)pos0 * sA) * 16u : kOob;
#pragma unroll
                    for (int r = 0; r < (1 << LRl); ++r) {
                        const int e = i * (1 << LRl) + r;
                        if ((!HALF || r < (1 << LRl) / 2) && wpos0 + (r << NSLl) < nout)
                            bstore16(pack2(v0[e], v1[e]), rs, o0 + r * step * 16u);
                    }
                }
            } else if (active) {
                float4* __restrict__ dst = reinterpret_cast<float4*>(const_cast<f2*>(obase)) + q;
#pragma unroll
                for (int i = 0; i < NBl; ++i) {
                    const int pos0 = out_pos<L, M - 1, RM>(g + G * i, 0);
                    float4* __restrict__ d0 = dst + (int64_t)pos0 * sA;
#pragma unroll
                    for (int r = 0; r < (1 << LRl); ++r) {
                        const int e = i * (1 << LRl) + r;
                        if ((!HALF || r < (1 << LRl) / 2) && pos0 + (r << NSLl) < nout)
                            st_nt16(pack2(v0[e], v1[e]), d0 + (int64_t)r * step);
                    }
                }
            }
        }
    };

    if constexpr (IN == IN_LOAD) {  // pass-0 operands at positions g + r*G
        // members b0 .. b1-1 of this tile; the operands of member b+1 are loaded while member b
        // is transformed and stored
        const int b0 = grp * a.bpb, b1 = min(b0 + a.bpb, a.nbatch);
        const uint32_t sA = (uint32_t)a.in_sA >> 1;  // in float4 units
        const uint32_t o0 = q + (uint32_t)g * sA, step = (uint32_t)G * sA;
        float4 nxt[EPT];
        auto load = [&](int b) {  // b >= b1: zeros, no memory access
            const int bb = min(b, b1 - 1);
            const __amdgpu_buffer_rsrc_t rs = rsrc(reinterpret_cast<const f2*>(a.in) + (int64_t)bb * a.in_bstride +
                                                   (int64_t)third * a.in_sT + cin);
            const bool ok = active && b < b1;
#pragma unroll
            for (int r = 0; r < EPT; ++r) nxt[r] = bload16(rs, ok ? (o0 + r * step) * 16u : kOob);
        };
        auto consume = [&] {
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                v0[r] = mk(nxt[r].x, nxt[r].y);
                v1[r] = mk(nxt[r].z, nxt[r].w);
            }
        };
        load(b0);
        {
            // as many out-of-range (dropped) stores as a member issues: the loop is entered with
            // the same vector-memory counter history as from its back edge, so the wait for the
            // prefetched operands does not also wait for the previous member's stores
            const __amdgpu_buffer_rsrc_t rs = rsrc(a.out);
#pragma unroll
            for (int e = 0; e < (HALF ? EPT / 2 : EPT); ++e)
                bstore16(make_float4(0.f, 0.f, 0.f, 0.f), rs, kOob + 16u * e);
        }
        consume();
        for (int b = b0; b < b1; ++b) {
            load(b + 1);  // in flight while member b is transformed and stored
            int gg = g, qq = q;
            asm volatile("" : "+v"(gg), "+v"(qq));
            transform_store(b, gg, qq);
            consume();
            __syncthreads();  // LDS reads of member b are done before the exchange of b+1
        }
    } else {  // IN_FILL32H: AXIS 2: (ky, kz) = (third, pos), AXIS 1: (pos, third)
        const int b0 = grp * a.bpb, b1 = min(b0 + a.bpb, a.nbatch);
        // Per element and column: row factor x column factor of the site, and the amplitude,
        // negated when the site is the partner (conjugate).  A site is self-conjugate only in
        // column 0 (kx = 0, and the Nyquist column it carries) on a self row (row == partner
        // row): selfm bit r; conjm bit r: column 0's site is the partner.
        uint32_t w0[EPT], w1[EPT];
        float2 ampv[EPT];
        uint32_t selfm = 0, conjm = 0;
        const bool col0 = active && c0 == 0, pack = col0 && a.npx >= 2;
        const int kxn = a.npx >> 1;
        const uint32_t txn = pack ? __ldg(a.txp + kxn) : 0u;
        // amplitude of element r of the pair at column cc, slot gg: amp32h[cc + hxa * row(pos)], row
        // affine in r
        auto amp_base = [&](const int cc, const int gg) __attribute__((always_inline)) {
            const int64_t arow0 =
                AXIS == 2 ? (int64_t)third + (int64_t)a.npy * gg : (int64_t)gg + (int64_t)a.npy * third;
            return a.amp32h + cc + a.hxa * arow0;
        };
        const int64_t arstep = AXIS == 2 ? (int64_t)a.npy * G : G;
        const float* __restrict__ abase = amp_base(c0, g);
        const int64_t astep = a.hxa * arstep;
        // the site values of column pair qq at slot gg
        auto site_values = [&](const int qq, const int gg) __attribute__((always_inline)) {
            const int cc = i0base + 2 * qq;
            selfm = conjm = 0;
            if (cc < a.n0) {
                const float* __restrict__ ab = amp_base(cc, gg);
                const int kxa = cc, kxb = cc + 1;
                const int pa = (a.npx - kxa) & (a.npx - 1), pb = (a.npx - kxb) & (a.npx - 1);
                const uint32_t txa = __ldg(a.txp + kxa), txpa = __ldg(a.txp + pa);
                const uint32_t txb = __ldg(a.txp + kxb), txpb = __ldg(a.txp + pb);
#pragma unroll
                for (int r = 0; r < EPT; ++r) {
                    const int pos = gg + r * G;
                    const Site sv = AXIS == 2 ? site_of(a, third, pos) : site_of(a, pos, third);
                    const uint32_t rf = __ldg(a.rowf + sv.site);
                    const float2 am = __ldg(reinterpret_cast<const float2*>(ab + r * astep));
                    // the site row is shared by every column; x order decides only inside self rows
                    const bool ja = sv.gt || (sv.eq && kxa > pa);
                    const bool jb = sv.gt || (sv.eq && kxb > pb);
                    w0[r] = pm_mul(rf, ja ? txpa : txa);
                    w1[r] = pm_mul(rf, jb ? txpb : txb);
                    ampv[r] = make_float2(ja ? -am.x : am.x, jb ? -am.y : am.y);
                    if (sv.eq) selfm |= 1u << r;
                    if (ja) conjm |= 1u << r;
                }
            } else {
                // inactive lanes: amplitude 0 and a valid state (coefficient 0), so that the member
                // loop runs without a branch on the lane's column
#pragma unroll
                for (int r = 0; r < EPT; ++r) {
                    w0[r] = w1[r] = 1u;
                    ampv[r] = make_float2(0.f, 0.f);
                }
            }
        };
        if constexpr (ZR) {
            // computed in the standard mapping (pair tid % PH, slot tid / PH), then moved to the
            // wave-local one through each pair's exchange slot: G * EPT values of 16 bytes, then
            // the two masks of each slot.  Value r of slot gg of pair qq sits at gg EPT +
            // (r + qq + gg / 2) mod EPT: the 16 lanes of a 16-byte LDS access (8 pairs x 2 slots
            // when written, 16 slots of one pair when read) then fall in distinct banks.
            uint4* xs = reinterpret_cast<uint4*>(fast_smem);
            static_assert(G * EPT * 16 + G * 8 <= ZSLOT * 16, "site values inside a pair's slot");
            auto vslot = [](const int qq, const int gg, const int r) __attribute__((always_inline)) {
                return qq * ZSLOT + gg * EPT + ((r + qq + (gg >> 1)) & (EPT - 1));
            };
            {
                const int qs = tid % PH, gs = tid / PH;
                site_values(qs, gs);
#pragma unroll
                for (int r = 0; r < EPT; ++r)
                    xs[vslot(qs, gs, r)] = make_uint4(w0[r], w1[r], __float_as_uint(ampv[r].x), __float_as_uint(ampv[r].y));
                uint32_t* dm = reinterpret_cast<uint32_t*>(xs + qs * ZSLOT + G * EPT) + 2 * gs;
                dm[0] = selfm;
                dm[1] = conjm;
            }
            __syncthreads();  // also orders the tangent table's writes before its first read
            {
#pragma unroll
                for (int r = 0; r < EPT; ++r) {
                    const uint4 u = xs[vslot(q, g, r)];
                    w0[r] = u.x;
                    w1[r] = u.y;
                    ampv[r] = make_float2(__uint_as_float(u.z), __uint_as_float(u.w));
                }
                const uint32_t* sm = reinterpret_cast<const uint32_t*>(xs + q * ZSLOT + G * EPT) + 2 * g;
                selfm = sm[0];
                conjm = sm[1];
            }
            // no block barrier before the first exchange: a wave writes only its own pairs' slots,
            // whose moved values only that wave reads, in order
        } else {
            site_values(q, g);
        }
        // the member loop, straight-line for the coefficient law IG (igauss) of the instantiation
        for (int b = b0; b < b1; ++b) {
            if (b > b0) xsync();  // LDS reads of the previous member are done
            // realization base: a uniform read through the scalar cache (written by an earlier
            // kernel, never by this one), so that no vector-memory wait (which would also wait
            // for the stores of the previous member) is needed
            const uint32_t rb = ((const __attribute__((address_space(4))) uint32_t*)a.rbase)[b];
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                // keep the site values opaque so that nothing derived from them is hoisted
                // out of the member loop (registers)
                asm volatile("" : "+v"(w0[r]), "+v"(w1[r]), "+v"(ampv[r].x), "+v"(ampv[r].y));
            }
            // The coefficients enter pass 0's single butterfly (radix EPT, no twiddles) through its
            // first radix-2 stage: elements j and j + EPT/2 give positions brev(j) and brev(j) + 1 of
            // the order idft_stages works in, their sum and difference, with the second element's
            // amplitude products fused into them.
            constexpr int H2 = EPT / 2;
            auto stage1_col = [&](f2* v, const bool second) __attribute__((always_inline)) {
#pragma unroll
                for (int j = 0; j < H2; ++j) {
                    float Ra, ca, sa, Rb, cb, sb;
                    coef_parts<IG>(second ? ampv[j].y : ampv[j].x, pm_mul_lazy(second ? w1[j] : w0[j], rb), Ra, ca, sa);
                    coef_parts<IG>(second ? ampv[j + H2].y : ampv[j + H2].x,
                                   pm_mul_lazy(second ? w1[j + H2] : w0[j + H2], rb), Rb, cb, sb);
                    const int k = brev(j, C::EL);
                    stage1_parts(v[k], v[k + 1], Ra, ca, sa, Rb, cb, sb);
                }
            };
            stage1_col(v1, true);
            if (col0) {  // kx = 0: self-conjugate sites; the Nyquist column (kx = npx/2, its own partner)
                int zero = 0;  // opaque: the addresses below are not kept across the loop
                asm volatile("" : "+v"(zero));
                const float* __restrict__ ab = abase + zero;
                f2 cv[EPT];
#pragma unroll
                for (int r = 0; r < EPT; ++r) {
                    cv[r] = coef_signed<IG>(ampv[r].x, pm_mul_lazy(w0[r], rb));
                    const uint32_t st = pm_mul(w0[r], rb);  // column factor of kx = 0 is 1
                    const bool cj = (conjm >> r) & 1u, sf = (selfm >> r) & 1u;
                    if (sf) cv[r] = coef_t<IG>(fabsf(ampv[r].x), st, false, true);
                    if (pack) {
                        const f2 zn = coef_t<IG>(__ldg(ab + r * astep + kxn), pm_mul(st, txn), cj, sf);
                        cv[r] = mk(cv[r].x - zn.y, cv[r].y + zn.x);
                    }
                }
#pragma unroll
                for (int j = 0; j < H2; ++j) {
                    const int k = brev(j, C::EL);
                    v0[k] = add(cv[j], cv[j + H2]);
                    v0[k + 1] = sub(cv[j], cv[j + H2]);
                }
            } else {
                stage1_col(v0, false);
            }
            // per-member recomputation of the thread's LDS, twiddle and store addresses (opaque
            // thread indices) instead of keeping them in registers across the member loop
            int gg = g, qq = q;
            asm volatile("" : "+v"(gg), "+v"(qq));
            transform_store(b, gg, qq);
        }
    }
}

template <int L, int P, int IN, int AXIS, int RM, int IG, bool PAD = (P < 16), bool ZR = false>
__global__ __launch_bounds__((StridedCfg<L, P, RM, PAD>::NT)) void strided_kernel(PassArgs a) {
    strided_tile<L, P, IN, AXIS, RM, IG, false, PAD, ZR>(a, xcd_remap(blockIdx.x, gridDim.x), blockIdx.y);
}

// ------------------------------------------------------------ fill pass beside a y pass
// The z fill pass of one batch (length 2^L, tile P, plan RM) and the y pass of the previous batch
// (length 2^LY, tile PY, plan RMY; tiles of the same thread count) in one launch.  Block roles
// interleave -- 8A fill blocks, then 8C y blocks, K times, then the rest of each -- so that every
// CU holds arithmetic-bound fill tiles and memory-bound y tiles at once.  Role-local block l runs
// tile l mod nblocks (in XCD order) of member group l / nblocks of its pass; 8A and 8C keep l and
// the block index equal mod 8 (the same XCD).  ZR: the fill role is the wave-local z-run fill
// (the y role then reads the z-run layout through its strides, as any other layout).
template <int L, int P, int RM, int IG, int LY = L, int PY = P, int RMY = RM, bool HALF = false, bool ZR = false>
__global__ __launch_bounds__((StridedCfg<L, P, RM>::NT)) void fill_load_kernel(PassArgs az, PassArgs ay, FuseMap m) {
    static_assert(StridedCfg<LY, PY, RMY>::NT == StridedCfg<L, P, RM>::NT, "tiles of one thread count");
    const int per = 8 * (m.A + m.C), bid = blockIdx.x, full = m.K * per;
    bool fill;
    int l;
    if (bid < full) {
        const int pr = bid / per, o = bid - pr * per;
        fill = o < 8 * m.A;
        l = fill ? pr * 8 * m.A + o : pr * 8 * m.C + (o - 8 * m.A);
    } else {
        const int o = bid - full, zr = m.nzb - 8 * m.A * m.K;
        fill = o < zr;
        l = fill ? 8 * m.A * m.K + o : 8 * m.C * m.K + (o - zr);
    }
    if (fill)
        strided_tile<L, P, IN_FILL32H, 2, RM, IG, HALF, (P < 16), ZR>(az, xcd_remap(l % az.nblocks, az.nblocks),
                                                                     l / az.nblocks);
    else
        strided_tile<LY, PY, IN_LOAD, 1, RMY, 0, HALF>(ay, xcd_remap(l % ay.nblocks, ay.nblocks), l / ay.nblocks);
}

// ------------------------------------------------------------ C2R pass along x
// Rows of the half spectrum X[0..n/2] (n = npx = 2^(L+1)) -> real row of n, cropped to nx,
// with FP64 statistics.  Thread tid: row q = tid / G, butterfly slot g = tid % G.
template <int L, int RM>
struct C2RCfg {
    static constexpr int N = 1 << L, EL = eptlog(L, RM), EPT = 1 << EL, G = N / EPT;
    static constexpr int P = G >= 256 ? 1 : 256 / G, NT = G * P, M = npass(L, RM);
    static constexpr int ROW = N + (N >> EL);
    static constexpr int ppos(int pos) { return pos + (pos >> EL); }
    static constexpr size_t lds_elems() { return (size_t)ROW * P; }
};

// X[N-k] for k = g + r G from the lanes of the same row (G <= 64: a row lies inside one wave):
// register EPT-1-r of lane G-g of the row (g >= 1), or this lane's register EPT-r (g = 0, r >= 1)
// -- each value read once from memory, exchanged
template <int EPT, int G>
__device__ __forceinline__ void c2r_mirror(const f2* xa, f2* xb, const int g) {
    static_assert(G <= 64, "a row inside one wave");
    const int lane = threadIdx.x & 63;
    const int plane = lane - g + ((G - g) & (G - 1));
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
        const f2 t = xa[EPT - 1 - r];
        const f2 sh = mk(__shfl(t.x, plane, 64), __shfl(t.y, plane, 64));
        xb[r] = g != 0 ? sh : (r == 0 ? mk(xa[0].y, 0.f) : xa[(EPT - r) & (EPT - 1)]);
    }
}

// Pass-0 operands of the length-N inverse transform of a real row of 2N: v[r] = E + i O from
// X[k] (xa) and X[N-k] (xb), k = g + r G; tw2 reads the W_{2N} table.
template <int EPT, int G, typename TW>
__device__ __forceinline__ void c2r_operands(const f2* xa, const f2* xb, TW tw2, const int g, f2* v) {
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
        const int k = g + r * G;
        const f2 A = xa[r];
        const f2 B = mk(xb[r].x, -xb[r].y);  // conj X[N-k]
        const f2 E = add(A, B);
        const f2 O = mul(tw2(k), sub(A, B));
        v[r] = mk(E.x - O.y, E.y + O.x);
    }
}

// The radix passes of the rows' transforms, exchanging through LDS between passes (the block
// waits at a barrier before each exchange after the first, and after each write).
template <int L, int RM, typename TW>
__device__ __forceinline__ void c2r_passes(f2* v, f2* lds, const int q, const int g, TW tw) {
    using C = C2RCfg<L, RM>;
    constexpr int N = C::N, EPT = C::EPT, G = C::G, M = C::M;
    [&]<int... J>(std::integer_sequence<int, J...>) {
        (
            [&] {
                constexpr int Jc = J;
                if constexpr (Jc > 0) {
                    constexpr int LRp = rlog(L, Jc - 1, RM), NBp = EPT >> LRp;
                    constexpr int LR = rlog(L, Jc, RM), NB = EPT >> LR;
                    if constexpr (Jc > 1) __syncthreads();
#pragma unroll
                    for (int i = 0; i < NBp; ++i)
#pragma unroll
                        for (int r = 0; r < (1 << LRp); ++r)
                            lds[q * C::ROW + C::ppos(out_pos<L, Jc - 1, RM>(g + G * i, r))] = v[i * (1 << LRp) + r];
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int r = 0; r < (1 << LR); ++r)
                            v[i * (1 << LR) + r] = lds[q * C::ROW + C::ppos(g + G * i + r * (N >> LR))];
                }
                radix_pass<L, Jc, G, RM>(v, g, tw);
            }(),
            ...);
    }(std::make_integer_sequence<int, M>{});
}

// Block statistics (sum, sum of squares; max|Im| = 0 for a real-output transform): wave sums
// into red[0 .. 2 NW) (the exchange buffer, once every wave is past its last LDS read); on return
// every thread may read them.
template <int NW>
__device__ __forceinline__ void c2r_wave_sums(double s1, double s2, double* red) {
    s1 = wsum(s1);
    s2 = wsum(s2);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[wid] = s1;
        red[NW + wid] = s2;
    }
    __syncthreads();
}

// One block of the C2R pass: rows rblock*P .. +P-1 of batch member b.
template <int L, int IN, int RM>
__device__ __forceinline__ void c2r_tile(const PassArgs& a, const int rblock, const int b) {
    using C = C2RCfg<L, RM>;
    constexpr int N = C::N, EPT = C::EPT, G = C::G, P = C::P, M = C::M;
    f2* lds = reinterpret_cast<f2*>(fast_smem);
    const f2* __restrict__ tw = reinterpret_cast<const f2*>(a.tw);    // W_N table
    const f2* __restrict__ tw2 = tw + N;                               // W_{2N} table (next length)
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;
    const int row = rblock * P + q;
    const bool active = row < a.nrows;
    f2 v[EPT];

    if (active) {
        f2 xa[EPT], xb[EPT];
        if constexpr (IN == IN_LOADH) {
            const f2* __restrict__ src = reinterpret_cast<const f2*>(a.in) + (int64_t)b * a.in_bstride +
                                         (int64_t)rblock * P * a.in_row;
            const uint32_t base = (uint32_t)q * (uint32_t)a.in_row;
#pragma unroll
            for (int r = 0; r < EPT; ++r) {  // read once: non-temporal
                const f2v t = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(src + base + g + r * G));
                xa[r] = mk(t.x, t.y);
            }
            if constexpr (G <= 64) {
                c2r_mirror<EPT, G>(xa, xb, g);
            } else {
#pragma unroll
                for (int r = 0; r < EPT; ++r) {
                    const int k = g + r * G;
                    xb[r] = k == 0 ? mk(xa[r].y, 0.f) : src[base + (N - k)];
                }
            }
            // column 0 = g0 + i gN (packed Nyquist column): X[0] = Re, X[N] = Im
            if (g == 0) xa[0].y = 0.f;
        } else {  // IN_FILLH: the 1D grid (npy = npz = 1) filled in place
            const uint32_t rb = __ldg(a.rbase + b);
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                const int k = g + r * G;
                xa[r] = fill_at(a, column_of(a, rb, k), 0, 0);
                xb[r] = fill_at(a, column_of(a, rb, N - k), 0, 0);
                if (k == 0) {  // DC and Nyquist kept real
                    xa[r].y = 0.f;
                    xb[r].y = 0.f;
                }
            }
        }
        c2r_operands<EPT, G>(xa, xb, TwGlobal{tw2}, g, v);
    } else {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = mk(0.f, 0.f);
    }

    c2r_passes<L, RM>(v, lds, q, g, TwGlobal{tw});

    double s1 = 0.0, s2 = 0.0;
    if (active) {
        float* __restrict__ dst = reinterpret_cast<float*>(a.out) + (int64_t)b * a.out_bstride +
                                  (int64_t)rblock * P * a.out_row;
        const uint32_t base = (uint32_t)q * (uint32_t)a.out_row;
        constexpr int LRl = M ? rlog(L, M > 0 ? M - 1 : 0, RM) : 0, NBl = EPT >> LRl;
        const int nx = a.n_out;
        // 8-byte aligned pairs: one non-temporal float2 store (the field is not read again by
        // this pass; streamed out, it leaves the cache to the fill's tables: fill 1.01 -> 0.94 ms)
        const bool even = ((a.out_row | a.out_bstride) & 1) == 0 && (reinterpret_cast<uintptr_t>(a.out) & 7) == 0;
#pragma unroll
        for (int i = 0; i < NBl; ++i)
#pragma unroll
            for (int r = 0; r < (1 << LRl); ++r) {
                int m = g;
                if constexpr (M > 0) m = out_pos<L, M - 1, RM>(g + G * i, r);
                const f2 z = v[i * (1 << LRl) + r];
                if (2 * m + 1 < nx) {
                    if (even) {
                        __builtin_nontemporal_store((f2v){z.x, z.y}, reinterpret_cast<f2v*>(dst + base + 2 * m));
                    } else {
                        dst[base + 2 * m] = z.x;
                        dst[base + 2 * m + 1] = z.y;
                    }
                    s1 += (double)z.x + (double)z.y;
                    s2 += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
                } else if (2 * m < nx) {
                    dst[base + 2 * m] = z.x;
                    s1 += (double)z.x;
                    s2 += (double)z.x * (double)z.x;
                }
            }
    }
    constexpr int NW = (C::NT + 63) / 64;
    static_assert(C::lds_elems() * sizeof(f2) >= 2 * NW * sizeof(double), "room for the block sums");
    double* red = reinterpret_cast<double*>(fast_smem);
    c2r_wave_sums<NW>(s1, s2, red);
    if (tid == 0) {
        double t1 = 0, t2 = 0;
        for (int w = 0; w < NW; ++w) {
            t1 += red[w];
            t2 += red[NW + w];
        }
        double* p = a.partial + 3 * ((int64_t)b * a.nblocks + rblock);
        p[0] = t1;
        p[1] = t2;
        p[2] = 0.0;
    }
}

template <int L, int IN, int RM>
__global__ __launch_bounds__((C2RCfg<L, RM>::NT)) void c2r_kernel(PassArgs a) {
    c2r_tile<L, IN, RM>(a, blockIdx.x, blockIdx.y);
}

// 8-byte raw buffer accesses (non-temporal), out-of-range offsets as for the 16-byte ones
__device__ __forceinline__ f2 bload8(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const unsigned int __attribute__((ext_vector_type(2))) w = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 2);
    return mk(__uint_as_float(w.x), __uint_as_float(w.y));
}
__device__ __forceinline__ void bstore8(f2 v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const unsigned int __attribute__((ext_vector_type(2))) w = {__float_as_uint(v.x), __float_as_uint(v.y)};
    __builtin_amdgcn_raw_buffer_store_b64(w, rs, off, 0, 2);
}
__device__ __forceinline__ void bstore4(float v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, off, 0, 2);
}
__device__ __forceinline__ void bstore8d(double v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const uint64_t u = __double_as_longlong(v);
    const unsigned int __attribute__((ext_vector_type(2))) w = {(unsigned int)u, (unsigned int)(u >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(w, rs, off, 0, 2);
}

// The C2R pass from the y output (IN_LOADH) over members b0 .. b1-1 of one tile of P rows (rows in
// one wave): the rows of member b+1 are loaded while member b is transformed and stored.  EVEN (an
// even cropped length and row and member strides, 8-byte aligned output): output pairs go out as
// one 8-byte store; otherwise as two 4-byte stores, the second dropped past an odd length.  Every
// access goes through a buffer resource with out-of-range offsets for rows past the end, members
// past the last and the lanes that write no partial, so every wave issues the same memory
// instructions per member and the wait for the prefetched rows never includes the stores of the
// member before.  Both twiddle tables are copied to LDS behind the exchange buffer (a load from
// memory would complete only after the prefetch).  Per member, the arithmetic and the statistics
// partials are those of c2r_tile.
template <int L, int RM>
struct C2RRowsCfg {
    using C = C2RCfg<L, RM>;
    static constexpr size_t lds_bytes() { return (C::lds_elems() + 2 * (size_t)C::N) * sizeof(f2); }
};
template <int L, int RM, bool EVEN>
__device__ __forceinline__ void c2r_rows(const PassArgs& a, const int rblock, const int b0, const int b1) {
    using C = C2RCfg<L, RM>;
    constexpr int N = C::N, EPT = C::EPT, G = C::G, P = C::P, M = C::M, NW = (C::NT + 63) / 64;
    static_assert(G <= 64 && M > 0, "rows inside one wave, at least one radix pass");
    static_assert(C::lds_elems() * sizeof(f2) >= 2 * NW * sizeof(double), "room for the block sums");
    f2* lds = reinterpret_cast<f2*>(fast_smem);
    f2* twl = lds + C::lds_elems();  // W_N (N entries), then the first N entries of W_{2N}
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;
    const bool active = rblock * P + q < a.nrows;
    {
        const f2* __restrict__ tw = reinterpret_cast<const f2*>(a.tw);
        for (int m = tid; m < 2 * N; m += C::NT) twl[m] = __ldg(tw + m);
    }
    const uint32_t ibase = (uint32_t)q * (uint32_t)a.in_row + g;
    f2 nxt[EPT];
    auto load = [&](int b) {  // b >= b1: zeros, no memory access
        const int bb = min(b, b1 - 1);
        const __amdgpu_buffer_rsrc_t rs =
            rsrc(reinterpret_cast<const f2*>(a.in) + (int64_t)bb * a.in_bstride + (int64_t)rblock * P * a.in_row);
        const bool ok = active && b < b1;
#pragma unroll
        for (int r = 0; r < EPT; ++r) nxt[r] = bload8(rs, ok ? (ibase + r * G) * 8u : kOob);
    };
    constexpr int LRl = rlog(L, M - 1, RM), NBl = EPT >> LRl;
    __syncthreads();  // twiddles in LDS
    load(b0);
    {
        // as many out-of-range (dropped) stores as a member issues: the loop is entered with the
        // same vector-memory counter history as from its back edge
        const __amdgpu_buffer_rsrc_t rs = rsrc(a.out);
#pragma unroll
        for (int e = 0; e < EPT + 3; ++e) bstore8(mk(0.f, 0.f), rs, kOob + 8u * e);
        if constexpr (!EVEN) {
#pragma unroll
            for (int e = 0; e < EPT; ++e) bstore4(0.f, rs, kOob + 4u * e);
        }
    }
    const uint32_t obase = (uint32_t)q * (uint32_t)a.out_row;
    const int nx = a.n_out;
    for (int b = b0; b < b1; ++b) {
        f2 xa[EPT], xb[EPT], v[EPT];
#pragma unroll
        for (int r = 0; r < EPT; ++r) xa[r] = nxt[r];
        load(b + 1);  // in flight while member b is transformed and stored
        if (b > b0) __syncthreads();  // the block sums of member b-1 are read
        c2r_mirror<EPT, G>(xa, xb, g);
        // column 0 = g0 + i gN (packed Nyquist column): X[0] = Re, X[N] = Im
        if (g == 0) xa[0].y = 0.f;
        c2r_operands<EPT, G>(xa, xb, TwLds{twl + N}, g, v);
        c2r_passes<L, RM>(v, lds, q, g, TwLds{twl});
        double s1 = 0.0, s2 = 0.0;
        const __amdgpu_buffer_rsrc_t ro =
            rsrc(reinterpret_cast<float*>(a.out) + (int64_t)b * a.out_bstride + (int64_t)rblock * P * a.out_row);
#pragma unroll
        for (int i = 0; i < NBl; ++i)
#pragma unroll
            for (int r = 0; r < (1 << LRl); ++r) {
                const int m = out_pos<L, M - 1, RM>(g + G * i, r);
                const f2 z = v[i * (1 << LRl) + r];
                const bool ok = active && 2 * m + 1 < nx;
                if constexpr (EVEN) {
                    bstore8(z, ro, ok ? (obase + 2 * m) * 4u : kOob);
                } else {
                    const bool okx = active && 2 * m < nx;
                    bstore4(z.x, ro, okx ? (obase + 2 * m) * 4u : kOob);
                    bstore4(z.y, ro, ok ? (obase + 2 * m + 1) * 4u : kOob);
                }
                if (ok) {
                    s1 += (double)z.x + (double)z.y;
                    s2 += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
                } else if (!EVEN && active && 2 * m < nx) {
                    s1 += (double)z.x;
                    s2 += (double)z.x * (double)z.x;
                }
            }
        double* red = reinterpret_cast<double*>(fast_smem);
        c2r_wave_sums<NW>(s1, s2, red);
        double t1 = 0, t2 = 0;
        for (int w = 0; w < NW; ++w) {
            t1 += red[w];
            t2 += red[NW + w];
        }
        const __amdgpu_buffer_rsrc_t rp = rsrc(a.partial + 3 * ((int64_t)b * a.nblocks + rblock));
        const uint32_t po = tid == 0 ? 0u : kOob;
        bstore8d(t1, rp, po);
        bstore8d(t2, rp, po + 8u);
        bstore8d(0.0, rp, po + 16u);
    }
}

// The C2R pass from the y output over a.bpb members per block (grid y: member groups).
template <int L, int RM, bool EVEN>
__global__ __launch_bounds__((C2RCfg<L, RM>::NT)) void c2r_loop_kernel(PassArgs a) {
    const int b0 = blockIdx.y * a.bpb;
    c2r_rows<L, RM, EVEN>(a, blockIdx.x, b0, min(b0 + a.bpb, a.nbatch));
}

// The C2R pass from a z-run y output: X[kx] of row (y, z) is component kx mod 2 of float4
// (z / 8) sZ + y sY + (kx / 2) sP + z mod 8, with sZ, sY, sP = in_sT, in_sA, in_sC in float4
// units.  A block's tile is P = 8 YB rows: the 8 z of one run at YB consecutive y.  Its loads
// read whole 128-byte runs (lanes along z mod 8, then the pair, then y) and put each pair into its
// row in LDS, from where thread (q, g) reads X[k] and X[N-k], k = g + r G, of row q -- the values
// c2r_mirror takes from the lanes of a row.  Operands, radix passes, stores and statistics are
// those of c2r_rows (the values of a row do not depend on its tile); the block sums have their own
// LDS slots, so a member waits at four block barriers: after the rows' writes, after their reads
// (before the first exchange writes), in the exchange, and for the block sums.
template <int L, int RM>
struct C2RZrCfg {
    using C = C2RCfg<L, RM>;
    static constexpr int N = C::N, P = C::P, HP = N / 2, YB = P / 8, NW = (C::NT + 63) / 64;
    static constexpr int NL = P * HP / C::NT;  // 16-byte loads per thread and member
    // row stride of the moved rows, complex: even (16-byte pair writes) and 2 RW = 40 mod 64
    // words, so that the 8 rows of a run written by one instruction fall in distinct banks
    static constexpr int RW = N + 20;
    static constexpr size_t xbuf_elems() {
        return C::lds_elems() > (size_t)P * RW ? C::lds_elems() : (size_t)P * RW;
    }
    static constexpr size_t lds_bytes() {
        return (xbuf_elems() + 2 * (size_t)N) * sizeof(f2) + 2 * (size_t)NW * sizeof(double);
    }
    static constexpr bool ok = C::G <= 64 && C::M > 0 && P % 8 == 0 && (P * HP) % C::NT == 0;
};
template <int L, int RM, bool EVEN>
__device__ __forceinline__ void c2r_rows_zr(const PassArgs& a, const int rblock, const int b0, const int b1) {
    using C = C2RCfg<L, RM>;
    using Z = C2RZrCfg<L, RM>;
    constexpr int N = C::N, EPT = C::EPT, G = C::G, M = C::M, NT = C::NT;
    constexpr int HP = Z::HP, YB = Z::YB, NL = Z::NL, RW = Z::RW, NW = Z::NW;
    static_assert(Z::ok, "rows inside one wave, whole runs per tile");
    f2* lds = reinterpret_cast<f2*>(fast_smem);
    f2* twl = lds + Z::xbuf_elems();  // W_N (N entries), then the first N entries of W_{2N}
    double* red = reinterpret_cast<double*>(twl + 2 * N);
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;  // row q of the tile: y0 + q / 8, z = 8 zb + q mod 8
    const int nyb = a.ny / YB, zb = rblock / nyb, y0 = (rblock - zb * nyb) * YB;
    {
        const f2* __restrict__ tw = reinterpret_cast<const f2*>(a.tw);
        for (int m = tid; m < 2 * N; m += NT) twl[m] = __ldg(tw + m);
    }
    const uint32_t sY = (uint32_t)(a.in_sA >> 1), sP = (uint32_t)(a.in_sC >> 1);
    const int64_t tile0 = 2 * ((int64_t)zb * (a.in_sT >> 1) + (int64_t)y0 * sY);  // complex
    // load j of thread tid: float4 f = j NT + tid of the tile, run position f mod 8, pair
    // (f / 8) mod HP, row y0 + f / (8 HP); the pair goes to LDS row 8 (f / (8 HP)) + f mod 8
    auto src_off = [&](const int j) __attribute__((always_inline)) {
        const int f = j * NT + tid;
        return (uint32_t)((f >> 3) / HP) * sY + (uint32_t)((f >> 3) & (HP - 1)) * sP + (uint32_t)(f & 7);
    };
    auto lds_pos = [&](const int j) __attribute__((always_inline)) {
        const int f = j * NT + tid;
        return (((f >> 3) / HP) * 8 + (f & 7)) * RW + 2 * ((f >> 3) & (HP - 1));
    };
    float4 nxt[NL];
    auto load = [&](int b) {  // b >= b1: zeros, no memory access
        const int bb = min(b, b1 - 1);
        const __amdgpu_buffer_rsrc_t rs =
            rsrc(reinterpret_cast<const f2*>(a.in) + (int64_t)bb * a.in_bstride + tile0);
        const bool ok = b < b1;
#pragma unroll
        for (int j = 0; j < NL; ++j) nxt[j] = bload16(rs, ok ? src_off(j) * 16u : kOob);
    };
    constexpr int LRl = rlog(L, M - 1, RM), NBl = EPT >> LRl;
    load(b0);
    {
        // as many out-of-range (dropped) stores as a member issues (see c2r_rows)
        const __amdgpu_buffer_rsrc_t rs = rsrc(a.out);
#pragma unroll
        for (int e = 0; e < EPT + 3; ++e) bstore8(mk(0.f, 0.f), rs, kOob + 8u * e);
        if constexpr (!EVEN) {
#pragma unroll
            for (int e = 0; e < EPT; ++e) bstore4(0.f, rs, kOob + 4u * e);
        }
    }
    const uint32_t obase = (uint32_t)((8 * zb + (q & 7)) * a.ny + y0 + (q >> 3)) * (uint32_t)a.out_row;
    const int nx = a.n_out;
    const int lane = tid & 63, wid = tid >> 6;
    for (int b = b0; b < b1; ++b) {
        f2 xa[EPT], xb[EPT], v[EPT];
#pragma unroll
        for (int j = 0; j < NL; ++j) *reinterpret_cast<float4*>(lds + lds_pos(j)) = nxt[j];
        load(b + 1);  // in flight while member b is transformed and stored
        __syncthreads();  // the rows are in LDS (and, before the first member, the twiddles)
#pragma unroll
        for (int r = 0; r < EPT; ++r) xa[r] = lds[q * RW + g + r * G];
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
            const int k = g + r * G;
            // X[N] rides in X[0].y (packed Nyquist column)
            xb[r] = k == 0 ? mk(xa[r].y, 0.f) : lds[q * RW + (N - k)];
        }
        if (g == 0) xa[0].y = 0.f;
        __syncthreads();  // the rows are read before the exchange buffer is written
        c2r_operands<EPT, G>(xa, xb, TwLds{twl + N}, g, v);
        c2r_passes<L, RM>(v, lds, q, g, TwLds{twl});
        double s1 = 0.0, s2 = 0.0;
        const __amdgpu_buffer_rsrc_t ro = rsrc(reinterpret_cast<float*>(a.out) + (int64_t)b * a.out_bstride);
#pragma unroll
        for (int i = 0; i < NBl; ++i)
#pragma unroll
            for (int r = 0; r < (1 << LRl); ++r) {
                const int m = out_pos<L, M - 1, RM>(g + G * i, r);
                const f2 z = v[i * (1 << LRl) + r];
                const bool ok = 2 * m + 1 < nx;
                if constexpr (EVEN) {
                    bstore8(z, ro, ok ? (obase + 2 * m) * 4u : kOob);
                } else {
                    const bool okx = 2 * m < nx;
                    bstore4(z.x, ro, okx ? (obase + 2 * m) * 4u : kOob);
                    bstore4(z.y, ro, ok ? (obase + 2 * m + 1) * 4u : kOob);
                }
                if (ok) {
                    s1 += (double)z.x + (double)z.y;
                    s2 += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
                } else if (!EVEN && 2 * m < nx) {
                    s1 += (double)z.x;
                    s2 += (double)z.x * (double)z.x;
                }
            }
        s1 = wsum(s1);
        s2 = wsum(s2);
        if (lane == 0) {
            red[wid] = s1;
            red[NW + wid] = s2;
        }
        __syncthreads();  // the block sums (also: every exchange read is done before the next rows)
        double t1 = 0, t2 = 0;
        for (int w = 0; w < NW; ++w) {
            t1 += red[w];
            t2 += red[NW + w];
        }
        const __amdgpu_buffer_rsrc_t rp = rsrc(a.partial + 3 * ((int64_t)b * a.nblocks + rblock));
        const uint32_t po = tid == 0 ? 0u : kOob;
        bstore8d(t1, rp, po);
        bstore8d(t2, rp, po + 8u);
        bstore8d(0.0, rp, po + 16u);
    }
}

template <int L, int RM, bool EVEN>
__global__ __launch_bounds__((C2RCfg<L, RM>::NT)) void c2r_zr_kernel(PassArgs a) {
    const int b0 = blockIdx.y * a.bpb;
    c2r_rows_zr<L, RM, EVEN>(a, blockIdx.x, b0, min(b0 + a.bpb, a.nbatch));
}

// 1-point x axis (npx = 1): real part of the only coefficient of each row.
template <int UNUSED = 0>
__global__ __launch_bounds__(256) void x1_kernel(PassArgs a) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    double s1 = 0, s2 = 0;
    if (row < a.nrows) {
        f2 x;
        if (a.in) {
            x = reinterpret_cast<const f2*>(a.in)[(int64_t)b * a.in_bstride + (int64_t)row * a.in_row];
        } else {
            x = fill_at(a, column_of(a, __ldg(a.rbase + b), 0), 0, 0);
        }
        reinterpret_cast<float*>(a.out)[(int64_t)b * a.out_bstride + (int64_t)row * a.out_row] = x.x;
        s1 = x.x;
        s2 = (double)x.x * x.x;
    }
    __shared__ double red[2][4];
    s1 = wsum(s1);
    s2 = wsum(s2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s1;
        red[1][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* p = a.partial + 3 * ((int64_t)b * a.nblocks + blockIdx.x);
        p[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        p[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        p[2] = 0.0;
    }
}

}  // namespace fast
}  // namespace ss3d
