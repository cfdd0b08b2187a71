// ss3d_fast64.h -- FP64 realization passes of the drop-in path with compile-time lengths.
//
// The arithmetic is that of the generic pass kernel (ss3d_pass.h) and of the reference's fft1d
// (src/specsim3d.f90:224-251): radix-2 decimation in time, V_s(c, k) kept at position
// c + k (n >> s), butterflies  t = W * up, lo + t, lo - t  with the table twiddle
// W = tw[k n / len], in chunks of up to 3 stages held in registers.  The results are bit for bit
// those of the generic kernel (checked by test); only the data movement differs:
//   - the inputs of the first chunk are read (or, in the x pass, generated) straight into
//     registers and the outputs of the last chunk written straight from registers: LDS is used
//     only for the exchange between chunks;
//   - the twiddle table is copied to LDS once per block;
//   - lengths, tiles and every index are compile-time constants;
//   - tiles in XCD-aware order, outputs stored non-temporally.
// Built in the -ffp-contract=off unit (ss3d_k64.hip), like every FP64 kernel.
#pragma once

#include "ss3d_pass.h"

namespace ss3d {
namespace f64 {

using c64 = cpx<double>;

constexpr int el(int L) { return L < 3 ? L : 3; }                        // log2 elements per thread and line
constexpr int nch(int L) { return L == 0 ? 0 : (L + el(L) - 1) / el(L); }  // chunks of stages
constexpr int ch_s(int L, int j) { return el(L) * j; }                     // stages done before chunk j
constexpr int ch_rc(int L, int j) { return L - el(L) * j < el(L) ? L - el(L) * j : el(L); }
constexpr int strided_P(int L) { return L <= 6 ? 64 : (4096 >> L); }     // columns per strided tile
// rows per contiguous tile: 4 rows of 512 in a 256-thread block; the row-pair kernel holds two
// rows and their draws in registers (158 VGPRs, 3 waves per SIMD), so smaller blocks let three
// of them share a CU (2.03 -> 1.58 ms per 512^2-row launch of 2 realizations)
constexpr int rows_P(int L) { return L <= 3 ? 64 : (L >= 12 ? 1 : (2048 >> L)); }

// twiddle table W_n^m, m < n/2, from HBM or from its LDS copy
struct TwG {
    const c64* __restrict__ p;
    __device__ __forceinline__ c64 operator()(int m) const { return p[m]; }
};
struct TwL {
    const c64* p;
    __device__ __forceinline__ c64 operator()(int m) const { return p[m]; }
};

// Stages S+1 .. S+RC on x = V_S(c0 + a nc, k0), a < 2^RC: the operations of radix_sub.
template <int L, int S, int RC, typename TW>
__device__ __forceinline__ void chunk(c64* x, int k0, TW tw) {
    constexpr int RR = 1 << RC;
#pragma unroll
    for (int t = 0; t < RC; ++t) {
        c64 y[RR];
        const int A = RR >> t, Ah = A >> 1;
#pragma unroll
        for (int be = 0; be < (1 << t); ++be) {
            const int k = k0 + (be << S);
            const c64 w = tw(k << (L - S - t - 1));
#pragma unroll
            for (int al = 0; al < Ah; ++al) {
                const c64 lo = x[al + A * be], up = x[al + Ah + A * be];
                const c64 tt = cmul(w, up);
                y[al + Ah * be] = cadd(lo, tt);
                y[al + Ah * (be + (1 << t))] = csub(lo, tt);
            }
        }
#pragma unroll
        for (int j = 0; j < RR; ++j) x[j] = y[j];
    }
}

// chunk j input / output position of element a (bb) of group w
template <int L, int J>
__device__ __forceinline__ int in_pos(int w, int a) {
    constexpr int S = ch_s(L, J), RC = ch_rc(L, J), LNC = L - S - RC;
    const int c0 = w & ((1 << LNC) - 1), k0 = w >> LNC;
    return c0 + (a << LNC) + (k0 << (L - S));
}
template <int L, int J>
__device__ __forceinline__ int out_pos(int w, int bb) {
    constexpr int S = ch_s(L, J), RC = ch_rc(L, J), LNC = L - S - RC;
    const int c0 = w & ((1 << LNC) - 1), k0 = w >> LNC;
    return c0 + ((k0 + (bb << S)) << LNC);
}

// All chunks on v[0 .. EPT) of one line (thread slot g of G); exchanges through LDS with
// `slot(pos)` the LDS index of position pos of this thread's line.  Before the first chunk v
// holds positions g + a G; after the last, v[i 2^RC + bb] holds position w + (bb << S),
// w = g + G i.
template <int L, int G, typename TW, typename SLOT>
__device__ __forceinline__ void transform(c64* v, int g, c64* lds, SLOT slot, TW tw) {
    constexpr int EPT = 1 << el(L), M = nch(L);
    [&]<int... J>(std::integer_sequence<int, J...>) {
        (
            [&] {
                constexpr int Jc = J, S = ch_s(L, Jc), RC = ch_rc(L, Jc), NB = EPT >> RC;
                if constexpr (Jc > 0) {
                    constexpr int RCp = ch_rc(L, Jc - 1), NBp = EPT >> RCp;
                    if constexpr (Jc > 1) __syncthreads();
#pragma unroll
                    for (int i = 0; i < NBp; ++i)
#pragma unroll
                        for (int bb = 0; bb < (1 << RCp); ++bb)
                            lds[slot(out_pos<L, Jc - 1>(g + G * i, bb))] = v[i * (1 << RCp) + bb];
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int a = 0; a < (1 << RC); ++a) v[i * (1 << RC) + a] = lds[slot(in_pos<L, Jc>(g + G * i, a))];
                }
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    constexpr int LNC = L - S - RC;
                    const int k0 = (g + G * i) >> LNC;
                    chunk<L, S, RC>(v + i * (1 << RC), k0, tw);
                }
            }(),
            ...);
    }(std::make_integer_sequence<int, M>{});
}

// ------------------------------------------------------------ strided passes (y, z): load
// Tile: P consecutive columns x the whole axis.  Thread: column q = tid % P, line slot
// g = tid / P.  OUT_STORE writes the complex result (positions < n_out); OUT_REAL writes its
// real part to the FP64 output with per-block (sum, sum of squares, max|Im|).
template <int L, int OUT>
__global__ __launch_bounds__((1 << L) / (1 << el(L)) * strided_P(L)) void strided_kernel(PassArgs a) {
    constexpr int N = 1 << L, EPT = 1 << el(L), G = N / EPT, P = strided_P(L), NT = G * P, M = nch(L);
    c64* lds = reinterpret_cast<c64*>(ss3d_smem);
    c64* twl = lds + (size_t)N * P;
    const int tid = threadIdx.x;
    const int q = tid % P, g = tid / P;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int i0base = (tile % a.ntiles0) * P;
    const int third = tile / a.ntiles0;
    const int b = blockIdx.y;
    const int i0 = i0base + q;
    const bool active = i0 < a.n0;
    const c64* __restrict__ tw = reinterpret_cast<const c64*>(a.tw);
    if constexpr (M > 0) {
        for (int m = tid; m < N / 2; m += NT) twl[m] = tw[m];
        __syncthreads();
    }
    c64 v[EPT];
    const c64* __restrict__ src =
        reinterpret_cast<const c64*>(a.in) + (int64_t)b * a.in_bstride + (int64_t)third * a.in_sT + i0;
    if (active) {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = src[(int64_t)(g + r * G) * a.in_sA];
    } else {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = c64{0.0, 0.0};
    }
    transform<L, G>(v, g, lds, [&](int pos) { return pos * P + q; }, TwL{twl});
    constexpr int RCl = M ? ch_rc(L, M - 1) : 0, Sl = M ? ch_s(L, M - 1) : 0, NBl = EPT >> RCl;
    const int nout = a.n_out;
    if constexpr (OUT == OUT_STORE) {
        if (active) {
            c64* __restrict__ dst = reinterpret_cast<c64*>(a.out) + (int64_t)b * a.out_bstride +
                                    (int64_t)third * a.out_sT + i0;
#pragma unroll
            for (int i = 0; i < NBl; ++i)
#pragma unroll
                for (int bb = 0; bb < (1 << RCl); ++bb) {
                    const int pos = g + G * i + (bb << Sl);
                    if (pos < nout) {
                        using d2 = double __attribute__((ext_vector_type(2)));
                        const c64 z = v[i * (1 << RCl) + bb];
                        __builtin_nontemporal_store((d2){z.x, z.y}, reinterpret_cast<d2*>(dst + (int64_t)pos * a.out_sA));
                    }
                }
        }
    } else if constexpr (OUT == OUT_SPEC) {
        // spectral density of the last forward pass: S = c/N, amp = sqrt(S) where Re S >= 0, else 0,
        // per-block max|Im S|, max|Re S|, sum of negative and of positive S (f90:1272-1288), the
        // operations of the generic kernel's OUT_SPEC stage
        const double rn = (double)a.npx * (double)a.npy * (double)a.npz;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        // columns kx <= npx/2 alone: column kx stands for itself and its mirror npx - kx
        const bool twice = a.spec_half && i0 != 0 && 2 * i0 != a.npx;
        if (active) {
#pragma unroll
            for (int i = 0; i < NBl; ++i)
#pragma unroll
                for (int bb = 0; bb < (1 << RCl); ++bb) {
                    const int pos = g + G * i + (bb << Sl);
                    const c64 c = v[i * (1 << RCl) + bb];
                    const int j = a.axis == 1 ? pos : third, k = a.axis == 1 ? third : pos;
                    const double sr = c.x / rn, si = c.y / rn;
                    acc[0] = fmax(acc[0], fabs(si));
                    acc[1] = fmax(acc[1], fabs(sr));
                    double amp;
                    const double sw = twice ? sr + sr : sr;
                    if (sr < 0.0) {
                        acc[2] += sw;
                        amp = 0.0;
                    } else {
                        acc[3] += sw;
                        amp = sqrt(sr);
                    }
                    const int64_t row = j + (int64_t)a.npy * k;
                    if (a.amp64_out) a.amp64_out[i0 + (int64_t)a.npx * row] = amp;
                    if (a.amp32h_out && i0 < a.hx) a.amp32h_out[i0 + (int64_t)a.hxa * row] = (float)(amp * kAmp32Scale);
                }
        }
        const int ops[4] = {1, 1, 0, 0};
        block_reduce_write<NT, 4>(acc, ops, a.partial + 4 * ((int64_t)b * a.nblocks + blockIdx.x));
    } else {  // OUT_REAL
        double acc[3] = {0.0, 0.0, 0.0};
        if (active) {
            double* __restrict__ dst = reinterpret_cast<double*>(a.out) + (int64_t)b * a.out_bstride +
                                       (int64_t)third * a.out_sT + i0;
#pragma unroll
            for (int i = 0; i < NBl; ++i)
#pragma unroll
                for (int bb = 0; bb < (1 << RCl); ++bb) {
                    const int pos = g + G * i + (bb << Sl);
                    if (pos < nout) {
                        const c64 z = v[i * (1 << RCl) + bb];
                        __builtin_nontemporal_store(z.x, dst + (int64_t)pos * a.out_sA);
                        acc[0] += z.x;
                        acc[1] += z.x * z.x;
                        acc[2] = fmax(acc[2], fabs(z.y));
                    }
                }
        }
        const int ops[3] = {0, 0, 1};
        block_reduce_write<NT, 3>(acc, ops, a.partial + 3 * ((int64_t)b * a.nblocks + blockIdx.x));
    }
}

// ------------------------------------------------------------ x pass: fill + transform of rows
// Tile: P rows.  Thread: row q = tid / G, line slot g = tid % G.  The coefficients are those of
// the generic kernel's IN_FILL64 stage (classify, site_state, coef64); outputs x < n_out.
template <int L>
__global__ __launch_bounds__((1 << L) / (1 << el(L)) * rows_P(L)) void rows_fill_kernel(PassArgs a) {
    constexpr int N = 1 << L, EPT = 1 << el(L), G = N / EPT, P = rows_P(L), NT = G * P, M = nch(L);
    constexpr int ROW = N + (N >> 3);  // one pad slot per 8 positions
    c64* lds = reinterpret_cast<c64*>(ss3d_smem);
    c64* twl = lds + (size_t)ROW * P;
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * P + q;
    const int b = blockIdx.y;
    const bool active = row < a.nrows;
    const c64* __restrict__ tw = reinterpret_cast<const c64*>(a.tw);
    if constexpr (M > 0) {
        for (int m = tid; m < N / 2; m += NT) twl[m] = tw[m];
        __syncthreads();
    }
    c64 v[EPT];
    if (active) {
        const uint32_t rb = __ldg(a.rbase + b);
        const int j = row % a.npy, k = row / a.npy;
        const double* __restrict__ amp = a.amp64 + (int64_t)N * row;
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
            const int pos = g + r * G;
            const DrawSite d = classify(pos, j, k, N, a.npy, a.npz);
            v[r] = coef64(__ldg(amp + pos), site_state(a, rb, d), a.igauss, d);
        }
    } else {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = c64{0.0, 0.0};
    }
    c64* lrow = lds + (size_t)q * ROW;
    transform<L, G>(v, g, lrow, [](int pos) { return pos + (pos >> 3); }, TwL{twl});
    if (active) {
        constexpr int RCl = M ? ch_rc(L, M - 1) : 0, Sl = M ? ch_s(L, M - 1) : 0, NBl = EPT >> RCl;
        c64* __restrict__ dst = reinterpret_cast<c64*>(a.out) + (int64_t)b * a.out_bstride + (int64_t)row * a.out_row;
        const int nout = a.n_out;
#pragma unroll
        for (int i = 0; i < NBl; ++i)
#pragma unroll
            for (int bb = 0; bb < (1 << RCl); ++bb) {
                const int pos = g + G * i + (bb << Sl);
                if (pos < nout) {
                    using d2 = double __attribute__((ext_vector_type(2)));
                    const c64 z = v[i * (1 << RCl) + bb];
                    __builtin_nontemporal_store((d2){z.x, z.y}, reinterpret_cast<d2*>(dst + pos));
                }
            }
    }
}

// ------------------------------------------------------------ x pass over Hermitian row pairs
// Row R = j + npy k and its partner row Rp = (-j, -k): for R < Rp every point of R draws and
// the point (-i) of Rp takes the conjugate of that draw (f90:1305-1334).  A unit of G threads
// (a line slot each) computes the draws of R once, applies the amplitudes of both rows, and
// transforms R, then Rp, in its LDS row.  The partner's value for slot g is the draw of slot
// G-g (register EPT-1-r), moved across the unit's lanes (the unit lies inside one wave).
// Self rows (R = Rp: j in {0, npy/2}, k in {0, npz/2}) take the pointwise path in blocks of
// their own.  Units: lower rows enumerated in closed form, self rows after them.
template <int L>
__global__ __launch_bounds__((1 << L) / (1 << el(L)) * rows_P(L)) void rows_pair_kernel(PassArgs a) {
    constexpr int N = 1 << L, EPT = 1 << el(L), G = N / EPT, P = rows_P(L), NT = G * P, M = nch(L);
    static_assert(G <= 64, "a unit lies inside one wave");
    constexpr int ROW = N + (N >> 3);
    c64* lds = reinterpret_cast<c64*>(ss3d_smem);
    c64* twl = lds + (size_t)ROW * P;
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;
    const int npy = a.npy, npz = a.npz;
    const int nlow = npy >= 2 ? npy / 2 - 1 : 0, nsp = npz >= 2 ? 2 : 1;
    const int npairs = nsp * nlow + (npz >= 2 ? (npz / 2 - 1) * npy : 0);
    const int nbp = (npairs + P - 1) / P;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const bool pairs = blk < nbp;  // block-uniform
    const int u = (pairs ? blk : blk - nbp) * P + q;
    const int nself = (npy >= 2 ? 2 : 1) * nsp;
    const bool active = pairs ? u < npairs : u < nself;
    const int b = blockIdx.y;
    const c64* __restrict__ tw = reinterpret_cast<const c64*>(a.tw);
    if constexpr (M > 0) {
        for (int m = tid; m < N / 2; m += NT) twl[m] = tw[m];
        __syncthreads();
    }
    int j = 0, k = 0;
    if (pairs) {
        if (u < nsp * nlow) {
            k = u / nlow == 0 ? 0 : npz / 2;
            j = 1 + u % nlow;
        } else {
            const int u2 = u - nsp * nlow;
            k = 1 + u2 / npy;
            j = u2 % npy;
        }
    } else {
        const int ny2 = npy >= 2 ? 2 : 1;
        j = (u % ny2) * (npy / 2);
        k = (u / ny2) * (npz / 2);
    }
    const int row = j + npy * k;
    const int rowp = ((npy - j) & (npy - 1)) + npy * ((npz - k) & (npz - 1));
    const uint32_t rb = __ldg(a.rbase + b);
    c64 v[EPT], w[EPT];
    if (active) {
        const double* __restrict__ amp = a.amp64 + (int64_t)N * row;
        if (pairs) {
            const double* __restrict__ ampp = a.amp64 + (int64_t)N * rowp;
            const uint32_t tr = pm_mul(rb, __ldg(a.rowf + row));
            Draw64 z[EPT];
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                const int pos = g + r * G;
                z[r] = draw64(pm_mul(tr, __ldg(a.txp + pos)), a.igauss);
                v[r] = coef_of_draw(__ldg(amp + pos), z[r], a.igauss, false);
            }
            const int lane = threadIdx.x & 63;
            const int plane = lane - g + ((G - g) & (G - 1));
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                const Draw64 t = z[EPT - 1 - r];
                const Draw64 sh{__shfl(t.a, plane, 64), __shfl(t.b, plane, 64)};
                const Draw64 zp = g != 0 ? sh : z[(EPT - r) & (EPT - 1)];
                w[r] = coef_of_draw(__ldg(ampp + g + r * G), zp, a.igauss, true);
            }
        } else {
#pragma unroll
            for (int r = 0; r < EPT; ++r) {
                const int pos = g + r * G;
                const DrawSite d = classify(pos, j, k, N, npy, npz);
                v[r] = coef64(__ldg(amp + pos), site_state(a, rb, d), a.igauss, d);
                w[r] = c64{0.0, 0.0};
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = w[r] = c64{0.0, 0.0};
    }
    c64* lrow = lds + (size_t)q * ROW;
    constexpr int RCl = M ? ch_rc(L, M - 1) : 0, Sl = M ? ch_s(L, M - 1) : 0, NBl = EPT >> RCl;
    const int nout = a.n_out;
    auto finish = [&](c64* x, int rw) {
        transform<L, G>(x, g, lrow, [](int pos) { return pos + (pos >> 3); }, TwL{twl});
        if (active) {
            c64* __restrict__ dst = reinterpret_cast<c64*>(a.out) + (int64_t)b * a.out_bstride + (int64_t)rw * a.out_row;
#pragma unroll
            for (int i = 0; i < NBl; ++i)
#pragma unroll
                for (int bb = 0; bb < (1 << RCl); ++bb) {
                    const int pos = g + G * i + (bb << Sl);
                    if (pos < nout) {
                        using d2 = double __attribute__((ext_vector_type(2)));
                        const c64 zz = x[i * (1 << RCl) + bb];
                        __builtin_nontemporal_store((d2){zz.x, zz.y}, reinterpret_cast<d2*>(dst + pos));
                    }
                }
        }
    };
    finish(v, row);
    if (pairs) {  // block-uniform
        __syncthreads();  // LDS reads of row R done
        finish(w, rowp);
    }
}

// ------------------------------------------------------------ x pass of the spectral density
// Rows of the covariance grid generated in registers at the minimum-image lags (f90:1256-1268,
// lag n/2 stays +n/2), as the generic kernel's IN_COV stage does, then transformed with the
// table in a.tw (forward) and stored complete.
template <int L>
__global__ __launch_bounds__((1 << L) / (1 << el(L)) * rows_P(L)) void rows_cov_kernel(PassArgs a) {
    constexpr int N = 1 << L, EPT = 1 << el(L), G = N / EPT, P = rows_P(L), NT = G * P, M = nch(L);
    constexpr int ROW = N + (N >> 3);
    c64* lds = reinterpret_cast<c64*>(ss3d_smem);
    c64* twl = lds + (size_t)ROW * P;
    const int tid = threadIdx.x;
    const int g = tid % G, q = tid / G;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * P + q;
    const bool active = row < a.nrows;
    const c64* __restrict__ tw = reinterpret_cast<const c64*>(a.tw);
    if constexpr (M > 0) {
        for (int m = tid; m < N / 2; m += NT) twl[m] = tw[m];
        __syncthreads();
    }
    c64 v[EPT];
    if (active) {
        const int j = row % a.npy, k = row / a.npy;
        const int jp = j > a.npy / 2 ? j - a.npy : j;
        const int kp = k > a.npz / 2 ? k - a.npz : k;
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
            const int pos = g + r * G;
            const int ip = pos > a.npx / 2 ? pos - a.npx : pos;
            v[r] = c64{cov64(a, (double)ip * a.dx, (double)jp * a.dy, (double)kp * a.dz), 0.0};
        }
    } else {
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = c64{0.0, 0.0};
    }
    c64* lrow = lds + (size_t)q * ROW;
    transform<L, G>(v, g, lrow, [](int pos) { return pos + (pos >> 3); }, TwL{twl});
    if (active) {
        constexpr int RCl = M ? ch_rc(L, M - 1) : 0, Sl = M ? ch_s(L, M - 1) : 0, NBl = EPT >> RCl;
        c64* __restrict__ dst = reinterpret_cast<c64*>(a.out) + (int64_t)row * a.out_row;
        const int nout = a.n_out;
#pragma unroll
        for (int i = 0; i < NBl; ++i)
#pragma unroll
            for (int bb = 0; bb < (1 << RCl); ++bb) {
                const int pos = g + G * i + (bb << Sl);
                if (pos < nout) {
                    using d2 = double __attribute__((ext_vector_type(2)));
                    const c64 z = v[i * (1 << RCl) + bb];
                    __builtin_nontemporal_store((d2){z.x, z.y}, reinterpret_cast<d2*>(dst + pos));
                }
            }
    }
}

// dynamic LDS of a tile: exchange buffer + twiddle table
constexpr size_t strided_lds(int L) { return ((size_t)(1 << L) * strided_P(L) + (1 << L) / 2) * sizeof(c64); }
constexpr size_t rows_lds(int L) { return ((size_t)((1 << L) + ((1 << L) >> 3)) * rows_P(L) + (1 << L) / 2) * sizeof(c64); }

}  // namespace f64
}  // namespace ss3d
