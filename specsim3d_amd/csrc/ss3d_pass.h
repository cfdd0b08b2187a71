// ss3d_pass.h -- the one FFT-pass kernel template of the engine, instantiated for FP64
// (ss3d_k64.hip, built with FMA contraction off) and FP32 (ss3d_k32.hip).
//
// A pass transforms whole pencils along one grid axis inside one workgroup:
//   1. an input stage fills the LDS tile in natural order: loaded from HBM, or generated in
//      registers (covariance model, Hermitian coefficient fill);
//   2. fft_core runs the radix-2 decimation-in-time recursion of the reference
//      (src/specsim3d.f90:208-241) in chunks of RC radix-2 stages held in registers, with one
//      LDS exchange per chunk;
//   3. an output stage writes the pruned result: complex to HBM, spectral amplitude, or the real
//      part of the cropped field with block statistics.
//
// Tiles: pass along x (axis 0) takes P = 2^logp whole rows (contiguous in HBM).  Passes along
// y or z take P consecutive x-columns of the pencil axis (P*16 or P*8 contiguous bytes per row),
// so every HBM access of a wavefront is a run of whole segments.
//
// Butterfly bookkeeping.  With V_s(c, k) the k-th DFT coefficient of the decimated subsequence
// x[c], x[c + n/2^s], ..., the reference's stage s+1 (len = 2^(s+1)) computes
//     V_{s+1}(c, k)       = V_s(c, k) + w_k V_s(c + n/2^(s+1), k)
//     V_{s+1}(c, k + 2^s) = V_s(c, k) - w_k V_s(c + n/2^(s+1), k),   w_k = tw[k * n/2^(s+1)]
// (u + t, u - t with t = tw(k*step+1)*x(i+k+half), f90:233-238).  V_s(c, k) is kept at position
// c + k*(n >> s): input and output are in natural order, and every butterfly is the reference's
// own (same operands, same twiddle value, same operation order), so an FP64 pass compiled
// without contraction reproduces the reference transform bit for bit.
#pragma once

#include "ss3d_exp.h"
#include "ss3d_internal.h"

namespace ss3d {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) {
    return {a.x - b.x, a.y - b.y};
}
// complex product w*x in the reference's order: (wr*xr - wi*xi, wr*xi + wi*xr)
template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> w, cpx<R> x) {
    return {w.x * x.x - w.y * x.y, w.x * x.y + w.y * x.x};
}

__device__ __forceinline__ int lds_slot(int idx) { return idx + (idx >> kPadShift); }

// ---------------------------------------------------------------- Park-Miller on the device
// state' = a*state mod (2^31-1) (f90:193), product folded with 2^31 == 1 (mod 2^31-1).
__device__ __forceinline__ uint32_t pm_mul(uint32_t a, uint32_t b) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    uint32_t s = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
    return s >= kPmM ? s - kPmM : s;
}

// Fill-order classification of (i,j,k) (f90:1307-1333): the coefficient comes from the draw of
// the point itself (lin < linp), is self-conjugate (lin == linp), or is the conjugate of the
// partner's (lin > linp).  Rows R = j + npy*k are compared first: lin < linp <=> R < Rp, or
// R == Rp and i < ip.
struct DrawSite {
    int i, j, k;
    bool conj, selfc;
};

__device__ __forceinline__ DrawSite classify(int i, int j, int k, int npx, int npy, int npz) {
    int ip = i ? npx - i : 0, jp = j ? npy - j : 0, kp = k ? npz - k : 0;
    int r = j + npy * k, rp = jp + npy * kp;
    DrawSite d{i, j, k, false, false};
    if (r > rp) {
        d.i = ip; d.j = jp; d.k = kp; d.conj = true;
    } else if (r == rp) {
        if (i > ip) {
            d.i = ip; d.conj = true;
        } else if (i == ip) {
            d.selfc = true;
        }
    }
    return d;
}

// Park-Miller state of the first draw of a drawing site in realization with base rb.
__device__ __forceinline__ uint32_t site_state(const PassArgs& a, uint32_t rb, const DrawSite& d) {
    uint32_t row = __ldg(a.rowf + (d.j + (int64_t)a.npy * d.k));
    return pm_mul(pm_mul(rb, row), __ldg(a.txp + d.i));
}

// The draw of a site, before its amplitude: igauss=0 (cos, sin) of the phase 2 pi u;
// igauss=1 the Box-Muller pair (g1, g2) (f90:1309-1332, gauss_pair f90:200-205).
struct Draw64 {
    double a, b;
};
__device__ __forceinline__ Draw64 draw64(uint32_t st, int igauss) {
    // u = state/2147483647 (f90:199) as a product with the rounded reciprocal, and
    // cos/sin(2 pi u) through sincospi (exact argument reduction): both within an ulp or two
    // of the reference's division and cos(2.0d0*pi*u), far inside the FP64 path's tolerance
    double sn, cs;
    if (!igauss) {
        const double u = (double)st * (1.0 / 2147483647.0);
        sincospi(2.0 * u, &sn, &cs);
        return {cs, sn};
    }
    const double u1 = (double)st * (1.0 / 2147483647.0);
    const double u2 = (double)pm_mul(st, kPmA) * (1.0 / 2147483647.0);
    const double rr = sqrt(-2.0 * log(u1));
    sincospi(2.0 * u2, &sn, &cs);
    return {rr * cs, rr * sn};
}

// Coefficient of a point that is not self-conjugate from its site's draw: amp (cos, +-sin) or
// amp (g1, +-g2) sqrt(0.5), the sign negative at the partner of the drawing point (conj).
__device__ __forceinline__ cpx<double> coef_of_draw(double amp, Draw64 z, int igauss, bool conj) {
    const double s05 = 0.70710678118654752440;  // sqrt(0.5d0)
    cpx<double> c;
    if (!igauss) {
        c.x = amp * z.a;
        c.y = conj ? amp * (-z.b) : amp * z.b;
    } else {
        c.x = amp * z.a * s05;
        c.y = conj ? amp * (-z.b) * s05 : amp * z.b * s05;
    }
    return c;
}

// FP64 coefficient exactly as f90:1309-1332 (own amplitude at every point).
__device__ __forceinline__ cpx<double> coef64(double amp, uint32_t st, int igauss, const DrawSite& d) {
    if (d.selfc) {
        cpx<double> c;
        if (!igauss) {
            // urand() < 0.5d0  <=>  state <= 1073741823 (exact integer form)
            c.x = (st <= 1073741823u) ? amp : -amp;
        } else {
            c.x = amp * draw64(st, 1).a;
        }
        c.y = 0.0;
        return c;
    }
    return coef_of_draw(amp, draw64(st, igauss), igauss, d.conj);
}

// Covariance at a lag, f90:153-178, in the reference's operation order, with the exp of the
// reference build's libm (ss3d_exp.h), so that the covariance grid is the reference's bit for bit.
__device__ __forceinline__ double cov64(const PassArgs& a, double hx, double hy, double hz) {
    double c = 0.0;
    if (hx == 0.0 && hy == 0.0 && hz == 0.0) c = a.nugget;
    for (int s = 0; s < a.nst; ++s) {
        const StructDev& st = a.st[s];
        double rx = st.rot[0] * hx + st.rot[1] * hy + st.rot[2] * hz;
        double ry = st.rot[3] * hx + st.rot[4] * hy + st.rot[5] * hz;
        double rz = st.rot[6] * hx + st.rot[7] * hy + st.rot[8] * hz;
        double h = sqrt(rx * rx + ry * ry + rz * rz);
        double hr = h / st.ahmax;
        if (st.itype == 1) {
            if (hr < 1.0) c = c + st.sill * (1.0 - 1.5 * hr + 0.5 * (hr * hr * hr));
        } else if (st.itype == 2) {
            c = c + st.sill * libm::exp(-3.0 * hr);
        } else {
            c = c + st.sill * libm::exp(-3.0 * hr * hr);
        }
    }
    return c;
}

// ---------------------------------------------------------------- FFT core
// RC radix-2 stages s+1 .. s+RC on the 2^RC inputs V_s(c0 + a*nc, k0), a < 2^RC, nc = n>>(s+RC).
// Register x[al + A*be] holds V(c0 + al*nc, k0 + be*2^s) with A = 2^RC >> t after t stages.
template <typename R, int RC>
__device__ __forceinline__ void radix_sub(cpx<R> (&x)[1 << RC], int k0, int s, int logn, const cpx<R>* __restrict__ tw) {
    constexpr int RR = 1 << RC;
#pragma unroll
    for (int t = 0; t < RC; ++t) {
        cpx<R> y[RR];
        const int A = RR >> t, Ah = A >> 1;
#pragma unroll
        for (int be = 0; be < (1 << t); ++be) {
            const int k = k0 + (be << s);
            const cpx<R> w = tw[k << (logn - s - t - 1)];
#pragma unroll
            for (int al = 0; al < Ah; ++al) {
                cpx<R> lo = x[al + A * be], up = x[al + Ah + A * be];
                cpx<R> tt = cmul(w, up);
                y[al + Ah * be] = cadd(lo, tt);
                y[al + Ah * (be + (1 << t))] = csub(lo, tt);
            }
        }
#pragma unroll
        for (int j = 0; j < RR; ++j) x[j] = y[j];
    }
}

template <typename R, int NT, int RC>
__device__ __forceinline__ void fft_chunk(cpx<R>* lds, int logn, int logp, bool strided,
                                          const cpx<R>* __restrict__ tw, int s) {
    constexpr int RR = 1 << RC;
    constexpr int GPT = kEPT / RR;
    cpx<R> v[GPT][RR];
    int c0s[GPT], k0s[GPT], qs[GPT];
    const int lognc = logn - s - RC;
    const int tid = threadIdx.x;
#pragma unroll
    for (int gi = 0; gi < GPT; ++gi) {
        const int g = tid + gi * NT;
        int q, w;
        if (strided) {
            q = g & ((1 << logp) - 1);
            w = g >> logp;
        } else {
            w = g & ((1 << (logn - RC)) - 1);
            q = g >> (logn - RC);
        }
        const int c0 = w & ((1 << lognc) - 1);
        const int k0 = w >> lognc;
        c0s[gi] = c0;
        k0s[gi] = k0;
        qs[gi] = q;
#pragma unroll
        for (int al = 0; al < RR; ++al) {
            const int pos = c0 + (al << lognc) + (k0 << (logn - s));
            const int idx = strided ? ((pos << logp) | q) : ((q << logn) | pos);
            v[gi][al] = lds[lds_slot(idx)];
        }
        radix_sub<R, RC>(v[gi], k0, s, logn, tw);
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < GPT; ++gi) {
#pragma unroll
        for (int bb = 0; bb < RR; ++bb) {
            const int pos = c0s[gi] + ((k0s[gi] + (bb << s)) << lognc);
            const int idx = strided ? ((pos << logp) | qs[gi]) : ((qs[gi] << logn) | pos);
            lds[lds_slot(idx)] = v[gi][bb];
        }
    }
    __syncthreads();
}

template <typename R, int NT>
__device__ __forceinline__ void fft_core(cpx<R>* lds, int logn, int logp, bool strided, const cpx<R>* __restrict__ tw) {
    constexpr int RL = sizeof(R) == 8 ? 3 : 4;  // radix-8 chunks in FP64, radix-16 in FP32
    int s = 0;
    while (s < logn) {
        const int rc = min(RL, logn - s);
        switch (rc) {
            case 1: fft_chunk<R, NT, 1>(lds, logn, logp, strided, tw, s); break;
            case 2: fft_chunk<R, NT, 2>(lds, logn, logp, strided, tw, s); break;
            case 3: fft_chunk<R, NT, 3>(lds, logn, logp, strided, tw, s); break;
            default:
                if constexpr (RL >= 4) fft_chunk<R, NT, 4>(lds, logn, logp, strided, tw, s);
                break;
        }
        s += rc;
    }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// K values per thread reduced over the block (ops: 0 sum, 1 max), thread 0 writes them.
template <int NT, int K>
__device__ __forceinline__ void block_reduce_write(double (&v)[K], const int (&ops)[K], double* dst) {
    __shared__ double red[K][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < K; ++c) {
        v[c] = ops[c] ? wave_max(v[c]) : wave_sum(v[c]);
        if (lane == 0) red[c][wid] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double r = red[c][0];
            for (int w = 1; w < NT / 64; ++w) r = ops[c] ? fmax(r, red[c][w]) : r + red[c][w];
            dst[c] = r;
        }
    }
}

extern __shared__ __attribute__((aligned(16))) unsigned char ss3d_smem[];

// ---------------------------------------------------------------- the pass kernel
template <typename R, int NT, int IN, int OUT>
__global__ __launch_bounds__(NT) void pass_kernel(PassArgs a) {
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(ss3d_smem);
    const bool strided = a.axis != 0;
    const int logn = a.logn, logp = a.logp;
    const int n = 1 << logn, P = 1 << logp;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int b = blockIdx.y;
    const int tid = threadIdx.x;
    int i0base = 0, third = 0, row0 = 0;
    if (strided) {
        i0base = (tile % a.ntiles0) * P;
        third = tile / a.ntiles0;
    } else {
        row0 = tile * P;
    }
    uint32_t rb = 0;
    if constexpr (IN == IN_FILL64) rb = __ldg(a.rbase + b);

    // ---- input stage
#pragma unroll 4
    for (int it = 0; it < kEPT; ++it) {
        const int e = tid + it * NT;
        int q, pos;
        if (strided) {
            q = e & (P - 1);
            pos = e >> logp;
        } else {
            pos = e & (n - 1);
            q = e >> logn;
        }
        cpx<R> v{R(0), R(0)};
        if (strided) {
            const int i0 = i0base + q;
            if (i0 < a.n0) {
                if constexpr (IN == IN_LOAD) {
                    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(a.in) + (int64_t)b * a.in_bstride;
                    v = src[i0 + pos * a.in_sA + third * a.in_sT];
                }
            }
        } else {
            const int row = row0 + q;
            if (row < a.nrows) {
                if constexpr (IN == IN_LOAD) {
                    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(a.in) + (int64_t)b * a.in_bstride;
                    v = src[(int64_t)row * a.in_row + pos];
                } else if constexpr (IN == IN_COV) {
                    // minimum-image lags, f90:1261-1265 (n/2 keeps +n/2)
                    const int j = row % a.npy, k = row / a.npy;
                    const int ip = pos > a.npx / 2 ? pos - a.npx : pos;
                    const int jp = j > a.npy / 2 ? j - a.npy : j;
                    const int kp = k > a.npz / 2 ? k - a.npz : k;
                    v.x = (R)cov64(a, (double)ip * a.dx, (double)jp * a.dy, (double)kp * a.dz);
                } else if constexpr (IN == IN_FILL64) {
                    const int j = row % a.npy, k = row / a.npy;
                    const double amp = a.amp64[pos + (int64_t)a.npx * row];
                    const DrawSite d = classify(pos, j, k, a.npx, a.npy, a.npz);
                    if constexpr (sizeof(R) == 8) v = coef64(amp, site_state(a, rb, d), a.igauss, d);
                }
            }
        }
        lds[lds_slot(e)] = v;
    }
    __syncthreads();

    // ---- transform
    fft_core<R, NT>(lds, logn, logp, strided, reinterpret_cast<const cpx<R>*>(a.tw));

    // ---- output stage
    if constexpr (OUT == OUT_STORE) {
        cpx<R>* dst = reinterpret_cast<cpx<R>*>(a.out) + (int64_t)b * a.out_bstride;
#pragma unroll 4
        for (int it = 0; it < kEPT; ++it) {
            const int e = tid + it * NT;
            if (strided) {
                const int q = e & (P - 1), pos = e >> logp;
                const int i0 = i0base + q;
                if (i0 < a.n0 && pos < a.n_out) dst[i0 + pos * a.out_sA + third * a.out_sT] = lds[lds_slot(e)];
            } else {
                const int pos = e & (n - 1), row = row0 + (e >> logn);
                if (row < a.nrows && pos < a.n_out) dst[(int64_t)row * a.out_row + pos] = lds[lds_slot(e)];
            }
        }
    } else if constexpr (OUT == OUT_SPEC) {
        // S = c/N, amp = sqrt(S) where Re S >= 0, else 0 (f90:1274-1288)
        const double rn = (double)a.npx * (double)a.npy * (double)a.npz;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // max|Im|, max|Re|, sneg, spos
#pragma unroll 4
        for (int it = 0; it < kEPT; ++it) {
            const int e = tid + it * NT;
            int i, j, k;
            bool ok;
            if (strided) {
                const int q = e & (P - 1), pos = e >> logp;
                i = i0base + q;
                ok = i < a.n0;
                if (a.axis == 1) { j = pos; k = third; } else { j = third; k = pos; }
            } else {
                const int pos = e & (n - 1), row = row0 + (e >> logn);
                ok = row < a.nrows;
                i = pos; j = row % a.npy; k = row / a.npy;
            }
            if (!ok) continue;
            const cpx<R> c = lds[lds_slot(e)];
            const double sr = (double)c.x / rn, si = (double)c.y / rn;
            acc[0] = fmax(acc[0], fabs(si));
            acc[1] = fmax(acc[1], fabs(sr));
            double amp;
            if (sr < 0.0) {
                acc[2] += sr;
                amp = 0.0;
            } else {
                acc[3] += sr;
                amp = sqrt(sr);
            }
            const int64_t row = j + (int64_t)a.npy * k;
            if (a.amp64_out) a.amp64_out[i + (int64_t)a.npx * row] = amp;
            // FP32 half spectrum: amplitude times sqrt(ln 2), the factor of the FP32 Gaussian law
            if (a.amp32h_out && i < a.hx) a.amp32h_out[i + (int64_t)a.hxa * row] = (float)(amp * kAmp32Scale);
        }
        const int ops[4] = {1, 1, 0, 0};
        block_reduce_write<NT, 4>(acc, ops, a.partial + 4 * ((int64_t)b * a.nblocks + blockIdx.x));
    } else {  // OUT_REAL: real part of the cropped corner + statistics
        R* dst = reinterpret_cast<R*>(a.out) + (int64_t)b * a.out_bstride;
        double acc[3] = {0.0, 0.0, 0.0};  // sum, sum of squares, max|Im|
#pragma unroll 4
        for (int it = 0; it < kEPT; ++it) {
            const int e = tid + it * NT;
            int64_t o;
            bool ok;
            if (strided) {
                const int q = e & (P - 1), pos = e >> logp;
                const int i0 = i0base + q;
                ok = i0 < a.n0 && pos < a.n_out;
                o = i0 + pos * a.out_sA + third * a.out_sT;
            } else {
                const int pos = e & (n - 1), row = row0 + (e >> logn);
                ok = row < a.nrows && pos < a.n_out;
                o = (int64_t)row * a.out_row + pos;
            }
            if (!ok) continue;
            const cpx<R> c = lds[lds_slot(e)];
            dst[o] = c.x;
            const double xv = (double)c.x;
            acc[0] += xv;
            acc[1] += xv * xv;
            acc[2] = fmax(acc[2], fabs((double)c.y));
        }
        const int ops[3] = {0, 0, 1};
        block_reduce_write<NT, 3>(acc, ops, a.partial + 3 * ((int64_t)b * a.nblocks + blockIdx.x));
    }
}

// Launch with the dynamic LDS the tile needs; raises the per-kernel LDS limit once when the
// tile needs more than the default 64 KB.
template <typename R, int NT, int IN, int OUT>
hipError_t launch_one(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    const size_t lds = pass_lds_bytes(NT, sizeof(cpx<R>));
    static bool attr_done = false;
    if (lds > 65536 && !attr_done) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pass_kernel<R, NT, IN, OUT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_done = true;
    }
    hipLaunchKernelGGL((pass_kernel<R, NT, IN, OUT>), dim3(nblocks, nbatch), dim3(NT), lds, s, a);
    return hipGetLastError();
}

template <typename R, int IN, int OUT>
hipError_t launch_nt(int nt, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    switch (nt) {
        case 64: return launch_one<R, 64, IN, OUT>(nblocks, nbatch, a, s);
        case 128: return launch_one<R, 128, IN, OUT>(nblocks, nbatch, a, s);
        case 256: return launch_one<R, 256, IN, OUT>(nblocks, nbatch, a, s);
        case 512: return launch_one<R, 512, IN, OUT>(nblocks, nbatch, a, s);
        case 1024:
            if constexpr (sizeof(R) == 4) return launch_one<R, 1024, IN, OUT>(nblocks, nbatch, a, s);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ss3d
