// ss3d_k64.hip -- FP64 instantiations of the pass kernel (spectral density and the drop-in
// realization path) and the reduction finalizers.  Built with -ffp-contract=off: every
// product and sum is rounded separately, as in the reference build, so butterflies,
// covariance and coefficients follow the reference's arithmetic.
#include "ss3d_fast64.h"

namespace ss3d {

// ---------------------------------------------------------------- compile-time FP64 passes
namespace {
template <typename K>
hipError_t launch_f64(K k, size_t lds, int nt, int nblocks, int nbatch, const PassArgs& a, hipStream_t s, bool& attr) {
    if (lds > 65536 && !attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k, dim3(nblocks, nbatch), dim3(nt), lds, s, a);
    return hipGetLastError();
}

template <int L>
hipError_t f64_L(int kind, int out, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    constexpr int G = (1 << L) >> f64::el(L);
    if (kind == 4) {
        if constexpr (G <= 64) {  // row pairs: one draw per Hermitian pair
            static bool attr = false;
            return launch_f64(f64::rows_pair_kernel<L>, f64::rows_lds(L), G * f64::rows_P(L), nblocks, nbatch, a, s,
                              attr);
        } else {
            static bool attr = false;
            return launch_f64(f64::rows_fill_kernel<L>, f64::rows_lds(L), G * f64::rows_P(L), nblocks, nbatch, a, s,
                              attr);
        }
    }
    if (kind == 6) {
        static bool attr = false;
        return launch_f64(f64::rows_cov_kernel<L>, f64::rows_lds(L), G * f64::rows_P(L), nblocks, nbatch, a, s, attr);
    }
    if (out == OUT_SPEC) {
        static bool attr = false;
        return launch_f64(f64::strided_kernel<L, OUT_SPEC>, f64::strided_lds(L), G * f64::strided_P(L), nblocks,
                          nbatch, a, s, attr);
    }
    if (out == OUT_STORE) {
        static bool attr = false;
        return launch_f64(f64::strided_kernel<L, OUT_STORE>, f64::strided_lds(L), G * f64::strided_P(L), nblocks,
                          nbatch, a, s, attr);
    }
    static bool attr = false;
    return launch_f64(f64::strided_kernel<L, OUT_REAL>, f64::strided_lds(L), G * f64::strided_P(L), nblocks, nbatch,
                      a, s, attr);
}

template <int... Ls>
hipError_t f64_dispatch(int L, int kind, int out, int nblocks, int nbatch, const PassArgs& a, hipStream_t s,
                        std::integer_sequence<int, Ls...>) {
    hipError_t e = hipErrorInvalidValue;
    ((L == Ls + 1 ? (e = f64_L<Ls + 1>(kind, out, nblocks, nbatch, a, s), 0) : 0), ...);
    return e;
}
}  // namespace

int fast64_tile(int kind, int L) { return kind == 4 || kind == 6 ? f64::rows_P(L) : f64::strided_P(L); }

// blocks of the x pass: row-pair kernel (lengths whose rows fit one wave) or one row per unit
int fast64_rows_blocks(int L, int npy, int npz) {
    const int P = f64::rows_P(L);
    const int G = (1 << L) >> f64::el(L);
    if (G > 64) return (npy * npz + P - 1) / P;
    const int nlow = npy >= 2 ? npy / 2 - 1 : 0, nsp = npz >= 2 ? 2 : 1;
    const int npairs = nsp * nlow + (npz >= 2 ? (npz / 2 - 1) * npy : 0);
    const int nself = (npy >= 2 ? 2 : 1) * nsp;
    return (npairs + P - 1) / P + (nself + P - 1) / P;
}

hipError_t launch_fast64(int kind, int L, int out, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    if (L < 1 || L > kMaxFast64Log) return hipErrorInvalidValue;
    if (kind == 5 && out != OUT_STORE && out != OUT_REAL && out != OUT_SPEC) return hipErrorInvalidValue;
    if (kind == 6 && out != OUT_STORE) return hipErrorInvalidValue;
    return f64_dispatch(L, kind, out, nblocks, nbatch, a, s, std::make_integer_sequence<int, kMaxFast64Log>{});
}

hipError_t launch_pass64(int in, int out, int nt, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    using R = double;
    if (in == IN_COV && out == OUT_STORE) return launch_nt<R, IN_COV, OUT_STORE>(nt, nblocks, nbatch, a, s);
    if (in == IN_COV && out == OUT_SPEC) return launch_nt<R, IN_COV, OUT_SPEC>(nt, nblocks, nbatch, a, s);
    if (in == IN_LOAD && out == OUT_STORE) return launch_nt<R, IN_LOAD, OUT_STORE>(nt, nblocks, nbatch, a, s);
    if (in == IN_LOAD && out == OUT_SPEC) return launch_nt<R, IN_LOAD, OUT_SPEC>(nt, nblocks, nbatch, a, s);
    if (in == IN_LOAD && out == OUT_REAL) return launch_nt<R, IN_LOAD, OUT_REAL>(nt, nblocks, nbatch, a, s);
    if (in == IN_FILL64 && out == OUT_STORE) return launch_nt<R, IN_FILL64, OUT_STORE>(nt, nblocks, nbatch, a, s);
    if (in == IN_FILL64 && out == OUT_REAL) return launch_nt<R, IN_FILL64, OUT_REAL>(nt, nblocks, nbatch, a, s);
    return hipErrorInvalidValue;
}

// Per batch member: fixed-order reduction of the per-block (sum, sum of squares, max|Im|)
// partials -> mean, population variance, max|Im| (f90:1338-1343).
__global__ __launch_bounds__(256) void stats_finalize_kernel(const double* __restrict__ partial, int nblocks,
                                                             double ncell, double* __restrict__ stats) {
    const int b = blockIdx.x;
    const double* p = partial + 3 * (int64_t)b * nblocks;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblocks; i += 256) {
        v[0] += p[3 * i];
        v[1] += p[3 * i + 1];
        v[2] = fmax(v[2], p[3 * i + 2]);
    }
    double out[3];
    const int ops[3] = {0, 0, 1};
    block_reduce_write<256, 3>(v, ops, out);
    if (threadIdx.x == 0) {
        const double mean = out[0] / ncell;
        double var = out[1] / ncell - mean * mean;
        if (var < 0.0) var = 0.0;
        stats[3 * b] = mean;
        stats[3 * b + 1] = var;
        stats[3 * b + 2] = out[2];
    }
}

// (max|Im S|, max|Re S|, sneg, spos) over the spectrum blocks (f90:1277-1288).
__global__ __launch_bounds__(1024) void diag_finalize_kernel(const double* __restrict__ partial, int nblocks,
                                                             double* __restrict__ diag4) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblocks; i += 1024) {
        v[0] = fmax(v[0], partial[4 * i]);
        v[1] = fmax(v[1], partial[4 * i + 1]);
        v[2] += partial[4 * i + 2];
        v[3] += partial[4 * i + 3];
    }
    const int ops[4] = {1, 1, 0, 0};
    double out[4];
    block_reduce_write<1024, 4>(v, ops, out);
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) diag4[c] = out[c];
}

hipError_t launch_stats_finalize(const double* partial, int nblocks, int nbatch, double ncell, double* stats,
                                 hipStream_t s) {
    hipLaunchKernelGGL(stats_finalize_kernel, dim3(nbatch), dim3(256), 0, s, partial, nblocks, ncell, stats);
    return hipGetLastError();
}

hipError_t launch_diag_finalize(const double* partial, int nblocks, double* diag4, hipStream_t s) {
    hipLaunchKernelGGL(diag_finalize_kernel, dim3(1), dim3(1024), 0, s, partial, nblocks, diag4);
    return hipGetLastError();
}

}  // namespace ss3d
