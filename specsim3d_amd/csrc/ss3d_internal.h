// ss3d_internal.h -- declarations shared by the host library (ss3d_host.cpp) and the kernel
// translation units (ss3d_k64.hip, ss3d_k32.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss3d {

constexpr int kMaxStruct = 10;     // maxst, src/specsim3d.f90:93
constexpr int kMaxLogN = 14;       // twiddle tables cover lengths up to 2^14
constexpr int kMaxFastLog = 13;    // longest axis of the FP32 passes: 2^13
constexpr uint32_t kPmM = 2147483647u;  // Park-Miller modulus 2^31-1 (f90:193)
constexpr uint32_t kPmA = 16807u;       // Park-Miller multiplier (f90:193)
constexpr int kEPT = 16;           // complex elements per thread in every FFT pass
constexpr int kPadShift = 4;       // LDS padding: one slot after every 16 complex elements
// The FP32 half-spectrum amplitudes are stored times sqrt(ln 2): amp sqrt(-2 ln u1)/sqrt(2) =
// (amp sqrt(ln 2)) sqrt(-log2 u1), one multiplication fewer per Gaussian coefficient.
constexpr double kAmp32Scale = 0.83255461115769775635;

// One covariance structure as the kernels use it (setrot already applied, f90:107-139).
struct StructDev {
    int itype;
    int pad_;
    double sill, ahmax;
    double rot[9];  // row-major
};

// Everything a pass needs.  Passed by value (about 1.3 KB of kernel arguments).
struct PassArgs {
    // FFT geometry
    int logn;   // transform length along the pass axis = 2^logn
    int logp;   // pencils per tile = 2^logp (tile holds 2^(logn+logp) = NT*kEPT elements)
    int n_out;  // positions kept along the axis on output (pruning)
    int axis;   // grid axis of the pass: 0 x (contiguous pencils), 1 y, 2 z (strided tiles)
    // strided tiles: pencils are consecutive along the contiguous dimension i0
    int n0;       // active extent of i0
    int ntiles0;  // ceil(n0 / P)
    int nthird;   // extent of the remaining dimension
    // contiguous tiles: whole rows
    int nrows;
    int64_t in_sA, in_sT, out_sA, out_sT;  // element strides (axis, third dimension)
    int64_t in_sC, out_sC;                 // strided passes: element stride between column tiles (0: P)
    int64_t in_row, out_row;               // row strides of contiguous tiles
    int64_t in_bstride, out_bstride;       // per batch member (realization)
    const void* in;
    void* out;
    const void* tw;  // twiddle table of length 2^(logn-1) for this axis and direction
    // periodic grid and crop
    int npx, npy, npz, hx;  // hx = npx/2+1 (npx >= 2) or 1: half-spectrum width
    int hxa;                // row stride of the FP32 amplitude: hx rounded up to even
    int nx, ny, nz;
    // covariance model (spectrum passes)
    int nst;
    double nugget, dx, dy, dz;
    StructDev st[kMaxStruct];
    // coefficient fill (realization passes)
    const double* amp64;     // npx*npy*npz
    const float* amp32h;     // hxa*npy*npz (columns kx <= npx/2 used), amplitude * kAmp32Scale
    const uint32_t* rowf;    // per drawing row: a^(g*(draw-point index of its first point))
    const uint32_t* txp;     // a^(g*i), i < npx
    const uint32_t* rbase;   // per realization: state before its first draw, times a
    int igauss;
    // outputs besides `out`
    double* amp64_out;
    float* amp32h_out;
    double* partial;  // per block: statistics (3 doubles) or spectrum diagnostics (4 doubles)
    int nblocks;      // blocks per batch member
    int nbatch;       // batch members of the launch
    int bpb;          // FP32 fill passes: batch members per block (grid y = ceil(nbatch / bpb))
    int zrun;         // FP32 z fill and C2R passes: z-run layout of the fill and y outputs (fast_zrun_*)
    int spec_half;    // last spectrum pass over the columns kx <= npx/2 alone: sums count 0 < kx < npx/2 twice
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous run of tiles so
// that neighbouring tiles (which share cache lines and table entries) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    constexpr int X = 8;
    const int per = nb / X, rem = nb % X;
    const int xcd = bid % X, loc = bid / X;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

enum InMode { IN_COV = 0, IN_FILL64 = 1, IN_LOAD = 2, IN_FILL32H = 3, IN_LOADH = 4, IN_FILLH = 5 };
enum OutMode { OUT_STORE = 0, OUT_SPEC = 1, OUT_REAL = 2 };

// Kernel launchers (defined in the .hip translation units).  nt: threads per block
// (64..1024), nbatch: batch members (grid y).  Return hipSuccess or the launch error.
hipError_t launch_pass64(int in_mode, int out_mode, int nt, int nblocks, int nbatch, const PassArgs& a,
                         hipStream_t s);
// Reduce per-block partials: statistics -> (mean, var, max|Im|) per batch member, and
// spectrum diagnostics -> (max|Im S|, max|Re S|, sneg, spos).
hipError_t launch_stats_finalize(const double* partial, int nblocks, int nbatch, double ncell, double* stats,
                                 hipStream_t s);
hipError_t launch_diag_finalize(const double* partial, int nblocks, double* diag4, hipStream_t s);

// FP32 realization passes (ss3d_fast.h).  `tw` of a strided pass points at the W_n table of
// its length (entry m = exp(+2 pi i m/n), FP32); a C2R pass gets the W_N table of its complex
// length N = npx/2, immediately followed by the W_2N table.
// rm: largest radix 2^rm of the pass plan (3, 4 or 5).
int fast_strided_tile(int L, int var = 0);     // columns per block of tile variant var
bool fast_strided_pad(int L, int var = 0);     // LDS position padding of tile variant var
bool fast_tile_variant_exists(int L, int rm);  // variant 1 instantiated for this length and plan
int fast_c2r_rows(int L, int rm);    // rows per block
hipError_t launch_fast_strided(int L, int in_mode, int axis, int rm, int var, int nblocks, int nbatch,
                               const PassArgs& a, hipStream_t s);
hipError_t launch_fast_c2r(int L, int in_mode, int rm, int nblocks, int nbatch, const PassArgs& a, hipStream_t s);
hipError_t launch_fast_x1(int nblocks, int nbatch, const PassArgs& a, hipStream_t s);
// z fill pass (az) beside the y pass (ay) of the previous batch in one launch of
// m.nzb + m.nyb blocks (both passes of length 2^L, largest radix 2^rm); the block roles
// interleave as 8A fill blocks, 8C y blocks, K times, then the rest of each.
struct FuseMap {
    int nzb, nyb;  // blocks of the fill and of the y pass (tiles x member groups)
    int A, C, K;
};
bool fast_same_plan(int L, int rm1, int rm2);  // the FP32 radix plans of length 2^L for rm1 and rm2 are one
// a fused launch exists for a fill of length 2^L (plan rm) and a y pass of length 2^LY (plan rmy):
// LY = L with rmy = rm, or LY = L + 1 with rmy = 3 where the tiles have one thread count
bool fast_fusable(int L, int rm, int LY, int rmy);
hipError_t launch_fast_fused(int L, int rm, int LY, const PassArgs& az, const PassArgs& ay, const FuseMap& m,
                             hipStream_t s);
// z-run layout (PassArgs::zrun = 1): the z fill of length 2^fast_zrun_log() (tile 16, any plan, tile
// variant 0; alone or fused with a y pass of that length) writes runs of 8 z per column pair, the
// y pass reads and writes them through its strides, and the C2R pass of length 2^L and plan rm
// reads them when fast_zrun_c2r_rows(L, rm) > 0 (its rows per block, a multiple of 8: the 8 z of a
// run at rows/8 consecutive y)
int fast_zrun_log();
int fast_zrun_c2r_rows(int L, int rm);

// FP64 realization passes with compile-time lengths 2^1 .. 2^kMaxFast64Log (ss3d_fast64.h), bit
// for bit the arithmetic of launch_pass64.  kind 4: x pass (fill + transform of rows, OUT_STORE),
// kind 5: strided y/z pass (IN_LOAD, OUT_STORE, OUT_REAL or OUT_SPEC), kind 6: x pass of the
// spectral density (IN_COV, OUT_STORE).  fast64_tile: rows or columns per block.
constexpr int kMaxFast64Log = 12;
int fast64_tile(int kind, int L);
int fast64_rows_blocks(int L, int npy, int npz);  // blocks of a kind-4 launch
hipError_t launch_fast64(int kind, int L, int out, int nblocks, int nbatch, const PassArgs& a, hipStream_t s);

// LDS bytes a pass with nt threads uses (dynamic shared memory).
inline size_t pass_lds_bytes(int nt, size_t elem_bytes) {
    size_t e = (size_t)nt * kEPT;
    return (e + (e >> kPadShift)) * elem_bytes;
}

}  // namespace ss3d
