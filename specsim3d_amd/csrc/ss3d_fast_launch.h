// ss3d_fast_launch.h -- launchers of the FP32 passes for one maximum radix (RM), instantiated
// by ss3d_k32_r{3,4,5}.hip so that the variants compile in parallel.
#pragma once

#include "ss3d_fast.h"

namespace ss3d {
namespace fastlaunch {

constexpr int strided_P(int L) { return L <= 7 ? 64 : (L >= 12 ? 2 : (8192 >> L)); }

// Tile variants of the strided passes, chosen per length and pass by the autotuner (results are
// the same bit for bit: each column's transform does not depend on the tile).  Variant 0 is the
// tile above; variant 1 halves the tile width for lengths up to 2^8 (twice the blocks for small
// grids) and drops the LDS position padding for length 2^10 (tiles of 8 columns).  Variant 1 is
// instantiated for radix plans 3 and 4; fused launches use variant 0.
constexpr bool tile_variant_exists(int L, int RM) { return (L <= 8 || L == 10) && RM <= 4; }
constexpr int tile_P(int L, int V) { return V == 1 && L <= 8 ? strided_P(L) / 2 : strided_P(L); }
constexpr bool tile_pad(int L, int V) { return V == 1 && L == 10 ? false : tile_P(L, V) < 16; }

template <typename K>
hipError_t launch_k(K k, size_t lds, int nt, int nblocks, int nbatch, const PassArgs& a, hipStream_t s, bool& attr) {
    if (lds > 65536 && !attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k, dim3(nblocks, nbatch), dim3(nt), lds, s, a);
    return hipGetLastError();
}

template <int L, int IN, int AXIS, int RM, int IG, int V = 0>
hipError_t strided_L(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    constexpr int P = tile_P(L, V);
    constexpr bool PAD = tile_pad(L, V);
    using C = fast::StridedCfg<L, P, RM, PAD>;
    static bool attr = false;
    return launch_k(fast::strided_kernel<L, P, IN, AXIS, RM, IG, PAD>, C::lds_bytes(), C::NT, nblocks, nbatch, a, s,
                    attr);
}

// y lengths of a fused launch beside a fill of length 2^L and plan RM: 2^L (plan RM) or, with
// plan 3, 2^(L+1) where the tiles have one thread count and two blocks fit the CU's LDS
template <int L, int RM>
constexpr bool fused_up() {
    if constexpr (L + 1 > kMaxFastLog) {
        return false;
    } else {
        using CZ = fast::StridedCfg<L, strided_P(L), RM>;
        using CY = fast::StridedCfg<L + 1, strided_P(L + 1), 3>;
        return CZ::NT == CY::NT && 2 * (CZ::lds_bytes() > CY::lds_bytes() ? CZ::lds_bytes() : CY::lds_bytes()) <= 163840;
    }
}

template <int L, int RM, int IG, int LY = L, int RMY = RM, bool HALF = false>
hipError_t fused_L(int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    using C = fast::StridedCfg<L, strided_P(L), RM>;
    using CY = fast::StridedCfg<LY, strided_P(LY), RMY>;
    static bool attr = false;
    const auto k = fast::fill_load_kernel<L, strided_P(L), RM, IG, LY, strided_P(LY), RMY, HALF>;
    const size_t lds = C::lds_bytes() > CY::lds_bytes() ? C::lds_bytes() : CY::lds_bytes();
    if (lds > 65536 && !attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(C::NT), lds, s, az, ay, m);
    return hipGetLastError();
}

// fills of length >= 2^kHalfMinLog whose z and y passes both keep at most half their axes use
// the HALF instance
constexpr int kHalfMinLog = 7;

template <int RM, int IG, int L>
hipError_t fused_one(int LY, int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    if constexpr (L >= kHalfMinLog) {
        if (2 * az.n_out <= (1 << L) && 2 * (int64_t)ay.n_out <= (int64_t(1) << LY)) {
            if (LY == L) return fused_L<L, RM, IG, L, RM, true>(nblocks, az, ay, m, s);
            if constexpr (fused_up<L, RM>()) {
                if (LY == L + 1) return fused_L<L, RM, IG, L + 1, 3, true>(nblocks, az, ay, m, s);
            }
            return hipErrorInvalidValue;
        }
    }
    if (LY == L) return fused_L<L, RM, IG>(nblocks, az, ay, m, s);
    if constexpr (fused_up<L, RM>()) {
        if (LY == L + 1) return fused_L<L, RM, IG, L + 1, 3>(nblocks, az, ay, m, s);
    }
    return hipErrorInvalidValue;
}

template <int RM, int IG, int... Ls>
hipError_t fused_dispatch(int L, int LY, int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m,
                          hipStream_t s, std::integer_sequence<int, Ls...>) {
    hipError_t e = hipErrorInvalidValue;
    ((L == Ls + 1 ? (e = fused_one<RM, IG, Ls + 1>(LY, nblocks, az, ay, m, s), 0) : 0), ...);
    return e;
}

// whether a fill of length 2^L (plan RM) and a y pass of length 2^LY (plan rmy) have a fused launch
template <int RM, int... Ls>
bool fusable(int L, int LY, int rmy, std::integer_sequence<int, Ls...>) {
    bool ok = false;
    ((L == Ls + 1 ? (ok = (LY == L && rmy == RM) || (LY == L + 1 && rmy == 3 && fused_up<Ls + 1, RM>()), 0) : 0), ...);
    return ok;
}

// ---- z-run layout (PassArgs::zrun): the wave-local z fill of length 2^kZrunLog (tile 16), alone or
// beside a y pass of the same length and plan, and the C2R pass reading the y pass's z-run output
// (lengths 2^kZrunC2rMin .. 2^kZrunC2rMax, plans whose tiles hold whole runs)
constexpr int kZrunLog = 9, kZrunC2rMin = 7, kZrunC2rMax = 10;

template <typename K>
hipError_t lds_attr(K k, size_t lds, bool& attr) {
    if (lds > 65536 && !attr) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    return hipSuccess;
}

template <int L, int RM, int IG>
hipError_t strided_zr_L(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    constexpr int P = strided_P(L);
    using C = fast::StridedCfg<L, P, RM, true>;
    static bool attr = false;
    return launch_k(fast::strided_kernel<L, P, IN_FILL32H, 2, RM, IG, (P < 16), true>, C::lds_bytes(), C::NT, nblocks,
                    nbatch, a, s, attr);
}
template <int RM>
hipError_t strided_zr(int L, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    if (L != kZrunLog) return hipErrorInvalidValue;
    return a.igauss ? strided_zr_L<kZrunLog, RM, 1>(nblocks, nbatch, a, s)
                    : strided_zr_L<kZrunLog, RM, 0>(nblocks, nbatch, a, s);
}

template <int L, int RM, int IG, bool HALF>
hipError_t fused_zr_L(int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    constexpr int P = strided_P(L);
    using C = fast::StridedCfg<L, P, RM, true>;
    using CY = fast::StridedCfg<L, P, RM>;
    constexpr size_t lds = C::lds_bytes() > CY::lds_bytes() ? C::lds_bytes() : CY::lds_bytes();
    static_assert(2 * lds <= 163840, "two blocks per CU");
    static bool attr = false;
    const auto k = fast::fill_load_kernel<L, P, RM, IG, L, P, RM, HALF, true>;
    hipError_t e = lds_attr(k, lds, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(C::NT), lds, s, az, ay, m);
    return hipGetLastError();
}
template <int RM>
hipError_t fused_zr(int L, int LY, int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    if (L != kZrunLog || LY != kZrunLog) return hipErrorInvalidValue;
    constexpr int Lz = kZrunLog;
    const bool half = 2 * az.n_out <= (1 << Lz) && 2 * (int64_t)ay.n_out <= (int64_t(1) << Lz);
    if (az.igauss)
        return half ? fused_zr_L<Lz, RM, 1, true>(nblocks, az, ay, m, s) : fused_zr_L<Lz, RM, 1, false>(nblocks, az, ay, m, s);
    return half ? fused_zr_L<Lz, RM, 0, true>(nblocks, az, ay, m, s) : fused_zr_L<Lz, RM, 0, false>(nblocks, az, ay, m, s);
}

// z fill (az, length 2^L, plan RM) beside the y pass (ay, length 2^LY) in nblocks = m.nzb + m.nyb blocks
template <int RM>
hipError_t fused_rm(int L, int LY, int nblocks, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    static_assert(2 * sizeof(PassArgs) + sizeof(FuseMap) <= 4096, "kernel arguments");
    if (az.zrun) return fused_zr<RM>(L, LY, nblocks, az, ay, m, s);
    using seq = std::make_integer_sequence<int, kMaxFastLog>;
    return az.igauss ? fused_dispatch<RM, 1>(L, LY, nblocks, az, ay, m, s, seq{})
                     : fused_dispatch<RM, 0>(L, LY, nblocks, az, ay, m, s, seq{});
}

// the member-loop C2R code (c2r_rows) serves a pass from the y output whose tiles of P rows span
// less than 2 GiB (buffer offsets); output pairs as 8-byte stores where c2r_rows_even holds
inline bool c2r_rows_fit(const PassArgs& a, int P) {
    return (int64_t)P * a.in_row * 8 < 0x7fffffffLL && (int64_t)P * a.out_row * 4 < 0x7fffffffLL;
}
inline bool c2r_rows_even(const PassArgs& a) {
    return ((a.n_out | a.out_row | a.out_bstride) & 1) == 0 && (reinterpret_cast<uintptr_t>(a.out) & 7) == 0;
}

template <int L, int IN, int RM>
hipError_t c2r_L(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    using C = fast::C2RCfg<L, RM>;
    if constexpr (IN == IN_LOADH && C::G <= 64 && C::M > 0) {
        if (c2r_rows_fit(a, C::P)) {
            static bool attr_even = false, attr_odd = false;
            const int groups = (nbatch + a.bpb - 1) / a.bpb;
            if (c2r_rows_even(a))
                return launch_k(fast::c2r_loop_kernel<L, RM, true>, fast::C2RRowsCfg<L, RM>::lds_bytes(), C::NT, nblocks,
                                groups, a, s, attr_even);
            return launch_k(fast::c2r_loop_kernel<L, RM, false>, fast::C2RRowsCfg<L, RM>::lds_bytes(), C::NT, nblocks,
                            groups, a, s, attr_odd);
        }
    }
    static bool attr = false;
    return launch_k(fast::c2r_kernel<L, IN, RM>, C::lds_elems() * sizeof(float2), C::NT, nblocks, nbatch, a, s, attr);
}

// the C2R pass reading the z-run layout
template <int L, int RM>
hipError_t c2r_zr_L(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    using Z = fast::C2RZrCfg<L, RM>;
    if constexpr (Z::ok) {
        static bool attr_even = false, attr_odd = false;
        const int groups = (nbatch + a.bpb - 1) / a.bpb;
        if (c2r_rows_even(a))
            return launch_k(fast::c2r_zr_kernel<L, RM, true>, Z::lds_bytes(), Z::C::NT, nblocks, groups, a, s, attr_even);
        return launch_k(fast::c2r_zr_kernel<L, RM, false>, Z::lds_bytes(), Z::C::NT, nblocks, groups, a, s, attr_odd);
    } else {
        return hipErrorInvalidValue;
    }
}
template <int RM, int... Ls>
hipError_t c2r_zr_dispatch(int L, int nblocks, int nbatch, const PassArgs& a, hipStream_t s,
                           std::integer_sequence<int, Ls...>) {
    hipError_t e = hipErrorInvalidValue;
    ((L == Ls + kZrunC2rMin ? (e = c2r_zr_L<Ls + kZrunC2rMin, RM>(nblocks, nbatch, a, s), 0) : 0), ...);
    return e;
}
// whether the C2R pass of length 2^L and plan RM reads the z-run layout: its rows per block, else 0
template <int L, int RM>
constexpr int c2r_zr_rows_L() {
    return fast::C2RZrCfg<L, RM>::ok ? fast::C2RCfg<L, RM>::P : 0;
}
template <int RM, int... Ls>
int c2r_zr_rows(int L, std::integer_sequence<int, Ls...>) {
    int p = 0;
    ((L == Ls + kZrunC2rMin ? (p = c2r_zr_rows_L<Ls + kZrunC2rMin, RM>(), 0) : 0), ...);
    return p;
}
using zr_c2r_seq = std::make_integer_sequence<int, kZrunC2rMax - kZrunC2rMin + 1>;

template <int L, int IN, int AXIS, int RM, int IG>
hipError_t strided_LV(int var, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    if (var == 0) return strided_L<L, IN, AXIS, RM, IG, 0>(nblocks, nbatch, a, s);
    if constexpr (tile_variant_exists(L, RM)) {
        if (var == 1) return strided_L<L, IN, AXIS, RM, IG, 1>(nblocks, nbatch, a, s);
    }
    return hipErrorInvalidValue;
}

template <int RM, int IN, int AXIS, int IG, int... Ls>
hipError_t strided_dispatch(int L, int var, int nblocks, int nbatch, const PassArgs& a, hipStream_t s,
                            std::integer_sequence<int, Ls...>) {
    hipError_t e = hipErrorInvalidValue;
    ((L == Ls + 1 ? (e = strided_LV<Ls + 1, IN, AXIS, RM, IG>(var, nblocks, nbatch, a, s), 0) : 0), ...);
    return e;
}

template <int RM, int IN, int... Ls>
hipError_t c2r_dispatch(int L, int nblocks, int nbatch, const PassArgs& a, hipStream_t s,
                        std::integer_sequence<int, Ls...>) {
    hipError_t e = hipErrorInvalidValue;
    ((L == Ls ? (e = c2r_L<Ls, IN, RM>(nblocks, nbatch, a, s), 0) : 0), ...);
    return e;
}

template <int RM>
hipError_t strided_rm(int L, int in_mode, int axis, int var, int nblocks, int nbatch, const PassArgs& a,
                      hipStream_t s) {
    using seq = std::make_integer_sequence<int, kMaxFastLog>;
    if (in_mode == IN_FILL32H && axis == 2 && a.zrun) return var == 0 ? strided_zr<RM>(L, nblocks, nbatch, a, s) : hipErrorInvalidValue;
    // fill passes: one instantiation per coefficient law (igauss 0 or 1)
    if (in_mode == IN_LOAD) return strided_dispatch<RM, IN_LOAD, 1, 0>(L, var, nblocks, nbatch, a, s, seq{});
    if (in_mode == IN_FILL32H && axis == 2)
        return a.igauss ? strided_dispatch<RM, IN_FILL32H, 2, 1>(L, var, nblocks, nbatch, a, s, seq{})
                        : strided_dispatch<RM, IN_FILL32H, 2, 0>(L, var, nblocks, nbatch, a, s, seq{});
    if (in_mode == IN_FILL32H && axis == 1)
        return a.igauss ? strided_dispatch<RM, IN_FILL32H, 1, 1>(L, var, nblocks, nbatch, a, s, seq{})
                        : strided_dispatch<RM, IN_FILL32H, 1, 0>(L, var, nblocks, nbatch, a, s, seq{});
    return hipErrorInvalidValue;
}

template <int RM>
hipError_t c2r_rm(int L, int in_mode, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    using seq = std::make_integer_sequence<int, kMaxFastLog>;
    if (in_mode == IN_LOADH && a.zrun) return c2r_zr_dispatch<RM>(L, nblocks, nbatch, a, s, zr_c2r_seq{});
    if (in_mode == IN_LOADH) return c2r_dispatch<RM, IN_LOADH>(L, nblocks, nbatch, a, s, seq{});
    if (in_mode == IN_FILLH) return c2r_dispatch<RM, IN_FILLH>(L, nblocks, nbatch, a, s, seq{});
    return hipErrorInvalidValue;
}

}  // namespace fastlaunch
}  // namespace ss3d
