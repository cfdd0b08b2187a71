// ss3d_k32.hip -- dispatch of the FP32 realization passes over their radix variants.
#include "ss3d_fast_launch.h"

namespace ss3d {

hipError_t launch_fast_strided_r3(int, int, int, int, int, int, const PassArgs&, hipStream_t);
hipError_t launch_fast_strided_r4(int, int, int, int, int, int, const PassArgs&, hipStream_t);
hipError_t launch_fast_strided_r5(int, int, int, int, int, int, const PassArgs&, hipStream_t);
hipError_t launch_fast_fused_r3(int, int, const PassArgs&, const PassArgs&, const FuseMap&, hipStream_t);
hipError_t launch_fast_fused_r4(int, int, const PassArgs&, const PassArgs&, const FuseMap&, hipStream_t);
hipError_t launch_fast_fused_r5(int, int, const PassArgs&, const PassArgs&, const FuseMap&, hipStream_t);
hipError_t launch_fast_c2r_r3(int, int, int, int, const PassArgs&, hipStream_t);
hipError_t launch_fast_c2r_r4(int, int, int, int, const PassArgs&, hipStream_t);
hipError_t launch_fast_c2r_r5(int, int, int, int, const PassArgs&, hipStream_t);

int fast_strided_tile(int L, int var) { return fastlaunch::tile_P(L, var); }
bool fast_strided_pad(int L, int var) { return fastlaunch::tile_pad(L, var); }
bool fast_tile_variant_exists(int L, int rm) { return fastlaunch::tile_variant_exists(L, rm); }

int fast_c2r_rows(int L, int rm) {
    const int el = rm == 3 ? fast::eptlog(L, 3) : rm == 4 ? fast::eptlog(L, 4) : fast::eptlog(L, 5);
    const int g = (1 << L) >> el;
    return g >= 256 ? 1 : 256 / g;
}

hipError_t launch_fast_strided(int L, int in_mode, int axis, int rm, int var, int nblocks, int nbatch,
                               const PassArgs& a, hipStream_t s) {
    if (L < 1 || L > kMaxFastLog) return hipErrorInvalidValue;
    switch (rm) {
        case 3: return launch_fast_strided_r3(L, in_mode, axis, var, nblocks, nbatch, a, s);
        case 4: return launch_fast_strided_r4(L, in_mode, axis, var, nblocks, nbatch, a, s);
        case 5: return launch_fast_strided_r5(L, in_mode, axis, var, nblocks, nbatch, a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_c2r(int L, int in_mode, int rm, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    if (L < 0 || L > kMaxFastLog - 1) return hipErrorInvalidValue;
    switch (rm) {
        case 3: return launch_fast_c2r_r3(L, in_mode, nblocks, nbatch, a, s);
        case 4: return launch_fast_c2r_r4(L, in_mode, nblocks, nbatch, a, s);
        case 5: return launch_fast_c2r_r5(L, in_mode, nblocks, nbatch, a, s);
        default: return hipErrorInvalidValue;
    }
}

int fast_zrun_log() { return fastlaunch::kZrunLog; }

int fast_zrun_c2r_rows(int L, int rm) {
    if (L < fastlaunch::kZrunC2rMin || L > fastlaunch::kZrunC2rMax) return 0;
    switch (rm) {
        case 3: return fastlaunch::c2r_zr_rows<3>(L, fastlaunch::zr_c2r_seq{});
        case 4: return fastlaunch::c2r_zr_rows<4>(L, fastlaunch::zr_c2r_seq{});
        case 5: return fastlaunch::c2r_zr_rows<5>(L, fastlaunch::zr_c2r_seq{});
        default: return 0;
    }
}

bool fast_same_plan(int L, int rm1, int rm2) {
    auto np = [L](int rm) { return rm == 3 ? fast::npass(L, 3) : rm == 4 ? fast::npass(L, 4) : fast::npass(L, 5); };
    auto rl = [L](int rm, int j) {
        return rm == 3 ? fast::rlog(L, j, 3) : rm == 4 ? fast::rlog(L, j, 4) : fast::rlog(L, j, 5);
    };
    if (np(rm1) != np(rm2)) return false;
    for (int j = 0; j < np(rm1); ++j)
        if (rl(rm1, j) != rl(rm2, j)) return false;
    return true;
}

bool fast_fusable(int L, int rm, int LY, int rmy) {
    using seq = std::make_integer_sequence<int, kMaxFastLog>;
    if (L < 1 || L > kMaxFastLog) return false;
    switch (rm) {
        case 3: return fastlaunch::fusable<3>(L, LY, rmy, seq{});
        case 4: return fastlaunch::fusable<4>(L, LY, rmy, seq{});
        case 5: return fastlaunch::fusable<5>(L, LY, rmy, seq{});
        default: return false;
    }
}

hipError_t launch_fast_fused(int L, int rm, int LY, const PassArgs& az, const PassArgs& ay, const FuseMap& m,
                             hipStream_t s) {
    if (L < 1 || L > kMaxFastLog) return hipErrorInvalidValue;
    switch (rm) {
        case 3: return launch_fast_fused_r3(L, LY, az, ay, m, s);
        case 4: return launch_fast_fused_r4(L, LY, az, ay, m, s);
        case 5: return launch_fast_fused_r5(L, LY, az, ay, m, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast_x1(int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(fast::x1_kernel<0>, dim3(nblocks, nbatch), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ss3d
