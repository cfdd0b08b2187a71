// ss3d_k32_c2r.hip -- the FP32 C2R x passes (every radix variant), compiled with multiply-adds
// fused only inside one source expression.
//
// Under the default fast contraction the compiler may fuse any product with any sum it feeds, and
// it chooses by the surrounding code: two kernels that hold the same C2R arithmetic (one member
// per block or several, 8-byte or 4-byte output stores) then rounded some outputs differently in
// the last bit.  With contraction limited to each expression (an FMA where the source writes
// a*b+c in one expression, nothing across statements) every C2R kernel rounds as its source says,
// so a realization's values do not depend on which of them ran.  The pass is bound by memory, so
// the few products left unfused cost no time.
#pragma clang fp contract(on)
#include "ss3d_fast_launch.h"

namespace ss3d {
hipError_t launch_fast_c2r_r3(int L, int in_mode, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    return fastlaunch::c2r_rm<3>(L, in_mode, nblocks, nbatch, a, s);
}
hipError_t launch_fast_c2r_r4(int L, int in_mode, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    return fastlaunch::c2r_rm<4>(L, in_mode, nblocks, nbatch, a, s);
}
hipError_t launch_fast_c2r_r5(int L, int in_mode, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    return fastlaunch::c2r_rm<5>(L, in_mode, nblocks, nbatch, a, s);
}
}  // namespace ss3d
