// ss3d_k32_r3.hip -- FP32 passes with radices up to 2^3 (see ss3d_fast_launch.h).
#include "ss3d_fast_launch.h"

namespace ss3d {
hipError_t launch_fast_strided_r3(int L, int in_mode, int axis, int var, int nblocks, int nbatch, const PassArgs& a,
                                   hipStream_t s) {
    return fastlaunch::strided_rm<3>(L, in_mode, axis, var, nblocks, nbatch, a, s);
}
hipError_t launch_fast_fused_r3(int L, int LY, const PassArgs& az, const PassArgs& ay, const FuseMap& m, hipStream_t s) {
    return fastlaunch::fused_rm<3>(L, LY, m.nzb + m.nyb, az, ay, m, s);
}
}  // namespace ss3d
