// ss3d_k32.hip -- FP32 instantiations of the pass kernel: the fast realization path
// (half-spectrum z and y passes, C2R x pass with FP64 statistics).
#include "ss3d_pass.h"

namespace ss3d {

hipError_t launch_pass32(int in, int out, int nt, int nblocks, int nbatch, const PassArgs& a, hipStream_t s) {
    using R = float;
    if (in == IN_FILL32H && out == OUT_STORE) return launch_nt<R, IN_FILL32H, OUT_STORE>(nt, nblocks, nbatch, a, s);
    if (in == IN_LOAD && out == OUT_STORE) return launch_nt<R, IN_LOAD, OUT_STORE>(nt, nblocks, nbatch, a, s);
    if (in == IN_LOADH && out == OUT_REAL) return launch_nt<R, IN_LOADH, OUT_REAL>(nt, nblocks, nbatch, a, s);
    if (in == IN_FILLH && out == OUT_REAL) return launch_nt<R, IN_FILLH, OUT_REAL>(nt, nblocks, nbatch, a, s);
    return hipErrorInvalidValue;
}

}  // namespace ss3d
