// fmmhip_gemm_d_f32leaf.hip -- instantiates the fused FMM kernel for
// FP64 operands, FP32 MFMA leaf, FP64 C.
#include "fmmhip_kernel.h"

int fmmk_launch_d_f32leaf(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
    return fmmk::launch_typed<double, float, double>(ka, shape, a_mn, b_mn, c_col, nwg, st);
}
