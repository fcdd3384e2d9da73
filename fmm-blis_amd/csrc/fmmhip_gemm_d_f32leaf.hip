// fmmhip_gemm_d_f32leaf.hip -- instantiates the fused FMM kernel for
// FP64 operands, FP32 MFMA leaf, FP64 C, and the kernels over the FP32
// pre-summed leaves of the same kind.
#include "fmmhip_kernel.h"

int fmmk_launch_d_f32leaf(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
    return fmmk::launch_typed<double, float, double>(ka, shape, a_mn, b_mn, c_col, nwg, st);
}

int fmmk_launch_f32leaves_d(const fmmk::KArgs& ka, int shape, int nwg, hipStream_t st)
{
    using namespace fmmk;
    if (ka.ksplit) return FMMHIP_ERR_UNSUPPORTED;
    if (shape == 3) {
        if (!fast_eligible<float>(ka, true, false, true)) return FMMHIP_ERR_UNSUPPORTED;
        hipLaunchKernelGGL((fmm_fast_kernel<float, float, double>), dim3(nwg), dim3(kFastNT), 0, st, ka);
    } else if (shape == 2) {
        hipLaunchKernelGGL((fmm_gemm_kernel<float, float, double, 2, true, true, false, true>), dim3(nwg),
                           dim3(Cfg<2>::NT), 0, st, ka);
    } else if (shape == 0) {
        hipLaunchKernelGGL((fmm_gemm_kernel<float, float, double, 0, true, true, false, true>), dim3(nwg),
                           dim3(Cfg<0>::NT), 0, st, ka);
    } else {
        return FMMHIP_ERR_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}
