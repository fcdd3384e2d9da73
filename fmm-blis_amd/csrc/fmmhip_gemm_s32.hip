// fmmhip_gemm_s32.hip -- instantiates the fused FMM kernel for
// FP32 operands, FP32 leaf, FP32 C.
#include "fmmhip_kernel.h"

int fmmk_launch_s32(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
    return fmmk::launch_typed<float, float, float>(ka, shape, a_mn, b_mn, c_col, nwg, st);
}

int fmmk_launch_batch_s32(const fmmk::KArgs* probs, const fmmk::BatchItem* items, const int* qoff, int* heads,
                          bool vec, int shape, int nwg, hipStream_t st)
{
    return fmmk::launch_batch_typed<float, float, float>(probs, items, qoff, heads, vec, shape, nwg, st);
}
