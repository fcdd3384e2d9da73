// fmmhip_gemm_d64.hip -- instantiates the fused FMM kernel for
// FP64 operands, FP64 leaf, FP64 C.
#include "fmmhip_kernel.h"

int fmmk_launch_d64(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
    return fmmk::launch_typed<double, double, double>(ka, shape, a_mn, b_mn, c_col, nwg, st);
}

int fmmk_launch_batch_d64(const fmmk::KArgs* probs, const fmmk::BatchItem* items, const int* qoff, int* heads,
                          bool vec, int shape, int nwg, hipStream_t st)
{
    return fmmk::launch_batch_typed<double, double, double>(probs, items, qoff, heads, vec, shape, nwg, st);
}
