// The 256 x 256 split-BF16x3 tile on four waves (128 x 128 per wave) with operands staged by direct
// global -> LDS requests.  Its 256 accumulator registers per lane live in AGPRs, so this file is built
// without the VGPR-accumulator form the other GEMM files use (see the Makefile).
#include "dcc_bf16x3.h"

namespace dcc_bf16x3 {

int run_bf16x3_dma_wide(ContractArgs<float> &args, int64_t nbatch, int64_t max_rows, bool c_mfast, int mi,
                        hipStream_t s) {
    if (mi == 32) return run_bf16x3_dma<256, 256, 32, 2, 4>(args, nbatch, max_rows, c_mfast, s);
    return run_bf16x3_dma<256, 256, 16, 2, 4>(args, nbatch, max_rows, c_mfast, s);
}

}  // namespace dcc_bf16x3
