// Grouped persistent gather-GEMM: one launch runs a list of independent contractions (for example
// the same residual term of many small CC systems).  Each problem is a dcc_contract_f64 call in
// descriptor form; the library tiles every problem, concatenates the tile ranges, and a fixed grid of
// workgroups (as many as are resident at once) walks the global tile list in windows of one grid:
// in window w, workgroup b takes tile w * grid + xcd_remap(b), so the tiles one XCD runs together are
// neighbours that share A and B panels in its L2.  The problem of a tile is found by binary search
// over the per-problem tile offsets, and the tile itself is the same code as the one-launch-per-GEMM
// kernel (dcc_gemm_tile.h: pair loads, one- or two-stage register prefetch, branch-free K maps).
// Every workgroup leaves the loop when the list is exhausted, so the launch always drains.  Problems
// of one launch must write disjoint C regions (no two problems accumulate into the same element).
#include "dcc_common.h"
#include "dcc_gemm_tile.h"

#include <algorithm>
#include <vector>

namespace {

using namespace dcc_gemm;

// flags of a prepared problem: bits 0..2 the loader/epilogue choice of the launch (A K-contiguous,
// B N-contiguous, C written along m; the same for every problem of a launch), bits 3..4 the pair
// vector loads allowed for A and B of this problem under that choice.
template <int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST, int STAGES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MinWaves<BM, BN, STAGES>::v)))
void grouped_kernel(const dcc_gemm_problem *__restrict__ probs, int nprob, int64_t total_tiles) {
    __shared__ double smem[TileSmem<BM, BN, C_MFAST>::v];
    const int64_t grid = gridDim.x;
    for (int64_t base = 0; base < total_tiles; base += grid) {
        const int64_t window = min(grid, total_tiles - base);
        if ((int64_t)blockIdx.x >= window) break;
        const int64_t t = base + xcd_remap((int64_t)blockIdx.x, window);
        // problem owning tile t: last p with tile_begin <= t
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (probs[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const dcc_gemm_problem &q = probs[lo];
        ContractArgs<double> a;
        a.M = 0;
        a.N = (int)q.N;
        a.K = (int)q.K;
        a.alpha = q.alpha;
        a.beta = q.beta;
        a.A = static_cast<const double *>(q.A);
        a.B = static_cast<const double *>(q.B);
        a.C = static_cast<double *>(q.C);
        a.am = q.am;
        a.bn = q.bn;
        a.cm = q.cm;
        a.cn = q.cn;
        a.ak = make_kmap(q.ak);
        a.bk = make_kmap(q.bk);
        a.batches = q.batches;
        a.tiles_m = (int)q.tiles_m;
        a.tiles_n = (int)q.tiles_n;
        a.total_tiles = 0;
        a.vec = (int)((q.flags >> 3) & 3);
        const int64_t local = t - q.tile_begin;
        const int64_t per_batch = q.tiles_m * q.tiles_n;
        const int64_t batch = local / per_batch;
        int tm, tn;
        tile_coords(local - batch * per_batch, a.tiles_m, a.tiles_n, tm, tn);
        __syncthreads();  // the previous tile's epilogue may still read LDS
        contract_tile<double, BM, BN, A_KFAST, B_NFAST, C_MFAST, STAGES>(a, batch, tm, tn, smem);
    }
}

template <int BM, int BN, int STAGES> struct GroupedLauncher {
    template <bool AK, bool BN_, bool CM>
    static int go(int64_t grid, const dcc_gemm_problem *probs, int nprob, int64_t total, hipStream_t s) {
        auto kern = grouped_kernel<BM, BN, AK, BN_, CM, STAGES>;
        if (grid == 0) {
            // resident workgroups per CU of this kernel (registers and LDS), queried once
            static int per_cu = 0;
            int dev = 0, cus = 0;
            DCC_HIP(hipGetDevice(&dev));
            DCC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (per_cu == 0) {
                int n = 0;
                DCC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kThreads, 0));
                per_cu = std::max(n, 1);
            }
            grid = (int64_t)cus * per_cu;
        }
        if (grid > total) grid = total;
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kThreads), 0, s, probs, nprob, total);
        DCC_LAUNCH_CHECK();
        return 0;
    }
    static int run(int flags, int64_t grid, const dcc_gemm_problem *probs, int nprob, int64_t total, hipStream_t s) {
        switch (flags & 7) {
            case 0: return go<false, false, false>(grid, probs, nprob, total, s);
            case 1: return go<true, false, false>(grid, probs, nprob, total, s);
            case 2: return go<false, true, false>(grid, probs, nprob, total, s);
            case 3: return go<true, true, false>(grid, probs, nprob, total, s);
            case 4: return go<false, false, true>(grid, probs, nprob, total, s);
            case 5: return go<true, false, true>(grid, probs, nprob, total, s);
            case 6: return go<false, true, true>(grid, probs, nprob, total, s);
            default: return go<true, true, true>(grid, probs, nprob, total, s);
        }
    }
};

// Loader/epilogue choice of one problem, as dcc_contract_f64 makes it.
int problem_flags(const dcc_gemm_problem &p) {
    const bool a_kfast = inner_stride(p.ak) == 1 || inner_stride(p.am) != 1;
    const bool b_nfast = inner_stride(p.bn) == 1 || inner_stride(p.bk) != 1;
    const bool c_mfast = inner_stride(p.cn) != 1;
    return (a_kfast ? 1 : 0) | (b_nfast ? 2 : 0) | (c_mfast ? 4 : 0);
}

}  // namespace

extern "C" {

int dcc_grouped_prepare(dcc_gemm_problem *probs, int64_t nprob, int64_t tile_shape_id, int64_t *total_tiles,
                        int64_t *variant) {
    DCC_CHECK(nprob >= 0 && total_tiles != nullptr && variant != nullptr, "bad arguments");
    DCC_CHECK(nprob == 0 || probs != nullptr, "null problems");
    int bm = 0, bn = 0, stages = 0;
    DCC_CHECK(tile_shape_of(tile_shape_id, bm, bn, stages), "tile shape must be 0..7");
    // The loaders and epilogue are correct for any strides; their choice only decides coalescing.
    // One launch has one choice: the one of the problems with the most multiply-adds.
    double weight[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < nprob; ++i) {
        const dcc_gemm_problem &p = probs[i];
        DCC_CHECK(p.N >= 0 && p.K >= 0 && p.nbatch >= 0 && p.max_rows >= 0, "negative extent");
        DCC_CHECK(p.N < ((int64_t)1 << 31) && p.K < ((int64_t)1 << 31), "extent exceeds 2^31");
        DCC_CHECK(p.ak.table == nullptr && p.bk.table == nullptr, "K index maps must be mixed-radix maps");
        for (const dcc_idxmap *m : {&p.am, &p.ak, &p.bk, &p.bn, &p.cm, &p.cn})
            DCC_CHECK(m->table != nullptr ? m->table_bstride >= 0
                                          : m->e1 >= 1 && m->e2 >= 1 && m->e1 * m->e2 < ((int64_t)1 << 32),
                      "invalid index map");
        weight[problem_flags(p)] += (double)p.N * (double)p.K * (double)p.max_rows * (double)p.nbatch;
    }
    const int flags = (int)(std::max_element(weight, weight + 8) - weight);
    const bool a_kfast = flags & 1, b_nfast = (flags & 2) != 0;
    int64_t acc = 0;
    for (int64_t i = 0; i < nprob; ++i) {
        dcc_gemm_problem &p = probs[i];
        p.tiles_m = (p.max_rows + bm - 1) / bm;
        p.tiles_n = (p.N + bn - 1) / bn;
        p.tile_begin = acc;
        const int vec = (pairs_contiguous(p.A, p.am, p.ak, a_kfast, 0, 16) ? 1 : 0) |
                        (pairs_contiguous(p.B, p.bn, p.bk, !b_nfast, p.N, 16) ? 2 : 0);
        p.flags = flags | (vec << 3);
        acc += p.nbatch * p.tiles_m * p.tiles_n;
    }
    DCC_CHECK(acc < ((int64_t)1 << 40), "too many tiles");
    *total_tiles = acc;
    *variant = tile_shape_id | ((int64_t)flags << 3);
    return 0;
}

int dcc_contract_grouped_f64(const dcc_gemm_problem *probs, int64_t nprob, int64_t total_tiles, int64_t variant,
                             int64_t nworkgroups, dcc_stream_t stream) {
    DCC_CHECK(nprob >= 0 && total_tiles >= 0 && nworkgroups >= 0, "negative argument");
    DCC_CHECK(nprob < (1 << 30), "too many problems");
    DCC_CHECK(variant >= 0 && variant < 64, "variant must come from dcc_grouped_prepare");
    if (nprob == 0 || total_tiles == 0) return 0;
    DCC_CHECK(probs != nullptr, "null problems");
    int bm = 0, bn = 0, stages = 0;
    tile_shape_of(variant & 7, bm, bn, stages);
    const int flags = (int)(variant >> 3);
    hipStream_t s = dcc::as_stream(stream);
    const int np = (int)nprob;
    if (bm == 64 && bn == 64)
        return stages == 2 ? GroupedLauncher<64, 64, 2>::run(flags, nworkgroups, probs, np, total_tiles, s)
                           : GroupedLauncher<64, 64, 1>::run(flags, nworkgroups, probs, np, total_tiles, s);
    if (bm == 128 && bn == 64)
        return stages == 2 ? GroupedLauncher<128, 64, 2>::run(flags, nworkgroups, probs, np, total_tiles, s)
                           : GroupedLauncher<128, 64, 1>::run(flags, nworkgroups, probs, np, total_tiles, s);
    if (bm == 64 && bn == 128)
        return stages == 2 ? GroupedLauncher<64, 128, 2>::run(flags, nworkgroups, probs, np, total_tiles, s)
                           : GroupedLauncher<64, 128, 1>::run(flags, nworkgroups, probs, np, total_tiles, s);
    return stages == 2 ? GroupedLauncher<128, 128, 2>::run(flags, nworkgroups, probs, np, total_tiles, s)
                       : GroupedLauncher<128, 128, 1>::run(flags, nworkgroups, probs, np, total_tiles, s);
}

int dcc_graph_capture_begin(dcc_stream_t stream) {
    DCC_CHECK(stream != nullptr, "capture needs a non-default stream");
    DCC_HIP(hipStreamBeginCapture(dcc::as_stream(stream), hipStreamCaptureModeRelaxed));
    return 0;
}

int dcc_graph_capture_end(dcc_stream_t stream, dcc_graph_t *graph) {
    DCC_CHECK(stream != nullptr && graph != nullptr, "null argument");
    hipGraph_t g = nullptr;
    DCC_HIP(hipStreamEndCapture(dcc::as_stream(stream), &g));
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    DCC_HIP(err);
    *graph = reinterpret_cast<dcc_graph_t>(e);
    return 0;
}

int dcc_graph_launch(dcc_graph_t graph, dcc_stream_t stream) {
    DCC_CHECK(graph != nullptr, "null graph");
    DCC_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), dcc::as_stream(stream)));
    return 0;
}

int dcc_graph_destroy(dcc_graph_t graph) {
    if (graph != nullptr) DCC_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)));
    return 0;
}

}  // extern "C"
