// Grouped persistent gather-GEMM: one launch runs a list of independent contractions (for example
// the same residual term of many small CC systems).  Each problem is a dcc_contract_f64 call in
// descriptor form; the library tiles every problem, concatenates the tile ranges, and a fixed grid of
// workgroups loops over the global tile list (tile t, t + gridDim.x, ...), locating the problem of
// each tile by binary search over the per-problem tile offsets.  Every workgroup leaves the loop when
// the list is exhausted, so the launch always drains.  Problems of one launch must write disjoint C
// regions (no two problems accumulate into the same element).
#include "dcc_common.h"

#include <algorithm>
#include <vector>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int kThreads = 256;
constexpr int kBK = 16;
constexpr int kGroupM = 8;

__device__ __forceinline__ int64_t gmap(const dcc_idxmap &mp, uint32_t x, int64_t batch) {
    if (mp.table != nullptr) return mp.table[batch * mp.table_bstride + x];
    const uint32_t e2 = (uint32_t)mp.e2;
    const uint32_t e12 = (uint32_t)(mp.e1 * mp.e2);
    const uint32_t q0 = x / e12;
    const uint32_t r0 = x - q0 * e12;
    const uint32_t q1 = r0 / e2;
    const uint32_t q2 = r0 - q1 * e2;
    return (int64_t)q0 * mp.s0 + (int64_t)q1 * mp.s1 + (int64_t)q2 * mp.s2;
}

template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void grouped_kernel(const dcc_gemm_problem *__restrict__ probs, int nprob,
                                                           int64_t total_tiles) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = BM + 1, LDB = BN + 1;
    constexpr int A_PER_T = BM * kBK / kThreads;
    constexpr int B_PER_T = BN * kBK / kThreads;
    constexpr int MAIN = 2 * kBK * LDA + 2 * kBK * LDB;
    constexpr int SMEM = MAIN > BN * (BM + 1) ? MAIN : BN * (BM + 1);
    __shared__ double smem[SMEM];
    double (*As)[kBK][LDA] = reinterpret_cast<double (*)[kBK][LDA]>(smem);
    double (*Bs)[kBK][LDB] = reinterpret_cast<double (*)[kBK][LDB]>(smem + 2 * kBK * LDA);
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int lrow = lane % 16, lk = lane / 16;

    for (int64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        // problem owning tile t: last p with tile_begin <= t
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (probs[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const dcc_gemm_problem &p = probs[lo];
        const bool a_kfast = (p.flags & 1) != 0, b_nfast = (p.flags & 2) != 0, c_mfast = (p.flags & 4) != 0;
        const int64_t local = t - p.tile_begin;
        const int64_t per_batch = p.tiles_m * p.tiles_n;
        const int64_t batch = local / per_batch;
        const int64_t pid = local - batch * per_batch;
        const int64_t group_span = (int64_t)kGroupM * p.tiles_n;
        const int64_t group = pid / group_span;
        const int first_m = (int)(group * kGroupM);
        const int gsize = (int)min((int64_t)kGroupM, p.tiles_m - first_m);
        const int64_t in_group = pid - group * group_span;
        const int tm = first_m + (int)(in_group % gsize);
        const int tn = (int)(in_group / gsize);
        const dcc_batch bt = p.batches[batch];
        const int64_t m0 = bt.m_begin + (int64_t)tm * BM;
        if (m0 >= bt.m_end) continue;  // uniform across the workgroup
        const int64_t m_end = bt.m_end;
        const int n0 = tn * BN;
        const int N = (int)p.N, K = (int)p.K;
        const double *A = static_cast<const double *>(p.A);
        const double *B = static_cast<const double *>(p.B);
        double *C = static_cast<double *>(p.C);

        const int a_kk = a_kfast ? (tid % kBK) : (tid / 16);
        const int a_m_base = a_kfast ? (tid / kBK) : (tid % 16);
        const int b_kk = b_nfast ? (tid / 16) : (tid % kBK);
        const int b_n_base = b_nfast ? (tid % 16) : (tid / kBK);
        int64_t a_row[A_PER_T];
        uint32_t a_ok = 0;
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r) {
            const int64_t m = m0 + a_m_base + 16 * r;
            const bool ok = m < m_end;
            a_ok |= (ok ? 1u : 0u) << r;
            a_row[r] = ok ? bt.a_off + gmap(p.am, (uint32_t)m, batch) : 0;
        }
        int64_t b_col[B_PER_T];
        uint32_t b_ok = 0;
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r) {
            const int n = n0 + b_n_base + 16 * r;
            const bool ok = n < N;
            b_ok |= (ok ? 1u : 0u) << r;
            b_col[r] = ok ? bt.b_off + gmap(p.bn, (uint32_t)n, batch) : 0;
        }
        const dcc::KMap akm = dcc::make_kmap(p.ak), bkm = dcc::make_kmap(p.bk);
        double a_reg[A_PER_T], b_reg[B_PER_T];
        // branch-free loads: out-of-range elements read offset 0 and are replaced by zero
        auto load_tiles = [&](int k0) {
            const uint32_t ka = (uint32_t)(k0 + a_kk), kb = (uint32_t)(k0 + b_kk);
            const bool ak_ok = ka < (uint32_t)K;
            const bool bk_ok = kb < (uint32_t)K;
            const int64_t aoff = dcc::koff(akm, ak_ok ? ka : 0u);
            const int64_t boff = dcc::koff(bkm, bk_ok ? kb : 0u);
#pragma unroll
            for (int r = 0; r < A_PER_T; ++r) {
                const bool ok = ak_ok && ((a_ok >> r) & 1u);
                const double x = A[ok ? a_row[r] + aoff : 0];
                a_reg[r] = ok ? x : 0.0;
            }
#pragma unroll
            for (int r = 0; r < B_PER_T; ++r) {
                const bool ok = bk_ok && ((b_ok >> r) & 1u);
                const double x = B[ok ? b_col[r] + boff : 0];
                b_reg[r] = ok ? x : 0.0;
            }
        };
        auto store_tiles = [&](int buf) {
#pragma unroll
            for (int r = 0; r < A_PER_T; ++r) As[buf][a_kk][a_m_base + 16 * r] = a_reg[r];
#pragma unroll
            for (int r = 0; r < B_PER_T; ++r) Bs[buf][b_kk][b_n_base + 16 * r] = b_reg[r];
        };
        dbl4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0, 0, 0, 0};
        const int nk = (K + kBK - 1) / kBK;
        __syncthreads();  // LDS of the previous tile is no longer read
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int buf = kt & 1;
            load_tiles((kt + 1) * kBK);  // past the last tile: offset-0 reads into the idle buffer
#pragma unroll
            for (int ks = 0; ks < kBK / 4; ++ks) {
                const int kr = ks * 4 + lk;
                double af[FM], bf[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i) af[i] = As[buf][kr][wm * WM + i * 16 + lrow];
#pragma unroll
                for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kr][wn * WN + j * 16 + lrow];
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
            store_tiles(buf ^ 1);
            __syncthreads();
        }
        if (c_mfast) {
            double *Cs = smem;
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Cs[(wn * WN + j * 16 + lrow) * (BM + 1) + wm * WM + i * 16 + lk + 4 * r] = acc[i][j][r];
            __syncthreads();
            constexpr int NSTEP = kThreads / BM;
            const int ml = tid % BM;
            const int64_t m = m0 + ml;
            if (m < m_end) {
                const int64_t crow = bt.c_off + gmap(p.cm, (uint32_t)m, batch);
                for (int nl = tid / BM; nl < BN; nl += NSTEP) {
                    const int n = n0 + nl;
                    if (n >= N) break;
                    double *c = C + crow + gmap(p.cn, (uint32_t)n, batch);
                    double v = p.alpha * Cs[nl * (BM + 1) + ml];
                    if (p.beta != 0.0) v += p.beta * *c;
                    *c = v;
                }
            }
            continue;  // the loop-top barrier protects Cs before the next tile's stores
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t m = m0 + wm * WM + i * 16 + lk + 4 * r;
                if (m >= m_end) continue;
                const int64_t crow = bt.c_off + gmap(p.cm, (uint32_t)m, batch);
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int n = n0 + wn * WN + j * 16 + lrow;
                    if (n >= N) continue;
                    double *c = C + crow + gmap(p.cn, (uint32_t)n, batch);
                    double v = p.alpha * acc[i][j][r];
                    if (p.beta != 0.0) v += p.beta * *c;
                    *c = v;
                }
            }
    }
}

int64_t inner_stride(const dcc_idxmap &m) {
    if (m.table != nullptr) return 0;
    if (m.e2 > 1) return m.s2;
    if (m.e1 > 1) return m.s1;
    return m.s0;
}

bool tile_shape(int64_t shape, int *bm, int *bn) {
    switch (shape) {
        case 0: *bm = 64; *bn = 64; return true;
        case 1: *bm = 128; *bn = 64; return true;
        case 2: *bm = 64; *bn = 128; return true;
        case 3: *bm = 128; *bn = 128; return true;
        default: return false;
    }
}

}  // namespace

extern "C" {

int dcc_grouped_prepare(dcc_gemm_problem *probs, int64_t nprob, int64_t tile_shape_id, int64_t *total_tiles) {
    DCC_CHECK(nprob >= 0 && total_tiles != nullptr, "bad arguments");
    DCC_CHECK(nprob == 0 || probs != nullptr, "null problems");
    int bm = 0, bn = 0;
    DCC_CHECK(tile_shape(tile_shape_id, &bm, &bn), "tile shape must be 0..3");
    int64_t acc = 0;
    for (int64_t i = 0; i < nprob; ++i) {
        dcc_gemm_problem &p = probs[i];
        DCC_CHECK(p.N >= 0 && p.K >= 0 && p.nbatch >= 0 && p.max_rows >= 0, "negative extent");
        DCC_CHECK(p.ak.table == nullptr && p.bk.table == nullptr, "K index maps must be mixed-radix maps");
        p.tiles_m = (p.max_rows + bm - 1) / bm;
        p.tiles_n = (p.N + bn - 1) / bn;
        p.tile_begin = acc;
        const bool a_kfast = inner_stride(p.ak) == 1 || inner_stride(p.am) != 1;
        const bool b_nfast = inner_stride(p.bn) == 1 || inner_stride(p.bk) != 1;
        const bool c_mfast = inner_stride(p.cn) != 1;
        p.flags = (a_kfast ? 1 : 0) | (b_nfast ? 2 : 0) | (c_mfast ? 4 : 0);
        acc += p.nbatch * p.tiles_m * p.tiles_n;
    }
    DCC_CHECK(acc < ((int64_t)1 << 40), "too many tiles");
    *total_tiles = acc;
    return 0;
}

int dcc_contract_grouped_f64(const dcc_gemm_problem *probs, int64_t nprob, int64_t total_tiles,
                             int64_t tile_shape_id, int64_t nworkgroups, dcc_stream_t stream) {
    DCC_CHECK(nprob >= 0 && total_tiles >= 0 && nworkgroups >= 0, "negative argument");
    DCC_CHECK(nprob < (1 << 30), "too many problems");
    if (nprob == 0 || total_tiles == 0) return 0;
    DCC_CHECK(probs != nullptr, "null problems");
    int bm = 0, bn = 0;
    DCC_CHECK(tile_shape(tile_shape_id, &bm, &bn), "tile shape must be 0..3");
    int64_t grid = nworkgroups;
    if (grid == 0) {
        int dev = 0, cus = 0;
        DCC_HIP(hipGetDevice(&dev));
        DCC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = (int64_t)cus * 4;  // persistent: a few resident workgroups per CU
    }
    if (grid > total_tiles) grid = total_tiles;
    const dim3 g((unsigned)grid), b(kThreads);
    hipStream_t s = dcc::as_stream(stream);
    if (bm == 64 && bn == 64)
        hipLaunchKernelGGL((grouped_kernel<64, 64>), g, b, 0, s, probs, (int)nprob, total_tiles);
    else if (bm == 128 && bn == 64)
        hipLaunchKernelGGL((grouped_kernel<128, 64>), g, b, 0, s, probs, (int)nprob, total_tiles);
    else if (bm == 64 && bn == 128)
        hipLaunchKernelGGL((grouped_kernel<64, 128>), g, b, 0, s, probs, (int)nprob, total_tiles);
    else
        hipLaunchKernelGGL((grouped_kernel<128, 128>), g, b, 0, s, probs, (int)nprob, total_tiles);
    DCC_LAUNCH_CHECK();
    return 0;
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
