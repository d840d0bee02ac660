// Batched gather-GEMM on FP64 MFMA for the triples contractions.
//
// Every contraction of the T3 residual is a GEMM once each tensor index is
// assigned to one of the M, N or K dimensions; the permuted index orders of
// the canonical i<=j<=k blocks only change *where* an element lives.  The
// kernel therefore takes, per operand and dimension, an index map (a lookup
// table or a three-level mixed-radix decomposition, see dcc_idxmap) and stages
// BMxBK / BKxBN tiles through LDS, loading along whichever of the two indices
// is contiguous in memory so every HBM/L2 access is coalesced.  This is the
// "scatter matrix" formulation of tensor contraction (as in TBLIS) mapped onto
// v_mfma_f64_16x16x4f64.
//
// Tile: BM x BN per workgroup of 256 threads (4 wave64), each wave owning a
// (BM/2) x (BN/2) sub-tile as (BM/32) x (BN/32) MFMA 16x16 blocks; BK = 16
// with a double-buffered LDS pipeline (one barrier per K step).
#include "dcc_common.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kBK = 16;
constexpr int kGroupM = 8;  // tiles of M visited together (L2 reuse of B panels)

__device__ __forceinline__ int64_t map_off(const dcc_idxmap &mp, uint32_t x, int64_t batch) {
    if (mp.table != nullptr) return mp.table[batch * mp.table_bstride + x];
    const uint32_t e2 = (uint32_t)mp.e2;
    const uint32_t e12 = (uint32_t)(mp.e1 * mp.e2);
    const uint32_t q0 = x / e12;
    const uint32_t r0 = x - q0 * e12;
    const uint32_t q1 = r0 / e2;
    const uint32_t q2 = r0 - q1 * e2;
    return (int64_t)q0 * mp.s0 + (int64_t)q1 * mp.s1 + (int64_t)q2 * mp.s2;
}

struct ContractArgs {
    int M, N, K;
    double alpha, beta;
    const double *A;
    dcc_idxmap am, ak;
    const double *B;
    dcc_idxmap bk, bn;
    double *C;
    dcc_idxmap cm, cn;
    const dcc_batch *batches;
    int tiles_m, tiles_n;
    int64_t total_tiles;
};

// Spread consecutive workgroups of one XCD over a contiguous range of tiles:
// the hardware deals workgroup i to XCD i % 8, so without this remap the
// neighbours that share an A or B panel would sit in eight different L2s.
__device__ __forceinline__ int64_t xcd_remap(int64_t L, int64_t total) {
    const int64_t nx = 8;
    const int64_t q = total / nx, r = total % nx;
    const int64_t xcd = L % nx, local = L / nx;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

template <int BM, int BN, bool A_KFAST, bool B_NFAST>
__global__ __launch_bounds__(kThreads) void contract_f64_kernel(ContractArgs p) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = BM + 1, LDB = BN + 1;
    constexpr int A_PER_T = BM * kBK / kThreads;
    constexpr int B_PER_T = BN * kBK / kThreads;
    __shared__ double As[2][kBK][LDA];
    __shared__ double Bs[2][kBK][LDB];

    const int tid = threadIdx.x;
    const int64_t L = xcd_remap((int64_t)blockIdx.x, p.total_tiles);
    const int64_t per_batch = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t batch = L / per_batch;
    const int64_t pid = L - batch * per_batch;
    // grouped ordering: kGroupM row tiles advance together along N
    const int64_t group_span = (int64_t)kGroupM * p.tiles_n;
    const int64_t group = pid / group_span;
    const int first_m = (int)(group * kGroupM);
    const int gsize = min(p.tiles_m - first_m, kGroupM);
    const int64_t in_group = pid - group * group_span;
    const int tm = first_m + (int)(in_group % gsize);
    const int tn = (int)(in_group / gsize);

    const dcc_batch bt = p.batches[batch];
    const int64_t m0 = bt.m_begin + (int64_t)tm * BM;
    if (m0 >= bt.m_end) return;
    const int64_t m_end = bt.m_end;
    const int n0 = tn * BN;

    // ---- loader coordinates (fixed per thread) ----
    int a_mm[A_PER_T], a_kk[A_PER_T];
    int64_t a_row[A_PER_T];
    bool a_ok[A_PER_T];
#pragma unroll
    for (int r = 0; r < A_PER_T; ++r) {
        if (A_KFAST) {
            a_kk[r] = tid % kBK;
            a_mm[r] = tid / kBK + r * (kThreads / kBK);
        } else {
            a_mm[r] = tid % BM;
            a_kk[r] = tid / BM + r * (kThreads / BM);
        }
        const int64_t m = m0 + a_mm[r];
        a_ok[r] = m < m_end;
        a_row[r] = a_ok[r] ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    }
    int b_nn[B_PER_T], b_kk[B_PER_T];
    int64_t b_col[B_PER_T];
    bool b_ok[B_PER_T];
#pragma unroll
    for (int r = 0; r < B_PER_T; ++r) {
        if (B_NFAST) {
            b_nn[r] = tid % BN;
            b_kk[r] = tid / BN + r * (kThreads / BN);
        } else {
            b_kk[r] = tid % kBK;
            b_nn[r] = tid / kBK + r * (kThreads / kBK);
        }
        const int n = n0 + b_nn[r];
        b_ok[r] = n < p.N;
        b_col[r] = b_ok[r] ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    }

    double a_reg[A_PER_T], b_reg[B_PER_T];
    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r) {
            const int k = k0 + a_kk[r];
            double v = 0.0;
            if (a_ok[r] && k < p.K) v = p.A[a_row[r] + map_off(p.ak, (uint32_t)k, batch)];
            a_reg[r] = v;
        }
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r) {
            const int k = k0 + b_kk[r];
            double v = 0.0;
            if (b_ok[r] && k < p.K) v = p.B[b_col[r] + map_off(p.bk, (uint32_t)k, batch)];
            b_reg[r] = v;
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r) As[buf][a_kk[r]][a_mm[r]] = a_reg[r];
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r) Bs[buf][b_kk[r]][b_nn[r]] = b_reg[r];
    };

    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int lrow = lane % 16, lk = lane / 16;
    dbl4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int nk = (p.K + kBK - 1) / kBK;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tiles((kt + 1) * kBK);
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
        if (kt + 1 < nk) store_tiles(buf ^ 1);
        __syncthreads();
    }

    // ---- epilogue: for v_mfma_f64_16x16x4f64 lane holds D[lane/16 + 4*r][lane%16], r = 0..3 ----
    int64_t ccol[FN];
    bool cok[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int n = n0 + wn * WN + j * 16 + lrow;
        cok[j] = n < p.N;
        ccol[j] = cok[j] ? map_off(p.cn, (uint32_t)n, batch) : 0;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t m = m0 + wm * WM + i * 16 + lk + 4 * r;
            if (m >= m_end) continue;
            const int64_t crow = bt.c_off + map_off(p.cm, (uint32_t)m, batch);
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (!cok[j]) continue;
                double *cp = p.C + crow + ccol[j];
                double v = p.alpha * acc[i][j][r];
                if (p.beta != 0.0) v += p.beta * (*cp);
                *cp = v;
            }
        }
    }
}

// Innermost stride of an index map (stride of consecutive indices), 0 if a table.
int64_t inner_stride(const dcc_idxmap &m) {
    if (m.table != nullptr) return 0;
    if (m.e2 > 1) return m.s2;
    if (m.e1 > 1) return m.s1;
    return m.s0;
}

template <int BM, int BN>
int launch(const ContractArgs &args, bool a_kfast, bool b_nfast, hipStream_t stream) {
    const dim3 grid((unsigned)args.total_tiles), block(kThreads);
    if (a_kfast && b_nfast)
        hipLaunchKernelGGL((contract_f64_kernel<BM, BN, true, true>), grid, block, 0, stream, args);
    else if (a_kfast)
        hipLaunchKernelGGL((contract_f64_kernel<BM, BN, true, false>), grid, block, 0, stream, args);
    else if (b_nfast)
        hipLaunchKernelGGL((contract_f64_kernel<BM, BN, false, true>), grid, block, 0, stream, args);
    else
        hipLaunchKernelGGL((contract_f64_kernel<BM, BN, false, false>), grid, block, 0, stream, args);
    DCC_LAUNCH_CHECK();
    return 0;
}

bool valid_map(const dcc_idxmap &m) {
    if (m.table != nullptr) return m.table_bstride >= 0;
    return m.e1 >= 1 && m.e2 >= 1 && m.e1 * m.e2 < (int64_t)1 << 32;
}

}  // namespace

extern "C" int dcc_contract_f64(int64_t M, int64_t N, int64_t K, double alpha, double beta, const double *A,
                                dcc_idxmap am, dcc_idxmap ak, const double *B, dcc_idxmap bk, dcc_idxmap bn,
                                double *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches,
                                int64_t nbatch, int64_t max_rows, dcc_stream_t stream) {
    DCC_CHECK(M >= 0 && N >= 0 && K >= 0 && nbatch >= 0 && max_rows >= 0, "negative extent");
    DCC_CHECK(M < ((int64_t)1 << 31) && N < ((int64_t)1 << 31) && K < ((int64_t)1 << 31),
              "extent exceeds 2^31");
    DCC_CHECK(valid_map(am) && valid_map(ak) && valid_map(bk) && valid_map(bn) && valid_map(cm) &&
                  valid_map(cn),
              "invalid index map");
    if (M == 0 || N == 0 || nbatch == 0 || max_rows == 0) return 0;
    DCC_CHECK(A != nullptr && B != nullptr && C != nullptr && batches != nullptr, "null pointer");

    ContractArgs args;
    args.M = (int)M;
    args.N = (int)N;
    args.K = (int)K;
    args.alpha = alpha;
    args.beta = beta;
    args.A = A;
    args.am = am;
    args.ak = ak;
    args.B = B;
    args.bk = bk;
    args.bn = bn;
    args.C = C;
    args.cm = cm;
    args.cn = cn;
    args.batches = batches;

    // choose the contiguous index of each operand as the lane-fastest one
    const bool a_kfast = inner_stride(ak) == 1 || inner_stride(am) != 1;
    const bool b_nfast = inner_stride(bn) == 1 || inner_stride(bk) != 1;

    constexpr int BM = 64, BN = 64;
    args.tiles_m = (int)((max_rows + BM - 1) / BM);
    args.tiles_n = (int)((N + BN - 1) / BN);
    args.total_tiles = (int64_t)nbatch * args.tiles_m * args.tiles_n;
    DCC_CHECK(args.total_tiles < ((int64_t)1 << 31), "too many tiles");
    if (K == 0) {
        // C = beta*C: handled by the same kernel with an empty K loop
    }
    return launch<BM, BN>(args, a_kfast, b_nfast, dcc::as_stream(stream));
}
