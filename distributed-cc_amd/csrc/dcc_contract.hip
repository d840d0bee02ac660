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
// with a double-buffered LDS pipeline (one barrier per K step).  Every thread
// loads its A and B elements at one fixed k row of the tile, so the K index
// map is evaluated once per operand and then advanced by BK with carries --
// no integer division inside the K loop.
#include "dcc_common.h"

#include <cstdio>
#include <cstdlib>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Element type -> MFMA instruction and accumulator layout.
//   double: v_mfma_f64_16x16x4f64, lane holds D[lane/16 + 4r][lane%16]
//   float : v_mfma_f32_16x16x4f32, lane holds D[4(lane/16) + r][lane%16]
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef dbl4 acc_t;
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
    typedef flt4 acc_t;
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

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

// K index map in incremental form: offset of k, advanced by kBK per step.
struct KMap {
    int64_t s0, s1, s2;   // strides of the three levels
    int64_t step;         // bk * s2
    int64_t carry1;       // s1 - e2 * s2  (q2 wraps into q1)
    int64_t carry0;       // s0 - e1 * s1  (q1 wraps into q0)
    uint32_t e1, e2;
    uint32_t bk;          // K advance per step
};

struct KIter {
    uint32_t q1, q2;
    int64_t off;
};

__device__ __forceinline__ KIter kinit(const KMap &m, uint32_t k) {
    const uint32_t e12 = m.e1 * m.e2;
    const uint32_t q0 = k / e12;
    const uint32_t r0 = k - q0 * e12;
    KIter it;
    it.q1 = r0 / m.e2;
    it.q2 = r0 - it.q1 * m.e2;
    it.off = (int64_t)q0 * m.s0 + (int64_t)it.q1 * m.s1 + (int64_t)it.q2 * m.s2;
    return it;
}

__device__ __forceinline__ void kadvance(const KMap &m, KIter &it) {
    it.q2 += m.bk;
    it.off += m.step;
    while (it.q2 >= m.e2) {  // at most ceil(kBK / e2) passes
        it.q2 -= m.e2;
        it.off += m.carry1;
        if (++it.q1 >= m.e1) {
            it.q1 = 0;
            it.off += m.carry0;
        }
    }
}

template <typename T>
struct ContractArgs {
    int M, N, K;
    double alpha, beta;
    const T *A;
    const T *B;
    T *C;
    dcc_idxmap am, bn, cm, cn;
    KMap ak, bk;
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

template <int a, int b> struct CMax { static constexpr int v = a > b ? a : b; };

// C_MFAST selects the epilogue: false writes C straight from the accumulators (16 consecutive lanes
// walk n, coalesced when C is contiguous along n); true stages the tile in LDS and writes it with
// consecutive lanes walking m, for outputs whose contiguous index is carried by M.
template <typename T, int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST>
__global__ __launch_bounds__(kThreads) void contract_kernel(const ContractArgs<T> p) {
    typedef Mfma<T> MF;
    typedef typename MF::acc_t acc_t;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = BM + 1, LDB = BN + 1;
    constexpr int A_PER_T = BM * kBK / kThreads;
    constexpr int B_PER_T = BN * kBK / kThreads;
    constexpr int MAIN = 2 * kBK * LDA + 2 * kBK * LDB;
    constexpr int SMEM = C_MFAST ? CMax<MAIN, BN * (BM + 1)>::v : MAIN;
    __shared__ T smem[SMEM];
    T (*As)[kBK][LDA] = reinterpret_cast<T (*)[kBK][LDA]>(smem);
    T (*Bs)[kBK][LDB] = reinterpret_cast<T (*)[kBK][LDB]>(smem + 2 * kBK * LDA);

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

    // ---- loader coordinates: one k row per thread and operand ----
    // K-fast: 16 consecutive lanes walk k; M/N-fast: 16 consecutive lanes walk m/n.
    const int a_kk = A_KFAST ? (tid % kBK) : (tid / 16);
    const int a_m_base = A_KFAST ? (tid / kBK) : (tid % 16);
    const int b_kk = B_NFAST ? (tid / 16) : (tid % kBK);
    const int b_n_base = B_NFAST ? (tid % 16) : (tid / kBK);
    int64_t a_row[A_PER_T];
    uint32_t a_ok = 0;
#pragma unroll
    for (int r = 0; r < A_PER_T; ++r) {
        const int64_t m = m0 + a_m_base + 16 * r;
        const bool ok = m < m_end;
        a_ok |= (ok ? 1u : 0u) << r;
        a_row[r] = ok ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    }
    int64_t b_col[B_PER_T];
    uint32_t b_ok = 0;
#pragma unroll
    for (int r = 0; r < B_PER_T; ++r) {
        const int n = n0 + b_n_base + 16 * r;
        const bool ok = n < p.N;
        b_ok |= (ok ? 1u : 0u) << r;
        b_col[r] = ok ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    }
    KIter ka = kinit(p.ak, (uint32_t)a_kk);
    KIter kb = kinit(p.bk, (uint32_t)b_kk);

    T a_reg[A_PER_T], b_reg[B_PER_T];
    auto load_tiles = [&](int k0) {
        const bool ak_ok = k0 + a_kk < p.K;
        const bool bk_ok = k0 + b_kk < p.K;
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r)
            a_reg[r] = (ak_ok && ((a_ok >> r) & 1u)) ? p.A[a_row[r] + ka.off] : T(0);
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r)
            b_reg[r] = (bk_ok && ((b_ok >> r) & 1u)) ? p.B[b_col[r] + kb.off] : T(0);
        kadvance(p.ak, ka);
        kadvance(p.bk, kb);
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r) As[buf][a_kk][a_m_base + 16 * r] = a_reg[r];
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r) Bs[buf][b_kk][b_n_base + 16 * r] = b_reg[r];
    };

    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int lrow = lane % 16, lk = lane / 16;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

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
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = As[buf][kr][wm * WM + i * 16 + lrow];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kr][wn * WN + j * 16 + lrow];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = MF::step(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < nk) store_tiles(buf ^ 1);
        __syncthreads();
    }

    if constexpr (C_MFAST) {
        // ---- LDS-staged epilogue: tile stored n-major, then written with lanes along m ----
        T *Cs = smem;  // the main loop ended with a barrier, so the staging buffers are free
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Cs[(wn * WN + j * 16 + lrow) * (BM + 1) + wm * WM + i * 16 + MF::row(lk, r)] = acc[i][j][r];
        __syncthreads();
        constexpr int NSTEP = kThreads / BM;
        const int ml = tid % BM;
        const int64_t m = m0 + ml;
        if (m >= m_end) return;
        const int64_t crow = bt.c_off + map_off(p.cm, (uint32_t)m, batch);
        for (int nl = tid / BM; nl < BN; nl += NSTEP) {
            const int n = n0 + nl;
            if (n >= p.N) break;
            T *c = p.C + crow + map_off(p.cn, (uint32_t)n, batch);
            double v = p.alpha * (double)Cs[nl * (BM + 1) + ml];
            if (p.beta != 0.0) v += p.beta * (double)*c;
            *c = (T)v;
        }
        return;
    }

    // ---- epilogue: accumulator row of (lane, r) from Mfma<T>::row, column lane%16 ----
    int64_t ccol[FN];
    uint32_t cok = 0;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int n = n0 + wn * WN + j * 16 + lrow;
        const bool ok = n < p.N;
        cok |= (ok ? 1u : 0u) << j;
        ccol[j] = ok ? map_off(p.cn, (uint32_t)n, batch) : 0;
    }
    // all C reads are issued before any store so their latencies overlap
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        int64_t crow[4];
        uint32_t rok = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t m = m0 + wm * WM + i * 16 + MF::row(lk, r);
            const bool ok = m < m_end;
            rok |= (ok ? 1u : 0u) << r;
            crow[r] = ok ? bt.c_off + map_off(p.cm, (uint32_t)m, batch) : 0;
        }
        double old[4][FN];  // C read in its own precision, combined in FP64
        if (p.beta != 0.0) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    old[r][j] = (((rok >> r) & (cok >> j)) & 1u) ? (double)p.C[crow[r] + ccol[j]] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (!(((rok >> r) & (cok >> j)) & 1u)) continue;
                double v = p.alpha * (double)acc[i][j][r];
                if (p.beta != 0.0) v += p.beta * old[r][j];
                p.C[crow[r] + ccol[j]] = (T)v;
            }
    }
}

// ---------------------------------------------------------------------------------------------
// Split-BF16x3: FP32 operands split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to
// nearest even); each product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation (the lo*lo term, ~2^-16 relative, is dropped).  Same index maps, batching and
// epilogues as contract_kernel; BK = 32 and the hi/lo tiles are stored k-contiguous in LDS so a
// lane reads its 8 consecutive k values of a row (A) or column (B) as one 16-byte load:
//   A fragment: lane holds A[lane%16][8*(lane/16) + 0..7];  B fragment: B[8*(lane/16) + 0..7][lane%16];
//   accumulator: lane holds D[4*(lane/16) + r][lane%16] (as v_mfma_f32_16x16x4f32).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kBK3 = 32;

__device__ __forceinline__ unsigned short bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST>
__global__ __launch_bounds__(kThreads) void contract_bf16x3_kernel(const ContractArgs<float> p) {
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int KP = kBK3 + 8;                  // padded k stride: conflict-free 16-byte reads
    constexpr int A_PER_T = BM * kBK3 / kThreads;
    constexpr int B_PER_T = BN * kBK3 / kThreads;
    constexpr int A_ELEMS = 2 * 2 * BM * KP, B_ELEMS = 2 * 2 * BN * KP;   // [buf][hi/lo][row][k]
    constexpr int MAIN_BYTES = (A_ELEMS + B_ELEMS) * 2;
    constexpr int EPI_BYTES = C_MFAST ? BN * (BM + 1) * 4 : 0;
    __shared__ __align__(16) unsigned char smem[MAIN_BYTES > EPI_BYTES ? MAIN_BYTES : EPI_BYTES];
    unsigned short *As = reinterpret_cast<unsigned short *>(smem);
    unsigned short *Bs = As + A_ELEMS;

    const int tid = threadIdx.x;
    const int64_t L = xcd_remap((int64_t)blockIdx.x, p.total_tiles);
    const int64_t per_batch = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t batch = L / per_batch;
    const int64_t pid = L - batch * per_batch;
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

    // one k row per thread and operand; rows/columns base + 8r
    const int a_kk = A_KFAST ? (tid % kBK3) : (tid / 8);
    const int a_m_base = A_KFAST ? (tid / kBK3) : (tid % 8);
    const int b_kk = B_NFAST ? (tid / 8) : (tid % kBK3);
    const int b_n_base = B_NFAST ? (tid % 8) : (tid / kBK3);
    int64_t a_row[A_PER_T];
    uint32_t a_ok = 0;
#pragma unroll
    for (int r = 0; r < A_PER_T; ++r) {
        const int64_t m = m0 + a_m_base + 8 * r;
        const bool ok = m < m_end;
        a_ok |= (ok ? 1u : 0u) << r;
        a_row[r] = ok ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    }
    int64_t b_col[B_PER_T];
    uint32_t b_ok = 0;
#pragma unroll
    for (int r = 0; r < B_PER_T; ++r) {
        const int n = n0 + b_n_base + 8 * r;
        const bool ok = n < p.N;
        b_ok |= (ok ? 1u : 0u) << r;
        b_col[r] = ok ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    }
    KIter ka = kinit(p.ak, (uint32_t)a_kk);
    KIter kb = kinit(p.bk, (uint32_t)b_kk);

    float a_reg[A_PER_T], b_reg[B_PER_T];
    auto load_tiles = [&](int k0) {
        const bool ak_ok = k0 + a_kk < p.K;
        const bool bk_ok = k0 + b_kk < p.K;
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r)
            a_reg[r] = (ak_ok && ((a_ok >> r) & 1u)) ? p.A[a_row[r] + ka.off] : 0.0f;
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r)
            b_reg[r] = (bk_ok && ((b_ok >> r) & 1u)) ? p.B[b_col[r] + kb.off] : 0.0f;
        kadvance(p.ak, ka);
        kadvance(p.bk, kb);
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int r = 0; r < A_PER_T; ++r) {
            const unsigned short hi = bf16_rne(a_reg[r]);
            const unsigned short lo = bf16_rne(a_reg[r] - __uint_as_float((uint32_t)hi << 16));
            const int row = a_m_base + 8 * r;
            As[((buf * 2 + 0) * BM + row) * KP + a_kk] = hi;
            As[((buf * 2 + 1) * BM + row) * KP + a_kk] = lo;
        }
#pragma unroll
        for (int r = 0; r < B_PER_T; ++r) {
            const unsigned short hi = bf16_rne(b_reg[r]);
            const unsigned short lo = bf16_rne(b_reg[r] - __uint_as_float((uint32_t)hi << 16));
            const int col = b_n_base + 8 * r;
            Bs[((buf * 2 + 0) * BN + col) * KP + b_kk] = hi;
            Bs[((buf * 2 + 1) * BN + col) * KP + b_kk] = lo;
        }
    };

    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int lrow = lane % 16, lk = lane / 16;
    flt4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = flt4{0, 0, 0, 0};

    const int nk = (p.K + kBK3 - 1) / kBK3;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tiles((kt + 1) * kBK3);
        bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int row = wm * WM + i * 16 + lrow;
            ah[i] = *reinterpret_cast<const bf16x8 *>(&As[((buf * 2 + 0) * BM + row) * KP + 8 * lk]);
            al[i] = *reinterpret_cast<const bf16x8 *>(&As[((buf * 2 + 1) * BM + row) * KP + 8 * lk]);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = wn * WN + j * 16 + lrow;
            bh[j] = *reinterpret_cast<const bf16x8 *>(&Bs[((buf * 2 + 0) * BN + col) * KP + 8 * lk]);
            bl[j] = *reinterpret_cast<const bf16x8 *>(&Bs[((buf * 2 + 1) * BN + col) * KP + 8 * lk]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if (kt + 1 < nk) store_tiles(buf ^ 1);
        __syncthreads();
    }

    if constexpr (C_MFAST) {
        float *Cs = reinterpret_cast<float *>(smem);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Cs[(wn * WN + j * 16 + lrow) * (BM + 1) + wm * WM + i * 16 + 4 * lk + r] = acc[i][j][r];
        __syncthreads();
        constexpr int NSTEP = kThreads / BM;
        const int ml = tid % BM;
        const int64_t m = m0 + ml;
        if (m >= m_end) return;
        const int64_t crow = bt.c_off + map_off(p.cm, (uint32_t)m, batch);
        for (int nl = tid / BM; nl < BN; nl += NSTEP) {
            const int n = n0 + nl;
            if (n >= p.N) break;
            float *c = p.C + crow + map_off(p.cn, (uint32_t)n, batch);
            double v = p.alpha * (double)Cs[nl * (BM + 1) + ml];
            if (p.beta != 0.0) v += p.beta * (double)*c;
            *c = (float)v;
        }
        return;
    }
    int64_t ccol[FN];
    uint32_t cok = 0;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int n = n0 + wn * WN + j * 16 + lrow;
        const bool ok = n < p.N;
        cok |= (ok ? 1u : 0u) << j;
        ccol[j] = ok ? map_off(p.cn, (uint32_t)n, batch) : 0;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        int64_t crow[4];
        uint32_t rok = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t m = m0 + wm * WM + i * 16 + 4 * lk + r;
            const bool ok = m < m_end;
            rok |= (ok ? 1u : 0u) << r;
            crow[r] = ok ? bt.c_off + map_off(p.cm, (uint32_t)m, batch) : 0;
        }
        double old[4][FN];
        if (p.beta != 0.0) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    old[r][j] = (((rok >> r) & (cok >> j)) & 1u) ? (double)p.C[crow[r] + ccol[j]] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (!(((rok >> r) & (cok >> j)) & 1u)) continue;
                double v = p.alpha * (double)acc[i][j][r];
                if (p.beta != 0.0) v += p.beta * old[r][j];
                p.C[crow[r] + ccol[j]] = (float)v;
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

KMap make_kmap(const dcc_idxmap &m, int bk = kBK) {
    KMap k;
    k.s0 = m.s0;
    k.s1 = m.s1;
    k.s2 = m.s2;
    k.e1 = (uint32_t)m.e1;
    k.e2 = (uint32_t)m.e2;
    k.step = (int64_t)bk * m.s2;
    k.bk = (uint32_t)bk;
    k.carry1 = m.s1 - m.e2 * m.s2;
    k.carry0 = m.s0 - m.e1 * m.s1;
    return k;
}

template <typename T, int BM, int BN, bool C_MFAST>
void launch_variant(ContractArgs<T> &args, bool a_kfast, bool b_nfast, dim3 grid, dim3 block, hipStream_t stream) {
    if (a_kfast && b_nfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, true, true, C_MFAST>), grid, block, 0, stream, args);
    else if (a_kfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, true, false, C_MFAST>), grid, block, 0, stream, args);
    else if (b_nfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, false, true, C_MFAST>), grid, block, 0, stream, args);
    else
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, false, false, C_MFAST>), grid, block, 0, stream, args);
}

template <typename T, int BM, int BN>
int launch(ContractArgs<T> &args, int64_t max_rows, bool a_kfast, bool b_nfast, bool c_mfast, hipStream_t stream) {
    args.tiles_m = (int)((max_rows + BM - 1) / BM);
    args.tiles_n = (int)((args.N + BN - 1) / BN);
    const int64_t nbatch = args.total_tiles;  // caller passes nbatch here
    args.total_tiles = nbatch * args.tiles_m * args.tiles_n;
    if (args.total_tiles >= ((int64_t)1 << 31)) {
        dcc::set_error("dcc_contract_f64: too many tiles");
        return 2;
    }
    const dim3 grid((unsigned)args.total_tiles), block(kThreads);
    if (c_mfast)
        launch_variant<T, BM, BN, true>(args, a_kfast, b_nfast, grid, block, stream);
    else
        launch_variant<T, BM, BN, false>(args, a_kfast, b_nfast, grid, block, stream);
    DCC_LAUNCH_CHECK();
    return 0;
}

bool valid_map(const dcc_idxmap &m) {
    if (m.table != nullptr) return m.table_bstride >= 0;
    return m.e1 >= 1 && m.e2 >= 1 && m.e1 * m.e2 < (int64_t)1 << 32;
}

template <typename T>
int contract_entry(int64_t M, int64_t N, int64_t K, double alpha, double beta, const T *A, dcc_idxmap am,
                   dcc_idxmap ak, const T *B, dcc_idxmap bk, dcc_idxmap bn, T *C, dcc_idxmap cm, dcc_idxmap cn,
                   const dcc_batch *batches, int64_t nbatch, int64_t max_rows, dcc_stream_t stream,
                   int64_t tile_shape = -1) {
    DCC_CHECK(M >= 0 && N >= 0 && K >= 0 && nbatch >= 0 && max_rows >= 0, "negative extent");
    DCC_CHECK(M < ((int64_t)1 << 31) && N < ((int64_t)1 << 31) && K < ((int64_t)1 << 31),
              "extent exceeds 2^31");
    DCC_CHECK(valid_map(am) && valid_map(ak) && valid_map(bk) && valid_map(bn) && valid_map(cm) &&
                  valid_map(cn),
              "invalid index map");
    DCC_CHECK(ak.table == nullptr && bk.table == nullptr, "K index maps must be mixed-radix maps");
    if (M == 0 || N == 0 || nbatch == 0 || max_rows == 0) return 0;
    DCC_CHECK(A != nullptr && B != nullptr && C != nullptr && batches != nullptr, "null pointer");

    ContractArgs<T> args;
    args.M = (int)M;
    args.N = (int)N;
    args.K = (int)K;
    args.alpha = alpha;
    args.beta = beta;
    args.A = A;
    args.B = B;
    args.C = C;
    args.am = am;
    args.bn = bn;
    args.cm = cm;
    args.cn = cn;
    args.ak = make_kmap(ak);
    args.bk = make_kmap(bk);
    args.batches = batches;
    args.total_tiles = nbatch;

    // choose the contiguous index of each operand as the lane-fastest one
    const bool a_kfast = inner_stride(ak) == 1 || inner_stride(am) != 1;
    const bool b_nfast = inner_stride(bn) == 1 || inner_stride(bk) != 1;
    // C written along m (through LDS) unless it is contiguous along n
    const bool c_mfast = inner_stride(cn) != 1;
    hipStream_t s = dcc::as_stream(stream);
    // Tile shape: DCC_CONTRACT_TILE=128x128|128x64|64x128|64x64 overrides the default.
    static const char *tile_env = getenv("DCC_CONTRACT_TILE");
    // 64x64 keeps four waves per SIMD resident, which hides HBM latency better
    // than the larger register tiles (measured on MI355X for every triples GEMM).
    int tbm = 64, tbn = 64;
    if (tile_shape >= 0) {
        DCC_CHECK(tile_shape <= 3, "tile shape must be 0..3");
        const int shapes[4][2] = {{64, 64}, {128, 64}, {64, 128}, {128, 128}};
        tbm = shapes[tile_shape][0];
        tbn = shapes[tile_shape][1];
    } else if (tile_env != nullptr && sscanf(tile_env, "%dx%d", &tbm, &tbn) != 2) {
        tbm = 64;
        tbn = 64;
    }
    if (tbm == 128 && tbn == 128) return launch<T, 128, 128>(args, max_rows, a_kfast, b_nfast, c_mfast, s);
    if (tbm == 128 && tbn == 64) return launch<T, 128, 64>(args, max_rows, a_kfast, b_nfast, c_mfast, s);
    if (tbm == 64 && tbn == 128) return launch<T, 64, 128>(args, max_rows, a_kfast, b_nfast, c_mfast, s);
    return launch<T, 64, 64>(args, max_rows, a_kfast, b_nfast, c_mfast, s);
}

template <bool C_MFAST>
void launch_bf16x3(ContractArgs<float> &args, bool a_kfast, bool b_nfast, dim3 grid, dim3 block,
                   hipStream_t stream) {
    if (a_kfast && b_nfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<64, 64, true, true, C_MFAST>), grid, block, 0, stream, args);
    else if (a_kfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<64, 64, true, false, C_MFAST>), grid, block, 0, stream, args);
    else if (b_nfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<64, 64, false, true, C_MFAST>), grid, block, 0, stream, args);
    else
        hipLaunchKernelGGL((contract_bf16x3_kernel<64, 64, false, false, C_MFAST>), grid, block, 0, stream, args);
}

int contract_bf16x3_entry(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                          dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn, float *C,
                          dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                          int64_t max_rows, dcc_stream_t stream) {
    DCC_CHECK(M >= 0 && N >= 0 && K >= 0 && nbatch >= 0 && max_rows >= 0, "negative extent");
    DCC_CHECK(M < ((int64_t)1 << 31) && N < ((int64_t)1 << 31) && K < ((int64_t)1 << 31),
              "extent exceeds 2^31");
    DCC_CHECK(valid_map(am) && valid_map(ak) && valid_map(bk) && valid_map(bn) && valid_map(cm) &&
                  valid_map(cn),
              "invalid index map");
    DCC_CHECK(ak.table == nullptr && bk.table == nullptr, "K index maps must be mixed-radix maps");
    if (M == 0 || N == 0 || nbatch == 0 || max_rows == 0) return 0;
    DCC_CHECK(A != nullptr && B != nullptr && C != nullptr && batches != nullptr, "null pointer");
    ContractArgs<float> args;
    args.M = (int)M;
    args.N = (int)N;
    args.K = (int)K;
    args.alpha = alpha;
    args.beta = beta;
    args.A = A;
    args.B = B;
    args.C = C;
    args.am = am;
    args.bn = bn;
    args.cm = cm;
    args.cn = cn;
    args.ak = make_kmap(ak, kBK3);
    args.bk = make_kmap(bk, kBK3);
    args.batches = batches;
    args.tiles_m = (int)((max_rows + 63) / 64);
    args.tiles_n = (int)((N + 63) / 64);
    args.total_tiles = nbatch * args.tiles_m * args.tiles_n;
    DCC_CHECK(args.total_tiles < ((int64_t)1 << 31), "too many tiles");
    const bool a_kfast = inner_stride(ak) == 1 || inner_stride(am) != 1;
    const bool b_nfast = inner_stride(bn) == 1 || inner_stride(bk) != 1;
    const bool c_mfast = inner_stride(cn) != 1;
    const dim3 grid((unsigned)args.total_tiles), block(kThreads);
    hipStream_t s = dcc::as_stream(stream);
    if (c_mfast)
        launch_bf16x3<true>(args, a_kfast, b_nfast, grid, block, s);
    else
        launch_bf16x3<false>(args, a_kfast, b_nfast, grid, block, s);
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" int dcc_contract_f32_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                                       const float *A, dcc_idxmap am, dcc_idxmap ak, const float *B,
                                       dcc_idxmap bk, dcc_idxmap bn, float *C, dcc_idxmap cm, dcc_idxmap cn,
                                       const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                                       dcc_stream_t stream) {
    return contract_bf16x3_entry(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch, max_rows,
                                 stream);
}

extern "C" int dcc_contract_f64(int64_t M, int64_t N, int64_t K, double alpha, double beta, const double *A,
                                dcc_idxmap am, dcc_idxmap ak, const double *B, dcc_idxmap bk, dcc_idxmap bn,
                                double *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches,
                                int64_t nbatch, int64_t max_rows, dcc_stream_t stream) {
    return contract_entry<double>(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch,
                                  max_rows, stream);
}

extern "C" int dcc_contract_tiled_f64(int64_t M, int64_t N, int64_t K, double alpha, double beta, const double *A,
                                      dcc_idxmap am, dcc_idxmap ak, const double *B, dcc_idxmap bk, dcc_idxmap bn,
                                      double *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches,
                                      int64_t nbatch, int64_t max_rows, int64_t tile_shape, dcc_stream_t stream) {
    return contract_entry<double>(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch,
                                  max_rows, stream, tile_shape);
}

extern "C" int dcc_contract_tiled_f32(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                                      dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn,
                                      float *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches,
                                      int64_t nbatch, int64_t max_rows, int64_t tile_shape, dcc_stream_t stream) {
    return contract_entry<float>(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch,
                                 max_rows, stream, tile_shape);
}

extern "C" int dcc_contract_f32(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                                dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn,
                                float *C, dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches,
                                int64_t nbatch, int64_t max_rows, dcc_stream_t stream) {
    return contract_entry<float>(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch,
                                 max_rows, stream);
}
