// One output tile of the batched gather-GEMM on FP64/FP32 MFMA, shared by the one-launch-per-GEMM
// kernel (dcc_contract.hip) and the persistent grouped kernel (dcc_grouped.hip).
//
// Every contraction of the T3 residual is a GEMM once each tensor index is assigned to one of the
// M, N or K dimensions; the permuted index orders of the canonical i<=j<=k blocks only change *where*
// an element lives.  The tile therefore takes, per operand and dimension, an index map (a lookup
// table or a three-level mixed-radix decomposition, see dcc_idxmap) and stages BMxBK / BKxBN tiles
// through LDS, loading along whichever of the two indices is contiguous in memory so every HBM/L2
// access is coalesced.  This is the "scatter matrix" formulation of tensor contraction (as in TBLIS)
// mapped onto v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.
//
// Tile: BM x BN per workgroup of 256 threads (4 wave64), each wave owning a (BM/2) x (BN/2) sub-tile
// as (BM/32) x (BN/32) MFMA 16x16 blocks; BK = 16 with a double-buffered LDS pipeline (one barrier
// per K step).  Every thread loads its A and B elements at fixed k rows of the tile, so the K offsets
// are decomposed once (by reciprocal multiplication) and then advanced from tile to tile without
// division; the only branches in the loop are uniform (vector or per-element loads).
#pragma once

#include "dcc_common.h"

#include <type_traits>

namespace dcc_gemm {

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

using dcc::KMap;
using dcc::koff;
using dcc::make_kmap;

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
    int vec;  // bit 0 (A), bit 1 (B): the element pairs of a thread may be loaded as one vector
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

// Tile (tm, tn) of tile pid within one batch: kGroupM row tiles advance together along N.
__device__ __forceinline__ void tile_coords(int64_t pid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int64_t group_span = (int64_t)kGroupM * tiles_n;
    const int64_t group = pid / group_span;
    const int first_m = (int)(group * kGroupM);
    const int gsize = min(tiles_m - first_m, kGroupM);
    const int64_t in_group = pid - group * group_span;
    tm = first_m + (int)(in_group % gsize);
    tn = (int)(in_group / gsize);
}

template <int a, int b> struct CMax { static constexpr int v = a > b ? a : b; };

// Loader geometry of one operand with R rows per tile (BK = 16, 256 threads), in element pairs:
// K-contiguous (KFAST): pair (2 kp, 2 kp + 1) with kp = tid % 8, rows tid / 8 + 32 r (NR = R/32, NK = 2);
// row-contiguous: rows (2 i, 2 i + 1) with i = tid % (R/2), k = tid / (R/2) + (512 / R) j (NR = 2,
// NK = R/32).  reg[r][j] holds row row(tid, r) at k kat(tid, r, j).
// NIT: K-offset iterators per thread (one per distinct k of the thread in a tile; the second k of a
// K-contiguous pair is the first plus one whenever pairs are loaded as vectors).  With R = 128 rows a
// row-contiguous operand's k is the wave index plus a constant, the same in every lane, so its offsets
// are computed once per wave in scalar registers.
template <int R, bool KFAST> struct StageP {
    static constexpr int NR = KFAST ? R / 32 : 2;
    static constexpr int NK = KFAST ? 2 : R / 32;
    static constexpr int NIT = KFAST ? 1 : NK;
    __device__ static int row(int tid, int r) { return KFAST ? tid / 8 + 32 * r : 2 * (tid % (R / 2)) + r; }
    __device__ static int kbase(int tid) {
        if constexpr (KFAST) return 2 * (tid % 8);
        else if constexpr (R / 2 == 64) return __builtin_amdgcn_readfirstlane(tid / 64);
        else return tid / (R / 2);
    }
    __device__ static int kidx(int tid, int j) { return KFAST ? kbase(tid) + j : kbase(tid) + (512 / R) * j; }
    __device__ static int kat(int tid, int r, int j) { return kidx(tid, j); }
    __device__ static int rat(int tid, int r, int j) { return row(tid, r); }
};

// Minimum resident waves per SIMD the register allocator must keep: the one-stage 64x64 tile keeps
// four (the LDS of four workgroups fits a CU); the two-stage one three, since its second register set
// does not fit 128 VGPRs without spilling; larger tiles are limited by LDS anyway.
template <int BM, int BN, int STAGES> struct MinWaves {
    static constexpr int v = (BM * BN <= 64 * 64) ? (STAGES == 1 ? 4 : 3) : (BM * BN <= 128 * 64) ? 3 : 2;
};

// LDS elements of one tile: two K buffers of A and B, or the C staging tile of the C_MFAST epilogue.
// The C_MFAST epilogue stages the tile in EPI_PASSES passes of BN / EPI_PASSES columns (one wave column
// per pass when the whole tile would not fit the main-loop buffers), so it never needs more LDS than
// the main loop and does not lower the number of resident workgroups.
template <int BM, int BN, bool C_MFAST> struct TileSmem {
    static constexpr int MAIN = 2 * kBK * (BM + 1) + 2 * kBK * (BN + 1);
    static constexpr int EPI_PASSES = BN * (BM + 1) > MAIN ? 2 : 1;
    static constexpr int v = C_MFAST ? CMax<MAIN, BN / EPI_PASSES * (BM + 1)>::v : MAIN;
};

// Output tile (tm, tn) of batch `batch`.  C_MFAST selects the epilogue: false writes C straight from
// the accumulators (16 consecutive lanes walk n, coalesced when C is contiguous along n); true stages
// the tile in LDS and writes it with consecutive lanes walking m, for outputs whose contiguous index is
// carried by M.  All threads of the workgroup call it with the same arguments; it leaves LDS in use by
// the epilogue, so a caller running several tiles places a barrier between them.
template <typename T, int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST, int STAGES>
__device__ __forceinline__ void contract_tile(const ContractArgs<T> &p, int64_t batch, int tm, int tn, T *smem) {
    typedef Mfma<T> MF;
    typedef typename MF::acc_t acc_t;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = BM + 1, LDB = BN + 1;
    T (*As)[kBK][LDA] = reinterpret_cast<T (*)[kBK][LDA]>(smem);
    T (*Bs)[kBK][LDB] = reinterpret_cast<T (*)[kBK][LDB]>(smem + 2 * kBK * LDA);
    const int tid = threadIdx.x;

    const dcc_batch bt = p.batches[batch];
    const int64_t m0 = bt.m_begin + (int64_t)tm * BM;
    if (m0 >= bt.m_end) return;
    const int64_t m_end = bt.m_end;
    const int n0 = tn * BN;

    // ---- loader coordinates: every thread holds element pairs ----
    // K-contiguous operand: pair (k, k+1), kp = tid % 8, rows tid / 8 + 32 r (R/32 rows);
    // row-contiguous operand: rows (2 i, 2 i + 1), i = tid % (R/2), at R/32 values of k.
    // A pair is loaded as one vector when the host and the batch checks allow (vec below).
    typedef StageP<BM, A_KFAST> SA;
    typedef StageP<BN, !B_NFAST> SB;
    int64_t a_row[SA::NR];
#pragma unroll
    for (int r = 0; r < SA::NR; ++r) {
        const int64_t m = m0 + SA::row(tid, r);
        a_row[r] = m < m_end ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    }
    int64_t b_col[SB::NR];
#pragma unroll
    for (int r = 0; r < SB::NR; ++r) {
        const int n = n0 + SB::row(tid, r);
        b_col[r] = n < p.N ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    }
    const bool vec = (p.vec & 3) == 3 && ((bt.a_off | bt.b_off) & 1) == 0 &&
                     (A_KFAST || ((bt.m_begin | bt.m_end) & 1) == 0);
    // STAGES = 2: two register sets hold the tiles k+1 and k+2 while tile k is multiplied out of LDS,
    // so every global load has two K steps of MFMA work (plus the other resident waves) to hide
    // behind; STAGES = 1 keeps one set (fewer registers, more resident waves).
    // Loads are branch-free: rows/columns outside the operand read its offset 0 (they only feed
    // outputs that are never stored), and k beyond K reads offset 0 and is replaced by zero.
    typedef T T2 __attribute__((ext_vector_type(2)));
    T a_reg[STAGES][SA::NR][SA::NK], b_reg[STAGES][SB::NR][SB::NK];
    // K offsets of this thread's k values, advanced by one K tile after every load (tiles are loaded
    // in increasing k0), so the main loop needs no index decomposition
    dcc::KIter a_it[SA::NIT], b_it[SB::NIT];
#pragma unroll
    for (int j = 0; j < SA::NIT; ++j) a_it[j] = dcc::kiter(p.ak, (uint32_t)SA::kidx(tid, j));
#pragma unroll
    for (int j = 0; j < SB::NIT; ++j) b_it[j] = dcc::kiter(p.bk, (uint32_t)SB::kidx(tid, j));
    // one operand: X with row offsets rows[], K map km and iterators it[], into reg[NR][NK]
    auto load_op = [&](const T *X, const int64_t *rows, const KMap &km, dcc::KIter *it, int k0, bool vfull,
                       auto &reg, auto kfast_tag, auto geom) {
        typedef decltype(geom) G;
        if constexpr (decltype(kfast_tag)::value) {
            const uint32_t k = (uint32_t)(k0 + G::kidx(tid, 0));
            if (vfull) {
                const int64_t o = it[0].off;
#pragma unroll
                for (int r = 0; r < G::NR; ++r) {
                    const T2 x = *reinterpret_cast<const T2 *>(X + rows[r] + o);
                    reg[r][0] = x[0];
                    reg[r][1] = x[1];
                }
            } else {
                // element by element: k from the iterator, k + 1 decomposed afresh
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const uint32_t kj = k + jj;
                    const bool ok = kj < (uint32_t)p.K;
                    const int64_t o = !ok ? 0 : jj == 0 ? it[0].off : koff(km, kj);
#pragma unroll
                    for (int r = 0; r < G::NR; ++r) {
                        const T x = X[rows[r] + o];
                        reg[r][jj] = ok ? x : T(0);
                    }
                }
            }
            dcc::kstep(km, it[0]);
        } else {
#pragma unroll
            for (int j = 0; j < G::NK; ++j) {
                const uint32_t k = (uint32_t)(k0 + G::kidx(tid, j));
                const bool ok = k < (uint32_t)p.K;
                const int64_t o = ok ? it[j].off : 0;
                dcc::kstep(km, it[j]);
                if (vfull) {
                    const T2 x = *reinterpret_cast<const T2 *>(X + rows[0] + o);
                    reg[0][j] = x[0];
                    reg[1][j] = x[1];
                } else {
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const T x = X[rows[r] + o];
                        reg[r][j] = ok ? x : T(0);
                    }
                }
            }
        }
    };
    auto load_tiles = [&](int s, int k0) {
        const bool vfull = vec && k0 + kBK <= p.K;   // the last, partial K tile loads per element
        load_op(p.A, a_row, p.ak, a_it, k0, vfull, a_reg[s], std::integral_constant<bool, A_KFAST>{}, SA{});
        load_op(p.B, b_col, p.bk, b_it, k0, vfull, b_reg[s], std::integral_constant<bool, !B_NFAST>{}, SB{});
    };
    auto store_tiles = [&](int s, int buf) {
#pragma unroll
        for (int r = 0; r < SA::NR; ++r)
#pragma unroll
            for (int j = 0; j < SA::NK; ++j)
                As[buf][SA::kat(tid, r, j)][SA::rat(tid, r, j)] = a_reg[s][r][j];
#pragma unroll
        for (int r = 0; r < SB::NR; ++r)
#pragma unroll
            for (int j = 0; j < SB::NK; ++j)
                Bs[buf][SB::kat(tid, r, j)][SB::rat(tid, r, j)] = b_reg[s][r][j];
    };

    const int wave = tid / 64, lane = tid % 64;
    const int wm = wave / 2, wn = wave % 2;
    const int lrow = lane % 16, lk = lane / 16;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
    auto compute = [&](int buf) {
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
    };

    // Prefetches past the last tile read offset 0 and land in the idle buffer, which nothing reads
    // afterwards.
    const int nk = (p.K + kBK - 1) / kBK;
    if constexpr (STAGES == 2) {
        // Step kt multiplies LDS buffer kt&1 while register set kt&1 receives tile kt+2 and set
        // (kt+1)&1 (tile kt+1) is written to the other buffer; one barrier per step.  The loop is
        // unrolled by two so register sets are indexed statically, and always runs whole pairs (an
        // odd last step multiplies a zero tile): a second exit made the register allocator copy every
        // accumulator once per iteration.
        load_tiles(0, 0);
        load_tiles(1, kBK);
        store_tiles(0, 0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += 2) {
            load_tiles(0, (kt + 2) * kBK);
            compute(0);
            store_tiles(1, 1);
            __syncthreads();
            load_tiles(1, (kt + 3) * kBK);
            compute(1);
            store_tiles(0, 0);
            __syncthreads();
        }
    } else {
        load_tiles(0, 0);
        store_tiles(0, 0);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            load_tiles(0, (kt + 1) * kBK);
            compute(kt & 1);
            store_tiles(0, (kt & 1) ^ 1);
            __syncthreads();
        }
    }

    if constexpr (C_MFAST) {
        // ---- LDS-staged epilogue: tile stored n-major, then written with lanes along m; in two
        // passes of one wave column each when the whole tile exceeds the main-loop LDS ----
        constexpr int NP = TileSmem<BM, BN, true>::EPI_PASSES;
        constexpr int NC = BN / NP;                 // columns per pass (NP = 2: NC = WN)
        constexpr int NSTEP = kThreads / BM;
        T *Cs = smem;  // the main loop ended with a barrier, so the staging buffers are free
        const int ml = tid % BM;
        const int64_t m = m0 + ml;
        const bool mok = m < m_end;
        const int64_t crow = mok ? bt.c_off + map_off(p.cm, (uint32_t)m, batch) : 0;
#pragma unroll
        for (int h = 0; h < NP; ++h) {
            if (h > 0) __syncthreads();  // the previous pass's columns have been written out
            if (NP == 1 || wn == h) {
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            Cs[(wn * WN + j * 16 + lrow - h * NC) * (BM + 1) + wm * WM + i * 16 + MF::row(lk, r)] =
                                acc[i][j][r];
            }
            __syncthreads();
            if (mok) {
                for (int nl = tid / BM; nl < NC; nl += NSTEP) {
                    const int n = n0 + h * NC + nl;
                    if (n >= p.N) break;
                    T *c = p.C + crow + map_off(p.cn, (uint32_t)n, batch);
                    double v = p.alpha * (double)Cs[nl * (BM + 1) + ml];
                    if (p.beta != 0.0) v += p.beta * (double)*c;
                    *c = (T)v;
                }
            }
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

// Innermost stride of an index map (stride of consecutive indices), 0 if a table.
inline int64_t inner_stride(const dcc_idxmap &m) {
    if (m.table != nullptr) return 0;
    if (m.e2 > 1) return m.s2;
    if (m.e1 > 1) return m.s1;
    return m.s0;
}

// Whether the vectors of q elements a loader thread holds -- (k .. k+q-1) of one row when the operand
// is staged K-contiguous, rows (q i .. q i + q-1) at one k otherwise -- are adjacent in memory and
// aligned to the vector size (vec_bytes) for every row and k (batch offsets and, for A, the row range
// are checked in the kernel).  nrows: the row extent (N for B) must be a multiple of q so no vector
// straddles its end (0 = checked per batch in the kernel).
inline bool vectors_contiguous(const void *base, const dcc_idxmap &rows, const dcc_idxmap &k, bool kfast,
                               int64_t nrows, int64_t q, size_t vec_bytes) {
    if (reinterpret_cast<uintptr_t>(base) % vec_bytes != 0 || rows.table != nullptr || k.table != nullptr)
        return false;
    const dcc_idxmap &vec = kfast ? k : rows;    // the map whose inner index the vector runs along
    const dcc_idxmap &other = kfast ? rows : k;
    if (inner_stride(vec) != 1) return false;
    // vectors start at multiples of q: the inner level must have an extent divisible by q, or be the
    // only level
    if (vec.e2 > 1 ? vec.e2 % q != 0 : false) return false;
    if (vec.e2 > 1 && (vec.s0 % q != 0 || vec.s1 % q != 0)) return false;
    if (vec.e2 <= 1 && vec.e1 > 1 && (vec.e1 % q != 0 || vec.s0 % q != 0)) return false;
    if (other.s0 % q != 0 || other.s1 % q != 0 || other.s2 % q != 0) return false;
    if (!kfast && nrows % q != 0) return false;
    return true;
}

inline bool pairs_contiguous(const void *base, const dcc_idxmap &rows, const dcc_idxmap &k, bool kfast,
                             int64_t nrows, size_t pair_bytes) {
    return vectors_contiguous(base, rows, k, kfast, nrows, 2, pair_bytes);
}

// Tile shape id 0..7: 0..3 = 64x64, 128x64, 64x128, 128x128 with two-stage register prefetch,
// 4..7 = the same shapes with one stage.
inline bool tile_shape_of(int64_t id, int &bm, int &bn, int &stages) {
    if (id < 0 || id > 7) return false;
    static const int shapes[4][2] = {{64, 64}, {128, 64}, {64, 128}, {128, 128}};
    bm = shapes[id % 4][0];
    bn = shapes[id % 4][1];
    stages = id < 4 ? 2 : 1;
    return true;
}

}  // namespace dcc_gemm
