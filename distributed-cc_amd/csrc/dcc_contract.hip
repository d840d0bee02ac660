// Batched gather-GEMM entry points for the triples contractions: one launch per GEMM on FP64/FP32
// MFMA (the tile itself is in dcc_gemm_tile.h) and the split-BF16x3 kernel on BF16 MFMA.
#include "dcc_common.h"
#include "dcc_gemm_tile.h"
#include "dcc_bf16x3.h"

#include <cstdio>
#include <cstdlib>
#include <type_traits>

namespace {

using namespace dcc_gemm;
using namespace dcc_bf16x3;

// One workgroup per output tile over all batches.
template <typename T, int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST, int STAGES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MinWaves<BM, BN, STAGES>::v)))
void contract_kernel(const ContractArgs<T> p) {
    __shared__ T smem[TileSmem<BM, BN, C_MFAST>::v];
    const int64_t L = xcd_remap((int64_t)blockIdx.x, p.total_tiles);
    const int64_t per_batch = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t batch = L / per_batch;
    int tm, tn;
    tile_coords(L - batch * per_batch, p.tiles_m, p.tiles_n, tm, tn);
    contract_tile<T, BM, BN, A_KFAST, B_NFAST, C_MFAST, STAGES>(p, batch, tm, tn, smem);
}

// ---------------------------------------------------------------------------------------------
// Split-BF16x3: FP32 operands split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to
// nearest even, v_cvt_pk_bf16_f32); each product is hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the lo*lo term, ~2^-16 relative, is dropped).
// Same index maps, batching and epilogues as contract_kernel; BK = 32 (one MFMA K step).
//
// MFMA operand layouts (wave64):
//   A fragment: lane holds A[lane%16][8*(lane/16) + 0..7];  B fragment: B[8*(lane/16) + 0..7][lane%16];
//   accumulator: lane holds D[4*(lane/16) + r][lane%16] (as v_mfma_f32_16x16x4f32).
// Each operand is staged in LDS as hi and lo planes in one of two layouts, chosen by which of its
// indices is contiguous in memory:
//   K-contiguous ([row][k], 16-byte chunks XOR-swizzled per row pair): every thread loads vectors of
//     Q consecutive k (Q = 2, or 4 for packed operands whose layout allows 16-byte loads) and writes
//     each as Q 16-bit values per plane; a fragment is one 16-byte ds_read_b128.
//   row-contiguous ([k][row], stride R + 16): every thread loads vectors of Q adjacent rows at one k
//     and writes each as Q 16-bit values per plane; a fragment is two ds_read_b64_tr_b16,
//     which on gfx950 gives lane 16g + i the four values at column i of the 4 x 16 block whose row j
//     the lanes 16g + 4j .. 16g + 4j + 3 address (four 16-bit values each).
// The K offsets of the next tile's 32 k values are computed once per step by wave 0 (reciprocal
// multiplication) into a small LDS table, so the loaders need no division; k >= K is marked -1 and
// masked only in the last tile.
// Staging geometry of one operand with R rows (m for A, n for B) per tile and NT threads, loaded in
// vectors of Q elements (Q = 2: pairs, Q = 4: quads of packed words, one 16-byte load each): along k
// for a K-contiguous operand (Q consecutive k of one row), along rows otherwise (Q adjacent rows at
// one k).
template <int R, bool KFAST, int Q, int NT> struct Stage3 {
    // K-contiguous planes are [row][32 k] with the four 16-byte chunks of a row XOR-swizzled by
    // (row >> 1) & 3: the 16-byte fragment reads of 8 consecutive rows and the vector writes of a
    // half-wave then touch distinct banks, without padding.
    static constexpr int PLANE = KFAST ? R * kBK3 : kBK3 * (R + 16);
    static constexpr int VPR = kBK3 / Q;                        // vectors per row (K-contiguous)
    static constexpr int RQ = R / Q;                            // row groups (row-contiguous)
    static constexpr int NR = KFAST ? R * VPR / NT : Q;         // rows per thread
    static constexpr int NK = KFAST ? Q : R * kBK3 / (NT * Q);  // k values per thread
    static_assert(NR * NK * NT == R * kBK3, "the threads cover the tile exactly once");
    __device__ static int row(int tid, int r) {
        return KFAST ? tid / VPR + (NT / VPR) * r : Q * (tid % RQ) + r;
    }
    __device__ static int kidx(int tid, int j) {
        return KFAST ? Q * (tid % VPR) + j : tid / RQ + (NT / RQ) * j;
    }
    // 16-bit element index of (row, k) in a plane
    __device__ static int at(int row, int k) {
        if constexpr (KFAST) return row * kBK3 + 8 * ((k >> 3) ^ ((row >> 1) & 3)) + (k & 7);
        else return k * (R + 16) + row;
    }
    // fragment of rows rbase..rbase+15 (16 rows x 8 k values per lane group) from one plane
    __device__ static bf16x8 frag(const unsigned short *plane, int rbase, int lane) {
        const int lrow = lane % 16, g = lane / 16;
        if constexpr (KFAST) {
            const int row = rbase + lrow;
            return *reinterpret_cast<const bf16x8 *>(&plane[row * kBK3 + 8 * (g ^ ((row >> 1) & 3))]);
        } else {
            const int kr = 8 * g + lrow / 4, c = rbase + 4 * (lrow % 4);
            typedef __attribute__((address_space(3))) unsigned short lds_u16;
            const lds_u16 *base = (const lds_u16 *)plane;
            const s16x4 lo4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(base + kr * (R + 16) + c));
            const s16x4 hi4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(base + (kr + 4) * (R + 16) + c));
            const bf16x4 a = __builtin_bit_cast(bf16x4, lo4), b = __builtin_bit_cast(bf16x4, hi4);
            return bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
        }
    }
    // fragment of v_mfma_f32_32x32x16_bf16 for k values 16 ksub .. 16 ksub + 15 of the plane: lane l
    // holds rows rbase + l % 32 at k = 16 ksub + 8 (l / 32) + 0..7
    __device__ static bf16x8 frag32(const unsigned short *plane, int rbase, int lane, int ksub) {
        if constexpr (KFAST) {
            const int row = rbase + lane % 32, chunk = 2 * ksub + lane / 32;
            return *reinterpret_cast<const bf16x8 *>(&plane[row * kBK3 + 8 * (chunk ^ ((row >> 1) & 3))]);
        } else {
            const int g = lane / 16, lrow = lane % 16;
            const int kr = 16 * ksub + 8 * (g / 2) + lrow / 4, c = rbase + 16 * (g % 2) + 4 * (lrow % 4);
            typedef __attribute__((address_space(3))) unsigned short lds_u16;
            const lds_u16 *base = (const lds_u16 *)plane;
            const s16x4 lo4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(base + kr * (R + 16) + c));
            const s16x4 hi4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(base + (kr + 4) * (R + 16) + c));
            const bf16x4 a = __builtin_bit_cast(bf16x4, lo4), b = __builtin_bit_cast(bf16x4, hi4);
            return bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
        }
    }
};

// MODE 0: FP32 operands loaded in pairs; 1: packed hi/lo words instead of FP32 values (the pointers
// of p are reinterpreted) loaded in pairs; 2: packed words loaded in quads (16-byte loads).
// NW waves per workgroup: 4 as a 2 x 2 grid, or 8 (256 x 128 and 128 x 256 tiles, 64 x 64 per wave,
// one workgroup per CU), which loads a quarter fewer operand bytes per flop than two 128 x 128 tiles;
// 256 x 256 on 8 waves as 4 x 2 (64 x 128 per wave) loads half the operand bytes per flop of
// 128 x 128 and reads a quarter fewer fragment bytes from LDS per MFMA.
// MI: the MFMA shape, 16 (v_mfma_f32_16x16x32_bf16, one K step of 32 per instruction) or 32
// (v_mfma_f32_32x32x16_bf16, two K steps of 16: half the instructions for the same work).
template <int BM, int BN, bool A_KFAST, bool B_NFAST, bool C_MFAST, int MODE, int NW, int MI = 16,
          bool CVEC = false>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void contract_bf16x3_kernel(const ContractArgs<float> p) {
    constexpr bool PACKED = MODE >= 1;
    constexpr int Q = MODE == 2 ? 4 : 2;
    constexpr int NT = NW * 64;
    typedef Stage3<BM, A_KFAST, Q, NT> SA;
    typedef Stage3<BN, !B_NFAST, Q, NT> SB;
    constexpr int WGM = NW == 4 ? 2 : (BM >= BN ? 4 : 2), WGN = NW / WGM;   // wave grid
    constexpr int WM = BM / WGM, WN = BN / WGN;
    constexpr int FM = WM / MI, FN = WN / MI;
    typedef typename AccOf<MI>::t acc_t;
    constexpr int NACC = AccOf<MI>::n;
    constexpr int BUF = 2 * (SA::PLANE + SB::PLANE);            // one buffer: hi + lo planes, 16-bit units
    constexpr int MAIN_BYTES = 2 * BUF * 2;                      // two buffers
    // the M-contiguous epilogue stages at most 128 columns of C at a time (two passes for BN = 256)
    constexpr int EBN = BN > 128 ? 128 : BN;
    constexpr int EPI_BYTES = C_MFAST ? EBN * (BM + 1) * 4 : 0;
    __shared__ __align__(16) unsigned char smem[MAIN_BYTES > EPI_BYTES ? MAIN_BYTES : EPI_BYTES];
    __shared__ int64_t ktab[2][2][kBK3];  // [tile parity][A, B][k]
    unsigned short *const lds = reinterpret_cast<unsigned short *>(smem);

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
    const int wave = tid / 64, lane = tid % 64;

    // row offsets; rows outside the operand read offset 0 (they only feed outputs never stored)
    int64_t a_row[SA::NR], b_col[SB::NR];
#pragma unroll
    for (int r = 0; r < SA::NR; ++r) {
        const int64_t m = m0 + SA::row(tid, r);
        a_row[r] = m < m_end ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    }
#pragma unroll
    for (int r = 0; r < SB::NR; ++r) {
        const int n = n0 + SB::row(tid, r);
        b_col[r] = n < p.N ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    }

    // K offsets of tile k0 into table slot t (-1 beyond K): lanes 0..15 of waves 0, 1 fill the A
    // entries, of waves 2, 3 the B entries
    auto fill_table = [&](int t, int k0) {
        if (wave < 4 && lane < 16) {
            const int e = (wave & 1) * 16 + lane;
            const uint32_t k = (uint32_t)(k0 + e);
            const int64_t o = koff(wave < 2 ? p.ak : p.bk, k);
            ktab[t][wave >> 1][e] = k < (uint32_t)p.K ? o : int64_t(-1);
        }
    };

    typedef typename std::conditional<PACKED, uint32_t, float>::type E;
    const E *Ap = reinterpret_cast<const E *>(p.A);
    const E *Bp = reinterpret_cast<const E *>(p.B);
    E a_reg[SA::NR][SA::NK], b_reg[SB::NR][SB::NK];
    typedef E EQ __attribute__((ext_vector_type(Q)));
    auto load_tiles = [&](int t, auto tail, auto vec) {
        constexpr bool TAIL = decltype(tail)::value;
        if constexpr (decltype(vec)::value && !TAIL) {
            // each thread's vectors are adjacent and aligned: one 4Q-byte load per vector
            if constexpr (A_KFAST) {
                const int64_t o = ktab[t][0][SA::kidx(tid, 0)];
#pragma unroll
                for (int r = 0; r < SA::NR; ++r) {
                    const EQ x = *reinterpret_cast<const EQ *>(Ap + a_row[r] + o);
#pragma unroll
                    for (int q = 0; q < Q; ++q) a_reg[r][q] = x[q];
                }
            } else {
#pragma unroll
                for (int j = 0; j < SA::NK; ++j) {
                    const EQ x = *reinterpret_cast<const EQ *>(Ap + a_row[0] + ktab[t][0][SA::kidx(tid, j)]);
#pragma unroll
                    for (int q = 0; q < Q; ++q) a_reg[q][j] = x[q];
                }
            }
            if constexpr (!B_NFAST) {
                const int64_t o = ktab[t][1][SB::kidx(tid, 0)];
#pragma unroll
                for (int r = 0; r < SB::NR; ++r) {
                    const EQ x = *reinterpret_cast<const EQ *>(Bp + b_col[r] + o);
#pragma unroll
                    for (int q = 0; q < Q; ++q) b_reg[r][q] = x[q];
                }
            } else {
#pragma unroll
                for (int j = 0; j < SB::NK; ++j) {
                    const EQ x = *reinterpret_cast<const EQ *>(Bp + b_col[0] + ktab[t][1][SB::kidx(tid, j)]);
#pragma unroll
                    for (int q = 0; q < Q; ++q) b_reg[q][j] = x[q];
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < SA::NK; ++j) {
            const int64_t o = ktab[t][0][SA::kidx(tid, j)];
#pragma unroll
            for (int r = 0; r < SA::NR; ++r) {
                if constexpr (TAIL) {
                    const E x = Ap[a_row[r] + (o < 0 ? 0 : o)];
                    a_reg[r][j] = o < 0 ? E(0) : x;
                } else {
                    a_reg[r][j] = Ap[a_row[r] + o];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < SB::NK; ++j) {
            const int64_t o = ktab[t][1][SB::kidx(tid, j)];
#pragma unroll
            for (int r = 0; r < SB::NR; ++r) {
                if constexpr (TAIL) {
                    const E x = Bp[b_col[r] + (o < 0 ? 0 : o)];
                    b_reg[r][j] = o < 0 ? E(0) : x;
                } else {
                    b_reg[r][j] = Bp[b_col[r] + o];
                }
            }
        }
    };
    // vectors along k (K-contiguous layout) or along rows (row-contiguous layout): Q/2 split pairs,
    // Q consecutive 16-bit values per plane, written as one 2Q-byte store
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    auto put = [&](unsigned short *P, int e, const uint32_t *w) {
        if constexpr (Q == 2) *reinterpret_cast<uint32_t *>(&P[e]) = w[0];
        else *reinterpret_cast<u32x2 *>(&P[e]) = u32x2{w[0], w[1]};
    };
    auto store_tiles = [&](int buf) {
        unsigned short *As = lds + buf * BUF;
        unsigned short *Bs = As + 2 * SA::PLANE;
        uint32_t hi[Q / 2], lo[Q / 2];
        if constexpr (A_KFAST) {
#pragma unroll
            for (int r = 0; r < SA::NR; ++r) {
#pragma unroll
                for (int h = 0; h < Q / 2; ++h) split_pair(a_reg[r][2 * h], a_reg[r][2 * h + 1], hi[h], lo[h]);
                const int e = SA::at(SA::row(tid, r), SA::kidx(tid, 0));
                put(As, e, hi);
                put(As + SA::PLANE, e, lo);
            }
        } else {
#pragma unroll
            for (int j = 0; j < SA::NK; ++j) {
#pragma unroll
                for (int h = 0; h < Q / 2; ++h) split_pair(a_reg[2 * h][j], a_reg[2 * h + 1][j], hi[h], lo[h]);
                const int e = SA::at(SA::row(tid, 0), SA::kidx(tid, j));
                put(As, e, hi);
                put(As + SA::PLANE, e, lo);
            }
        }
        if constexpr (!B_NFAST) {
#pragma unroll
            for (int r = 0; r < SB::NR; ++r) {
#pragma unroll
                for (int h = 0; h < Q / 2; ++h) split_pair(b_reg[r][2 * h], b_reg[r][2 * h + 1], hi[h], lo[h]);
                const int e = SB::at(SB::row(tid, r), SB::kidx(tid, 0));
                put(Bs, e, hi);
                put(Bs + SB::PLANE, e, lo);
            }
        } else {
#pragma unroll
            for (int j = 0; j < SB::NK; ++j) {
#pragma unroll
                for (int h = 0; h < Q / 2; ++h) split_pair(b_reg[2 * h][j], b_reg[2 * h + 1][j], hi[h], lo[h]);
                const int e = SB::at(SB::row(tid, 0), SB::kidx(tid, j));
                put(Bs, e, hi);
                put(Bs + SB::PLANE, e, lo);
            }
        }
    };

    const int wm = wave / WGN, wn = wave % WGN;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
    // row (of the workgroup tile) of accumulator element r of MFMA tile i, column of MFMA tile j:
    // 16 x 16: D[4 (lane / 16) + r][lane % 16]; 32 x 32: D[8 (r / 4) + 4 (lane / 32) + r % 4][lane % 32]
    auto acc_row = [&](int i, int r) {
        return MI == 16 ? wm * WM + i * 16 + 4 * (lane / 16) + r
                        : wm * WM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    };
    auto acc_col = [&](int j) { return wn * WN + j * MI + lane % MI; };

    // Two LDS buffers and one register stage: step kt loads tile kt+1 into registers, multiplies
    // tile kt from buffer kt&1 and writes tile kt+1 into the other buffer (last read in step kt-1);
    // one barrier per step.  Table slot t&1 holds the K offsets of tile t; slot kt&1 is refilled
    // for tile kt+2 during step kt (its last reader, the load of tile kt, ran in step kt-1).
    const int nk = (p.K + kBK3 - 1) / kBK3;
    const int nfull = p.K / kBK3;  // tiles with every k < K
    // vector loads: the host checked strides, extents and pointer alignment; the batch offsets
    // (and, for row vectors of A, the row range) are checked here, uniformly per workgroup
    const bool vec = (p.vec & 3) == 3 && ((bt.a_off | bt.b_off) & (Q - 1)) == 0 &&
                     (A_KFAST || ((bt.m_begin | bt.m_end) & (Q - 1)) == 0);
    auto mainloop = [&](auto vtag) {
    fill_table(0, 0);
    fill_table(1, kBK3);
    __syncthreads();
    if (nfull > 0) load_tiles(0, std::false_type{}, vtag);
    else load_tiles(0, std::true_type{}, vtag);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int t1 = (kt + 1) & 1;
        if (kt + 1 < nfull) load_tiles(t1, std::false_type{}, vtag);
        else load_tiles(t1, std::true_type{}, vtag);
        fill_table(kt & 1, (kt + 2) * kBK3);
        const unsigned short *As = lds + (kt & 1) * BUF;
        const unsigned short *Bs = As + 2 * SA::PLANE;
        if constexpr (MI == 16) {
            bf16x8 ah[FM], al[FM];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                ah[i] = SA::frag(As, wm * WM + i * 16, lane);
                al[i] = SA::frag(As + SA::PLANE, wm * WM + i * 16, lane);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const bf16x8 bh = SB::frag(Bs, wn * WN + j * 16, lane);
                const bf16x8 bl = SB::frag(Bs + SB::PLANE, wn * WN + j * 16, lane);
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                bf16x8 ah[FM], al[FM];
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    ah[i] = SA::frag32(As, wm * WM + i * 32, lane, ks);
                    al[i] = SA::frag32(As + SA::PLANE, wm * WM + i * 32, lane, ks);
                }
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const bf16x8 bh = SB::frag32(Bs, wn * WN + j * 32, lane, ks);
                    const bf16x8 bl = SB::frag32(Bs + SB::PLANE, wn * WN + j * 32, lane, ks);
#pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
                    }
                }
            }
        }
        store_tiles(t1);
        __syncthreads();
    }
    };
    if (vec) mainloop(std::true_type{});
    else mainloop(std::false_type{});

    if constexpr (C_MFAST) {
        // C staged through LDS column block by column block (EBN columns per pass), then written along
        // its contiguous M index; every thread reaches every barrier
        float *Cs = reinterpret_cast<float *>(smem);
        constexpr int NSTEP = NT / BM;
        const int ml = tid % BM;
        const int64_t m = m0 + ml;
        const bool mok = m < m_end;
        const int64_t crow = mok ? bt.c_off + map_off(p.cm, (uint32_t)m, batch) : 0;
#pragma unroll
        for (int pass = 0; pass < BN / EBN; ++pass) {
            if (pass > 0) __syncthreads();   // the previous pass has read its columns
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int col = acc_col(j) - pass * EBN;
                    if (col < 0 || col >= EBN) continue;
#pragma unroll
                    for (int r = 0; r < NACC; ++r) Cs[col * (BM + 1) + acc_row(i, r)] = acc[i][j][r];
                }
            __syncthreads();
            if (mok) {
                for (int nl = tid / BM; nl < EBN; nl += NSTEP) {
                    const int n = n0 + pass * EBN + nl;
                    if (n >= p.N) break;
                    float *c = p.C + crow + map_off(p.cn, (uint32_t)n, batch);
                    double v = p.alpha * (double)Cs[nl * (BM + 1) + ml];
                    if (p.beta != 0.0) v += p.beta * (double)*c;
                    *c = (float)v;
                }
            }
        }
        return;
    }
    if constexpr (MI == 16 && CVEC) {
        // C contiguous along n with rows of 16-byte vectors (checked on the host; the batch offset
        // here): a band of rows -- fragment i of every wave row -- is staged in LDS, then each wave
        // writes whole rows with 16-byte stores (written directly from the accumulator layout, a store
        // covers 64-byte pieces of four rows).  A separate instance: in the long-K kernels the extra
        // code cost a spill reload per k step (the register-staged W_vvvv launch ran 8 % longer).
        if ((bt.c_off & 3) == 0) {
            constexpr int ER = 16 * WGM, LDC = BN + 16, NV = BN / 4;
            static_assert(ER * LDC * 4 <= (int)sizeof(smem), "epilogue band fits the stage buffers");
            float *Cs = reinterpret_cast<float *>(smem);
            const int brow = wm * 16 + 4 * (lane / 16);   // band row of this lane's r = 0
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                __syncthreads();   // the previous band (first pass: the main loop's buffers) is read
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < NACC; ++r) Cs[(brow + r) * LDC + acc_col(j)] = acc[i][j][r];
                __syncthreads();
                for (int g = tid; g < ER * NV; g += NT) {
                    const int lr = g / NV, c4 = g % NV;
                    const int64_t m = m0 + (lr / 16) * WM + i * 16 + lr % 16;
                    const int n = n0 + 4 * c4;
                    if (m >= m_end || n >= p.N) continue;
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    const f4 x = *reinterpret_cast<const f4 *>(Cs + lr * LDC + 4 * c4);
                    const int64_t crow = bt.c_off + map_off(p.cm, (uint32_t)m, batch);
                    float *c = p.C + crow + map_off(p.cn, (uint32_t)n, batch);
                    if (n + 3 < p.N) {
                        f4 old = {0.f, 0.f, 0.f, 0.f};
                        if (p.beta != 0.0) old = *reinterpret_cast<const f4 *>(c);
                        f4 out;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            double v = p.alpha * (double)x[e];
                            if (p.beta != 0.0) v += p.beta * (double)old[e];
                            out[e] = (float)v;
                        }
                        *reinterpret_cast<f4 *>(c) = out;
                    } else {
                        for (int e = 0; n + e < p.N; ++e) {
                            float *ce = p.C + crow + map_off(p.cn, (uint32_t)(n + e), batch);
                            double v = p.alpha * (double)x[e];
                            if (p.beta != 0.0) v += p.beta * (double)*ce;
                            *ce = (float)v;
                        }
                    }
                }
            }
            return;
        }
    }
    int64_t ccol[FN];
    uint32_t cok = 0;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int n = n0 + acc_col(j);
        const bool ok = n < p.N;
        cok |= (ok ? 1u : 0u) << j;
        ccol[j] = ok ? map_off(p.cn, (uint32_t)n, batch) : 0;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        // four accumulator rows at a time (for 32 x 32 tiles: the four groups of four)
#pragma unroll
        for (int r0 = 0; r0 < NACC; r0 += 4) {
            int64_t crow[4];
            uint32_t rok = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t m = m0 + acc_row(i, r0 + r);
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
                    double v = p.alpha * (double)acc[i][j][r0 + r];
                    if (p.beta != 0.0) v += p.beta * old[r][j];
                    p.C[crow[r] + ccol[j]] = (float)v;
                }
        }
    }
}

template <typename T, int BM, int BN, bool C_MFAST, int STAGES>
void launch_variant(ContractArgs<T> &args, bool a_kfast, bool b_nfast, dim3 grid, dim3 block, hipStream_t stream) {
    if (a_kfast && b_nfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, true, true, C_MFAST, STAGES>), grid, block, 0, stream, args);
    else if (a_kfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, true, false, C_MFAST, STAGES>), grid, block, 0, stream, args);
    else if (b_nfast)
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, false, true, C_MFAST, STAGES>), grid, block, 0, stream, args);
    else
        hipLaunchKernelGGL((contract_kernel<T, BM, BN, false, false, C_MFAST, STAGES>), grid, block, 0, stream, args);
}

template <typename T, int BM, int BN>
int launch(ContractArgs<T> &args, int64_t max_rows, bool a_kfast, bool b_nfast, bool c_mfast, int stages,
           hipStream_t stream) {
    args.tiles_m = (int)((max_rows + BM - 1) / BM);
    args.tiles_n = (int)((args.N + BN - 1) / BN);
    const int64_t nbatch = args.total_tiles;  // caller passes nbatch here
    args.total_tiles = nbatch * args.tiles_m * args.tiles_n;
    if (args.total_tiles >= ((int64_t)1 << 31)) {
        dcc::set_error("dcc_contract_f64: too many tiles");
        return 2;
    }
    const dim3 grid((unsigned)args.total_tiles), block(kThreads);
    if (c_mfast && stages == 2)
        launch_variant<T, BM, BN, true, 2>(args, a_kfast, b_nfast, grid, block, stream);
    else if (c_mfast)
        launch_variant<T, BM, BN, true, 1>(args, a_kfast, b_nfast, grid, block, stream);
    else if (stages == 2)
        launch_variant<T, BM, BN, false, 2>(args, a_kfast, b_nfast, grid, block, stream);
    else
        launch_variant<T, BM, BN, false, 1>(args, a_kfast, b_nfast, grid, block, stream);
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
    args.vec = (pairs_contiguous(A, am, ak, a_kfast, 0, 2 * sizeof(T)) ? 1 : 0) |
               (pairs_contiguous(B, bn, bk, !b_nfast, N, 2 * sizeof(T)) ? 2 : 0);
    // C written along m (through LDS) unless it is contiguous along n
    const bool c_mfast = inner_stride(cn) != 1;
    hipStream_t s = dcc::as_stream(stream);
    // Workgroup tile: tile_shape 0..3 = 64x64, 128x64, 64x128, 128x128 with the two-stage register
    // prefetch, 4..7 = the same shapes with one stage; -1 = 64x64 two-stage, or DCC_CONTRACT_TILE
    // (e.g. "128x64" or "64x64s1") when set.
    static const char *tile_env = getenv("DCC_CONTRACT_TILE");
    int tbm = 64, tbn = 64, stages = 2;
    if (tile_shape >= 0) {
        DCC_CHECK(tile_shape_of(tile_shape, tbm, tbn, stages), "tile shape must be 0..7");
    } else if (tile_env != nullptr) {
        int st = 2;
        const int got = sscanf(tile_env, "%dx%ds%d", &tbm, &tbn, &st);
        if (got < 2) {
            tbm = 64;
            tbn = 64;
        }
        stages = st == 1 ? 1 : 2;
    }
    if (tbm == 128 && tbn == 128) return launch<T, 128, 128>(args, max_rows, a_kfast, b_nfast, c_mfast, stages, s);
    if (tbm == 128 && tbn == 64) return launch<T, 128, 64>(args, max_rows, a_kfast, b_nfast, c_mfast, stages, s);
    if (tbm == 64 && tbn == 128) return launch<T, 64, 128>(args, max_rows, a_kfast, b_nfast, c_mfast, stages, s);
    return launch<T, 64, 64>(args, max_rows, a_kfast, b_nfast, c_mfast, stages, s);
}

template <int BM, int BN, bool C_MFAST, int MODE, int NW, int MI, bool CVEC = false>
void launch_bf16x3(ContractArgs<float> &args, bool a_kfast, bool b_nfast, dim3 grid, dim3 block,
                   hipStream_t stream) {
    if (a_kfast && b_nfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<BM, BN, true, true, C_MFAST, MODE, NW, MI, CVEC>), grid, block,
                           0, stream, args);
    else if (a_kfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<BM, BN, true, false, C_MFAST, MODE, NW, MI, CVEC>), grid, block,
                           0, stream, args);
    else if (b_nfast)
        hipLaunchKernelGGL((contract_bf16x3_kernel<BM, BN, false, true, C_MFAST, MODE, NW, MI, CVEC>), grid, block,
                           0, stream, args);
    else
        hipLaunchKernelGGL((contract_bf16x3_kernel<BM, BN, false, false, C_MFAST, MODE, NW, MI, CVEC>), grid,
                           block, 0, stream, args);
}

// K up to which the register-staged tiles of a C contiguous along n write C through the vector epilogue
// (short-K GEMMs, whose time the epilogue's stores dominate)
constexpr int64_t kVecEpilogueMaxK = 4096;

template <int BM, int BN, int NW = 4, int MI = 16>
int run_bf16x3(ContractArgs<float> &args, int64_t nbatch, int64_t max_rows, bool a_kfast, bool b_nfast,
               bool c_mfast, int mode, hipStream_t s) {
    args.tiles_m = (int)((max_rows + BM - 1) / BM);
    args.tiles_n = (int)((args.N + BN - 1) / BN);
    args.total_tiles = nbatch * args.tiles_m * args.tiles_n;
    DCC_CHECK(args.total_tiles < ((int64_t)1 << 31), "too many tiles");
    const dim3 grid((unsigned)args.total_tiles), block(NW * 64);
    if (c_mfast) {
        if (mode == 2) launch_bf16x3<BM, BN, true, 2, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
        else if (mode == 1) launch_bf16x3<BM, BN, true, 1, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
        else launch_bf16x3<BM, BN, true, 0, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
    } else {
        bool cvec = false;
        if constexpr (MI == 16) {
            // packed operands with a short K: the vector epilogue instance
            cvec = mode >= 1 && (args.vec & 4) && args.K <= kVecEpilogueMaxK;
            if (cvec) {
                if (mode == 2) launch_bf16x3<BM, BN, false, 2, NW, MI, true>(args, a_kfast, b_nfast, grid, block, s);
                else launch_bf16x3<BM, BN, false, 1, NW, MI, true>(args, a_kfast, b_nfast, grid, block, s);
            }
        }
        if (!cvec) {
            if (mode == 2) launch_bf16x3<BM, BN, false, 2, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
            else if (mode == 1) launch_bf16x3<BM, BN, false, 1, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
            else launch_bf16x3<BM, BN, false, 0, NW, MI>(args, a_kfast, b_nfast, grid, block, s);
        }
    }
    DCC_LAUNCH_CHECK();
    return 0;
}

int contract_bf16x3_entry(int64_t M, int64_t N, int64_t K, double alpha, double beta, const float *A,
                          dcc_idxmap am, dcc_idxmap ak, const float *B, dcc_idxmap bk, dcc_idxmap bn, float *C,
                          dcc_idxmap cm, dcc_idxmap cn, const dcc_batch *batches, int64_t nbatch,
                          int64_t max_rows, int64_t tile_shape, dcc_stream_t stream, bool packed = false) {
    DCC_CHECK(M >= 0 && N >= 0 && K >= 0 && nbatch >= 0 && max_rows >= 0, "negative extent");
    DCC_CHECK(M < ((int64_t)1 << 31) && N < ((int64_t)1 << 31) && K < ((int64_t)1 << 31),
              "extent exceeds 2^31");
    DCC_CHECK(valid_map(am) && valid_map(ak) && valid_map(bk) && valid_map(bn) && valid_map(cm) &&
                  valid_map(cn),
              "invalid index map");
    DCC_CHECK(ak.table == nullptr && bk.table == nullptr, "K index maps must be mixed-radix maps");
    DCC_CHECK(tile_shape >= -1 && tile_shape <= 13, "tile shape must be -1..13");
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
    args.ak = make_kmap(ak);
    args.bk = make_kmap(bk);
    args.batches = batches;
    const bool a_kfast = inner_stride(ak) == 1 || inner_stride(am) != 1;
    const bool b_nfast = inner_stride(bn) == 1 || inner_stride(bk) != 1;
    const bool c_mfast = inner_stride(cn) != 1;
    // packed operands whose layout allows it are loaded in quads (16 bytes), the rest in pairs
    const bool quads = packed && vectors_contiguous(A, am, ak, a_kfast, 0, 4, 16) &&
                       vectors_contiguous(B, bn, bk, !b_nfast, N, 4, 16);
    const int mode = quads ? 2 : packed ? 1 : 0;
    args.vec = quads ? 3
                     : (pairs_contiguous(A, am, ak, a_kfast, 0, 8) ? 1 : 0) |
                           (pairs_contiguous(B, bn, bk, !b_nfast, N, 8) ? 2 : 0);
    // bit 2: C contiguous along n in aligned runs of 16-byte vectors (rows staged through LDS and
    // written as vectors by the register-staged tiles)
    if (!c_mfast && vectors_contiguous(C, cm, cn, true, 0, 4, 16)) args.vec |= 4;
    hipStream_t s = dcc::as_stream(stream);
    // default: 128x128 unless one side of the problem is small
    int shape = (int)tile_shape;
    if (shape < 0) shape = (max_rows <= 64 ? 0 : 1) + (N <= 64 ? 0 : 2);
    // shapes 8..13: direct-to-LDS staging, for packed quads with both operands K-contiguous and K a
    // multiple of 32 x stages; otherwise the register-staged kernel of the same tile
    const int stages = (shape == 9 || shape == 10) ? 3 : 2;
    const bool dma = mode == 2 && a_kfast && !b_nfast && K >= stages * kBK3 && K % (stages * kBK3) == 0;
    if (shape >= 8 && dma) {
        switch (shape) {
            case 8: return run_bf16x3_dma<256, 256, 16, 2>(args, nbatch, max_rows, c_mfast, s);
            case 9: return run_bf16x3_dma<256, 128, 16, 3>(args, nbatch, max_rows, c_mfast, s);
            case 10: return run_bf16x3_dma<128, 256, 16, 3>(args, nbatch, max_rows, c_mfast, s);
            case 11: return run_bf16x3_dma<256, 256, 32, 2>(args, nbatch, max_rows, c_mfast, s);
            case 12: return run_bf16x3_dma_wide(args, nbatch, max_rows, c_mfast, 16, s);
            default: return run_bf16x3_dma_wide(args, nbatch, max_rows, c_mfast, 32, s);
        }
    }
    if (shape >= 8) shape = shape == 9 ? 4 : shape == 10 ? 5 : (shape == 11 || shape == 13) ? 7 : 6;
    switch (shape) {
        case 0: return run_bf16x3<64, 64>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 1: return run_bf16x3<128, 64>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 2: return run_bf16x3<64, 128>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 4: return run_bf16x3<256, 128, 8>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 5: return run_bf16x3<128, 256, 8>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 6: return run_bf16x3<256, 256, 8>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        case 7: return run_bf16x3<256, 256, 8, 32>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
        default: return run_bf16x3<128, 128>(args, nbatch, max_rows, a_kfast, b_nfast, c_mfast, mode, s);
    }
}

}  // namespace

extern "C" int dcc_contract_f32_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                                       const float *A, dcc_idxmap am, dcc_idxmap ak, const float *B,
                                       dcc_idxmap bk, dcc_idxmap bn, float *C, dcc_idxmap cm, dcc_idxmap cn,
                                       const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                                       dcc_stream_t stream) {
    return contract_bf16x3_entry(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch, max_rows,
                                 -1, stream);
}

extern "C" int dcc_contract_packed_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                                          const uint32_t *A, dcc_idxmap am, dcc_idxmap ak, const uint32_t *B,
                                          dcc_idxmap bk, dcc_idxmap bn, float *C, dcc_idxmap cm, dcc_idxmap cn,
                                          const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                                          int64_t tile_shape, dcc_stream_t stream) {
    return contract_bf16x3_entry(M, N, K, alpha, beta, reinterpret_cast<const float *>(A), am, ak,
                                 reinterpret_cast<const float *>(B), bk, bn, C, cm, cn, batches, nbatch, max_rows,
                                 tile_shape, stream, true);
}

extern "C" int dcc_contract_tiled_f32_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, double beta,
                                             const float *A, dcc_idxmap am, dcc_idxmap ak, const float *B,
                                             dcc_idxmap bk, dcc_idxmap bn, float *C, dcc_idxmap cm, dcc_idxmap cn,
                                             const dcc_batch *batches, int64_t nbatch, int64_t max_rows,
                                             int64_t tile_shape, dcc_stream_t stream) {
    return contract_bf16x3_entry(M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch, max_rows,
                                 tile_shape, stream);
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
