// Split-BF16x3 building blocks shared by the register-staged kernel (dcc_contract.hip) and the kernels
// whose operands are staged by direct global -> LDS requests (dcc_contract.hip for eight waves,
// dcc_contract_wide.hip for four waves with accumulators in AGPRs).
//
// FP32 operands are split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to nearest even);
// each product is hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 (or 32x32x16) with FP32
// accumulation (the lo*lo term, ~2^-16 relative, is dropped).
#pragma once

#include "dcc_common.h"
#include "dcc_gemm_tile.h"

namespace dcc_bf16x3 {

using namespace dcc_gemm;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;
constexpr int kBK3 = 32;

// hi/lo bf16 pairs of two floats, element 0 in the low half of each word.
__device__ __forceinline__ void split_pair(float x0, float x1, uint32_t &hi, uint32_t &lo) {
    const bf16x2 h = {(__bf16)x0, (__bf16)x1};
    hi = __builtin_bit_cast(uint32_t, h);
    const float h0 = __uint_as_float(hi << 16), h1 = __uint_as_float(hi & 0xffff0000u);
    const bf16x2 l = {(__bf16)(x0 - h0), (__bf16)(x1 - h1)};
    lo = __builtin_bit_cast(uint32_t, l);
}

// Packed operands hold one element per 32-bit word: hi = bf16(x) in the low half, lo = bf16(x - hi) in
// the high half (written by dcc_pack_bf16x3 with the same conversion as split_pair).  Two words give
// the hi pair and the lo pair directly.
__device__ __forceinline__ void split_pair(uint32_t w0, uint32_t w1, uint32_t &hi, uint32_t &lo) {
    hi = __builtin_amdgcn_perm(w1, w0, 0x05040100u);
    lo = __builtin_amdgcn_perm(w1, w0, 0x07060302u);
}

typedef float flt16 __attribute__((ext_vector_type(16)));
// Accumulator of one MI x MI MFMA tile: 4 floats per lane (16 x 16) or 16 (32 x 32).
template <int MI> struct AccOf { typedef flt4 t; static constexpr int n = 4; };
template <> struct AccOf<32> { typedef flt16 t; static constexpr int n = 16; };

// Split-BF16x3 on packed operands that are both K-contiguous, staged by direct global -> LDS loads
// (global_load_lds_dwordx4: each lane's 16-byte vector lands at the wave's LDS base + 16 x lane,
// without passing through registers).  S stages of LDS are in flight: tile kt + S - 1 is requested
// while tile kt is multiplied, so a load has S - 1 steps to arrive (one step in the register-staged
// kernel), and no VGPRs hold staged operands.  Each stage is its own __shared__ array and the loop is
// unrolled by S, so the compiler sees which stage every fragment read uses and does not wait for the
// loads of the other stages; the waits are explicit (s_waitcnt vmcnt) before a raw s_barrier.
// Stage layout: [row][32 packed words], the eight 16-byte chunks of a row XOR-swizzled by
// swz(row) = (row & 7) ^ ((row >> 3) & 1), which keeps the 16-byte fragment reads of 8 (32 banks) or
// 16 (64 banks) consecutive rows on distinct banks;
// fragments are two 16-byte reads of packed words, split into hi and lo halves (v_perm_b32).
// Requirements (checked on the host): A and B K-contiguous with 16-byte quads, K a positive multiple of
// 32 S (whole passes of S steps).  A batch whose offsets are not multiples of four words is staged with 4-byte loads into the same
// layout.  NW waves: 8 (two per SIMD, 64 x 128 or 64 x 64 per wave, accumulators in VGPRs) or 4 (one
// per SIMD, 128 x 128 per wave: a third less fragment traffic from LDS per MFMA; its 256 accumulator
// registers need the AGPRs, so that variant is built in a translation unit of its own).

template <int BM, int BN, int S> struct DmaSmem {
    static constexpr int AW = BM * kBK3, BW = BN * kBK3;   // 32-bit words per stage and operand
};

// The requests are written as inline asm rather than with __builtin_amdgcn_global_load_lds: for the
// builtin the compiler cannot tell the stages apart once it has merged the unrolled steps, and waits for
// every outstanding request (vmcnt(0)) before each fragment read.  The asm has no register results
// (its destination is LDS at M0 + 16 x lane, M0 set in the same block), so nothing the compiler
// allocates is written by the request; completion is awaited explicitly with wait_vm before the
// barrier that precedes every read of a stage.
__device__ __forceinline__ uint32_t lds_addr(const uint32_t *p) {
    return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(__attribute__((address_space(3))) const uint32_t *)p);
}
__device__ __forceinline__ void dma16(const uint32_t *g, uint32_t *lds_base) {
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" : : "s"(lds_addr(lds_base)), "v"(g) : "memory");
}
__device__ __forceinline__ void dma4(const uint32_t *g, uint32_t *lds_base) {
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, off" : : "s"(lds_addr(lds_base)), "v"(g) : "memory");
}

// Wait until at most n vector-memory operations of this wave are outstanding (n a constant).
template <int n> __device__ __forceinline__ void wait_vm() {
    static_assert(n >= 0 && n < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | (0x7 << 4) | (0x3F << 8));
}

template <int BM, int BN, bool C_MFAST, int MI, int S, int NW, bool QUAD>
__device__ __forceinline__ void dma_tile(const ContractArgs<float> &p, int64_t batch, const dcc_batch &bt, int tm,
                                         int tn, uint32_t *sa0, uint32_t *sb0, uint32_t *sa1, uint32_t *sb1,
                                         uint32_t *sa2, uint32_t *sb2, int64_t (*ktab)[2][kBK3]) {
    static_assert(S == 2 || S == 3, "two or three stages");
    static_assert(NW == 4 || NW == 8, "four or eight waves");
    constexpr int WGM = NW == 4 ? 2 : (BM >= BN ? 4 : 2), WGN = NW / WGM;
    constexpr int WM = BM / WGM, WN = BN / WGN;
    constexpr int FM = WM / MI, FN = WN / MI;
    typedef typename AccOf<MI>::t acc_t;
    constexpr int NACC = AccOf<MI>::n;
    constexpr int AW = DmaSmem<BM, BN, S>::AW;
    // 16-byte loads: a wave request covers 8 rows of 128 bytes; per wave and tile DA + DB requests
    constexpr int DA = BM / (8 * NW), DB = BN / (8 * NW), DPT = DA + DB;
    static_assert(DA * 8 * NW == BM && DB * 8 * NW == BN, "rows per request");
    constexpr int DA4 = BM / (2 * NW), DB4 = BN / (2 * NW);  // 4-byte loads: 2 rows per request
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t m0 = bt.m_begin + (int64_t)tm * BM;
    const int64_t m_end = bt.m_end;
    const int n0 = tn * BN;
    const uint32_t *Ap = reinterpret_cast<const uint32_t *>(p.A);
    const uint32_t *Bp = reinterpret_cast<const uint32_t *>(p.B);

    // this lane's rows of its requests; rows outside the operand read row offset 0 (they only feed
    // outputs that are never stored)
    auto arow = [&](int r) __attribute__((always_inline)) -> int64_t {
        const int64_t m = m0 + r;
        return m < m_end ? bt.a_off + map_off(p.am, (uint32_t)m, batch) : 0;
    };
    auto brow = [&](int r) __attribute__((always_inline)) -> int64_t {
        const int n = n0 + r;
        return n < p.N ? bt.b_off + map_off(p.bn, (uint32_t)n, batch) : 0;
    };
    // 16-byte path: request i of this wave covers rows (NW i + wave) 8 .. + 7, lane -> row lane / 8,
    // LDS chunk lane % 8 holding global chunk (lane % 8) ^ swz(row); the
    // row offsets are kept.  4-byte path (misaligned batches only): request i covers rows
    // (NW i + wave) 2 .. + 1, lane -> row lane / 32, word lane % 32; row offsets recomputed per request
    int64_t ra[QUAD ? DA : 1], rb[QUAD ? DB : 1];
    if constexpr (QUAD) {
#pragma unroll
        for (int i = 0; i < DA; ++i) ra[i] = arow((NW * i + wave) * 8 + lane / 8);
#pragma unroll
        for (int i = 0; i < DB; ++i) rb[i] = brow((NW * i + wave) * 8 + lane / 8);
    }
    auto swz = [](int row) __attribute__((always_inline)) { return (row & 7) ^ ((row >> 3) & 1); };
    // 16-byte path: row = (NW i + wave) 8 + lane / 8 with NW even, so swz(row) = (lane / 8) ^ (wave & 1)
    // for every request i
    const int gchunk = (lane % 8) ^ (lane / 8) ^ (wave & 1);
    const int w4 = lane % 32, row4 = lane / 32;                       // 4-byte path: word, row in pair

    auto fill_table = [&](int t, int k0) __attribute__((always_inline)) {
        if (wave < 4 && lane < 16) {
            const int e = (wave & 1) * 16 + lane;
            const uint32_t k = (uint32_t)(k0 + e);
            ktab[t][wave >> 1][e] = k < (uint32_t)p.K ? koff(wave < 2 ? p.ak : p.bk, k) : int64_t(0);
        }
    };
    // requests of tile kt (K offsets in table slot kt % S) into stage arrays (sa, sb)
    auto issue = [&](int slot, uint32_t *sa, uint32_t *sb) __attribute__((always_inline)) {
        if constexpr (QUAD) {
            const int64_t ka = ktab[slot][0][4 * gchunk], kb = ktab[slot][1][4 * gchunk];
#pragma unroll
            for (int i = 0; i < DA; ++i) dma16(Ap + ra[i] + ka, sa + (NW * i + wave) * 8 * kBK3);
#pragma unroll
            for (int i = 0; i < DB; ++i) dma16(Bp + rb[i] + kb, sb + (NW * i + wave) * 8 * kBK3);
        } else {
            for (int i = 0; i < DA4; ++i) {
                const int row = (NW * i + wave) * 2 + row4;
                const int k = 4 * ((w4 / 4) ^ swz(row)) + (w4 % 4);
                dma4(Ap + arow(row) + ktab[slot][0][k], sa + (NW * i + wave) * 2 * kBK3);
            }
            for (int i = 0; i < DB4; ++i) {
                const int row = (NW * i + wave) * 2 + row4;
                const int k = 4 * ((w4 / 4) ^ swz(row)) + (w4 % 4);
                dma4(Bp + brow(row) + ktab[slot][1][k], sb + (NW * i + wave) * 2 * kBK3);
            }
        }
    };

    const int wm = wave / WGN, wn = wave % WGN;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
    auto acc_row = [&](int i, int r) __attribute__((always_inline)) {
        return MI == 16 ? wm * WM + i * 16 + 4 * (lane / 16) + r
                        : wm * WM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    };
    auto acc_col = [&](int j) __attribute__((always_inline)) { return wn * WN + j * MI + lane % MI; };

    // fragment (hi, lo) of rows rbase.. of a stage for k chunks c, c + 1 (k = 4 c .. 4 c + 7)
    auto frag = [&](const uint32_t *st, int row, int c, bf16x8 &hi, bf16x8 &lo) __attribute__((always_inline)) {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 x = *reinterpret_cast<const u32x4 *>(st + row * kBK3 + 4 * (c ^ swz(row)));
        const u32x4 y = *reinterpret_cast<const u32x4 *>(st + row * kBK3 + 4 * ((c + 1) ^ swz(row)));
        uint32_t h[4], l[4];
        split_pair(x[0], x[1], h[0], l[0]);
        split_pair(x[2], x[3], h[1], l[1]);
        split_pair(y[0], y[1], h[2], l[2]);
        split_pair(y[2], y[3], h[3], l[3]);
        typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
        hi = __builtin_bit_cast(bf16x8, u32x4v{h[0], h[1], h[2], h[3]});
        lo = __builtin_bit_cast(bf16x8, u32x4v{l[0], l[1], l[2], l[3]});
    };
    auto compute = [&](const uint32_t *sa, const uint32_t *sb) __attribute__((always_inline)) {
        if constexpr (MI == 16) {
            bf16x8 ah[FM], al[FM];
#pragma unroll
            for (int i = 0; i < FM; ++i) frag(sa, wm * WM + i * 16 + lane % 16, 2 * (lane / 16), ah[i], al[i]);
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                bf16x8 bh, bl;
                frag(sb, wn * WN + j * 16 + lane % 16, 2 * (lane / 16), bh, bl);
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
                }
                // four waves: keep the scheduler from hoisting every B fragment of the step (their
                // registers would exceed the 256 VGPRs left beside the AGPR accumulators)
                if constexpr (NW == 4) __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                bf16x8 ah[FM], al[FM];
#pragma unroll
                for (int i = 0; i < FM; ++i)
                    frag(sa, wm * WM + i * 32 + lane % 32, 4 * ks + 2 * (lane / 32), ah[i], al[i]);
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    bf16x8 bh, bl;
                    frag(sb, wn * WN + j * 32 + lane % 32, 4 * ks + 2 * (lane / 32), bh, bl);
#pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
                    }
                    if constexpr (NW == 4) __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    };

    const int nk = p.K / kBK3;   // K % 32 == 0
    // prologue: tables of tiles 0 .. S-1, requests of tiles 0 .. S-2
#pragma unroll
    for (int t = 0; t < S; ++t) fill_table(t, t * kBK3);
    __syncthreads();
    issue(0, sa0, sb0);
    if (S == 3 && nk > 1) issue(1, sa1, sb1);
    // One step: wait for tile t (the requests of later tiles may stay in flight), barrier (all waves
    // have their part of tile t in LDS and have finished reading the stage refilled next), request
    // tile t + S - 1 into that stage, refill table slot t % S for tile t + S, multiply tile t.
    auto step = [&](int t, uint32_t *ca, uint32_t *cb, uint32_t *na, uint32_t *nb, int slot_now) __attribute__((always_inline)) {
        if constexpr (S == 3) {
            if (t + 1 < nk) wait_vm<QUAD ? DPT : DA4 + DB4>();
            else wait_vm<0>();
        } else {
            wait_vm<0>();
        }
        __builtin_amdgcn_s_barrier();
        if (t + S - 1 < nk) issue((t + S - 1) % S, na, nb);
        fill_table(slot_now, (t + S) * kBK3);
        compute(ca, cb);
    };
    // nk is a multiple of S (checked on the host), so every pass runs all S steps: no step is
    // conditional, and the accumulators have a single definition per pass
    for (int kt = 0; kt < nk; kt += S) {
        if constexpr (S == 2) {
            step(kt, sa0, sb0, sa1, sb1, 0);
            step(kt + 1, sa1, sb1, sa0, sb0, 1);
        } else {
            step(kt, sa0, sb0, sa2, sb2, 0);
            step(kt + 1, sa1, sb1, sa0, sb0, 1);
            step(kt + 2, sa2, sb2, sa1, sb1, 2);
        }
    }
    wait_vm<0>();
    __syncthreads();

    if constexpr (C_MFAST) {
        // C staged through LDS (stage array sa0 reused) EBN columns at a time, then written along M
        constexpr int EBN = (AW / (BM + 1)) / 16 * 16;
        static_assert(EBN >= 16 && BN % EBN == 0, "epilogue columns per pass");
        float *Cs = reinterpret_cast<float *>(sa0);
        constexpr int NSTEP = NW * 64 / BM;
        const int ml = tid % BM;
        const int64_t m = m0 + ml;
        const bool mok = m < m_end;
        const int64_t crow = mok ? bt.c_off + map_off(p.cm, (uint32_t)m, batch) : 0;
        for (int pass = 0; pass < BN / EBN; ++pass) {
            if (pass > 0) __syncthreads();
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

template <int BM, int BN, bool C_MFAST, int MI, int S, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4)))
void contract_bf16x3_dma_kernel(const ContractArgs<float> p) {
    constexpr int AW = DmaSmem<BM, BN, S>::AW, BW = DmaSmem<BM, BN, S>::BW;
    __shared__ __align__(16) uint32_t sa0[AW], sb0[BW], sa1[AW], sb1[BW];
    __shared__ __align__(16) uint32_t sa2[S == 3 ? AW : 4], sb2[S == 3 ? BW : 4];
    __shared__ int64_t ktab[S][2][kBK3];
    const int64_t L = xcd_remap((int64_t)blockIdx.x, p.total_tiles);
    const int64_t per_batch = (int64_t)p.tiles_m * p.tiles_n;
    const int64_t batch = L / per_batch;
    int tm, tn;
    tile_coords(L - batch * per_batch, p.tiles_m, p.tiles_n, tm, tn);
    const dcc_batch bt = p.batches[batch];
    if (bt.m_begin + (int64_t)tm * BM >= bt.m_end) return;
    // 16-byte requests need every vector 16-byte aligned: batch offsets in multiples of four words
    if (((bt.a_off | bt.b_off) & 3) == 0)
        dma_tile<BM, BN, C_MFAST, MI, S, NW, true>(p, batch, bt, tm, tn, sa0, sb0, sa1, sb1, sa2, sb2, ktab);
    else
        dma_tile<BM, BN, C_MFAST, MI, S, NW, false>(p, batch, bt, tm, tn, sa0, sb0, sa1, sb1, sa2, sb2, ktab);
}

template <int BM, int BN, int MI, int S, int NW = 8>
int run_bf16x3_dma(ContractArgs<float> &args, int64_t nbatch, int64_t max_rows, bool c_mfast, hipStream_t s) {
    args.tiles_m = (int)((max_rows + BM - 1) / BM);
    args.tiles_n = (int)((args.N + BN - 1) / BN);
    args.total_tiles = nbatch * args.tiles_m * args.tiles_n;
    DCC_CHECK(args.total_tiles < ((int64_t)1 << 31), "too many tiles");
    const dim3 grid((unsigned)args.total_tiles), block(NW * 64);
    if (c_mfast)
        hipLaunchKernelGGL((contract_bf16x3_dma_kernel<BM, BN, true, MI, S, NW>), grid, block, 0, s, args);
    else
        hipLaunchKernelGGL((contract_bf16x3_dma_kernel<BM, BN, false, MI, S, NW>), grid, block, 0, s, args);
    DCC_LAUNCH_CHECK();
    return 0;
}

// The 256 x 256 tile on four waves (128 x 128 per wave, 16 x 16 x 32 or 32 x 32 x 16 MFMAs), defined in
// dcc_contract_wide.hip, which keeps MFMA accumulators in AGPRs.
int run_bf16x3_dma_wide(ContractArgs<float> &args, int64_t nbatch, int64_t max_rows, bool c_mfast, int mi,
                        hipStream_t s);


}  // namespace dcc_bf16x3
