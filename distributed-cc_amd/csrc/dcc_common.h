// Internal helpers shared by the libdcc translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "dcc.h"

namespace dcc {

void set_error(const std::string &msg);
void clear_error();

inline hipStream_t as_stream(dcc_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// K index map of the gather-GEMMs, evaluated without integer division or branches: the two
// quotients of the three-level mixed-radix decomposition use round-up reciprocals,
// q = (n * mul) >> sh with sh = 31 + ceil(log2 d) and mul = ceil(2^sh / d), which is exact for every
// n < 2^31 (the error n * (mul * d - 2^sh) / (d * 2^sh) stays below 1/d).
struct KMap {
    int64_t s0, s1, s2;   // strides of the three levels
    uint32_t e2, e12;     // extents of the inner level and of the two inner levels
    uint32_t mul2, mul12; // reciprocals of e2 and e12
    uint32_t sh2, sh12;
    // Advancing an index by a fixed step (the K tile of a GEMM) without division: the step's digits
    // d1, d2 in the radix (e1, e2), the offset the step adds, and the corrections of a carry out of
    // the inner (c2) and the middle (c1) level.
    uint32_t e1, d1, d2;
    int64_t doff, c2, c1;
};

__host__ __device__ inline void magic(uint32_t d, uint32_t &mul, uint32_t &sh) {
    uint32_t l = 0;
    while (((uint64_t)1 << l) < d) ++l;
    sh = 31 + l;
    mul = (uint32_t)((((uint64_t)1 << sh) + d - 1) / d);
}

__host__ __device__ inline KMap make_kmap(const dcc_idxmap &m, uint32_t step = 16) {
    KMap k;
    k.s0 = m.s0;
    k.s1 = m.s1;
    k.s2 = m.s2;
    k.e2 = (uint32_t)m.e2;
    k.e12 = (uint32_t)(m.e1 * m.e2);
    magic(k.e2, k.mul2, k.sh2);
    magic(k.e12, k.mul12, k.sh12);
    k.e1 = (uint32_t)m.e1;
    k.d2 = step % k.e2;
    k.d1 = (step / k.e2) % k.e1;
    const uint32_t d0 = step / k.e12;
    k.doff = (int64_t)d0 * m.s0 + (int64_t)k.d1 * m.s1 + (int64_t)k.d2 * m.s2;
    k.c2 = m.s1 - (int64_t)k.e2 * m.s2;
    k.c1 = m.s0 - (int64_t)k.e1 * m.s1;
    return k;
}

// Offset of index k under the map, kept up to date while k advances by the map's step.
struct KIter {
    int64_t off;
    uint32_t q1, q2;   // middle and inner digits of k
};

__device__ __forceinline__ uint32_t magic_div(uint32_t n, uint32_t mul, uint32_t sh) {
    return (uint32_t)(((uint64_t)n * mul) >> sh);
}

// Offset of index k (< 2^31) under the map.
__device__ __forceinline__ int64_t koff(const KMap &m, uint32_t k) {
    const uint32_t q0 = magic_div(k, m.mul12, m.sh12);
    const uint32_t r0 = k - q0 * m.e12;
    const uint32_t q1 = magic_div(r0, m.mul2, m.sh2);
    const uint32_t q2 = r0 - q1 * m.e2;
    return (int64_t)q0 * m.s0 + (int64_t)q1 * m.s1 + (int64_t)q2 * m.s2;
}

__device__ __forceinline__ KIter kiter(const KMap &m, uint32_t k) {
    const uint32_t q0 = magic_div(k, m.mul12, m.sh12);
    const uint32_t r0 = k - q0 * m.e12;
    const uint32_t q1 = magic_div(r0, m.mul2, m.sh2);
    const uint32_t q2 = r0 - q1 * m.e2;
    KIter it;
    it.off = (int64_t)q0 * m.s0 + (int64_t)q1 * m.s1 + (int64_t)q2 * m.s2;
    it.q1 = q1;
    it.q2 = q2;
    return it;
}

// k += step: the digits d2 < e2 and d1 < e1 carry at most once per level.
__device__ __forceinline__ void kstep(const KMap &m, KIter &it) {
    it.off += m.doff;
    it.q2 += m.d2;
    it.q1 += m.d1;
    if (it.q2 >= m.e2) {
        it.q2 -= m.e2;
        it.q1 += 1;
        it.off += m.c2;
    }
    if (it.q1 >= m.e1) {
        it.q1 -= m.e1;
        it.off += m.c1;
    }
}

// Split-BF16x3 word of x: bf16(x) in the low half, bf16(x - bf16(x)) in the high half, both rounded
// to nearest even (the v_cvt_pk_bf16_f32 conversion the split-BF16x3 GEMM applies to FP32 operands).
__device__ __forceinline__ uint32_t pack_bf16x3(float x) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    const __bf16 h = (__bf16)x;
    const float hf = __uint_as_float((uint32_t)__builtin_bit_cast(unsigned short, h) << 16);
    const bf2 w = {h, (__bf16)(x - hf)};
    return __builtin_bit_cast(uint32_t, w);
}

}  // namespace dcc

// Return a nonzero status from an entry point when a HIP call fails.
#define DCC_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t err_ = (call);                                                          \
        if (err_ != hipSuccess) {                                                          \
            dcc::set_error(std::string(#call) + ": " + hipGetErrorString(err_));            \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

#define DCC_CHECK(cond, msg)                                                               \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            dcc::set_error(std::string(__func__) + ": " + (msg));                           \
            return 2;                                                                      \
        }                                                                                  \
    } while (0)

// Launch status: a launch error is returned as status 1.
#define DCC_LAUNCH_CHECK()                                                                 \
    do {                                                                                   \
        hipError_t err_ = hipGetLastError();                                               \
        if (err_ != hipSuccess) {                                                          \
            dcc::set_error(std::string(__func__) + " launch: " + hipGetErrorString(err_));  \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)
