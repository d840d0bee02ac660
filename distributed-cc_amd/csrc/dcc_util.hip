// Bandwidth-bound helpers of the RCCSDT iteration around the triples update:
// strided tensor axpby (block views, transposes, packing), deterministic dot
// products and the singles/doubles denominators.
#include "dcc_common.h"
#include <type_traits>

namespace {

constexpr int kDotBlocks = 1024;  // partial sums per dot product (fixed: deterministic order)

// dst[idx] = beta * dst[idx] + alpha * src[idx] over an up-to-6-d index space.
// The innermost axis is mapped to consecutive lanes so both sides coalesce
// when it is unit-stride.
template <typename TD, typename TS>
__global__ __launch_bounds__(256) void axpby_kernel(TD *__restrict__ dst, const TS *__restrict__ src,
                                                    dcc_strided lay, int64_t total, double alpha, double beta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        int64_t rem = t, od = 0, os = 0;
        for (int d = (int)lay.ndim - 1; d >= 0; --d) {
            const int64_t n = lay.shape[d];
            const int64_t x = rem % n;
            rem /= n;
            od += x * lay.dst_stride[d];
            os += x * lay.src_stride[d];
        }
        const double s = alpha * (double)src[os];
        dst[od] = (TD)((beta == 0.0) ? s : beta * (double)dst[od] + s);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void dot_partial_kernel(const T *__restrict__ x, const T *__restrict__ y,
                                                          int64_t n, double *__restrict__ partials) {
    __shared__ double sh[256];
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc += (double)x[i] * (double)y[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void dot_final_kernel(const double *__restrict__ partials, int nparts,
                                                        double *__restrict__ out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) acc += partials[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// rank 1: x[i,a] /= eia[i,a];  rank 2: x[i,j,a,b] /= eia[i,a] + eia[j,b]
template <typename T>
__global__ __launch_bounds__(256) void denom_kernel(T *__restrict__ x, int64_t no, int64_t nv, int rank,
                                                    const double *__restrict__ eia) {
    const int64_t total = rank == 1 ? no * nv : no * no * nv * nv;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        double d;
        if (rank == 1) {
            d = eia[t];
        } else {
            const int64_t b = t % nv, a = (t / nv) % nv, j = (t / (nv * nv)) % no, i = t / (nv * nv * no);
            d = eia[i * nv + a] + eia[j * nv + b];
        }
        x[t] = (T)((double)x[t] / d);
    }
}

// dst[i] = beta * dst[i] + alpha * sum_s src[s * n + i], s in fixed order
template <typename T>
__global__ __launch_bounds__(256) void sum_slices_kernel(T *__restrict__ dst, const T *__restrict__ src,
                                                         int64_t nslices, int64_t n, double alpha, double beta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double acc = 0.0;
        for (int64_t s = 0; s < nslices; ++s) acc += (double)src[s * n + i];
        dst[i] = (T)((beta == 0.0 ? 0.0 : beta * (double)dst[i]) + alpha * acc);
    }
}

// Split-BF16x3 packing: dst[i] = bf16(x) | bf16(x - bf16(x)) << 16 with x = src[i] (round to nearest
// even, the v_cvt_pk_bf16_f32 conversion of the GEMM's own split).  dst may alias src.
__global__ __launch_bounds__(256) void pack_bf16x3_kernel(const float *src, uint32_t *dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = dcc::pack_bf16x3(src[i]);
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *__restrict__ p) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kDotBlocks) p[i] = 0.0;
}

inline unsigned grid_for(int64_t total) {
    int64_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

// Operands of the fused T2 -> T3 base-term GEMM for the first occupied indices i0 .. i0+n-1, K = (d, then l):
//   Ap[i,b,j,a,K] = W_ovvv[j,a,b,K] (K < v), -W_oovo[i0+i,j,a,K-v] (K >= v)
//   Bp[i,b,K,k,c] = t2[i0+i,k,K,c]   (K < v),  t2[K-v,k,b,c]      (K >= v)
// One element per thread, consecutive threads along the contiguous index of both outputs (K for Ap,
// (k,c) for Bp), so every write and most reads coalesce; O = uint32_t writes split-BF16x3 words.
template <typename T, typename O>
__global__ __launch_bounds__(256) void t2t3_operands_kernel(O *__restrict__ Ap, O *__restrict__ Bp,
                                                            const T *__restrict__ W_ovvv,
                                                            const T *__restrict__ W_oovo,
                                                            const T *__restrict__ t2, int o, int v, int i0,
                                                            int64_t na, int64_t total) {
    const int K2 = v + o, ov = o * v;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        T x;
        O *dst;
        if (t < na) {                      // Ap: row (i,b,j,a), column K
            const int row = (int)(t / K2), K = (int)(t - (int64_t)row * K2);
            const int a = row % v, r1 = row / v, j = r1 % o, r2 = r1 / o, b = r2 % v, i = r2 / v;
            x = K < v ? W_ovvv[(((int64_t)j * v + a) * v + b) * v + K]
                      : -W_oovo[(((int64_t)(i0 + i) * o + j) * v + a) * o + (K - v)];
            dst = Ap + t;
        } else {                           // Bp: row (i,b,K), column (k,c)
            const int64_t u = t - na;
            const int row = (int)(u / ov), col = (int)(u - (int64_t)row * ov);
            const int c = col % v, k = col / v, K = row % K2, r1 = row / K2, b = r1 % v, i = r1 / v;
            x = K < v ? t2[(((int64_t)(i0 + i) * o + k) * v + K) * v + c]
                      : t2[(((int64_t)(K - v) * o + k) * v + b) * v + c];
            dst = Bp + u;
        }
        if constexpr (std::is_same<O, uint32_t>::value) *dst = dcc::pack_bf16x3((float)x);
        else *dst = (O)x;
    }
}

template <typename T, typename O>
int t2t3_operands(O *Ap, O *Bp, const T *W_ovvv, const T *W_oovo, const T *t2, int64_t nocc, int64_t nvir,
                  int64_t i0, int64_t n, dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nvir >= 0 && i0 >= 0 && n >= 0 && i0 + n <= nocc, "bad occupied range");
    if (n == 0 || nvir == 0 || nocc == 0) return 0;
    DCC_CHECK(Ap != nullptr && Bp != nullptr && W_ovvv != nullptr && W_oovo != nullptr && t2 != nullptr,
              "null pointer");
    const int64_t rows_a = n * nvir * nocc * nvir, rows_b = n * nvir * (nvir + nocc);
    DCC_CHECK(rows_a < ((int64_t)1 << 31) && rows_b < ((int64_t)1 << 31) && nocc * nvir < ((int64_t)1 << 31),
              "extent exceeds 2^31");
    const int64_t na = rows_a * (nvir + nocc), total = na + rows_b * nocc * nvir;
    hipLaunchKernelGGL((t2t3_operands_kernel<T, O>), dim3(grid_for(total)), dim3(256), 0, dcc::as_stream(stream),
                       Ap, Bp, W_ovvv, W_oovo, t2, (int)nocc, (int)nvir, (int)i0, na, total);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename TD, typename TS>
int tensor_axpby(TD *dst, const TS *src, const dcc_strided *layout, double alpha, double beta, dcc_stream_t stream) {
    DCC_CHECK(layout != nullptr, "null layout");
    DCC_CHECK(layout->ndim >= 1 && layout->ndim <= 6, "ndim must be 1..6");
    int64_t total = 1;
    for (int d = 0; d < layout->ndim; ++d) {
        DCC_CHECK(layout->shape[d] >= 0, "negative extent");
        total *= layout->shape[d];
    }
    if (total == 0) return 0;
    DCC_CHECK(dst != nullptr && src != nullptr, "null pointer");
    hipLaunchKernelGGL((axpby_kernel<TD, TS>), dim3(grid_for(total)), dim3(256), 0, dcc::as_stream(stream), dst,
                       src, *layout, total, alpha, beta);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int dot(const T *x, const T *y, int64_t n, double *partials, double *out, dcc_stream_t stream) {
    DCC_CHECK(n >= 0, "negative length");
    DCC_CHECK(partials != nullptr && out != nullptr, "null pointer");
    if (n > 0) {
        DCC_CHECK(x != nullptr && y != nullptr, "null pointer");
        hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(kDotBlocks), dim3(256), 0, dcc::as_stream(stream), x, y, n,
                           partials);
        DCC_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(fill_zero_kernel, dim3(4), dim3(256), 0, dcc::as_stream(stream), partials);
        DCC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, dcc::as_stream(stream), partials, kDotBlocks, out);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int denom_divide(T *x, int64_t nocc, int64_t nvir, int64_t rank, const double *eia, dcc_stream_t stream) {
    DCC_CHECK(rank == 1 || rank == 2, "rank must be 1 or 2");
    DCC_CHECK(nocc >= 0 && nvir >= 0, "negative extent");
    const int64_t total = rank == 1 ? nocc * nvir : nocc * nocc * nvir * nvir;
    if (total == 0) return 0;
    DCC_CHECK(x != nullptr && eia != nullptr, "null pointer");
    hipLaunchKernelGGL(denom_kernel<T>, dim3(grid_for(total)), dim3(256), 0, dcc::as_stream(stream), x, nocc, nvir,
                       (int)rank, eia);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int sum_slices(T *dst, const T *src, int64_t nslices, int64_t n, double alpha, double beta, dcc_stream_t stream) {
    DCC_CHECK(nslices >= 0 && n >= 0, "negative extent");
    if (n == 0) return 0;
    DCC_CHECK(dst != nullptr && (nslices == 0 || src != nullptr), "null pointer");
    hipLaunchKernelGGL(sum_slices_kernel<T>, dim3(grid_for(n)), dim3(256), 0, dcc::as_stream(stream), dst, src,
                       nslices, n, alpha, beta);
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" {

int64_t dcc_dot_partials(void) { return kDotBlocks; }

int dcc_pack_bf16x3(const float *src, uint32_t *dst, int64_t n, dcc_stream_t stream) {
    DCC_CHECK(n >= 0, "negative length");
    if (n == 0) return 0;
    DCC_CHECK(src != nullptr && dst != nullptr, "null pointer");
    hipLaunchKernelGGL(pack_bf16x3_kernel, dim3(grid_for(n)), dim3(256), 0, dcc::as_stream(stream), src, dst, n);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_t2t3_operands_f64(double *Ap, double *Bp, const double *W_ovvv, const double *W_oovo, const double *t2,
                          int64_t nocc, int64_t nvir, int64_t i0, int64_t n, dcc_stream_t stream) {
    return t2t3_operands<double, double>(Ap, Bp, W_ovvv, W_oovo, t2, nocc, nvir, i0, n, stream);
}

int dcc_t2t3_operands_f32(float *Ap, float *Bp, const float *W_ovvv, const float *W_oovo, const float *t2,
                          int64_t nocc, int64_t nvir, int64_t i0, int64_t n, dcc_stream_t stream) {
    return t2t3_operands<float, float>(Ap, Bp, W_ovvv, W_oovo, t2, nocc, nvir, i0, n, stream);
}

int dcc_t2t3_operands_packed_f32(uint32_t *Ap, uint32_t *Bp, const float *W_ovvv, const float *W_oovo,
                                 const float *t2, int64_t nocc, int64_t nvir, int64_t i0, int64_t n,
                                 dcc_stream_t stream) {
    return t2t3_operands<float, uint32_t>(Ap, Bp, W_ovvv, W_oovo, t2, nocc, nvir, i0, n, stream);
}

int dcc_sum_slices_f64(double *dst, const double *src, int64_t nslices, int64_t n, double alpha, double beta,
                       dcc_stream_t stream) {
    return sum_slices<double>(dst, src, nslices, n, alpha, beta, stream);
}

int dcc_sum_slices_f32(float *dst, const float *src, int64_t nslices, int64_t n, double alpha, double beta,
                       dcc_stream_t stream) {
    return sum_slices<float>(dst, src, nslices, n, alpha, beta, stream);
}

int dcc_tensor_axpby_f64(double *dst, const double *src, const dcc_strided *layout, double alpha, double beta,
                         dcc_stream_t stream) {
    return tensor_axpby<double, double>(dst, src, layout, alpha, beta, stream);
}

int dcc_tensor_axpby_f32(float *dst, const float *src, const dcc_strided *layout, double alpha, double beta,
                         dcc_stream_t stream) {
    return tensor_axpby<float, float>(dst, src, layout, alpha, beta, stream);
}

int dcc_tensor_axpby_f32_f64(float *dst, const double *src, const dcc_strided *layout, double alpha, double beta,
                             dcc_stream_t stream) {
    return tensor_axpby<float, double>(dst, src, layout, alpha, beta, stream);
}

int dcc_tensor_axpby_f64_f32(double *dst, const float *src, const dcc_strided *layout, double alpha, double beta,
                             dcc_stream_t stream) {
    return tensor_axpby<double, float>(dst, src, layout, alpha, beta, stream);
}

int dcc_dot_f64(const double *x, const double *y, int64_t n, double *partials, double *out, dcc_stream_t stream) {
    return dot<double>(x, y, n, partials, out, stream);
}

int dcc_dot_f32(const float *x, const float *y, int64_t n, double *partials, double *out, dcc_stream_t stream) {
    return dot<float>(x, y, n, partials, out, stream);
}

int dcc_denom_divide_f64(double *x, int64_t nocc, int64_t nvir, int64_t rank, const double *eia,
                         dcc_stream_t stream) {
    return denom_divide<double>(x, nocc, nvir, rank, eia, stream);
}

}  // extern "C"
