// Bandwidth-bound helpers of the RCCSDT iteration around the triples update:
// strided tensor axpby (block views, transposes, packing), deterministic dot
// products and the singles/doubles denominators.
#include "dcc_common.h"

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
