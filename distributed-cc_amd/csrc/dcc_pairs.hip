// Pair-symmetric form of the W_vvvv x T3 terms of the triples residual (distr_cc/rccsdt.py:424-426).
//
// Each of the three terms is r[.. x .. y ..] += sum_{d,e} W1[x,y,d,e] T[.. d .. e ..] with the pair (x, y)
// of output indices and the pair (d, e) of T3 indices at the same two positions of the block and the
// third index z shared.  The dressed W_vvvv (and W1, which folds in the F_vv terms) is pair-symmetric,
// W1[x,y,d,e] = W1[y,x,e,d], so with
//   TS[q(d<=e)] = T[de] + T[ed],   TA[q(d<e)] = T[de] - T[ed],
//   WS[p(x<=y), q(d<=e)] = W1[xy,de] + W1[xy,ed] (d<e), W1[xy,dd] (d=e),
//   WA[p(x<y), q(d<e)]   = W1[xy,de] - W1[xy,ed],
// the products US = WS TS and UA = WA TA hold r[xy] + r[yx] and r[xy] - r[yx] over the triangular pairs,
// and r[xy] = (US + UA) / 2, r[yx] = (US - UA) / 2, r[xx] = US / 2.  The two products have half the rows
// and half the K of the plain v^2 x v^2 one, so the term costs half the flops.
//
// Pair indices are row-major over the upper triangle: q(d<=e) = d v - d(d-1)/2 + (e - d) (nS = v(v+1)/2
// of them), q(d<e) = d (v-1) - d(d-1)/2 + (e - d - 1) (nA = v(v-1)/2).  Layouts (kp >= nS: the padded K,
// mp = nS: rows of U; the second half of Tp / Up holds the antisymmetric part):
//   Wp [nS + nA][kp]                       WS rows, then WA rows; columns past nS / nA are zero
//   Tp [2][nb][kp][v]  (pair axes 0,1)     TS / TA of each block, the shared index z innermost
//   Tp [2][nb][v][kp]  (pair axes 0,2 / 1,2)  z outer, pairs contiguous (K-contiguous GEMM operand)
//   Up [2][nb][mp][v]  (pair axes 0,1)     US / UA as the GEMM writes them
//   Up [2][nb][v][mp]  (pair axes 0,2 / 1,2)
// Padding entries of Tp are written as zeros by every split, since the two layouts share one buffer.
#include "dcc_common.h"
#include <type_traits>

namespace {

// plane tile edge (64 x 64 tiles of 4-byte elements ran the plane passes about 20 % slower: fewer
// workgroups in flight per CU)
template <typename T> constexpr int plane_tile() { return 32; }
constexpr int kPThreads = 256;
// flat element loops run on 32-bit indices below this extent (the loop index plus one grid stride stays
// below 2^32)
constexpr int64_t kIndex32 = (int64_t)1 << 31;

__device__ __forceinline__ int64_t row_start(int64_t d, int v, bool strict) {
    return strict ? d * (2 * (int64_t)v - d - 1) / 2 : d * (2 * (int64_t)v - d + 1) / 2;
}

// (d, e) of pair index q (d <= e, or d < e when strict); q must be below the pair count.  A single-precision
// estimate of the row, corrected by whole steps against the exact integer row starts (one step at most
// for any nvir this library accepts, where the estimate is off by far less than a row).
__device__ __forceinline__ void decode_pair(int64_t q, int v, bool strict, int &d, int &e) {
    const float w = strict ? 2.0f * v - 1.0f : 2.0f * v + 1.0f;
    int x = (int)((w - sqrtf(fmaxf(w * w - 8.0f * (float)q, 0.0f))) * 0.5f);
    x = x < 0 ? 0 : (x >= v ? v - 1 : x);
    while (x > 0 && row_start(x, v, strict) > q) --x;
    while (row_start(x + 1, v, strict) <= q) ++x;
    d = x;
    e = (int)(q - row_start(x, v, strict)) + x + (strict ? 1 : 0);
}

template <typename O, typename T>
__device__ __forceinline__ O out_value(T x) {
    if constexpr (std::is_same<O, uint32_t>::value) return dcc::pack_bf16x3((float)x);
    else return (O)x;
}

inline unsigned grid_for(int64_t total) {
    int64_t blocks = (total + kPThreads - 1) / kPThreads;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

// Wp[r][q], r < nS: WS, else WA (row r - nS); one element per thread, q fastest.
// I: the index type of the flat element loop (uint32_t when the extent allows: cheap divisions).
template <typename T, typename O, typename I>
__global__ __launch_bounds__(kPThreads) void pair_weights_kernel(const T *__restrict__ W1, int v, int64_t nS,
                                                                int64_t nA, I kp, O *__restrict__ Wp,
                                                                I total) {
    const I stride = (I)gridDim.x * blockDim.x;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const I r_ = t / kp;
        const int64_t r = r_, q = t - r_ * kp;
        const bool anti = r >= nS;
        T x = (T)0;
        if (q < (anti ? nA : nS)) {
            int a, b, d, e;
            decode_pair(anti ? r - nS : r, v, anti, a, b);
            decode_pair(q, v, anti, d, e);
            const int64_t row = ((int64_t)a * v + b) * v;
            const T w_de = W1[(row + d) * v + e];
            if (anti) x = w_de - W1[(row + e) * v + d];
            else x = d == e ? w_de : w_de + W1[(row + e) * v + d];
        }
        Wp[t] = out_value<O>(x);
    }
}

// Pair axes (0, 1): Tp[part][b][q][z] = T[b,d,e,z] +- T[b,e,d,z], z innermost on both sides.
template <typename T, typename O, typename I>
__global__ __launch_bounds__(kPThreads) void pair_split_row_kernel(const T *__restrict__ t3, int v, int64_t nS,
                                                                  int64_t nA, I kp, I half,
                                                                  O *__restrict__ Tp) {
    const int64_t v3 = (int64_t)v * v * v;
    const I stride = (I)gridDim.x * blockDim.x, vv = (I)v;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < 2 * half; t += stride) {
        const bool anti = t >= half;
        const I u = anti ? t - half : t;
        const I row_ = u / vv, b_ = row_ / kp;
        const int64_t z = u - row_ * vv, b = b_, q = row_ - b_ * kp;
        T x = (T)0;
        if (q < (anti ? nA : nS)) {
            int d, e;
            decode_pair(q, v, anti, d, e);
            const T *blk = t3 + b * v3;
            const T p = blk[((int64_t)d * v + e) * v + z], m = blk[((int64_t)e * v + d) * v + z];
            x = anti ? p - m : p + m;
        }
        Tp[t] = out_value<O>(x);
    }
}

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
template <> struct Vec4<uint32_t> { using type = uint4; };

// The row split with four consecutive z per thread as one vector (v % 4 == 0, 16-byte aligned blocks):
// one pair decode and two vector loads per four elements.  g runs over groups of four.
template <typename T, typename O, typename I>
__global__ __launch_bounds__(kPThreads) void pair_split_row4_kernel(const T *__restrict__ t3, int v, int64_t nS,
                                                                   int64_t nA, I kp, I half4,
                                                                   O *__restrict__ Tp) {
    using TV = typename Vec4<T>::type;
    using OV = typename Vec4<O>::type;
    const int64_t v3 = (int64_t)v * v * v;
    const I stride = (I)gridDim.x * blockDim.x, v4 = (I)(v / 4);
    for (I g = (I)blockIdx.x * blockDim.x + threadIdx.x; g < 2 * half4; g += stride) {
        const bool anti = g >= half4;
        const I u = anti ? g - half4 : g;
        const I row_ = u / v4, b_ = row_ / kp;
        const int64_t z = 4 * (int64_t)(u - row_ * v4), b = b_, q = row_ - b_ * kp;
        OV out;
        if (q < (anti ? nA : nS)) {
            int d, e;
            decode_pair(q, v, anti, d, e);
            const T *blk = t3 + b * v3;
            const TV p = *reinterpret_cast<const TV *>(blk + ((int64_t)d * v + e) * v + z);
            const TV m = *reinterpret_cast<const TV *>(blk + ((int64_t)e * v + d) * v + z);
            out.x = out_value<O>(anti ? p.x - m.x : p.x + m.x);
            out.y = out_value<O>(anti ? p.y - m.y : p.y + m.y);
            out.z = out_value<O>(anti ? p.z - m.z : p.z + m.z);
            out.w = out_value<O>(anti ? p.w - m.w : p.w + m.w);
        } else {
            out.x = out.y = out.z = out.w = out_value<O>((T)0);
        }
        reinterpret_cast<OV *>(Tp)[g] = out;
    }
}

// Pair axes (0, 2) or (1, 2): the pair spans a plane P[x][y] = T[b, x sx + z sz + y] whose columns are
// contiguous.  One workgroup per (block, z, tile pair dt <= et): it stages the tiles P[dt][et] and
// P[et][dt] in LDS and writes TS[q(d,e)] = P[d][e] + P[e][d] (d <= e) and TA[q(d,e)] = P[d][e] - P[e][d]
// (d < e) along e, so reads and writes both run along contiguous rows.  The extra workgroup per (block,
// z) (tile index ntp) writes the zero padding of both rows.
template <typename T, typename O, int PT>
__global__ __launch_bounds__(kPThreads) void pair_split_plane_kernel(const T *__restrict__ t3, int v, int64_t sx,
                                                                    int64_t sz, int64_t nS, int64_t nA,
                                                                    int64_t kp, int64_t half, int ntile, int ntp,
                                                                    O *__restrict__ Tp) {
    __shared__ T s1[PT][PT + 1], s2[PT][PT + 1];
    const int64_t wg = blockIdx.x;
    const int tp = (int)(wg % (ntp + 1));
    const int64_t bz = wg / (ntp + 1);
    const int64_t b = bz / v, z = bz - b * v;
    O *rowS = Tp + bz * kp, *rowA = Tp + half + bz * kp;
    const int tid = threadIdx.x;
    if (tp == ntp) {
        for (int64_t q = nS + tid; q < kp; q += kPThreads) rowS[q] = out_value<O>((T)0);
        for (int64_t q = nA + tid; q < kp; q += kPThreads) rowA[q] = out_value<O>((T)0);
        return;
    }
    int dt = 0, rem = tp;
    while (rem >= ntile - dt) {
        rem -= ntile - dt;
        ++dt;
    }
    const int et = dt + rem;
    const int d0 = dt * PT, e0 = et * PT;
    const T *plane = t3 + b * (int64_t)v * v * v + z * sz;
    const int c = tid % PT, r0 = tid / PT;
    for (int r = r0; r < PT; r += kPThreads / PT) {
        const int x1 = d0 + r, y1 = e0 + c, x2 = e0 + r, y2 = d0 + c;
        s1[r][c] = (x1 < v && y1 < v) ? plane[x1 * sx + y1] : (T)0;
        s2[r][c] = (x2 < v && y2 < v) ? plane[x2 * sx + y2] : (T)0;
    }
    __syncthreads();
    for (int r = r0; r < PT; r += kPThreads / PT) {
        const int d = d0 + r, e = e0 + c;
        if (d >= v || e >= v || d > e) continue;
        const T p = s1[r][c], m = s2[c][r];
        rowS[row_start(d, v, false) + (e - d)] = out_value<O>(p + m);
        if (d < e) rowA[row_start(d, v, true) + (e - d - 1)] = out_value<O>(p - m);
    }
}

// Pair axes (0, 1): R[b,x,y,z] += (US[b][p][z] +- UA[b][p][z]) / 2 with p the pair of (min, max) of (x, y).
template <typename T, typename I>
__global__ __launch_bounds__(kPThreads) void pair_unfold_row_kernel(const T *__restrict__ Up, int v, int64_t mp,
                                                                   int64_t half, T *__restrict__ R, I total) {
    const I stride = (I)gridDim.x * blockDim.x, vv = (I)v;
    for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const I r1_ = t / vv, r2_ = r1_ / vv, b_ = r2_ / vv;
        const int64_t z = t - r1_ * vv, y = r1_ - r2_ * vv, b = b_, x = r2_ - b_ * vv;
        const int64_t lo = x < y ? x : y, hi = x < y ? y : x;
        const T *U = Up + b * mp * v;
        T s = U[(row_start(lo, v, false) + (hi - lo)) * v + z];
        if (lo < hi) {
            const T a = U[half + (row_start(lo, v, true) + (hi - lo - 1)) * v + z];
            s = x < y ? s + a : s - a;
        }
        R[t] += (T)0.5 * s;
    }
}

// The row unfold with four consecutive z per thread as one vector (v % 4 == 0, 16-byte aligned).
template <typename T, typename I>
__global__ __launch_bounds__(kPThreads) void pair_unfold_row4_kernel(const T *__restrict__ Up, int v, int64_t mp,
                                                                    int64_t half, T *__restrict__ R, I total4) {
    using TV = typename Vec4<T>::type;
    const I stride = (I)gridDim.x * blockDim.x, v4 = (I)(v / 4), vv = (I)v;
    for (I g = (I)blockIdx.x * blockDim.x + threadIdx.x; g < total4; g += stride) {
        const I r1_ = g / v4, r2_ = r1_ / vv, b_ = r2_ / vv;
        const int64_t z = 4 * (int64_t)(g - r1_ * v4), y = r1_ - r2_ * vv, b = b_, x = r2_ - b_ * vv;
        const int64_t lo = x < y ? x : y, hi = x < y ? y : x;
        const T *U = Up + b * mp * v;
        TV s = *reinterpret_cast<const TV *>(U + (row_start(lo, v, false) + (hi - lo)) * v + z);
        if (lo < hi) {
            const TV a = *reinterpret_cast<const TV *>(U + half + (row_start(lo, v, true) + (hi - lo - 1)) * v + z);
            if (x < y) { s.x += a.x; s.y += a.y; s.z += a.z; s.w += a.w; }
            else { s.x -= a.x; s.y -= a.y; s.z -= a.z; s.w -= a.w; }
        }
        TV r = reinterpret_cast<TV *>(R)[g];
        r.x += (T)0.5 * s.x;
        r.y += (T)0.5 * s.y;
        r.z += (T)0.5 * s.z;
        r.w += (T)0.5 * s.w;
        reinterpret_cast<TV *>(R)[g] = r;
    }
}

// Pair axes (0, 2) or (1, 2): R[b, x sx + z sz + y] += (US +- UA)[b][z][p] / 2.  One workgroup per
// (block, z, tile xt, tile yt); elements with x <= y read their pair along y directly, those with x > y
// come through an LDS tile loaded along x (the contiguous direction of their pairs), so every read and
// the read-modify-write of R run along contiguous rows.
template <typename T, int PT>
__global__ __launch_bounds__(kPThreads) void pair_unfold_plane_kernel(const T *__restrict__ Up, int v, int64_t sx,
                                                                     int64_t sz, int64_t mp, int64_t half,
                                                                     int ntile, T *__restrict__ R) {
    __shared__ T sl[PT][PT + 1];
    const int64_t wg = blockIdx.x;
    const int nt2 = ntile * ntile;
    const int tt = (int)(wg % nt2);
    const int64_t bz = wg / nt2;
    const int64_t b = bz / v, z = bz - b * v;
    const int xt = tt / ntile, yt = tt - xt * ntile;
    const int x0 = xt * PT, y0 = yt * PT;
    const T *US = Up + bz * mp, *UA = Up + half + bz * mp;
    T *plane = R + b * (int64_t)v * v * v + z * sz;
    const int tid = threadIdx.x, c = tid % PT, r0 = tid / PT;
    if (xt >= yt) {
        // lower part: pair (y, x) with y = y0 + r, x = x0 + c, y < x
        for (int r = r0; r < PT; r += kPThreads / PT) {
            const int y = y0 + r, x = x0 + c;
            T val = (T)0;
            if (x < v && y < v && y < x)
                val = (T)0.5 * (US[row_start(y, v, false) + (x - y)] - UA[row_start(y, v, true) + (x - y - 1)]);
            sl[r][c] = val;
        }
        __syncthreads();
    }
    for (int r = r0; r < PT; r += kPThreads / PT) {
        const int x = x0 + r, y = y0 + c;
        if (x >= v || y >= v) continue;
        T val;
        if (x < y) val = (T)0.5 * (US[row_start(x, v, false) + (y - x)] + UA[row_start(x, v, true) + (y - x - 1)]);
        else if (x == y) val = (T)0.5 * US[row_start(x, v, false)];
        else val = sl[c][r];
        plane[x * sx + y] += val;
    }
}

// strides (x, z) of the plane for pair axes 1 = (0, 2) and 2 = (1, 2)
inline void plane_strides(int axes, int64_t v, int64_t &sx, int64_t &sz) {
    if (axes == 1) { sx = v * v; sz = v; }
    else { sx = v; sz = v * v; }
}

// four consecutive z as one vector: rows of nvir elements in multiples of four and a buffer aligned to
// four elements (every row, block and half offset is then aligned too)
inline bool vec4_ok(int64_t nvir, const void *p, size_t elem) {
    return nvir % 4 == 0 && ((uintptr_t)p % (4 * elem)) == 0;
}

int check_sizes(int64_t nblk, int64_t nvir, int32_t axes, int64_t kpad) {
    DCC_CHECK(nblk >= 0 && nvir >= 0, "negative extent");
    DCC_CHECK(axes >= 0 && axes <= 2, "pair axes must be 0 = (0,1), 1 = (0,2) or 2 = (1,2)");
    DCC_CHECK(kpad >= nvir * (nvir + 1) / 2, "kpad below the pair count nvir (nvir + 1) / 2");
    DCC_CHECK(nvir < 46341 && nblk * nvir < ((int64_t)1 << 31), "extent exceeds 2^31");
    return 0;
}

template <typename T, typename O>
int pair_weights(const T *W1, int64_t nvir, int64_t kpad, O *Wp, dcc_stream_t stream) {
    if (int rc = check_sizes(0, nvir, 0, kpad)) return rc;
    if (nvir == 0) return 0;
    DCC_CHECK(W1 != nullptr && Wp != nullptr, "null pointer");
    const int64_t nS = nvir * (nvir + 1) / 2, nA = nvir * (nvir - 1) / 2, total = (nS + nA) * kpad;
    if (total < kIndex32)
        hipLaunchKernelGGL((pair_weights_kernel<T, O, uint32_t>), dim3(grid_for(total)), dim3(kPThreads), 0,
                           dcc::as_stream(stream), W1, (int)nvir, nS, nA, (uint32_t)kpad, Wp, (uint32_t)total);
    else
        hipLaunchKernelGGL((pair_weights_kernel<T, O, int64_t>), dim3(grid_for(total)), dim3(kPThreads), 0,
                           dcc::as_stream(stream), W1, (int)nvir, nS, nA, kpad, Wp, total);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T, typename O>
int pair_split(const T *t3, int64_t nblk, int64_t nvir, int32_t axes, int64_t kpad, O *Tp, dcc_stream_t stream) {
    if (int rc = check_sizes(nblk, nvir, axes, kpad)) return rc;
    if (nblk == 0 || nvir == 0) return 0;
    DCC_CHECK(t3 != nullptr && Tp != nullptr, "null pointer");
    const int64_t nS = nvir * (nvir + 1) / 2, nA = nvir * (nvir - 1) / 2, half = nblk * kpad * nvir;
    hipStream_t s = dcc::as_stream(stream);
    if (axes == 0) {
        if (vec4_ok(nvir, t3, sizeof(T)) && vec4_ok(nvir, Tp, sizeof(O)) && 2 * half < kIndex32)
            hipLaunchKernelGGL((pair_split_row4_kernel<T, O, uint32_t>), dim3(grid_for(half / 2)), dim3(kPThreads),
                               0, s, t3, (int)nvir, nS, nA, (uint32_t)kpad, (uint32_t)(half / 4), Tp);
        else if (2 * half < kIndex32)
            hipLaunchKernelGGL((pair_split_row_kernel<T, O, uint32_t>), dim3(grid_for(2 * half)), dim3(kPThreads), 0,
                               s, t3, (int)nvir, nS, nA, (uint32_t)kpad, (uint32_t)half, Tp);
        else
            hipLaunchKernelGGL((pair_split_row_kernel<T, O, int64_t>), dim3(grid_for(2 * half)), dim3(kPThreads), 0,
                               s, t3, (int)nvir, nS, nA, kpad, half, Tp);
    } else {
        int64_t sx, sz;
        plane_strides(axes, nvir, sx, sz);
        constexpr int PT = plane_tile<T>();
        const int ntile = (int)((nvir + PT - 1) / PT), ntp = ntile * (ntile + 1) / 2;
        const int64_t grid = nblk * nvir * (ntp + 1);
        DCC_CHECK(grid < ((int64_t)1 << 31), "grid exceeds 2^31 workgroups");
        hipLaunchKernelGGL((pair_split_plane_kernel<T, O, PT>), dim3((unsigned)grid), dim3(kPThreads), 0, s, t3,
                           (int)nvir, sx, sz, nS, nA, kpad, half, ntile, ntp, Tp);
    }
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int pair_unfold(const T *Up, int64_t nblk, int64_t nvir, int32_t axes, T *R, dcc_stream_t stream) {
    const int64_t mp = nvir * (nvir + 1) / 2;
    if (int rc = check_sizes(nblk, nvir, axes, mp)) return rc;
    if (nblk == 0 || nvir == 0) return 0;
    DCC_CHECK(Up != nullptr && R != nullptr, "null pointer");
    const int64_t half = nblk * mp * nvir;
    hipStream_t s = dcc::as_stream(stream);
    if (axes == 0) {
        const int64_t total = nblk * nvir * nvir * nvir;
        if (vec4_ok(nvir, Up, sizeof(T)) && vec4_ok(nvir, R, sizeof(T)) && total < kIndex32)
            hipLaunchKernelGGL((pair_unfold_row4_kernel<T, uint32_t>), dim3(grid_for(total / 4)), dim3(kPThreads), 0,
                               s, Up, (int)nvir, mp, half, R, (uint32_t)(total / 4));
        else if (total < kIndex32)
            hipLaunchKernelGGL((pair_unfold_row_kernel<T, uint32_t>), dim3(grid_for(total)), dim3(kPThreads), 0, s,
                               Up, (int)nvir, mp, half, R, (uint32_t)total);
        else
            hipLaunchKernelGGL((pair_unfold_row_kernel<T, int64_t>), dim3(grid_for(total)), dim3(kPThreads), 0, s,
                               Up, (int)nvir, mp, half, R, total);
    } else {
        int64_t sx, sz;
        plane_strides(axes, nvir, sx, sz);
        constexpr int PT = plane_tile<T>();
        const int ntile = (int)((nvir + PT - 1) / PT);
        const int64_t grid = nblk * nvir * ntile * ntile;
        DCC_CHECK(grid < ((int64_t)1 << 31), "grid exceeds 2^31 workgroups");
        hipLaunchKernelGGL((pair_unfold_plane_kernel<T, PT>), dim3((unsigned)grid), dim3(kPThreads), 0, s, Up,
                           (int)nvir, sx, sz, mp, half, ntile, R);
    }
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" {

int dcc_vvvv_pair_weights_f64(const double *W1, int64_t nvir, int64_t kpad, double *Wp, dcc_stream_t stream) {
    return pair_weights<double, double>(W1, nvir, kpad, Wp, stream);
}

int dcc_vvvv_pair_weights_f32(const float *W1, int64_t nvir, int64_t kpad, float *Wp, dcc_stream_t stream) {
    return pair_weights<float, float>(W1, nvir, kpad, Wp, stream);
}

int dcc_vvvv_pair_weights_packed_f32(const float *W1, int64_t nvir, int64_t kpad, uint32_t *Wp,
                                     dcc_stream_t stream) {
    return pair_weights<float, uint32_t>(W1, nvir, kpad, Wp, stream);
}

int dcc_vvvv_pair_split_f64(const double *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                            double *Tp, dcc_stream_t stream) {
    return pair_split<double, double>(t3, nblk, nvir, pair_axes, kpad, Tp, stream);
}

int dcc_vvvv_pair_split_f32(const float *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                            float *Tp, dcc_stream_t stream) {
    return pair_split<float, float>(t3, nblk, nvir, pair_axes, kpad, Tp, stream);
}

int dcc_vvvv_pair_split_packed_f32(const float *t3, int64_t nblk, int64_t nvir, int32_t pair_axes, int64_t kpad,
                                   uint32_t *Tp, dcc_stream_t stream) {
    return pair_split<float, uint32_t>(t3, nblk, nvir, pair_axes, kpad, Tp, stream);
}

int dcc_vvvv_pair_unfold_f64(const double *Up, int64_t nblk, int64_t nvir, int32_t pair_axes, double *R,
                             dcc_stream_t stream) {
    return pair_unfold<double>(Up, nblk, nvir, pair_axes, R, stream);
}

int dcc_vvvv_pair_unfold_f32(const float *Up, int64_t nblk, int64_t nvir, int32_t pair_axes, float *R,
                             dcc_stream_t stream) {
    return pair_unfold<float>(Up, nblk, nvir, pair_axes, R, stream);
}

}  // extern "C"
