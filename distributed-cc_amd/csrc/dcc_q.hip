// Kernels of the (Q)/[Q] energy (distr_cc/rccsdt_q.py:943-1177).
#include "dcc_common.h"

namespace {

// out[a,b,c,i,j,k] (a,b,c < vpad) = T[ijk][abc] for a,b,c < nvir, 0 in the padding.
// T3 is the canonical i<=j<=k block array; the ordered (i,j,k) maps onto its
// canonical block with the simultaneous permutation of (a,b,c).
__global__ void t3_to_abc_kernel(const double *__restrict__ t3, int no, int nv, int vp, double *__restrict__ out) {
    const int64_t o = no, v = nv, p = vp;
    const int64_t total = p * p * p * o * o * o;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int k = (int)(t % o), j = (int)((t / o) % o), i = (int)((t / (o * o)) % o);
    const int64_t abc = t / (o * o * o);
    const int c = (int)(abc % p), b = (int)((abc / p) % p), a = (int)(abc / (p * p));
    double val = 0.0;
    if (a < nv && b < nv && c < nv) {
        // sort (i,j,k) carrying (a,b,c) along
        int oi[3] = {i, j, k}, vi[3] = {a, b, c};
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int u = 0; u < 2 - s; ++u)
                if (oi[u] > oi[u + 1]) {
                    int x = oi[u]; oi[u] = oi[u + 1]; oi[u + 1] = x;
                    x = vi[u]; vi[u] = vi[u + 1]; vi[u + 1] = x;
                }
        const int64_t n = no;
        const int64_t ci = oi[0], cj = oi[1], ck = oi[2];
        const int64_t ni = n - ci;
        const int64_t lin = (n * (n + 1) * (n + 2) - ni * (ni + 1) * (ni + 2)) / 6 +
                            (cj - ci) * (2 * n - cj - ci + 1) / 2 + (ck - cj);
        val = t3[lin * v * v * v + ((int64_t)vi[0] * v + vi[1]) * v + vi[2]];
    }
    out[t] = val;
}

// Canonical index of a sorted triple x <= y <= z of n values, lexicographic (distribute_t3.py:51-60,
// the same closed form for occupied and virtual triples).
__device__ __forceinline__ int64_t tri_lin(int64_t n, int64_t x, int64_t y, int64_t z) {
    const int64_t m = n - x;
    return (n * (n + 1) * (n + 2) - m * (m + 1) * (m + 2)) / 6 + (y - x) * (2 * n - y - x + 1) / 2 + (z - y);
}

// Inverse of tri_lin (x <= y <= z < n) by counting.
__device__ __forceinline__ void tri_decode(int64_t n, int64_t lin, int &x, int &y, int &z) {
    x = 0;
    while (x + 1 < n && tri_lin(n, x + 1, x + 1, x + 1) <= lin) ++x;
    y = x;
    while (y + 1 < n && tri_lin(n, x, y + 1, y + 1) <= lin) ++y;
    z = (int)(lin - tri_lin(n, x, y, y)) + y;
}

// Sort three keys ascending, carrying three values along (ties keep their order).
__device__ __forceinline__ void sort3_carry(int k[3], int w[3]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 2 - s; ++u)
            if (k[u] > k[u + 1]) {
                int t = k[u]; k[u] = k[u + 1]; k[u + 1] = t;
                t = w[u]; w[u] = w[u + 1]; w[u + 1] = t;
            }
}

// IJK -> ABC: out[lin(a,b,c), i, j, k] = T[ijk][abc] for a <= b <= c and every ordered (i,j,k), from the
// canonical i <= j <= k blocks.  One workgroup per virtual triple; threads run over (i,j,k), so every
// output row of o^3 values is written contiguously.
__global__ __launch_bounds__(256) void t3_tri_to_abc_packed_kernel(const double *__restrict__ t3, int no, int nv,
                                                                   double *__restrict__ out) {
    __shared__ int sabc[3];
    const int64_t lin = blockIdx.x;
    if (threadIdx.x == 0) tri_decode(nv, lin, sabc[0], sabc[1], sabc[2]);
    __syncthreads();
    const int64_t o = no, v = nv, o3 = o * o * o;
    for (int64_t t = threadIdx.x; t < o3; t += blockDim.x) {
        int occ[3] = {(int)(t / (o * o)), (int)((t / o) % o), (int)(t % o)};
        int vir[3] = {sabc[0], sabc[1], sabc[2]};
        sort3_carry(occ, vir);
        out[lin * o3 + t] = t3[tri_lin(o, occ[0], occ[1], occ[2]) * v * v * v + ((int64_t)vir[0] * v + vir[1]) * v +
                               vir[2]];
    }
}

// ABC -> dense: out[a,b,c,i,j,k] (a,b,c < vpad) = T_{ijk}^{abc} for a,b,c < nvir (0 in the padding) from
// the ABC array t3_abc[lin(a<=b<=c), i, j, k]; (a,b,c) is sorted carrying (i,j,k) along.  One workgroup
// per (a,b,c) of the dense output.
__global__ __launch_bounds__(256) void t3_abc_packed_to_dense_kernel(const double *__restrict__ t3_abc, int no,
                                                                     int nv, int vp, double *__restrict__ out) {
    const int64_t abc = blockIdx.x;
    const int a = (int)(abc / ((int64_t)vp * vp)), b = (int)((abc / vp) % vp), c = (int)(abc % vp);
    const int64_t o = no, o3 = o * o * o;
    double *dst = out + abc * o3;
    if (a >= nv || b >= nv || c >= nv) {
        for (int64_t t = threadIdx.x; t < o3; t += blockDim.x) dst[t] = 0.0;
        return;
    }
    int key[3] = {a, b, c}, pos[3] = {0, 1, 2};
    sort3_carry(key, pos);
    const double *src = t3_abc + tri_lin(nv, key[0], key[1], key[2]) * o3;
    for (int64_t t = threadIdx.x; t < o3; t += blockDim.x) {
        const int y[3] = {(int)(t / (o * o)), (int)((t / o) % o), (int)(t % o)};
        dst[t] = src[((int64_t)y[pos[0]] * o + y[pos[1]]) * o + y[pos[2]]];
    }
}

// factor of the restricted a<=b<=c<=d sum (distr_cc/rccsdt_q.py:410-428)
__device__ __forceinline__ double q_factor(int a, int b, int c, int d) {
    if (a < b && b < c && c < d) return 24.0;
    if ((a == b && b < c && c < d) || (a < b && b == c && c < d) || (a < b && b < c && c == d)) return 12.0;
    if (a == b && b < c && c == d) return 6.0;
    return 0.0;
}

constexpr int kQThreads = 256;
constexpr int kQMaxO4 = 6561;  // nocc <= 9: two o^4 FP64 buffers in LDS

// Task block start letters and the (a,b,c,d) of workgroup blockIdx.x of a chunk of tasks.
struct QQuad {
    int a, b, c, d;
};

__device__ __forceinline__ QQuad q_quad(const int *__restrict__ tasks, int bs) {
    const int bs4 = bs * bs * bs * bs;
    const int64_t task = blockIdx.x / bs4;
    const int e = (int)(blockIdx.x % bs4);
    const int ia = e / (bs * bs * bs), ib = (e / (bs * bs)) % bs, ic = (e / bs) % bs, id = e % bs;
    return {tasks[4 * task] + ia, tasks[4 * task + 1] + ib, tasks[4 * task + 2] + ic, tasks[4 * task + 3] + id};
}

// z4_sq[i,j,k,l] of (a,b,c,d) from <ab|ij> and t2, both (vp, vp, o, o) (rccsdt_q.py:1143-1150).
__device__ __forceinline__ double q_zsq(const double *__restrict__ vvoo, const double *__restrict__ t2, int64_t v,
                                        int o, const QQuad &q, int i, int j, int k, int l) {
    auto VV = [&](int p, int r, int s, int u) { return vvoo[((p * v + r) * o + s) * o + u]; };
    auto T2 = [&](int p, int r, int s, int u) { return t2[((p * v + r) * o + s) * o + u]; };
    const int a = q.a, b = q.b, c = q.c, d = q.d;
    return VV(a, b, i, j) * T2(c, d, k, l) + VV(a, c, i, k) * T2(b, d, j, l) + VV(a, d, i, l) * T2(b, c, j, k) +
           VV(b, c, j, k) * T2(a, d, i, l) + VV(b, d, j, l) * T2(a, c, i, k) + VV(c, d, k, l) * T2(a, b, i, j);
}

// Workgroup sums of (paren, bracket) into partials[2 * blockIdx.x + {0, 1}].
__device__ __forceinline__ void q_store_partials(double paren, double bracket, double *__restrict__ partials) {
    __shared__ double red[2][kQThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        paren += __shfl_down(paren, off, 64);
        bracket += __shfl_down(bracket, off, 64);
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    if (lane == 0) {
        red[0][wave] = paren;
        red[1][wave] = bracket;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double p = 0.0, q = 0.0;
        for (int w = 0; w < kQThreads / 64; ++w) {
            p += red[0][w];
            q += red[1][w];
        }
        partials[2 * blockIdx.x] = p;
        partials[2 * blockIdx.x + 1] = q;
    }
}

// One workgroup per (task, a, b, c, d): t4 = P4_444[(t4 + z4_p) / D] * factor on
// the o^4 occupied block, then partial sums <t4, z4_p> and <t4, z4_sq> with
// z4_sq built on the fly from <ab|ij> (vvoo) and t2 (vvoo layout).  nocc <= 9:
// both intermediates of P4_444 live in LDS.
__global__ __launch_bounds__(kQThreads) void q_finalize_kernel(
    const double *__restrict__ t4, const double *__restrict__ z4, const int *__restrict__ tasks, int bs, int no,
    int nv, int vp, const double *__restrict__ e_occ, const double *__restrict__ e_vir,
    const double *__restrict__ vvoo, const double *__restrict__ t2, double *__restrict__ partials) {
    extern __shared__ double lds[];
    const int o = no, o2 = o * o, o3 = o2 * o, o4 = o3 * o;
    double *b0 = lds, *b1 = lds + o4;
    const QQuad q = q_quad(tasks, bs);
    double paren = 0.0, bracket = 0.0;
    const double fac = (q.a < nv && q.b < nv && q.c < nv && q.d < nv) ? q_factor(q.a, q.b, q.c, q.d) : 0.0;
    if (fac != 0.0) {  // uniform over the workgroup
        const int64_t base = blockIdx.x * (int64_t)o4;
        const double ev = e_vir[q.a] + e_vir[q.b] + e_vir[q.c] + e_vir[q.d];
        for (int x = threadIdx.x; x < o4; x += kQThreads) {
            const int i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            const double den = e_occ[i] + e_occ[j] + e_occ[k] + e_occ[l] - ev;
            const double num = t4[base + x] + z4[base + x];
            b0[x] = fabs(den) > 1e-15 ? num / den : 0.0;
        }
        __syncthreads();
        // (2 - P_ab - P_ac - P_ad) on the occupied positions
        for (int x = threadIdx.x; x < o4; x += kQThreads) {
            const int i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            b1[x] = 2.0 * b0[x] - b0[((j * o + i) * o + k) * o + l] - b0[((k * o + j) * o + i) * o + l] -
                    b0[((l * o + j) * o + k) * o + i];
        }
        __syncthreads();
        // (2 - P_bc - P_bd)
        for (int x = threadIdx.x; x < o4; x += kQThreads) {
            const int i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            b0[x] = 2.0 * b1[x] - b1[((i * o + k) * o + j) * o + l] - b1[((i * o + l) * o + k) * o + j];
        }
        __syncthreads();
        for (int x = threadIdx.x; x < o4; x += kQThreads) {
            const int i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            // (2 - P_cd), then the factor
            const double t = fac * (2.0 * b0[x] - b0[((i * o + j) * o + l) * o + k]);
            paren += t * z4[base + x];
            bracket += t * q_zsq(vvoo, t2, vp, o, q, i, j, k, l);
        }
    }
    q_store_partials(paren, bracket, partials);
}

// The same update for any nocc, with one o^4 intermediate per workgroup in the global workspace ws
// (o^4 doubles at ws + blockIdx.x * o^4).  The denominator is invariant under permutations of
// (i,j,k,l), so b1 = (2 - P_ab - P_ac - P_ad)(t4 + z4) / D is formed in one pass from t4 and z4 at
// the four positions; (2 - P_cd)(2 - P_bc - P_bd) b1 expands to
//   4 b1[ijkl] - 2 b1[ikjl] - 2 b1[ilkj] - 2 b1[ijlk] + b1[iljk] + b1[iklj].
__global__ __launch_bounds__(kQThreads) void q_finalize_ws_kernel(
    const double *__restrict__ t4, const double *__restrict__ z4, const int *__restrict__ tasks, int bs, int no,
    int nv, int vp, const double *__restrict__ e_occ, const double *__restrict__ e_vir,
    const double *__restrict__ vvoo, const double *__restrict__ t2, double *__restrict__ ws,
    double *__restrict__ partials) {
    const int64_t o = no, o2 = o * o, o3 = o2 * o, o4 = o3 * o;
    const QQuad q = q_quad(tasks, bs);
    double paren = 0.0, bracket = 0.0;
    const double fac = (q.a < nv && q.b < nv && q.c < nv && q.d < nv) ? q_factor(q.a, q.b, q.c, q.d) : 0.0;
    if (fac != 0.0) {  // uniform over the workgroup
        const int64_t base = blockIdx.x * o4;
        const double *s4 = t4 + base, *z = z4 + base;
        double *b1 = ws + base;
        const double ev = e_vir[q.a] + e_vir[q.b] + e_vir[q.c] + e_vir[q.d];
        auto S = [&](int64_t y) { return s4[y] + z[y]; };
        for (int64_t x = threadIdx.x; x < o4; x += kQThreads) {
            const int64_t i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            const double den = e_occ[i] + e_occ[j] + e_occ[k] + e_occ[l] - ev;
            const double num = 2.0 * S(x) - S(((j * o + i) * o + k) * o + l) - S(((k * o + j) * o + i) * o + l) -
                               S(((l * o + j) * o + k) * o + i);
            b1[x] = fabs(den) > 1e-15 ? num / den : 0.0;
        }
        __syncthreads();
        for (int64_t x = threadIdx.x; x < o4; x += kQThreads) {
            const int64_t i = x / o3, j = (x / o2) % o, k = (x / o) % o, l = x % o;
            const double t = fac * (4.0 * b1[x] - 2.0 * b1[((i * o + k) * o + j) * o + l] -
                                    2.0 * b1[((i * o + l) * o + k) * o + j] - 2.0 * b1[((i * o + j) * o + l) * o + k] +
                                    b1[((i * o + l) * o + j) * o + k] + b1[((i * o + k) * o + l) * o + j]);
            paren += t * z[x];
            bracket += t * q_zsq(vvoo, t2, vp, (int)o, q, (int)i, (int)j, (int)k, (int)l);
        }
    }
    q_store_partials(paren, bracket, partials);
}

}  // namespace

extern "C" {

int dcc_t3_to_abc_f64(const double *t3_tri, int64_t nocc, int64_t nvir, int64_t vpad, double *out,
                      dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nvir >= 0 && vpad >= nvir, "bad extents");
    const int64_t total = vpad * vpad * vpad * nocc * nocc * nocc;
    if (total == 0) return 0;
    DCC_CHECK(t3_tri != nullptr && out != nullptr, "null pointer");
    DCC_CHECK((total + 255) / 256 < ((int64_t)1 << 31), "grid too large");
    hipLaunchKernelGGL(t3_to_abc_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       dcc::as_stream(stream), t3_tri, (int)nocc, (int)nvir, (int)vpad, out);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_t3_tri_to_abc_packed_f64(const double *t3_tri, int64_t nocc, int64_t nvir, double *t3_abc,
                                 dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nvir >= 0, "bad extents");
    const int64_t nabc = nvir * (nvir + 1) * (nvir + 2) / 6;
    if (nabc == 0 || nocc == 0) return 0;
    DCC_CHECK(t3_tri != nullptr && t3_abc != nullptr, "null pointer");
    DCC_CHECK(nabc < ((int64_t)1 << 31), "grid too large");
    hipLaunchKernelGGL(t3_tri_to_abc_packed_kernel, dim3((unsigned)nabc), dim3(256), 0, dcc::as_stream(stream),
                       t3_tri, (int)nocc, (int)nvir, t3_abc);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_t3_abc_packed_to_dense_f64(const double *t3_abc, int64_t nocc, int64_t nvir, int64_t vpad, double *out,
                                   dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nvir >= 0 && vpad >= nvir, "bad extents");
    const int64_t n = vpad * vpad * vpad;
    if (n == 0 || nocc == 0) return 0;
    DCC_CHECK(t3_abc != nullptr && out != nullptr, "null pointer");
    DCC_CHECK(n < ((int64_t)1 << 31), "grid too large");
    hipLaunchKernelGGL(t3_abc_packed_to_dense_kernel, dim3((unsigned)n), dim3(256), 0, dcc::as_stream(stream),
                       t3_abc, (int)nocc, (int)nvir, (int)vpad, out);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_q_finalize_f64(const double *t4, const double *z4, const int32_t *tasks, int64_t ntask, int64_t blksize,
                       int64_t nocc, int64_t nvir, int64_t vpad, const double *e_occ, const double *e_vir,
                       const double *vvoo, const double *t2, double *ws, double *partials, dcc_stream_t stream) {
    DCC_CHECK(ntask >= 0 && blksize > 0 && nocc > 0 && nvir > 0 && vpad >= nvir, "bad extents");
    const int64_t o4 = nocc * nocc * nocc * nocc;
    DCC_CHECK(ws != nullptr || o4 <= kQMaxO4, "nocc > 9 needs the workspace ws (ntask * blksize^4 * nocc^4 doubles)");
    const int64_t nwg = ntask * blksize * blksize * blksize * blksize;
    if (nwg == 0) return 0;
    DCC_CHECK(t4 && z4 && tasks && e_occ && e_vir && vvoo && t2 && partials, "null pointer");
    DCC_CHECK(nwg < ((int64_t)1 << 31), "grid too large");
    if (ws == nullptr) {
        hipLaunchKernelGGL(q_finalize_kernel, dim3((unsigned)nwg), dim3(kQThreads), (size_t)(2 * o4 * sizeof(double)),
                           dcc::as_stream(stream), t4, z4, tasks, (int)blksize, (int)nocc, (int)nvir, (int)vpad,
                           e_occ, e_vir, vvoo, t2, partials);
    } else {
        hipLaunchKernelGGL(q_finalize_ws_kernel, dim3((unsigned)nwg), dim3(kQThreads), 0, dcc::as_stream(stream), t4,
                           z4, tasks, (int)blksize, (int)nocc, (int)nvir, (int)vpad, e_occ, e_vir, vvoo, t2, ws,
                           partials);
    }
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
