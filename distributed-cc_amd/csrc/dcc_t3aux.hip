// Small O(n^2) builders that turn the T3 intermediates into GEMM operands.
#include "dcc_common.h"

namespace {

// Prefactor matrices of the W_oooo / F_oo terms.  For source canonical block p
// = (o00,o10,o20) and target canonical block q = (i,j,k), six coefficients
// pfs[s][p][q] multiply the source block transposed by permutation s
// (s = 0..5: abc, acb, bac, and the outputs qcab, qbca, qcba, see
// distr_cc/rccsdt.py:726-731).  The conditions are those of
// distr_cc/rccsdt.py:630-723: each ordered pair (l,m) of the sum over
// W_oooo[l,m,.,.] T[l,m,.] (and each l of F_oo) is reached exactly once.
template <typename T>
__global__ void oooo_prefactors_kernel(const T *__restrict__ W, const T *__restrict__ F, int no,
                                       const int *__restrict__ ijk, int n3, T *__restrict__ pfs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n3 * n3) return;
    const int p = (int)(t / n3), q = (int)(t % n3);
    const int a = ijk[3 * p], b = ijk[3 * p + 1], c = ijk[3 * p + 2];  // o00, o10, o20
    const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
    const int64_t n = no;
    auto w = [&](int l, int m, int x, int y) { return (double)W[((l * n + m) * n + x) * n + y]; };
    auto f = [&](int l, int x) { return (double)F[l * n + x]; };
    double s[6] = {0, 0, 0, 0, 0, 0};
    const bool ab = a != b, bc = b != c;
    // part 0: W_oooo[l,m,i,j] T[l,m,k]
    if (k == c) s[0] += w(a, b, i, j);
    if (bc && k == b) s[1] += w(a, c, i, j);
    if (ab && k == c) s[2] += w(b, a, i, j);
    if (bc && k == b) s[3] += w(c, a, i, j);
    if (ab && k == a) s[4] += w(b, c, i, j);
    if (ab && bc && k == a) s[5] += w(c, b, i, j);
    // part 1: W_oooo[l,m,i,k] T[l,m,j] (acb)
    if (j == c) s[1] += w(a, b, i, k);
    if (bc && j == b) s[0] += w(a, c, i, k);
    if (ab && j == c) s[4] += w(b, a, i, k);
    if (bc && j == b) s[5] += w(c, a, i, k);
    if (ab && j == a) s[2] += w(b, c, i, k);
    if (ab && bc && j == a) s[3] += w(c, b, i, k);
    // part 2: W_oooo[l,m,j,k] T[l,m,i] (bca)
    if (i == c) s[3] += w(a, b, j, k);
    if (bc && i == b) s[2] += w(a, c, j, k);
    if (ab && i == c) s[5] += w(b, a, j, k);
    if (bc && i == b) s[4] += w(c, a, j, k);
    if (ab && i == a) s[0] += w(b, c, j, k);
    if (ab && bc && i == a) s[1] += w(c, b, j, k);
    // part 3: -F_oo[l,i] T[l,j,k]
    if (j == a && k == b) s[3] -= f(c, i);
    if (bc && j == a && k == c) s[2] -= f(b, i);
    if (ab && j == b && k == c) s[0] -= f(a, i);
    // part 4: -F_oo[l,j] T[i,l,k]
    if (i == a && k == b) s[1] -= f(c, j);
    if (bc && i == a && k == c) s[0] -= f(b, j);
    if (ab && i == b && k == c) s[2] -= f(a, j);
    // part 5: -F_oo[l,k] T[i,j,l]
    if (i == a && j == b) s[0] -= f(c, k);
    if (bc && i == a && j == c) s[1] -= f(b, k);
    if (ab && i == b && j == c) s[4] -= f(a, k);
    const int64_t nn = (int64_t)n3 * n3;
#pragma unroll
    for (int u = 0; u < 6; ++u) pfs[u * nn + t] = (T)s[u];
}

// Weight matrices of the ring terms, all (o*v) x (o*v) with rows (r, x) and
// columns (l, d):
//   WR1[r,x,l,d]  = -W_ovov[l,x,r,d]
//   WR3a[r,x,l,d] =  W_ovvo[l,x,d,r]
//   WR3b[r,x,l,d] = -1/2 (W_ovov[l,x,r,d] + W_ovvo[l,x,d,r])
template <typename T>
__global__ void ring_weights_kernel(const T *__restrict__ Wovov, const T *__restrict__ Wovvo, int no, int nv,
                                    T *__restrict__ WR1, T *__restrict__ WR3a, T *__restrict__ WR3b) {
    const int64_t o = no, v = nv;
    const int64_t total = o * v * o * v;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t d = t % v, l = (t / v) % o, x = (t / (v * o)) % v, r = t / (v * o * v);
    const double ovov = Wovov[((l * v + x) * o + r) * v + d];
    const double ovvo = Wovvo[((l * v + x) * v + d) * o + r];
    WR1[t] = (T)(-ovov);
    WR3a[t] = (T)ovvo;
    WR3b[t] = (T)(-0.5 * (ovov + ovvo));
}

}  // namespace

namespace {

template <typename T>
int oooo_prefactors(const T *W_oooo, const T *F_oo, int64_t nocc, const int32_t *blk_ijk, int64_t nblk, T *pfs,
                    dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nblk >= 0, "negative extent");
    if (nblk == 0) return 0;
    DCC_CHECK(W_oooo && F_oo && blk_ijk && pfs, "null pointer");
    const int64_t total = nblk * nblk;
    hipLaunchKernelGGL(oooo_prefactors_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       dcc::as_stream(stream), W_oooo, F_oo, (int)nocc, blk_ijk, (int)nblk, pfs);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int ring_weights(const T *W_ovov, const T *W_ovvo, int64_t nocc, int64_t nvir, T *WR1, T *WR3a, T *WR3b,
                 dcc_stream_t stream) {
    DCC_CHECK(nocc >= 0 && nvir >= 0, "negative extent");
    const int64_t total = nocc * nvir * nocc * nvir;
    if (total == 0) return 0;
    DCC_CHECK(W_ovov && W_ovvo && WR1 && WR3a && WR3b, "null pointer");
    hipLaunchKernelGGL(ring_weights_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       dcc::as_stream(stream), W_ovov, W_ovvo, (int)nocc, (int)nvir, WR1, WR3a, WR3b);
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" {

int dcc_oooo_prefactors_f64(const double *W_oooo, const double *F_oo, int64_t nocc, const int32_t *blk_ijk,
                            int64_t nblk, double *pfs, dcc_stream_t stream) {
    return oooo_prefactors<double>(W_oooo, F_oo, nocc, blk_ijk, nblk, pfs, stream);
}

int dcc_oooo_prefactors_f32(const float *W_oooo, const float *F_oo, int64_t nocc, const int32_t *blk_ijk,
                            int64_t nblk, float *pfs, dcc_stream_t stream) {
    return oooo_prefactors<float>(W_oooo, F_oo, nocc, blk_ijk, nblk, pfs, stream);
}

int dcc_ring_weights_f64(const double *W_ovov, const double *W_ovvo, int64_t nocc, int64_t nvir, double *WR1,
                         double *WR3a, double *WR3b, dcc_stream_t stream) {
    return ring_weights<double>(W_ovov, W_ovvo, nocc, nvir, WR1, WR3a, WR3b, stream);
}

int dcc_ring_weights_f32(const float *W_ovov, const float *W_ovvo, int64_t nocc, int64_t nvir, float *WR1,
                         float *WR3a, float *WR3b, dcc_stream_t stream) {
    return ring_weights<float>(W_ovov, W_ovvo, nocc, nvir, WR1, WR3a, WR3b, stream);
}

}  // extern "C"
