// HBM-bound transforms of v^3 blocks under the symmetric group S3 of (a,b,c).
//
// Every permutation-type operation on a T3 block -- taking a permuted block, spin summation, the
// P3_full projection -- reads element x of the output from elements p(x) of the input for the six
// axis permutations p.  A workgroup computes one 16x16x16 output tile.  For every permutation p with
// a nonzero coefficient it reads the source tile at the p-permuted tile coordinates, along that
// tile's contiguous axis (rows of 16 elements: whole 128-byte lines in FP64), into LDS; the permuted
// accesses are LDS accesses, and the next source tile is loaded while the current one is read.  The
// output tile accumulates in registers and is written once along its contiguous axis.  The six tiles
// of a tile orbit (a tile and its five permuted tiles) run as consecutive workgroups on one XCD, so
// a source block that serves all six permutations is read from HBM once and re-read from L2.
#include "dcc_common.h"
#include <mutex>
#include <type_traits>
#include "dcc_gemm_tile.h"   // xcd_remap

namespace {

constexpr int E = 16;                  // tile edge
constexpr int TE = E * E * E;          // elements per tile
constexpr int kThreads = 512;
constexpr int kPer = TE / kThreads;    // elements per thread (8)
// LDS strides (elements) of a staged tile, x slowest: rows padded to 17 and planes to 275 elements,
// so that the lanes of a wave reading the tile along any of its three axes spread over the banks (in
// an unpadded 16^3 tile a read along the slowest axis puts every lane of a row on one bank)
constexpr int kSY = E + 1, kSX = E * kSY + 3, kLdsTile = E * kSX;
// workgroups per 2x2x2 block of tile orbits: 8 orbit slots x 6 permuted tiles
constexpr int kRun = 48;

// The six permutations in the reference's tp_t3 order (src/rccsdt.c:32-39);
// X.transpose(p)[x] = X[y] with y[p[n]] = x[n].
__constant__ int kP[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Tile coordinates (x slowest).  Three named values rather than an array: picking one by a run-time
// axis index is then a select of values, not an indexed access that would place the array in
// scratch memory.
struct I3 {
    int x, y, z;
};
__device__ __forceinline__ int pick3(int i, int a, int b, int c) { return i == 0 ? a : (i == 1 ? b : c); }

// Orbit o -> tile triple (A >= B >= C) with o = A(A+1)(A+2)/6 + B(B+1)/2 + C.
__device__ __forceinline__ void decode_orbit(int o, int &A, int &B, int &C) {
    A = 0;
    while ((A + 1) * (A + 2) * (A + 3) / 6 <= o) ++A;
    int r = o - A * (A + 1) * (A + 2) / 6;
    B = 0;
    while ((B + 1) * (B + 2) / 2 <= r) ++B;
    C = r - B * (B + 1) / 2;
}

__host__ __device__ __forceinline__ int64_t n_orbits(int64_t nt) { return nt * (nt + 1) * (nt + 2) / 6; }

// Workgroup b of a launch over jobs of nt^3 tiles -> (job, output tile coordinates Tc).  Workgroups
// come in runs of 48 per 2x2x2 block of tile orbits (nso such blocks per job): 8 orbit slots (A >= B
// >= C and A < nt, else exit) times the six permuted tiles s of the orbit, tile s having coordinates
// (R[p_s[0]], R[p_s[1]], R[p_s[2]]) (a tile that repeats an earlier one of its orbit exits).
// Consecutive workgroups run on one XCD (xcd_remap): the six tiles of an orbit read the same source
// tiles while they are in that XCD's L2, and in FP32 (rows of 64 bytes, half a line) the tiles of a
// 2x2x2 block read and write the two halves of their lines together.
__device__ __forceinline__ bool tile_of(int64_t b, int64_t total, int nt, int nso, int64_t &job, I3 &Tc) {
    const int64_t L = dcc_gemm::xcd_remap(b, total);
    job = L / (kRun * (int64_t)nso);
    const int r = (int)(L - job * kRun * (int64_t)nso);
    const int slot = (r / 6) & 7, s = r % 6;
    int SA, SB, SC;
    decode_orbit(r / kRun, SA, SB, SC);
    const int A = 2 * SA + ((slot >> 2) & 1), B = 2 * SB + ((slot >> 1) & 1), C = 2 * SC + (slot & 1);
    if (!(A >= B && B >= C && A < nt)) return false;
    Tc = I3{pick3(kP[s][0], A, B, C), pick3(kP[s][1], A, B, C), pick3(kP[s][2], A, B, C)};
    for (int u = 0; u < s; ++u)
        if (pick3(kP[u][0], A, B, C) == Tc.x && pick3(kP[u][1], A, B, C) == Tc.y && pick3(kP[u][2], A, B, C) == Tc.z)
            return false;
    return true;
}

// The elements of a tile a thread owns (tile-local x slowest, z contiguous).  PAIR (v even, blocks
// aligned to two elements): z = 2 (t % 8) + h, h < 2, of row y = (t / 8) % 16 in the planes
// x = t / 128 + 4u, u < 4; each pair is moved with one vector access, and a wave covers 8 whole rows.
// Scalar: z = t % 16, y = (t / 16) % 16, x = t / 256 + 2k, k < 8.
template <bool PAIR> struct Own {
    int x0, y, z0;
    __device__ __forceinline__ Own() {
        const int t = threadIdx.x;
        x0 = PAIR ? t >> 7 : t >> 8;
        y = PAIR ? (t >> 3) & (E - 1) : (t >> 4) & (E - 1);
        z0 = PAIR ? 2 * (t & 7) : t & (E - 1);
    }
    static constexpr int dx(int k) { return PAIR ? 4 * (k >> 1) : 2 * k; }
    static constexpr int dz(int k) { return PAIR ? (k & 1) : 0; }
};

template <typename T> struct Vec2 { typedef T t __attribute__((ext_vector_type(2))); };

// This thread's elements of the tile at tile coordinates Tc of the block at blk, zero outside the
// block.  Elements outside read the block's first element instead, so that every load is issued
// unconditionally and all of them are in flight together.
template <typename T, bool PAIR>
__device__ __forceinline__ void load_tile(const T *__restrict__ blk, int nvir, const I3 Tc, const Own<PAIR> &w,
                                          T v[kPer]) {
    const int gx0 = Tc.x * E + w.x0, gy = Tc.y * E + w.y, gz = Tc.z * E + w.z0;
    const bool yz = gy < nvir && gz < nvir;   // PAIR: gz even and nvir even, so gz + 1 < nvir too
    const int64_t row = (int64_t)gy * nvir + gz, plane = (int64_t)nvir * nvir;
    if constexpr (PAIR) {
#pragma unroll
        for (int u = 0; u < kPer / 2; ++u) {
            const int gx = gx0 + Own<PAIR>::dx(2 * u);
            const bool in = yz && gx < nvir;
            const typename Vec2<T>::t x =
                *reinterpret_cast<const typename Vec2<T>::t *>(blk + (in ? gx * plane + row : 0));
            v[2 * u] = in ? x[0] : T(0);
            v[2 * u + 1] = in ? x[1] : T(0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int gx = gx0 + Own<PAIR>::dx(k);
            const bool in = yz && gx < nvir;
            const T x = blk[in ? gx * plane + row : 0];
            v[k] = in ? x : T(0);
        }
    }
}

// Store this thread's elements of tile Tc of the block at blk (elements outside the block skipped).
template <typename O, bool PAIR>
__device__ __forceinline__ void store_tile(O *blk, int nvir, const I3 Tc, const Own<PAIR> &w, const O v[kPer]) {
    const int gx0 = Tc.x * E + w.x0, gy = Tc.y * E + w.y, gz = Tc.z * E + w.z0;
    if (!(gy < nvir && gz < nvir)) return;
    const int64_t row = (int64_t)gy * nvir + gz, plane = (int64_t)nvir * nvir;
    if constexpr (PAIR) {
#pragma unroll
        for (int u = 0; u < kPer / 2; ++u) {
            const int gx = gx0 + Own<PAIR>::dx(2 * u);
            if (gx < nvir)
                *reinterpret_cast<typename Vec2<O>::t *>(blk + gx * plane + row) =
                    typename Vec2<O>::t{v[2 * u], v[2 * u + 1]};
        }
    } else {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int gx = gx0 + Own<PAIR>::dx(k);
            if (gx < nvir) blk[gx * plane + row] = v[k];
        }
    }
}

// Write this thread's elements of a source tile into the LDS tile (its own layout).
template <typename T, bool PAIR>
__device__ __forceinline__ void stage(T *lds, const Own<PAIR> &w, const T v[kPer]) {
    const int base = w.x0 * kSX + w.y * kSY + w.z0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) lds[base + Own<PAIR>::dx(k) * kSX + Own<PAIR>::dz(k)] = v[k];
}

__device__ __forceinline__ int lds_stride(int axis) { return axis == 0 ? kSX : (axis == 1 ? kSY : 1); }

// This thread's output elements from the staged source tile through permutation p: output coordinate
// n is the staged tile's coordinate along its axis p[n].
template <typename T, bool PAIR>
__device__ __forceinline__ void read_permuted(const T *lds, const Own<PAIR> &w, int p, double x[kPer]) {
    const int s0 = lds_stride(kP[p][0]), s1 = lds_stride(kP[p][1]), s2 = lds_stride(kP[p][2]);
    const int base = w.x0 * s0 + w.y * s1 + w.z0 * s2;
#pragma unroll
    for (int k = 0; k < kPer; ++k) x[k] = (double)lds[base + Own<PAIR>::dx(k) * s0 + Own<PAIR>::dz(k) * s2];
}

// Tile coordinates of the source tile that output tile Tc reads through permutation p (source axis
// p[n] carries output tile coordinate n).
__device__ __forceinline__ I3 source_tile(const I3 Tc, int p) {
    const int p0 = kP[p][0], p1 = kP[p][1];
    return I3{p0 == 0 ? Tc.x : (p1 == 0 ? Tc.y : Tc.z), p0 == 1 ? Tc.x : (p1 == 1 ? Tc.y : Tc.z),
              p0 == 2 ? Tc.x : (p1 == 2 ? Tc.y : Tc.z)};
}

// acc[k] += coef[p] * (block src_of(p)).transpose(p) at this thread's elements of output tile Tc,
// over the permutations p with coef[p] != 0 in ascending order (coef and src_of uniform over the
// workgroup).  Source tiles pass through the one LDS tile; the next is loaded while the current one
// is read.
template <typename T, bool PAIR, typename SrcOf>
__device__ __forceinline__ void accumulate(T *lds, int nvir, const I3 Tc, const Own<PAIR> &w, const double *coef,
                                           SrcOf src_of, double acc[kPer]) {
    int p = 0;
    while (p < 6 && coef[p] == 0.0) ++p;
    T nxt[kPer];
    if (p < 6) load_tile<T, PAIR>(src_of(p), nvir, source_tile(Tc, p), w, nxt);
#pragma unroll 1
    while (p < 6) {
        int pn = p + 1;
        while (pn < 6 && coef[pn] == 0.0) ++pn;
        __syncthreads();   // every wave has finished reading the previous source tile
        stage<T, PAIR>(lds, w, nxt);
        __syncthreads();
        if (pn < 6) load_tile<T, PAIR>(src_of(pn), nvir, source_tile(Tc, pn), w, nxt);
        double x[kPer];
        read_permuted<T, PAIR>(lds, w, p, x);
        const double c = coef[p];
#pragma unroll
        for (int k = 0; k < kPer; ++k) acc[k] += c * x[k];
        p = pn;
    }
}

// O: the output element type -- T, or uint32_t for FP32 input written as split-BF16x3 words
// (dcc::pack_bf16x3 of the float result; beta must then be 0).
template <typename T, typename O, bool PAIR>
__device__ __forceinline__ void s3_transform_tile(T *lds, O *__restrict__ dst, const T *__restrict__ src, int nvir,
                                                  double beta, const dcc_s3_job *jb, const I3 Tc) {
    constexpr bool same = std::is_same<O, T>::value;
    const Own<PAIR> w;
    const bool rmw = same && beta != 0.0;
    T old[kPer] = {};
    if constexpr (same) {
        if (rmw) load_tile<T, PAIR>(dst, nvir, Tc, w, old);
    }
    double acc[kPer] = {};
    accumulate<T, PAIR>(lds, nvir, Tc, w, jb->coef, [&](int) { return src; }, acc);
    O res[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        double val = acc[k];
        if constexpr (same) {
            if (rmw) val += beta * (double)old[k];
            res[k] = (T)val;
        } else {
            res[k] = dcc::pack_bf16x3((float)val);
        }
    }
    store_tile<O, PAIR>(dst, nvir, Tc, w, res);
}

template <typename T, typename O = T, bool PAIR = false>
__global__ __launch_bounds__(kThreads) void s3_transform_kernel(O *__restrict__ out, const T *__restrict__ src,
                                                                int nvir, double beta,
                                                                const dcc_s3_job *__restrict__ jobs, int nt,
                                                                int nso) {
    __shared__ T lds[kLdsTile];
    int64_t job;
    I3 Tc;
    if (!tile_of(blockIdx.x, gridDim.x, nt, nso, job, Tc)) return;
    const dcc_s3_job *jb = jobs + job;
    // the paired form needs both blocks to start on an even element (uniform over the workgroup)
    if (PAIR && ((jb->dst | jb->src) & 1) == 0)
        s3_transform_tile<T, O, true>(lds, out + jb->dst, src + jb->src, nvir, beta, jb, Tc);
    else
        s3_transform_tile<T, O, false>(lds, out + jb->dst, src + jb->src, nvir, beta, jb, Tc);
}

// Multi-source S3 transform: out[x] = beta*out[x] + sum_p c[p] * src_p.transpose(p)[x], with its own
// source block src_p per permutation (for example the six permuted base blocks X[pi(ijk)] of a
// symmetrised term); the output tile is read and written once.
template <typename T, bool PAIR>
__device__ __forceinline__ void s3_gather_tile(T *lds, T *__restrict__ out, const T *__restrict__ src, int nvir,
                                               double beta, const dcc_s3_gather_job *jb, const I3 Tc) {
    const Own<PAIR> w;
    T *dst = out + jb->dst;
    T old[kPer] = {};
    if (beta != 0.0) load_tile<T, PAIR>(dst, nvir, Tc, w, old);
    double acc[kPer] = {};
    accumulate<T, PAIR>(lds, nvir, Tc, w, jb->coef, [&](int p) { return src + jb->src[p]; }, acc);
    T res[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        double val = acc[k];
        if (beta != 0.0) val += beta * (double)old[k];
        res[k] = (T)val;
    }
    store_tile<T, PAIR>(dst, nvir, Tc, w, res);
}

template <typename T, bool PAIR = false>
__global__ __launch_bounds__(kThreads) void s3_gather_kernel(T *__restrict__ out, const T *__restrict__ src,
                                                             int nvir, double beta,
                                                             const dcc_s3_gather_job *__restrict__ jobs,
                                                             int nt, int nso) {
    __shared__ T lds[kLdsTile];
    int64_t job;
    I3 Tc;
    if (!tile_of(blockIdx.x, gridDim.x, nt, nso, job, Tc)) return;
    const dcc_s3_gather_job *jb = jobs + job;
    int64_t odd = jb->dst;
    for (int p = 0; p < 6; ++p) odd |= jb->coef[p] != 0.0 ? jb->src[p] : 0;
    if (PAIR && (odd & 1) == 0) s3_gather_tile<T, true>(lds, out, src, nvir, beta, jb, Tc);
    else s3_gather_tile<T, false>(lds, out, src, nvir, beta, jb, Tc);
}

// The fused projection/update keeps the 8x8x8 tile orbits: one workgroup reads the six tiles of an
// orbit of R from HBM once and writes the six output tiles.  In the per-output-tile form above the
// six tiles of an orbit each re-read all six source tiles through L2, one after another, and the
// projection ran 1.3x longer (nocc=10, nvir=90, FP64) and 1.2x longer (nocc=16, nvir=160, FP32).
namespace orbit8 {

constexpr int E = 8;            // tile edge
constexpr int TE = E * E * E;   // elements per tile
constexpr int kThreads = 256;
constexpr int kPer = TE / kThreads;  // elements per thread per tile

__device__ __forceinline__ int perm_index(int a, int b, int c) {
    // index in kP of the permutation (a,b,c)
    return a == 0 ? (b == 1 ? 0 : 1) : (a == 1 ? (b == 0 ? 2 : 3) : (b == 0 ? 4 : 5));
}


// Workgroup b of a launch over njobs blocks -> (job, orbit index, tile triple R).  A tile row is 64
// bytes of doubles (32 of floats), so a 128-byte line is shared with tiles whose coordinate along the
// contiguous axis differs in the low bits; those belong to orbits that differ from this one in the
// same bits of one tile coordinate.  Workgroups therefore walk 2x2x2 blocks of tile coordinates (nso
// block orbits of 8 slots; slots with A >= B >= C violated or A >= nt exit at once), and consecutive
// slots run on one XCD (xcd_remap), so the parts of a line are read while it is in that XCD's L2.
// In plain orbit order the FP32 placement passes ran 1.3x longer at nocc=16, nvir=160 (4x4x4 blocks
// for floats measured no better than 2x2x2).
__device__ __forceinline__ bool orbit_of(int64_t b, int64_t total, int nt, int nso, int64_t &job, int &orbit,
                                         int R[3]) {
    const int64_t L = dcc_gemm::xcd_remap(b, total);
    job = L / (8 * (int64_t)nso);
    const int r = (int)(L - job * 8 * (int64_t)nso);
    int SA, SB, SC;
    decode_orbit(r >> 3, SA, SB, SC);
    const int A = 2 * SA + ((r >> 2) & 1), B = 2 * SB + ((r >> 1) & 1), C = 2 * SC + (r & 1);
    if (!(A >= B && B >= C && A < nt)) return false;
    R[0] = A;
    R[1] = B;
    R[2] = C;
    orbit = A * (A + 1) * (A + 2) / 6 + B * (B + 1) / 2 + C;
    return true;
}

// Two element mappings of a tile to the 256 threads.  Scalar: thread t holds elements t and t + 256
// (one element per load).  PAIR (v even, operands 2-element aligned): thread t holds elements 2t and
// 2t + 1, adjacent along the contiguous axis, moved with one vector load or store (16 bytes for doubles,
// 8 for floats); with v even a pair lies entirely inside or outside the block.
template <bool PAIR> __device__ __forceinline__ int elem_of(int h) {
    return PAIR ? 2 * (int)threadIdx.x + h : (int)threadIdx.x + h * kThreads;
}

// Local element e of orbit tile s: block coordinates g[3], offset in the v^3 block and whether it lies
// inside the block.  Elements outside (partial edge tiles) get offset 0, the block's first element, so
// that loads can be issued unconditionally: all of a thread's loads are then in flight together (with
// a branch per element the compiler waited for each load before the next).
__device__ __forceinline__ bool tile_elem(int nvir, const int R[3], int s, int e, int g[3], int64_t &off) {
    g[0] = R[kP[s][0]] * E + (e >> 6);
    g[1] = R[kP[s][1]] * E + ((e >> 3) & 7);
    g[2] = R[kP[s][2]] * E + (e & 7);
    const bool in = g[0] < nvir && g[1] < nvir && g[2] < nvir;
    off = in ? ((int64_t)g[0] * nvir + g[1]) * nvir + g[2] : 0;
    return in;
}

// This thread's two elements of orbit tile s of a block, loaded together (zero outside the block).
template <typename T, bool PAIR>
__device__ __forceinline__ void load2(const T *__restrict__ src, int nvir, const int R[3], int s, T out[2]) {
    if constexpr (PAIR) {
        int g[3];
        int64_t off;
        const bool in = tile_elem(nvir, R, s, elem_of<true>(0), g, off);
        const typename Vec2<T>::t x = *reinterpret_cast<const typename Vec2<T>::t *>(src + off);
        out[0] = in ? x[0] : T(0);
        out[1] = in ? x[1] : T(0);
    } else {
#pragma unroll
        for (int h = 0; h < kPer; ++h) {
            int g[3];
            int64_t off;
            const bool in = tile_elem(nvir, R, s, elem_of<false>(h), g, off);
            const T x = src[off];
            out[h] = in ? x : T(0);
        }
    }
}

// Load the six permuted tiles of orbit (R0,R1,R2) of block `src` into LDS: all loads first, then
// the LDS stores.
template <typename T, bool PAIR>
__device__ __forceinline__ void load_orbit(T (*tiles)[TE], const T *__restrict__ src, int nvir, const int R[3]) {
    T val[6][kPer];
#pragma unroll
    for (int s = 0; s < 6; ++s) load2<T, PAIR>(src, nvir, R, s, val[s]);
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int h = 0; h < kPer; ++h) tiles[s][elem_of<PAIR>(h)] = val[s][h];
}

// Is output tile s a repeat of an earlier tile of the same orbit?
__device__ __forceinline__ bool duplicate_tile(int s, const int R[3]) {
    const int a = R[kP[s][0]], b = R[kP[s][1]], c = R[kP[s][2]];
    for (int u = 0; u < s; ++u)
        if (R[kP[u][0]] == a && R[kP[u][1]] == b && R[kP[u][2]] == c) return true;
    return false;
}

// This thread's current values of the six output tiles of orbit R in block dst (loads issued
// together; elements outside the block read the block's first element and are never stored).
template <typename T, bool PAIR>
__device__ __forceinline__ void prefetch_out(const T *dst, int nvir, const int R[3], T old[6][kPer]) {
#pragma unroll
    for (int s = 0; s < 6; ++s) load2<T, PAIR>(dst, nvir, R, s, old[s]);
}

// Store this thread's two values of output tile s (elements outside the block are skipped).
template <typename T, bool PAIR>
__device__ __forceinline__ void store2(T *dst, int nvir, const int R[3], int s, const T val[2]) {
    if constexpr (PAIR) {
        int g[3];
        int64_t off;
        if (tile_elem(nvir, R, s, elem_of<true>(0), g, off))
            *reinterpret_cast<typename Vec2<T>::t *>(dst + off) = typename Vec2<T>::t{val[0], val[1]};
    } else {
#pragma unroll
        for (int h = 0; h < kPer; ++h) {
            int g[3];
            int64_t off;
            if (tile_elem(nvir, R, s, elem_of<false>(h), g, off)) dst[off] = val[h];
        }
    }
}

// Value of X.transpose(p)[x] for x = local element e of output tile s.
// y[p[n]] = x[n]; y lives in tile s' with ranges R[s'[m]] = R[s[pinv[m]]].
template <typename T>
__device__ __forceinline__ double permuted(const T (*tiles)[TE], int s, int p, int e) {
    const int x[3] = {e >> 6, (e >> 3) & 7, e & 7};
    int y[3], sy[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        y[kP[p][n]] = x[n];
        sy[kP[p][n]] = kP[s][n];
    }
    const int sp = perm_index(sy[0], sy[1], sy[2]);
    return (double)tiles[sp][(y[0] << 6) | (y[1] << 3) | y[2]];
}

// Fused P3_full projection + denominator + t3 update + norms (one tile orbit).
// Storage in T, arithmetic (projection, division, update, norms) in FP64.  PAIR: v even and the
// three arrays aligned to two elements (checked on the host).
template <typename T, bool PAIR = false>
__global__ __launch_bounds__(kThreads) void r3_finalize_kernel(const T *__restrict__ R, T *__restrict__ t3,
                                                               T *__restrict__ r3_out, int nvir,
                                                               const int *__restrict__ blk_ijk,
                                                               const double *__restrict__ eia, int nvir_diis,
                                                               double *__restrict__ partials, int nt, int nso) {
    __shared__ T tiles[6][TE];
    __shared__ double red[2][kThreads / 64];
    __shared__ double sden[3][3][E];
    int64_t blk;
    int orbit, Rg[3];
    if (!orbit_of(blockIdx.x, gridDim.x, nt, nso, blk, orbit, Rg)) return;
    const int64_t norb = n_orbits(nt);
    const int64_t v = nvir, v2 = v * v, v3 = v2 * v;
    const T *Rb = R + blk * v3;
    T *tb = t3 + blk * v3;
    const int oi = blk_ijk[3 * blk], oj = blk_ijk[3 * blk + 1], ok = blk_ijk[3 * blk + 2];
    const double *ei = eia + (int64_t)oi * v, *ej = eia + (int64_t)oj * v, *ek = eia + (int64_t)ok * v;
    // the orbital energies of the three tile ranges of ei, ej, ek (3 x 3 x 8 values, index clamped at
    // the block edge) in LDS; this thread's t3 elements, loaded together with the R orbit
    const int w = threadIdx.x / (3 * E), c = (threadIdx.x / E) % 3, i = threadIdx.x % E;
    const bool has_e = threadIdx.x < 9 * E;
    const double *ew = w == 0 ? ei : (w == 1 ? ej : ek);
    const double eval = ew[has_e ? min(Rg[c] * E + i, nvir - 1) : 0];
    T old[6][kPer];
    prefetch_out<T, PAIR>(tb, nvir, Rg, old);
    load_orbit<T, PAIR>(tiles, Rb, nvir, Rg);
    if (has_e) sden[w][c][i] = eval;
    __syncthreads();
    double nrm = 0.0, nact = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (duplicate_tile(s, Rg)) continue;
        T tnew[kPer], rout[kPer];
        bool any_in = false;
#pragma unroll
        for (int h = 0; h < kPer; ++h) {
            const int e = elem_of<PAIR>(h);
            int g[3];
            int64_t off;
            const bool in = tile_elem(nvir, Rg, s, e, g, off);
            any_in |= in;
            double sym = 0.0;
#pragma unroll
            for (int p = 0; p < 6; ++p) sym += permuted(tiles, s, p, e);
            const double own = permuted(tiles, s, 0, e);
            const double den = sden[0][kP[s][0]][e >> 6] + sden[1][kP[s][1]][(e >> 3) & 7] + sden[2][kP[s][2]][e & 7];
            const double r = (own - sym / 6.0) / den;
            tnew[h] = (T)((double)old[s][h] + r);
            rout[h] = (T)r;
            if (in) {
                nrm += r * r;
                if (g[0] < nvir_diis && g[1] < nvir_diis && g[2] < nvir_diis) nact += r * r;
            }
        }
        if (!any_in) continue;
        store2<T, PAIR>(tb, nvir, Rg, s, tnew);
        if (r3_out != nullptr) store2<T, PAIR>(r3_out + blk * v3, nvir, Rg, s, rout);
    }
    // fixed-order reduction: wave shuffles, then the four wave sums in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nrm += __shfl_down(nrm, off, 64);
        nact += __shfl_down(nact, off, 64);
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    if (lane == 0) {
        red[0][wave] = nrm;
        red[1][wave] = nact;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        partials[2 * (blk * norb + orbit)] = a;
        partials[2 * (blk * norb + orbit) + 1] = b;
    }
}

}  // namespace orbit8

__global__ __launch_bounds__(256) void reduce_pairs_kernel(const double *__restrict__ x, int64_t n,
                                                            double *__restrict__ out) {
    __shared__ double sa[256], sb[256];
    double a = 0.0, b = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        a += x[2 * i];
        b += x[2 * i + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

}  // namespace

namespace {

// The paired element mapping needs v even (pairs never straddle a row end) and both arrays aligned
// to two elements; per-job offsets are checked in the kernels.
template <typename A, typename B> bool pair_ok(int64_t nvir, const A *a, const B *b) {
    return nvir % 2 == 0 && reinterpret_cast<uintptr_t>(a) % (2 * sizeof(A)) == 0 &&
           reinterpret_cast<uintptr_t>(b) % (2 * sizeof(B)) == 0;
}

template <typename T>
int s3_gather(T *out, const T *src, int64_t nvir, double beta, const dcc_s3_gather_job *jobs, int64_t njobs,
              dcc_stream_t stream) {
    DCC_CHECK(nvir >= 0 && njobs >= 0, "negative extent");
    if (nvir == 0 || njobs == 0) return 0;
    DCC_CHECK(out != nullptr && src != nullptr && jobs != nullptr, "null pointer");
    const int64_t nt = (nvir + E - 1) / E, nso = n_orbits((nt + 1) / 2);
    DCC_CHECK(kRun * nso * njobs < ((int64_t)1 << 31), "grid too large");
    if (pair_ok(nvir, out, src))
        hipLaunchKernelGGL((s3_gather_kernel<T, true>), dim3((unsigned)(kRun * nso * njobs)), dim3(kThreads), 0,
                           dcc::as_stream(stream), out, src, (int)nvir, beta, jobs, (int)nt, (int)nso);
    else
        hipLaunchKernelGGL((s3_gather_kernel<T, false>), dim3((unsigned)(kRun * nso * njobs)), dim3(kThreads), 0,
                           dcc::as_stream(stream), out, src, (int)nvir, beta, jobs, (int)nt, (int)nso);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T, typename O = T>
int s3_transform(O *out, const T *src, int64_t nvir, double beta, const dcc_s3_job *jobs, int64_t njobs,
                 dcc_stream_t stream) {
    DCC_CHECK(nvir >= 0 && njobs >= 0, "negative extent");
    if (nvir == 0 || njobs == 0) return 0;
    DCC_CHECK(out != nullptr && src != nullptr && jobs != nullptr, "null pointer");
    const int64_t nt = (nvir + E - 1) / E, nso = n_orbits((nt + 1) / 2);
    DCC_CHECK(kRun * nso * njobs < ((int64_t)1 << 31), "grid too large");
    if (pair_ok(nvir, out, src))
        hipLaunchKernelGGL((s3_transform_kernel<T, O, true>), dim3((unsigned)(kRun * nso * njobs)), dim3(kThreads), 0,
                           dcc::as_stream(stream), out, src, (int)nvir, beta, jobs, (int)nt, (int)nso);
    else
        hipLaunchKernelGGL((s3_transform_kernel<T, O, false>), dim3((unsigned)(kRun * nso * njobs)), dim3(kThreads),
                           0, dcc::as_stream(stream), out, src, (int)nvir, beta, jobs, (int)nt, (int)nso);
    DCC_LAUNCH_CHECK();
    return 0;
}

template <typename T>
int r3_finalize(const T *R, T *t3, T *r3_out, int64_t nblk, int64_t nvir, const int32_t *blk_ijk, const double *eia,
                int64_t nocc, int64_t nvir_diis, double *partials, dcc_stream_t stream) {
    DCC_CHECK(nblk >= 0 && nvir >= 0 && nocc >= 0, "negative extent");
    if (nblk == 0 || nvir == 0) return 0;
    DCC_CHECK(R != nullptr && t3 != nullptr && blk_ijk != nullptr && eia != nullptr && partials != nullptr,
              "null pointer");
    const int64_t nt = (nvir + orbit8::E - 1) / orbit8::E, nso = n_orbits((nt + 1) / 2);
    DCC_CHECK(8 * nso * nblk < ((int64_t)1 << 31), "grid too large");
    if (nvir_diis < 0 || nvir_diis > nvir) nvir_diis = nvir;
    if (pair_ok(nvir, R, t3) && (r3_out == nullptr || pair_ok(nvir, r3_out, r3_out)))
        hipLaunchKernelGGL((orbit8::r3_finalize_kernel<T, true>), dim3((unsigned)(8 * nso * nblk)),
                           dim3(orbit8::kThreads), 0, dcc::as_stream(stream), R, t3, r3_out, (int)nvir, blk_ijk, eia,
                           (int)nvir_diis, partials, (int)nt, (int)nso);
    else
        hipLaunchKernelGGL((orbit8::r3_finalize_kernel<T, false>), dim3((unsigned)(8 * nso * nblk)),
                           dim3(orbit8::kThreads), 0, dcc::as_stream(stream), R, t3, r3_out, (int)nvir, blk_ijk, eia,
                           (int)nvir_diis, partials, (int)nt, (int)nso);
    DCC_LAUNCH_CHECK();
    return 0;
}

// Allocation-free finalize.  The (block, tile orbit) partial norms live in this file-scope device
// array instead of a caller workspace; each device holds its own copy from module load until the
// process ends.  In host code the name is a stand-in: the device address is looked up once per device
// with hipGetSymbolAddress and only that address is ever passed to a kernel.
constexpr int64_t kScratchDoubles = (int64_t)3 << 20;  // 24 MiB: 1572864 (block, orbit) pairs
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices];

int scratch_on_current_device(double **out) {
    int dev = -1;
    DCC_HIP(hipGetDevice(&dev));
    DCC_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal out of range for the finalize scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        DCC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int dcc_s3_transform_f64(double *out, const double *src, int64_t nvir, double beta, const dcc_s3_job *jobs,
                         int64_t njobs, dcc_stream_t stream) {
    return s3_transform<double>(out, src, nvir, beta, jobs, njobs, stream);
}

int dcc_s3_transform_f32(float *out, const float *src, int64_t nvir, double beta, const dcc_s3_job *jobs,
                         int64_t njobs, dcc_stream_t stream) {
    return s3_transform<float>(out, src, nvir, beta, jobs, njobs, stream);
}

int dcc_s3_transform_pack_f32(uint32_t *out, const float *src, int64_t nvir, const dcc_s3_job *jobs, int64_t njobs,
                              dcc_stream_t stream) {
    return s3_transform<float, uint32_t>(out, src, nvir, 0.0, jobs, njobs, stream);
}

int dcc_s3_gather_f64(double *out, const double *src, int64_t nvir, double beta, const dcc_s3_gather_job *jobs,
                      int64_t njobs, dcc_stream_t stream) {
    return s3_gather<double>(out, src, nvir, beta, jobs, njobs, stream);
}

int dcc_s3_gather_f32(float *out, const float *src, int64_t nvir, double beta, const dcc_s3_gather_job *jobs,
                      int64_t njobs, dcc_stream_t stream) {
    return s3_gather<float>(out, src, nvir, beta, jobs, njobs, stream);
}

int64_t dcc_r3_finalize_tiles(int64_t nvir) {
    const int64_t nt = (nvir + orbit8::E - 1) / orbit8::E;
    return n_orbits(nt);
}

int dcc_r3_finalize_f64(const double *R, double *t3, double *r3_out, int64_t nblk, int64_t nvir,
                        const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                        double *partials, dcc_stream_t stream) {
    return r3_finalize<double>(R, t3, r3_out, nblk, nvir, blk_ijk, eia, nocc, nvir_diis, partials, stream);
}

int dcc_r3_finalize_f32(const float *R, float *t3, float *r3_out, int64_t nblk, int64_t nvir,
                        const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                        double *partials, dcc_stream_t stream) {
    return r3_finalize<float>(R, t3, r3_out, nblk, nvir, blk_ijk, eia, nocc, nvir_diis, partials, stream);
}

int dcc_r3_finalize_noalloc_f64(const double *R, double *t3, double *r3_out, int64_t nblk, int64_t nvir,
                                const int32_t *blk_ijk, const double *eia, int64_t nocc, int64_t nvir_diis,
                                double *norms, dcc_stream_t stream) {
    DCC_CHECK(nblk >= 0 && nvir >= 0 && nocc >= 0, "negative extent");
    DCC_CHECK(norms != nullptr, "null norms");
    const int64_t tiles = dcc_r3_finalize_tiles(nvir);
    DCC_CHECK(tiles == 0 || nblk <= kScratchDoubles / (2 * tiles),
              "nblk * dcc_r3_finalize_tiles(nvir) exceeds the 1572864 (block, tile orbit) partial sums of the "
              "allocation-free scratch; use dcc_r3_finalize_f64 with a workspace");
    double *ws = nullptr;
    if (int st = scratch_on_current_device(&ws)) return st;
    if (int st = r3_finalize<double>(R, t3, r3_out, nblk, nvir, blk_ijk, eia, nocc, nvir_diis, ws, stream))
        return st;
    return dcc_reduce_pairs_f64(ws, nblk * tiles, norms, stream);
}

int dcc_reduce_pairs_f64(const double *x, int64_t n, double *out, dcc_stream_t stream) {
    DCC_CHECK(n >= 0 && out != nullptr, "bad arguments");
    DCC_CHECK(n == 0 || x != nullptr, "null pointer");
    hipLaunchKernelGGL(reduce_pairs_kernel, dim3(1), dim3(256), 0, dcc::as_stream(stream), x, n, out);
    DCC_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
