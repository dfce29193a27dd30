// jg_device.h -- device-side building blocks for the gfx950 pressure solver.
//
// Data layout in HBM (same as Jet's Array3 / FdmMatrix3, so an upload is a
// plain copy): a matrix level is an array of rows {center, right, up, front}
// with x fastest, index = i + nx*(j + ny*k).  FP64 rows are 32 B, the FP32
// shadow rows of the mixed-precision preconditioner are 16 B; one row is one
// (or two) 16-byte vector loads per lane.  The stencil is symmetric and stored
// as upper couplings only: the -x/-y/-z couplings of a cell are the right/up/
// front of the neighbour on that side (src/jet/fdm_linear_system3.cpp:79-96).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jg {

template <typename T>
struct alignas(4 * sizeof(T)) Row {
    T c, r, u, f;  // center, right (+x), up (+y), front (+z)
};

using idx_t = long;

struct Dims {
    int nx, ny, nz;
    __host__ __device__ idx_t n() const { return (idx_t)nx * ny * nz; }
};

// Device-resident scalars of one solve.  Only the finalize kernels write
// them; every other kernel reads them, so an iteration needs no host round
// trip and can be captured in a graph.
struct Ctl {
    double sigma;      // r.z of the current iterate
    double sigma_old;
    double dq;         // d.q
    double alpha;
    double beta;
    double rr;         // r.r of the recursively updated residual
    double bb;         // b.b of the current right-hand side
    double tol2;       // Jet's stop rule: sigma <= tol2
    double rel2;       // relative rule: rr <= rel2 * bb (0 disables)
    double scale;      // mixed precision: scale of the FP32 right-hand side
    unsigned iter;
    unsigned max_iter;
    int trigger;       // Jet: recompute the true residual next iteration
    int refresh;       // this iteration recomputes the true residual
    int done;          // stop flag: set once converged or out of iterations
    int stop_rule;     // 0: Jet (sigma), 1: relative residual, 2: both
    int nonfinite;     // a reduction produced NaN/Inf
    int pad;
};

constexpr int kWave = 64;
constexpr int NT = 256;  // threads per block of the streaming kernels
// Per reduction slot.  A plane is reduced in tiles of 64 x 4 cells, one partial
// each, so this bounds a plane (a 2-D grid) at about 2^24 cells.
constexpr int kMaxPartials = 65536;

// Several systems of one hierarchy solved by the same launches (a batch
// group) keep their arrays back to back: per system every level's rows and
// vectors at a stride of that level's cell count, kSysSlots reduction slots
// of kMaxPartials partials, kSysCtl control blocks and kSysScal device
// scalars.  A launch covers the group by repeating its grid: the z extent of
// a stencil kernel (nsb z-blocks per system), the y extent of an elementwise
// kernel, the block count of a one-block kernel.  A single system is the
// group of one.
constexpr int kSysSlots = 4;
constexpr idx_t kSysPart = (idx_t)kSysSlots * kMaxPartials;
constexpr int kSysCtl = 2;
constexpr int kSysScal = 16;

// System of a stencil block whose grid has nsb z-blocks per system: bz
// becomes the block's z index within its system, part the index of its
// reduction partial within the system's slot: the tile's own index, so a
// system's partials land in the same slots (and are summed in the same
// order) whether it is launched alone or in a batch group.
__device__ __forceinline__ int block_system(int bx, int by, int& bz, int nsb, idx_t& part) {
    const int sys = bz / nsb;
    bz -= sys * nsb;
    part = bx + (idx_t)gridDim.x * (by + (idx_t)gridDim.y * bz);
    return sys;
}

// ------------------------------------------------------------ lane shift

constexpr int kDppFromLeft = 0x138;   // wave_shr:1 -- lane i reads lane i - 1
constexpr int kDppFromRight = 0x130;  // wave_shl:1 -- lane i reads lane i + 1

// v of the lane to the left / right; the lane without one reads 0.
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ---------------------------------------------------------------- reduce

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
    if ((lane & (kWave - 1)) == 0) sh[lane / kWave] = v;
    __syncthreads();
    double s = 0.0;
    if (lane == 0) {
#pragma unroll
        for (int w = 0; w < NT / kWave; ++w) s += sh[w];
    }
    __syncthreads();
    return s;  // valid in thread 0 only
}

// Writes this block's partial sum of v at index b of the partials (one
// slot per block, deterministic order) for a later finalize kernel.
template <int NT>
__device__ __forceinline__ void block_partial(double v, double* partials, idx_t b) {
    __shared__ double sh[NT / kWave];
    double s = block_sum<NT>(v, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[b] = s;
}

template <int NT>
__device__ __forceinline__ void block_partial2(double v0, double v1, double* p0, double* p1,
                                               idx_t b) {
    __shared__ double sh[2][NT / kWave];
    double s0 = block_sum<NT>(v0, sh[0]);
    double s1 = block_sum<NT>(v1, sh[1]);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        p0[b] = s0;
        p1[b] = s1;
    }
}

// Block size of the one-block finalize kernels.
constexpr int kFinThreads = 1024;

// Sums of n partials of one or two arrays by one block of NT threads, fixed
// order (deterministic): thread t adds p[t], p[t + NT], ... in that order
// (four loads in flight before the adds), then the waves and the block.
template <int NT, int K>
__device__ __forceinline__ void reduce_partials_k(const double* const (&p)[K], int n, double* sh,
                                                  double (&r)[K]) {
    const int t = threadIdx.x;
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    int i = t;
    for (; i + 3 * NT < n; i += 4 * NT) {
        double v[K][4];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < 4; ++u) v[k][u] = p[k][i + u * NT];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < 4; ++u) s[k] += v[k][u];
    }
    for (; i < n; i += NT)
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] += p[k][i];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) s[k] += __shfl_down(s[k], o, kWave);
        if ((t & (kWave - 1)) == 0) sh[k * (NT / kWave) + t / kWave] = s[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        r[k] = 0.0;
#pragma unroll
        for (int w = 0; w < NT / kWave; ++w) r[k] += sh[k * (NT / kWave) + w];
    }
    __syncthreads();  // results valid in every thread
}

template <int NT>
__device__ __forceinline__ double reduce_partials(const double* p, int n, double* sh) {
    const double* const a[1] = {p};
    double r[1];
    reduce_partials_k<NT, 1>(a, n, sh, r);
    return r[0];
}

// XCD-aware tile order.  Workgroups are handed to the 8 XCDs round-robin
// (workgroup id mod 8), and each XCD has its own L2.  Tiles overlap by
// their halos, so spatial neighbours should share an L2: workgroup id w is
// mapped to logical tile (w mod 8) * (W / 8) + w / 8, which gives every XCD a
// contiguous run of tiles (x fastest, then y, then z).  The last W mod 8
// workgroups keep the identity mapping.
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
    const int nbx = gridDim.x, nby = gridDim.y;
    const int total = nbx * nby * (int)gridDim.z;
    int id = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const int per = total / 8;
    if (id < per * 8) id = (id % 8) * per + id / 8;
    bx = id % nbx;
    by = (id / nbx) % nby;
    bz = id / (nbx * nby);
}

// --------------------------------------------------------------- stencil

// Off-diagonal sum  sum_nbr a(c,nbr) * x(nbr)  of a 7-point row at (i,j,k).
template <typename TA, typename TX, typename TC>
__device__ __forceinline__ TC offdiag(const Row<TA>* __restrict__ A, const TX* __restrict__ x,
                                      const Row<TA>& row, idx_t c, int i, int j, int k,
                                      const Dims& d) {
#pragma clang fp contract(off)
    // Jet's relax operand order (src/jet/fdm_gauss_seidel_solver3.cpp:110-116)
    const idx_t sx = 1, sy = d.nx, sz = (idx_t)d.nx * d.ny;
    TC s = (i > 0) ? (TC)A[c - sx].r * (TC)x[c - sx] : TC(0);
    s += (i + 1 < d.nx) ? (TC)row.r * (TC)x[c + sx] : TC(0);
    s += (j > 0) ? (TC)A[c - sy].u * (TC)x[c - sy] : TC(0);
    s += (j + 1 < d.ny) ? (TC)row.u * (TC)x[c + sy] : TC(0);
    s += (k > 0) ? (TC)A[c - sz].f * (TC)x[c - sz] : TC(0);
    s += (k + 1 < d.nz) ? (TC)row.f * (TC)x[c + sz] : TC(0);
    return s;
}

}  // namespace jg
