// jg_kernels.h -- HIP kernels of the gfx950 pressure solver.
//
// Geometry: a 256-thread block covers a bx*by tile of (i, j) columns and
// marches over a chunk of zc planes, keeping the z-neighbours of the current
// plane in registers.  x/y neighbours come from the block's own loads through
// the vector L1 / L2.  Reductions write one FP64 partial per block (fixed
// order) that a one-block finalize kernel sums, so every dot product is
// deterministic and no cross-XCD fence is needed inside a kernel.
#pragma once

#include "jg_device.h"

namespace jg {

// ----------------------------------------------------------- SpMV (+ dot)

// Everything one cell of a z-marching stencil kernel reads in its plane:
// the row, the -x / -y couplings (neighbours' r and u) and, for each of NV
// source vectors, the centre value and the four lateral neighbour values
// (slots 0 centre, 1 -x, 2 +x, 3 -y, 4 +y).  Loads at the domain edge use the
// cell's own address (always valid) and are masked at use.  Loads are stored
// raw and combined where used, so the next plane's loads can all be issued
// before the current plane is computed (software pipelining) without the
// compiler waiting on them early.
template <typename TA, typename TX, int NV>
struct StPlane {
    Row<TA> row;
    TA rl, ud;
    TX v[NV][5];
};

// Column flags of one thread: which lateral neighbours exist.
struct StCol {
    bool l, r, d, u;
};

// crow: cell whose row and lateral values are loaded; cx: cell whose centre
// values are loaded (they differ on the last plane of a chunk, where only the
// +z value of the next plane is needed and the rest re-reads cached data).
template <typename TA, typename TX, int NV>
__device__ __forceinline__ void st_load(StPlane<TA, TX, NV>& p, const Row<TA>* __restrict__ A,
                                        const TX* const (&src)[NV], idx_t crow, idx_t cx,
                                        const StCol& h, int nx) {
    p.row = A[crow];
    p.rl = A[h.l ? crow - 1 : crow].r;
    p.ud = A[h.d ? crow - nx : crow].u;
#pragma unroll
    for (int s = 0; s < NV; ++s) {
        p.v[s][0] = src[s][cx];
        p.v[s][1] = src[s][h.l ? crow - 1 : crow];
        p.v[s][2] = src[s][h.r ? crow + 1 : crow];
        p.v[s][3] = src[s][h.d ? crow - nx : crow];
        p.v[s][4] = src[s][h.u ? crow + nx : crow];
    }
}

// Value of slot q: the source itself (NV = 1) or the CG direction
// z + beta d_old (NV = 2; Jet's axpy(beta, d, s, d) with s = z).
template <typename TX, int NV>
__device__ __forceinline__ TX st_val(const TX (&v)[NV][5], int q, double beta) {
#pragma clang fp contract(off)
    if constexpr (NV == 1) {
        return v[0][q];
    } else {
        return (TX)((double)v[0][q] + beta * (double)v[1][q]);
    }
}

// Jet's mvm operand order (src/jet/fdm_linear_system3.cpp:87-95): centre,
// -x, +x, -y, +y, -z, +z, each product rounded before the add (contraction
// off), so FP64 results are bitwise Jet's.  fm/xm and xp are the z
// neighbours (zero outside the domain).
template <typename TA, typename TX, int NV>
__device__ __forceinline__ TX st_sum(const StPlane<TA, TX, NV>& p, const StCol& h, double beta,
                                     TA fm, TX xm, TX xp) {
#pragma clang fp contract(off)
    TX s = (TX)p.row.c * st_val<TX, NV>(p.v, 0, beta);
    s += h.l ? (TX)p.rl * st_val<TX, NV>(p.v, 1, beta) : TX(0);
    s += h.r ? (TX)p.row.r * st_val<TX, NV>(p.v, 2, beta) : TX(0);
    s += h.d ? (TX)p.ud * st_val<TX, NV>(p.v, 3, beta) : TX(0);
    s += h.u ? (TX)p.row.u * st_val<TX, NV>(p.v, 4, beta) : TX(0);
    s += (TX)fm * xm;
    s += (TX)p.row.f * xp;
    return s;
}

// The z-march shared by the stencil kernels: for each plane k of the chunk
// calls op(c, plane, h, fm, xm, xp) with the next plane's loads already
// issued.  xm / xp are the -z / +z neighbours' values (zero outside).
template <typename TA, typename TX, int NV, typename Op>
__device__ __forceinline__ void st_march(const Row<TA>* __restrict__ A, const TX* const (&src)[NV],
                                         double beta, int i, int j, int k0, int k1, const Dims& d,
                                         Op&& op) {
    const idx_t sz = (idx_t)d.nx * d.ny;
    const StCol h{i > 0, i + 1 < d.nx, j > 0, j + 1 < d.ny};
    idx_t c = i + (idx_t)d.nx * (j + (idx_t)d.ny * k0);
    TX xm = TX(0);
    TA fm = TA(0);
    if (k0 > 0) {
        TX vm[NV][5];
#pragma unroll
        for (int s = 0; s < NV; ++s) vm[s][0] = src[s][c - sz];
        xm = st_val<TX, NV>(vm, 0, beta);
        fm = A[c - sz].f;
    }
    StPlane<TA, TX, NV> cur;
    st_load(cur, A, src, c, c, h, d.nx);
    for (int k = k0; k < k1; ++k, c += sz) {
        StPlane<TA, TX, NV> nxt;
        const bool zp = k + 1 < d.nz;
        st_load(nxt, A, src, k + 1 < k1 ? c + sz : c, zp ? c + sz : c, h, d.nx);
        const TX xp = zp ? st_val<TX, NV>(nxt.v, 0, beta) : TX(0);
        op(c, cur, h, fm, xm, xp);
        xm = st_val<TX, NV>(cur.v, 0, beta);
        fm = cur.row.f;
        cur = nxt;
    }
}

// y = A x  (FdmBlas3::mvm, src/jet/fdm_linear_system3.cpp:79-96); optional
// partial sum of x.y per block (the d.q of CG).
template <typename TA, typename TX, bool DOT>
__global__ __launch_bounds__(NT) void k_spmv(const Row<TA>* __restrict__ A,
                                             const TX* __restrict__ x, TX* __restrict__ y,
                                             Dims d, int zc, double* __restrict__ partial,
                                             const Ctl* __restrict__ ctl) {
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        x += sys * n;
        y += sys * n;
        if (DOT) partial += sys * kSysPart;
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    double acc = 0.0;
    if (i < d.nx && j < d.ny) {
        const TX* const src[1] = {x};
        st_march<TA, TX, 1>(A, src, 0.0, i, j, k0, k1, d,
                            [&](idx_t c, const StPlane<TA, TX, 1>& p, const StCol& h, TA fm, TX xm,
                                TX xp) {
                                const TX s = st_sum<TA, TX, 1>(p, h, 0.0, fm, xm, xp);
                                y[c] = s;
                                if (DOT) acc += (double)p.v[0][0] * (double)s;
                            });
    }
    if (DOT) block_partial<NT>(acc, partial, part);
}

// Fused CG direction update and SpMV:  d = z + beta*d_old (Jet's
// axpy(beta, d, s, d)), q = A d, partial d.q.  d_old and d_new are separate
// buffers because neighbours read d_old while d_new is written.
template <typename TA, typename TX>
__global__ __launch_bounds__(NT) void k_dir_spmv(const Row<TA>* __restrict__ A,
                                                 const TX* __restrict__ z,
                                                 const TX* __restrict__ dold,
                                                 TX* __restrict__ dnew, TX* __restrict__ q,
                                                 Dims d, int zc, double* __restrict__ partial,
                                                 const Ctl* __restrict__ ctl) {
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        z += sys * n;
        dold += sys * n;
        dnew += sys * n;
        q += sys * n;
        partial += sys * kSysPart;
        ctl += sys * kSysCtl;
    }
    if (ctl->done) return;
    const double beta = ctl->beta;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    double acc = 0.0;
    if (i < d.nx && j < d.ny) {
        const TX* const src[2] = {z, dold};
        st_march<TA, TX, 2>(A, src, beta, i, j, k0, k1, d,
                            [&](idx_t c, const StPlane<TA, TX, 2>& p, const StCol& h, TA fm, TX xm,
                                TX xp) {
                                const TX s = st_sum<TA, TX, 2>(p, h, beta, fm, xm, xp);
                                const TX dc = st_val<TX, 2>(p.v, 0, beta);
                                q[c] = s;
                                dnew[c] = dc;
                                acc += (double)dc * (double)s;
                            });
    }
    block_partial<NT>(acc, partial, part);
}

// Geometry of the tiled z-march (tm_march): one wave per x row of a
// 64 x kTmY tile of each plane, kTmZ planes per block (on the 256^3 sphere
// pool 8 x 16 is within noise of the best of 4, 8, 16 rows and 8, 16, 32
// planes for the SpMV).
constexpr int kTmX = 64, kTmY = 8, kTmZ = 16;

// What a cell of the tiled z-march sees in plane k: its row, its value v
// (the stencil source of the kernel), the values of its six neighbours
// (zero outside the grid), the -x / -y / -z couplings and cen, the value
// of an optional centre-only vector (b of a residual).
template <typename T>
struct TmCell {
    Row<T> row;
    T v, l, r, d, u, m, p;  // centre, -x, +x, -y, +y, -z, +z
    T rl, ud, fm;           // A.r of -x, A.u of -y, A.f of -z
    T cen;
};

// The tiled z-march shared by the stencil kernels, reading each cell's row
// and sources once: NV source vectors combine into the stencil value
// val(s0[, s1]) (x itself, or the CG direction z + beta d_old).  x
// neighbours come from the lanes either side (DPP), y neighbours from the
// adjacent waves through a double-buffered LDS row per wave (one barrier
// per plane), z neighbours from registers; the tile's edge lanes and waves
// load the sources beyond the tile from memory, a plane ahead like the
// rows (the centre sources two planes ahead: plane k needs the value of
// plane k + 1).  op(c, in, cell) runs for every thread of every plane k0 ..
// k1 - 1 of the block's column (in: the cell is inside the grid; c is its
// index when it is).  All threads of the block call this with the same
// (bx, by, k0, k1).
template <typename T, int NV, typename Val, typename Op>
__device__ __forceinline__ void tm_march(const Row<T>* __restrict__ A,
                                         const T* const (&src)[NV], const T* __restrict__ cen,
                                         const Dims& d, int bx, int by, int k0, int k1,
                                         T (&sv)[2][kTmY][kTmX], T (&su)[2][kTmY][kTmX],
                                         Val&& val, Op&& op) {
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int gx = bx * kTmX + t, gy = by * kTmY + w;
    const bool in = gx < d.nx && gy < d.ny;
    const bool hl = gx > 0, hr = gx + 1 < d.nx, hd = gy > 0, hu = gy + 1 < d.ny;
    const idx_t sz = (idx_t)d.nx * d.ny;
    const idx_t col = in ? gx + (idx_t)d.nx * gy : 0;
    // edge roles: the first / last lane loads the x neighbour beyond the
    // tile, the first / last wave the y neighbour
    const bool xl = t == 0 && in && hl, xr = t == kTmX - 1 && in && hr;
    const bool yd = w == 0 && in && hd, yu = w == kTmY - 1 && in && hu;
    const idx_t xoff = t == 0 ? -1 : 1;
    const idx_t yoff = w == 0 ? -(idx_t)d.nx : (idx_t)d.nx;
    struct Raw {
        T s[NV];
    };
    struct Pl {
        Row<T> row;
        T cen, ex_r, ey_u;
        T ex[NV], ey[NV];
    };
    auto vof = [&](const T (&s)[NV]) __attribute__((always_inline)) {
        if constexpr (NV == 1)
            return val(s[0]);
        else
            return val(s[0], s[1]);
    };
    auto load_raw = [&](int k) __attribute__((always_inline)) {
        Raw v;
#pragma unroll
        for (int i = 0; i < NV; ++i) v.s[i] = T(0);
        if (in && k < d.nz) {
            const idx_t c = col + sz * k;
#pragma unroll
            for (int i = 0; i < NV; ++i) v.s[i] = src[i][c];
        }
        return v;
    };
    auto load_pl = [&](int k) __attribute__((always_inline)) {
        Pl p;
        p.row = Row<T>{T(0), T(0), T(0), T(0)};
        p.cen = p.ex_r = p.ey_u = T(0);
#pragma unroll
        for (int i = 0; i < NV; ++i) p.ex[i] = p.ey[i] = T(0);
        if (in && k < k1) {
            const idx_t c = col + sz * k;
            p.row = A[c];
            if (cen) p.cen = cen[c];
            if (xl || xr) {
#pragma unroll
                for (int i = 0; i < NV; ++i) p.ex[i] = src[i][c + xoff];
                if (xl) p.ex_r = A[c - 1].r;
            }
            if (yd || yu) {
#pragma unroll
                for (int i = 0; i < NV; ++i) p.ey[i] = src[i][c + yoff];
                if (yd) p.ey_u = A[c - d.nx].u;
            }
        }
        return p;
    };
    T vm = T(0), fm = T(0);  // value and A.f of plane k - 1 (zero below the grid)
    if (in && k0 > 0) {
        const Raw r = [&] {
            Raw v;
            const idx_t c = col + sz * (k0 - 1);
#pragma unroll
            for (int i = 0; i < NV; ++i) v.s[i] = src[i][c];
            return v;
        }();
        vm = vof(r.s);
        fm = A[col + sz * (k0 - 1)].f;
    }
    Pl cur = load_pl(k0);
    const Raw r0 = load_raw(k0);
    Raw rn1 = load_raw(k0 + 1);
    T vk = vof(r0.s);
    for (int k = k0; k < k1; ++k) {
        const Pl nxt = load_pl(k + 1);
        const Raw rn2 = load_raw(k + 2);
        const T vp = k + 1 < d.nz ? vof(rn1.s) : T(0);  // value of plane k + 1
        const int buf = k & 1;
        sv[buf][w][t] = vk;
        su[buf][w][t] = cur.row.u;
        TmCell<T> e;
        e.row = cur.row;
        e.v = vk;
        e.l = dpp_shift<kDppFromLeft>(vk);
        e.r = dpp_shift<kDppFromRight>(vk);
        e.rl = dpp_shift<kDppFromLeft>(cur.row.r);
        if (xl) {
            e.l = vof(cur.ex);
            e.rl = cur.ex_r;
        }
        if (xr) e.r = vof(cur.ex);
        __syncthreads();
        e.d = sv[buf][w > 0 ? w - 1 : 0][t];
        e.u = sv[buf][w + 1 < kTmY ? w + 1 : w][t];
        e.ud = su[buf][w > 0 ? w - 1 : 0][t];
        if (yd) {
            e.d = vof(cur.ey);
            e.ud = cur.ey_u;
        }
        if (yu) e.u = vof(cur.ey);
        e.m = vm;
        e.p = vp;
        e.fm = fm;
        e.cen = cur.cen;
        // (masks: neighbours outside the grid hold zero or a clamped value;
        // op tests hl / hr / hd / hu like the per-cell kernels)
        op(in ? col + sz * k : 0, in, e, hl, hr, hd, hu);
        vm = vk;
        fm = cur.row.f;
        vk = vp;
        rn1 = rn2;
        cur = nxt;
    }
}

// Grid of a tm_march kernel on one system of shape d: 64 x 8 columns,
// z-chunks of at most kTmZ planes, shortened on small grids so a system has
// about 2048 blocks, and at most kMaxPartials blocks.  The chunking depends
// on the shape alone, so a member of a batch group is cut (and its
// reductions summed) exactly as the same system solved alone.
inline void tm_grid(const Dims& d, dim3& grid, int& zc) {
    const long cols = (long)((d.nx + kTmX - 1) / kTmX) * ((d.ny + kTmY - 1) / kTmY);
    zc = (int)std::max<long>(1, std::min<long>(std::min(d.nz, kTmZ), cols * d.nz / 2048));
    while (cols * ((d.nz + zc - 1) / zc) > kMaxPartials) ++zc;
    grid = dim3((d.nx + kTmX - 1) / kTmX, (d.ny + kTmY - 1) / kTmY, (unsigned)((d.nz + zc - 1) / zc));
}

// k_dir_spmv through the tiled z-march: the same values, operand order and
// per-cell products (the block sums its d.q partial in its own order).  A
// grid z extent of nsb z-blocks per system covers a batch group.
template <typename T>
__global__ __launch_bounds__(kTmX * kTmY) void k_dir_spmv_tile(
    const Row<T>* __restrict__ A, const T* __restrict__ z, const T* __restrict__ dold,
    T* __restrict__ dnew, T* __restrict__ q, Dims d, int zc, double* __restrict__ partial,
    const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        z += sys * n;
        dold += sys * n;
        dnew += sys * n;
        q += sys * n;
        partial += sys * kSysPart;
        ctl += sys * kSysCtl;
    }
    if (ctl->done) return;
    const double beta = ctl->beta;
    __shared__ T sv[2][kTmY][kTmX];
    __shared__ T su[2][kTmY][kTmX];
    const T* const src[2] = {z, dold};
    double acc = 0.0;
    tm_march<T, 2>(
        A, src, nullptr, d, bx, by, bz * zc, min(d.nz, bz * zc + zc), sv, su,
        [&](T zv, T ov) { return (T)((double)zv + beta * (double)ov); },
        [&](idx_t c, bool in, const TmCell<T>& e, bool hl, bool hr, bool hd, bool hu) {
            // Jet's mvm order: centre, -x, +x, -y, +y, -z, +z
            T s = e.row.c * e.v;
            s += hl ? e.rl * e.l : T(0);
            s += hr ? e.row.r * e.r : T(0);
            s += hd ? e.ud * e.d : T(0);
            s += hu ? e.row.u * e.u : T(0);
            s += e.fm * e.m;
            s += e.row.f * e.p;
            if (in) {
                q[c] = s;
                dnew[c] = e.v;
                acc += (double)e.v * (double)s;
            }
        });
    block_partial<kTmX * kTmY>(acc, partial, part);
}

// r = b - A x (k_residual: the same values in Jet's residual operand order)
// through the tiled z-march, stored as TR; DOT: one partial of r.r per
// block; REFRESH / GATED: k_residual's control-block gates.  A grid z
// extent of nsb z-blocks per system covers a batch group.
template <typename T, typename TR, bool DOT, bool REFRESH, bool GATED>
__global__ __launch_bounds__(kTmX * kTmY) void k_residual_tile(
    const Row<T>* __restrict__ A, const T* __restrict__ x, const T* __restrict__ b,
    TR* __restrict__ r, Dims d, int zc, double* __restrict__ partial,
    const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        x += sys * n;
        b += sys * n;
        r += sys * n;
        if (DOT) partial += sys * kSysPart;
        if (ctl) ctl += sys * kSysCtl;
    }
    if (REFRESH && (ctl->done || !ctl->refresh)) return;
    if (GATED && ctl && ctl->done) return;
    __shared__ T sv[2][kTmY][kTmX];
    __shared__ T su[2][kTmY][kTmX];
    const T* const src[1] = {x};
    double acc = 0.0;
    tm_march<T, 1>(
        A, src, b, d, bx, by, bz * zc, min(d.nz, bz * zc + zc), sv, su, [](T xv) { return xv; },
        [&](idx_t c, bool in, const TmCell<T>& e, bool hl, bool hr, bool hd, bool hu) {
            // Jet's residual operand order (src/jet/fdm_linear_system3.cpp:107-115)
            T rv = e.cen - e.row.c * e.v;
            rv -= hl ? e.rl * e.l : T(0);
            rv -= hr ? e.row.r * e.r : T(0);
            rv -= hd ? e.ud * e.d : T(0);
            rv -= hu ? e.row.u * e.u : T(0);
            rv -= e.fm * e.m;
            rv -= e.row.f * e.p;
            if (in) {
                r[c] = (TR)rv;
                if (DOT) acc += (double)rv * (double)rv;
            }
        });
    if (DOT) block_partial<kTmX * kTmY>(acc, partial, part);
}

// r = b - A x (FdmBlas3::residual, src/jet/fdm_linear_system3.cpp:98-116);
// TR may be narrower than TX (FP32 right-hand side of the mixed-precision
// inner solve).  DOT: partial r.r.  REFRESH: only on iterations where the
// control block asks for a true residual (Jet's every-50 / sigma rule).
// GATED: skipped once the control block says the solve has stopped (a
// V-cycle residual inside a replayed batch graph).
template <typename TA, typename TX, typename TR, bool DOT, bool REFRESH, bool GATED = false>
__global__ __launch_bounds__(NT) void k_residual(const Row<TA>* __restrict__ A,
                                                 const TX* __restrict__ x,
                                                 const TX* __restrict__ b, TR* __restrict__ r,
                                                 Dims d, int zc, double* __restrict__ partial,
                                                 const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        x += sys * n;
        b += sys * n;
        r += sys * n;
        if (DOT) partial += sys * kSysPart;
        if (ctl) ctl += sys * kSysCtl;
    }
    if (REFRESH && (ctl->done || !ctl->refresh)) return;
    if (GATED && ctl && ctl->done) return;
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    double acc = 0.0;
    if (i < d.nx && j < d.ny) {
        const idx_t sz = (idx_t)d.nx * d.ny;
        const TX* const src[1] = {x};
        TX bn = b[i + (idx_t)d.nx * (j + (idx_t)d.ny * k0)];
        st_march<TA, TX, 1>(A, src, 0.0, i, j, k0, k1, d,
                            [&](idx_t c, const StPlane<TA, TX, 1>& p, const StCol& h, TA fm, TX xm,
                                TX xp) {
                                const TX bc = bn;
                                bn = b[c + sz < (idx_t)d.nz * sz ? c + sz : c];  // next plane's b
                                // Jet's residual operand order (src/jet/fdm_linear_system3.cpp:107-115)
                                TX rv = bc - (TX)p.row.c * p.v[0][0];
                                rv -= h.l ? (TX)p.rl * p.v[0][1] : TX(0);
                                rv -= h.r ? (TX)p.row.r * p.v[0][2] : TX(0);
                                rv -= h.d ? (TX)p.ud * p.v[0][3] : TX(0);
                                rv -= h.u ? (TX)p.row.u * p.v[0][4] : TX(0);
                                rv -= (TX)fm * xm;
                                rv -= (TX)p.row.f * xp;
                                r[c] = (TR)rv;
                                if (DOT) acc += (double)rv * (double)rv;
                            });
    }
    if (DOT) block_partial<NT>(acc, partial, part);
}

// ---------------------------------------------------- elementwise updates

// CG step: x += alpha d; r -= alpha q (skipped on true-residual refresh
// iterations, where k_residual<..., REFRESH> recomputes r from the new x).
template <typename TX>
__global__ __launch_bounds__(NT) void k_update_xr(TX* __restrict__ x, TX* __restrict__ r,
                                                  const TX* __restrict__ dd,
                                                  const TX* __restrict__ q, idx_t n,
                                                  const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int sys = blockIdx.y;  // the system of a batch group
    x += sys * n;
    r += sys * n;
    dd += sys * n;
    q += sys * n;
    ctl += sys * kSysCtl;
    if (ctl->done) return;
    const double alpha = ctl->alpha;
    const bool upd_r = !ctl->refresh;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT) {
        x[c] = (TX)((double)x[c] + alpha * (double)dd[c]);
        if (upd_r) r[c] = (TX)((double)r[c] - alpha * (double)q[c]);
    }
}

template <typename TA, typename TB>
__global__ __launch_bounds__(NT) void k_dot(const TA* __restrict__ a, const TB* __restrict__ b,
                                            idx_t n, double* __restrict__ partial,
                                            const Ctl* __restrict__ ctl) {
    const int sys = blockIdx.y;  // the system of a batch group
    a += sys * n;
    b += sys * n;
    partial += sys * kSysPart;
    if (ctl) ctl += sys * kSysCtl;
    if (ctl && ctl->done) return;
    double acc = 0.0;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT)
        acc += (double)a[c] * (double)b[c];
    block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

// r.z and r.r partials (slots p_rz, p_rr) in one pass over r and z.
template <typename T>
__global__ __launch_bounds__(NT) void k_dot_rz_rr(const T* __restrict__ r, const T* __restrict__ z,
                                                  idx_t n, double* __restrict__ p_rz,
                                                  double* __restrict__ p_rr,
                                                  const Ctl* __restrict__ ctl) {
    const int sys = blockIdx.y;  // the system of a batch group
    r += sys * n;
    z += sys * n;
    p_rz += sys * kSysPart;
    p_rr += sys * kSysPart;
    if (ctl) ctl += sys * kSysCtl;
    if (ctl && ctl->done) return;
    double a0 = 0.0, a1 = 0.0;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT) {
        const double rv = (double)r[c];
        a0 += rv * (double)z[c];
        a1 += rv * rv;
    }
    block_partial2<NT>(a0, a1, p_rz, p_rr, (idx_t)blockIdx.x);
}

template <typename TX>
__global__ __launch_bounds__(NT) void k_axpy(double a, const TX* __restrict__ x,
                                             const TX* __restrict__ y, TX* __restrict__ out,
                                             idx_t n) {
#pragma clang fp contract(off)
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT)
        out[c] = a * x[c] + y[c];
}

template <typename TX>
__global__ __launch_bounds__(NT) void k_fill(TX* __restrict__ x, TX v, idx_t n) {
    x += blockIdx.y * n;  // the system of a batch group
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT)
        x[c] = v;
}

template <typename TS, typename TD>
__global__ __launch_bounds__(NT) void k_convert(const TS* __restrict__ s, TD* __restrict__ dst,
                                                idx_t n) {
    s += blockIdx.y * n;  // the system of a batch group
    dst += blockIdx.y * n;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT)
        dst[c] = (TD)s[c];
}

// FP32 shadow of an FP64 matrix (mixed-precision preconditioner).
__global__ __launch_bounds__(NT) void k_rows_to_f32(const Row<double>* __restrict__ s,
                                                    Row<float>* __restrict__ dst, idx_t n) {
    s += blockIdx.y * n;  // the system of a batch group
    dst += blockIdx.y * n;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT) {
        const Row<double> r = s[c];
        Row<float> o;
        o.c = (float)r.c;
        o.r = (float)r.r;
        o.u = (float)r.u;
        o.f = (float)r.f;
        dst[c] = o;
    }
}

// Mixed precision: x64 += e32 (iterative-refinement correction).
__global__ __launch_bounds__(NT) void k_add_correction(double* __restrict__ x,
                                                       const float* __restrict__ e, idx_t n) {
    x += blockIdx.y * n;  // the system of a batch group
    e += blockIdx.y * n;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT)
        x[c] += (double)e[c];
}

// ------------------------------------------------------------- smoothers

// One colour of a red-black SOR sweep (FdmGaussSeidelSolver3::relaxRedBlack,
// src/jet/fdm_gauss_seidel_solver3.cpp:150-221): red = (i+j+k) even.
// ZERO_INIT: x is treated as all zeros before this red pass (first
// pre-smoothing sweep from a zero initial guess): red cells get omega*b/c,
// black cells 0 -- no read of x.
template <typename TA, typename TX, bool ZERO_INIT>
__global__ __launch_bounds__(NT) void k_rbgs(const Row<TA>* __restrict__ A,
                                             const TX* __restrict__ b, TX* __restrict__ x,
                                             Dims d, int zc, int color, TX omega,
                                             const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= d.nx || j >= d.ny) return;
    const int k0 = blockIdx.z * zc;
    const int k1 = min(d.nz, k0 + zc);
    idx_t c = i + (idx_t)d.nx * (j + (idx_t)d.ny * k0);
    const idx_t sz = (idx_t)d.nx * d.ny;
    for (int k = k0; k < k1; ++k, c += sz) {
        const bool mine = ((i + j + k) & 1) == color;
        if (ZERO_INIT) {
            x[c] = mine ? omega * b[c] / (TX)A[c].c : TX(0);
        } else if (mine) {
            const Row<TA> row = A[c];
            const TX s = offdiag<TA, TX, TX>(A, x, row, c, i, j, k, d);
            x[c] = (TX(1) - omega) * x[c] + omega * (b[c] - s) / (TX)row.c;
        }
    }
}

// Weighted Jacobi sweep: xout = x + omega ((b - s) / c - x)  (omega = 1 is
// FdmJacobiSolver3::relax, src/jet/fdm_jacobi_solver3.cpp:95-112, computed
// as its (b - s) / c).  ZC: a zero-centre row gives 0 (the V-cycle smoother;
// Jet's relax divides by the zero).
template <typename TA, typename TX, bool ZC = false>
__global__ __launch_bounds__(NT) void k_jacobi(const Row<TA>* __restrict__ A,
                                               const TX* __restrict__ b,
                                               const TX* __restrict__ x, TX* __restrict__ xo,
                                               Dims d, int zc, TX omega,
                                               const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= d.nx || j >= d.ny) return;
    const int k0 = blockIdx.z * zc;
    const int k1 = min(d.nz, k0 + zc);
    idx_t c = i + (idx_t)d.nx * (j + (idx_t)d.ny * k0);
    const idx_t sz = (idx_t)d.nx * d.ny;
    for (int k = k0; k < k1; ++k, c += sz) {
        const Row<TA> row = A[c];
        if (ZC && row.c == TA(0)) {
            xo[c] = TX(0);
            continue;
        }
        const TX s = offdiag<TA, TX, TX>(A, x, row, c, i, j, k, d);
        if (omega == TX(1))
            xo[c] = (b[c] - s) / (TX)row.c;
        else
            xo[c] = x[c] + omega * ((b[c] - s) / (TX)row.c - x[c]);
    }
}

// ------------------------------------------------------------- transfers

// FdmMgUtils3::restrict (src/jet/fdm_mg_linear_system3.cpp:43-98): coarse
// (i,j,k) = sum w_x w_y w_z fine(idx), 1-D weights {1/8,3/8,3/8,1/8} over
// {2i-1, 2i, 2i+1, 2i+2} clamped at both ends.  One thread per coarse cell.
template <typename TX>
__global__ __launch_bounds__(NT) void k_restrict(const TX* __restrict__ fine,
                                                 TX* __restrict__ coarse, Dims cd,
                                                 const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int sys = blockIdx.z / cd.nz;  // the system of a batch group (one plane per z-block)
    const int k = blockIdx.z - sys * cd.nz;
    if (sys > 0) {
        fine += 8 * sys * cd.n();
        coarse += sys * cd.n();
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= cd.nx || j >= cd.ny) return;
    const int fnx = 2 * cd.nx, fny = 2 * cd.ny;
    const TX kern[4] = {TX(0.125), TX(0.375), TX(0.375), TX(0.125)};
    const int ii[4] = {i > 0 ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1, i + 1 < cd.nx ? 2 * i + 2 : 2 * i + 1};
    const int jj[4] = {j > 0 ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1, j + 1 < cd.ny ? 2 * j + 2 : 2 * j + 1};
    const int kk[4] = {k > 0 ? 2 * k - 1 : 2 * k, 2 * k, 2 * k + 1, k + 1 < cd.nz ? 2 * k + 2 : 2 * k + 1};
    // Jet's accumulation order (z, y, x; weight product first).
    TX sum = 0;
#pragma unroll
    for (int z = 0; z < 4; ++z)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const TX* row = fine + (idx_t)fnx * (jj[y] + (idx_t)fny * kk[z]);
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const TX w = kern[x] * kern[y] * kern[z];
                sum += w * row[ii[x]];
            }
        }
    coarse[i + (idx_t)cd.nx * (j + (idx_t)cd.ny * k)] = sum;
}

// Coarse rows of the flat-entry MGPCG from the next finer level (no
// counterpart in Jet, whose multigrid callers build each level from markers).
// Coarse cell C aggregates the 2 x 2 x DZ fine cells it covers (DZ = 1: a
// one-plane level).  With A = M + L (M the row sums, L the coupling part) the
// coarse operator is kappa (P'AP + P'MP), kappa = 1/16 (1/8 on one plane), P
// the piecewise-constant injection: a uniform stencil keeps its mass and
// quarters its couplings, i.e. it is rediscretised at twice the spacing.
//   right/up/front_C = kappa * (fine couplings leaving C in +x / +y / +z)
//   center_C = kappa * ((2 sum centers + 4 sum inner couplings) + sum of
//              couplings leaving C through its faces)
// Couplings to cells outside the grid are not counted.  FP64, one thread per
// coarse cell, every sum in a fixed order.
template <int DZ>
__global__ __launch_bounds__(NT) void k_mg_coarsen(const Row<double>* __restrict__ fine,
                                                   Row<double>* __restrict__ coarse, Dims cd) {
#pragma clang fp contract(off)
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= cd.nx || J >= cd.ny) return;
    const int fnx = 2 * cd.nx, fny = 2 * cd.ny;
    const int i0 = 2 * I, j0 = 2 * J, k0 = DZ * K;
    auto at = [&](int i, int j, int k) -> const Row<double>& {
        return fine[i + (idx_t)fnx * (j + (idx_t)fny * k)];
    };
    double sc = 0.0, si = 0.0;
    double px = 0.0, mx = 0.0, py = 0.0, my = 0.0, pz = 0.0, mz = 0.0;
#pragma unroll
    for (int c = 0; c < DZ; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) sc += at(i0 + a, j0 + b, k0 + c).c;
#pragma unroll
    for (int c = 0; c < DZ; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b) si += at(i0, j0 + b, k0 + c).r;
#pragma unroll
    for (int c = 0; c < DZ; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a) si += at(i0 + a, j0, k0 + c).u;
    if constexpr (DZ == 2) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) si += at(i0 + a, j0 + b, k0).f;
    }
#pragma unroll
    for (int c = 0; c < DZ; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if (I + 1 < cd.nx) px += at(i0 + 1, j0 + b, k0 + c).r;
            if (I > 0) mx += at(i0 - 1, j0 + b, k0 + c).r;
        }
#pragma unroll
    for (int c = 0; c < DZ; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (J + 1 < cd.ny) py += at(i0 + a, j0 + 1, k0 + c).u;
            if (J > 0) my += at(i0 + a, j0 - 1, k0 + c).u;
        }
    if constexpr (DZ == 2) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (K + 1 < cd.nz) pz += at(i0 + a, j0 + b, k0 + 1).f;
                if (K > 0) mz += at(i0 + a, j0 + b, k0 - 1).f;
            }
    }
    const double kappa = DZ == 2 ? 0.0625 : 0.125;
    const double se = ((((px + mx) + py) + my) + pz) + mz;
    Row<double> r;
    r.c = kappa * ((2.0 * sc + 4.0 * si) + se);
    r.r = kappa * px;
    r.u = kappa * py;
    r.f = kappa * pz;
    coarse[I + (idx_t)cd.nx * (J + (idx_t)cd.ny * K)] = r;
}

// FdmMgUtils3::correct (src/jet/fdm_mg_linear_system3.cpp:100-179): fine +=
// trilinear (1/4, 3/4) prolongation of coarse, clamped.  Jet clamps the k
// upper index against n.y (:159); this kernel clamps against min(n.y, n.z),
// which equals Jet wherever Jet stays in bounds (n.y <= n.z) and stays in
// bounds where Jet would read past the coarse array (n.y > n.z).
template <typename TX>
__global__ __launch_bounds__(NT) void k_correct(const TX* __restrict__ coarse,
                                                TX* __restrict__ fine, Dims fd,
                                                const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    const int sys = blockIdx.z / fd.nz;  // the system of a batch group (one plane per z-block)
    const int k = blockIdx.z - sys * fd.nz;
    if (sys > 0) {
        fine += sys * fd.n();
        coarse += sys * (fd.n() / 8);
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= fd.nx || j >= fd.ny) return;
    const int cnx = fd.nx / 2, cny = fd.ny / 2;
    const int kclamp = min(fd.ny, fd.nz);
    const int ci = i / 2, cj = j / 2, ck = k / 2;
    int i0, i1, j0, j1, k0, k1;
    TX wi0, wi1, wj0, wj1, wk0, wk1;
    if ((i & 1) == 0) { i0 = i > 1 ? ci - 1 : ci; i1 = ci; wi0 = TX(0.25); wi1 = TX(0.75); }
    else { i0 = ci; i1 = i + 1 < fd.nx ? ci + 1 : ci; wi0 = TX(0.75); wi1 = TX(0.25); }
    if ((j & 1) == 0) { j0 = j > 1 ? cj - 1 : cj; j1 = cj; wj0 = TX(0.25); wj1 = TX(0.75); }
    else { j0 = cj; j1 = j + 1 < fd.ny ? cj + 1 : cj; wj0 = TX(0.75); wj1 = TX(0.25); }
    if ((k & 1) == 0) { k0 = k > 1 ? ck - 1 : ck; k1 = ck; wk0 = TX(0.25); wk1 = TX(0.75); }
    else { k0 = ck; k1 = k + 1 < kclamp ? ck + 1 : ck; wk0 = TX(0.75); wk1 = TX(0.25); }
    auto C = [&](int a, int b_, int c_) { return coarse[a + (idx_t)cnx * (b_ + (idx_t)cny * c_)]; };
    const int iI[2] = {i0, i1}, jI[2] = {j0, j1}, kI[2] = {k0, k1};
    const TX iW[2] = {wi0, wi1}, jW[2] = {wj0, wj1}, kW[2] = {wk0, wk1};
    // Jet's accumulation order: eight separate += of (wi*wj*wk*coarse).
    const idx_t f = i + (idx_t)fd.nx * (j + (idx_t)fd.ny * k);
    TX v = fine[f];
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int x = 0; x < 2; ++x) v += iW[x] * jW[y] * kW[z] * C(iI[x], jI[y], kI[z]);
    fine[f] = v;
}

// r = b - A x on a one-plane (2-D) level, one thread per cell, in
// FdmBlas3::residual's operand order without the z terms (src/jet/
// fdm_linear_system2.cpp's FdmBlas2::residual has the same four): the
// z-marching k_residual would spend most of its pipeline on a single plane.
template <typename T>
__device__ __forceinline__ T residual_plane_at(const Row<T>* __restrict__ A,
                                               const T* __restrict__ x, const T* __restrict__ b,
                                               int i, int j, const Dims& d) {
#pragma clang fp contract(off)
    const idx_t c = i + (idx_t)d.nx * j;
    const Row<T> row = A[c];
    T rv = b[c] - row.c * x[c];
    rv -= i > 0 ? A[c - 1].r * x[c - 1] : T(0);
    rv -= i + 1 < d.nx ? row.r * x[c + 1] : T(0);
    rv -= j > 0 ? A[c - d.nx].u * x[c - d.nx] : T(0);
    rv -= j + 1 < d.ny ? row.u * x[c + d.nx] : T(0);
    return rv;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_residual_plane(const Row<T>* __restrict__ A,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ b,
                                                       T* __restrict__ r, Dims d,
                                                       const Ctl* __restrict__ ctl) {
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= d.nx || j >= d.ny) return;
    r[i + (idx_t)d.nx * j] = residual_plane_at(A, x, b, i, j, d);
}

// PCG's periodic true residual (k_residual's REFRESH) on one plane: the
// same values, over a bounded grid (a grid-stride loop), so the iterations
// whose control block skips it cost one small launch rather than one
// early-exit block per stencil tile.
template <typename T>
__global__ __launch_bounds__(NT) void k_residual_plane_refresh(const Row<T>* __restrict__ A,
                                                               const T* __restrict__ x,
                                                               const T* __restrict__ b,
                                                               T* __restrict__ r, Dims d,
                                                               const Ctl* __restrict__ ctl) {
    if (ctl->done || !ctl->refresh) return;
    const idx_t n = (idx_t)d.nx * d.ny;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT) {
        const int j = (int)(c / d.nx), i = (int)(c - (idx_t)j * d.nx);
        r[c] = residual_plane_at(A, x, b, i, j, d);
    }
}

// r = b - A x of an FP64 one-plane level (k_residual_plane's values) stored
// as TR, with one partial of r.r per block: the finest-level residual of
// mixed refinement and of the final report, over a bounded grid.
template <typename TR>
__global__ __launch_bounds__(NT) void k_residual_plane_dot(const Row<double>* __restrict__ A,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ b,
                                                           TR* __restrict__ r, Dims d,
                                                           double* __restrict__ partial) {
    const idx_t n = (idx_t)d.nx * d.ny;
    double acc = 0.0;
    for (idx_t c = blockIdx.x * (idx_t)NT + threadIdx.x; c < n; c += (idx_t)gridDim.x * NT) {
        const int j = (int)(c / d.nx), i = (int)(c - (idx_t)j * d.nx);
        const double rv = residual_plane_at(A, x, b, i, j, d);
        r[c] = (TR)rv;
        acc += rv * rv;
    }
    block_partial<NT>(acc, partial, (idx_t)blockIdx.x);
}

// FdmMgUtils2::restrict (src/jet/fdm_mg_linear_system2.cpp:44-84) on one
// plane: the 3-D taps and clamps in x and y, the 16 products summed y-major
// with each weight kernel[x] * kernel[y] formed first.  One thread per
// coarse cell.
template <typename TX>
__global__ __launch_bounds__(NT) void k_restrict2(const TX* __restrict__ fine,
                                                  TX* __restrict__ coarse, Dims cd,
                                                  const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= cd.nx || j >= cd.ny) return;
    const int fnx = 2 * cd.nx;
    const TX kern[4] = {TX(0.125), TX(0.375), TX(0.375), TX(0.125)};
    const int ii[4] = {i > 0 ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1, i + 1 < cd.nx ? 2 * i + 2 : 2 * i + 1};
    const int jj[4] = {j > 0 ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1, j + 1 < cd.ny ? 2 * j + 2 : 2 * j + 1};
    TX sum = 0;
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const TX* row = fine + (idx_t)fnx * jj[y];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const TX w = kern[x] * kern[y];
            sum += w * row[ii[x]];
        }
    }
    coarse[i + (idx_t)cd.nx * j] = sum;
}

// FdmMgUtils2::correct (src/jet/fdm_mg_linear_system2.cpp:86-143) at fine
// cell (i, j) of a plane: v + the bilinear (1/4, 3/4) prolongation of
// coarse, clamped at both ends in x and y, four separate adds of
// (wi * wj) * coarse, y-major.
template <typename TX>
__device__ __forceinline__ TX correct2_at(const TX* __restrict__ coarse, TX v, int i, int j,
                                          const Dims& fd) {
#pragma clang fp contract(off)
    const int cnx = fd.nx / 2;
    const int ci = i / 2, cj = j / 2;
    int i0, i1, j0, j1;
    TX wi0, wi1, wj0, wj1;
    if ((i & 1) == 0) { i0 = i > 1 ? ci - 1 : ci; i1 = ci; wi0 = TX(0.25); wi1 = TX(0.75); }
    else { i0 = ci; i1 = i + 1 < fd.nx ? ci + 1 : ci; wi0 = TX(0.75); wi1 = TX(0.25); }
    if ((j & 1) == 0) { j0 = j > 1 ? cj - 1 : cj; j1 = cj; wj0 = TX(0.25); wj1 = TX(0.75); }
    else { j0 = cj; j1 = j + 1 < fd.ny ? cj + 1 : cj; wj0 = TX(0.75); wj1 = TX(0.25); }
    const int iI[2] = {i0, i1}, jI[2] = {j0, j1};
    const TX iW[2] = {wi0, wi1}, jW[2] = {wj0, wj1};
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int x = 0; x < 2; ++x) v += iW[x] * jW[y] * coarse[iI[x] + (idx_t)cnx * jI[y]];
    return v;
}

// fine += the prolongation of coarse on one plane, one thread per fine cell.
template <typename TX>
__global__ __launch_bounds__(NT) void k_correct2(const TX* __restrict__ coarse,
                                                 TX* __restrict__ fine, Dims fd,
                                                 const Ctl* __restrict__ ctl) {
    if (ctl && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= fd.nx || j >= fd.ny) return;
    const idx_t f = i + (idx_t)fd.nx * j;
    fine[f] = correct2_at(coarse, fine[f], i, j, fd);
}

// ------------------------------------------------------------- finalize

// Initial PCG scalars: sigma = r.z, rr = r.r, bb = b.b (for the relative
// rule), beta = 0 so the first fused direction update gives d = z.
__global__ __launch_bounds__(kFinThreads) void k_fin_init(Ctl* ctl, const double* p_rz,
                                                          const double* p_rr, int n) {
    ctl += blockIdx.x * kSysCtl;  // one block per system of a batch group
    p_rz += blockIdx.x * kSysPart;
    p_rr += blockIdx.x * kSysPart;
    __shared__ double sh[2 * kFinThreads / kWave];
    const double* const pp[2] = {p_rz, p_rr};
    double rs[2];
    reduce_partials_k<kFinThreads, 2>(pp, n, sh, rs);
    const double rz = rs[0], rr = rs[1];
    if (threadIdx.x == 0) {
        ctl->sigma = rz;
        ctl->sigma_old = rz;
        ctl->rr = rr;
        ctl->bb = rr;
        ctl->beta = 0.0;
        ctl->alpha = 0.0;
        ctl->iter = 0;
        ctl->trigger = 0;
        ctl->refresh = 0;
        ctl->nonfinite = !isfinite(rz) || !isfinite(rr);
        const bool go = (ctl->stop_rule == 1) ? (rr > ctl->rel2 * ctl->bb) : (rz > ctl->tol2);
        ctl->done = !(go && ctl->iter < ctl->max_iter) || ctl->nonfinite;
    }
}

// alpha = sigma / d.q ; Jet's refresh rule for this iteration.
__global__ __launch_bounds__(kFinThreads) void k_fin_alpha(Ctl* ctl, const double* p_dq, int n) {
    ctl += blockIdx.x * kSysCtl;  // one block per system of a batch group
    p_dq += blockIdx.x * kSysPart;
    if (ctl->done) return;
    __shared__ double sh[kFinThreads / kWave];
    const double dq = reduce_partials<kFinThreads>(p_dq, n, sh);
    if (threadIdx.x == 0) {
        ctl->dq = dq;
        ctl->alpha = ctl->sigma / dq;
        ctl->refresh = ctl->trigger || (ctl->iter % 50 == 0 && ctl->iter > 0);
        ctl->trigger = 0;
        if (!isfinite(ctl->alpha)) {
            ctl->nonfinite = 1;
            ctl->done = 1;
        }
    }
}

// End of a PCG iteration (include/jet/detail/cg-inl.h:65-90): sigma_new =
// r.z, trigger on growth, beta, iteration count and the stop test.
__global__ __launch_bounds__(kFinThreads) void k_fin_sigma(Ctl* ctl, const double* p_rz,
                                                           const double* p_rr, int n) {
    ctl += blockIdx.x * kSysCtl;  // one block per system of a batch group
    p_rz += blockIdx.x * kSysPart;
    p_rr += blockIdx.x * kSysPart;
    if (ctl->done) return;
    __shared__ double sh[2 * kFinThreads / kWave];
    const double* const pp[2] = {p_rz, p_rr};
    double rs[2];
    reduce_partials_k<kFinThreads, 2>(pp, n, sh, rs);
    const double rz = rs[0], rr = rs[1];
    if (threadIdx.x == 0) {
        const double sold = ctl->sigma;
        ctl->sigma_old = sold;
        ctl->sigma = rz;
        ctl->rr = rr;
        if (rz > sold) ctl->trigger = 1;
        ctl->beta = rz / sold;
        ctl->iter += 1;
        bool go;
        if (ctl->stop_rule == 1)
            go = rr > ctl->rel2 * ctl->bb;
        else
            go = rz > ctl->tol2;
        if (!isfinite(rz) || !isfinite(rr)) ctl->nonfinite = 1;
        ctl->done = !(go && ctl->iter < ctl->max_iter) || ctl->nonfinite;
    }
}

// Generic: result = sum of partials (used by the BLAS entry points).
__global__ __launch_bounds__(kFinThreads) void k_fin_sum(const double* p, int n, double* out) {
    p += blockIdx.x * kSysPart;  // one block per system of a batch group
    out += blockIdx.x * kSysScal;
    __shared__ double sh[kFinThreads / kWave];
    const double s = reduce_partials<kFinThreads>(p, n, sh);
    if (threadIdx.x == 0) *out = s;
}

}  // namespace jg
