// jg_plane2.h -- two red-black sweeps of a one-plane (2-D) level per launch.
//
// FdmGaussSeidelSolver2::relaxRedBlack (src/jet/fdm_gauss_seidel_solver2.cpp:
// 147-187) applied twice in one colour order, in one pass over HBM, with
// the V-cycle's neighbours fused in: the residual b - A x of the result
// (the restriction's input, FdmBlas2::residual's terms) after the pre-
// smoothing pair, or FdmMgUtils2::correct's prolonged coarse correction
// (src/jet/fdm_mg_linear_system2.cpp:86-143) before the post-smoothing pair.
//
// A block of NW waves owns a 128 x (NW RY) tile of cells.  Lane t of wave w
// keeps the cells of columns 2t, 2t + 1 and rows w RY .. w RY + RY - 1 in
// registers: x, b, A.c (FP32: omega / A.c), A.r, A.u, and the -x coupling of
// column 0 (the left lane's A.r).  Tile origins have an even coordinate sum
// and RY is even, so in every row the cell of a given colour sits in a fixed
// column of the lane's pair: each lane updates exactly one cell per row and
// half-sweep, and no lane idles.  Neighbours come from the lane's own
// registers, from the lanes either side (a DPP shift of the whole wave) and,
// for the first and last rows of a wave's strip, from the waves above and
// below through LDS (one barrier per half-sweep).  The valid region shrinks
// by one cell per half-sweep from the tile edge, so the halo is 4 cells (5
// with the residual), recomputed rather than exchanged; cells nearer the
// edge may hold garbage that never reaches the written interior.
//
// Each update is k_rb_sweep_plane's expression and the residual is
// k_residual_plane's, operand for operand with contraction off, so one
// launch equals two k_rb_sweep_plane launches (and k_residual_plane, and a
// preceding k_correct2) bit for bit.  The launch reads xin and writes xout
// (xin != xout): neighbouring tiles read each other's cells as halo.
#pragma once

#include <type_traits>

namespace jg {

constexpr int kP2Lanes = 64;       // lanes per tile row
constexpr int kP2X = 2 * kP2Lanes;  // tile width in cells
// Waves per block (NW) and rows per lane (RY) of the sweep kernel (P2Cfg)
// and of the direction-update SpMV (SpCfg): 8 x 4, a 128 x 32 tile, in both
// precisions.  A block loads its tile, then relaxes it between barriers, so
// a CU must hold several blocks for one block's loads to overlap another's
// half-sweeps: WPE (waves per SIMD for two blocks) bounds the VGPRs.  On the
// 4096^2 MGPCG, 16 x 4 / 16 x 6 tiles (one block per CU) took 34.4 ms FP64
// and 31.3 ms mixed, 8 x 4 takes 31.1 and 25.8; taller FP32 strips (8 x 6,
// 8 x 8) and shorter ones (8 x 2) are slower, as is 16 x 2 in FP64.
template <typename T>
struct P2Cfg {
    static constexpr int NW = 8;
    static constexpr int RY = 4;
    static constexpr int TY = NW * RY;
    static constexpr int WPE = 2 * NW / 4;
};
template <typename T>
struct SpCfg {
    static constexpr int NW = 8;
    static constexpr int RY = 4;
    static constexpr int TY = NW * RY;
    static constexpr int WPE = 2 * NW / 4;
};
constexpr int p2_halo(bool res) { return res ? 5 : 4; }

// FIRST: the colour updated first in both sweeps ((i + j) & 1 == FIRST).
// XIN: x starts from xin (else from zero).  CORR: the prolongation of
// coarse is added to the starting x.  RES: the residual of the result is
// written to res.  DOT: one partial of b.x (PCG's r.z when b is r and x is
// z) to p_bx and of b.b to p_bb per block.
template <typename T, int FIRST, bool XIN, bool RES, bool CORR, bool DOT = false>
__global__ __launch_bounds__(kP2Lanes * P2Cfg<T>::NW, P2Cfg<T>::WPE) void k_rb_sweep2_plane(
    const Row<T>* __restrict__ A, const T* __restrict__ b, const T* __restrict__ xin,
    const T* __restrict__ coarse, T* __restrict__ xout, T* __restrict__ res, Dims d, T omega,
    const Ctl* __restrict__ ctl, double* __restrict__ p_bx = nullptr,
    double* __restrict__ p_bb = nullptr) {
#pragma clang fp contract(off)
    constexpr bool kW = sizeof(T) == 4;
    constexpr int RY = P2Cfg<T>::RY, TY = P2Cfg<T>::TY, NW = P2Cfg<T>::NW, H = p2_halo(RES);
    constexpr int IX = kP2X - 2 * H, IY = TY - 2 * H;
    constexpr bool kZeroStart = !XIN && !CORR;
    static_assert(RY % 2 == 0 && IX % 2 == 0 && IY % 2 == 0, "colour columns must stay fixed");
    if (ctl && ctl->done) return;
    // x of the first and last row of every wave's strip (by column parity,
    // then lane), and A.u of the last row
    __shared__ T s_top[NW][2][kP2Lanes];
    __shared__ T s_bot[NW][2][kP2Lanes];
    __shared__ T s_u[NW][2][kP2Lanes];
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    // (gx0, gy0): the lane's cell in column 0, row 0 of its strip; the tile
    // origin (bx IX - H, by IY - H) has an even coordinate sum, so cell (j, k)
    // has colour (j + k) & 1
    const int gx0 = bx * IX - H + 2 * t;
    const int gy0 = by * IY - H + w * RY;
    const int wd = w > 0 ? w - 1 : 0, wu = w + 1 < NW ? w + 1 : w;
    T x[2][RY], cw[2][RY], bb[2][RY], rr[2][RY], uu[2][RY], cc[2][RY];
    T rl[RY];  // -x coupling of column 0: A.r of the left lane's column 1
    T ud[2];   // -y coupling of row 0: A.u of the lower wave's last row
    bool hl[2], hr[2], hd[RY], hu[RY];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        hl[j] = gx0 + j > 0;
        hr[j] = gx0 + j + 1 < d.nx;
    }
#pragma unroll
    for (int k = 0; k < RY; ++k) {
        const int gy = gy0 + k;
        hd[k] = gy > 0;
        hu[k] = gy + 1 < d.ny;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int gx = gx0 + j;
            T xv = T(0), c = T(0), bv = T(0), r = T(0), u = T(0);
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny) {
                const idx_t i = gx + (idx_t)d.nx * gy;
                const Row<T> row = A[i];
                c = row.c;
                r = row.r;
                u = row.u;
                bv = b[i];
                if constexpr (XIN) xv = xin[i];
                if constexpr (CORR) xv = correct2_at(coarse, xv, gx, gy, d);
            }
            x[j][k] = xv;
            bb[j][k] = bv;
            rr[j][k] = r;
            uu[j][k] = u;
            cc[j][k] = c;
            cw[j][k] = kW ? (c == T(0) ? T(0) : omega / c) : c;  // w = omega / c
        }
    }
#pragma unroll
    for (int k = 0; k < RY; ++k) rl[k] = dpp_shift<kDppFromLeft>(rr[1][k]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        s_u[w][j][t] = uu[j][RY - 1];
        s_top[w][j][t] = x[j][0];
        s_bot[w][j][t] = x[j][RY - 1];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; ++j) ud[j] = s_u[wd][j][t];

    // neighbours of cell (j, k): x to the left / right / below / above and
    // the -x / -y couplings
    auto x_left = [&](int j, int k) __attribute__((always_inline)) {
        return j == 0 ? dpp_shift<kDppFromLeft>(x[1][k]) : x[0][k];
    };
    auto x_right = [&](int j, int k) __attribute__((always_inline)) {
        return j == 1 ? dpp_shift<kDppFromRight>(x[0][k]) : x[1][k];
    };
    auto x_down = [&](int j, int k) __attribute__((always_inline)) {
        return k > 0 ? x[j][k - 1] : s_bot[wd][j][t];
    };
    auto x_up = [&](int j, int k) __attribute__((always_inline)) {
        return k + 1 < RY ? x[j][k + 1] : s_top[wu][j][t];
    };
    auto c_left = [&](int j, int k) __attribute__((always_inline)) {
        return j == 0 ? rl[k] : rr[0][k];
    };
    auto c_down = [&](int j, int k) __attribute__((always_inline)) {
        return k > 0 ? uu[j][k - 1] : ud[j];
    };

    // one half-sweep: every cell of colour C, then publish the strip ends
    auto half = [&](auto colour, auto zero) __attribute__((always_inline)) {
        constexpr int C = decltype(colour)::value;
#pragma unroll
        for (int k = 0; k < RY; ++k) {
            const int j = (C + k) & 1;
            T v;
            if constexpr (decltype(zero)::value) {
                // every neighbour is still zero
                v = cw[j][k] == T(0) ? T(0)
                    : kW             ? cw[j][k] * (bb[j][k] - T(0))
                                     : omega * (bb[j][k] - T(0)) / cw[j][k];
            } else {
                const T xl = x_left(j, k), xr = x_right(j, k), xd = x_down(j, k), xu = x_up(j, k);
                T s = (hl[j] ? c_left(j, k) * xl : T(0)) + (hr[j] ? rr[j][k] * xr : T(0));
                s += hd[k] ? c_down(j, k) * xd : T(0);
                s += hu[k] ? uu[j][k] * xu : T(0);
                const T xo = x[j][k];
                v = cw[j][k] == T(0) ? T(0)
                    : kW             ? (T(1) - omega) * xo + cw[j][k] * (bb[j][k] - s)
                                     : (T(1) - omega) * xo + omega * (bb[j][k] - s) / cw[j][k];
            }
            x[j][k] = v;
        }
        s_top[w][C & 1][t] = x[C & 1][0];
        s_bot[w][(C + RY - 1) & 1][t] = x[(C + RY - 1) & 1][RY - 1];
        __syncthreads();
    };
    using C0 = std::integral_constant<int, FIRST>;
    using C1 = std::integral_constant<int, 1 - FIRST>;
    half(C0{}, std::integral_constant<bool, kZeroStart>{});
    half(C1{}, std::false_type{});
    half(C0{}, std::false_type{});
    half(C1{}, std::false_type{});

    // the interior: x, and the residual in k_residual_plane's order
    const int lx0 = 2 * t, ly0 = w * RY;
    double a_bx = 0.0, a_bb = 0.0;
#pragma unroll
    for (int k = 0; k < RY; ++k) {
        const int ly = ly0 + k, gy = gy0 + k;
        const bool row_in = ly >= H && ly < TY - H && gy < d.ny;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            T rv = T(0);
            if constexpr (RES) {
                const T xl = x_left(j, k), xr = x_right(j, k), xd = x_down(j, k), xu = x_up(j, k);
                rv = bb[j][k] - cc[j][k] * x[j][k];
                rv -= hl[j] ? c_left(j, k) * xl : T(0);
                rv -= hr[j] ? rr[j][k] * xr : T(0);
                rv -= hd[k] ? c_down(j, k) * xd : T(0);
                rv -= hu[k] ? uu[j][k] * xu : T(0);
            }
            const int lx = lx0 + j, gx = gx0 + j;
            if (row_in && lx >= H && lx < kP2X - H && gx < d.nx) {
                const idx_t i = gx + (idx_t)d.nx * gy;
                xout[i] = x[j][k];
                if constexpr (RES) res[i] = rv;
                if constexpr (DOT) {
                    const double bv = (double)bb[j][k];
                    a_bx += bv * (double)x[j][k];
                    a_bb += bv * bv;
                }
            }
        }
    }
    if constexpr (DOT)
        block_partial2<kP2Lanes * NW>(a_bx, a_bb, p_bx, p_bb, bx + (idx_t)gridDim.x * by);
}

// q = A d with d = z + beta d_old on one plane (FdmBlas2::mvm's terms in
// k_dir_spmv's order: centre, -x, +x, -y, +y; d as st_val forms it), d
// written to dnew, one partial of d.q per block.  The lane layout of
// k_rb_sweep2_plane without a halo: lateral neighbours from the lanes
// either side, vertical ones from the adjacent waves through LDS; the
// block's first and last lanes and waves load the neighbour beyond the
// tile from memory.
template <typename T>
__global__ __launch_bounds__(kP2Lanes * SpCfg<T>::NW, SpCfg<T>::WPE) void k_dir_spmv_plane(
    const Row<T>* __restrict__ A, const T* __restrict__ z, const T* __restrict__ dold,
    T* __restrict__ dnew, T* __restrict__ q, Dims d, double* __restrict__ partial,
    const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    constexpr int RY = SpCfg<T>::RY, TY = SpCfg<T>::TY, NW = SpCfg<T>::NW;
    if (ctl->done) return;
    const double beta = ctl->beta;
    __shared__ T s_top[NW][2][kP2Lanes];
    __shared__ T s_bot[NW][2][kP2Lanes];
    __shared__ T s_u[NW][2][kP2Lanes];
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int gx0 = bx * kP2X + 2 * t, gy0 = by * TY + w * RY;
    const int wd = w > 0 ? w - 1 : 0, wu = w + 1 < NW ? w + 1 : w;
    auto dval = [&](idx_t i) __attribute__((always_inline)) {
        return (T)((double)z[i] + beta * (double)dold[i]);
    };
    T dv[2][RY], cc[2][RY], rr[2][RY], uu[2][RY];
    T rl[RY], dl[RY], dr[RY];  // -x coupling and d to the left of column 0, d right of column 1
    T ud[2], dd[2], du[2];     // -y coupling and d below row 0, d above row RY - 1
    bool hl[2], hr[2], hd[RY], hu[RY], in[2][RY];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        hl[j] = gx0 + j > 0;
        hr[j] = gx0 + j + 1 < d.nx;
    }
#pragma unroll
    for (int k = 0; k < RY; ++k) {
        const int gy = gy0 + k;
        hd[k] = gy > 0;
        hu[k] = gy + 1 < d.ny;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            in[j][k] = gx0 + j < d.nx && gy < d.ny;
            T c = T(0), r = T(0), u = T(0), v = T(0);
            if (in[j][k]) {
                const idx_t i = gx0 + j + (idx_t)d.nx * gy;
                const Row<T> row = A[i];
                c = row.c;
                r = row.r;
                u = row.u;
                v = dval(i);
            }
            cc[j][k] = c;
            rr[j][k] = r;
            uu[j][k] = u;
            dv[j][k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < RY; ++k) {
        rl[k] = dpp_shift<kDppFromLeft>(rr[1][k]);
        dl[k] = dpp_shift<kDppFromLeft>(dv[1][k]);
        dr[k] = dpp_shift<kDppFromRight>(dv[0][k]);
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < RY; ++k)
            if (in[0][k] && hl[0]) {
                const idx_t i = gx0 + (idx_t)d.nx * (gy0 + k);
                rl[k] = A[i - 1].r;
                dl[k] = dval(i - 1);
            }
    }
    if (t == kP2Lanes - 1) {
#pragma unroll
        for (int k = 0; k < RY; ++k)
            if (in[1][k] && hr[1]) dr[k] = dval(gx0 + 1 + (idx_t)d.nx * (gy0 + k) + 1);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        s_u[w][j][t] = uu[j][RY - 1];
        s_top[w][j][t] = dv[j][0];
        s_bot[w][j][t] = dv[j][RY - 1];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        ud[j] = s_u[wd][j][t];
        dd[j] = s_bot[wd][j][t];
        du[j] = s_top[wu][j][t];
    }
    if (w == 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (in[j][0] && hd[0]) {
                const idx_t i = gx0 + j + (idx_t)d.nx * gy0;
                ud[j] = A[i - d.nx].u;
                dd[j] = dval(i - d.nx);
            }
    }
    if (w == NW - 1) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (in[j][RY - 1] && hu[RY - 1]) du[j] = dval(gx0 + j + (idx_t)d.nx * (gy0 + RY));
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < RY; ++k) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const T xl = j == 0 ? dl[k] : dv[0][k], cl = j == 0 ? rl[k] : rr[0][k];
            const T xr = j == 1 ? dr[k] : dv[1][k];
            const T xd = k > 0 ? dv[j][k - 1] : dd[j], cd = k > 0 ? uu[j][k - 1] : ud[j];
            const T xu = k + 1 < RY ? dv[j][k + 1] : du[j];
            T s = cc[j][k] * dv[j][k];
            s += hl[j] ? cl * xl : T(0);
            s += hr[j] ? rr[j][k] * xr : T(0);
            s += hd[k] ? cd * xd : T(0);
            s += hu[k] ? uu[j][k] * xu : T(0);
            if (in[j][k]) {
                const idx_t i = gx0 + j + (idx_t)d.nx * (gy0 + k);
                q[i] = s;
                dnew[i] = dv[j][k];
                acc += (double)dv[j][k] * (double)s;
            }
        }
    }
    block_partial<kP2Lanes * NW>(acc, partial, bx + (idx_t)gridDim.x * by);
}

// The coarsest level of a 2-D V-cycle in one workgroup: fwd forward sweeps
// (first colour 0) then fwd backward ones (first colour 1) from x = 0, each
// half-sweep k_rb_sweep_plane's expression, so the result equals the 2 fwd
// single-sweep launches bit for bit.  The level (at most kCoarseMaxCells
// cells) lives in LDS; thread t holds cells t, t + 1024, ... with their
// couplings in registers; one barrier per half-sweep.
constexpr int kCoarseThreads = 1024, kCoarseMC = 4;
constexpr int kCoarseMaxCells = kCoarseThreads * kCoarseMC;

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_coarsest_plane(
    const Row<T>* __restrict__ A, const T* __restrict__ b, T* __restrict__ xout, Dims d, int fwd,
    T omega, const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    constexpr bool kW = sizeof(T) == 4;
    constexpr int MC = kCoarseMC;
    if (ctl && ctl->done) return;
    __shared__ T sx[kCoarseMaxCells];
    const int t = threadIdx.x;
    const int n = d.nx * d.ny;
    T xv[MC], cw[MC], bb[MC], rr[MC], uu[MC], rl[MC], ud[MC];
    bool hl[MC], hr[MC], hd[MC], hu[MC], ok[MC];
    int par[MC];
#pragma unroll
    for (int m = 0; m < MC; ++m) {
        const int c = t + kCoarseThreads * m;
        ok[m] = c < n;
        const int j = ok[m] ? c / d.nx : 0, i = ok[m] ? c - j * d.nx : 0;
        hl[m] = i > 0;
        hr[m] = i + 1 < d.nx;
        hd[m] = j > 0;
        hu[m] = j + 1 < d.ny;
        par[m] = (i + j) & 1;
        T c0 = T(0), bv = T(0), r = T(0), u = T(0), l = T(0), dn = T(0);
        if (ok[m]) {
            const Row<T> row = A[c];
            c0 = row.c;
            r = row.r;
            u = row.u;
            bv = b[c];
            if (hl[m]) l = A[c - 1].r;
            if (hd[m]) dn = A[c - d.nx].u;
            sx[c] = T(0);
        }
        xv[m] = T(0);
        cw[m] = kW ? (c0 == T(0) ? T(0) : omega / c0) : c0;
        bb[m] = bv;
        rr[m] = r;
        uu[m] = u;
        rl[m] = l;
        ud[m] = dn;
    }
    __syncthreads();
    auto half = [&](int colour, bool zero) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < MC; ++m) {
            if (!ok[m] || par[m] != colour) continue;
            const int c = t + kCoarseThreads * m;
            T v;
            if (cw[m] == T(0)) {
                v = T(0);  // zero-centre row
            } else if (zero) {
                v = kW ? cw[m] * (bb[m] - T(0)) : omega * (bb[m] - T(0)) / cw[m];
            } else {
                // (indices clamped to the cell itself where the neighbour is absent)
                const T xl = sx[hl[m] ? c - 1 : c], xr = sx[hr[m] ? c + 1 : c];
                const T xd = sx[hd[m] ? c - d.nx : c], xu = sx[hu[m] ? c + d.nx : c];
                T s = (hl[m] ? rl[m] * xl : T(0)) + (hr[m] ? rr[m] * xr : T(0));
                s += hd[m] ? ud[m] * xd : T(0);
                s += hu[m] ? uu[m] * xu : T(0);
                v = kW ? (T(1) - omega) * xv[m] + cw[m] * (bb[m] - s)
                       : (T(1) - omega) * xv[m] + omega * (bb[m] - s) / cw[m];
            }
            xv[m] = v;
            sx[c] = v;
        }
        __syncthreads();
    };
    for (int k = 0; k < 2 * fwd; ++k) {
        const int first = k < fwd ? 0 : 1;
        half(first, k == 0);
        half(1 - first, false);
    }
#pragma unroll
    for (int m = 0; m < MC; ++m)
        if (ok[m]) xout[t + kCoarseThreads * m] = xv[m];
}

}  // namespace jg
