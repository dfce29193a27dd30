// jg_sweep.h -- fused red-black sweep: both colours of one Gauss-Seidel/SOR
// sweep in a single pass over HBM.
//
// A 256-thread block owns a 64 x 16 tile of (i, j) positions (60 x 12
// interior plus a 2-cell halo that is recomputed, not exchanged) and marches
// through a chunk of z planes.  x lives in a 4-plane LDS ring; the +x / +y
// couplings (A.r, A.u) of the planes in flight live in LDS rings as well
// (they are the -x / -y couplings of the neighbours); A.c, A.f and b stay in
// registers.  At step t plane t is loaded, the first colour is updated on
// plane t-1 (interior + 1-cell ring, which needs only original values of the
// other colour) and the second colour on plane t-2 (interior), whose values
// are then final and written back.  Each colour update is Jet's
// relaxRedBlack expression (src/jet/fdm_gauss_seidel_solver3.cpp:163-190) in
// Jet's operand order with contraction off, so one forward sweep is
// bit-identical to Jet's relaxRedBlack.
//
// The sweep reads xin and writes xout (xin != xout): tiles overlap, so a
// block must never overwrite values a neighbouring block may still load as
// its halo.
//
// Options: ZERO treats the incoming x as zero (no read of xin); CORRECT adds
// the trilinear prolongation of a coarse correction to x as it is loaded
// (FdmMgUtils3::correct, src/jet/fdm_mg_linear_system3.cpp:100-179); DOT
// accumulates b.x and b.b of the written x (r.z and r.r of PCG when b = r,
// x = z) into two partial slots.
#pragma once

#include "jg_device.h"

namespace jg {

constexpr int kSwLX = 64, kSwTY = 8, kSwRows = 2, kSwLY = kSwTY * kSwRows, kSwH = 2;
constexpr int kSwIX = kSwLX - 2 * kSwH;  // 60 interior columns
constexpr int kSwIY = kSwLY - 2 * kSwH;  // 12 interior rows
constexpr int kSwThreads = kSwLX * kSwTY;

// Prolongation value at fine (i, j, k) (clamped trilinear 1/4-3/4, with the
// k clamp bounded by min(ny, nz) like k_correct).
template <typename T>
__device__ __forceinline__ T prolong_at(const T* __restrict__ coarse, int i, int j, int k,
                                        const Dims& fd) {
#pragma clang fp contract(off)
    const int cnx = fd.nx / 2, cny = fd.ny / 2;
    const int kclamp = min(fd.ny, fd.nz);
    const int ci = i / 2, cj = j / 2, ck = k / 2;
    int iI[2], jI[2], kI[2];
    T iW[2], jW[2], kW[2];
    if ((i & 1) == 0) { iI[0] = i > 1 ? ci - 1 : ci; iI[1] = ci; iW[0] = T(0.25); iW[1] = T(0.75); }
    else { iI[0] = ci; iI[1] = i + 1 < fd.nx ? ci + 1 : ci; iW[0] = T(0.75); iW[1] = T(0.25); }
    if ((j & 1) == 0) { jI[0] = j > 1 ? cj - 1 : cj; jI[1] = cj; jW[0] = T(0.25); jW[1] = T(0.75); }
    else { jI[0] = cj; jI[1] = j + 1 < fd.ny ? cj + 1 : cj; jW[0] = T(0.75); jW[1] = T(0.25); }
    if ((k & 1) == 0) { kI[0] = k > 1 ? ck - 1 : ck; kI[1] = ck; kW[0] = T(0.25); kW[1] = T(0.75); }
    else { kI[0] = ck; kI[1] = k + 1 < kclamp ? ck + 1 : ck; kW[0] = T(0.75); kW[1] = T(0.25); }
    T v = 0;
#pragma unroll
    for (int z = 0; z < 2; ++z)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int x = 0; x < 2; ++x)
                v += iW[x] * jW[y] * kW[z] * coarse[iI[x] + (idx_t)cnx * (jI[y] + (idx_t)cny * kI[z])];
    return v;
}

template <typename T, int FIRST, bool ZERO, bool CORRECT, bool DOT>
__global__ __launch_bounds__(kSwThreads) void k_rb_sweep(const Row<T>* __restrict__ A,
                                                         const T* __restrict__ b,
                                                         const T* __restrict__ xin,
                                                         T* __restrict__ xout, Dims d, int zc,
                                                         T omega, const T* __restrict__ coarse,
                                                         double* __restrict__ p_bx,
                                                         double* __restrict__ p_bb,
                                                         const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    if (ctl && ctl->done) return;
    __shared__ T sx[4][kSwLY][kSwLX];  // x, planes t .. t-3 (slot = plane & 3)
    __shared__ T sr[3][kSwLY][kSwLX];  // A.r, planes t .. t-2 (rotating slots)
    __shared__ T su[3][kSwLY][kSwLX];  // A.u

    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int gx = bx * kSwIX - kSwH + tx;
    const int gy0 = by * kSwIY - kSwH + ty;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    const idx_t sxy = (idx_t)d.nx * d.ny;
    const bool xok = gx >= 0 && gx < d.nx;

    // Each thread owns two positions, rows ly = ty and ty + kSwTY.  Per
    // position: c and b of planes t, t-1, t-2, f of planes t .. t-3
    // (committed) and the prefetched plane t+1.  Plain structs (no arrays)
    // keep everything in registers.
    struct Pos {
        T c0, c1, c2, f0, f1, f2, f3, b0, b1, b2;
        T pc, pf, pr, pu, pb, px;
        int ly, gy;
    };
    Pos P0, P1;
    P0.ly = ty;
    P1.ly = ty + kSwTY;
    P0.gy = gy0;
    P1.gy = gy0 + kSwTY;
    auto init = [&](Pos& P) __attribute__((always_inline)) {
        P.c0 = P.c1 = P.c2 = T(1);
        P.f0 = P.f1 = P.f2 = P.f3 = T(0);
        P.b0 = P.b1 = P.b2 = T(0);
#pragma unroll
        for (int r = 0; r < 4; ++r) sx[r][P.ly][tx] = T(0);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            sr[r][P.ly][tx] = T(0);
            su[r][P.ly][tx] = T(0);
        }
    };
    // Branch-free: lanes outside the domain load cell 0 (always valid) and
    // select the neutral values (c = 1, everything else 0).
    auto fetch = [&](Pos& P, int t) __attribute__((always_inline)) {
        const bool ok = xok && t >= 0 && t < d.nz && P.gy >= 0 && P.gy < d.ny;
        const idx_t c = ok ? gx + (idx_t)d.nx * P.gy + sxy * t : 0;
        const Row<T> row = A[c];
        const T bv = b[c];
        T xv = ZERO ? T(0) : xin[c];
        if (CORRECT) xv += prolong_at<T>(coarse, ok ? gx : 0, ok ? P.gy : 0, ok ? t : 0, d);
        P.pc = ok ? row.c : T(1);
        P.pf = ok ? row.f : T(0);
        P.pr = ok ? row.r : T(0);
        P.pu = ok ? row.u : T(0);
        P.pb = ok ? bv : T(0);
        P.px = ok ? xv : T(0);
    };
    auto commit = [&](Pos& P, int st, int s0) __attribute__((always_inline)) {
        P.c2 = P.c1;
        P.c1 = P.c0;
        P.c0 = P.pc;
        P.f3 = P.f2;
        P.f2 = P.f1;
        P.f1 = P.f0;
        P.f0 = P.pf;
        P.b2 = P.b1;
        P.b1 = P.b0;
        P.b0 = P.pb;
        sx[st][P.ly][tx] = P.px;
        sr[s0][P.ly][tx] = P.pr;
        su[s0][P.ly][tx] = P.pu;
    };
    // First colour on plane p (interior + 1-cell ring); r/u slot s1.
    auto first_colour = [&](Pos& P, int p, int s1) __attribute__((always_inline)) {
        const int ly = P.ly;
        const bool act = tx >= 1 && tx < kSwLX - 1 && ly >= 1 && ly < kSwLY - 1 && xok &&
                         P.gy >= 0 && P.gy < d.ny && ((gx + P.gy + p) & 1) == FIRST;
        if (!act) return;
        const int sp = p & 3, sm = (p - 1) & 3, sn = (p + 1) & 3;
        if (ZERO && !CORRECT) {
            // x is zero everywhere before this colour: (1-w)0 + w(b - 0)/c
            sx[sp][ly][tx] = omega * (P.b1 - T(0)) / P.c1;
        } else {
            const T xo = sx[sp][ly][tx];
            T s = sr[s1][ly][tx - 1] * sx[sp][ly][tx - 1];
            s += sr[s1][ly][tx] * sx[sp][ly][tx + 1];
            s += su[s1][ly - 1][tx] * sx[sp][ly - 1][tx];
            s += su[s1][ly][tx] * sx[sp][ly + 1][tx];
            s += P.f2 * sx[sm][ly][tx];
            s += P.f1 * sx[sn][ly][tx];
            sx[sp][ly][tx] = (T(1) - omega) * xo + omega * (P.b1 - s) / P.c1;
        }
    };
    double acc = 0.0, acc2 = 0.0;
    // Second colour on plane q (interior), write back; r/u slot s2.
    auto second_colour = [&](Pos& P, int q, int s2) __attribute__((always_inline)) {
        const int ly = P.ly;
        const bool act = tx >= kSwH && tx < kSwLX - kSwH && ly >= kSwH && ly < kSwLY - kSwH &&
                         xok && P.gy >= 0 && P.gy < d.ny;
        if (!act) return;
        const int sq = q & 3, sm = (q - 1) & 3, sn = (q + 1) & 3;
        T xv = sx[sq][ly][tx];
        if (((gx + P.gy + q) & 1) != FIRST) {
            T s = sr[s2][ly][tx - 1] * sx[sq][ly][tx - 1];
            s += sr[s2][ly][tx] * sx[sq][ly][tx + 1];
            s += su[s2][ly - 1][tx] * sx[sq][ly - 1][tx];
            s += su[s2][ly][tx] * sx[sq][ly + 1][tx];
            s += P.f3 * sx[sm][ly][tx];
            s += P.f2 * sx[sn][ly][tx];
            xv = (T(1) - omega) * xv + omega * (P.b2 - s) / P.c2;
            sx[sq][ly][tx] = xv;
        }
        xout[gx + (idx_t)d.nx * P.gy + sxy * q] = xv;
        if (DOT) {
            acc += (double)P.b2 * (double)xv;
            acc2 += (double)P.b2 * (double)P.b2;
        }
    };

    init(P0);
    init(P1);
    int s0 = 0, s1 = 2, s2 = 1;  // r/u slots of planes t, t-1, t-2
    fetch(P0, k0 - 2);
    fetch(P1, k0 - 2);
    __syncthreads();

    for (int t = k0 - 2; t <= k1 + 1; ++t) {
        // ---- commit plane t (prefetched), then prefetch plane t+1
        {
            const int tmp = s2;
            s2 = s1;
            s1 = s0;
            s0 = tmp;
        }
        commit(P0, t & 3, s0);
        commit(P1, t & 3, s0);
        if (t + 1 <= k1 + 1) {
            fetch(P0, t + 1);
            fetch(P1, t + 1);
        }
        __syncthreads();
        // ---- first colour on plane t - 1
        if (t - 1 >= 0 && t - 1 < d.nz) {
            first_colour(P0, t - 1, s1);
            first_colour(P1, t - 1, s1);
        }
        __syncthreads();
        // ---- second colour on plane t - 2, final values written back
        if (t - 2 >= k0 && t - 2 < k1) {
            second_colour(P0, t - 2, s2);
            second_colour(P1, t - 2, s2);
        }
        __syncthreads();
    }
    if (DOT) block_partial2<kSwThreads>(acc, acc2, p_bx, p_bb);
}

}  // namespace jg
