// jg_sweep.h -- fused red-black sweep: both colours of one Gauss-Seidel/SOR
// sweep in a single pass over HBM.
//
// A block owns a 64 x 16 tile of (i, j) positions (60 x 12
// interior plus a 2-cell halo that is recomputed, not exchanged) and marches
// through a chunk of z planes.  x lives in a 4-plane LDS ring; the +x / +y
// couplings (A.r, A.u) of the two planes being relaxed live in 2-slot LDS
// rings (they are the -x / -y couplings of the neighbours); A.c, A.f and b
// stay in registers.  At step t plane t is committed from the registers it
// was prefetched into (and plane t+1 is prefetched), the first colour is
// updated on plane t-1 (interior + 1-cell ring, which needs only original
// values of the other colour) and the second colour on plane t-2 (interior),
// whose values are then final and written back.  Each colour update is
// Jet's relaxRedBlack expression (src/jet/fdm_gauss_seidel_solver3.cpp:
// 163-190) in Jet's operand order with contraction off, so one forward sweep
// is bit-identical to Jet's relaxRedBlack.
//
// The sweep reads xin and writes xout (xin != xout): tiles overlap, so a
// block must never overwrite values a neighbouring block may still load as
// its halo.
//
// A row with a zero centre (a fluid cell whose neighbours are all solid,
// which Jet's single-phase build leaves as an all-zero row with b = 0, and
// which Jet's smoothers divide by) is set to zero instead, so the smoother
// stays a symmetric, finite operator.
//
// FP64 colour updates are Jet's expression; FP32 ones (the mixed-precision
// inner solve) use rb_sum6 / rb_coef / rb_update below: the same operator,
// rounded differently.
//
// Options: ZERO treats the incoming x as zero (no read of xin); CORRECT adds
// the trilinear prolongation of a coarse correction to x as it is loaded
// (the operator of FdmMgUtils3::correct, src/jet/fdm_mg_linear_system3.cpp:
// 100-179, evaluated separably: the block interpolates each coarse plane pair
// in z into an LDS tile, then every position interpolates that tile in x and
// y); DOT accumulates b.x and b.b of the written x (r.z and r.r of PCG when
// b = r, x = z) into two partial slots.
#pragma once

#include "jg_device.h"

namespace jg {

// Smoother arithmetic shared by the sweep kernels (single sweep, two sweeps,
// bricks, bottom), so that their results stay bit-identical to each other.
// FP64 is Jet's relaxRedBlack update (src/jet/fdm_gauss_seidel_solver3.cpp:
// 163-190): the neighbour sum -x, +x, -y, +y, -z, +z, then
// (1 - omega) x + omega (b - s) / c, contraction off.  FP32 (the mixed-
// precision preconditioner) is the same operator rounded for throughput: the
// sum in the same order as a chain of fused multiply-adds, w = omega / c from
// the hardware reciprocal (v_rcp_f32, within 1 ulp) once per cell, and
// fma(w, b - s, (1 - omega) x).
template <typename T>
__device__ __forceinline__ T rb_sum6(T a0, T x0, T a1, T x1, T a2, T x2, T a3, T x3, T a4, T x4,
                                     T a5, T x5) {
#pragma clang fp contract(off)
    if constexpr (sizeof(T) == 4) {
        T s = a0 * x0;
        s = __builtin_fmaf(a1, x1, s);
        s = __builtin_fmaf(a2, x2, s);
        s = __builtin_fmaf(a3, x3, s);
        s = __builtin_fmaf(a4, x4, s);
        return __builtin_fmaf(a5, x5, s);
    } else {
        T s = a0 * x0;
        s += a1 * x1;
        s += a2 * x2;
        s += a3 * x3;
        s += a4 * x4;
        s += a5 * x5;
        return s;
    }
}
// The per-cell coefficient a sweep keeps: FP32 w = omega / c (0 for a
// zero-centre row), FP64 c itself.
template <typename T>
__device__ __forceinline__ T rb_coef(T omega, T c) {
    if constexpr (sizeof(T) == 4)
        return c == T(0) ? T(0) : omega * __builtin_amdgcn_rcpf(c);
    else
        return c;
}
// The relaxed value from the old value xo, the coefficient of rb_coef, b and
// the neighbour sum s (callers handle zero-centre rows).
template <typename T>
__device__ __forceinline__ T rb_update(T omega, T xo, T cw, T b, T s) {
#pragma clang fp contract(off)
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf(cw, b - s, (T(1) - omega) * xo);
    else
        return (T(1) - omega) * xo + omega * (b - s) / cw;
}

constexpr int kSwLX = 64, kSwLY = 16, kSwH = 2;  // tile of positions, halo
constexpr int kSwIX = kSwLX - 2 * kSwH;  // 60 interior columns
constexpr int kSwIY = kSwLY - 2 * kSwH;  // 12 interior rows
// Coarse tile behind one block's positions: columns gx0/2 - 1 .. gx0/2 + 32.
constexpr int kCzX = kSwLX / 2 + 2, kCzY = kSwLY / 2 + 2;
static_assert(kCzX * kCzY <= kSwLX * kSwLY / 2, "one coarse entry per thread");

// Prolongation value at fine (i, j, k) (clamped trilinear 1/4-3/4, with the
// k clamp bounded by min(ny, nz) like k_correct), in Jet's accumulation order.
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

// z taps of the prolongation for fine plane t (clamped into the domain).
template <typename T>
__device__ __forceinline__ void prolong_z(int t, const Dims& fd, int& k0, int& k1, T& w0, T& w1) {
    const int tc = min(max(t, 0), fd.nz - 1);
    const int ck = tc / 2;
    if ((tc & 1) == 0) {
        k0 = tc > 1 ? ck - 1 : ck;
        k1 = ck;
        w0 = T(0.25);
        w1 = T(0.75);
    } else {
        k0 = ck;
        k1 = tc + 1 < min(fd.ny, fd.nz) ? ck + 1 : ck;
        w0 = T(0.75);
        w1 = T(0.25);
    }
}

// Launch shapes of the fused sweep: RP tile rows per thread (block of
// 64 x 16/RP threads) and PD planes of loads in flight ahead of the plane
// being committed.  The autotuner picks one per level and precision.
struct SweepShape {
    int rp, pd;
};
constexpr SweepShape kSweepShapes[] = {{2, 1}, {2, 2}, {1, 1}, {1, 2}};
constexpr int kNumSweepShapes = 4;

// Occupancy each shape is compiled for (waves per SIMD): as many blocks per
// CU as the VGPR budget allows without spilling.  The correction variant
// holds the coarse taps too; in FP32 with two rows per thread it gets a
// larger budget (one block per CU fewer) instead of spilling.
constexpr int sweep_waves_per_eu(int tsize, int rp, int pd, bool correct) {
    return rp == 2 ? (pd == 1 ? (tsize == 4 ? (correct ? 5 : 6) : 4)
                              : (tsize == 4 ? (correct ? 4 : 5) : 2))
                   : (tsize == 4 ? 8 : 4);
}

template <typename T, int FIRST, bool ZERO, bool CORRECT, bool DOT, int RP, int PD>
__global__ __launch_bounds__(kSwLX * kSwLY / RP, sweep_waves_per_eu(sizeof(T), RP, PD, CORRECT)) void k_rb_sweep(
    const Row<T>* __restrict__ A, const T* __restrict__ b, const T* __restrict__ xin,
    T* __restrict__ xout, Dims d, int zc, T omega, const T* __restrict__ coarse,
    double* __restrict__ p_bx, double* __restrict__ p_bb, const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    static_assert(RP == 1 || RP == 2, "rows per thread");
    static_assert(PD == 1 || PD == 2, "prefetch depth");
    constexpr int TY = kSwLY / RP;
    constexpr int NTH = kSwLX * TY;
    // FP32: the c ring holds w = omega / c (0 for a zero-centre row) and the
    // update is (1 - omega) x + w (b - s), one division per cell and sweep;
    // FP64 keeps Jet's omega (b - s) / c bit for bit.
    constexpr bool kWeighted = sizeof(T) == 4;
    __shared__ __align__(16) T sx[4][kSwLY][kSwLX];  // x, planes t .. t-3 (slot = plane & 3)
    __shared__ T sr[2][kSwLY][kSwLX];                // A.r, planes t-1, t-2 (slot = plane & 1)
    __shared__ T su[2][kSwLY][kSwLX];                // A.u
    __shared__ T sz[CORRECT ? kCzY : 1][CORRECT ? kCzX : 1];  // coarse plane pair, z-interpolated

    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (d.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        b += sys * n;
        if (!ZERO) xin += sys * n;
        xout += sys * n;
        if (CORRECT) coarse += sys * (n / 8);
        if (DOT) {
            p_bx += sys * kSysPart;
            p_bb += sys * kSysPart;
        }
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int gx0 = bx * kSwIX - kSwH, gy0 = by * kSwIY - kSwH;  // both even
    const int gx = gx0 + tx;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    const idx_t sxy = (idx_t)d.nx * d.ny;
    const bool xok = gx >= 0 && gx < d.nx;

    // CORRECT: thread e < kCzX * kCzY stages coarse column (gx0/2 - 1 + e % kCzX,
    // gy0/2 - 1 + e / kCzX), clamped to the coarse grid (the clamp is exactly
    // the prolongation's edge rule).  Each position then reads the 2 x 2 entries
    // around it: x taps (tx >> 1) + {0, 1} for odd tx, (tx >> 1) + {-1, 0} + 1 for even.
    const int e = tx + kSwLX * ty;
    const bool zact = CORRECT && e < kCzX * kCzY;
    unsigned zoff = 0;  // column offset within a coarse plane (< 2^32 cells)
    const idx_t csz = (idx_t)(d.nx / 2) * (d.ny / 2);
    if (CORRECT) {
        const int ci = min(max(gx0 / 2 - 1 + e % kCzX, 0), d.nx / 2 - 1);
        const int cj = min(max(gy0 / 2 - 1 + e / kCzX, 0), d.ny / 2 - 1);
        zoff = (unsigned)ci + (unsigned)(d.nx / 2) * (unsigned)cj;
    }
    T zc0 = 0, zc1 = 0;  // prefetched coarse pair of the next plane
    auto zfetch = [&](int t) __attribute__((always_inline)) {
        if (!zact) return;
        int c0, c1;
        T w0, w1;
        prolong_z<T>(t, d, c0, c1, w0, w1);
        zc0 = coarse[zoff + csz * c0];
        zc1 = coarse[zoff + csz * c1];
    };
    auto zstore = [&](int t) __attribute__((always_inline)) {
        if (!zact) return;
        int c0, c1;
        T w0, w1;
        prolong_z<T>(t, d, c0, c1, w0, w1);
        sz[e / kCzX][e % kCzX] = w0 * zc0 + w1 * zc1;
    };

    // Each thread owns RP positions, rows ly = ty + TY * r.  Per position: c
    // and b of planes t, t-1, t-2, f of planes t .. t-3, the held r/u of
    // plane t (written to LDS one step later) and PD prefetched planes.
    // Prefetched plane: raw loads (lanes outside the domain load cell 0),
    // masked only when committed, so nothing waits on them before then.
    struct Pre {
        Row<T> row;
        T b, x;
        bool ok;
    };
    struct Pos {
        T c0, c1, c2, f0, f1, f2, f3, b0, b1, b2, hr, hu;
        T px;      // final value of the last second-colour plane, stored one step later
        bool pst;  // px is pending
        Pre p[PD];
        int ly, gy;
    };
    Pos P[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        P[r].ly = ty + TY * r;
        P[r].gy = gy0 + ty + TY * r;
        P[r].c0 = P[r].c1 = P[r].c2 = T(1);
        P[r].f0 = P[r].f1 = P[r].f2 = P[r].f3 = T(0);
        P[r].b0 = P[r].b1 = P[r].b2 = T(0);
        P[r].hr = P[r].hu = T(0);
        P[r].px = T(0);
        P[r].pst = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) sx[s][P[r].ly][tx] = T(0);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            sr[s][P[r].ly][tx] = T(0);
            su[s][P[r].ly][tx] = T(0);
        }
    }
    // Branch-free: lanes outside the domain load cell 0 (always valid); the
    // commit selects the neutral values for them (c = 1, everything else 0).
    auto fetch = [&](Pre& q, int gy, int t) __attribute__((always_inline)) {
        const bool ok = xok && t >= 0 && t < d.nz && gy >= 0 && gy < d.ny;
        const idx_t c = ok ? gx + (idx_t)d.nx * gy + sxy * t : 0;
        q.row = A[c];
        q.b = b[c];
        q.x = ZERO ? T(0) : xin[c];
        q.ok = ok;
    };
    auto commit = [&](Pos& Q, const Pre& q, int t) __attribute__((always_inline)) {
        Q.c2 = Q.c1;
        Q.c1 = Q.c0;
        Q.c0 = q.ok ? q.row.c : T(1);
        Q.c0 = rb_coef<T>(omega, Q.c0);  // FP32: w = omega / c; FP64: c
        Q.f3 = Q.f2;
        Q.f2 = Q.f1;
        Q.f1 = Q.f0;
        Q.f0 = q.ok ? q.row.f : T(0);
        Q.b2 = Q.b1;
        Q.b1 = Q.b0;
        Q.b0 = q.ok ? q.b : T(0);
        T xv = q.ok ? q.x : T(0);
        if (CORRECT) {
            const int ax0 = (tx >> 1) + (tx & 1);  // left x tap in the coarse tile
            const T wx0 = (tx & 1) ? T(0.75) : T(0.25), wx1 = (tx & 1) ? T(0.25) : T(0.75);
            const int ay0 = (Q.ly >> 1) + (Q.ly & 1);
            const T wy0 = (Q.ly & 1) ? T(0.75) : T(0.25), wy1 = (Q.ly & 1) ? T(0.25) : T(0.75);
            const T lo = wx0 * sz[ay0][ax0] + wx1 * sz[ay0][ax0 + 1];
            const T hi = wx0 * sz[ay0 + 1][ax0] + wx1 * sz[ay0 + 1][ax0 + 1];
            const T corr = wy0 * lo + wy1 * hi;
            xv += q.ok ? corr : T(0);
        }
        sx[t & 3][Q.ly][tx] = xv;
        sr[(t - 1) & 1][Q.ly][tx] = Q.hr;
        su[(t - 1) & 1][Q.ly][tx] = Q.hu;
        Q.hr = q.ok ? q.row.r : T(0);
        Q.hu = q.ok ? q.row.u : T(0);
    };
    // First colour on plane p (interior + 1-cell ring).
    auto first_colour = [&](Pos& Q, int p) __attribute__((always_inline)) {
        const int ly = Q.ly;
        const bool act = tx >= 1 && tx < kSwLX - 1 && ly >= 1 && ly < kSwLY - 1 && xok &&
                         Q.gy >= 0 && Q.gy < d.ny && ((gx + Q.gy + p) & 1) == FIRST;
        if (!act) return;
        const int sp = p & 3, sm = (p - 1) & 3, sn = (p + 1) & 3, s1 = p & 1;
        if (Q.c1 == T(0)) {
            sx[sp][ly][tx] = T(0);  // zero-centre row (see the file comment)
        } else if (ZERO && !CORRECT) {
            // x is zero everywhere before this colour: (1-w)0 + w(b - 0)/c
            sx[sp][ly][tx] = kWeighted ? Q.c1 * (Q.b1 - T(0)) : omega * (Q.b1 - T(0)) / Q.c1;
        } else {
            const T xo = sx[sp][ly][tx];
            const T s = rb_sum6<T>(sr[s1][ly][tx - 1], sx[sp][ly][tx - 1], sr[s1][ly][tx],
                                   sx[sp][ly][tx + 1], su[s1][ly - 1][tx], sx[sp][ly - 1][tx],
                                   su[s1][ly][tx], sx[sp][ly + 1][tx], Q.f2, sx[sm][ly][tx], Q.f1,
                                   sx[sn][ly][tx]);
            sx[sp][ly][tx] = rb_update<T>(omega, xo, Q.c1, Q.b1, s);
        }
    };
    double acc = 0.0, acc2 = 0.0;
    // Second colour on plane q (interior); the final value is stored by
    // flush() at the next step, after that step's loads are issued, so the
    // wait for those loads never includes the store.
    auto second_colour = [&](Pos& Q, int q) __attribute__((always_inline)) {
        const int ly = Q.ly;
        const bool act = tx >= kSwH && tx < kSwLX - kSwH && ly >= kSwH && ly < kSwLY - kSwH &&
                         xok && Q.gy >= 0 && Q.gy < d.ny;
        Q.pst = act;
        if (!act) return;
        const int sq = q & 3, sm = (q - 1) & 3, sn = (q + 1) & 3, s2 = q & 1;
        T xv = sx[sq][ly][tx];
        if (((gx + Q.gy + q) & 1) != FIRST) {
            const T s = rb_sum6<T>(sr[s2][ly][tx - 1], sx[sq][ly][tx - 1], sr[s2][ly][tx],
                                   sx[sq][ly][tx + 1], su[s2][ly - 1][tx], sx[sq][ly - 1][tx],
                                   su[s2][ly][tx], sx[sq][ly + 1][tx], Q.f3, sx[sm][ly][tx], Q.f2,
                                   sx[sn][ly][tx]);
            xv = Q.c2 == T(0) ? T(0) : rb_update<T>(omega, xv, Q.c2, Q.b2, s);
            sx[sq][ly][tx] = xv;
        }
        Q.px = xv;
        if (DOT) {
            acc += (double)Q.b2 * (double)xv;
            acc2 += (double)Q.b2 * (double)Q.b2;
        }
    };

    auto flush = [&](Pos& Q, int q) __attribute__((always_inline)) {
        if (Q.pst) xout[gx + (idx_t)d.nx * Q.gy + sxy * q] = Q.px;
        Q.pst = false;
    };

    // Prologue: planes k0-2 .. k0-3+PD in flight, the coarse tile of k0-2.
    if (CORRECT) zfetch(k0 - 2);
#pragma unroll
    for (int s = 0; s < PD; ++s)
#pragma unroll
        for (int r = 0; r < RP; ++r) fetch(P[r].p[s], P[r].gy, k0 - 2 + s);
    if (CORRECT) zstore(k0 - 2);
    __syncthreads();

    // Step t: commit plane t from prefetch slot (t - k0 + 2) % PD and refill
    // that slot with plane t + PD.  The loop is unrolled PD times so every
    // slot index is static: no register moves wait on loads in flight.
    // The step count is rounded up to a multiple of PD so the unrolled body
    // has no exit (which would keep the compiler from counting the loads in
    // flight); a step past k1 + 1 commits a re-read plane and its colour
    // stages touch only planes no later stage reads.
    const int tend = k0 - 2 + (k1 + 4 - k0 + PD - 1) / PD * PD;
    for (int t0 = k0 - 2; t0 < tend; t0 += PD) {
#pragma unroll
        for (int u = 0; u < PD; ++u) {
            const int t = t0 + u;
            // ---- commit plane t (prefetched), then prefetch plane t+PD
#pragma unroll
            for (int r = 0; r < RP; ++r) commit(P[r], P[r].p[u], t);
            if (CORRECT && t + 1 <= k1 + 1) zfetch(t + 1);
            // unconditional (past the end it re-reads plane k1+1, unused) so
            // the compiler can count the loads in flight
#pragma unroll
            for (int r = 0; r < RP; ++r) fetch(P[r].p[u], P[r].gy, min(t + PD, k1 + 1));
            // the final values of plane t - 3 (second colour of the last step)
#pragma unroll
            for (int r = 0; r < RP; ++r) flush(P[r], t - 3);
            __syncthreads();
            // ---- first colour on plane t - 1
            if (t - 1 >= 0 && t - 1 < d.nz) {
#pragma unroll
                for (int r = 0; r < RP; ++r) first_colour(P[r], t - 1);
            }
            __syncthreads();
            // ---- second colour on plane t - 2, final values written back
            if (t - 2 >= k0 && t - 2 < k1) {
#pragma unroll
                for (int r = 0; r < RP; ++r) second_colour(P[r], t - 2);
            }
            // the coarse tile of plane t+1 (read by the next commit)
            if (CORRECT && t + 1 <= k1 + 1) zstore(t + 1);
            __syncthreads();
        }
    }
    // the last second-colour plane, k1 - 1
#pragma unroll
    for (int r = 0; r < RP; ++r) flush(P[r], k1 - 1);
    if (DOT) {
        // Block partials of b.x and b.b; the x ring is free after the last barrier.
        double* sh = reinterpret_cast<double*>(&sx[0][0][0]);
        constexpr int W = NTH / kWave;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            acc += __shfl_down(acc, o, kWave);
            acc2 += __shfl_down(acc2, o, kWave);
        }
        if ((e & (kWave - 1)) == 0) {
            sh[e / kWave] = acc;
            sh[W + e / kWave] = acc2;
        }
        __syncthreads();
        if (e == 0) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                s0 += sh[w];
                s1 += sh[W + w];
            }
            p_bx[part] = s0;
            p_bb[part] = s1;
        }
    }
}

// One red-black sweep of a one-plane (2-D) level: FdmGaussSeidelSolver2::
// relaxRedBlack (src/jet/fdm_gauss_seidel_solver2.cpp:147-187), both colours
// in one pass.  The z-marching k_rb_sweep would spend most of its pipeline on
// filling and draining for a single plane; here a block loads a 64 x 16 tile
// of x and the couplings once into LDS (2-cell halo, recomputed), updates the
// first colour on the interior + 1-cell ring, then the second colour on the
// interior, and writes the interior.  Same conventions as k_rb_sweep: FIRST
// is the colour updated first, ZERO treats xin as zero, zero-centre rows are
// set to zero, FP64 uses Jet's expression and FP32 the w = omega / c form.
// The neighbour sum is Jet's 2-D one, (((-x) + (+x)) + (-y)) + (+y), with
// the terms outside the grid masked as Jet masks them.
template <typename T, int FIRST, bool ZERO>
__global__ __launch_bounds__(kSwLX * kSwLY) void k_rb_sweep_plane(
    const Row<T>* __restrict__ A, const T* __restrict__ b, const T* __restrict__ xin,
    T* __restrict__ xout, Dims d, T omega, const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    constexpr bool kWeighted = sizeof(T) == 4;
    if (ctl && ctl->done) return;
    __shared__ T sx[kSwLY][kSwLX];
    __shared__ T sr[kSwLY][kSwLX];
    __shared__ T su[kSwLY][kSwLX];
    const int tx = threadIdx.x, ly = threadIdx.y;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int gx = bx * kSwIX - kSwH + tx, gy = by * kSwIY - kSwH + ly;
    const bool ok = gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny;
    const idx_t c = ok ? gx + (idx_t)d.nx * gy : 0;
    T xv = T(0), cc = T(0), bb = T(0);
    if (ok) {
        const Row<T> row = A[c];
        if (!ZERO) xv = xin[c];
        cc = row.c;
        bb = b[c];
        sr[ly][tx] = row.r;
        su[ly][tx] = row.u;
        if (kWeighted) cc = cc == T(0) ? T(0) : omega / cc;  // w = omega / c
    } else {
        sr[ly][tx] = T(0);
        su[ly][tx] = T(0);
    }
    sx[ly][tx] = xv;
    __syncthreads();
    const bool hl = gx > 0, hr = gx + 1 < d.nx, hd = gy > 0, hu = gy + 1 < d.ny;
    auto relax = [&](T xo) __attribute__((always_inline)) {
        T s = (hl ? sr[ly][tx - 1] * sx[ly][tx - 1] : T(0)) +
              (hr ? sr[ly][tx] * sx[ly][tx + 1] : T(0));
        s += hd ? su[ly - 1][tx] * sx[ly - 1][tx] : T(0);
        s += hu ? su[ly][tx] * sx[ly + 1][tx] : T(0);
        return kWeighted ? (T(1) - omega) * xo + cc * (bb - s)
                         : (T(1) - omega) * xo + omega * (bb - s) / cc;
    };
    const bool mine = ((gx + gy) & 1) == FIRST;
    // first colour: interior and the 1-cell ring (reads only the other colour)
    if (ok && mine && tx >= 1 && tx < kSwLX - 1 && ly >= 1 && ly < kSwLY - 1) {
        T v;
        if (cc == T(0))
            v = T(0);  // zero-centre row
        else if (ZERO)
            v = kWeighted ? cc * (bb - T(0)) : omega * (bb - T(0)) / cc;
        else
            v = relax(xv);
        xv = v;
        sx[ly][tx] = v;
    }
    __syncthreads();
    // second colour on the interior (reads the updated first colour); the
    // interior of both colours is written
    if (ok && tx >= kSwH && tx < kSwLX - kSwH && ly >= kSwH && ly < kSwLY - kSwH) {
        if (!mine) xv = cc == T(0) ? T(0) : relax(xv);
        xout[c] = xv;
    }
}

}  // namespace jg
