// jg_sweep2.h -- two red-black sweeps in one pass over HBM (temporal
// blocking of the fused sweep in jg_sweep.h).
//
// Two consecutive sweeps with the same colour order are four colour
// half-sweeps h = 1..4 (colours c1, c2, c1, c2).  A block owns a 64 x 32 tile
// of (i, j) positions -- a 56 x 24 interior plus a 4-cell halo that is
// recomputed, not exchanged -- and marches through a chunk of z planes.  At
// step t it commits plane t from the registers it was prefetched into
// (prefetching plane t + 2) and runs h1 on plane t-1 in the same phase (its
// +z neighbour is the thread's own position, just committed by itself;
// everything else it reads is a step old), then h2 on t-2, h3 on t-3 and h4 on t-4, each after a
// barrier: four barriers per step.  Plane t-4 is then final and written
// back.  Half-sweep h is valid on positions at least h cells inside the tile
// (it reads neighbours updated by h-1, valid one cell further out), so the
// interior (ring >= 4) of plane t-4 holds exactly the values two full
// sweeps give.  Each block also recomputes the 4 planes below and above its
// chunk.
//
// x lives in an LDS ring (planes t-5 .. t in 8 slots), the +x / +y couplings in
// 5-plane rings (planes t .. t-4; they are the -x / -y couplings of the
// neighbours; plane t's are written after the step's first barrier); w =
// omega / c, A.f and b stay in registers.  A thread owns two
// vertically adjacent positions, which have opposite colours, so every
// half-sweep updates one position per thread.  Each update is the single
// sweep's FP32 arithmetic (rb_sum6 / rb_coef / rb_update of jg_sweep.h: Jet's
// relaxRedBlack operand order for s, src/jet/fdm_gauss_seidel_solver3.cpp:
// 163-190), so the result is bit-identical to two single FP32 fused sweeps.  Zero-centre rows are set to zero, as in the single sweep.
//
// Options: ZERO (x starts at zero, xin not read), CORRECT (the prolongation
// of a coarse correction is added to x as it is loaded), RES (the residual
// b - A x of the two-sweep result is written too, for the restriction that
// follows pre-smoothing).  RES needs final x one cell around every written
// residual, so its tile has a 5-cell halo (a 54 x 22 interior, 5 extra planes
// below and above the chunk).  The residual of plane q is evaluated in the
// operand order of k_residual_restrict (Jet's FdmBlas3::residual order), so it
// is bit-identical to that kernel's: all terms but the last once plane q is
// final (after h4 on q), the +z term one step later (after h4 on q + 1).
#pragma once

#include <type_traits>

#include "jg_device.h"
#include "jg_sweep.h"

namespace jg {

constexpr int kS2LX = 64, kS2LY = 32;
constexpr int kS2TY = 16;  // thread rows (two tile rows each)
constexpr int kS2Threads = kS2LX * kS2TY;
// Halo (in x, y and z) and interior of the tile: 4 cells for the two sweeps,
// 5 when the residual is written too; one less when x starts at zero with no
// correction (Z), where the first half-sweep, w b, reads no neighbour and is
// valid up to the tile's edge.
template <bool RES, bool Z = false>
struct S2Tile {
    static constexpr int H = (RES ? 5 : 4) - (Z ? 1 : 0);
    static constexpr int IX = kS2LX - 2 * H, IY = kS2LY - 2 * H;
    // r / u coupling ring: planes t .. t-4.  The couplings of plane t are
    // written after the first barrier of step t (h1 on plane t-1 runs before
    // it and reads those of t-1), and RES reads those of plane t-4 after the
    // last barrier of step t, so five slots never collide.
    static constexpr int CR = 5;
    static_assert(IX % 2 == 0 && IY % 2 == 0, "tile interiors stay even");
};
// x ring: planes t-7 .. t are live (8 slots make the slot a mask)
constexpr int kS2XR = 8;
constexpr int kS2CzX = kS2LX / 2 + 2, kS2CzY = kS2LY / 2 + 2;
static_assert(kS2CzX * kS2CzY <= kS2Threads, "one coarse entry per thread");

// FP32 only: the FP64 rings would need 234 KB of LDS.
template <typename T, int FIRST, bool ZERO, bool CORRECT, bool RES>
__global__ __launch_bounds__(kS2Threads, 1) void k_rb_sweep2(
    const Row<T>* __restrict__ A, const T* __restrict__ b, const T* __restrict__ xin,
    T* __restrict__ xout, Dims d, int zc, T omega, const T* __restrict__ coarse,
    T* __restrict__ res, const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    static_assert(sizeof(T) == 4, "two-sweep kernel is FP32 only");
    // CORRECT's prolongation weights take the parity of x and y from the
    // tile-local index, which needs the even origins of the 4-cell halo
    static_assert(!(RES && CORRECT), "the residual variant has no correction");
    // ZH: x starts at zero and nothing is added to it, so half-sweep H is
    // valid on ring >= H - 1 (the first reads no neighbour)
    constexpr int ZH = ZERO && !CORRECT ? 1 : 0;
    using G = S2Tile<RES, ZH != 0>;
    constexpr int kH = G::H, kCR = G::CR;
    if (ctl && ctl->done) return;
    __shared__ __align__(16) T sx[kS2XR][kS2LY][kS2LX];  // x, slot = plane mod 8
    __shared__ T sr[kCR][kS2LY][kS2LX];                  // A.r, slot = plane mod kCR
    __shared__ T su[kCR][kS2LY][kS2LX];                  // A.u
    __shared__ T sz[CORRECT ? kS2CzY : 1][CORRECT ? kS2CzX : 1];

    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int gx0 = bx * G::IX - kH, gy0 = by * G::IY - kH;  // even unless RES
    const int gx = gx0 + tx;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    const idx_t sxy = (idx_t)d.nx * d.ny;
    const bool xok = gx >= 0 && gx < d.nx;
    // planes are >= k0 - 6 >= -6: slot arithmetic on p + 24 stays non-negative
    auto xs = [](int p) { return (unsigned)(p + 24) & (kS2XR - 1); };
    auto cs = [](int p) { return (unsigned)(p + 24) % (unsigned)kCR; };

    // CORRECT: as in k_rb_sweep, thread e < kS2CzX * kS2CzY stages one coarse
    // column, z-interpolated for the plane about to be committed.
    const int e = tx + kS2LX * ty;
    const bool zact = CORRECT && e < kS2CzX * kS2CzY;
    unsigned zoff = 0;
    const idx_t csz = (idx_t)(d.nx / 2) * (d.ny / 2);
    if (CORRECT) {
        const int ci = min(max(gx0 / 2 - 1 + e % kS2CzX, 0), d.nx / 2 - 1);
        const int cj = min(max(gy0 / 2 - 1 + e / kS2CzX, 0), d.ny / 2 - 1);
        zoff = (unsigned)ci + (unsigned)(d.nx / 2) * (unsigned)cj;
    }
    T zc0 = 0, zc1 = 0;
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
        sz[e / kS2CzX][e % kS2CzX] = w0 * zc0 + w1 * zc1;
    };

    struct Pre {
        Row<T> row;
        T b, x;
        bool ok;
    };
    // A thread owns the two vertically adjacent positions (tx, 2 ty) and
    // (tx, 2 ty + 1), which always have opposite colours: every half-sweep
    // updates exactly one of them per thread (no idle lanes).  Per position:
    // w = omega / c (0 for a zero-centre row) and b of planes t .. t-4, f of
    // planes t .. t-5, the held r / u of plane t (written to LDS at the next
    // commit), the pending final value, and the prefetched planes.
    // RES also keeps A.c of planes t .. t-4 and the partial residual of
    // plane t-4 (rv).  To stay within the registers of 1024 threads, RES and
    // CORRECT prefetch one plane ahead instead of two.
    constexpr int kC = RES ? 5 : 1;
    constexpr int kPD = RES || CORRECT ? 1 : 2;
    struct Pos {
        T w[5], bb[5], f[6], cc[kC];
        T hr, hu, px, rv;
        Pre p[kPD];  // planes in flight: slot (t - t0) % kPD holds plane t
        int ly, gy;
        bool in4;  // in the written interior
    };
    Pos P[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        Pos& Q = P[r];
        Q.ly = 2 * ty + r;
        Q.gy = gy0 + Q.ly;
        Q.in4 = tx >= kH && tx < kS2LX - kH && Q.ly >= kH && Q.ly < kS2LY - kH && xok &&
                Q.gy >= 0 && Q.gy < d.ny;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            Q.w[s] = omega;
            Q.bb[s] = T(0);
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) Q.f[s] = T(0);
#pragma unroll
        for (int s = 0; s < kC; ++s) Q.cc[s] = T(0);
        Q.hr = Q.hu = Q.px = Q.rv = T(0);
#pragma unroll
        for (int s = 0; s < kS2XR; ++s) sx[s][Q.ly][tx] = T(0);
#pragma unroll
        for (int s = 0; s < kCR; ++s) {
            sr[s][Q.ly][tx] = T(0);
            su[s][Q.ly][tx] = T(0);
        }
    }
    auto fetch = [&](Pre& q, int gy, int t) __attribute__((always_inline)) {
        const bool ok = xok && t >= 0 && t < d.nz && gy >= 0 && gy < d.ny;
        const idx_t c = ok ? gx + (idx_t)d.nx * gy + sxy * t : 0;
        q.row = A[c];
        q.b = b[c];
        q.x = ZERO ? T(0) : xin[c];
        q.ok = ok;
    };
    auto commit = [&](Pos& Q, const Pre& q, int t) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 4; s > 0; --s) {
            Q.w[s] = Q.w[s - 1];
            Q.bb[s] = Q.bb[s - 1];
        }
#pragma unroll
        for (int s = 5; s > 0; --s) Q.f[s] = Q.f[s - 1];
        const T c = q.ok ? q.row.c : T(1);
        Q.w[0] = rb_coef<T>(omega, c);  // w = omega / c, 0 for a zero-centre row
        if (RES) {
#pragma unroll
            for (int s = kC - 1; s > 0; --s) Q.cc[s] = Q.cc[s - 1];
            Q.cc[0] = q.row.c;
        }
        Q.f[0] = q.ok ? q.row.f : T(0);
        Q.bb[0] = q.ok ? q.b : T(0);
        T xv = q.ok ? q.x : T(0);
        if (CORRECT) {
            const int ax0 = (tx >> 1) + (tx & 1);
            const T wx0 = (tx & 1) ? T(0.75) : T(0.25), wx1 = (tx & 1) ? T(0.25) : T(0.75);
            const int ay0 = (Q.ly >> 1) + (Q.ly & 1);
            const T wy0 = (Q.ly & 1) ? T(0.75) : T(0.25), wy1 = (Q.ly & 1) ? T(0.25) : T(0.75);
            const T lo = wx0 * sz[ay0][ax0] + wx1 * sz[ay0][ax0 + 1];
            const T hi = wx0 * sz[ay0 + 1][ax0] + wx1 * sz[ay0 + 1][ax0 + 1];
            const T corr = wy0 * lo + wy1 * hi;
            xv += q.ok ? corr : T(0);
        }
        sx[xs(t)][Q.ly][tx] = xv;
        Q.hr = q.ok ? q.row.r : T(0);
        Q.hu = q.ok ? q.row.u : T(0);
    };
    // The couplings of plane t, after the first barrier of step t.
    auto put_couplings = [&](int t) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            sr[cs(t)][P[r].ly][tx] = P[r].hr;
            su[cs(t)][P[r].ly][tx] = P[r].hu;
        }
    };
    // Half-sweep H (1..4) on plane p = t - H: colour FIRST for odd H, the
    // other for even H, on positions at least H cells inside the tile.  The
    // row of this thread's position of that colour is 2 ty + sel.
    auto half = [&](auto hc, int p) __attribute__((always_inline)) {
        constexpr int H = decltype(hc)::value;
        const int colour = (H & 1) ? FIRST : 1 - FIRST;
        const int sel = ((gx + gy0 + p) & 1) != colour;  // 0: row 2ty, 1: row 2ty + 1
        const int ly = 2 * ty + sel;
        const int gy = gy0 + ly;
        constexpr int R = H - ZH;  // innermost ring this half-sweep is valid on
        const bool in = tx >= R && tx < kS2LX - R && ly >= R && ly < kS2LY - R && xok && gy >= 0 &&
                        gy < d.ny;
        if (in) {
            const T w = sel ? P[1].w[H] : P[0].w[H];
            const T bv = sel ? P[1].bb[H] : P[0].bb[H];
            const T fm = sel ? P[1].f[H + 1] : P[0].f[H + 1];
            const T fp = sel ? P[1].f[H] : P[0].f[H];
            const int sp = xs(p), sm = xs(p - 1), sn = xs(p + 1), s1 = cs(p);
            const T xo = sx[sp][ly][tx];
            T xv;
            if (w == T(0)) {
                xv = T(0);  // zero-centre row
            } else if (ZERO && !CORRECT && H == 1) {
                // x is zero everywhere before the first half-sweep
                xv = w * (bv - T(0));
            } else {
                // (h1 runs right after the commit of plane p + 1 with no
                // barrier between: the +z neighbour is this thread's own
                // position, which it has just written itself)
                const T s = rb_sum6<T>(sr[s1][ly][tx - 1], sx[sp][ly][tx - 1], sr[s1][ly][tx],
                                       sx[sp][ly][tx + 1], su[s1][ly - 1][tx], sx[sp][ly - 1][tx],
                                       su[s1][ly][tx], sx[sp][ly + 1][tx], fm, sx[sm][ly][tx], fp,
                                       sx[sn][ly][tx]);
                xv = rb_update<T>(omega, xo, w, bv, s);
            }
            sx[sp][ly][tx] = xv;
        }
    };
    // After h4 on plane q: the final values of both positions (the one h4
    // updated and the one h3 left) become the pending stores.
    auto finish = [&](int q) __attribute__((always_inline)) {
        const unsigned sq = xs(q);
#pragma unroll
        for (int r = 0; r < 2; ++r) P[r].px = sx[sq][P[r].ly][tx];
    };
    auto flush = [&](int q) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (P[r].in4) xout[gx + (idx_t)d.nx * P[r].gy + sxy * q] = P[r].px;
    };
    // RES at step t, after h4 on plane t - 4: the +z term of plane t - 5's
    // residual, which is then stored, and plane t - 4's without that term, at
    // both positions.  k_residual_restrict's operand order and masks (Jet's
    // FdmBlas3::residual, src/jet/fdm_linear_system3.cpp:107-115); outside
    // the grid x, A.f and the couplings in LDS are zero.
    auto resid_finish = [&](int t) __attribute__((always_inline)) {
        const unsigned sn = xs(t - 4);
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (P[r].in4)
                res[gx + (idx_t)d.nx * P[r].gy + sxy * (t - 5)] =
                    P[r].rv - P[r].f[5] * sx[sn][P[r].ly][tx];
    };
    // The -x / -y terms need no mask: past the grid edge the LDS coupling and
    // x are both +0, so the product is the +0 the masked form subtracts.
    auto resid_start = [&](int t) __attribute__((always_inline)) {
        const unsigned sq = xs(t - 4), sm = xs(t - 5), s1 = cs(t - 4);
        const bool hr = gx + 1 < d.nx;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            Pos& Q = P[r];
            if (!Q.in4) continue;  // the written interior: ring >= 5, inside the grid
            const int ly = Q.ly;
            const bool hu = Q.gy + 1 < d.ny;
            T rv = Q.bb[4] - Q.cc[kC - 1] * sx[sq][ly][tx];
            rv -= sr[s1][ly][tx - 1] * sx[sq][ly][tx - 1];
            rv -= hr ? sr[s1][ly][tx] * sx[sq][ly][tx + 1] : T(0);
            rv -= su[s1][ly - 1][tx] * sx[sq][ly - 1][tx];
            rv -= hu ? su[s1][ly][tx] * sx[sq][ly + 1][tx] : T(0);
            rv -= Q.f[5] * sx[sm][ly][tx];
            Q.rv = rv;
        }
    };

    using H1 = std::integral_constant<int, 1>;
    using H2 = std::integral_constant<int, 2>;
    using H3 = std::integral_constant<int, 3>;
    using H4 = std::integral_constant<int, 4>;

    // Planes k0-kH .. k1+kH-1 are committed; half-sweep H covers planes
    // k0-kH+H-ZH .. k1+kH-1-H+ZH (those within the grid): exactly the planes
    // the next half-sweep reads for the chunk's planes k0 .. k1-1 (and, for
    // RES, h4 covers k0-1 .. k1, the planes the residual reads).  Half-sweep
    // H runs on plane t - H at step t, so the last live one is at step
    // t1 + ZH: with ZH, one step past the last commit.
    // Two planes of loads are in flight ahead of the committed one; the loop
    // is unrolled twice so the prefetch slot of every step is static.
    const int t0 = k0 - kH, t1 = k1 + kH - 1;  // t1 - t0 >= 6
    const int tl = t1 + ZH;                    // last step with a live half-sweep
    auto live = [&](int p, int H) {
        return p >= max(k0 - kH + H - ZH, 0) && p < min(k1 + kH - H + ZH, d.nz);
    };
    if (CORRECT) zfetch(t0);
    // plane by plane, in the order the loop issues them
#pragma unroll
    for (int r = 0; r < 2; ++r) fetch(P[r].p[0], P[r].gy, t0);
    if (kPD == 2) {
#pragma unroll
        for (int r = 0; r < 2; ++r) fetch(P[r].p[kPD - 1], P[r].gy, t0 + 1);
    }
    if (CORRECT) zstore(t0);
    __syncthreads();

    // The body is unrolled kPD times (the prefetch slot of every step is
    // static) and has no exit: a step past t1 commits a re-read plane into
    // the slot of a plane that is no longer read, and runs only the
    // half-sweeps still live (none past tl).
    const int tend = t0 + (tl - t0 + kPD) / kPD * kPD;
    bool pending = false;  // uniform: the pending x stores hold plane t - 5
    for (int t2 = t0; t2 < tend; t2 += kPD) {
#pragma unroll
        for (int u = 0; u < kPD; ++u) {
            const int t = t2 + u;
#pragma unroll
            for (int r = 0; r < 2; ++r) commit(P[r], P[r].p[u % kPD], t);
            if (CORRECT && t + 1 <= t1) zfetch(t + 1);
            // unconditional (past the end it re-reads plane t1, unused) so the
            // compiler can count the loads in flight and wait only for plane t
#pragma unroll
            for (int r = 0; r < 2; ++r) fetch(P[r].p[u % kPD], P[r].gy, min(t + kPD, t1));
            if (pending) flush(t - 5);
            // h1 on plane t - 1 needs no barrier after the commit: its in-plane
            // neighbours and couplings (plane t - 1) are a step old, and its
            // +z neighbour is this thread's own committed value
            if (live(t - 1, 1)) half(H1{}, t - 1);
            __syncthreads();
            put_couplings(t);
            if (live(t - 2, 2)) half(H2{}, t - 2);
            __syncthreads();
            if (live(t - 3, 3)) half(H3{}, t - 3);
            __syncthreads();
            pending = t - 4 >= k0 && t - 4 < k1;
            if (live(t - 4, 4)) half(H4{}, t - 4);
            if (CORRECT && t + 1 <= t1) zstore(t + 1);
            __syncthreads();
            if (pending) finish(t - 4);
            if (RES) {
                if (t - 5 >= k0 && t - 5 < k1) resid_finish(t);
                if (pending) resid_start(t);
            }
        }
    }
    if (pending) flush(tend - 5);
}

}  // namespace jg
