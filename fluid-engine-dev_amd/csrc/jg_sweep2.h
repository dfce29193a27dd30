// jg_sweep2.h -- two red-black sweeps in one pass over HBM (temporal
// blocking of the fused sweep in jg_sweep.h).
//
// Two consecutive sweeps with the same colour order are four colour
// half-sweeps h = 1..4 (colours c1, c2, c1, c2).  A block owns a 64 x 32 tile
// of (i, j) positions -- a 56 x 24 interior plus a 4-cell halo that is
// recomputed, not exchanged -- and marches through a chunk of z planes.  At
// step t it commits plane t from the registers it was prefetched into
// (prefetching plane t + 1 or t + 2) and runs h1 on plane t-1 in the same
// phase (its +z neighbour is the thread's own position, just committed by
// itself; everything else it reads is a step old), then h2 on t-2, h3 on t-3
// and h4 on t-4, each after a barrier: four barriers per step.  Plane t-4 is
// then final and written back.  Half-sweep h is valid on positions at least h
// cells inside the tile (it reads neighbours updated by h-1, valid one cell
// further out), so the interior (ring >= 4) of plane t-4 holds exactly the
// values two full sweeps give.  Each block also recomputes the 4 planes below
// and above its chunk.
//
// Every ring has six slots and the step loop is unrolled six times, so the
// slot of every plane a step touches is a compile-time constant: LDS
// addresses are a per-thread base plus an immediate offset, and the register
// rings are renamed, never shifted.  x lives in an LDS ring (planes t-5 .. t),
// the +x / +y couplings in LDS rings (planes t .. t-5; they are the -x / -y
// couplings of the neighbours; plane t's are written after the step's first
// barrier).  A thread owns two vertically adjacent positions, which have
// opposite colours, so every half-sweep updates one position per thread; its
// register rings hold w = omega / c, b and A.f per colour rather than per
// position (the commit sorts them), so a half-sweep of a given colour reads
// fixed registers.  Each update is the single sweep's FP32 arithmetic
// (rb_sum6 / rb_coef / rb_update of jg_sweep.h: Jet's relaxRedBlack operand
// order for s, src/jet/fdm_gauss_seidel_solver3.cpp:163-190), so the result
// is bit-identical to two single FP32 fused sweeps.  Zero-centre rows are set
// to zero, as in the single sweep.
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
// Steps per unrolled loop body, and the slot count of every ring.
constexpr int kS2U = 6;
// Halo (in x, y and z) and interior of the tile: 4 cells for the two sweeps,
// 5 when the residual is written too; one less when x starts at zero with no
// correction (Z), where the first half-sweep, w b, reads no neighbour and is
// valid up to the tile's edge.
template <bool RES, bool Z = false>
struct S2Tile {
    static constexpr int H = (RES ? 5 : 4) - (Z ? 1 : 0);
    static constexpr int IX = kS2LX - 2 * H, IY = kS2LY - 2 * H;
    static_assert(IX % 2 == 0 && IY % 2 == 0, "tile interiors stay even");
};
constexpr int kS2CzX = kS2LX / 2 + 2, kS2CzY = kS2LY / 2 + 2;
static_assert(kS2CzX * kS2CzY <= kS2Threads, "one coarse entry per thread");

// FP32 only: the FP64 rings would need 288 KB of LDS.
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
    constexpr int kH = G::H;
    constexpr int U = kS2U;
    // slot of plane p: (p - t0) mod U, a constant at every unrolled step.
    // Tile row ly is array row ly + 1: a guard row above and below each
    // slot keeps the neighbour reads of every lane inside the arrays (a
    // half-sweep computes on all lanes and keeps the result where valid).
    __shared__ __align__(16) T sx[U][kS2LY + 2][kS2LX];  // x
    __shared__ T sr[U][kS2LY + 2][kS2LX];                 // A.r
    __shared__ T su[U][kS2LY + 2][kS2LX];                 // A.u
    __shared__ T sz[CORRECT ? kS2CzY : 1][CORRECT ? kS2CzX : 1];

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
        if (RES) res += sys * n;
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int gx0 = bx * G::IX - kH, gy0 = by * G::IY - kH;  // even unless RES
    const int gx = gx0 + tx;
    const int k0 = bz * zc;
    const int k1 = min(d.nz, k0 + zc);
    const idx_t sxy = (idx_t)d.nx * d.ny;
    const bool xok = gx >= 0 && gx < d.nx;
    const int t0 = k0 - kH, t1 = k1 + kH - 1;  // t1 - t0 >= 6
    const int tl = t1 + ZH;                    // last step with a live half-sweep
    // Colour of row 0 (tile row 2 ty) in plane t0; in plane t0 + m it is
    // q ^ (m & 1), and the row of colour c in that plane is c ^ q ^ (m & 1).
    const int q = (gx + gy0 + t0) & 1;
    // the tile rows a colour can occupy: row lyA = 2 ty + q has colour 0 in
    // the planes t0 + even, row lyB the other one
    const int lyA = 2 * ty + q, lyB = 2 * ty + 1 - q;

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
    // To stay within the registers of 1024 threads, RES prefetches one plane
    // ahead instead of two.
    constexpr int kPD = RES ? 1 : 2;
    static_assert(U % kPD == 0, "prefetch slots are static");
    Pre pf[2][kPD];  // [row][(t - t0) mod kPD]
    // Per colour c and slot: w = omega / c (0 for a zero-centre row), b and
    // A.f of the position of colour c in that plane.  RES keeps A.c there
    // instead of w (the residual needs it; w is formed as the half-sweep
    // reads it), to stay within the registers of 1024 threads.
    T W[2][U], BB[2][U], F[2][U];
    // RES: the partial residual of plane t-4 per row
    T rv[2] = {T(0), T(0)};
    T hr[2], hu[2];                  // couplings of plane t, written after the first barrier
    T px[2] = {T(0), T(0)};          // final x of the last finished plane, per row
    T pxm[2] = {T(0), T(0)};         // and of the one before it (RES)
    bool in4[2];                     // row in the written interior
    int gyr[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int ly = 2 * ty + r;
        gyr[r] = gy0 + ly;
        in4[r] = tx >= kH && tx < kS2LX - kH && ly >= kH && ly < kS2LY - kH && xok &&
                 gyr[r] >= 0 && gyr[r] < d.ny;
#pragma unroll
        for (int s = 0; s < U; ++s) {
            W[r][s] = omega;
            BB[r][s] = T(0);
            F[r][s] = T(0);
            sx[s][ly + 1][tx] = T(0);
            sr[s][ly + 1][tx] = T(0);
            su[s][ly + 1][tx] = T(0);
        }
        hr[r] = hu[r] = T(0);
    }
    // In-tile masks of the two rows a colour can occupy, per half-sweep ring.
    auto inside = [&](int ly, int R) {
        const int gy = gy0 + ly;
        return tx >= R && tx < kS2LX - R && ly >= R && ly < kS2LY - R && xok && gy >= 0 &&
               gy < d.ny;
    };
    bool inA[4], inB[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        inA[h] = inside(lyA, h + 1 - ZH);
        inB[h] = inside(lyB, h + 1 - ZH);
    }

    auto fetch = [&](Pre& pr, int gy, int t) __attribute__((always_inline)) {
        const bool ok = xok && t >= 0 && t < d.nz && gy >= 0 && gy < d.ny;
        const idx_t c = ok ? gx + (idx_t)d.nx * gy + sxy * t : 0;
        pr.row = A[c];
        pr.b = b[c];
        pr.x = ZERO ? T(0) : xin[c];
        pr.ok = ok;
    };
    // Commit plane t at unrolled step u: x of both rows to LDS, the row values
    // sorted by colour into the register rings, the couplings held.
    auto commit = [&](auto uc) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        T wr[2], br[2], fr[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const Pre& pr = pf[r][u % kPD];
            const T c = pr.ok ? pr.row.c : T(1);
            // w = omega / c, 0 for a zero-centre row (RES: c itself)
            wr[r] = RES ? c : rb_coef<T>(omega, c);
            br[r] = pr.ok ? pr.b : T(0);
            fr[r] = pr.ok ? pr.row.f : T(0);
            T xv = pr.ok ? pr.x : T(0);
            if (CORRECT) {
                const int ly = 2 * ty + r;
                const int ax0 = (tx >> 1) + (tx & 1);
                const T wx0 = (tx & 1) ? T(0.75) : T(0.25), wx1 = (tx & 1) ? T(0.25) : T(0.75);
                const int ay0 = (ly >> 1) + (ly & 1);
                const T wy0 = (ly & 1) ? T(0.75) : T(0.25), wy1 = (ly & 1) ? T(0.25) : T(0.75);
                const T lo = wx0 * sz[ay0][ax0] + wx1 * sz[ay0][ax0 + 1];
                const T hi = wx0 * sz[ay0 + 1][ax0] + wx1 * sz[ay0 + 1][ax0 + 1];
                const T corr = wy0 * lo + wy1 * hi;
                xv += pr.ok ? corr : T(0);
            }
            sx[u][2 * ty + r + 1][tx] = xv;
            hr[r] = pr.ok ? pr.row.r : T(0);
            hu[r] = pr.ok ? pr.row.u : T(0);
        }
        // row 0 has colour q ^ (u & 1) in this plane
        const bool swap = (q ^ (u & 1)) != 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int r0 = c, r1 = 1 - c;  // the row of colour c without / with the swap
            W[c][u] = swap ? wr[r1] : wr[r0];
            BB[c][u] = swap ? br[r1] : br[r0];
            F[c][u] = swap ? fr[r1] : fr[r0];
        }
    };
    auto put_couplings = [&](auto uc) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            sr[u][2 * ty + r + 1][tx] = hr[r];
            su[u][2 * ty + r + 1][tx] = hu[r];
        }
    };
    // Half-sweep H (1..4) on plane p = t - H at unrolled step u: colour FIRST
    // for odd H, the other for even H, on positions at least H cells inside
    // the tile.  The position of colour cH in p is on row lyA or lyB, fixed
    // by the parity of u - H.
    auto half = [&](auto uc, auto hc) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        constexpr int H = decltype(hc)::value;
        constexpr int cH = (H & 1) ? FIRST : 1 - FIRST;
        constexpr int sp = (u - H + 2 * U) % U, sm = (u - H - 1 + 2 * U) % U,
                      sn = (u - H + 1 + 2 * U) % U;
        // row of colour cH in plane p: cH ^ q ^ ((u - H) & 1), i.e. lyA when
        // cH ^ ((u - H) & 1) is 0
        constexpr bool onA = (cH ^ ((u - H + 2 * U) & 1)) == 0;
        const int ly = onA ? lyA : lyB;
        const bool in = onA ? inA[H - 1] : inB[H - 1];
        // every lane computes; a lane outside the valid ring writes its old
        // value back (a half-sweep reads only the other colour, so nothing
        // reads this position in this phase)
        const T w = RES ? rb_coef<T>(omega, W[cH][sp]) : W[cH][sp];
        const T bv = BB[cH][sp];
        const T fp = F[cH][sp];       // A.f of this position in plane p
        const T fm = F[1 - cH][sm];   // in plane p - 1, where it has the other colour
        const int row = ly + 1;       // array row of tile row ly
        const T xo = sx[sp][row][tx];
        T xv;
        if (ZERO && !CORRECT && H == 1) {
            // x is zero everywhere before the first half-sweep
            xv = w * (bv - T(0));
        } else {
            // (h1 runs right after the commit of plane p + 1 with no
            // barrier between: the +z neighbour is this thread's own
            // position, which it has just written itself)
            const T s = rb_sum6<T>(sr[sp][row][tx - 1], sx[sp][row][tx - 1], sr[sp][row][tx],
                                   sx[sp][row][tx + 1], su[sp][row - 1][tx], sx[sp][row - 1][tx],
                                   su[sp][row][tx], sx[sp][row + 1][tx], fm, sx[sm][row][tx], fp,
                                   sx[sn][row][tx]);
            xv = rb_update<T>(omega, xo, w, bv, s);
        }
        xv = w == T(0) ? T(0) : xv;  // zero-centre row
        sx[sp][row][tx] = in ? xv : xo;
    };
    // After h4 on plane t - 4: the final values of both rows (the one h4
    // updated and the one h3 left); they are the pending stores when plane
    // t - 4 is in the chunk, and the residual's -z values one step later.
    auto finish = [&](auto uc) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        constexpr int s4 = (u - 4 + 2 * U) % U;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            pxm[r] = px[r];
            px[r] = sx[s4][2 * ty + r + 1][tx];
        }
    };
    auto flush = [&](int p) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (in4[r]) xout[gx + (idx_t)d.nx * gyr[r] + sxy * p] = px[r];
    };
    // A.f of row r in the plane p with p - t0 = m (mod U), slot sk, from the
    // colour-sorted ring: the row's colour there is r ^ q ^ (m & 1).
    auto f_row = [&](int r, int m, int sk) __attribute__((always_inline)) {
        const bool c1 = ((r ^ q ^ (m & 1)) & 1) != 0;
        return c1 ? F[1][sk] : F[0][sk];
    };
    // RES at step t (unrolled u), after h4 on plane t - 4: the +z term of
    // plane t - 5's residual, which is then stored, and plane t - 4's without
    // that term, at both rows.  k_residual_restrict's operand order and masks
    // (Jet's FdmBlas3::residual, src/jet/fdm_linear_system3.cpp:107-115);
    // outside the grid x, A.f and the couplings in LDS are zero.  x of this
    // thread's own positions in planes t - 4 and t - 5 come from registers.
    auto resid_finish = [&](auto uc, int t) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        constexpr int s5 = (u - 5 + 2 * U) % U;
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (in4[r])
                res[gx + (idx_t)d.nx * gyr[r] + sxy * (t - 5)] =
                    rv[r] - f_row(r, u - 5, s5) * px[r];
    };
    // The -x / -y terms need no mask: past the grid edge the LDS coupling and
    // x are both +0, so the product is the +0 the masked form subtracts.
    auto resid_start = [&](auto uc) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        constexpr int s4 = (u - 4 + 2 * U) % U, s5 = (u - 5 + 2 * U) % U;
        const bool hr_ = gx + 1 < d.nx;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (!in4[r]) continue;  // the written interior: ring >= 5, inside the grid
            const int ly = 2 * ty + r;
            const bool hu_ = gyr[r] + 1 < d.ny;
            const bool c1 = ((r ^ q ^ ((u - 4) & 1)) & 1) != 0;
            const T bq = c1 ? BB[1][s4] : BB[0][s4];
            const T cq = c1 ? W[1][s4] : W[0][s4];  // A.c (RES keeps it in W)
            T v = bq - cq * px[r];
            v -= sr[s4][ly + 1][tx - 1] * sx[s4][ly + 1][tx - 1];
            v -= hr_ ? sr[s4][ly + 1][tx] * sx[s4][ly + 1][tx + 1] : T(0);
            v -= su[s4][ly - 1 + 1][tx] * sx[s4][ly - 1 + 1][tx];
            v -= hu_ ? su[s4][ly + 1][tx] * sx[s4][ly + 1 + 1][tx] : T(0);
            v -= f_row(r, u - 5, s5) * pxm[r];
            rv[r] = v;
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
    auto live = [&](int p, int H) {
        return p >= max(k0 - kH + H - ZH, 0) && p < min(k1 + kH - H + ZH, d.nz);
    };
    if (CORRECT) zfetch(t0);
#pragma unroll
    for (int r = 0; r < 2; ++r) fetch(pf[r][0], gyr[r], t0);
    if (kPD == 2) {
#pragma unroll
        for (int r = 0; r < 2; ++r) fetch(pf[r][kPD - 1], gyr[r], t0 + 1);
    }
    if (CORRECT) zstore(t0);
    __syncthreads();

    // One step; t is uniform, so every thread leaves the loop together.
    bool pending = false;  // uniform: the pending x stores hold plane t - 5
    auto step = [&](auto uc, int t) __attribute__((always_inline)) {
        constexpr int u = decltype(uc)::value;
        commit(uc);
        if (CORRECT && t + 1 <= t1) zfetch(t + 1);
        // unconditional (past the end it re-reads plane t1, unused) so the
        // compiler can count the loads in flight and wait only for plane t
#pragma unroll
        for (int r = 0; r < 2; ++r) fetch(pf[r][u % kPD], gyr[r], min(t + kPD, t1));
        if (pending) flush(t - 5);
        // h1 on plane t - 1 needs no barrier after the commit: its in-plane
        // neighbours and couplings (plane t - 1) are a step old, and its
        // +z neighbour is this thread's own committed value
        if (live(t - 1, 1)) half(uc, H1{});
        __syncthreads();
        put_couplings(uc);
        if (live(t - 2, 2)) half(uc, H2{});
        __syncthreads();
        if (live(t - 3, 3)) half(uc, H3{});
        __syncthreads();
        pending = t - 4 >= k0 && t - 4 < k1;
        if (live(t - 4, 4)) half(uc, H4{});
        if (CORRECT && t + 1 <= t1) zstore(t + 1);
        __syncthreads();
        finish(uc);
        if constexpr (RES) {
            if (t - 5 >= k0 && t - 5 < k1) resid_finish(uc, t);
            if (pending) resid_start(uc);
        }
    };
    int t = t0;
    while (true) {
        // unrolled U times: the step index u is a constant in each copy
        if (t > tl) break;
        step(std::integral_constant<int, 0>{}, t);
        if (++t > tl) break;
        step(std::integral_constant<int, 1>{}, t);
        if (++t > tl) break;
        step(std::integral_constant<int, 2>{}, t);
        if (++t > tl) break;
        step(std::integral_constant<int, 3>{}, t);
        if (++t > tl) break;
        step(std::integral_constant<int, 4>{}, t);
        if (++t > tl) break;
        step(std::integral_constant<int, 5>{}, t);
        ++t;
    }
    if (pending) flush(t - 5);
}

}  // namespace jg
