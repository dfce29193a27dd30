// jg_brick.h -- two red-black sweeps of a small level in one launch, by
// bricks (no z-march).
//
// On the coarser levels of the V-cycle the z-marching sweeps are latency-
// bound: a launch costs about as much as its serial chain of plane steps,
// not its bytes.  Here a 1024-thread block owns an 8^3 brick of cells and
// loads the 16^3 region around it (a 4-cell halo on every side, recomputed,
// not exchanged) into LDS once: x and the three upper couplings A.r, A.u,
// A.f (the lower ones are the neighbours' upper ones); A.c and b stay in
// registers, four cells per thread along z.  It then runs the four colour
// half-sweeps of two sweeps (colours c1, c2, c1, c2) with one barrier each:
// half-sweep h is valid on cells at least h inside the region, so the brick
// (ring >= 4) ends with exactly the values of two single sweeps, which it
// writes.  Every update is the single fused sweep's (jg_sweep.h: rb_sum6,
// rb_coef, rb_update; Jet's relaxRedBlack expression for FP64); zero-centre
// rows are set to zero; the correction
// variant adds the prolongation of the coarse correction evaluated as the
// single sweep does (z first, then x, then y), so the result is
// bit-identical to two single sweeps.  Reads xin, writes xout (xin != xout).
#pragma once

#include "jg_device.h"
#include "jg_sweep.h"

namespace jg {

constexpr int kBkB = 8, kBkH = 4, kBkR = kBkB + 2 * kBkH;  // brick, halo, region edge
constexpr int kBkThreads = kBkR * kBkR * (kBkR / 4);        // four cells per thread along z
static_assert(kBkThreads == 1024, "one 1024-thread block per brick");

template <typename T, int FIRST, bool ZERO, bool CORRECT>
__global__ __launch_bounds__(kBkThreads, sizeof(T) == 4 ? 2 : 1) void k_brick_sweep2(
    const Row<T>* __restrict__ A, const T* __restrict__ b, const T* __restrict__ xin,
    T* __restrict__ xout, Dims d, T omega, const T* __restrict__ coarse,
    const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    constexpr bool kWeighted = sizeof(T) == 4;
    // the system of a batch group: (nz + 7) / 8 z-blocks per system
    const int nbz = (d.nz + kBkB - 1) / kBkB;
    const int sys = (int)blockIdx.z / nbz;
    const int bzl = (int)blockIdx.z - sys * nbz;
    if (sys > 0) {
        const idx_t n = d.n();
        A += sys * n;
        b += sys * n;
        if (!ZERO) xin += sys * n;
        xout += sys * n;
        if (CORRECT) coarse += sys * (n / 8);
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    __shared__ T sx[kBkR][kBkR][kBkR];
    __shared__ T sr[kBkR][kBkR][kBkR];
    __shared__ T su[kBkR][kBkR][kBkR];
    __shared__ T sf[kBkR][kBkR][kBkR];
    // CORRECT: the coarse correction around the region, 10^3 values
    constexpr int kCz = kBkR / 2 + 2;
    __shared__ T scz[CORRECT ? kCz : 1][CORRECT ? kCz : 1][CORRECT ? kCz : 1];

    const int t = threadIdx.x;
    const int li = t % kBkR, lj = (t / kBkR) % kBkR, lk0 = (t / (kBkR * kBkR)) * 4;
    const int gi = (int)blockIdx.x * kBkB - kBkH + li;
    const int gj = (int)blockIdx.y * kBkB - kBkH + lj;
    const int gk0 = bzl * kBkB - kBkH + lk0;
    const bool xyok = gi >= 0 && gi < d.nx && gj >= 0 && gj < d.ny;
    const int ring_xy = min(min(li, kBkR - 1 - li), min(lj, kBkR - 1 - lj));

    // Coarse taps of the prolongation in x and y (the single sweep's clamped
    // staging rule, which is Jet's edge rule), relative to the staged block
    // of coarse cells whose corner is (gi0/2 - 1, gj0/2 - 1, gk0/2 - 1).
    int ci0 = 0, ci1 = 0, cj0 = 0, cj1 = 0;
    T wx0 = 0, wx1 = 0, wy0 = 0, wy1 = 0;
    const int ib = (int)blockIdx.x * (kBkB / 2) - kBkH / 2 - 1;
    const int jb = (int)blockIdx.y * (kBkB / 2) - kBkH / 2 - 1;
    const int kb = bzl * (kBkB / 2) - kBkH / 2 - 1;
    if (CORRECT) {
        // stage: thread t < 1000 loads coarse cell (ib, jb, kb) + (t % 10, t / 10 % 10, t / 100),
        // each index clamped to the coarse grid
        if (t < kCz * kCz * kCz) {
            const int qi = t % kCz, qj = (t / kCz) % kCz, qk = t / (kCz * kCz);
            const int cnx = d.nx / 2, cny = d.ny / 2, cnz = d.nz / 2;
            const int ci = min(max(ib + qi, 0), cnx - 1);
            const int cj = min(max(jb + qj, 0), cny - 1);
            const int ck = min(max(kb + qk, 0), cnz - 1);
            scz[qk][qj][qi] = coarse[ci + (idx_t)cnx * (cj + (idx_t)cny * ck)];
        }
        __syncthreads();
        const int cnx = d.nx / 2, cny = d.ny / 2;
        const int hi = gi >> 1, hj = gj >> 1;  // gi, gj may be negative outside the grid
        if ((gi & 1) == 0) {
            ci0 = hi - 1; ci1 = hi; wx0 = T(0.25); wx1 = T(0.75);
        } else {
            ci0 = hi; ci1 = hi + 1; wx0 = T(0.75); wx1 = T(0.25);
        }
        if ((gj & 1) == 0) {
            cj0 = hj - 1; cj1 = hj; wy0 = T(0.25); wy1 = T(0.75);
        } else {
            cj0 = hj; cj1 = hj + 1; wy0 = T(0.75); wy1 = T(0.25);
        }
        ci0 = min(max(ci0, 0), cnx - 1) - ib;
        ci1 = min(max(ci1, 0), cnx - 1) - ib;
        cj0 = min(max(cj0, 0), cny - 1) - jb;
        cj1 = min(max(cj1, 0), cny - 1) - jb;
    }

    T cw[4], bb[4];  // FP64: c; FP32: w = omega / c (0 for a zero-centre row)
    bool ok[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int lk = lk0 + m, gk = gk0 + m;
        ok[m] = xyok && gk >= 0 && gk < d.nz;
        const idx_t c = ok[m] ? gi + (idx_t)d.nx * (gj + (idx_t)d.ny * gk) : 0;
        const Row<T> row = A[c];
        const T bv = b[c];
        T xv = ZERO ? T(0) : xin[c];
        if (CORRECT && ok[m]) {
            int k0c, k1c;
            T w0, w1;
            prolong_z<T>(gk, d, k0c, k1c, w0, w1);
            k0c -= kb;
            k1c -= kb;
            auto Z = [&](int ci, int cj) { return w0 * scz[k0c][cj][ci] + w1 * scz[k1c][cj][ci]; };
            const T lo = wx0 * Z(ci0, cj0) + wx1 * Z(ci1, cj0);
            const T hi = wx0 * Z(ci0, cj1) + wx1 * Z(ci1, cj1);
            const T corr = wy0 * lo + wy1 * hi;
            xv += corr;
        }
        const T cc = ok[m] ? row.c : T(1);
        cw[m] = rb_coef<T>(omega, cc);  // FP32: w = omega / c; FP64: c
        bb[m] = ok[m] ? bv : T(0);
        sx[lk][lj][li] = ok[m] ? xv : T(0);
        sr[lk][lj][li] = ok[m] ? row.r : T(0);
        su[lk][lj][li] = ok[m] ? row.u : T(0);
        sf[lk][lj][li] = ok[m] ? row.f : T(0);
    }
    __syncthreads();

#pragma unroll
    for (int h = 1; h <= 4; ++h) {
        const int colour = (h & 1) ? FIRST : 1 - FIRST;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int lk = lk0 + m;
            const int ring = min(ring_xy, min(lk, kBkR - 1 - lk));
            if (ring < h || !ok[m] || ((gi + gj + gk0 + m) & 1) != colour) continue;
            const T xo = sx[lk][lj][li];
            T xv;
            if (cw[m] == T(0)) {
                xv = T(0);  // zero-centre row
            } else if (ZERO && !CORRECT && h == 1) {
                // x is zero everywhere before the first half-sweep
                xv = kWeighted ? cw[m] * (bb[m] - T(0)) : omega * (bb[m] - T(0)) / cw[m];
            } else {
                // Jet's relax operand order: -x, +x, -y, +y, -z, +z
                const T s = rb_sum6<T>(sr[lk][lj][li - 1], sx[lk][lj][li - 1], sr[lk][lj][li],
                                       sx[lk][lj][li + 1], su[lk][lj - 1][li], sx[lk][lj - 1][li],
                                       su[lk][lj][li], sx[lk][lj + 1][li], sf[lk - 1][lj][li],
                                       sx[lk - 1][lj][li], sf[lk][lj][li], sx[lk + 1][lj][li]);
                xv = rb_update<T>(omega, xo, cw[m], bb[m], s);
            }
            sx[lk][lj][li] = xv;
        }
        __syncthreads();
    }

    // the brick: ring >= 4 in every direction
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int lk = lk0 + m;
        const int ring = min(ring_xy, min(lk, kBkR - 1 - lk));
        if (ring >= kBkH && ok[m])
            xout[gi + (idx_t)d.nx * (gj + (idx_t)d.ny * (gk0 + m))] = sx[lk][lj][li];
    }
}

}  // namespace jg
