// jg_transfer.h -- fused residual + restriction of the V-cycle's downward
// pass: coarse b = restrict(b - A x) in one pass over the fine level, without
// writing or re-reading the fine residual.
//
// A 512-thread block owns a 31 x 7 tile of coarse (I, J) columns and marches
// through a chunk of coarse planes.  Restriction tap t of coarse index I is
// fine index 2I - 1 + t (clamped at the grid edges), so the tile's taps cover
// fine columns 2*I0 - 1 .. 2*I0 + 62 and rows 2*J0 - 1 .. 2*J0 + 14: a 64 x 16
// fine tile, two rows per thread.  For every fine plane the block computes
// the residual (FdmBlas3::residual's expression, src/jet/fdm_linear_system3.cpp:
// 98-116, in Jet's operand order) into LDS; threads e < 217 then restrict that
// plane in x and y (16 taps) for their coarse column and keep the result in a
// 4-plane register window, from which every second plane the z taps give the
// coarse value.  This is FdmMgUtils3::restrict's 4 x 4 x 4 kernel
// (src/jet/fdm_mg_linear_system3.cpp:43-98) evaluated separably.  Tiles
// overlap by one fine column/row on each side; each block recomputes one fine
// plane at each end of its chunk.
#pragma once

#include "jg_device.h"
#include "jg_kernels.h"

namespace jg {

constexpr int kRrCX = 31, kRrCY = 7;                     // coarse tile
constexpr int kRrLX = 64, kRrTY = 8, kRrLY = 2 * kRrTY;  // fine tile, thread rows
constexpr int kRrThreads = kRrLX * kRrTY;
static_assert(2 * kRrCX + 2 == kRrLX && 2 * kRrCY + 2 == kRrLY, "tile taps");

// Element o of a plane that starts at a block-uniform base: the byte offset
// is formed in 32 bits, so the load takes the base in scalar registers and
// only a 32-bit offset per lane (a plane holds far fewer than 2^32 bytes).
template <typename U>
__device__ __forceinline__ U plane_at(const U* base, unsigned o) {
    return *reinterpret_cast<const U*>(reinterpret_cast<const char*>(base) + o * (unsigned)sizeof(U));
}

// FROMT: b already holds the fine residual (written by the pre-smoothing
// launch); A and x are not read, and the restriction is the same.
template <typename T, bool FROMT = false>
__global__ __launch_bounds__(kRrThreads, sizeof(T) == 4 ? 4 : 2) void k_residual_restrict(
    const Row<T>* __restrict__ A, const T* __restrict__ x, const T* __restrict__ b,
    T* __restrict__ cb, Dims d, Dims cd, int zc, const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ T sres[2][kRrLY][kRrLX];  // fine residual of plane p in buffer p & 1

    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    idx_t part;
    const int sys = block_system(bx, by, bz, (cd.nz + zc - 1) / zc, part);
    if (sys > 0) {
        const idx_t n = d.n();
        if (!FROMT) {
            A += sys * n;
            x += sys * n;
        }
        b += sys * n;
        cb += sys * cd.n();
        if (ctl) ctl += sys * kSysCtl;
    }
    if (ctl && ctl->done) return;
    const int I0 = bx * kRrCX, J0 = by * kRrCY;
    const int fx0 = 2 * I0 - 1, fy0 = 2 * J0 - 1;  // fine origin of the tile
    const int K0 = bz * zc, K1 = min(cd.nz, K0 + zc);
    const idx_t sz = (idx_t)d.nx * d.ny;
    const int gx = fx0 + tx;
    const bool xok = gx >= 0 && gx < d.nx;

    // Two fine positions per thread (rows ty and ty + kRrTY of the tile).
    // Planes are double-buffered in registers (buf[p & 1] for plane p, with
    // the plane counted from p0): plane p + 1 is loaded while plane p is
    // reduced, and the +z value x(p + 1) that plane p's residual needs was
    // loaded one plane earlier, so no load is waited for in the plane it is
    // issued.
    struct Cell {
        int ly;
        bool ok;  // inside the fine grid
        StCol h;
        unsigned col;  // index within a plane (0 outside the grid)
        StPlane<T, T, 1> buf[2];
        T bb[2], xz[2];
        T xm, fm;
    };
    Cell C[2];
    // Fine planes of this chunk: the taps of coarse planes K0 .. K1-1.
    const int p0 = max(2 * K0 - 1, 0), p1 = min(2 * K1, d.nz - 1);
    // Plane q of this chunk (clamped to p1: the load past the chunk end reads
    // cached data and is never used).  Addresses are a plane base common to
    // the block plus a 32-bit offset within the plane.
    auto load_plane = [&](Cell& c, int slot, int q) __attribute__((always_inline)) {
        const idx_t base = sz * min(q, p1);
        c.bb[slot] = plane_at(b + base, c.col);
        if (FROMT) return;
        const Row<T>* Ap = A + base;
        const T* xp = x + base;
        const unsigned o = c.col;
        const unsigned ol = c.h.l ? o - 1 : o, orr = c.h.r ? o + 1 : o;
        const unsigned od = c.h.d ? o - d.nx : o, ou = c.h.u ? o + d.nx : o;
        StPlane<T, T, 1>& P = c.buf[slot];
        P.row = plane_at(Ap, o);
        P.rl = plane_at(Ap, ol).r;
        P.ud = plane_at(Ap, od).u;
        P.v[0][0] = plane_at(xp, o);
        P.v[0][1] = plane_at(xp, ol);
        P.v[0][2] = plane_at(xp, orr);
        P.v[0][3] = plane_at(xp, od);
        P.v[0][4] = plane_at(xp, ou);
    };
    auto load_xz = [&](Cell& c, int slot, int q) __attribute__((always_inline)) {
        // x at plane q + 1 for plane q's residual (masked past the grid at use)
        if (!FROMT) c.xz[slot] = plane_at(x + sz * min(q + 1, d.nz - 1), c.col);
    };
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        Cell& c = C[r];
        c.ly = ty + r * kRrTY;
        const int gy = fy0 + c.ly;
        c.ok = xok && gy >= 0 && gy < d.ny;
        c.h = c.ok ? StCol{gx > 0, gx + 1 < d.nx, gy > 0, gy + 1 < d.ny}
                   : StCol{false, false, false, false};
        c.col = c.ok ? (unsigned)(gx + d.nx * gy) : 0u;  // outside: column 0, result unused
        load_plane(c, 0, p0);
        load_xz(c, 0, p0);
        c.xm = c.fm = T(0);
        if (!FROMT && c.ok && p0 > 0) {
            const idx_t below = sz * (p0 - 1);
            c.xm = plane_at(x + below, c.col);
            c.fm = plane_at(A + below, c.col).f;
        }
    }

    // Coarse column of threads e < kRrCX * kRrCY and its x / y taps in the
    // tile (local index 2*ci + t is fine index 2I - 1 + t; clamped like Jet's).
    const int e = tx + kRrLX * ty;
    const int ci = e % kRrCX, cj = min(e / kRrCX, kRrCY - 1);
    const int I = I0 + ci, J = J0 + cj;
    const bool cact = e < kRrCX * kRrCY && I < cd.nx && J < cd.ny;
    const int li = 2 * ci, lj = 2 * cj;
    const int i0 = I > 0 ? li : li + 1, i3 = I + 1 < cd.nx ? li + 3 : li + 2;
    const int j0 = J > 0 ? lj : lj + 1, j3 = J + 1 < cd.ny ? lj + 3 : lj + 2;
    const T wa = T(0.125), wb = T(0.375);
    T w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // x-y restricted planes p-3 .. p

    int K = K0;
    // An even number of planes, so the unrolled body has no exit (which would
    // keep the compiler from counting the loads in flight): a plane past p1
    // re-reads plane p1 and produces no coarse value.
    const int pend = p0 + ((p1 - p0 + 2) & ~1);
    for (int p2 = p0; p2 < pend; p2 += 2) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int p = p2 + u;
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                Cell& c = C[r];
                // loads of plane p + 1 into the other slot
                load_plane(c, u ^ 1, p + 1);
                load_xz(c, u ^ 1, p + 1);
                T rv = c.bb[u];
                if (!FROMT) {
                    const T xp = c.ok && p + 1 < d.nz ? c.xz[u] : T(0);
                    // Jet's residual operand order (src/jet/fdm_linear_system3.cpp:107-115)
                    const StPlane<T, T, 1>& q = c.buf[u];
                    rv = rv - q.row.c * q.v[0][0];
                    rv -= c.h.l ? q.rl * q.v[0][1] : T(0);
                    rv -= c.h.r ? q.row.r * q.v[0][2] : T(0);
                    rv -= c.h.d ? q.ud * q.v[0][3] : T(0);
                    rv -= c.h.u ? q.row.u * q.v[0][4] : T(0);
                    rv -= c.fm * c.xm;
                    rv -= q.row.f * xp;
                    c.xm = q.v[0][0];
                    c.fm = q.row.f;
                }
                sres[p & 1][c.ly][tx] = c.ok ? rv : T(0);
            }
            __syncthreads();
            // x-y restriction of plane p for this thread's coarse column (the
            // buffer is rewritten two planes later, after the next barrier).
            if (e < kRrCX * kRrCY) {
                const T(*R)[kRrLX] = sres[p & 1];
                auto rowsum = [&](int y) {
                    return (wa * R[y][i0] + wb * R[y][li + 1]) + (wb * R[y][li + 2] + wa * R[y][i3]);
                };
                const T P = (wa * rowsum(j0) + wb * rowsum(lj + 1)) + (wb * rowsum(lj + 2) + wa * rowsum(j3));
                w0 = w1;
                w1 = w2;
                w2 = w3;
                w3 = P;
            }
            // coarse plane K once its last tap plane is in the window
            if (K < K1 && p == min(2 * K + 2, d.nz - 1)) {
                // z taps 2K-1 .. 2K+2 (clamped) as window slots q - (p - 3)
                const int k0 = (K > 0 ? 2 * K - 1 : 2 * K) - (p - 3);
                const int k1 = 2 * K - (p - 3), k2 = 2 * K + 1 - (p - 3);
                const int k3 = (K + 1 < cd.nz ? 2 * K + 2 : 2 * K + 1) - (p - 3);
                auto win = [&](int q) { return q == 0 ? w0 : (q == 1 ? w1 : (q == 2 ? w2 : w3)); };
                const T v = (wa * win(k0) + wb * win(k1)) + (wb * win(k2) + wa * win(k3));
                if (cact) cb[I + (idx_t)cd.nx * (J + (idx_t)cd.ny * K)] = v;
                ++K;
            }
        }
    }
}

}  // namespace jg
