// jg_bottom.h -- the bottom of the V-cycle in one workgroup.
//
// Below a few thousand cells a level's kernels are pure launch and memory
// latency (a fused sweep over a 16^3 level costs as much as one over 64^3).
// k_bottom runs the whole lower part of the V-cycle -- every level from the
// first one with at most BotCfg<T>::MaxCells cells down to the coarsest -- as one
// 1024-thread workgroup: vectors of every level in LDS, the stencil of the
// top level (BotCfg<T>::MC cells per thread) and of the coarsest level (one
// cell per thread) in registers, the rows of the levels between staged in
// LDS with precomputed cell coordinates, __syncthreads between phases.
// Zero-centre rows are set to zero by the smoother, as in k_rb_sweep.
// The sequence is exactly the multi-kernel V-cycle's: `pre` forward (red,
// black) sweeps from zero, residual, Jet's restriction, recursion, Jet's
// prolongation-correction, `post` backward (black, red) sweeps; on the
// coarsest level `cfwd` forward then `cfwd` backward sweeps.  Restriction and
// prolongation are FdmMgUtils3::restrict / correct
// (src/jet/fdm_mg_linear_system3.cpp:43-179); the smoother is Jet's
// relaxRedBlack update (src/jet/fdm_gauss_seidel_solver3.cpp:163-190).
#pragma once

#include "jg_device.h"
#include "jg_sweep.h"

namespace jg {

constexpr int kBotThreads = 1024;
constexpr int kBotMaxLevels = 6;  // 4096 -> 512 -> 64 -> 8 -> 1 (+1 spare)

// Cells per thread on the top level: 4 in FP32; 2 in FP64, where four
// cells' stencils no longer fit in the 128 VGPRs of a 1024-thread block.
template <typename T>
struct BotCfg {
    static constexpr int MC = sizeof(T) == 4 ? 4 : 2;
    static constexpr int MaxCells = MC * kBotThreads;
    static constexpr int Pool = MaxCells + MaxCells / 7 + 8;  // >= sum of the level sizes
    static constexpr int Rows = MaxCells / 7 + 8;             // >= sizes below the top
};

template <typename T>
struct BotLevels {
    int L;  // levels in this bottom solve (>= 1); level 0 is the top, each halves exactly
    Dims d[1];
    const Row<T>* A[kBotMaxLevels];
};

// Stencil of MC cells per thread, held in registers: omega / center and the
// six couplings (zero where the neighbour is outside the grid), plus one flag
// word per cell (bits 0-5: the -x, +x, -y, +y, -z, +z neighbour exists; bit
// 6: parity; bit 7: the cell exists; bit 8: see below).  With an even nx the
// cells are taken in x-adjacent pairs (2p, 2p + 1), p = tid + kBotThreads * h,
// which always have opposite colours; slot 2h holds the pair's red cell and
// slot 2h + 1 its black cell (bit 8 set: the red cell is 2p + 1), so every
// lane updates exactly one cell of each pair per half-sweep.  Otherwise slot
// m of a thread is cell tid + kBotThreads * m.
template <typename T, int MC, bool PAIRS>
struct BotRegs {
    static_assert(!PAIRS || MC % 2 == 0, "pairs need an even cell count");
    T wc[MC], xp[MC], xm[MC], yp[MC], ym[MC], zp[MC], zm[MC];
    int fl[MC];
    int nx, nxy;

    __device__ __forceinline__ int cell(int m) const {
        if (PAIRS)
            return 2 * ((int)threadIdx.x + kBotThreads * (m >> 1)) + ((m & 1) ^ ((fl[m] >> 8) & 1));
        return (int)threadIdx.x + kBotThreads * m;
    }

    __device__ __forceinline__ void load(const Row<T>* __restrict__ A, const Dims& d, T omega) {
        const int n = d.nx * d.ny * d.nz;
        nx = d.nx;
        nxy = d.nx * d.ny;
#pragma unroll
        for (int m = 0; m < MC; ++m) {
            int cl = (int)threadIdx.x + kBotThreads * m, swap = 0;
            if (PAIRS) {
                const int p = (int)threadIdx.x + kBotThreads * (m >> 1);
                const int i0 = (2 * p) % d.nx, j0 = ((2 * p) / d.nx) % d.ny, k0 = (2 * p) / nxy;
                swap = (i0 + j0 + k0) & 1;  // cell 2p is black: the red one is 2p + 1
                cl = 2 * p + ((m & 1) ^ swap);
            }
            const bool ok = cl < n;
            const int cc = ok ? cl : 0;
            const int i = cc % d.nx, j = (cc / d.nx) % d.ny, k = cc / nxy;
            const bool hxm = i > 0, hxp = i + 1 < d.nx, hym = j > 0, hyp = j + 1 < d.ny,
                       hzm = k > 0, hzp = k + 1 < d.nz;
            fl[m] = (hxm ? 1 : 0) | (hxp ? 2 : 0) | (hym ? 4 : 0) | (hyp ? 8 : 0) | (hzm ? 16 : 0) |
                    (hzp ? 32 : 0) | (((i + j + k) & 1) << 6) | (ok ? 128 : 0) | (swap << 8);
            const Row<T> row = A[cc];
            wc[m] = (ok && row.c != T(0)) ? omega / row.c : T(0);  // 0: zero-centre row
            xp[m] = (ok && hxp) ? row.r : T(0);
            yp[m] = (ok && hyp) ? row.u : T(0);
            zp[m] = (ok && hzp) ? row.f : T(0);
            xm[m] = (ok && hxm) ? A[cc - 1].r : T(0);
            ym[m] = (ok && hym) ? A[cc - d.nx].u : T(0);
            zm[m] = (ok && hzm) ? A[cc - nxy].f : T(0);
        }
    }
    __device__ __forceinline__ T offdiag(int m, const T* X, int c) const {
#pragma clang fp contract(off)
        const int f = fl[m];
        T s = xm[m] * X[c - ((f & 1) ? 1 : 0)];
        s += xp[m] * X[c + ((f & 2) ? 1 : 0)];
        s += ym[m] * X[c - ((f & 4) ? nx : 0)];
        s += yp[m] * X[c + ((f & 8) ? nx : 0)];
        s += zm[m] * X[c - ((f & 16) ? nxy : 0)];
        s += zp[m] * X[c + ((f & 32) ? nxy : 0)];
        return s;
    }
    // (1 - w) x + (w / c)(b - s): the relaxRedBlack update with w / c
    // precomputed (one rounding apart from Jet's w (b - s) / c).  A
    // half-sweep reads only the other colour, so all loads are issued before
    // any store.  Callers pass a constant colour.
    __device__ __forceinline__ void relax(T* X, const T* B, int colour, T omega) const {
#pragma clang fp contract(off)
        if (PAIRS) {
            constexpr int H = PAIRS ? MC / 2 : 1;
            T nv[H];
            int cs[H];
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int m = 2 * h + colour;
                int c = (fl[m] & 128) ? cell(m) : 0;  // 0: no cell (flags of cell 0)
                // an empty asm makes c opaque: the neighbour addresses are
                // formed at each use instead of being hoisted out of the
                // kernel's loops for every cell slot (which spilled to scratch)
                asm volatile("" : "+v"(c));
                cs[h] = c;
                nv[h] = wc[m] == T(0) ? T(0) : (T(1) - omega) * X[c] + wc[m] * (B[c] - offdiag(m, X, c));
            }
#pragma unroll
            for (int h = 0; h < H; ++h)
                if (fl[2 * h + colour] & 128) X[cs[h]] = nv[h];
        } else {
            const int want = 128 | (colour << 6);
#pragma unroll
            for (int m = 0; m < MC; ++m) {
                if ((fl[m] & 192) != want) continue;
                int c = cell(m);
                asm volatile("" : "+v"(c));
                X[c] = wc[m] == T(0) ? T(0) : (T(1) - omega) * X[c] + wc[m] * (B[c] - offdiag(m, X, c));
            }
        }
    }
    // The center coefficient is read from A (once per V-cycle, cache-hot).
    __device__ __forceinline__ void residual(const Row<T>* __restrict__ A, const T* X, const T* B,
                                             T* R) const {
#pragma clang fp contract(off)
#pragma unroll
        for (int m = 0; m < MC; ++m) {
            if (!(fl[m] & 128)) continue;
            int c = cell(m);
            asm volatile("" : "+v"(c));
            R[c] = B[c] - (A[c].c * X[c] + offdiag(m, X, c));
        }
    }
};

// Cell coordinates packed 10 bits each (every level below a top level that
// has a coarser level has dimensions <= 1024).
__device__ __forceinline__ int bot_pack(int i, int j, int k) { return i | (j << 10) | (k << 20); }
__device__ __forceinline__ void bot_unpack(int v, int& i, int& j, int& k) {
    i = v & 1023;
    j = (v >> 10) & 1023;
    k = v >> 20;
}

template <typename T>
__device__ __forceinline__ void bot_relax_lds(const Row<T>* R, const int* Q, T* X, const T* B,
                                              const Dims& d, int colour, T omega) {
#pragma clang fp contract(off)
    const int n = d.nx * d.ny * d.nz;
    for (int c = threadIdx.x; c < n; c += kBotThreads) {
        int i, j, k;
        bot_unpack(Q[c], i, j, k);
        if (((i + j + k) & 1) != colour) continue;
        const Row<T> row = R[c];
        const T s = offdiag<T, T, T>(R, X, row, c, i, j, k, d);
        X[c] = row.c == T(0) ? T(0) : (T(1) - omega) * X[c] + omega * (B[c] - s) / row.c;
    }
}

template <typename T>
__device__ __forceinline__ void bot_residual_lds(const Row<T>* R, const int* Q, const T* X,
                                                 const T* B, T* Rs, const Dims& d) {
#pragma clang fp contract(off)
    const int n = d.nx * d.ny * d.nz;
    for (int c = threadIdx.x; c < n; c += kBotThreads) {
        int i, j, k;
        bot_unpack(Q[c], i, j, k);
        const Row<T> row = R[c];
        Rs[c] = B[c] - (row.c * X[c] + offdiag<T, T, T>(R, X, row, c, i, j, k, d));
    }
}

// coarse = restriction of fine (Jet's 4x4x4 kernel, clamped), cd = coarse dims.
// Tap t of coarse index i is fine index 2i - 1 + t, clamped to 2i at the low
// edge and to 2i + 1 at the high edge; weights 1/8, 3/8, 3/8, 1/8.  One z tap
// per loop iteration keeps 16 loads in flight, not 64.
template <typename T>
__device__ __forceinline__ void bot_restrict(const T* fine, T* coarse, const int* Qc,
                                             const Dims& cd) {
#pragma clang fp contract(off)
    const int n = cd.nx * cd.ny * cd.nz;
    const int fnx = 2 * cd.nx, fny = 2 * cd.ny;
    auto tap = [](int i, int t, int cn) {
        const int f = 2 * i - 1 + t;
        return (t == 0 && i == 0) ? 2 * i : ((t == 3 && i + 1 == cn) ? 2 * i + 1 : f);
    };
    auto wt = [](int t) { return (t == 0 || t == 3) ? T(0.125) : T(0.375); };
    for (int c = threadIdx.x; c < n; c += kBotThreads) {
        int i, j, k;
        bot_unpack(Qc[c], i, j, k);
        T sum = 0;
#pragma unroll 1
        for (int z = 0; z < 4; ++z) {
            const int kz = tap(k, z, cd.nz);
            const T wz = wt(z);
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const T* row = fine + fnx * (tap(j, y, cd.ny) + fny * kz);
#pragma unroll
                for (int x = 0; x < 4; ++x) sum += wt(x) * wt(y) * wz * row[tap(i, x, cd.nx)];
            }
        }
        coarse[c] = sum;
    }
}

template <typename T>
__device__ __forceinline__ void bot_correct(const T* coarse, T* fine, const int* Qf,
                                            const Dims& fd) {
    const int n = fd.nx * fd.ny * fd.nz;
    for (int c = threadIdx.x; c < n; c += kBotThreads) {
        int i, j, k;
        bot_unpack(Qf[c], i, j, k);
        fine[c] += prolong_at<T>(coarse, i, j, k, fd);
    }
}

// x0 = bottom V-cycle applied to b0 (level 0 of `bl`).  post0: post-sweeps
// on the top level, post: on the levels below it.  PAIRS: the top level's nx
// is even (pair-slotted registers, see BotRegs).
template <typename T, bool PAIRS>
__global__ __launch_bounds__(kBotThreads) void k_bottom(BotLevels<T> bl, const T* __restrict__ b0,
                                                        T* __restrict__ x0, int pre, int post0,
                                                        int post, int cfwd, T omega,
                                                        const Ctl* __restrict__ ctl) {
#pragma clang fp contract(off)
    // one workgroup per system of a batch group; AL(l): its rows of level l
    const int sys = blockIdx.x;
    if (sys > 0) {
        b0 += sys * bl.d[0].n();
        x0 += sys * bl.d[0].n();
        if (ctl) ctl += sys * kSysCtl;
    }
    auto AL = [&](int l) { return bl.A[l] + sys * (bl.d[0].n() >> (3 * l)); };
    if (ctl && ctl->done) return;
    __shared__ T sv[3][BotCfg<T>::Pool];    // x, b, residual of every level
    __shared__ int sq[BotCfg<T>::Pool];     // packed cell coordinates (L >= 2)
    __shared__ Row<T> sA[BotCfg<T>::Rows];  // rows of the levels below the top
    const int L = bl.L;
    // Levels halve exactly, so level l has n0 >> 3l cells; the vectors of
    // level l start at the sum of the sizes above it.
    const int n0 = bl.d[0].nx * bl.d[0].ny * bl.d[0].nz;
    auto off = [&](int l) {
        int o = 0;
        for (int m = 0; m < l; ++m) o += n0 >> (3 * m);
        return o;
    };
    auto D = [&](int l) { return Dims{bl.d[0].nx >> l, bl.d[0].ny >> l, bl.d[0].nz >> l}; };
    auto X = [&](int l) { return &sv[0][off(l)]; };
    auto B = [&](int l) { return &sv[1][off(l)]; };
    auto Rv = [&](int l) { return &sv[2][off(l)]; };
    auto Q = [&](int l) { return &sq[off(l)]; };
    auto RA = [&](int l) { return &sA[off(l) - n0]; };  // rows of level l >= 1

    // Prologue: every global load is issued before the LDS is filled, so
    // their latencies overlap.
    constexpr int MC = BotCfg<T>::MC;
    const int tid = threadIdx.x;
    T bv[MC];
#pragma unroll
    for (int m = 0; m < MC; ++m) {
        const int c = tid + kBotThreads * m;
        bv[m] = c < n0 ? b0[c] : T(0);
    }
    BotRegs<T, MC, PAIRS> top;
    top.load(AL(0), bl.d[0], omega);
    BotRegs<T, 1, false> bot;  // coarsest level when it is not the top
    if (L >= 2) bot.load(AL(L - 1), D(L - 1), omega);

    for (int c = tid; c < off(L); c += kBotThreads) sv[0][c] = T(0);
#pragma unroll
    for (int m = 0; m < MC; ++m) {
        const int c = tid + kBotThreads * m;
        if (c < n0) sv[1][c] = bv[m];
    }
    if (L >= 2) {
        for (int l = 0; l < L; ++l) {
            const Dims d = D(l);
            int* q = Q(l);
            for (int c = tid; c < d.nx * d.ny * d.nz; c += kBotThreads)
                q[c] = bot_pack(c % d.nx, (c / d.nx) % d.ny, c / (d.nx * d.ny));
        }
    }
    for (int l = 1; l + 1 < L; ++l)  // rows of the levels between (<= 512 cells each)
        if (tid < (n0 >> (3 * l))) RA(l)[tid] = AL(l)[tid];
    __syncthreads();

    auto relax = [&](int l, int colour) __attribute__((always_inline)) {
        if (l == 0) top.relax(X(0), B(0), colour, omega);
        else if (l == L - 1) bot.relax(X(l), B(l), colour, omega);
        else bot_relax_lds<T>(RA(l), Q(l), X(l), B(l), D(l), colour, omega);
        __syncthreads();
    };
    // down
    for (int l = 0; l + 1 < L; ++l) {
        for (int s = 0; s < pre; ++s) {
            relax(l, 0);
            relax(l, 1);
        }
        if (l == 0) top.residual(AL(0), X(0), B(0), Rv(0));
        else bot_residual_lds<T>(RA(l), Q(l), X(l), B(l), Rv(l), D(l));
        __syncthreads();
        bot_restrict<T>(Rv(l), B(l + 1), Q(l + 1), D(l + 1));
        __syncthreads();
    }
    // coarsest
    for (int s = 0; s < cfwd; ++s) {
        relax(L - 1, 0);
        relax(L - 1, 1);
    }
    for (int s = 0; s < cfwd; ++s) {
        relax(L - 1, 1);
        relax(L - 1, 0);
    }
    // up
    for (int l = L - 2; l >= 0; --l) {
        bot_correct<T>(X(l + 1), X(l), Q(l), D(l));
        __syncthreads();
        const int np = l == 0 ? post0 : post;
        for (int s = 0; s < np; ++s) {
            relax(l, 1);
            relax(l, 0);
        }
    }
#pragma unroll
    for (int m = 0; m < MC; ++m) {
        const int c = tid + kBotThreads * m;
        if (c < n0) x0[c] = sv[0][c];
    }
}

}  // namespace jg
