// jg_solver.hip -- solver objects, CG / MGPCG drivers and the C ABI.
//
// Host code only launches kernels; every scalar of the Krylov loop (alpha,
// beta, sigma, the stop test) lives in device memory (jg::Ctl) and is
// computed by one-block finalize kernels, so iterations run back to back on
// one stream and the host only polls the stop flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "jg_internal.h"
#include "jg_kernels.h"
#include "jg_sweep.h"
#include "jg_sweep2.h"
#include "jg_plane2.h"
#include "jg_brick.h"
#include "jg_bottom.h"
#include "jg_transfer.h"
#include "jg_csr.h"

#include <cstdlib>

using namespace jg;

// ------------------------------------------- scratch of the no-allocation entry
//
// jg_relax_red_black_noalloc keeps the copy of x that its sweep reads in this
// array instead of a hipMalloc'd temporary.  Every GPU holds its own copy of
// the array (16 MiB) from module load until the process ends.  Host code
// never passes the array's name to a kernel or a copy: that name is a host-side
// stand-in.  The device address comes from hipGetSymbolAddress, resolved once
// per device by jg_solver_create with that device current.
constexpr size_t kNoAllocMaxCells = size_t(1) << 21;
static __device__ __align__(256) double scratch[kNoAllocMaxCells];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};
hipError_t g_scratch_err[kScratchDevices] = {};
bool g_scratch_done[kScratchDevices] = {};

// Resolves the address of `scratch` on `device` (which must be current) the
// first time it is asked for that device.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? static_cast<double*>(p) : nullptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_err[device];
}

hipError_t scratch_address(int device, double** out) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) return hipErrorNotInitialized;
    *out = g_scratch_ptr[device];
    return g_scratch_err[device];
}

// ------------------------------------------------------- level kernels

// z-chunk of the fused sweep when no tuned value exists: enough blocks to
// fill the chip, chunks long enough that the 4-plane pipeline fill stays a
// small overhead.  JG_SWEEP_ZC overrides everything (tuning experiments).
int sweep_zc_env() {
    static const int env = [] {
        const char* e = std::getenv("JG_SWEEP_ZC");
        return e ? std::atoi(e) : 0;
    }();
    return env;
}

// JG_TUNE_LOG=1 prints every autotuner timing to stderr.
bool tune_log() {
    static const bool on = [] {
        const char* e = std::getenv("JG_TUNE_LOG");
        return e && *e && *e != '0';
    }();
    return on;
}

int sweep_zc(const Dims& d, int tuned) {
    if (sweep_zc_env() > 0) return std::min(sweep_zc_env(), d.nz);
    if (tuned > 0) return std::min(tuned, d.nz);
    const long tiles = (long)((d.nx + kSwIX - 1) / kSwIX) * ((d.ny + kSwIY - 1) / kSwIY);
    const long gz = std::max<long>(1, (1024 + tiles - 1) / tiles);
    long zc = (d.nz + gz - 1) / gz;
    zc = std::max<long>(8, std::min<long>(64, zc));
    return (int)std::min<long>(zc, d.nz);
}

// Grids of a launch over every system of s (one, or a batch group): the
// stencil kernels repeat their z extent, the elementwise kernels take the
// system from the grid's y index, the one-block kernels one block each.
inline dim3 grid_sys(dim3 g, const jg_solver* s) {
    g.z *= (unsigned)s->nsys;
    return g;
}
inline dim3 grid_1d(int nb, const jg_solver* s) { return dim3((unsigned)nb, (unsigned)s->nsys, 1); }

// One fused red-black sweep (see jg_sweep.h).  first: colour updated first
// (0 = red, forward; 1 = black, backward).  xin == nullptr: x starts at zero.
// coarse != nullptr: add the prolongation of coarse to x on load.
// p_bx/p_bb != nullptr: emit b.x and b.b partials.  shape: index into
// kSweepShapes.  Returns the number of b.x / b.b partials written (the
// block count; 0 without partials, or when the grid has more blocks than a
// slot holds -- the caller then reduces with k_dot).
template <typename T, int RP, int PD>
void launch_sweep_shape(jg_solver* s, int code, dim3 grid, const Row<T>* A, const T* b,
                        const T* xin, T* xout, const Dims& d, int zc, T omega, const T* coarse,
                        double* p_bx, double* p_bb, const Ctl* ctl) {
    const dim3 block(kSwLX, kSwLY / RP, 1);
#define JG_SWEEP_CASE(C, F, Z, K, D)                                                            \
    case C:                                                                                     \
        k_rb_sweep<T, F, Z, K, D, RP, PD><<<grid, block, 0, s->stream>>>(A, b, xin, xout, d, zc, \
                                                                       omega, coarse, p_bx,    \
                                                                       p_bb, ctl);             \
        break;
    switch (code) {
        JG_SWEEP_CASE(0, 0, false, false, false)
        JG_SWEEP_CASE(1, 0, false, false, true)
        JG_SWEEP_CASE(2, 0, false, true, false)
        JG_SWEEP_CASE(3, 0, false, true, true)
        JG_SWEEP_CASE(4, 0, true, false, false)
        JG_SWEEP_CASE(5, 0, true, false, true)
        JG_SWEEP_CASE(6, 0, true, true, false)
        JG_SWEEP_CASE(7, 0, true, true, true)
        JG_SWEEP_CASE(8, 1, false, false, false)
        JG_SWEEP_CASE(9, 1, false, false, true)
        JG_SWEEP_CASE(10, 1, false, true, false)
        JG_SWEEP_CASE(11, 1, false, true, true)
        JG_SWEEP_CASE(12, 1, true, false, false)
        JG_SWEEP_CASE(13, 1, true, false, true)
        JG_SWEEP_CASE(14, 1, true, true, false)
        JG_SWEEP_CASE(15, 1, true, true, true)
    }
#undef JG_SWEEP_CASE
}

template <typename T>
int launch_sweep(jg_solver* s, const Row<T>* A, const T* b, const T* xin, T* xout, const Dims& d,
                 int first, const T* coarse, double* p_bx, double* p_bb, T omega, const Ctl* ctl,
                 int tuned_zc = 0, int shape = 0) {
    int zc = sweep_zc(d, tuned_zc);
    const long cols = (long)((d.nx + kSwIX - 1) / kSwIX) * ((d.ny + kSwIY - 1) / kSwIY);
    // one b.x / b.b partial per block: keep the block count within a slot
    while (p_bx && zc < d.nz && cols * ((d.nz + zc - 1) / zc) > kMaxPartials) ++zc;
    dim3 grid((d.nx + kSwIX - 1) / kSwIX, (d.ny + kSwIY - 1) / kSwIY, (d.nz + zc - 1) / zc);
    if ((long)grid.x * grid.y * grid.z > kMaxPartials) p_bx = p_bb = nullptr;  // caller uses k_dot
    const int nb = (int)(grid.x * grid.y * grid.z);  // per system
    grid = grid_sys(grid, s);
    const int code = (first ? 8 : 0) | (xin ? 0 : 4) | (coarse ? 2 : 0) | (p_bx ? 1 : 0);
    switch (shape) {
        case 1:
            launch_sweep_shape<T, 2, 2>(s, code, grid, A, b, xin, xout, d, zc, omega, coarse, p_bx,
                                        p_bb, ctl);
            break;
        case 2:
            launch_sweep_shape<T, 1, 1>(s, code, grid, A, b, xin, xout, d, zc, omega, coarse, p_bx,
                                        p_bb, ctl);
            break;
        case 3:
            launch_sweep_shape<T, 1, 2>(s, code, grid, A, b, xin, xout, d, zc, omega, coarse, p_bx,
                                        p_bb, ctl);
            break;
        default:
            launch_sweep_shape<T, 2, 1>(s, code, grid, A, b, xin, xout, d, zc, omega, coarse, p_bx,
                                        p_bb, ctl);
    }
    return p_bx ? nb : 0;
}

// Two sweeps in one launch (jg_sweep2.h, FP32 only).  res != nullptr: also
// write the residual b - A x of the result there (not with a coarse
// correction).  Returns JG_OK, or JG_ERR_UNSUPPORTED for FP64 or for res with
// a coarse correction.
template <typename T>
int launch_sweep2(jg_solver* s, const Row<T>* A, const T* b, const T* xin, T* xout, const Dims& d,
                  int first, const T* coarse, T* res, T omega, const Ctl* ctl, int tuned_zc) {
    if constexpr (sizeof(T) != 4) {
        return JG_ERR_UNSUPPORTED;
    } else {
        if (res && coarse) return JG_ERR_UNSUPPORTED;
        const int zc = std::min(tuned_zc > 0 ? tuned_zc : 32, d.nz);
        // the tile of a zero-start launch without a correction has a one cell
        // narrower halo (jg_sweep2.h)
        const bool z = !xin && !coarse;
        const int ix = res ? (z ? S2Tile<true, true>::IX : S2Tile<true>::IX)
                           : (z ? S2Tile<false, true>::IX : S2Tile<false>::IX);
        const int iy = res ? (z ? S2Tile<true, true>::IY : S2Tile<true>::IY)
                           : (z ? S2Tile<false, true>::IY : S2Tile<false>::IY);
        const dim3 grid = grid_sys(dim3((d.nx + ix - 1) / ix, (d.ny + iy - 1) / iy, (d.nz + zc - 1) / zc), s);
        const dim3 block(kS2LX, kS2TY, 1);
        const int code = (first ? 8 : 0) | (xin ? 0 : 4) | (coarse ? 2 : 0) | (res ? 1 : 0);
#define JG_SWEEP2_CASE(C, F, Z, K, R)                                                     \
    case C:                                                                               \
        k_rb_sweep2<T, F, Z, K, R><<<grid, block, 0, s->stream>>>(A, b, xin, xout, d, zc, \
                                                                  omega, coarse, res,     \
                                                                  ctl);                   \
        break;
        switch (code) {
            JG_SWEEP2_CASE(0, 0, false, false, false)
            JG_SWEEP2_CASE(1, 0, false, false, true)
            JG_SWEEP2_CASE(2, 0, false, true, false)
            JG_SWEEP2_CASE(4, 0, true, false, false)
            JG_SWEEP2_CASE(5, 0, true, false, true)
            JG_SWEEP2_CASE(6, 0, true, true, false)
            JG_SWEEP2_CASE(8, 1, false, false, false)
            JG_SWEEP2_CASE(9, 1, false, false, true)
            JG_SWEEP2_CASE(10, 1, false, true, false)
            JG_SWEEP2_CASE(12, 1, true, false, false)
            JG_SWEEP2_CASE(13, 1, true, false, true)
            JG_SWEEP2_CASE(14, 1, true, true, false)
        }
#undef JG_SWEEP2_CASE
        return JG_OK;
    }
}

// Two sweeps of a small level in one launch by 8^3 bricks (jg_brick.h).
template <typename T>
void launch_brick2(jg_solver* s, const Row<T>* A, const T* b, const T* xin, T* xout, const Dims& d,
                   int first, const T* coarse, T omega, const Ctl* ctl) {
    const dim3 grid = grid_sys(
        dim3((d.nx + kBkB - 1) / kBkB, (d.ny + kBkB - 1) / kBkB, (d.nz + kBkB - 1) / kBkB), s);
    const int code = (first ? 4 : 0) | (xin ? 0 : 2) | (coarse ? 1 : 0);
#define JG_BRICK_CASE(C, F, Z, K)                                                               \
    case C:                                                                                     \
        k_brick_sweep2<T, F, Z, K><<<grid, kBkThreads, 0, s->stream>>>(A, b, xin, xout, d, omega, \
                                                                       coarse, ctl);            \
        break;
    switch (code) {
        JG_BRICK_CASE(0, 0, false, false)
        JG_BRICK_CASE(1, 0, false, true)
        JG_BRICK_CASE(2, 0, true, false)
        JG_BRICK_CASE(3, 0, true, true)
        JG_BRICK_CASE(4, 1, false, false)
        JG_BRICK_CASE(5, 1, false, true)
        JG_BRICK_CASE(6, 1, true, false)
        JG_BRICK_CASE(7, 1, true, true)
    }
#undef JG_BRICK_CASE
}

// Levels up to this size may use the brick kernel (it reloads a 16^3 region
// per 8^3 brick, which only pays where launches are latency-bound).
constexpr idx_t kBrickMaxCells = idx_t(1) << 21;

// Sweeps of one smoothing step, grouped into launches: 1 = one sweep,
// kLaunchMarch / kLaunchBrick = two sweeps in one launch of k_rb_sweep2 /
// k_brick_sweep2, where the tuner found that faster for the kind of the
// launch's first sweep (kind_first for the first launch, plain afterwards).
// pairs > 0 forces two-sweep launches wherever two sweeps follow (FP32:
// bricks up to 2^18 cells, else the z-marching kernel; FP64: bricks up to
// kBrickMaxCells cells), pairs < 0 forbids them.
constexpr int kLaunchSingle = 1, kLaunchMarch = 2, kLaunchBrick = 3;
template <typename T>
std::vector<int> sweep_plan(const Level& lv, unsigned n, int kind_first, int pairs) {
    std::vector<int> v;
    const bool big = lv.d.n() > BotCfg<T>::MaxCells;
    for (unsigned k = 0; k < n;) {
        int mode = Prec<T>::pair(lv, k == 0 ? kind_first : kSweepPlain);
        if (pairs > 0)  // FP32: bricks up to 64^3 cells, the z-marching kernel above
            mode = sizeof(T) == 4 ? (lv.d.n() <= (idx_t(1) << 18) ? kPairBrick : kPairMarch)
                                  : (lv.d.n() <= kBrickMaxCells ? kPairBrick : kPairNone);
        if (pairs < 0 || !big || n - k < 2) mode = kPairNone;
        v.push_back(mode == kPairBrick ? kLaunchBrick : mode == kPairMarch ? kLaunchMarch : kLaunchSingle);
        k += mode == kPairNone ? 1 : 2;
    }
    return v;
}

// Coarse planes per block of the fused residual-restriction: about `blocks`
// blocks, to fill the chip; each block recomputes one fine plane per chunk.
int rr_zc(const Dims& cd, long blocks = 1024) {
    const long tiles = (long)((cd.nx + kRrCX - 1) / kRrCX) * ((cd.ny + kRrCY - 1) / kRrCY);
    long zc = (long)cd.nz * tiles / blocks;
    zc = std::max<long>(2, std::min<long>(32, zc));
    return (int)std::min<long>(zc, cd.nz);
}

// Whether the stencil kernels take the tiled z-march on shape d: rows of at
// least 48 cells (a 64-lane row would idle too many lanes below that).  On
// small grids its chunks shorten to keep the chip busy (tm_grid), which
// measured as fast or faster than the per-cell z-march down to 64^3 (64^3
// CG 15.5 -> 11.9 ms, 128^3 MGPCG and the 64-system batch level or better).
inline bool use_tm(const Dims& d) { return d.nx >= 48; }

// r = b - A x on shape d (nsys systems back to back) stored as TR, with
// k_residual's DOT / REFRESH / GATED options; returns the partial count per
// system.  Rows of at least 48 cells take the tiled z-march
// (k_residual_tile), narrower grids the per-cell one (k_residual).
template <typename T, typename TR, bool DOT, bool REFRESH, bool GATED>
int launch_residual(jg_solver* s, const Row<T>* A, const T* x, const T* b, TR* r, const Dims& d,
                    double* partial, const Ctl* ctl) {
    if (use_tm(d)) {
        dim3 grid;
        int zc;
        tm_grid(d, grid, zc);
        k_residual_tile<T, TR, DOT, REFRESH, GATED><<<grid_sys(grid, s), dim3(kTmX, kTmY, 1), 0,
                                                      s->stream>>>(A, x, b, r, d, zc, partial, ctl);
        return (int)(grid.x * grid.y * grid.z);
    }
    const Geo g = geo_stencil(d);
    k_residual<T, T, TR, DOT, REFRESH, GATED><<<grid_sys(g.grid, s), g.block, 0, s->stream>>>(
        A, x, b, r, d, g.zc, partial, ctl);
    return nblocks(g);
}

template <typename T>
struct MgCfg {
    unsigned pre, post0, post, coarsest;
    T omega;
    int pairs = 0;         // jg_gpu_options.sweep_pairs
    bool jacobi = false;   // weighted-Jacobi smoothing (JG_SMOOTH_JACOBI)
};

// The V-cycle with weighted Jacobi smoothing (JG_SMOOTH_JACOBI): zero initial
// guess, `pre` sweeps x += omega ((b - s) / c - x) (FdmJacobiSolver3::relax,
// src/jet/fdm_jacobi_solver3.cpp:95-112, weighted), the residual and Jet's
// restriction, recursion, Jet's correction, `post` sweeps; `coarsest` sweeps
// on the coarsest level.  Every level keeps its current iterate in x and
// sweeps out of place into y, copying back after an odd count; 2-D
// hierarchies take FdmMgUtils2's transfers and the one-plane residual.
template <typename T>
int vcycle_jacobi(jg_solver* s, const T* r0, T* z0, const MgCfg<T>& c, const Ctl* ctl) {
    const size_t L = s->lv.size();
    auto B = [&](size_t l) -> const T* { return l == 0 ? r0 : Prec<T>::b(s->lv[l]); };
    auto X = [&](size_t l) -> T* { return l == 0 ? z0 : Prec<T>::x(s->lv[l]); };
    // `count` sweeps of level l's x (through y; the result ends in x)
    auto sweeps = [&](size_t l, unsigned count) {
        Level& lv = s->lv[l];
        const Dims& d = lv.d;
        const Geo g = geo_stencil(d);
        T* x = X(l);
        T* y = Prec<T>::y(lv);
        const T* in = x;
        for (unsigned k = 0; k < count; ++k) {
            T* out = in == x ? y : x;
            k_jacobi<T, T, true><<<g.grid, g.block, 0, s->stream>>>(Prec<T>::A(lv), B(l), in, out, d,
                                                                   g.zc, c.omega, ctl);
            in = out;
        }
        if (in != x) k_convert<T, T><<<nblocks_1d(d.n()), NT, 0, s->stream>>>(in, x, d.n());
    };
    for (size_t l = 0; l < L; ++l) {
        Level& lv = s->lv[l];
        const Dims& d = lv.d;
        k_fill<T><<<nblocks_1d(d.n()), NT, 0, s->stream>>>(X(l), T(0), d.n());
        if (l + 1 == L) {
            sweeps(l, c.coarsest);
            break;
        }
        sweeps(l, c.pre);
        const Dims& cd = s->lv[l + 1].d;
        const Geo gc = geo_planes(cd);
        if (s->twod) {
            const Geo gf = geo_planes(d);
            k_residual_plane<T><<<gf.grid, gf.block, 0, s->stream>>>(Prec<T>::A(lv), X(l), B(l),
                                                                    Prec<T>::t(lv), d, ctl);
            k_restrict2<T><<<gc.grid, gc.block, 0, s->stream>>>(Prec<T>::t(lv),
                                                                Prec<T>::b(s->lv[l + 1]), cd, ctl);
        } else {
            const Geo gf = geo_stencil(d);
            k_residual<T, T, T, false, false, true><<<gf.grid, gf.block, 0, s->stream>>>(
                Prec<T>::A(lv), X(l), B(l), Prec<T>::t(lv), d, gf.zc, nullptr, ctl);
            k_restrict<T><<<gc.grid, gc.block, 0, s->stream>>>(Prec<T>::t(lv),
                                                               Prec<T>::b(s->lv[l + 1]), cd, ctl);
        }
        JG_TRY(check_launch(s, "jacobi vcycle down"));
    }
    for (size_t l = L - 1; l-- > 0;) {
        const Dims& d = s->lv[l].d;
        const Geo gf = geo_planes(d);
        if (s->twod)
            k_correct2<T><<<gf.grid, gf.block, 0, s->stream>>>(X(l + 1), X(l), d, ctl);
        else
            k_correct<T><<<gf.grid, gf.block, 0, s->stream>>>(X(l + 1), X(l), d, ctl);
        sweeps(l, l == 0 ? c.post0 : c.post);
        JG_TRY(check_launch(s, "jacobi vcycle up"));
    }
    return JG_OK;
}

// Two sweeps of a one-plane level in one launch (k_rb_sweep2_plane): xin
// nullptr starts from zero, coarse adds its prolongation to the start, res
// receives the residual of the result.  The residual is fused only into
// first-colour-0 (pre-smoothing) pairs and the correction only into
// first-colour-1 (post-smoothing) pairs, never both; dot (post-smoothing
// only) writes the b.x / b.b partials to slots 0 and 1 and returns their
// count (else 0).
template <typename T>
int launch_plane2(jg_solver* s, const Row<T>* A, const T* b, const T* xin, const T* coarse,
                  T* xout, T* res, const Dims& d, int first, T omega, const Ctl* ctl,
                  bool dot = false) {
    const int H = p2_halo(res != nullptr);
    const int IX = kP2X - 2 * H, IY = P2Cfg<T>::TY - 2 * H;
    const dim3 grid((d.nx + IX - 1) / IX, (d.ny + IY - 1) / IY, 1);
    const dim3 block(kP2Lanes, P2Cfg<T>::NW, 1);
    hipStream_t st = s->stream;
    double *p0 = s->slot(0), *p1 = s->slot(1);
#define JG_P2(F, XI, RE, CO, DO)                                                                \
    k_rb_sweep2_plane<T, F, XI, RE, CO, DO><<<grid, block, 0, st>>>(A, b, xin, coarse, xout, res, \
                                                                    d, omega, ctl, p0, p1)
    if (first == 0) {
        if (xin) {
            if (res) JG_P2(0, true, true, false, false);
            else JG_P2(0, true, false, false, false);
        } else {
            if (res) JG_P2(0, false, true, false, false);
            else JG_P2(0, false, false, false, false);
        }
        return 0;
    }
    if (dot) {
        if (xin) {
            if (coarse) JG_P2(1, true, false, true, true);
            else JG_P2(1, true, false, false, true);
        } else {
            if (coarse) JG_P2(1, false, false, true, true);
            else JG_P2(1, false, false, false, true);
        }
        return (int)(grid.x * grid.y);
    }
    if (xin) {
        if (coarse) JG_P2(1, true, false, true, false);
        else JG_P2(1, true, false, false, false);
    } else {
        if (coarse) JG_P2(1, false, false, true, false);
        else JG_P2(1, false, false, false, false);
    }
#undef JG_P2
    return 0;
}

// The V-cycle on a 2-D hierarchy (FdmMgLinearSystem2): the same symmetric
// cycle as vcycle() below, with FdmMgUtils2's transfers and Jet's 2-D relax.
// Consecutive sweeps of a level run two per launch (k_rb_sweep2_plane): the
// last pre-smoothing pair also writes the residual the restriction reads,
// and the first post-smoothing pair adds the prolonged correction as it
// loads x.  Leftover single sweeps (odd counts, or pairs < 0) take
// k_rb_sweep_plane, with separate k_residual_plane / k_correct2 launches
// where nothing fuses them; a coarsest level of at most kCoarseMaxCells
// cells runs all its sweeps in one workgroup (k_coarsest_plane).  The
// result is bit-identical either way.
template <typename T>
int vcycle2d(jg_solver* s, const T* r0, T* z0, const MgCfg<T>& c, const Ctl* ctl, int* dot_blocks) {
    const size_t L = s->lv.size();
    auto B = [&](size_t l) -> const T* { return l == 0 ? r0 : Prec<T>::b(s->lv[l]); };
    auto X = [&](size_t l) -> T* { return l == 0 ? z0 : Prec<T>::x(s->lv[l]); };
    const unsigned fwd = std::max(1u, (c.coarsest + 1) / 2);
    // one sweep of a one-plane level: k_rb_sweep_plane, 60 x 12 interiors
    auto sweep = [&](size_t l, const T* in, T* out, int first) {
        Level& lv = s->lv[l];
        const Dims& d = lv.d;
        const dim3 grid((d.nx + kSwIX - 1) / kSwIX, (d.ny + kSwIY - 1) / kSwIY, 1);
        const dim3 block(kSwLX, kSwLY, 1);
        const Row<T>* A = Prec<T>::A(lv);
        if (first == 0) {
            if (in)
                k_rb_sweep_plane<T, 0, false><<<grid, block, 0, s->stream>>>(A, B(l), in, out, d, c.omega, ctl);
            else
                k_rb_sweep_plane<T, 0, true><<<grid, block, 0, s->stream>>>(A, B(l), in, out, d, c.omega, ctl);
        } else {
            if (in)
                k_rb_sweep_plane<T, 1, false><<<grid, block, 0, s->stream>>>(A, B(l), in, out, d, c.omega, ctl);
            else
                k_rb_sweep_plane<T, 1, true><<<grid, block, 0, s->stream>>>(A, B(l), in, out, d, c.omega, ctl);
        }
    };
    // n sweeps of one colour order as launches of 2 and 1 sweeps: a leftover
    // single first (lead = true: the pre-smoothing, whose last pair carries
    // the residual) or last (the post-smoothing, whose first pair carries the
    // correction)
    auto plan = [&](unsigned n, bool lead) {
        std::vector<int> p;
        if (c.pairs < 0) return std::vector<int>(n, 1);
        if (lead && n % 2) p.push_back(1);
        for (unsigned k = 0; k + 1 < n; k += 2) p.push_back(2);
        if (!lead && n % 2) p.push_back(1);
        return p;
    };
    // Launch m of a level's M launches writes ((M-1-m) even ? x : y), so the
    // last lands in x; an unfused correction is added in place to the buffer
    // the pre-smoothing left (zeroed first when there is none).
    std::vector<const T*> pre_out(L, nullptr);
    for (size_t l = 0; l < L; ++l) {
        Level& lv = s->lv[l];
        const Dims& d = lv.d;
        const Row<T>* A = Prec<T>::A(lv);
        T* Y = Prec<T>::y(lv);
        if (l + 1 == L && c.pairs >= 0 && d.n() <= (idx_t)kCoarseMaxCells) {
            // every coarsest sweep in one workgroup
            k_coarsest_plane<T><<<1, kCoarseThreads, 0, s->stream>>>(A, B(l), X(l), d, (int)fwd,
                                                                     c.omega, ctl);
            JG_TRY(check_launch(s, "coarsest 2-D"));
            break;
        }
        if (l + 1 == L) {
            const std::vector<int> fw = plan(fwd, true), bw = plan(fwd, false);
            const size_t M = fw.size() + bw.size();
            const T* in = nullptr;
            for (size_t m = 0; m < M; ++m) {
                T* out = ((M - 1 - m) % 2 == 0) ? X(l) : Y;
                const bool f = m < fw.size();
                const int n = f ? fw[m] : bw[m - fw.size()];
                if (n == 2)
                    launch_plane2<T>(s, A, B(l), in, nullptr, out, nullptr, d, f ? 0 : 1, c.omega, ctl);
                else
                    sweep(l, in, out, f ? 0 : 1);
                in = out;
            }
            JG_TRY(check_launch(s, "coarsest 2-D"));
            break;
        }
        const unsigned post = l == 0 ? c.post0 : c.post;
        const std::vector<int> pre = plan(c.pre, true);
        const size_t M = pre.size() + plan(post, false).size();
        const T* in = nullptr;
        bool res_done = false;
        for (size_t m = 0; m < pre.size(); ++m) {
            T* out = ((M - 1 - m) % 2 == 0) ? X(l) : Y;
            res_done = pre[m] == 2 && m + 1 == pre.size();
            if (pre[m] == 2)
                launch_plane2<T>(s, A, B(l), in, nullptr, out, res_done ? Prec<T>::t(lv) : nullptr, d,
                                 0, c.omega, ctl);
            else
                sweep(l, in, out, 0);
            in = out;
        }
        pre_out[l] = in;
        const Dims& cd = s->lv[l + 1].d;
        const Geo gc = geo_planes(cd);
        if (in) {
            if (!res_done) {
                const Geo gf = geo_planes(d);
                k_residual_plane<T><<<gf.grid, gf.block, 0, s->stream>>>(A, in, B(l), Prec<T>::t(lv),
                                                                        d, ctl);
            }
            k_restrict2<T><<<gc.grid, gc.block, 0, s->stream>>>(Prec<T>::t(lv),
                                                                Prec<T>::b(s->lv[l + 1]), cd, ctl);
        } else {
            k_restrict2<T><<<gc.grid, gc.block, 0, s->stream>>>(B(l), Prec<T>::b(s->lv[l + 1]), cd,
                                                                ctl);
        }
        JG_TRY(check_launch(s, "vcycle 2-D down"));
    }
    for (size_t l = L - 1; l-- > 0;) {
        Level& lv = s->lv[l];
        const Dims& d = lv.d;
        const Row<T>* A = Prec<T>::A(lv);
        T* Y = Prec<T>::y(lv);
        const unsigned post_n = l == 0 ? c.post0 : c.post;
        const size_t P = plan(c.pre, true).size();
        const std::vector<int> post = plan(post_n, false);
        const size_t M = P + post.size();
        const T* in = pre_out[l];
        const bool fused = !post.empty() && post[0] == 2;
        if (!fused) {
            T* cur = const_cast<T*>(in);
            if (!cur) {
                cur = (M % 2 == 0) ? X(l) : Y;
                k_fill<T><<<nblocks_1d(d.n()), NT, 0, s->stream>>>(cur, T(0), d.n());
            }
            const Geo gf = geo_planes(d);
            k_correct2<T><<<gf.grid, gf.block, 0, s->stream>>>(X(l + 1), cur, d, ctl);
            in = cur;
        }
        for (size_t q = 0; q < post.size(); ++q) {
            T* out = ((M - 1 - (P + q)) % 2 == 0) ? X(l) : Y;
            // the last level-0 launch also forms PCG's r.z and r.r partials
            const bool dot = l == 0 && q + 1 == post.size();
            if (post[q] == 2)
                *dot_blocks = launch_plane2<T>(s, A, B(l), in, q == 0 ? X(l + 1) : nullptr, out,
                                               nullptr, d, 1, c.omega, ctl, dot);
            else
                sweep(l, in, out, 1);
            in = out;
        }
        JG_TRY(check_launch(s, "vcycle 2-D up"));
    }
    return JG_OK;
}

// z = M r: one symmetric V-cycle with zero initial guess.  Level 0 reads
// its right-hand side from r0 and writes its result to z0.  Each level runs
// `pre` forward sweeps (red then black) from x = 0, restricts the residual,
// recurses, then `post` backward sweeps (black then red), the first of which
// adds the prolonged coarse correction as it loads x.  Sweeps ping-pong
// between the level's x and y vectors so that the last one lands in x.
// Returns the number of b.x / b.b partials the last level-0 sweep wrote to
// slots 0 and 1 (0 if it wrote none).
template <typename T>
int vcycle(jg_solver* s, const T* r0, T* z0, const MgCfg<T>& c, const Ctl* ctl, int* dot_blocks) {
    const size_t L = s->lv.size();
    *dot_blocks = 0;
    if ((c.jacobi || s->twod) && s->nsys > 1) return JG_ERR_UNSUPPORTED;  // batches are 3-D red-black
    if (c.jacobi) return vcycle_jacobi<T>(s, r0, z0, c, ctl);
    if (s->twod) return vcycle2d<T>(s, r0, z0, c, ctl, dot_blocks);
    auto B = [&](size_t l) -> const T* { return l == 0 ? r0 : Prec<T>::b(s->lv[l]); };
    auto X = [&](size_t l) -> T* { return l == 0 ? z0 : Prec<T>::x(s->lv[l]); };
    std::vector<const T*> pre_out(L, nullptr);
    // Levels from `bot` down run in one workgroup (k_bottom).
    size_t bot = L;
    for (size_t l = 0; l < L; ++l)
        if (s->lv[l].d.n() <= BotCfg<T>::MaxCells) {
            bot = l;
            break;
        }
    const int fwd = (int)std::max(1u, (c.coarsest + 1) / 2);
    size_t top = L;  // first level of the upward pass
    for (size_t l = 0; l < L; ++l) {
        Level& lv = s->lv[l];
        const Row<T>* A = Prec<T>::A(lv);
        const Dims& d = lv.d;
        T* Y = Prec<T>::y(lv);
        if (l == bot) {
            if (L - l > (size_t)kBotMaxLevels) return JG_ERR_UNSUPPORTED;
            BotLevels<T> bl;
            bl.L = (int)(L - l);
            bl.d[0] = d;
            for (size_t m = l; m < L; ++m) bl.A[m - l] = Prec<T>::A(s->lv[m]);
            const int post_top = (int)(l == 0 ? c.post0 : c.post);
            if (d.nx % 2 == 0)
                k_bottom<T, true><<<s->nsys, kBotThreads, 0, s->stream>>>(bl, B(l), X(l), (int)c.pre,
                                                                    post_top, (int)c.post, fwd,
                                                                    c.omega, ctl);
            else
                k_bottom<T, false><<<s->nsys, kBotThreads, 0, s->stream>>>(bl, B(l), X(l), (int)c.pre,
                                                                     post_top, (int)c.post, fwd,
                                                                     c.omega, ctl);
            JG_TRY(check_launch(s, "bottom"));
            top = l;
            break;
        }
        if (l + 1 == L) {
            // coarsest level too large for one workgroup: symmetric fused sweeps
            const int S = 2 * fwd;
            const T* in = nullptr;
            for (int k = 0; k < S; ++k) {
                T* out = ((S - 1 - k) % 2 == 0) ? X(l) : Y;
                const int kind = in ? kSweepPlain : kSweepZero;
                launch_sweep<T>(s, A, B(l), in, out, d, k < fwd ? 0 : 1, nullptr, nullptr, nullptr,
                                c.omega, ctl, Prec<T>::zc(lv, kind), Prec<T>::sh(lv, kind));
                in = out;
            }
            JG_TRY(check_launch(s, "coarsest"));
            top = l;
            break;
        }
        const unsigned post = l == 0 ? c.post0 : c.post;
        // launches of the level ping-pong between x and y so the last lands in x
        const std::vector<int> pre_plan = sweep_plan<T>(lv, c.pre, kSweepZero, c.pairs);
        const size_t M = pre_plan.size() + sweep_plan<T>(lv, post, kSweepCorrect, c.pairs).size();
        auto buf = [&](size_t m) -> T* { return ((M - 1 - m) % 2 == 0) ? X(l) : Y; };
        const T* in = nullptr;
        for (size_t m = 0; m < pre_plan.size(); ++m) {
            const int kind = in ? kSweepPlain : kSweepZero;
            if (pre_plan[m] == kLaunchBrick)
                launch_brick2<T>(s, A, B(l), in, buf(m), d, 0, nullptr, c.omega, ctl);
            else if (pre_plan[m] == kLaunchMarch) {
                const bool res = m + 1 == pre_plan.size();
                JG_TRY(launch_sweep2<T>(s, A, B(l), in, buf(m), d, 0, nullptr,
                                        res ? Prec<T>::t(lv) : nullptr, c.omega, ctl,
                                        res && kind == kSweepZero ? Prec<T>::zc2res(lv)
                                                                  : Prec<T>::zc2(lv, kind)));
            }
            else
                launch_sweep<T>(s, A, B(l), in, buf(m), d, 0, nullptr, nullptr, nullptr, c.omega, ctl,
                                Prec<T>::zc(lv, kind), Prec<T>::sh(lv, kind));
            in = buf(m);
        }
        pre_out[l] = in;  // nullptr: x is still zero
        const Dims& cd = s->lv[l + 1].d;
        Geo gc = geo_planes(cd);
        if (in && sizeof(T) == 8) {
            // FP64: residual into t, then restriction (the fused kernel's
            // register footprint halves its occupancy in FP64, and the pair
            // measures faster)
            launch_residual<T, T, false, false, true>(s, A, in, B(l), Prec<T>::t(lv), d, nullptr, ctl);
            k_restrict<T><<<grid_sys(gc.grid, s), gc.block, 0, s->stream>>>(Prec<T>::t(lv), Prec<T>::b(s->lv[l + 1]),
                                                               cd, ctl);
        } else if (in) {
            const bool from_t = pre_plan.back() == kLaunchMarch;  // that launch wrote the residual into t
            // restricting a stored residual is one load per cell: twice the
            // blocks (shorter chunks) keep more loads in flight
            const int zc = from_t ? rr_zc(cd, 2048) : rr_zc(cd);
            const dim3 grid = grid_sys(
                dim3((cd.nx + kRrCX - 1) / kRrCX, (cd.ny + kRrCY - 1) / kRrCY, (cd.nz + zc - 1) / zc), s);
            if (from_t)
                k_residual_restrict<T, true><<<grid, dim3(kRrLX, kRrTY, 1), 0, s->stream>>>(
                    nullptr, nullptr, Prec<T>::t(lv), Prec<T>::b(s->lv[l + 1]), d, cd, zc, ctl);
            else
                k_residual_restrict<T><<<grid, dim3(kRrLX, kRrTY, 1), 0, s->stream>>>(
                    A, in, B(l), Prec<T>::b(s->lv[l + 1]), d, cd, zc, ctl);
        } else {
            k_restrict<T><<<grid_sys(gc.grid, s), gc.block, 0, s->stream>>>(B(l), Prec<T>::b(s->lv[l + 1]), cd,
                                                               ctl);
        }
        JG_TRY(check_launch(s, "vcycle down"));
    }
    for (size_t l = top; l-- > 0;) {
        Level& lv = s->lv[l];
        const Row<T>* A = Prec<T>::A(lv);
        const Dims& d = lv.d;
        T* Y = Prec<T>::y(lv);
        const unsigned post = l == 0 ? c.post0 : c.post;
        const size_t P0 = sweep_plan<T>(lv, c.pre, kSweepZero, c.pairs).size();
        const std::vector<int> post_plan = sweep_plan<T>(lv, post, kSweepCorrect, c.pairs);
        const size_t M = P0 + post_plan.size();
        auto buf = [&](size_t m) -> T* { return ((M - 1 - m) % 2 == 0) ? X(l) : Y; };
        const T* coarse = X(l + 1);
        if (post == 0) {
            // No post-smoothing: x = pre-smoothed x (already in X) + P e.
            if (!pre_out[l]) k_fill<T><<<grid_1d(nblocks_1d(d.n()), s), NT, 0, s->stream>>>(X(l), T(0), d.n());
            Geo gf = geo_planes(d);
            k_correct<T><<<grid_sys(gf.grid, s), gf.block, 0, s->stream>>>(coarse, X(l), d, ctl);
        } else {
            const T* in = pre_out[l];
            for (size_t q = 0; q < post_plan.size(); ++q) {
                const size_t m = P0 + q;
                const bool last0 = (l == 0 && m + 1 == M);
                const int kind = q == 0 ? kSweepCorrect : kSweepPlain;
                const T* cor = q == 0 ? coarse : nullptr;
                // the two-sweep kernel leaves r.z / r.r to one fused dot kernel
                // (accumulating them in it costs more than that pass)
                const bool dot = last0 && post_plan[q] == kLaunchSingle;
                double* pbx = dot ? s->slot(0) : nullptr;
                double* pbb = dot ? s->slot(1) : nullptr;
                int nb = 0;
                if (post_plan[q] == kLaunchBrick)
                    launch_brick2<T>(s, A, B(l), in, buf(m), d, 1, cor, c.omega, ctl);
                else if (post_plan[q] == kLaunchMarch)
                    JG_TRY(launch_sweep2<T>(s, A, B(l), in, buf(m), d, 1, cor, nullptr, c.omega,
                                            ctl, Prec<T>::zc2(lv, kind)));
                else
                    nb = launch_sweep<T>(s, A, B(l), in, buf(m), d, 1, cor, pbx, pbb, c.omega, ctl,
                                         Prec<T>::zc(lv, kind), Prec<T>::sh(lv, kind));
                if (dot) *dot_blocks = nb;
                in = buf(m);
            }
        }
        JG_TRY(check_launch(s, "vcycle up"));
    }
    return JG_OK;
}

// Launch-shape and z-chunk autotuner of the fused sweep: times each
// candidate on the level with the level's own matrix and scratch vectors,
// separately for the three sweep kinds (plain, zero initial x, fused
// correction), which differ in loads and registers.  Runs once per hierarchy
// and precision, before a solve, so nothing live is overwritten.
template <typename T>
int autotune_sweeps(jg_solver* s) {
    if (sweep_zc_env() > 0) return JG_OK;
    const size_t L = s->lv.size();
    for (size_t l = 0; l < L; ++l) {
        Level& lv = s->lv[l];
        if (Prec<T>::zc(lv) > 0) continue;
        const Dims& d = lv.d;
        if (d.n() <= BotCfg<T>::MaxCells) {  // handled by k_bottom
            for (int kind = 0; kind < kNumSweepKinds; ++kind) Prec<T>::zc(lv, kind) = sweep_zc(d, 0);
            continue;
        }
        const idx_t n = d.n();
        k_fill<T><<<grid_1d(nblocks_1d(n), s), NT, 0, s->stream>>>(Prec<T>::t(lv), T(1), n);
        k_fill<T><<<grid_1d(nblocks_1d(n), s), NT, 0, s->stream>>>(Prec<T>::y(lv), T(0.5), n);
        // the correction kind reads the next level's x as its coarse vector
        // (a 2-D hierarchy corrects in a separate launch: no correction kind)
        const T* coarse = nullptr;
        if (l + 1 < L && !s->twod) {
            const idx_t cn = s->lv[l + 1].d.n();
            coarse = Prec<T>::x(s->lv[l + 1]);
            k_fill<T><<<grid_1d(nblocks_1d(cn), s), NT, 0, s->stream>>>(Prec<T>::x(s->lv[l + 1]), T(0.25), cn);
        }
        const int cands[] = {2, 4, 8, 12, 16, 24, 32, 48, 64};
        float tbest[kNumSweepKinds] = {1e30f, 1e30f, 1e30f};
        for (int kind = 0; kind < kNumSweepKinds; ++kind) {
            if (kind == kSweepCorrect && !coarse) {
                Prec<T>::zc(lv, kind) = Prec<T>::zc(lv, kSweepPlain);
                Prec<T>::sh(lv, kind) = Prec<T>::sh(lv, kSweepPlain);
                continue;
            }
            float best = 1e30f;
            int best_zc = sweep_zc(d, 0), best_sh = 0;
            for (int sh = 0; sh < kNumSweepShapes; ++sh) {
                for (int zc : cands) {
                    if (zc > d.nz) break;
                    auto run = [&]() {
                        launch_sweep<T>(s, Prec<T>::A(lv), Prec<T>::t(lv),
                                        kind == kSweepZero ? nullptr : Prec<T>::y(lv), Prec<T>::x(lv),
                                        d, 1, kind == kSweepCorrect ? coarse : nullptr, nullptr,
                                        nullptr, T(1), nullptr, zc, sh);
                    };
                    run();
                    JG_HIP(hipEventRecord(s->ev0, s->stream));
                    for (int r = 0; r < 3; ++r) run();
                    JG_HIP(hipEventRecord(s->ev1, s->stream));
                    JG_TRY(sync(s));
                    float ms = 0.f;
                    JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
                    if (tune_log())
                        std::fprintf(stderr,
                                     "jetgpu tune: %dx%dx%d %s kind %d shape %d (rp %d pd %d) zc %d: %.1f us\n",
                                     d.nx, d.ny, d.nz, sizeof(T) == 4 ? "f32" : "f64", kind, sh,
                                     kSweepShapes[sh].rp, kSweepShapes[sh].pd, zc, ms * 1e3f / 3);
                    if (ms < best) {
                        best = ms;
                        best_zc = zc;
                        best_sh = sh;
                    }
                }
            }
            Prec<T>::sh(lv, kind) = best_sh;
            Prec<T>::zc(lv, kind) = best_zc;
            tbest[kind] = best;
        }
        // Two sweeps per launch: one launch replaces a sweep of this kind
        // followed by a plain sweep, when it is faster than that pair.
        auto time3 = [&](auto&& run) -> float {
            run();
            float ms = 0.f;
            if (hipEventRecord(s->ev0, s->stream) != hipSuccess) return -1.f;
            for (int r = 0; r < 3; ++r) run();
            if (hipEventRecord(s->ev1, s->stream) != hipSuccess || sync(s) != JG_OK ||
                hipEventElapsedTime(&ms, s->ev0, s->ev1) != hipSuccess)
                return -1.f;
            return ms;
        };
        for (int kind = 0; kind < kNumSweepKinds; ++kind) {
            Prec<T>::pair(lv, kind) = kPairNone;
            if (s->twod) continue;  // the 2-D cycle launches single sweeps
            if (kind == kSweepCorrect && !coarse) continue;
            const T* xin = kind == kSweepZero ? nullptr : Prec<T>::y(lv);
            const T* cor = kind == kSweepCorrect ? coarse : nullptr;
            float best = tbest[kind] + tbest[kSweepPlain];
            if constexpr (sizeof(T) == 4) {
                const int cands2[] = {8, 12, 16, 24, 32, 48, 64};
                float best2 = 1e30f;
                int best_zc = 32;
                for (int zc : cands2) {
                    if (zc > d.nz && zc != cands2[0]) break;
                    const float ms = time3([&]() {
                        launch_sweep2<T>(s, Prec<T>::A(lv), Prec<T>::t(lv), xin, Prec<T>::x(lv), d, 1,
                                         cor, nullptr, T(1), nullptr, zc);
                    });
                    if (ms < 0.f) return JG_ERR_HIP;
                    if (tune_log())
                        std::fprintf(stderr, "jetgpu tune: %dx%dx%d f32 kind %d two-sweep zc %d: %.1f us\n",
                                     d.nx, d.ny, d.nz, kind, zc, ms * 1e3f / 3);
                    if (ms < best2) {
                        best2 = ms;
                        best_zc = zc;
                    }
                }
                lv.zc2_32[kind] = best_zc;
                if (best2 < best) {
                    best = best2;
                    Prec<T>::pair(lv, kind) = kPairMarch;
                }
                if (kind == kSweepZero) {
                    // the residual-writing form of the zero-start launch (the
                    // last pre-smoothing launch of a level), residual into y
                    float best3 = 1e30f;
                    for (int zc : cands2) {
                        if (zc > d.nz && zc != cands2[0]) break;
                        const float ms = time3([&]() {
                            launch_sweep2<T>(s, Prec<T>::A(lv), Prec<T>::t(lv), nullptr,
                                             Prec<T>::x(lv), d, 0, nullptr, Prec<T>::y(lv), T(1),
                                             nullptr, zc);
                        });
                        if (ms < 0.f) return JG_ERR_HIP;
                        if (tune_log())
                            std::fprintf(stderr,
                                         "jetgpu tune: %dx%dx%d f32 two-sweep with residual zc %d: %.1f us\n",
                                         d.nx, d.ny, d.nz, zc, ms * 1e3f / 3);
                        if (ms < best3) {
                            best3 = ms;
                            lv.zc2res_32 = zc;
                        }
                    }
                }
            }
            if (d.n() <= kBrickMaxCells) {
                const float ms = time3([&]() {
                    launch_brick2<T>(s, Prec<T>::A(lv), Prec<T>::t(lv), xin, Prec<T>::x(lv), d, 1, cor,
                                     T(1), nullptr);
                });
                if (ms < 0.f) return JG_ERR_HIP;
                if (tune_log())
                    std::fprintf(stderr, "jetgpu tune: %dx%dx%d %s kind %d brick: %.1f us\n", d.nx, d.ny,
                                 d.nz, sizeof(T) == 4 ? "f32" : "f64", kind, ms * 1e3f / 3);
                if (ms < best) {
                    best = ms;
                    Prec<T>::pair(lv, kind) = kPairBrick;
                }
            }
        }
    }
    return check_launch(s, "autotune");
}

// --------------------------------------------------------------- PCG

// Vectors and control block of one PCG solve.
template <typename T>
struct Pcg {
    const Row<T>* A;
    const T* b;
    T *x, *r, *z, *dA, *dB, *q;
    Dims d;
    Ctl* ctl;
    const MgCfg<T>* mg;  // nullptr: no preconditioner (z aliases r)
    const CsrDev* csr = nullptr;  // compressed system (FP64, no preconditioner): A unused, d = {rows, 1, 1}
    const EllDev* ell = nullptr;  // its ELL copy, when it has one
};

// z = M r and the r.z / r.r partials in slots 0 / 1 (from the V-cycle's last
// sweep when it provides them, else from dot kernels); returns the partial
// count, or minus a status.
// The fused direction update and SpMV of a PCG iteration on level-0 shape d
// (nsys systems back to back), its d.q partials in `partial`; returns their
// count per system.  One plane: k_dir_spmv_plane; rows of at least 48
// cells: the tiled z-march k_dir_spmv_tile; narrower grids keep the
// per-cell z-march k_dir_spmv (a 64-lane row would idle too many lanes).
template <typename T>
int launch_dir_spmv(jg_solver* s, const Row<T>* A, const T* z, const T* dold, T* dnew, T* q,
                    const Dims& d, double* partial, const Ctl* ctl) {
    if (d.nz == 1 && s->nsys == 1) {
        const dim3 grid((d.nx + kP2X - 1) / kP2X, (d.ny + SpCfg<T>::TY - 1) / SpCfg<T>::TY, 1);
        k_dir_spmv_plane<T><<<grid, dim3(kP2Lanes, SpCfg<T>::NW, 1), 0, s->stream>>>(
            A, z, dold, dnew, q, d, partial, ctl);
        return (int)(grid.x * grid.y);
    }
    if (use_tm(d)) {
        dim3 grid;
        int zc;
        tm_grid(d, grid, zc);
        k_dir_spmv_tile<T><<<grid_sys(grid, s), dim3(kTmX, kTmY, 1), 0, s->stream>>>(
            A, z, dold, dnew, q, d, zc, partial, ctl);
        return (int)(grid.x * grid.y * grid.z);
    }
    const Geo g = geo_stencil(d);
    k_dir_spmv<T, T><<<grid_sys(g.grid, s), g.block, 0, s->stream>>>(A, z, dold, dnew, q, d, g.zc,
                                                                     partial, ctl);
    return nblocks(g);
}

template <typename T>
int pcg_precondition(jg_solver* s, const Pcg<T>& v, const Ctl* c) {
    const idx_t n = v.d.n();
    const int nb1 = nblocks_1d(n);
    int nb = 0;
    if (v.mg) {
        const int st = vcycle<T>(s, v.r, v.z, *v.mg, c, &nb);
        if (st != JG_OK) return -st;
    }
    if (nb == 0) {
        if (v.mg)
            k_dot_rz_rr<T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.r, v.z, n, s->slot(0), s->slot(1), c);
        else
            k_dot<T, T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.r, v.z, n, s->slot(0), c);
        nb = nb1;
    }
    return nb;
}

// Start of Jet's pcg (include/jet/detail/cg-inl.h:15-50): r = b - A x, d = 0,
// z = M r, sigma = r.z.  Launches only; the control block says whether to go on.
template <typename T>
int pcg_init(jg_solver* s, const Pcg<T>& v) {
    const idx_t n = v.d.n();
    const int nb1 = nblocks_1d(n);
    // r = b - A x with x = 0 (every caller zeroes x first, as Jet's solvers
    // do): b - A 0 is b bit for bit, so a copy
    k_convert<T, T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.b, v.r, n);
    k_fill<T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.dA, T(0), n);
    k_fill<T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.dB, T(0), n);
    // (a batch group gates the preconditioner by the control blocks, which
    // say which members take part; a single system's is not set up yet)
    const int nb = pcg_precondition<T>(s, v, s->nsys > 1 ? v.ctl : nullptr);
    if (nb < 0) return -nb;
    k_fin_init<<<s->nsys, kFinThreads, 0, s->stream>>>(v.ctl, s->slot(0), v.mg ? s->slot(1) : s->slot(0), nb);
    return check_launch(s, "pcg init");
}

// One iteration of Jet's pcg loop (include/jet/detail/cg-inl.h:52-91) on the
// device; every kernel is gated by the control block, so iterations can be
// launched (or replayed from a graph) past convergence.  odd: the two
// direction buffers swap roles every iteration.
template <typename T>
int pcg_iteration(jg_solver* s, const Pcg<T>& v, bool odd) {
    const idx_t n = v.d.n();
    const int nb1 = nblocks_1d(n);
    T* dold = odd ? v.dB : v.dA;
    T* dnew = odd ? v.dA : v.dB;
    if constexpr (sizeof(T) == 8) {
        // (no stencil geometry here: geo_stencil of a {rows, 1, 1} shape
        // never meets its partial-count bound once rows > 64 * kMaxPartials)
        if (v.csr) {
            if (v.ell) {
                // d first, then q = A d gathering d alone: 477.8 ms for the 256^3
                // scene's compressed CG against 493.9 with d formed per entry
                k_ell_dir<<<nb1, NT, 0, s->stream>>>(v.z, dold, dnew, n, v.ctl);
                k_ell_spmv_dot<<<nb1, NT, 0, s->stream>>>(*v.ell, dnew, v.q, s->slot(2), v.ctl);
            } else
                k_csr_dir_spmv<<<nb1, NT, 0, s->stream>>>(*v.csr, v.z, dold, dnew, v.q, n,
                                                          s->slot(2), v.ctl);
            k_fin_alpha<<<1, kFinThreads, 0, s->stream>>>(v.ctl, s->slot(2), nb1);
            k_update_xr<T><<<nb1, NT, 0, s->stream>>>(v.x, v.r, dnew, v.q, n, v.ctl);
            if (v.ell)
                k_ell_residual<false, true><<<nb1, NT, 0, s->stream>>>(*v.ell, v.x, v.b, v.r,
                                                                       nullptr, v.ctl);
            else
                k_csr_residual<false, true><<<nb1, NT, 0, s->stream>>>(*v.csr, v.x, v.b, v.r, n,
                                                                       nullptr, v.ctl);
            const int nb = pcg_precondition<T>(s, v, v.ctl);
            if (nb < 0) return -nb;
            k_fin_sigma<<<1, kFinThreads, 0, s->stream>>>(v.ctl, s->slot(0), s->slot(0), nb);
            return check_launch(s, "csr pcg iteration");
        }
    }
    Geo g = geo_stencil(v.d);
    const int nsp = launch_dir_spmv<T>(s, v.A, v.z, dold, dnew, v.q, v.d, s->slot(2), v.ctl);
    k_fin_alpha<<<s->nsys, kFinThreads, 0, s->stream>>>(v.ctl, s->slot(2), nsp);
    k_update_xr<T><<<grid_1d(nb1, s), NT, 0, s->stream>>>(v.x, v.r, dnew, v.q, n, v.ctl);
    if (v.d.nz == 1 && s->nsys == 1)
        k_residual_plane_refresh<T><<<std::min(nb1, 1024), NT, 0, s->stream>>>(v.A, v.x, v.b, v.r,
                                                                                v.d, v.ctl);
    else
        launch_residual<T, T, false, true, false>(s, v.A, v.x, v.b, v.r, v.d, nullptr, v.ctl);
    const int nb = pcg_precondition<T>(s, v, v.ctl);
    if (nb < 0) return -nb;
    k_fin_sigma<<<s->nsys, kFinThreads, 0, s->stream>>>(v.ctl, s->slot(0), v.mg ? s->slot(1) : s->slot(0), nb);
    return check_launch(s, "pcg iteration");
}

// Jet's pcg control flow (include/jet/detail/cg-inl.h:15-94) on the device.
// Vectors of type T; the matrix of level 0 is A; M is either the identity
// (z aliases r) or the V-cycle.  x must be zero on entry (Jet zeroes it).
// The host only polls the stop flag, every check_every iterations.
template <typename T>
int pcg_device(jg_solver* s, const Pcg<T>& v, unsigned check_every) {
    JG_TRY(pcg_init<T>(s, v));
    JG_HIP(hipMemcpyAsync(s->ctl_host, v.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    if (s->ctl_host->done) return JG_OK;
    const unsigned max_iter = s->ctl_host->max_iter;
    // One iteration of lookahead: iteration k+1 is enqueued before the host
    // waits for the stop flag iteration k left (every kernel of an iteration
    // returns at once once the control block says the solve has stopped), so
    // the device does not idle while the host polls.  When the last two
    // polled iterations predict (by their ratio) that the iteration in flight
    // meets the stop test, the host waits first instead: a speculative
    // iteration after the last one still costs its launches.
    auto metric = [](const Ctl& c) { return c.stop_rule == 1 ? c.rr : c.sigma; };
    const Ctl& c0 = *s->ctl_host;  // stable until the first poll below
    const double lim = c0.stop_rule == 1 ? c0.rel2 * c0.bb : c0.tol2;
    double m_prev = -1.0, m_last = metric(c0);
    unsigned launched = 0;
    int slot = 0;
    JG_TRY(pcg_iteration<T>(s, v, false));
    ++launched;
    while (true) {
        const bool poll = launched % check_every == 0 || launched >= max_iter;
        if (poll) {
            JG_HIP(hipMemcpyAsync(s->ctl_host + slot, v.ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                                  s->stream));
            JG_HIP(hipEventRecord(s->evp[slot], s->stream));
        }
        bool more = launched <= max_iter;  // the device stops by itself at max_iter
        const bool likely_last = poll && m_prev > 0.0 && m_last > 0.0 && m_last < m_prev &&
                                 m_last * (m_last / m_prev) <= lim;
        if (more && !likely_last) {
            JG_TRY(pcg_iteration<T>(s, v, (launched & 1) != 0));
            ++launched;
            more = false;  // already launched this pass
        }
        if (poll) {
            JG_HIP(hipEventSynchronize(s->evp[slot]));
            const Ctl& c = s->ctl_host[slot];
            if (c.done) break;
            m_prev = m_last;
            m_last = metric(c);
            slot ^= 1;
        }
        if (more) {  // the prediction was wrong: go on
            JG_TRY(pcg_iteration<T>(s, v, (launched & 1) != 0));
            ++launched;
        } else if (launched > max_iter) {
            break;
        }
    }
    // final state in slot 0 for the callers (the look-ahead iteration, a
    // no-op, has drained by then)
    JG_HIP(hipMemcpyAsync(s->ctl_host, v.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    return sync(s);
}

void init_ctl(Ctl* h, unsigned max_iter, double tol, int stop_rule, double rel_tol) {
    std::memset(h, 0, sizeof(Ctl));
    h->max_iter = max_iter;
    h->tol2 = tol * tol;
    h->stop_rule = stop_rule;
    h->rel2 = rel_tol * rel_tol;
}

int upload_ctl(jg_solver* s, Ctl* dev, const Ctl& h) {
    JG_HIP(hipMemcpyAsync(dev, &h, sizeof(Ctl), hipMemcpyHostToDevice, s->stream));
    return JG_OK;
}

int ensure_f32(jg_solver* s) {
    if (s->f32_valid) return JG_OK;
    for (auto& l : s->lv)
        k_rows_to_f32<<<grid_1d(nblocks_1d(l.d.n()), s), NT, 0, s->stream>>>(l.A64, l.A32, l.d.n());
    JG_TRY(check_launch(s, "rows_to_f32"));
    s->f32_valid = true;
    return JG_OK;
}

// r = b - A x of the FP64 finest level, stored as TR, with the partials of
// r.r written to `partial`; returns their count.  One-plane levels take the
// plane kernel (the z-marching one would spend its pipeline on one plane).
template <typename TR>
int residual_dot64(jg_solver* s, const Level& l0, TR* r, double* partial) {
    if (l0.d.nz == 1) {
        const int nb = nblocks_1d(l0.d.n());
        k_residual_plane_dot<TR><<<nb, NT, 0, s->stream>>>(l0.A64, s->x, s->b, r, l0.d, partial);
        return nb;
    }
    return launch_residual<double, TR, true, false, false>(s, l0.A64, s->x, s->b, r, l0.d, partial,
                                                          nullptr);
}

// FP64 ||b - A x|| / ||b|| of the finest level (reporting).
int true_rel_residual(jg_solver* s, double* out) {
    Level& l0 = s->lv[0];
    const int nb1 = nblocks_1d(l0.d.n());
    const int nr = residual_dot64<double>(s, l0, s->r, s->slot(0));
    k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(0), nr, s->scal + 0);
    k_dot<double, double><<<nb1, NT, 0, s->stream>>>(s->b, s->b, l0.d.n(), s->slot(1), nullptr);
    k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(1), nb1, s->scal + 1);
    JG_TRY(check_launch(s, "true residual"));
    double h[2];
    JG_HIP(hipMemcpyAsync(h, s->scal, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    *out = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    return JG_OK;
}

bool valid_dims(size_t nx, size_t ny, size_t nz) {
    const size_t lim = (size_t)1 << 30;
    return nx > 0 && ny > 0 && nz > 0 && nx < 65536 && ny < 65536 && nz < 65536 &&
           nx * ny * nz < lim * 2;
}

}  // namespace

namespace {

// Options of an MGPCG solve or V-cycle.  Every level of the caller's
// hierarchy is used, as Jet's mgVCycle recurses over A.levels.size()
// (include/jet/detail/mg-inl.h:26); max_levels only tells callers how deep to
// build it.  Relaxation factor of the red-black smoother: the GPU option if
// given, else the caller's Jet sorFactor (FdmMgSolver3's, default 1.5).  Any
// 0 < omega < 2 keeps the symmetric V-cycle positive definite.  Weight of the
// Jacobi smoother: the GPU option if given, else 6/7 (the weight that damps
// the high-frequency half of the 7-point Laplacian's spectrum best); it must
// lie in (0, 1], where 2 D / omega - A stays positive definite on diagonally
// dominant rows (plain Jacobi, omega = 1, barely smooths the checkerboard
// mode).
int mg_options(jg_solver* s, const jg_mg_params* p, const jg_gpu_options* o, jg_gpu_options* opt,
               double* omega) {
    if (o)
        *opt = *o;
    else
        jg_gpu_options_default(opt);
    if (opt->check_every == 0) opt->check_every = 1;
    if (opt->smoother == JG_SMOOTH_JACOBI) {
        *omega = opt->omega > 0 ? opt->omega : 6.0 / 7.0;
        if (!(*omega > 0.0 && *omega <= 1.0)) {
            s->err = "Jacobi smoother weight must lie in (0, 1]";
            return JG_ERR_INVALID_ARG;
        }
        return JG_OK;
    }
    if (opt->smoother != JG_SMOOTH_RBGS) {
        s->err = "unsupported smoother";
        return JG_ERR_UNSUPPORTED;
    }
    *omega = opt->omega > 0 ? opt->omega : p->sor;
    if (!(*omega > 0.0 && *omega < 2.0)) {
        s->err = "smoother relaxation factor must lie in (0, 2)";
        return JG_ERR_INVALID_ARG;
    }
    return JG_OK;
}

}  // namespace

// ====================================================================== C ABI

// The ELL copy of the resident compressed system when its rows are short
// enough (the longest row is found on the device); otherwise ell_w = 0 and
// the CG runs the CSR kernels.
static int csr_prepare_ell(jg_solver* s) {
    jg_solver::Csr& c = s->csr;
    if (c.ell_valid) return JG_OK;
    c.ell_w = 0;
    const int w = c.max_row;
    if (c.rows == 0 || w <= 0 || w > kEllMaxW) {
        c.ell_valid = true;  // the CSR kernels run
        return JG_OK;
    }
    const size_t cells = (size_t)w * c.rows;
    if (cells > c.ell_cap) {
        JG_TRY(dalloc(s, &c.ell_val, cells));
        JG_TRY(dalloc(s, &c.ell_ci, cells));
        c.ell_cap = cells;
    }
    if (c.rows > c.ell_rows_cap) {
        JG_TRY(dalloc(s, &c.ell_len, c.rows));
        c.ell_rows_cap = c.rows;
    }
    k_csr_to_ell<<<nblocks_1d((idx_t)c.rows), NT, 0, s->stream>>>(
        CsrDev{c.rp, c.ci, c.val, (int)c.nnz}, (idx_t)c.rows, c.ell_val, c.ell_ci, c.ell_len, w);
    JG_TRY(check_launch(s, "csr to ell"));
    c.ell_w = w;
    c.ell_valid = true;
    return JG_OK;
}

// FdmCgSolver3::solveCompressed on the solver's resident compressed system
// (A, b; x is overwritten): Jet's cg through pcg_device with the CSR kernels.
static int csr_solve(jg_solver* s, unsigned max_iter, double tol, jg_solve_info* info) {
    jg_solver::Csr& c = s->csr;
    JG_TRY(csr_prepare_ell(s));
    const CsrDev m{c.rp, c.ci, c.val, (int)c.nnz};
    const idx_t n = (idx_t)c.rows;
    const int nb1 = nblocks_1d(n);
    Ctl h;
    init_ctl(&h, max_iter, tol, 0, 0.0);
    JG_TRY(upload_ctl(s, s->ctl, h));
    k_fill<double><<<nb1, NT, 0, s->stream>>>(c.x, 0.0, n);  // Jet zeroes x (fdm_cg_solver3.cpp:62)
    JG_HIP(hipEventRecord(s->ev0, s->stream));
    Pcg<double> v{nullptr, c.b, c.x, c.r, c.r, c.dA, c.dB, c.q, Dims{(int)c.rows, 1, 1}, s->ctl, nullptr};
    v.csr = &m;
    const EllDev e{c.ell_val, c.ell_ci, c.ell_len, (idx_t)c.rows};
    if (c.ell_w > 0) v.ell = &e;
    JG_TRY(pcg_device<double>(s, v, 1));
    JG_HIP(hipEventRecord(s->ev1, s->stream));
    JG_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    const Ctl ch = *s->ctl_host;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        float ms = 0.f;
        JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        info->solve_ms = ms;
        info->iterations = ch.iter;
        info->last_residual = std::sqrt(std::fabs(ch.sigma));
        // FdmCgSolver3::solveCompressed's return (src/jet/fdm_cg_solver3.cpp:72-73)
        info->converged = (info->last_residual <= tol) || (ch.iter < max_iter);
        // ||b - A x|| / ||b|| in FP64 (reporting)
        k_csr_residual<true, false><<<nb1, NT, 0, s->stream>>>(m, c.x, c.b, c.r, n, s->slot(0), nullptr);
        k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(0), nb1, s->scal + 0);
        k_dot<double, double><<<nb1, NT, 0, s->stream>>>(c.b, c.b, n, s->slot(1), nullptr);
        k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(1), nb1, s->scal + 1);
        JG_TRY(check_launch(s, "csr true residual"));
        double sc[2];
        JG_HIP(hipMemcpyAsync(sc, s->scal, sizeof(sc), hipMemcpyDeviceToHost, s->stream));
        JG_TRY(sync(s));
        info->rel_residual = sc[1] > 0 ? std::sqrt(sc[0] / sc[1]) : std::sqrt(sc[0]);
    }
    if (ch.nonfinite) {
        s->err = "non-finite value in CG reduction";
        return JG_ERR_NONFINITE;
    }
    return JG_OK;
}

// Room for a compressed system of rows x nnz and its CG vectors (grown when
// needed: the old arrays stay owned by the solver until it is released), and
// the reduction slots / control block if the solver has no hierarchy yet.
int csr_reserve(jg_solver* s, size_t rows, size_t nnz) {
    jg_solver::Csr& c = s->csr;
    if (rows > c.cap_rows || nnz > c.cap_nnz) {
        c.val = nullptr;
        JG_TRY(dalloc(s, &c.val, nnz));
        JG_TRY(dalloc(s, &c.rp, rows + 1));
        JG_TRY(dalloc(s, &c.ci, nnz));
        for (double** p : {&c.b, &c.x, &c.r, &c.dA, &c.dB, &c.q}) JG_TRY(dalloc(s, p, rows));
        c.cap_rows = rows;
        c.cap_nnz = nnz;
    }
    if (!s->part) {
        JG_TRY(dalloc(s, &s->part, (size_t)jg_solver::kSlots * kMaxPartials));
        JG_TRY(dalloc(s, &s->scal, 16));
        JG_TRY(dalloc(s, &s->ctl, 2));
    }
    c.rows = rows;
    c.nnz = nnz;
    return JG_OK;
}

// Validates a host MatrixCsr (32-bit sizes, row pointers from 0 to nnz and
// non-decreasing, columns inside the square matrix) and uploads it to the
// solver's compressed-system arrays, growing them when needed; the CG
// vectors and reduction slots come with it.  Synchronous on return.
static int csr_upload(jg_solver* s, size_t rows, size_t nnz, const double* values, const size_t* row_ptr,
               const size_t* col_idx) {
    if (rows == 0 || !row_ptr || (nnz > 0 && (!values || !col_idx))) return JG_ERR_INVALID_ARG;
    if (rows >= (size_t(1) << 31) || nnz >= (size_t(1) << 31) || row_ptr[0] != 0 ||
        row_ptr[rows] != nnz) {
        s->err = "compressed matrix: row pointers must run from 0 to nnz (< 2^31)";
        return JG_ERR_INVALID_ARG;
    }
    std::vector<int> rp(rows + 1), ci(nnz);
    int max_row = 0;
    for (size_t i = 0; i <= rows; ++i) {
        if (i > 0 && row_ptr[i] < row_ptr[i - 1]) {
            s->err = "compressed matrix: row pointers decrease";
            return JG_ERR_INVALID_ARG;
        }
        rp[i] = (int)row_ptr[i];
        if (i > 0) max_row = std::max(max_row, rp[i] - rp[i - 1]);
    }
    for (size_t k = 0; k < nnz; ++k) {
        if (col_idx[k] >= rows) {
            s->err = "compressed matrix: column index outside the square matrix";
            return JG_ERR_INVALID_ARG;
        }
        ci[k] = (int)col_idx[k];
    }
    JG_HIP(hipSetDevice(s->device));
    JG_TRY(csr_reserve(s, rows, nnz));
    jg_solver::Csr& c = s->csr;
    c.from_grid = false;
    c.max_row = max_row;
    c.ell_valid = false;
    JG_HIP(hipMemcpyAsync(c.rp, rp.data(), (rows + 1) * sizeof(int), hipMemcpyHostToDevice, s->stream));
    if (nnz > 0) {
        JG_HIP(hipMemcpyAsync(c.ci, ci.data(), nnz * sizeof(int), hipMemcpyHostToDevice, s->stream));
        JG_HIP(hipMemcpyAsync(c.val, values, nnz * sizeof(double), hipMemcpyHostToDevice, s->stream));
    }
    return sync(s);  // the staging vectors go out of scope
}

namespace {

// Allocates the hierarchy (levels halving as checked by the caller) for
// nsys systems held back to back (jg_device.h, kSys*).
int levels_alloc(jg_solver* s, size_t nlev, const size_t* nx, const size_t* ny, const size_t* nz,
                 int nsys) {
    JG_HIP(hipSetDevice(s->device));
    JG_HIP(hipStreamSynchronize(s->stream));
    s->release();
    s->nsys = nsys;
    s->twod = nlev > 1 && nz[0] == 1;
    s->lv.resize(nlev);
    const size_t m = (size_t)nsys;
    for (size_t l = 0; l < nlev; ++l) {
        Level& lv = s->lv[l];
        lv.d = Dims{(int)nx[l], (int)ny[l], (int)nz[l]};
        const size_t n = lv.d.n() * m;
        JG_TRY(dalloc(s, &lv.A64, n));
        JG_TRY(dalloc(s, &lv.A32, n));
        JG_TRY(dalloc(s, &lv.t64, n));
        JG_TRY(dalloc(s, &lv.t32, n));
        JG_TRY(dalloc(s, &lv.markers, n));
        JG_TRY(dalloc(s, &lv.y64, n));
        JG_TRY(dalloc(s, &lv.y32, n));
        JG_TRY(dalloc(s, &lv.x64, n));
        JG_TRY(dalloc(s, &lv.x32, n));
        if (l > 0) {
            JG_TRY(dalloc(s, &lv.b64, n));
            JG_TRY(dalloc(s, &lv.b32, n));
        }
    }
    const size_t n0 = s->lv[0].d.n() * m;
    for (double** p : {&s->b, &s->x, &s->r, &s->z, &s->dA, &s->dB, &s->q}) JG_TRY(dalloc(s, p, n0));
    for (float** p : {&s->rhs32, &s->r32, &s->z32, &s->d32A, &s->d32B, &s->q32, &s->e32})
        JG_TRY(dalloc(s, p, n0));
    JG_TRY(dalloc(s, &s->part, (size_t)kSysPart * m));
    JG_TRY(dalloc(s, &s->scal, (size_t)kSysScal * m));
    JG_TRY(dalloc(s, &s->ctl, (size_t)kSysCtl * m));
    return JG_OK;
}

// Makes v a view of member i of group g: g's hierarchy and tuning, every
// pointer offset to member i's part of g's arrays, g's stream.
void make_view(jg_solver* g, int i, jg_solver* v) {
    v->view = true;
    v->nsys = 1;
    v->twod = g->twod;
    v->lv = g->lv;
    for (Level& l : v->lv) {
        const idx_t n = l.d.n() * i;
        l.A64 += n;
        l.A32 += n;
        l.t64 += n;
        l.t32 += n;
        l.markers += n;
        l.y64 += n;
        l.y32 += n;
        l.x64 += n;
        l.x32 += n;
        if (l.b64) l.b64 += n;
        if (l.b32) l.b32 += n;
    }
    const idx_t n0 = g->lv[0].d.n() * i;
    v->b = g->b + n0;
    v->x = g->x + n0;
    v->r = g->r + n0;
    v->z = g->z + n0;
    v->dA = g->dA + n0;
    v->dB = g->dB + n0;
    v->q = g->q + n0;
    v->rhs32 = g->rhs32 + n0;
    v->r32 = g->r32 + n0;
    v->z32 = g->z32 + n0;
    v->d32A = g->d32A + n0;
    v->d32B = g->d32B + n0;
    v->q32 = g->q32 + n0;
    v->e32 = g->e32 + n0;
    v->part = g->part + kSysPart * i;
    v->scal = g->scal + kSysScal * i;
    v->ctl = g->ctl + kSysCtl * i;
    if (v->stream) (void)hipStreamDestroy(v->stream);
    v->stream = g->stream;
}

}  // namespace

extern "C" {

void jg_mg_params_default(jg_mg_params* p) {
    // MgParameters defaults (include/jet/mg.h:58-86) and FdmMgSolver3's
    // sorFactor / ordering defaults (include/jet/fdm_mg_solver3.h:24-30).
    p->max_levels = 1;
    p->restrict_iter = 5;
    p->correct_iter = 5;
    p->coarsest_iter = 20;
    p->final_iter = 20;
    p->max_tol = 1e-9;
    p->sor = 1.5;
    p->red_black = 0;
}

void jg_gpu_options_default(jg_gpu_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->precision = JG_FP64;
    o->stop_rule = JG_STOP_JET;
    o->rel_tol = 1e-8;
    o->smoother = JG_SMOOTH_RBGS;
    o->omega = 0.0;
    o->max_refine = 8;
    o->inner_rel_tol = 1e-4;
    o->check_every = 1;
}

int jg_solver_create(int device, jg_solver** out) {
    if (!out) return JG_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
        return JG_ERR_NO_DEVICE;
    jg_solver* s = new jg_solver();
    s->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&s->evp[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->evp[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&s->ctl_host, 2 * sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
        delete s;
        return JG_ERR_HIP;
    }
    // this device's address of the no-allocation scratch array (a failure is
    // reported by jg_relax_red_black_noalloc, the only user)
    (void)scratch_init(device);
    *out = s;
    return JG_OK;
}

void jg_solver_destroy(jg_solver* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    delete s;
}

const char* jg_last_error(const jg_solver* s) { return s ? s->err.c_str() : "null solver"; }

void* jg_solver_stream(jg_solver* s) { return s ? (void*)s->stream : nullptr; }

int jg_solver_set_levels(jg_solver* s, size_t nlev, const size_t* nx, const size_t* ny,
                         const size_t* nz) {
    if (!s || nlev == 0 || nlev > 16 || !nx || !ny || !nz) return JG_ERR_INVALID_ARG;
    for (size_t l = 0; l < nlev; ++l) {
        if (!valid_dims(nx[l], ny[l], nz[l])) {
            s->err = "invalid level size";
            return JG_ERR_INVALID_ARG;
        }
        // a finest level one plane thick with coarser levels is a 2-D
        // hierarchy (a 3-D one cannot halve an odd nz): every level one plane
        const bool twod = nlev > 1 && nz[0] == 1;
        if (l > 0 && (nx[l - 1] != 2 * nx[l] || ny[l - 1] != 2 * ny[l] ||
                      (twod ? nz[l] != 1 : nz[l - 1] != 2 * nz[l]))) {
            s->err = "level sizes must halve exactly (in x and y only for one-plane levels)";
            return JG_ERR_INVALID_ARG;
        }
        // one reduction partial per stencil tile of a plane at least
        const Geo g = geo_stencil(Dims{(int)nx[l], (int)ny[l], (int)nz[l]});
        if ((long)g.grid.x * g.grid.y * g.grid.z > kMaxPartials) {
            s->err = "level plane too large for the reduction partials (more than 65536 tiles of 64 x 4 cells)";
            return JG_ERR_UNSUPPORTED;
        }
    }
    bool same = s->lv.size() == nlev;
    for (size_t l = 0; same && l < nlev; ++l)
        same = s->lv[l].d.nx == (int)nx[l] && s->lv[l].d.ny == (int)ny[l] && s->lv[l].d.nz == (int)nz[l];
    if (same) return JG_OK;
    if (s->view) {
        s->err = "a batch member keeps the hierarchy its batch was created with";
        return JG_ERR_INVALID_ARG;
    }
    return levels_alloc(s, nlev, nx, ny, nz, 1);
}

double* jg_solver_matrix(jg_solver* s, size_t level) {
    if (!s || level >= s->lv.size()) return nullptr;
    return (double*)s->lv[level].A64;
}

double* jg_solver_rhs(jg_solver* s) { return s ? s->b : nullptr; }

double* jg_solver_solution(jg_solver* s) { return s ? s->x : nullptr; }

int jg_solver_upload(jg_solver* s, size_t level, const double* rows, const double* b) {
    if (!s || level >= s->lv.size() || !rows) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    const size_t n = s->lv[level].d.n();
    JG_HIP(hipMemcpyAsync(s->lv[level].A64, rows, n * sizeof(Row<double>), hipMemcpyHostToDevice,
                          s->stream));
    if (b && level == 0)
        JG_HIP(hipMemcpyAsync(s->b, b, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    s->f32_valid = false;
    return sync(s);
}

int jg_solver_download_x(jg_solver* s, double* x) {
    if (!s || s->lv.empty() || !x) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    JG_HIP(hipMemcpyAsync(x, s->x, s->lv[0].d.n() * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    return sync(s);
}

int jg_solver_matrices_changed(jg_solver* s) {
    if (!s) return JG_ERR_INVALID_ARG;
    s->f32_valid = false;
    return JG_OK;
}

int jg_solver_cg(jg_solver* s, unsigned max_iter, double tol, jg_solve_info* info) {
    if (!s || s->lv.empty()) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Level& l0 = s->lv[0];
    Ctl h;
    init_ctl(&h, max_iter, tol, 0, 0.0);
    JG_TRY(upload_ctl(s, s->ctl, h));
    k_fill<double><<<nblocks_1d(l0.d.n()), NT, 0, s->stream>>>(s->x, 0.0, l0.d.n());
    JG_HIP(hipEventRecord(s->ev0, s->stream));
    JG_TRY(pcg_device<double>(s, Pcg<double>{l0.A64, s->b, s->x, s->r, s->r, s->dA, s->dB, s->q, l0.d,
                                              s->ctl, nullptr},
                              1));
    JG_HIP(hipEventRecord(s->ev1, s->stream));
    JG_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    const Ctl& c = *s->ctl_host;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        float ms = 0.f;
        JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        info->solve_ms = ms;
        info->iterations = c.iter;
        info->last_residual = std::sqrt(std::fabs(c.sigma));
        // FdmCgSolver3::solve's return (src/jet/fdm_cg_solver3.cpp:45-46)
        info->converged = (info->last_residual <= tol) || (c.iter < max_iter);
        JG_TRY(true_rel_residual(s, &info->rel_residual));
    }
    if (c.nonfinite) {
        s->err = "non-finite value in CG reduction";
        return JG_ERR_NONFINITE;
    }
    return JG_OK;
}

int jg_mg_vcycle(jg_solver* s, const jg_mg_params* p, const jg_gpu_options* o, const double* r,
                 double* z) {
    if (!s || s->lv.empty() || !p || !r || !z || r == z) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    jg_gpu_options opt;
    double omega = 0.0;
    JG_TRY(mg_options(s, p, o, &opt, &omega));
    const idx_t n = s->lv[0].d.n();
    const int nb1 = nblocks_1d(n);
    int nb = 0;
    if (opt.precision == JG_MIXED) {
        JG_TRY(ensure_f32(s));
        JG_TRY(autotune_sweeps<float>(s));
        MgCfg<float> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter,
                        (float)omega};
        mg.pairs = opt.sweep_pairs;
        mg.jacobi = opt.smoother == JG_SMOOTH_JACOBI;
        k_convert<double, float><<<nb1, NT, 0, s->stream>>>(r, s->r32, n);
        JG_TRY(vcycle<float>(s, s->r32, s->z32, mg, nullptr, &nb));
        k_convert<float, double><<<nb1, NT, 0, s->stream>>>(s->z32, z, n);
    } else {
        JG_TRY(autotune_sweeps<double>(s));
        MgCfg<double> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter, omega};
        mg.pairs = opt.sweep_pairs;
        mg.jacobi = opt.smoother == JG_SMOOTH_JACOBI;
        JG_TRY(vcycle<double>(s, r, z, mg, nullptr, &nb));
    }
    JG_TRY(check_launch(s, "vcycle"));
    return sync(s);
}

// FdmMgSolver3::solve(FdmMgLinearSystem3*) (src/jet/fdm_mg_solver3.cpp:69-74):
// one V-cycle from the resident x.  For the linear V-cycle this is
// x += M (b - A x) with M the zero-start cycle of jg_mg_vcycle; Jet's result
// is lastResidualNorm < maxTolerance, the norm of b - A x afterwards
// (include/jet/detail/mg-inl.h:62-66).
int jg_solver_mg(jg_solver* s, const jg_mg_params* p, const jg_gpu_options* o, jg_solve_info* info) {
    if (!s || s->lv.empty() || !p) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Level& l0 = s->lv[0];
    const idx_t n = l0.d.n();
    const int nb1 = nblocks_1d(n);
    JG_HIP(hipEventRecord(s->ev0, s->stream));
    launch_residual<double, double, false, false, false>(s, l0.A64, s->x, s->b, s->r, l0.d, nullptr,
                                                         nullptr);
    JG_TRY(check_launch(s, "mg residual"));
    JG_TRY(jg_mg_vcycle(s, p, o, s->r, s->z));
    k_axpy<double><<<nb1, NT, 0, s->stream>>>(1.0, s->z, s->x, s->x, n);
    const int nr = launch_residual<double, double, true, false, false>(s, l0.A64, s->x, s->b, s->r,
                                                                       l0.d, s->slot(0), nullptr);
    k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(0), nr, s->scal + 0);
    k_dot<double, double><<<nb1, NT, 0, s->stream>>>(s->b, s->b, n, s->slot(1), nullptr);
    k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(1), nb1, s->scal + 1);
    JG_HIP(hipEventRecord(s->ev1, s->stream));
    JG_TRY(check_launch(s, "mg solve"));
    double h[2];
    JG_HIP(hipMemcpyAsync(h, s->scal, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    if (!std::isfinite(h[0])) {
        s->err = "non-finite value in the V-cycle";
        if (info) info->converged = 0;
        return JG_ERR_NONFINITE;
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        float ms = 0.f;
        JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        info->solve_ms = ms;
        info->iterations = 1;
        info->last_residual = std::sqrt(h[0]);
        info->rel_residual = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
        info->converged = info->last_residual < p->max_tol;
    }
    return JG_OK;
}

int jg_solver_mgpcg(jg_solver* s, unsigned max_iter, const jg_mg_params* p,
                    const jg_gpu_options* o, jg_solve_info* info) {
    if (!s || s->lv.empty() || !p) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    jg_gpu_options opt;
    double omega = 0.0;
    JG_TRY(mg_options(s, p, o, &opt, &omega));
    Level& l0 = s->lv[0];
    const idx_t n = l0.d.n();
    int status = JG_OK;
    unsigned iters = 0, refinements = 0;
    double last_res = 0.0;
    double bnorm = 0.0;  // mixed: ||b||, for the relative residual of the final x
    double mixed_goal = -1.0;  // mixed: the FP64 residual norm the refinement had to reach
    Ctl* ctl = s->ctl;
    if (opt.precision == JG_MIXED) {
        JG_TRY(ensure_f32(s));
        JG_TRY(autotune_sweeps<float>(s));
    } else {
        JG_TRY(autotune_sweeps<double>(s));
    }
    k_fill<double><<<nblocks_1d(n), NT, 0, s->stream>>>(s->x, 0.0, n);
    JG_HIP(hipEventRecord(s->ev0, s->stream));
    if (opt.precision != JG_MIXED) {
        MgCfg<double> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter, omega};
        mg.pairs = opt.sweep_pairs;
        mg.jacobi = opt.smoother == JG_SMOOTH_JACOBI;
        Ctl h;
        init_ctl(&h, max_iter, p->max_tol, opt.stop_rule == JG_STOP_RELATIVE ? 1 : 0, opt.rel_tol);
        JG_TRY(upload_ctl(s, ctl, h));
        JG_TRY(pcg_device<double>(s, Pcg<double>{l0.A64, s->b, s->x, s->r, s->z, s->dA, s->dB, s->q,
                                                  l0.d, ctl, &mg},
                                  opt.check_every));
        JG_HIP(hipMemcpyAsync(s->ctl_host, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
        JG_TRY(sync(s));
        iters = s->ctl_host->iter;
        last_res = std::sqrt(std::fabs(s->ctl_host->sigma));
        if (s->ctl_host->nonfinite) status = JG_ERR_NONFINITE;
    } else {
        // Iterative refinement: FP64 residual and correction, FP32 MGPCG
        // (FP32 stencil, smoother and V-cycle; FP64 dot accumulation).
        MgCfg<float> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter,
                        (float)omega};
        mg.pairs = opt.sweep_pairs;
        mg.jacobi = opt.smoother == JG_SMOOTH_JACOBI;
        const int nb1 = nblocks_1d(n);
        k_dot<double, double><<<nb1, NT, 0, s->stream>>>(s->b, s->b, n, s->slot(3), nullptr);
        k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(3), nb1, s->scal + 2);
        double bb = 0.0;
        JG_HIP(hipMemcpyAsync(&bb, s->scal + 2, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        JG_TRY(sync(s));
        bnorm = std::sqrt(bb);
        const double goal = opt.stop_rule == JG_STOP_RELATIVE ? opt.rel_tol * bnorm : p->max_tol;
        Ctl* ictl = s->ctl + 1;
        for (unsigned outer = 0;; ++outer) {
            double rr = 0.0;
            if (outer == 0) {
                // x = 0: b - A x is b itself (bit for bit), and its norm is known
                k_convert<double, float><<<nb1, NT, 0, s->stream>>>(s->b, s->rhs32, n);
                rr = bb;
            } else {
                const int nr = residual_dot64<float>(s, l0, s->rhs32, s->slot(3));
                k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(3), nr, s->scal + 3);
                JG_HIP(hipMemcpyAsync(&rr, s->scal + 3, sizeof(double), hipMemcpyDeviceToHost,
                                      s->stream));
                JG_TRY(sync(s));
            }
            last_res = std::sqrt(rr);
            if (!std::isfinite(rr)) {
                status = JG_ERR_NONFINITE;
                break;
            }
            if (last_res <= goal || outer >= opt.max_refine || iters >= max_iter) break;
            Ctl h;
            // The last refinement only needs to close the remaining gap to
            // the goal (with a factor 2 margin for the FP32 inner residual).
            const double inner = std::max(opt.inner_rel_tol, std::min(0.5, 0.5 * goal / last_res));
            init_ctl(&h, max_iter - iters, 0.0, 1, inner);
            JG_TRY(upload_ctl(s, ictl, h));
            k_fill<float><<<nb1, NT, 0, s->stream>>>(s->e32, 0.f, n);
            JG_TRY(pcg_device<float>(s, Pcg<float>{l0.A32, s->rhs32, s->e32, s->r32, s->z32, s->d32A,
                                                   s->d32B, s->q32, l0.d, ictl, &mg},
                                     opt.check_every));
            iters += s->ctl_host->iter;
            if (s->ctl_host->nonfinite) {
                // a value beyond FP32 range (or a NaN) in the inner solve: the
                // correction is meaningless, and x keeps the last finite one
                status = JG_ERR_NONFINITE;
                break;
            }
            k_add_correction<<<nb1, NT, 0, s->stream>>>(s->x, s->e32, n);
            JG_TRY(check_launch(s, "refinement"));
            ++refinements;
        }
        mixed_goal = goal;
    }
    JG_HIP(hipEventRecord(s->ev1, s->stream));
    JG_TRY(sync(s));
    if (info) {
        std::memset(info, 0, sizeof(*info));
        float ms = 0.f;
        JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        info->solve_ms = ms;
        info->iterations = iters;
        info->refinements = refinements;
        info->last_residual = last_res;
        // FP64: Jet's rule (cg-inl.h:93 and fdm_mgpcg_solver3.cpp:78). Mixed: the
        // outer loop stops on the true residual, so converged means it reached
        // the goal; running out of refinements or iterations above it is not.
        info->converged = opt.precision == JG_MIXED ? last_res <= mixed_goal
                                                    : (last_res <= p->max_tol) || (iters < max_iter);
        if (status != JG_OK) info->converged = 0;
        if (opt.precision == JG_MIXED && bnorm > 0.0)
            info->rel_residual = last_res / bnorm;  // the FP64 b - A x of the final x
        else
            JG_TRY(true_rel_residual(s, &info->rel_residual));
        if (opt.stop_rule == JG_STOP_RELATIVE && status == JG_OK)
            info->converged = info->rel_residual <= opt.rel_tol;
    }
    if (status == JG_ERR_NONFINITE) s->err = "non-finite value in MGPCG reduction";
    return status;
}

int jg_cg_solve_host(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                     const double* b, double* x, unsigned max_iter, double tol,
                     jg_solve_info* info) {
    if (!s || !rows || !b || !x) return JG_ERR_INVALID_ARG;
    size_t sx[1] = {nx}, sy[1] = {ny}, sz[1] = {nz};
    JG_TRY(jg_solver_set_levels(s, 1, sx, sy, sz));
    JG_TRY(jg_solver_upload(s, 0, rows, b));
    JG_TRY(jg_solver_cg(s, max_iter, tol, info));
    return jg_solver_download_x(s, x);
}

int jg_csr_cg_solve_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                         const size_t* row_ptr, const size_t* col_idx, const double* b, double* x,
                         unsigned max_iter, double tol, jg_solve_info* info) {
    if (!s || !b || !x) return JG_ERR_INVALID_ARG;
    JG_TRY(csr_upload(s, rows, nnz, values, row_ptr, col_idx));
    jg_solver::Csr& c = s->csr;
    JG_HIP(hipMemcpyAsync(c.b, b, rows * sizeof(double), hipMemcpyHostToDevice, s->stream));
    const int st = csr_solve(s, max_iter, tol, info);
    if (st != JG_OK && st != JG_ERR_NONFINITE) return st;
    JG_HIP(hipMemcpyAsync(x, c.x, rows * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    return st;
}

int jg_solver_csr_cg(jg_solver* s, unsigned max_iter, double tol, jg_solve_info* info) {
    if (!s || s->csr.rows == 0 || !s->csr.rp) {
        if (s) s->err = "jg_solver_csr_cg before a compressed system was built or uploaded";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    return csr_solve(s, max_iter, tol, info);
}

int jg_csr_mvm_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                    const size_t* row_ptr, const size_t* col_idx, const double* v, double* result) {
    if (!s || !v || !result) return JG_ERR_INVALID_ARG;
    JG_TRY(csr_upload(s, rows, nnz, values, row_ptr, col_idx));
    jg_solver::Csr& c = s->csr;
    JG_HIP(hipMemcpyAsync(c.x, v, rows * sizeof(double), hipMemcpyHostToDevice, s->stream));
    k_csr_mvm<<<nblocks_1d((idx_t)rows), NT, 0, s->stream>>>(CsrDev{c.rp, c.ci, c.val, (int)c.nnz}, c.x, c.q,
                                                             (idx_t)rows);
    JG_TRY(check_launch(s, "csr mvm"));
    JG_HIP(hipMemcpyAsync(result, c.q, rows * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return sync(s);
}

int jg_csr_residual_host(jg_solver* s, size_t rows, size_t nnz, const double* values,
                         const size_t* row_ptr, const size_t* col_idx, const double* x,
                         const double* b, double* result) {
    if (!s || !x || !b || !result) return JG_ERR_INVALID_ARG;
    JG_TRY(csr_upload(s, rows, nnz, values, row_ptr, col_idx));
    jg_solver::Csr& c = s->csr;
    JG_HIP(hipMemcpyAsync(c.x, x, rows * sizeof(double), hipMemcpyHostToDevice, s->stream));
    JG_HIP(hipMemcpyAsync(c.b, b, rows * sizeof(double), hipMemcpyHostToDevice, s->stream));
    k_csr_residual<false, false><<<nblocks_1d((idx_t)rows), NT, 0, s->stream>>>(
        CsrDev{c.rp, c.ci, c.val, (int)c.nnz}, c.x, c.b, c.r, (idx_t)rows, nullptr, nullptr);
    JG_TRY(check_launch(s, "csr residual"));
    JG_HIP(hipMemcpyAsync(result, c.r, rows * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return sync(s);
}

int jg_mgpcg_solve_host(jg_solver* s, size_t nlev, const size_t* nx, const size_t* ny,
                        const size_t* nz, const double* const* rows, const double* b, double* x,
                        unsigned max_iter, const jg_mg_params* p, const jg_gpu_options* o,
                        jg_solve_info* info) {
    if (!s || !rows || !b || !x || !p) return JG_ERR_INVALID_ARG;
    JG_TRY(jg_solver_set_levels(s, nlev, nx, ny, nz));
    for (size_t l = 0; l < nlev; ++l) JG_TRY(jg_solver_upload(s, l, rows[l], l == 0 ? b : nullptr));
    JG_TRY(jg_solver_mgpcg(s, max_iter, p, o, info));
    return jg_solver_download_x(s, x);
}

int jg_mg_solve_host(jg_solver* s, size_t nlev, const size_t* nx, const size_t* ny, const size_t* nz,
                     const double* const* rows, const double* b, double* x, const jg_mg_params* p,
                     const jg_gpu_options* o, jg_solve_info* info) {
    if (!s || !rows || !b || !x || !p) return JG_ERR_INVALID_ARG;
    JG_TRY(jg_solver_set_levels(s, nlev, nx, ny, nz));
    for (size_t l = 0; l < nlev; ++l) JG_TRY(jg_solver_upload(s, l, rows[l], l == 0 ? b : nullptr));
    // the V-cycle starts from the caller's x, as Jet's does
    JG_HIP(hipMemcpyAsync(s->x, x, s->lv[0].d.n() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    JG_TRY(jg_solver_mg(s, p, o, info));
    return jg_solver_download_x(s, x);
}

int jg_solver_coarsen(jg_solver* s) {
    if (!s || s->lv.empty()) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    for (size_t l = 1; l < s->lv.size(); ++l) {
        const Dims cd = s->lv[l].d;
        const Geo g = geo_planes(cd);
        if (s->twod)
            k_mg_coarsen<1><<<g.grid, g.block, 0, s->stream>>>(s->lv[l - 1].A64, s->lv[l].A64, cd);
        else
            k_mg_coarsen<2><<<g.grid, g.block, 0, s->stream>>>(s->lv[l - 1].A64, s->lv[l].A64, cd);
    }
    JG_TRY(check_launch(s, "mg_coarsen"));
    s->f32_valid = false;
    return sync(s);
}

int jg_mgpcg_flat_solve_host(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                             const double* b, double* x, unsigned max_iter, const jg_mg_params* p,
                             const jg_gpu_options* o, jg_solve_info* info) {
    if (!s || !rows || !b || !x || !p) return JG_ERR_INVALID_ARG;
    // FdmMgUtils3::resizeArrayWithFinest (include/jet/detail/fdm_mg_linear_system3-inl.h:
    // 35-50): halve while every extent is even, up to max_levels; one-plane
    // systems halve in x and y (FdmMgUtils2's rule)
    size_t lx[16] = {nx}, ly[16] = {ny}, lz[16] = {nz};
    const bool plane = nz == 1;
    const size_t cap = std::min<size_t>(std::max<size_t>(p->max_levels, 1), 16);
    size_t nlev = 1;
    while (nlev < cap && lx[nlev - 1] % 2 == 0 && ly[nlev - 1] % 2 == 0 &&
           (plane || lz[nlev - 1] % 2 == 0)) {
        lx[nlev] = lx[nlev - 1] / 2;
        ly[nlev] = ly[nlev - 1] / 2;
        lz[nlev] = plane ? 1 : lz[nlev - 1] / 2;
        ++nlev;
    }
    JG_TRY(jg_solver_set_levels(s, nlev, lx, ly, lz));
    JG_TRY(jg_solver_upload(s, 0, rows, b));
    JG_TRY(jg_solver_coarsen(s));
    JG_TRY(jg_solver_mgpcg(s, max_iter, p, o, info));
    return jg_solver_download_x(s, x);
}

}  // extern "C"

// ------------------------------------------------------ kernel timing

namespace {

template <typename T>
struct Vecs {
    T *x, *b, *r, *z, *dA, *dB, *q;
};

Vecs<double> vecs64(jg_solver* s) { return {s->x, s->b, s->r, s->z, s->dA, s->dB, s->q}; }
Vecs<float> vecs32(jg_solver* s) { return {s->e32, s->rhs32, s->r32, s->z32, s->d32A, s->d32B, s->q32}; }

// The first level-0 launch of the pre-smoothing (post = false) or of the
// post-smoothing (post = true) of a V(2,2) cycle, exactly as vcycle() issues
// it: the sweep plan and tuning of the solve, b = r, the result in z (pre:
// x starts at zero and, for a two-sweep launch, the residual is written for
// the restriction; post: the coarse correction of level 1 is added as x is
// loaded).
template <typename T>
int level0_plan(jg_solver* s, bool post) {
    return sweep_plan<T>(s->lv[0], 2, post ? kSweepCorrect : kSweepZero, 0).front();
}

template <typename T>
void level0_launch(jg_solver* s, const Vecs<T>& v, bool post, T omega) {
    Level& l0 = s->lv[0];
    const Row<T>* A = Prec<T>::A(l0);
    const int mode = level0_plan<T>(s, post);
    const T* xin = post ? Prec<T>::y(l0) : nullptr;
    const T* coarse = post && s->lv.size() > 1 ? Prec<T>::x(s->lv[1]) : nullptr;
    const int kind = post ? kSweepCorrect : kSweepZero;
    if (mode == kLaunchBrick)
        launch_brick2<T>(s, A, v.r, xin, v.z, l0.d, post ? 1 : 0, coarse, omega, s->ctl);
    else if (mode == kLaunchMarch)
        launch_sweep2<T>(s, A, v.r, xin, v.z, l0.d, post ? 1 : 0, coarse,
                         post ? nullptr : Prec<T>::t(l0), omega, s->ctl,
                         post ? Prec<T>::zc2(l0, kind) : Prec<T>::zc2res(l0));
    else
        launch_sweep<T>(s, A, v.r, xin, v.z, l0.d, post ? 1 : 0, coarse, nullptr, nullptr, omega,
                        s->ctl, Prec<T>::zc(l0, kind), Prec<T>::sh(l0, kind));
}

// Algorithmic bytes per cell of that launch under the 4-coefficient model:
// a sweep reads A, b and x and writes x; the residual written by a pre-
// smoothing two-sweep launch reads A, x and b (4 coefficients, 2 vectors)
// and writes r; the prolongation-correction reads the coarse vector (1/8)
// and reads and writes x.
template <typename T>
double level0_launch_bytes(jg_solver* s, bool post) {
    const double sa = sizeof(Row<T>), sv = sizeof(T);
    const int mode = level0_plan<T>(s, post);
    const double sweeps = mode == kLaunchSingle ? 1.0 : 2.0;
    double b = sweeps * (sa + 3 * sv);
    if (post && s->lv.size() > 1) b += 2 * sv + sv / 8;
    if (!post && mode == kLaunchMarch) b += sa + 3 * sv;
    return b;
}

template <typename T>
int time_kernel_t(jg_solver* s, const Vecs<T>& v, int kernel, unsigned reps, double* avg_ms,
                  double* bytes) {
    Level& l0 = s->lv[0];
    const Dims d = l0.d;
    const idx_t n = d.n();
    const int nb1 = nblocks_1d(n);
    Geo g = geo_stencil(d);
    const Row<T>* A = Prec<T>::A(l0);
    Ctl h;
    init_ctl(&h, 1u << 30, 0.0, 0, 0.0);
    h.alpha = 1e-3;
    h.beta = 0.5;
    h.sigma = 1.0;
    JG_TRY(upload_ctl(s, s->ctl, h));
    // the kernels that read a right-hand side take level 0's y, a V-cycle
    // scratch vector, so the system's b survives the measurement
    T* const bt = Prec<T>::y(l0);
    for (T* p : {v.x, bt, v.r, v.z, v.dA, v.dB, v.q})
        k_fill<T><<<nb1, NT, 0, s->stream>>>(p, T(1), n);
    const double sa = sizeof(Row<T>), sv = sizeof(T);
    auto launch = [&]() {
        switch (kernel) {
            case JG_KERNEL_RBGS:
                k_rbgs<T, T, false><<<g.grid, g.block, 0, s->stream>>>(A, bt, v.x, d, g.zc, 1, T(1),
                                                                       s->ctl);
                break;
            case JG_KERNEL_DIR_SPMV:
                launch_dir_spmv<T>(s, A, v.z, v.dA, v.dB, v.q, d, s->slot(2), s->ctl);
                break;
            case JG_KERNEL_RESIDUAL:
                launch_residual<T, T, false, false, false>(s, A, v.x, bt, v.r, d, nullptr, nullptr);
                break;
            case JG_KERNEL_UPDATE_XR:
                k_update_xr<T><<<nb1, NT, 0, s->stream>>>(v.x, v.r, v.dA, v.q, n, s->ctl);
                break;
            case JG_KERNEL_DOT:
                k_dot<T, T><<<nb1, NT, 0, s->stream>>>(v.r, v.z, n, s->slot(0), s->ctl);
                break;
            case JG_KERNEL_SWEEP:
                launch_sweep<T>(s, A, bt, v.r, v.x, d, 1, nullptr, nullptr, nullptr, T(1), s->ctl,
                                Prec<T>::zc(l0), Prec<T>::sh(l0));
                break;
            case JG_KERNEL_SWEEP2:
                launch_sweep2<T>(s, A, bt, v.r, v.x, d, 1, nullptr, nullptr, T(1), s->ctl,
                                 Prec<T>::zc2(l0, kSweepPlain));
                break;
            case JG_KERNEL_PRE0:
            case JG_KERNEL_POST0:
                level0_launch<T>(s, v, kernel == JG_KERNEL_POST0, T(1.3));
                break;
            default:
                k_convert<T, T><<<nb1, NT, 0, s->stream>>>(v.r, v.q, n);
                break;
        }
    };
    // Compulsory bytes per launch under Jet's 4-coefficient data model.
    double per_cell;
    switch (kernel) {
        case JG_KERNEL_RBGS: per_cell = sa + 2 * sv; break;      // A (own + neighbour couplings), x, b/2, x/2 written
        case JG_KERNEL_DIR_SPMV: per_cell = sa + 4 * sv; break;  // A, z, d_old -> d_new, q
        case JG_KERNEL_RESIDUAL: per_cell = sa + 3 * sv; break;  // A, x, b -> r
        case JG_KERNEL_UPDATE_XR: per_cell = 6 * sv; break;      // x, d, r, q -> x, r
        case JG_KERNEL_DOT: per_cell = 2 * sv; break;
        case JG_KERNEL_SWEEP: per_cell = sa + 3 * sv; break;     // A, b, x -> x (both colours)
        case JG_KERNEL_SWEEP2: per_cell = 2 * (sa + 3 * sv); break;  // two sweeps, each as above
        case JG_KERNEL_PRE0:
        case JG_KERNEL_POST0: per_cell = level0_launch_bytes<T>(s, kernel == JG_KERNEL_POST0); break;
        default: per_cell = 2 * sv; break;
    }
    launch();  // warm
    JG_TRY(check_launch(s, "time_kernel"));
    JG_HIP(hipEventRecord(s->ev0, s->stream));
    for (unsigned i = 0; i < reps; ++i) launch();
    JG_HIP(hipEventRecord(s->ev1, s->stream));
    JG_TRY(sync(s));
    float ms = 0.f;
    JG_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *avg_ms = (double)ms / reps;
    *bytes = per_cell * (double)n;
    return JG_OK;
}

}  // namespace

extern "C" int jg_solver_time_kernel(jg_solver* s, int kernel, int precision, unsigned reps,
                                     double* avg_ms, double* bytes) {
    if (!s || s->lv.empty() || !avg_ms || !bytes || reps == 0 || kernel < 0 ||
        kernel > JG_KERNEL_POST0)
        return JG_ERR_INVALID_ARG;
    if (kernel == JG_KERNEL_SWEEP2 && precision != JG_MIXED) {
        s->err = "the two-sweep kernel is FP32 only";
        return JG_ERR_UNSUPPORTED;
    }
    JG_HIP(hipSetDevice(s->device));
    if (precision == JG_MIXED) {
        JG_TRY(ensure_f32(s));
        JG_TRY(autotune_sweeps<float>(s));
        return time_kernel_t<float>(s, vecs32(s), kernel, reps, avg_ms, bytes);
    }
    JG_TRY(autotune_sweeps<double>(s));
    return time_kernel_t<double>(s, vecs64(s), kernel, reps, avg_ms, bytes);
}

extern "C" {

// --------------------------------------------------------- BLAS entries

int jg_blas_mvm(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                const double* x, double* y) {
    if (!s || !valid_dims(nx, ny, nz) || !rows || !x || !y) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims d{(int)nx, (int)ny, (int)nz};
    Geo g = geo_stencil(d);
    k_spmv<double, double, false><<<g.grid, g.block, 0, s->stream>>>(
        (const Row<double>*)rows, x, y, d, g.zc, nullptr, nullptr);
    JG_TRY(check_launch(s, "mvm"));
    return sync(s);
}

int jg_blas_residual(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                     const double* x, const double* b, double* r) {
    if (!s || !valid_dims(nx, ny, nz) || !rows || !x || !b || !r) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims d{(int)nx, (int)ny, (int)nz};
    launch_residual<double, double, false, false, false>(s, (const Row<double>*)rows, x, b, r, d,
                                                         nullptr, nullptr);
    JG_TRY(check_launch(s, "residual"));
    return sync(s);
}

int jg_blas_dot(jg_solver* s, size_t n, const double* a, const double* b, double* result) {
    if (!s || !a || !b || !result) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    if (!s->part) {
        JG_TRY(dalloc(s, &s->part, (size_t)jg_solver::kSlots * kMaxPartials));
        JG_TRY(dalloc(s, &s->scal, 16));
    }
    const int nb = nblocks_1d((idx_t)n);
    k_dot<double, double><<<nb, NT, 0, s->stream>>>(a, b, (idx_t)n, s->slot(0), nullptr);
    k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(0), nb, s->scal + 4);
    JG_TRY(check_launch(s, "dot"));
    JG_HIP(hipMemcpyAsync(result, s->scal + 4, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return sync(s);
}

int jg_blas_axpy(jg_solver* s, size_t n, double a, const double* x, const double* y, double* out) {
    if (!s || !x || !y || !out) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    k_axpy<double><<<nblocks_1d((idx_t)n), NT, 0, s->stream>>>(a, x, y, out, (idx_t)n);
    JG_TRY(check_launch(s, "axpy"));
    return sync(s);
}

int jg_relax_red_black(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                       const double* b, double sor, double* x) {
    if (!s || !valid_dims(nx, ny, nz) || !rows || !b || !x) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims d{(int)nx, (int)ny, (int)nz};
    // One fused forward sweep (red then black) reading a copy of x: the
    // sweep is out of place because its tiles overlap.
    double* tmp = nullptr;
    JG_HIP(hipMalloc(&tmp, d.n() * sizeof(double)));
    hipError_t e = hipMemcpyAsync(tmp, x, d.n() * sizeof(double), hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) {
        launch_sweep<double>(s, (const Row<double>*)rows, b, tmp, x, d, 0, nullptr, nullptr, nullptr,
                             sor, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) {
        s->err = std::string("relax_red_black: ") + hipGetErrorString(e);
        return JG_ERR_HIP;
    }
    return JG_OK;
}

int jg_relax_red_black_noalloc(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                               const double* b, double sor, double* x) {
    if (!s || !valid_dims(nx, ny, nz) || !rows || !b || !x) return JG_ERR_INVALID_ARG;
    const size_t n = nx * ny * nz;
    if (n > kNoAllocMaxCells) {
        s->err = "relax_red_black_noalloc: " + std::to_string(n) + " cells exceed the " +
                 std::to_string(kNoAllocMaxCells) + "-cell scratch array";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    double* tmp = nullptr;
    const hipError_t se = scratch_address(s->device, &tmp);
    if (se != hipSuccess) {
        s->err = std::string("relax_red_black_noalloc: no scratch address: ") + hipGetErrorString(se);
        return JG_ERR_HIP;
    }
    Dims d{(int)nx, (int)ny, (int)nz};
    // the copy of x that the out-of-place sweep reads, then the sweep itself:
    // the kernels of jg_relax_red_black, with its temporary replaced by scratch
    k_convert<double, double><<<nblocks_1d((idx_t)n), NT, 0, s->stream>>>(x, tmp, (idx_t)n);
    launch_sweep<double>(s, (const Row<double>*)rows, b, tmp, x, d, 0, nullptr, nullptr, nullptr, sor,
                         nullptr);
    JG_TRY(check_launch(s, "relax_red_black_noalloc"));
    return sync(s);
}

int jg_relax_jacobi(jg_solver* s, size_t nx, size_t ny, size_t nz, const double* rows,
                    const double* b, const double* x, double* xtemp) {
    if (!s || !valid_dims(nx, ny, nz) || !rows || !b || !x || !xtemp || x == xtemp)
        return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims d{(int)nx, (int)ny, (int)nz};
    Geo g = geo_stencil(d);
    k_jacobi<double, double><<<g.grid, g.block, 0, s->stream>>>((const Row<double>*)rows, b, x,
                                                               xtemp, d, g.zc, 1.0, nullptr);
    JG_TRY(check_launch(s, "relax_jacobi"));
    return sync(s);
}

int jg_mg_restrict(jg_solver* s, size_t cnx, size_t cny, size_t cnz, const double* fine,
                   double* coarse) {
    if (!s || !valid_dims(cnx, cny, cnz) || !fine || !coarse) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims cd{(int)cnx, (int)cny, (int)cnz};
    Geo g = geo_planes(cd);
    k_restrict<double><<<g.grid, g.block, 0, s->stream>>>(fine, coarse, cd, nullptr);
    JG_TRY(check_launch(s, "restrict"));
    return sync(s);
}

int jg_mg_correct(jg_solver* s, size_t fnx, size_t fny, size_t fnz, const double* coarse,
                  double* fine) {
    if (!s || !valid_dims(fnx, fny, fnz) || !coarse || !fine || (fnx | fny | fnz) & 1)
        return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims fd{(int)fnx, (int)fny, (int)fnz};
    Geo g = geo_planes(fd);
    k_correct<double><<<g.grid, g.block, 0, s->stream>>>(coarse, fine, fd, nullptr);
    JG_TRY(check_launch(s, "correct"));
    return sync(s);
}

int jg_mg_restrict2(jg_solver* s, size_t cnx, size_t cny, const double* fine, double* coarse) {
    if (!s || !valid_dims(cnx, cny, 1) || !fine || !coarse) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims cd{(int)cnx, (int)cny, 1};
    Geo g = geo_planes(cd);
    k_restrict2<double><<<g.grid, g.block, 0, s->stream>>>(fine, coarse, cd, nullptr);
    JG_TRY(check_launch(s, "restrict2"));
    return sync(s);
}

int jg_mg_correct2(jg_solver* s, size_t fnx, size_t fny, const double* coarse, double* fine) {
    if (!s || !valid_dims(fnx, fny, 1) || !coarse || !fine || (fnx | fny) & 1)
        return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Dims fd{(int)fnx, (int)fny, 1};
    Geo g = geo_planes(fd);
    k_correct2<double><<<g.grid, g.block, 0, s->stream>>>(coarse, fine, fd, nullptr);
    JG_TRY(check_launch(s, "correct2"));
    return sync(s);
}

}  // extern "C"

// ============================================================ batched solves
//
// Many independent systems of uneven sizes.  Systems with the same hierarchy
// form a group: one jg_solver holds all of them back to back (nsys > 1), and
// every V-cycle and PCG launch covers the whole group, each block taking its
// system from its grid index and exiting at once when that system's control
// block says it has stopped -- the per-system convergence mask.  A member is
// a view into its group (its own handle for building, uploading and reading
// its system).  The PCG iterations of all groups are captured once into one
// hipGraph (a fork from the batch stream, one branch per group, K
// iterations on that group's stream, a join); the host replays it and
// gathers every member's stop flag with one kernel and one copy per replay.

namespace {

constexpr int kBatchIters = 4;  // PCG iterations per graph replay (even: direction buffers realign)

// done flag and iteration count of every system's control block
__global__ void k_gather_ctl(Ctl* const* ctls, int n, int* out) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        out[2 * i] = ctls[i]->done | ctls[i]->nonfinite << 1;
        out[2 * i + 1] = (int)ctls[i]->iter;
    }
}

__global__ void k_gather_scal(double* const* scal, int idx, int n, double* out) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = scal[i][idx];
}

}  // namespace

struct jg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    std::vector<jg_solver*> sys;     // member views, in the caller's order
    std::vector<jg_solver*> groups;  // one solver per distinct hierarchy, owning the arrays
    std::vector<std::vector<int>> members;  // per group: its members in group order
    std::vector<hipEvent_t> join;    // per group
    Ctl** dctl = nullptr;       // per member: ctl (FP64 solve), then ctl + 1 (mixed inner solve)
    double** dscal = nullptr;   // per member: scal
    int* dflags = nullptr;      // gathered (flags, iteration) pairs
    double* dvals = nullptr;    // gathered scalars
    int* hflags = nullptr;      // pinned
    double* hvals = nullptr;    // pinned
    struct Graph {
        int precision;
        unsigned pre, post0, post, coarsest;
        double omega;
        int pairs;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    std::string err;

    ~jg_batch() {
        for (auto& g : graphs) {
            (void)hipGraphExecDestroy(g.exec);
            (void)hipGraphDestroy(g.graph);
        }
        for (auto* s : sys) jg_solver_destroy(s);  // views first: they use the groups' streams
        for (auto* g : groups) jg_solver_destroy(g);
        for (auto e : join) (void)hipEventDestroy(e);
        for (void* p : {(void*)dctl, (void*)dscal, (void*)dflags, (void*)dvals}) (void)hipFree(p);
        if (hflags) (void)hipHostFree(hflags);
        if (hvals) (void)hipHostFree(hvals);
        for (auto e : {fork, ev0, ev1})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define JG_HIPB(call)                                                    \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) {                                          \
            B->err = std::string(#call) + ": " + hipGetErrorString(e_);  \
            return JG_ERR_HIP;                                           \
        }                                                                \
    } while (0)

namespace {

// The batch stream waits for the work queued so far on every group stream.
int batch_join(jg_batch* B) {
    for (size_t g = 0; g < B->groups.size(); ++g) {
        JG_HIPB(hipEventRecord(B->join[g], B->groups[g]->stream));
        JG_HIPB(hipStreamWaitEvent(B->stream, B->join[g], 0));
    }
    return JG_OK;
}

// Every group stream waits for the work queued so far on the batch stream.
int batch_fork(jg_batch* B) {
    JG_HIPB(hipEventRecord(B->fork, B->stream));
    for (auto* g : B->groups) JG_HIPB(hipStreamWaitEvent(g->stream, B->fork, 0));
    return JG_OK;
}

// Gathers (flags, iteration) of control block `which` (0: FP64, 1: inner) of
// every member into hflags; synchronises.
int batch_poll(jg_batch* B, int which) {
    const int S = (int)B->sys.size();
    k_gather_ctl<<<1, 256, 0, B->stream>>>(B->dctl + which * S, S, B->dflags);
    JG_HIPB(hipGetLastError());
    JG_HIPB(hipMemcpyAsync(B->hflags, B->dflags, 2 * S * sizeof(int), hipMemcpyDeviceToHost,
                           B->stream));
    JG_HIPB(hipStreamSynchronize(B->stream));
    return JG_OK;
}

int batch_scal(jg_batch* B, int idx) {
    const int S = (int)B->sys.size();
    k_gather_scal<<<1, 256, 0, B->stream>>>(B->dscal, idx, S, B->dvals);
    JG_HIPB(hipGetLastError());
    JG_HIPB(hipMemcpyAsync(B->hvals, B->dvals, S * sizeof(double), hipMemcpyDeviceToHost, B->stream));
    JG_HIPB(hipStreamSynchronize(B->stream));
    return JG_OK;
}

// Precision-specific views of a solver's level-0 PCG buffers (a group's:
// its first member's pointers, which every launch extends to all members).
Pcg<double> pcg64(jg_solver* s, const MgCfg<double>* mg) {
    Level& l0 = s->lv[0];
    return Pcg<double>{l0.A64, s->b, s->x, s->r, s->z, s->dA, s->dB, s->q, l0.d, s->ctl, mg};
}
Pcg<float> pcg32(jg_solver* s, const MgCfg<float>* mg) {
    Level& l0 = s->lv[0];
    return Pcg<float>{l0.A32, s->rhs32, s->e32, s->r32, s->z32, s->d32A, s->d32B, s->q32, l0.d,
                      s->ctl + 1, mg};
}

// Captures kBatchIters PCG iterations of every group into one graph (or
// returns the cached one for the same precision and V-cycle parameters).
template <typename T>
int batch_graph(jg_batch* B, int precision, const MgCfg<T>& mg, hipGraphExec_t* out) {
    for (auto& g : B->graphs)
        if (g.precision == precision && g.pre == mg.pre && g.post0 == mg.post0 && g.post == mg.post &&
            g.coarsest == mg.coarsest && g.omega == (double)mg.omega && g.pairs == mg.pairs) {
            *out = g.exec;
            return JG_OK;
        }
    JG_HIPB(hipStreamBeginCapture(B->stream, hipStreamCaptureModeRelaxed));
    int st = JG_OK;
    if (hipEventRecord(B->fork, B->stream) != hipSuccess) st = JG_ERR_HIP;
    for (size_t gi = 0; gi < B->groups.size() && st == JG_OK; ++gi) {
        jg_solver* g = B->groups[gi];
        if (hipStreamWaitEvent(g->stream, B->fork, 0) != hipSuccess) {
            st = JG_ERR_HIP;
            break;
        }
        for (int k = 0; k < kBatchIters && st == JG_OK; ++k) {
            if constexpr (sizeof(T) == 8)
                st = pcg_iteration<double>(g, pcg64(g, &mg), (k & 1) != 0);
            else
                st = pcg_iteration<float>(g, pcg32(g, &mg), (k & 1) != 0);
        }
        if (st == JG_OK && (hipEventRecord(B->join[gi], g->stream) != hipSuccess ||
                            hipStreamWaitEvent(B->stream, B->join[gi], 0) != hipSuccess))
            st = JG_ERR_HIP;
    }
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &graph);
    if (st != JG_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        B->err = "graph capture failed";
        return st != JG_OK ? st : JG_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    JG_HIPB(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    B->graphs.push_back(
        {precision, mg.pre, mg.post0, mg.post, mg.coarsest, (double)mg.omega, mg.pairs, graph, exec});
    *out = exec;
    return JG_OK;
}

// Replays the iteration graph until every member's control block `which`
// is done (or the iteration budget is spent).
int batch_replay(jg_batch* B, hipGraphExec_t exec, int which, unsigned max_iter) {
    const size_t S = B->sys.size();
    for (unsigned launched = 0;; launched += kBatchIters) {
        JG_TRY(batch_poll(B, which));
        bool all = true;
        for (size_t i = 0; i < S; ++i) all = all && (B->hflags[2 * i] & 1);
        if (all || launched > max_iter + kBatchIters) break;
        JG_HIPB(hipGraphLaunch(exec, B->stream));
    }
    return JG_OK;
}

// Sweep tuning once per group, with launches that cover the whole group.
template <typename T>
int batch_tune(jg_batch* B) {
    for (auto* g : B->groups) {
        const int st = autotune_sweeps<T>(g);
        if (st != JG_OK) {
            B->err = g->err;
            return st;
        }
    }
    return JG_OK;
}

// The FP32 copies of every group whose members' matrices changed.
int batch_f32(jg_batch* B) {
    for (size_t gi = 0; gi < B->groups.size(); ++gi) {
        jg_solver* g = B->groups[gi];
        for (int m : B->members[gi])
            if (!B->sys[m]->f32_valid) g->f32_valid = false;
        const int st = ensure_f32(g);
        if (st != JG_OK) {
            B->err = g->err;
            return st;
        }
        for (int m : B->members[gi]) B->sys[m]->f32_valid = true;
    }
    return JG_OK;
}

}  // namespace

extern "C" {

int jg_batch_create(int device, size_t nsys, const size_t* nx, const size_t* ny, const size_t* nz,
                    const size_t* nlev, jg_batch** out) {
    if (!out || nsys == 0 || nsys > 4096 || !nx || !ny || !nz || !nlev) return JG_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
        return JG_ERR_NO_DEVICE;
    jg_batch* B = new jg_batch();
    B->device = device;
    auto fail = [&](int st) {
        delete B;
        return st;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&B->stream) != hipSuccess ||
        hipEventCreateWithFlags(&B->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&B->ev0) != hipSuccess || hipEventCreate(&B->ev1) != hipSuccess)
        return fail(JG_ERR_HIP);
    // group the systems by hierarchy (finest size and level count)
    std::vector<int> group_of(nsys, -1), index_in(nsys, 0);
    std::vector<size_t> rep;  // a member of each group
    for (size_t i = 0; i < nsys; ++i) {
        if (nlev[i] == 0 || nlev[i] > 16) return fail(JG_ERR_INVALID_ARG);
        for (size_t g = 0; g < rep.size() && group_of[i] < 0; ++g) {
            const size_t r = rep[g];
            if (nx[r] == nx[i] && ny[r] == ny[i] && nz[r] == nz[i] && nlev[r] == nlev[i])
                group_of[i] = (int)g;
        }
        if (group_of[i] < 0) {
            group_of[i] = (int)rep.size();
            rep.push_back(i);
            B->members.emplace_back();
        }
        index_in[i] = (int)B->members[group_of[i]].size();
        B->members[group_of[i]].push_back((int)i);
    }
    for (size_t g = 0; g < rep.size(); ++g) {
        const size_t r = rep[g];
        std::vector<size_t> lx(nlev[r]), ly(nlev[r]), lz(nlev[r]);
        for (size_t l = 0; l < nlev[r]; ++l) {
            lx[l] = nx[r] >> l;
            ly[l] = ny[r] >> l;
            lz[l] = nz[r] >> l;
        }
        jg_solver* s = nullptr;
        int st = jg_solver_create(device, &s);
        if (st != JG_OK) return fail(st);
        B->groups.push_back(s);
        // the checks of a single hierarchy, then the arrays of every member
        st = jg_solver_set_levels(s, nlev[r], lx.data(), ly.data(), lz.data());
        if (st == JG_OK && (s->twod || lz[0] == 1)) st = JG_ERR_UNSUPPORTED;  // batches are 3-D
        if (st == JG_OK)
            st = levels_alloc(s, nlev[r], lx.data(), ly.data(), lz.data(),
                              (int)B->members[g].size());
        if (st != JG_OK) return fail(st);
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(JG_ERR_HIP);
        B->join.push_back(e);
    }
    for (size_t i = 0; i < nsys; ++i) {
        jg_solver* v = nullptr;
        const int st = jg_solver_create(device, &v);
        if (st != JG_OK) return fail(st);
        make_view(B->groups[group_of[i]], index_in[i], v);
        B->sys.push_back(v);
    }
    std::vector<Ctl*> ctls(2 * nsys);
    std::vector<double*> scals(nsys);
    for (size_t i = 0; i < nsys; ++i) {
        ctls[i] = B->sys[i]->ctl;
        ctls[nsys + i] = B->sys[i]->ctl + 1;
        scals[i] = B->sys[i]->scal;
    }
    if (hipMalloc((void**)&B->dctl, 2 * nsys * sizeof(Ctl*)) != hipSuccess ||
        hipMalloc((void**)&B->dscal, nsys * sizeof(double*)) != hipSuccess ||
        hipMalloc((void**)&B->dflags, 2 * nsys * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&B->dvals, nsys * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&B->hflags, 2 * nsys * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&B->hvals, nsys * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipMemcpy(B->dctl, ctls.data(), 2 * nsys * sizeof(Ctl*), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->dscal, scals.data(), nsys * sizeof(double*), hipMemcpyHostToDevice) != hipSuccess)
        return fail(JG_ERR_OUT_OF_MEMORY);
    *out = B;
    return JG_OK;
}

void jg_batch_destroy(jg_batch* B) {
    if (!B) return;
    (void)hipSetDevice(B->device);
    (void)hipStreamSynchronize(B->stream);
    for (auto* g : B->groups) (void)hipStreamSynchronize(g->stream);
    delete B;
}

const char* jg_batch_last_error(const jg_batch* B) { return B ? B->err.c_str() : "null batch"; }

size_t jg_batch_size(const jg_batch* B) { return B ? B->sys.size() : 0; }

jg_solver* jg_batch_system(jg_batch* B, size_t i) {
    return (B && i < B->sys.size()) ? B->sys[i] : nullptr;
}

int jg_batch_mgpcg(jg_batch* B, unsigned max_iter, const jg_mg_params* p, const jg_gpu_options* o,
                   jg_solve_info* infos, double* total_ms) {
    if (!B || !p) return JG_ERR_INVALID_ARG;
    JG_HIPB(hipSetDevice(B->device));
    const size_t S = B->sys.size();
    jg_gpu_options opt;
    double omega = 0.0;
    {
        const int st = mg_options(B->sys[0], p, o, &opt, &omega);
        if (st != JG_OK) {
            B->err = B->sys[0]->err;
            return st;
        }
    }
    if (opt.smoother == JG_SMOOTH_JACOBI) {
        B->err = "batched solves use the red-black smoother";
        return JG_ERR_UNSUPPORTED;
    }
    const bool mixed = opt.precision == JG_MIXED;
    if (mixed) JG_TRY(batch_f32(B));
    JG_TRY(mixed ? batch_tune<float>(B) : batch_tune<double>(B));
    for (auto* g : B->groups) JG_TRY(sync(g));
    std::vector<unsigned> iters(S, 0), refinements(S, 0);
    std::vector<double> last(S, 0.0);
    std::vector<int> status(S, JG_OK);
    JG_TRY(batch_join(B));
    JG_HIPB(hipEventRecord(B->ev0, B->stream));
    JG_TRY(batch_fork(B));
    if (!mixed) {
        MgCfg<double> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter, omega};
        mg.pairs = opt.sweep_pairs;
        hipGraphExec_t exec = nullptr;
        JG_TRY(batch_graph<double>(B, JG_FP64, mg, &exec));
        JG_TRY(batch_fork(B));
        for (size_t gi = 0; gi < B->groups.size(); ++gi) {
            jg_solver* g = B->groups[gi];
            const Dims& d0 = g->lv[0].d;
            k_fill<double><<<grid_1d(nblocks_1d(d0.n()), g), NT, 0, g->stream>>>(g->x, 0.0, d0.n());
            for (int m : B->members[gi]) {
                Ctl h;
                init_ctl(&h, max_iter, p->max_tol, opt.stop_rule == JG_STOP_RELATIVE ? 1 : 0, opt.rel_tol);
                JG_TRY(upload_ctl(g, B->sys[m]->ctl, h));
            }
            JG_TRY(pcg_init<double>(g, pcg64(g, &mg)));
        }
        JG_TRY(batch_join(B));
        JG_TRY(batch_replay(B, exec, 0, max_iter));
        JG_TRY(batch_poll(B, 0));
        for (size_t i = 0; i < S; ++i) {
            iters[i] = (unsigned)B->hflags[2 * i + 1];
            if (B->hflags[2 * i] & 2) status[i] = JG_ERR_NONFINITE;
        }
    } else {
        MgCfg<float> mg{p->restrict_iter, p->final_iter, p->correct_iter, p->coarsest_iter, (float)omega};
        mg.pairs = opt.sweep_pairs;
        hipGraphExec_t exec = nullptr;
        JG_TRY(batch_graph<float>(B, JG_MIXED, mg, &exec));
        JG_TRY(batch_fork(B));
        // ||b|| of every member
        for (auto* s : B->sys) {
            const Dims& d0 = s->lv[0].d;
            const int nb1 = nblocks_1d(d0.n());
            k_fill<double><<<nb1, NT, 0, s->stream>>>(s->x, 0.0, d0.n());
            k_dot<double, double><<<nb1, NT, 0, s->stream>>>(s->b, s->b, d0.n(), s->slot(3), nullptr);
            k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(3), nb1, s->scal + 2);
        }
        JG_TRY(batch_join(B));
        JG_TRY(batch_scal(B, 2));
        std::vector<double> goal(S);
        for (size_t i = 0; i < S; ++i)
            goal[i] = opt.stop_rule == JG_STOP_RELATIVE ? opt.rel_tol * std::sqrt(B->hvals[i]) : p->max_tol;
        std::vector<char> done(S, 0);
        for (unsigned outer = 0;; ++outer) {
            JG_TRY(batch_fork(B));
            // FP64 residual r = b - A x of every unfinished member, into the FP32 right-hand side
            for (size_t i = 0; i < S; ++i) {
                if (done[i]) continue;
                jg_solver* s = B->sys[i];
                const int nr = residual_dot64<float>(s, s->lv[0], s->rhs32, s->slot(3));
                k_fin_sum<<<1, kFinThreads, 0, s->stream>>>(s->slot(3), nr, s->scal + 3);
            }
            JG_TRY(batch_join(B));
            JG_TRY(batch_scal(B, 3));
            bool any = false;
            for (size_t i = 0; i < S; ++i) {
                if (done[i]) continue;
                const double rr = B->hvals[i];
                last[i] = std::sqrt(rr);
                if (!std::isfinite(rr)) {
                    status[i] = JG_ERR_NONFINITE;
                    done[i] = 1;
                } else if (last[i] <= goal[i] || outer >= opt.max_refine || iters[i] >= max_iter) {
                    done[i] = 1;
                }
                any = any || !done[i];
            }
            if (!any) break;
            JG_TRY(batch_fork(B));
            for (size_t gi = 0; gi < B->groups.size(); ++gi) {
                jg_solver* g = B->groups[gi];
                // each member's inner control block: a finished member's
                // blocks exit at once
                for (int m : B->members[gi]) {
                    Ctl h;
                    if (done[m]) {
                        init_ctl(&h, 0, 0.0, 1, 0.0);
                        h.done = 1;
                    } else {
                        const double inner =
                            std::max(opt.inner_rel_tol, std::min(0.5, 0.5 * goal[m] / last[m]));
                        init_ctl(&h, max_iter - iters[m], 0.0, 1, inner);
                    }
                    JG_TRY(upload_ctl(g, B->sys[m]->ctl + 1, h));
                }
                const Dims& d0 = g->lv[0].d;
                k_fill<float><<<grid_1d(nblocks_1d(d0.n()), g), NT, 0, g->stream>>>(g->e32, 0.f, d0.n());
                JG_TRY(pcg_init<float>(g, pcg32(g, &mg)));
            }
            JG_TRY(batch_join(B));
            JG_TRY(batch_replay(B, exec, 1, max_iter));
            JG_TRY(batch_poll(B, 1));
            JG_TRY(batch_fork(B));
            for (size_t i = 0; i < S; ++i) {
                if (done[i]) continue;
                jg_solver* s = B->sys[i];
                iters[i] += (unsigned)B->hflags[2 * i + 1];
                if (B->hflags[2 * i] & 2) {  // the inner FP32 solve met a non-finite value
                    status[i] = JG_ERR_NONFINITE;
                    done[i] = 1;
                    continue;
                }
                const Dims& d0 = s->lv[0].d;
                k_add_correction<<<nblocks_1d(d0.n()), NT, 0, s->stream>>>(s->x, s->e32, d0.n());
                ++refinements[i];
            }
        }
    }
    JG_TRY(batch_join(B));
    JG_HIPB(hipEventRecord(B->ev1, B->stream));
    JG_HIPB(hipStreamSynchronize(B->stream));
    float ms = 0.f;
    JG_HIPB(hipEventElapsedTime(&ms, B->ev0, B->ev1));
    if (total_ms) *total_ms = ms;
    int worst = JG_OK;
    for (size_t i = 0; i < S; ++i) {
        jg_solver* s = B->sys[i];
        if (status[i] != JG_OK) worst = status[i];
        if (!infos) continue;
        jg_solve_info& info = infos[i];
        std::memset(&info, 0, sizeof(info));
        info.iterations = iters[i];
        info.refinements = refinements[i];
        info.solve_ms = ms;
        JG_TRY(true_rel_residual(s, &info.rel_residual));
        if (!mixed) {
            JG_HIPB(hipMemcpy(s->ctl_host, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
            last[i] = std::sqrt(std::fabs(s->ctl_host->sigma));
        }
        info.last_residual = last[i];
        if (status[i] != JG_OK)
            info.converged = 0;
        else if (opt.stop_rule == JG_STOP_RELATIVE)
            info.converged = info.rel_residual <= opt.rel_tol;
        else if (mixed)
            info.converged = last[i] <= p->max_tol;  // the outer loop's goal
        else
            info.converged = last[i] <= p->max_tol || iters[i] < max_iter;
    }
    if (worst == JG_ERR_NONFINITE) B->err = "non-finite value in a batched solve";
    return worst;
}

}  // extern "C"
