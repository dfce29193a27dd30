// jg_internal.h -- solver state shared by the translation units of
// libjetgpu (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jetgpu.h"
#include "jg_device.h"

#define JG_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            s->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return e_ == hipErrorOutOfMemory ? JG_ERR_OUT_OF_MEMORY : JG_ERR_HIP; \
        }                                                                        \
    } while (0)

#define JG_TRY(expr)            \
    do {                        \
        int st_ = (expr);       \
        if (st_ != JG_OK) return st_; \
    } while (0)

namespace jg {

struct Level {
    Dims d{0, 0, 0};
    Row<double>* A64 = nullptr;
    Row<float>* A32 = nullptr;
    // Preconditioner work vectors (levels > 0 use x/b; every level uses t).
    double *x64 = nullptr, *b64 = nullptr, *t64 = nullptr;
    float *x32 = nullptr, *b32 = nullptr, *t32 = nullptr;
    double* y64 = nullptr;  // ping-pong partner of x for the fused sweeps
    float* y32 = nullptr;
    int8_t* markers = nullptr;  // pressure-system markers of this level
    // Variational pressure data (allocated by the first jg_fr_build): fluid
    // SDF at cell centres and face weights, float as Jet stores them.
    float *fs = nullptr, *uw = nullptr, *vw = nullptr, *ww = nullptr;
    // Tuned z-chunk (0: untuned) and launch shape of the fused sweep, per
    // sweep kind (kSweepPlain, kSweepZero, kSweepCorrect).
    int zc64[3] = {0, 0, 0}, zc32[3] = {0, 0, 0};
    int sh64[3] = {0, 0, 0}, sh32[3] = {0, 0, 0};
    // Two sweeps per launch, per kind: kPairNone (two single sweeps),
    // kPairMarch (k_rb_sweep2, FP32) or kPairBrick (k_brick_sweep2), as the
    // tuner measured fastest; and the z-chunk of k_rb_sweep2.
    int pair64[3] = {0, 0, 0}, pair32[3] = {0, 0, 0};
    int zc2_32[3] = {0, 0, 0};
    // z-chunk of the zero-start two-sweep launch that also writes the
    // residual (its 5-cell halo changes the tile and block counts); 0: use
    // the zero-start kind's
    int zc2res_32 = 0;
};

enum { kPairNone = 0, kPairMarch = 1, kPairBrick = 2 };

// Sweep kinds tuned separately: they differ in registers and loads.
enum { kSweepPlain = 0, kSweepZero = 1, kSweepCorrect = 2, kNumSweepKinds = 3 };

template <typename T>
struct Prec;
template <>
struct Prec<double> {
    static Row<double>* A(Level& l) { return l.A64; }
    static double* x(Level& l) { return l.x64; }
    static double* b(Level& l) { return l.b64; }
    static double* t(Level& l) { return l.t64; }
    static double* y(Level& l) { return l.y64; }
    static int& zc(Level& l, int kind = kSweepPlain) { return l.zc64[kind]; }
    static int& sh(Level& l, int kind = kSweepPlain) { return l.sh64[kind]; }
    static int& pair(Level& l, int kind) { return l.pair64[kind]; }
    static int pair(const Level& l, int kind) { return l.pair64[kind]; }
    static int zc2(const Level&, int) { return 0; }  // no FP64 z-marching two-sweep kernel
    static int zc2res(const Level&) { return 0; }
};
template <>
struct Prec<float> {
    static Row<float>* A(Level& l) { return l.A32; }
    static float* x(Level& l) { return l.x32; }
    static float* b(Level& l) { return l.b32; }
    static float* t(Level& l) { return l.t32; }
    static float* y(Level& l) { return l.y32; }
    static int& zc(Level& l, int kind = kSweepPlain) { return l.zc32[kind]; }
    static int& sh(Level& l, int kind = kSweepPlain) { return l.sh32[kind]; }
    static int& pair(Level& l, int kind) { return l.pair32[kind]; }
    static int pair(const Level& l, int kind) { return l.pair32[kind]; }
    static int zc2(const Level& l, int kind) { return l.zc2_32[kind]; }
    static int zc2res(const Level& l) {
        return l.zc2res_32 > 0 ? l.zc2res_32 : l.zc2_32[kSweepZero];
    }
};

struct Geo {
    dim3 grid, block;
    int zc;
};

inline int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// 2.5-D launch geometry: bx*by = 256 columns per block, zc planes per block.
inline Geo geo_stencil(const Dims& d) {
    Geo g;
    const int bx = std::min(64, pow2ceil(std::max(1, d.nx)));
    const int by = NT / bx;
    const int gx = (d.nx + bx - 1) / bx;
    const int gy = (d.ny + by - 1) / by;
    const long cols = (long)gx * gy;
    // Aim for >= 2048 blocks; at most 8 planes per block; bounded partials.
    long zc = std::max<long>(1, std::min<long>(8, cols * d.nz / 2048));
    // (bounded by nz: a plane of more than kMaxPartials tiles is rejected
    // by jg_solver_set_levels, so the bound is always met there)
    while ((long)gx * gy * ((d.nz + zc - 1) / zc) > kMaxPartials && zc < d.nz) ++zc;
    g.zc = (int)zc;
    g.block = dim3(bx, by, 1);
    g.grid = dim3(gx, gy, (d.nz + g.zc - 1) / g.zc);
    return g;
}

// One thread per cell of a plane, one block row of planes per z.
inline Geo geo_planes(const Dims& d) {
    Geo g;
    const int bx = std::min(64, pow2ceil(std::max(1, d.nx)));
    const int by = NT / bx;
    g.zc = 1;
    g.block = dim3(bx, by, 1);
    g.grid = dim3((d.nx + bx - 1) / bx, (d.ny + by - 1) / by, d.nz);
    return g;
}

inline int nblocks_1d(idx_t n) {
    idx_t b = (n + NT - 1) / NT;
    return (int)std::max<idx_t>(1, std::min<idx_t>(b, 4096));
}

inline int nblocks(const Geo& g) { return (int)(g.grid.x * g.grid.y * g.grid.z); }

}  // namespace jg

struct jg_solver {
    int device = 0;
    // Systems held back to back (a batch group of one hierarchy): every
    // array below holds nsys of them (jg_device.h, kSys*), and the V-cycle
    // and PCG launches cover all of them.  A view is one member of a group:
    // its pointers address that member's part of the group's arrays, it owns
    // no array and shares the group's stream.
    int nsys = 1;
    bool view = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evp[2] = {nullptr, nullptr};  // stop-flag polls (two in flight)
    std::string err;
    std::vector<jg::Level> lv;
    // Levels are one plane thick and halve in x and y only (Jet's 2-D
    // hierarchy, FdmMgLinearSystem2): the V-cycle takes the 2-D transfers.
    bool twod = false;
    // Finest-level Krylov vectors.
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *dA = nullptr, *dB = nullptr,
           *q = nullptr;
    float *rhs32 = nullptr, *r32 = nullptr, *z32 = nullptr, *d32A = nullptr, *d32B = nullptr,
          *q32 = nullptr, *e32 = nullptr;
    // Velocity boundary conditions: projected faces, two validity masks.
    double* bc_tmp = nullptr;
    int8_t* bc_valid = nullptr;
    // Compressed (CSR) system of jg_csr_cg_solve_host and its CG vectors.
    struct Csr {
        size_t rows = 0, nnz = 0, cap_rows = 0, cap_nnz = 0;
        double* val = nullptr;
        int *rp = nullptr, *ci = nullptr;
        double *b = nullptr, *x = nullptr, *r = nullptr, *dA = nullptr, *dB = nullptr, *q = nullptr;
        // built on the device from the finest level (jg_sp_build_compressed):
        // per cell its row (ro) and first entry (eo), the counts, scan scratch
        bool from_grid = false;
        // ELL copy for the CG kernels (rows of at most kEllMaxW entries)
        double* ell_val = nullptr;
        int* ell_ci = nullptr;
        int8_t* ell_len = nullptr;
        size_t ell_cap = 0, ell_rows_cap = 0;
        int ell_w = 0;  // 0: no ELL copy (a longer row), the CSR kernels run
        int max_row = 0;         // longest row (an upper bound for a device build)
        bool ell_valid = false;  // the ELL copy matches the CSR arrays
        int *ro = nullptr, *eo = nullptr, *cr = nullptr, *ce = nullptr;
        void* scan_tmp = nullptr;
        size_t scan_bytes = 0;
    } csr;
    // Reduction slots and device scalars.
    double* part = nullptr;  // kSlots * kMaxPartials
    double* scal = nullptr;  // small result area
    jg::Ctl* ctl = nullptr;      // [0]: FP64 / outer, [1]: FP32 inner
    jg::Ctl* ctl_host = nullptr; // pinned mirror, two slots
    bool f32_valid = false;
    std::vector<void*> allocs;

    static constexpr int kSlots = jg::kSysSlots;
    double* slot(int i) { return part + (size_t)i * jg::kMaxPartials; }

    ~jg_solver() {
        release();
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : evp)
            if (e) (void)hipEventDestroy(e);
        if (stream && !view) (void)hipStreamDestroy(stream);
    }
    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        lv.clear();
        b = x = r = z = dA = dB = q = nullptr;
        rhs32 = r32 = z32 = d32A = d32B = q32 = e32 = nullptr;
        bc_tmp = nullptr;
        bc_valid = nullptr;
        csr = Csr{};
        part = scal = nullptr;
        ctl = nullptr;
        f32_valid = false;
    }
};

namespace jg {

template <typename T>
inline int dalloc(jg_solver* s, T** p, size_t count) {
    void* q = nullptr;
    JG_HIP(hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T)));
    s->allocs.push_back(q);
    *p = (T*)q;
    return JG_OK;
}

inline int check_launch(jg_solver* s, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        s->err = std::string(what) + ": " + hipGetErrorString(e);
        return JG_ERR_HIP;
    }
    return JG_OK;
}

inline int sync(jg_solver* s) {
    JG_HIP(hipStreamSynchronize(s->stream));
    return JG_OK;
}

}  // namespace jg

// Room on the device for a compressed system of `rows` rows and `nnz`
// entries with its CG vectors (jg_solver.hip); sets csr.rows / csr.nnz.
int csr_reserve(jg_solver* s, size_t rows, size_t nnz);
