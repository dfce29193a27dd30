// jg_pressure.h -- GPU assembly of Jet's single-phase pressure system and
// the pressure-gradient projection, plus the synthetic-scene kernels.
//
// Face arrays: u is (nx+1)*ny*nz, v is nx*(ny+1)*nz, w is nx*ny*(nz+1),
// x fastest (FaceCenteredGrid3, src/jet/face_centered_grid3.cpp:429-444).
// Markers: 0 fluid, 1 air, 2 boundary (src/jet/grid_single_phase_pressure_solver3.cpp:17-19).
#pragma once

#include "jg_device.h"
#include "jg_synth.h"

namespace jg {

enum : int8_t { kFluid = 0, kAir = 1, kBoundary = 2 };

struct H3 {
    double x, y, z;
};

// buildMarkers, finest level (src/jet/grid_single_phase_pressure_solver3.cpp:258-269).
__global__ __launch_bounds__(256) void k_sp_markers(const double* __restrict__ bsdf,
                                                    const double* __restrict__ fsdf,
                                                    int8_t* __restrict__ m, idx_t n) {
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256)
        m[c] = bsdf[c] < 0 ? kBoundary : (fsdf[c] < 0 ? kFluid : kAir);
}

// Coarse markers: majority vote over the clamped 4^3 neighbourhood, ties to
// the higher code (argmax3, include/jet/detail/math_utils-inl.h:107-113;
// src/jet/grid_single_phase_pressure_solver3.cpp:271-327).
__global__ __launch_bounds__(256) void k_sp_coarsen(const int8_t* __restrict__ fine,
                                                    int8_t* __restrict__ coarse, Dims cd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= cd.nx || j >= cd.ny) return;
    const int fnx = 2 * cd.nx, fny = 2 * cd.ny;
    const int ii[4] = {i > 0 ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1, i + 1 < cd.nx ? 2 * i + 2 : 2 * i + 1};
    const int jj[4] = {j > 0 ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1, j + 1 < cd.ny ? 2 * j + 2 : 2 * j + 1};
    const int kk[4] = {k > 0 ? 2 * k - 1 : 2 * k, 2 * k, 2 * k + 1, k + 1 < cd.nz ? 2 * k + 2 : 2 * k + 1};
    int cnt[3] = {0, 0, 0};
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const int8_t f = fine[ii[x] + (idx_t)fnx * (jj[y] + (idx_t)fny * kk[z])];
                ++cnt[f == kBoundary ? 2 : (f == kFluid ? 0 : 1)];
            }
    int8_t r;
    if (cnt[0] > cnt[1])
        r = cnt[0] > cnt[2] ? 0 : 2;
    else
        r = cnt[1] > cnt[2] ? 1 : 2;
    coarse[i + (idx_t)cd.nx * (j + (idx_t)cd.ny * k)] = r;
}

// Coarse markers of a 2-D hierarchy: the vote over the clamped 4^2
// neighbourhood (src/jet/grid_single_phase_pressure_solver2.cpp:238-283).
__global__ __launch_bounds__(256) void k_sp_coarsen2(const int8_t* __restrict__ fine,
                                                     int8_t* __restrict__ coarse, Dims cd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= cd.nx || j >= cd.ny) return;
    const int fnx = 2 * cd.nx;
    const int ii[4] = {i > 0 ? 2 * i - 1 : 2 * i, 2 * i, 2 * i + 1, i + 1 < cd.nx ? 2 * i + 2 : 2 * i + 1};
    const int jj[4] = {j > 0 ? 2 * j - 1 : 2 * j, 2 * j, 2 * j + 1, j + 1 < cd.ny ? 2 * j + 2 : 2 * j + 1};
    int cnt[3] = {0, 0, 0};
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
            const int8_t f = fine[ii[x] + (idx_t)fnx * jj[y]];
            ++cnt[f == kBoundary ? 2 : (f == kFluid ? 0 : 1)];
        }
    int8_t r;
    if (cnt[0] > cnt[1])
        r = cnt[0] > cnt[2] ? 0 : 2;
    else
        r = cnt[1] > cnt[2] ? 1 : 2;
    coarse[i + (idx_t)cd.nx * j] = r;
}

// buildSingleSystem (src/jet/grid_single_phase_pressure_solver3.cpp:25-79).
// b (divergence, FaceCenteredGrid3::divergenceAtCellCenter,
// src/jet/face_centered_grid3.cpp:137-152) is written only when u is given.
// w == nullptr: a one-plane 2-D system (src/jet/grid_single_phase_pressure_
// solver2.cpp:25-67, FaceCenteredGrid2::divergenceAtCellCenter, src/jet/
// face_centered_grid2.cpp:107-118): no w term, and nz == 1 leaves no z
// neighbour.
__global__ __launch_bounds__(256) void k_sp_build(const int8_t* __restrict__ m,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const double* __restrict__ w, Dims d, H3 h,
                                                  Row<double>* __restrict__ A,
                                                  double* __restrict__ b) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const idx_t sy = d.nx, sz = (idx_t)d.nx * d.ny;
    const idx_t c = i + sy * j + sz * k;
    const double ihx = 1.0 / h.x, ihy = 1.0 / h.y, ihz = 1.0 / h.z;
    const double ihx2 = ihx * ihx, ihy2 = ihy * ihy, ihz2 = ihz * ihz;
    Row<double> row{0.0, 0.0, 0.0, 0.0};
    double bc = 0.0;
    if (m[c] == kFluid) {
        if (u) {
            const double lu = u[i + (idx_t)(d.nx + 1) * (j + (idx_t)d.ny * k)];
            const double ru = u[i + 1 + (idx_t)(d.nx + 1) * (j + (idx_t)d.ny * k)];
            const double bv = v[i + (idx_t)d.nx * (j + (idx_t)(d.ny + 1) * k)];
            const double tv = v[i + (idx_t)d.nx * (j + 1 + (idx_t)(d.ny + 1) * k)];
            bc = (ru - lu) / h.x + (tv - bv) / h.y;
            if (w) {
                const double bw = w[c];
                const double fw = w[c + sz];
                bc = bc + (fw - bw) / h.z;
            }
        }
        if (i + 1 < d.nx && m[c + 1] != kBoundary) {
            row.c += ihx2;
            if (m[c + 1] == kFluid) row.r -= ihx2;
        }
        if (i > 0 && m[c - 1] != kBoundary) row.c += ihx2;
        if (j + 1 < d.ny && m[c + sy] != kBoundary) {
            row.c += ihy2;
            if (m[c + sy] == kFluid) row.u -= ihy2;
        }
        if (j > 0 && m[c - sy] != kBoundary) row.c += ihy2;
        if (k + 1 < d.nz && m[c + sz] != kBoundary) {
            row.c += ihz2;
            if (m[c + sz] == kFluid) row.f -= ihz2;
        }
        if (k > 0 && m[c - sz] != kBoundary) row.c += ihz2;
    } else {
        row.c = 1.0;
    }
    A[c] = row;
    if (b) b[c] = bc;
}

// applyPressureGradient (src/jet/grid_single_phase_pressure_solver3.cpp:401-431):
// the +x/+y/+z face of every fluid cell whose neighbour there is not
// Boundary gets u + (p(nbr) - p)/h.  Each face has exactly one writer, so
// the output may alias the input.
__global__ __launch_bounds__(256) void k_sp_gradient(const int8_t* __restrict__ m,
                                                     const double* __restrict__ p,
                                                     const double* u, const double* v,
                                                     const double* w, double* uo, double* vo,
                                                     double* wo, Dims d, H3 h) {
    // u + (1/h) (p1 - p0) with the product rounded before the add (Jet's
    // expression; a fused multiply-add differs wherever 1/h is not a power of 2)
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const idx_t sy = d.nx, sz = (idx_t)d.nx * d.ny;
    const idx_t c = i + sy * j + sz * k;
    if (m[c] != kFluid) return;
    const double ihx = 1.0 / h.x, ihy = 1.0 / h.y, ihz = 1.0 / h.z;
    if (i + 1 < d.nx && m[c + 1] != kBoundary) {
        const idx_t f = i + 1 + (idx_t)(d.nx + 1) * (j + (idx_t)d.ny * k);
        uo[f] = u[f] + ihx * (p[c + 1] - p[c]);
    }
    if (j + 1 < d.ny && m[c + sy] != kBoundary) {
        const idx_t f = i + (idx_t)d.nx * (j + 1 + (idx_t)(d.ny + 1) * k);
        vo[f] = v[f] + ihy * (p[c + sy] - p[c]);
    }
    if (k + 1 < d.nz && m[c + sz] != kBoundary) {
        const idx_t f = c + sz;
        wo[f] = w[f] + ihz * (p[c + sz] - p[c]);
    }
}

// ---------------------------------------------------------- synthetic

__global__ __launch_bounds__(256) void k_synth_faces(synth::Scene s, int axis, double* __restrict__ out) {
    const int n = s.n;
    const int sx = n + (axis == 0), sy = n + (axis == 1), sz = n + (axis == 2);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= sx || j >= sy || k >= sz) return;
    out[i + (idx_t)sx * (j + (idx_t)sy * k)] = synth::face_velocity(s, axis, i, j, k);
}

__global__ __launch_bounds__(256) void k_synth_cells(synth::Scene s, double* __restrict__ bsdf,
                                                     double* __restrict__ fsdf) {
    const int n = s.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= n || j >= n) return;
    const double x = synth::cc(s, i), y = synth::cc(s, j), z = synth::cc(s, k);
    const idx_t c = i + (idx_t)n * (j + (idx_t)n * k);
    bsdf[c] = synth::boundary_sdf(s, x, y, z);
    fsdf[c] = synth::fluid_sdf(s, x, y, z);
}

__global__ __launch_bounds__(256) void k_synth_vertices(synth::Scene s, double* __restrict__ vs) {
    const int n1 = s.n + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= n1 || j >= n1) return;
    vs[i + (idx_t)n1 * (j + (idx_t)n1 * k)] =
        synth::boundary_sdf(s, s.h * (double)i, s.h * (double)j, s.h * (double)k);
}

// Non-solid cells per y row (integer atomics: deterministic).
__global__ __launch_bounds__(256) void k_synth_row_open(synth::Scene s, unsigned* __restrict__ rows) {
    const int n = s.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= n || j >= n) return;
    if (!synth::solid_cell(s, i, j, k)) atomicAdd(&rows[j], 1u);
}

// Compressed buildSingleSystem (src/jet/grid_single_phase_pressure_solver3.cpp:
// 81-171) in three passes over the finest level's markers (0 fluid, 2
// boundary): count, exclusive scans of the counts (row number and first
// entry of every fluid cell), fill.

// Pass 1: rows[c] = 1 for a fluid cell, nnz[c] = 1 + its fluid neighbours.
__global__ __launch_bounds__(256) void k_csr_count(const int8_t* __restrict__ m, Dims d,
                                                   int* __restrict__ rows, int* __restrict__ nnz) {
    const idx_t n = d.n(), sy = d.nx, sz = (idx_t)d.nx * d.ny;
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        if (m[c] != 0) {
            rows[c] = 0;
            nnz[c] = 0;
            continue;
        }
        const int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / sz);
        int e = 1;
        e += (i + 1 < d.nx && m[c + 1] == 0) + (i > 0 && m[c - 1] == 0);
        e += (j + 1 < d.ny && m[c + sy] == 0) + (j > 0 && m[c - sy] == 0);
        e += (k + 1 < d.nz && m[c + sz] == 0) + (k > 0 && m[c - sz] == 0);
        rows[c] = 1;
        nnz[c] = e;
    }
}

// Pass 3: the row of fluid cell c (row ro[c], entries from eo[c]) in
// ascending column order -- for x-fastest numbering -z, -y, -x, centre, +x,
// +y, +z, the order MatrixCsr::addRow sorts Jet's row into
// (include/jet/detail/matrix_csr-inl.h:253-275) -- with the centre summed in
// Jet's order +x, -x, +y, -y, +z, -z over non-boundary neighbours, and
// b = FaceCenteredGrid3::divergenceAtCellCenter (as k_sp_build).
__global__ __launch_bounds__(256) void k_csr_fill(const int8_t* __restrict__ m,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const double* __restrict__ w, Dims d, double hx,
                                                  double hy, double hz, const int* __restrict__ ro,
                                                  const int* __restrict__ eo, int* __restrict__ rp,
                                                  int* __restrict__ ci, double* __restrict__ val,
                                                  double* __restrict__ b) {
#pragma clang fp contract(off)
    const idx_t n = d.n(), sy = d.nx, sz = (idx_t)d.nx * d.ny;
    const double ihx = 1.0 / hx, ihy = 1.0 / hy, ihz = 1.0 / hz;
    const double ih2[3] = {ihx * ihx, ihy * ihy, ihz * ihz};
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        if (m[c] != 0) continue;
        const int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / sz);
        // neighbours in Jet's order +x, -x, +y, -y, +z, -z
        const bool in[6] = {i + 1 < d.nx, i > 0, j + 1 < d.ny, j > 0, k + 1 < d.nz, k > 0};
        const idx_t off[6] = {1, -1, sy, -sy, sz, -sz};
        double centre = 0.0;
        bool fl[6];
        for (int q = 0; q < 6; ++q) {
            const int8_t mq = in[q] ? m[c + off[q]] : int8_t(2);
            if (mq != 2) centre += ih2[q / 2];
            fl[q] = mq == 0;
        }
        const int row = ro[c];
        int e = eo[c];
        rp[row] = e;
        const int order[6] = {5, 3, 1, 0, 2, 4};
        for (int o = 0; o < 6; ++o) {
            if (o == 3) {
                val[e] = centre;
                ci[e] = row;
                ++e;
            }
            const int q = order[o];
            if (!fl[q]) continue;
            val[e] = -ih2[q / 2];
            ci[e] = ro[c + off[q]];
            ++e;
        }
        const double lu = u[i + (idx_t)(d.nx + 1) * (j + (idx_t)d.ny * k)];
        const double ru = u[i + 1 + (idx_t)(d.nx + 1) * (j + (idx_t)d.ny * k)];
        const double bv = v[i + (idx_t)d.nx * (j + (idx_t)(d.ny + 1) * k)];
        const double tv = v[i + (idx_t)d.nx * (j + 1 + (idx_t)(d.ny + 1) * k)];
        double bc = (ru - lu) / hx + (tv - bv) / hy;
        if (w) {  // w == nullptr: one-plane 2-D system, no w term
            const double bw = w[c], fw = w[c + sz];
            bc = bc + (fw - bw) / hz;
        }
        b[row] = bc;
    }
}

// decompressSolution (src/jet/grid_single_phase_pressure_solver3.cpp:330-341):
// the solution of fluid row ro[c] at its cell, 0 at every other cell.
__global__ __launch_bounds__(256) void k_csr_scatter(const int8_t* __restrict__ m,
                                                     const int* __restrict__ ro,
                                                     const double* __restrict__ xc,
                                                     double* __restrict__ x, idx_t n) {
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256)
        x[c] = m[c] == 0 ? xc[ro[c]] : 0.0;
}

// GridBackwardEulerDiffusionSolver3::buildMarkers / buildMatrix /
// buildVectors (src/jet/grid_backward_euler_diffusion_solver3.cpp:175-330) for
// one data grid: markers from the SDFs at the data points (boundary first,
// then fluid, else air), rows (I + dt mu L) with the neighbour rules of the
// Neumann / Dirichlet boundary types, b = f plus, for Dirichlet, c f of
// boundary neighbours.  c = dt mu / h^2 per axis.
__global__ __launch_bounds__(256) void k_diffusion_build(const double* __restrict__ f,
                                                         const double* __restrict__ bsdf,
                                                         const double* __restrict__ fsdf, Dims d,
                                                         double cx, double cy, double cz,
                                                         int dirichlet, Row<double>* __restrict__ A,
                                                         double* __restrict__ b) {
#pragma clang fp contract(off)
    const idx_t n = d.n(), oy = d.nx, oz = (idx_t)d.nx * d.ny;
    auto mk = [&](idx_t q) -> int8_t { return bsdf[q] < 0 ? kBoundary : (fsdf[q] < 0 ? kFluid : kAir); };
    for (idx_t q = blockIdx.x * (idx_t)256 + threadIdx.x; q < n; q += (idx_t)gridDim.x * 256) {
        const int i = (int)(q % d.nx), j = (int)((q / d.nx) % d.ny), k = (int)(q / oz);
        Row<double> row{1.0, 0.0, 0.0, 0.0};
        double bv = f[q];
        if (mk(q) == kFluid) {
            auto counts = [&](idx_t qq) {
                const int8_t m = mk(qq);
                return (dirichlet && m != kAir) || m == kFluid;
            };
            if (i + 1 < d.nx) {
                if (counts(q + 1)) row.c += cx;
                if (mk(q + 1) == kFluid) row.r -= cx;
            }
            if (i > 0 && counts(q - 1)) row.c += cx;
            if (j + 1 < d.ny) {
                if (counts(q + oy)) row.c += cy;
                if (mk(q + oy) == kFluid) row.u -= cy;
            }
            if (j > 0 && counts(q - oy)) row.c += cy;
            if (k + 1 < d.nz) {
                if (counts(q + oz)) row.c += cz;
                if (mk(q + oz) == kFluid) row.f -= cz;
            }
            if (k > 0 && counts(q - oz)) row.c += cz;
            if (dirichlet) {
                if (i + 1 < d.nx && mk(q + 1) == kBoundary) bv += cx * f[q + 1];
                if (i > 0 && mk(q - 1) == kBoundary) bv += cx * f[q - 1];
                if (j + 1 < d.ny && mk(q + oy) == kBoundary) bv += cy * f[q + oy];
                if (j > 0 && mk(q - oy) == kBoundary) bv += cy * f[q - oy];
                if (k + 1 < d.nz && mk(q + oz) == kBoundary) bv += cz * f[q + oz];
                if (k > 0 && mk(q - oz) == kBoundary) bv += cz * f[q - oz];
            }
        }
        A[q] = row;
        b[q] = bv;
    }
}

}  // namespace jg
