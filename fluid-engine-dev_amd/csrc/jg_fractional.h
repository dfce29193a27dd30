// jg_fractional.h -- GPU assembly of Jet's variational (fractional) pressure
// system and its pressure-gradient projection.
//
// Face weights are float (Jet stores them as Array3<float>), the fluid SDF is
// float, the system is FP64.  Every expression is Jet's, in Jet's operand
// order, with contraction off, so results are bit-identical to the CPU
// restatement of GridFractionalSinglePhasePressureSolver3.
#pragma once

#include "jg_device.h"
#include "jg_pressure.h"

namespace jg {

// fractionInsideSdf (include/jet/detail/level_set_utils-inl.h:51-62).
__device__ __forceinline__ double fr_inside_sdf(double phi0, double phi1) {
#pragma clang fp contract(off)
    if (phi0 < 0 && phi1 < 0) return 1;
    if (phi0 < 0 && !(phi1 < 0)) return phi0 / (phi0 - phi1);
    if (!(phi0 < 0) && phi1 < 0) return phi1 / (phi1 - phi0);
    return 0;
}

// One step of the corner rotation list[0..3] <- list[1], list[2], list[3], list[0].
__device__ __forceinline__ void fr_rotate(double& a, double& b, double& c, double& d) {
    const double t = a;
    a = b;
    b = c;
    c = d;
    d = t;
}

// fractionInside: area fraction of a square cell with corner values bl, br,
// tl, tr that lies inside (phi < 0), Batty's corner-case formulas
// (include/jet/detail/level_set_utils-inl.h:71-147).  Corners are walked in
// the cyclic order bl, br, tr, tl; every rotation search ends within four
// steps (at least one corner has the required sign in each branch).
__device__ inline double fr_inside(double bl, double br, double tl, double tr) {
#pragma clang fp contract(off)
    const int inside = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
    double a = bl, b = br, c = tr, d = tl;
    if (inside == 4) return 1;
    if (inside == 3) {
        for (int s = 0; s < 4 && a < 0; ++s) fr_rotate(a, b, c, d);
        const double side0 = 1 - fr_inside_sdf(a, d);
        const double side1 = 1 - fr_inside_sdf(a, b);
        return 1 - 0.5 * side0 * side1;
    }
    if (inside == 2) {
        for (int s = 0; s < 4 && (a >= 0 || !(b < 0 || c < 0)); ++s) fr_rotate(a, b, c, d);
        if (b < 0) {
            const double left = fr_inside_sdf(a, d);
            const double right = fr_inside_sdf(b, c);
            return 0.5 * (left + right);
        }
        const double middle = 0.25 * (a + b + c + d);
        if (middle < 0) {
            double area = 0;
            const double side1 = 1 - fr_inside_sdf(a, d);
            const double side3 = 1 - fr_inside_sdf(c, d);
            area += 0.5 * side1 * side3;
            const double side2 = 1 - fr_inside_sdf(c, b);
            const double side0 = 1 - fr_inside_sdf(a, b);
            area += 0.5 * side0 * side2;
            return 1 - area;
        }
        double area = 0;
        const double side0 = fr_inside_sdf(a, b);
        const double side1 = fr_inside_sdf(a, d);
        area += 0.5 * side0 * side1;
        const double side2 = fr_inside_sdf(c, b);
        const double side3 = fr_inside_sdf(c, d);
        area += 0.5 * side2 * side3;
        return area;
    }
    if (inside == 1) {
        for (int s = 0; s < 4 && a >= 0; ++s) fr_rotate(a, b, c, d);
        const double side0 = fr_inside_sdf(a, d);
        const double side1 = fr_inside_sdf(a, b);
        return 0.5 * side0 * side1;
    }
    return 0;
}

// Face weight: 1 - fractionInside of the collider, clamped to [0, 1],
// floored to kMinWeight = 0.01 when positive, stored as float
// (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:553-617).
__device__ __forceinline__ float fr_weight(double p0, double p1, double p2, double p3) {
#pragma clang fp contract(off)
    double w = 1.0 - fr_inside(p0, p1, p2, p3);
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    if (w < 0.01 && w > 0.0) w = 0.01;
    return (float)w;
}

// Face weights of one axis from the boundary SDF at the grid vertices
// ((nx+1)*(ny+1)*(nz+1), vertex (i,j,k) at origin + h*(i,j,k)); the four
// corners of a face are vertices (buildWeights samples them, :553-617).
template <int AXIS>
__global__ __launch_bounds__(256) void k_fr_weights(const double* __restrict__ vs, Dims d,
                                                    float* __restrict__ out) {
    const int fx = d.nx + (AXIS == 0), fy = d.ny + (AXIS == 1), fz = d.nz + (AXIS == 2);
    const idx_t n = (idx_t)fx * fy * fz;
    const idx_t vx = d.nx + 1, vxy = vx * (d.ny + 1);
    for (idx_t f = blockIdx.x * (idx_t)256 + threadIdx.x; f < n; f += (idx_t)gridDim.x * 256) {
        const int i = (int)(f % fx), j = (int)((f / fx) % fy), k = (int)(f / ((idx_t)fx * fy));
        auto V = [&](int a, int b, int c) { return vs[a + vx * b + vxy * c]; };
        float w;
        if (AXIS == 0)
            w = fr_weight(V(i, j, k), V(i, j + 1, k), V(i, j, k + 1), V(i, j + 1, k + 1));
        else if (AXIS == 1)
            w = fr_weight(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k), V(i + 1, j, k + 1));
        else
            w = fr_weight(V(i, j, k), V(i, j + 1, k), V(i + 1, j, k), V(i + 1, j + 1, k));
        out[f] = w;
    }
}

// Level-0 fluid SDF as Jet stores it (float).
__global__ __launch_bounds__(256) void k_fr_sdf_to_float(const double* __restrict__ s,
                                                         float* __restrict__ o, idx_t n) {
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256)
        o[c] = (float)s[c];
}

// Float restriction of a cell or face array (the anonymous restrict() of
// src/jet/grid_fractional_single_phase_pressure_solver3.cpp:29-115): per
// axis the centred kernel {1/8, 3/8, 3/8, 1/8} over fine 2c-1 .. 2c+2
// (clamped), or injection {0, 1, 0} over 2c-1 .. 2c+1 along an axis whose
// fine size is not twice the coarse size (the staggered axis of a face
// array).  Jet's z, y, x accumulation order, float arithmetic.
__global__ __launch_bounds__(256) void k_fr_restrict(const float* __restrict__ fine, Dims fd,
                                                     float* __restrict__ coarse, Dims cd) {
#pragma clang fp contract(off)
    const idx_t n = cd.n();
    const int stx = fd.nx != 2 * cd.nx, sty = fd.ny != 2 * cd.ny, stz = fd.nz != 2 * cd.nz;
    for (idx_t e = blockIdx.x * (idx_t)256 + threadIdx.x; e < n; e += (idx_t)gridDim.x * 256) {
        const int c3[3] = {(int)(e % cd.nx), (int)((e / cd.nx) % cd.ny), (int)(e / ((idx_t)cd.nx * cd.ny))};
        const int cn[3] = {cd.nx, cd.ny, cd.nz};
        const int st[3] = {stx, sty, stz};
        int idx[3][4];
        float kern[3][4];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int c = c3[a];
            if (st[a]) {
                idx[a][0] = c > 0 ? 2 * c - 1 : 2 * c;
                idx[a][1] = 2 * c;
                idx[a][2] = c + 1 < cn[a] ? 2 * c + 1 : 2 * c;
                idx[a][3] = 2 * c;
                kern[a][0] = 0.f;
                kern[a][1] = 1.f;
                kern[a][2] = 0.f;
                kern[a][3] = 0.f;
            } else {
                idx[a][0] = c > 0 ? 2 * c - 1 : 2 * c;
                idx[a][1] = 2 * c;
                idx[a][2] = 2 * c + 1;
                idx[a][3] = c + 1 < cn[a] ? 2 * c + 2 : 2 * c + 1;
                kern[a][0] = 0.125f;
                kern[a][1] = 0.375f;
                kern[a][2] = 0.375f;
                kern[a][3] = 0.125f;
            }
        }
        const int kx = stx ? 3 : 4, ky = sty ? 3 : 4, kz = stz ? 3 : 4;
        float sum = 0.0f;
        for (int z = 0; z < kz; ++z)
            for (int y = 0; y < ky; ++y)
                for (int x = 0; x < kx; ++x) {
                    const float w = kern[0][x] * kern[1][y] * kern[2][z];
                    sum += w * fine[idx[0][x] + (idx_t)fd.nx * (idx[1][y] + (idx_t)fd.ny * idx[2][z])];
                }
        coarse[e] = sum;
    }
}

// Ghost-fluid fraction at a fluid/air face, floored at 0.01 (std::max).
__device__ __forceinline__ double fr_theta(double a, double b) {
    const double t = fr_inside_sdf(a, b);
    return t < 0.01 ? 0.01 : t;
}

// buildSingleSystem, variational form
// (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:117-267): for a
// fluid cell (phi < 0) each face inside the domain adds w/h^2 to the centre
// (divided by the ghost-fluid theta toward air) and, toward a +side fluid
// cell, -w/h^2 to the coupling; b is the weighted divergence (domain faces
// unweighted) plus the collider term sum (1 - w) u_collider / h.  A centre
// below machine epsilon gives the row 1, 0.  b is written only when u is
// given; bu/bv/bw may be null (zero collider velocity).
__global__ __launch_bounds__(256) void k_fr_build(
    const float* __restrict__ fs, const float* __restrict__ uw, const float* __restrict__ vw,
    const float* __restrict__ ww, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ bu, const double* __restrict__ bv,
    const double* __restrict__ bw, Dims d, H3 h, Row<double>* __restrict__ A,
    double* __restrict__ b) {
#pragma clang fp contract(off)
    const double invH[3] = {1.0 / h.x, 1.0 / h.y, 1.0 / h.z};
    const double invHSqr[3] = {invH[0] * invH[0], invH[1] * invH[1], invH[2] * invH[2]};
    const idx_t n = d.n();
    const idx_t ux = d.nx + 1, vy = d.ny + 1, sxy = (idx_t)d.nx * d.ny;
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        const int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / sxy);
        Row<double> row{0.0, 0.0, 0.0, 0.0};
        double bc = 0.0;
        const double cphi = fs[c];
        if (cphi < 0) {
            const idx_t uR = (i + 1) + ux * (j + (idx_t)d.ny * k), uL = uR - 1;
            const idx_t vU = i + (idx_t)d.nx * ((j + 1) + vy * k), vD = vU - d.nx;
            const idx_t wF = i + (idx_t)d.nx * (j + (idx_t)d.ny * (k + 1)), wB = wF - sxy;
            const bool rhs = u != nullptr;
            double term;
            if (i + 1 < d.nx) {
                term = uw[uR] * invHSqr[0];
                const double phi = fs[c + 1];
                if (phi < 0) {
                    row.c += term;
                    row.r -= term;
                } else {
                    row.c += term / fr_theta(cphi, phi);
                }
                if (rhs) bc += uw[uR] * u[uR] * invH[0];
            } else if (rhs) {
                bc += u[uR] * invH[0];
            }
            if (i > 0) {
                term = uw[uL] * invHSqr[0];
                const double phi = fs[c - 1];
                if (phi < 0)
                    row.c += term;
                else
                    row.c += term / fr_theta(cphi, phi);
                if (rhs) bc -= uw[uL] * u[uL] * invH[0];
            } else if (rhs) {
                bc -= u[uL] * invH[0];
            }
            if (j + 1 < d.ny) {
                term = vw[vU] * invHSqr[1];
                const double phi = fs[c + d.nx];
                if (phi < 0) {
                    row.c += term;
                    row.u -= term;
                } else {
                    row.c += term / fr_theta(cphi, phi);
                }
                if (rhs) bc += vw[vU] * v[vU] * invH[1];
            } else if (rhs) {
                bc += v[vU] * invH[1];
            }
            if (j > 0) {
                term = vw[vD] * invHSqr[1];
                const double phi = fs[c - d.nx];
                if (phi < 0)
                    row.c += term;
                else
                    row.c += term / fr_theta(cphi, phi);
                if (rhs) bc -= vw[vD] * v[vD] * invH[1];
            } else if (rhs) {
                bc -= v[vD] * invH[1];
            }
            if (k + 1 < d.nz) {
                term = ww[wF] * invHSqr[2];
                const double phi = fs[c + sxy];
                if (phi < 0) {
                    row.c += term;
                    row.f -= term;
                } else {
                    row.c += term / fr_theta(cphi, phi);
                }
                if (rhs) bc += ww[wF] * w[wF] * invH[2];
            } else if (rhs) {
                bc += w[wF] * invH[2];
            }
            if (k > 0) {
                term = ww[wB] * invHSqr[2];
                const double phi = fs[c - sxy];
                if (phi < 0)
                    row.c += term;
                else
                    row.c += term / fr_theta(cphi, phi);
                if (rhs) bc -= ww[wB] * w[wB] * invH[2];
            } else if (rhs) {
                bc -= w[wB] * invH[2];
            }
            if (rhs) {
                const double buR = bu ? bu[uR] : 0.0, buL = bu ? bu[uL] : 0.0;
                const double bvU = bv ? bv[vU] : 0.0, bvD = bv ? bv[vD] : 0.0;
                const double bwF = bw ? bw[wF] : 0.0, bwB = bw ? bw[wB] : 0.0;
                const double contrib = (1.0 - uw[uR]) * buR * invH[0] - (1.0 - uw[uL]) * buL * invH[0] +
                                       (1.0 - vw[vU]) * bvU * invH[1] - (1.0 - vw[vD]) * bvD * invH[1] +
                                       (1.0 - ww[wF]) * bwF * invH[2] - (1.0 - ww[wB]) * bwB * invH[2];
                bc += contrib;
            }
            if (row.c < 2.220446049250313080847e-16) {  // DBL_EPSILON (include/jet/constants.h:74)
                row.c = 1.0;
                bc = 0.0;
            }
        } else {
            row.c = 1.0;
        }
        A[c] = row;
        if (b) b[c] = bc;
    }
}

// applyPressureGradient, variational form
// (src/jet/grid_fractional_single_phase_pressure_solver3.cpp:715-765): a +side
// face of cell c is updated when its weight is positive and either side is
// fluid, by (p(+) - p(c)) / (h theta).  The output may alias the input.
__global__ __launch_bounds__(256) void k_fr_gradient(
    const float* __restrict__ fs, const float* __restrict__ uw, const float* __restrict__ vw,
    const float* __restrict__ ww, const double* __restrict__ p, const double* u,
    const double* v, const double* w, double* uo, double* vo, double* wo, Dims d, H3 h) {
#pragma clang fp contract(off)
    const double invH[3] = {1.0 / h.x, 1.0 / h.y, 1.0 / h.z};
    const idx_t n = d.n();
    const idx_t ux = d.nx + 1, vy = d.ny + 1, sxy = (idx_t)d.nx * d.ny;
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        const int i = (int)(c % d.nx), j = (int)((c / d.nx) % d.ny), k = (int)(c / sxy);
        const double cphi = fs[c];
        if (i + 1 < d.nx) {
            const idx_t f = (i + 1) + ux * (j + (idx_t)d.ny * k);
            const double nphi = fs[c + 1];
            if (uw[f] > 0.0 && (cphi < 0 || nphi < 0))
                uo[f] = u[f] + invH[0] / fr_theta(cphi, nphi) * (p[c + 1] - p[c]);
        }
        if (j + 1 < d.ny) {
            const idx_t f = i + (idx_t)d.nx * ((j + 1) + vy * k);
            const double nphi = fs[c + d.nx];
            if (vw[f] > 0.0 && (cphi < 0 || nphi < 0))
                vo[f] = v[f] + invH[1] / fr_theta(cphi, nphi) * (p[c + d.nx] - p[c]);
        }
        if (k + 1 < d.nz) {
            const idx_t f = i + (idx_t)d.nx * (j + (idx_t)d.ny * (k + 1));
            const double nphi = fs[c + sxy];
            if (ww[f] > 0.0 && (cphi < 0 || nphi < 0))
                wo[f] = w[f] + invH[2] / fr_theta(cphi, nphi) * (p[c + sxy] - p[c]);
        }
    }
}

// ------------------------------------------------------------------ 2-D
//
// GridFractionalSinglePhasePressureSolver2 on a one-plane level (d.nz == 1).

// Face weights of one axis, 2-D (buildWeights,
// src/jet/grid_fractional_single_phase_pressure_solver2.cpp:440-480): 1 -
// fractionInsideSdf of the boundary SDF at the centres of the two cells the
// face separates, clamped, floored to 0.01 when positive, as float.  bsdf is
// the boundary SDF at the cell centres of the grid plus one ghost cell per
// side: cell (i, j) at bsdf[(i + 1) + (nx + 2)(j + 1)].
template <int AXIS>
__global__ __launch_bounds__(256) void k_fr_weights2(const double* __restrict__ bsdf, Dims d,
                                                     float* __restrict__ out) {
#pragma clang fp contract(off)
    const int fx = d.nx + (AXIS == 0), fy = d.ny + (AXIS == 1);
    const idx_t n = (idx_t)fx * fy;
    const idx_t gx = d.nx + 2;
    for (idx_t f = blockIdx.x * (idx_t)256 + threadIdx.x; f < n; f += (idx_t)gridDim.x * 256) {
        const int i = (int)(f % fx), j = (int)(f / fx);
        // the cell before the face (i - 1 or j - 1) and the one after it
        const idx_t g1 = (idx_t)(i + 1) + gx * (j + 1);
        const idx_t g0 = AXIS == 0 ? g1 - 1 : g1 - gx;
        double w = 1.0 - fr_inside_sdf(bsdf[g0], bsdf[g1]);
        w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
        if (w < 0.01 && w > 0.0) w = 0.01;
        out[f] = (float)w;
    }
}

// Float restriction of a one-plane cell or face array (the anonymous
// restrict() of src/jet/grid_fractional_single_phase_pressure_solver2.cpp:
// 29-93): the centred kernel, or injection along an axis whose fine size is
// not twice the coarse size; y outer, x inner, weight kx * ky, float sums.
__global__ __launch_bounds__(256) void k_fr_restrict2(const float* __restrict__ fine, int fnx, int fny,
                                                      float* __restrict__ coarse, int cnx, int cny) {
#pragma clang fp contract(off)
    const idx_t n = (idx_t)cnx * cny;
    const bool stx = fnx != 2 * cnx, sty = fny != 2 * cny;
    for (idx_t e = blockIdx.x * (idx_t)256 + threadIdx.x; e < n; e += (idx_t)gridDim.x * 256) {
        const int i = (int)(e % cnx), j = (int)(e / cnx);
        int ii[4], jj[4];
        float kx[4], ky[4];
        ii[0] = i > 0 ? 2 * i - 1 : 2 * i;
        ii[1] = 2 * i;
        jj[0] = j > 0 ? 2 * j - 1 : 2 * j;
        jj[1] = 2 * j;
        if (stx) {
            ii[2] = i + 1 < cnx ? 2 * i + 1 : 2 * i;
            ii[3] = 2 * i;
            kx[0] = 0.f, kx[1] = 1.f, kx[2] = 0.f, kx[3] = 0.f;
        } else {
            ii[2] = 2 * i + 1;
            ii[3] = i + 1 < cnx ? 2 * i + 2 : 2 * i + 1;
            kx[0] = 0.125f, kx[1] = 0.375f, kx[2] = 0.375f, kx[3] = 0.125f;
        }
        if (sty) {
            jj[2] = j + 1 < cny ? 2 * j + 1 : 2 * j;
            jj[3] = 2 * j;
            ky[0] = 0.f, ky[1] = 1.f, ky[2] = 0.f, ky[3] = 0.f;
        } else {
            jj[2] = 2 * j + 1;
            jj[3] = j + 1 < cny ? 2 * j + 2 : 2 * j + 1;
            ky[0] = 0.125f, ky[1] = 0.375f, ky[2] = 0.375f, ky[3] = 0.125f;
        }
        const int nkx = stx ? 3 : 4, nky = sty ? 3 : 4;
        float sum = 0.0f;
        for (int y = 0; y < nky; ++y)
            for (int x = 0; x < nkx; ++x) {
                const float w = kx[x] * ky[y];
                sum += w * fine[ii[x] + (idx_t)fnx * jj[y]];
            }
        coarse[e] = sum;
    }
}

// buildSingleSystem, 2-D variational form
// (src/jet/grid_fractional_single_phase_pressure_solver2.cpp:95-203) on one
// plane: the +x, -x, +y, -y faces in Jet's order, the collider term
// sum (1 - w) u_collider / h, and the row {1, 0} below machine epsilon or
// outside the fluid.  Rows are {center, right, up, 0}.  b is written only
// when u is given; bu / bv may be null (collider at rest).
__global__ __launch_bounds__(256) void k_fr_build2(
    const float* __restrict__ fs, const float* __restrict__ uw, const float* __restrict__ vw,
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ bu,
    const double* __restrict__ bv, Dims d, double hx, double hy, Row<double>* __restrict__ A,
    double* __restrict__ b) {
#pragma clang fp contract(off)
    const double invH[2] = {1.0 / hx, 1.0 / hy};
    const double invHSqr[2] = {invH[0] * invH[0], invH[1] * invH[1]};
    const idx_t n = (idx_t)d.nx * d.ny;
    const idx_t ux = d.nx + 1;
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        const int i = (int)(c % d.nx), j = (int)(c / d.nx);
        Row<double> row{0.0, 0.0, 0.0, 0.0};
        double bc = 0.0;
        const double cphi = fs[c];
        if (cphi < 0) {
            const idx_t uR = (i + 1) + ux * j, uL = uR - 1;
            const idx_t vU = i + (idx_t)d.nx * (j + 1), vD = vU - d.nx;
            const bool rhs = u != nullptr;
            double term;
            if (i + 1 < d.nx) {
                term = uw[uR] * invHSqr[0];
                const double phi = fs[c + 1];
                if (phi < 0) {
                    row.c += term;
                    row.r -= term;
                } else {
                    row.c += term / fr_theta(cphi, phi);
                }
                if (rhs) bc += uw[uR] * u[uR] * invH[0];
            } else if (rhs) {
                bc += u[uR] * invH[0];
            }
            if (i > 0) {
                term = uw[uL] * invHSqr[0];
                const double phi = fs[c - 1];
                if (phi < 0)
                    row.c += term;
                else
                    row.c += term / fr_theta(cphi, phi);
                if (rhs) bc -= uw[uL] * u[uL] * invH[0];
            } else if (rhs) {
                bc -= u[uL] * invH[0];
            }
            if (j + 1 < d.ny) {
                term = vw[vU] * invHSqr[1];
                const double phi = fs[c + d.nx];
                if (phi < 0) {
                    row.c += term;
                    row.u -= term;
                } else {
                    row.c += term / fr_theta(cphi, phi);
                }
                if (rhs) bc += vw[vU] * v[vU] * invH[1];
            } else if (rhs) {
                bc += v[vU] * invH[1];
            }
            if (j > 0) {
                term = vw[vD] * invHSqr[1];
                const double phi = fs[c - d.nx];
                if (phi < 0)
                    row.c += term;
                else
                    row.c += term / fr_theta(cphi, phi);
                if (rhs) bc -= vw[vD] * v[vD] * invH[1];
            } else if (rhs) {
                bc -= v[vD] * invH[1];
            }
            if (rhs) {
                const double buR = bu ? bu[uR] : 0.0, buL = bu ? bu[uL] : 0.0;
                const double bvU = bv ? bv[vU] : 0.0, bvD = bv ? bv[vD] : 0.0;
                const double contrib = (1.0 - uw[uR]) * buR * invH[0] - (1.0 - uw[uL]) * buL * invH[0] +
                                       (1.0 - vw[vU]) * bvU * invH[1] - (1.0 - vw[vD]) * bvD * invH[1];
                bc += contrib;
            }
            if (row.c < 2.220446049250313080847e-16) {  // DBL_EPSILON (include/jet/constants.h:74)
                row.c = 1.0;
                bc = 0.0;
            }
        } else {
            row.c = 1.0;
        }
        A[c] = row;
        if (b) b[c] = bc;
    }
}

// applyPressureGradient, 2-D variational form
// (src/jet/grid_fractional_single_phase_pressure_solver2.cpp:565-600): the
// +x / +y face of cell c is updated when its weight is positive and either
// side is fluid.  The output may alias the input.
__global__ __launch_bounds__(256) void k_fr_gradient2(
    const float* __restrict__ fs, const float* __restrict__ uw, const float* __restrict__ vw,
    const double* __restrict__ p, const double* u, const double* v, double* uo, double* vo,
    Dims d, double hx, double hy) {
#pragma clang fp contract(off)
    const double invH[2] = {1.0 / hx, 1.0 / hy};
    const idx_t n = (idx_t)d.nx * d.ny;
    for (idx_t c = blockIdx.x * (idx_t)256 + threadIdx.x; c < n; c += (idx_t)gridDim.x * 256) {
        const int i = (int)(c % d.nx), j = (int)(c / d.nx);
        const double cphi = fs[c];
        if (i + 1 < d.nx) {
            const idx_t f = (i + 1) + (idx_t)(d.nx + 1) * j;
            const double nphi = fs[c + 1];
            if (uw[f] > 0.0 && (cphi < 0 || nphi < 0))
                uo[f] = u[f] + invH[0] / fr_theta(cphi, nphi) * (p[c + 1] - p[c]);
        }
        if (j + 1 < d.ny) {
            const idx_t f = i + (idx_t)d.nx * (j + 1);
            const double nphi = fs[c + d.nx];
            if (vw[f] > 0.0 && (cphi < 0 || nphi < 0))
                vo[f] = v[f] + invH[1] / fr_theta(cphi, nphi) * (p[c + d.nx] - p[c]);
        }
    }
}

}  // namespace jg
