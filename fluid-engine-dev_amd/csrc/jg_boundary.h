// jg_boundary.h -- the velocity boundary conditions Jet applies after the
// pressure projection: GridFractionalBoundaryConditionSolver3::
// constrainVelocity (src/jet/grid_fractional_boundary_condition_solver3.cpp:
// 26-232) and the blocked solver's extra pass (src/jet/grid_blocked_boundary_
// condition_solver3.cpp:23-69), for a rigid-body collider.
//
// Every stage is elementwise over the faces of one axis (u, v or w), one
// thread per face, so each is one launch per axis:
//   init       marker of the face (1 where any part of it is outside the
//              collider), collider velocity on fully covered faces;
//   extrapolate one pass of extrapolateToRegion (include/jet/detail/
//              array_utils-inl.h:172-241) -- a pass writes only faces that were
//              invalid and reads only faces that were valid at its start, so
//              its faces are independent; validity ping-pongs between two
//              byte arrays;
//   project    no-flux projection with friction inside the collider, into a
//              scratch copy (it samples all three extrapolated components);
//   finish     copy back, zeroing the closed domain walls;
//   blocked    faces between a collider cell and a fluid cell take the
//              collider velocity.
// FP64 with contraction off, in Jet's operand order: the result is bitwise
// Jet's.
#pragma once

#include "jg_device.h"

namespace jg {

// Cells nx*ny*nz, spacing and origin of the face-centred grid.
struct BcGrid {
    Dims d;
    double h[3], o[3];
};

// RigidBodyCollider3 (velocities, surface translation) and Collider3's friction.
struct BcCollider {
    double lin[3], ang[3], tr[3], friction;
};

// Face grid of axis A: sizes and data origin o + h/2 off the axis
// (src/jet/face_centered_grid3.cpp:216-238, 437-447).
template <int A>
__device__ __forceinline__ void bc_face_dims(const Dims& d, int s[3]) {
    s[0] = d.nx + (A == 0);
    s[1] = d.ny + (A == 1);
    s[2] = d.nz + (A == 2);
}
template <int A>
__device__ __forceinline__ void bc_face_origin(const BcGrid& g, double fo[3]) {
#pragma clang fp contract(off)
    for (int b = 0; b < 3; ++b) fo[b] = g.o[b] + 0.5 * (b == A ? 0.0 : g.h[b]);
}
__device__ __forceinline__ void bc_pos(const double fo[3], const double h[3], int i, int j, int k,
                                       double p[3]) {
#pragma clang fp contract(off)
    p[0] = fo[0] + h[0] * (double)i;
    p[1] = fo[1] + h[1] * (double)j;
    p[2] = fo[2] + h[2] * (double)k;
}

// RigidBodyCollider3::velocityAt (src/jet/rigid_body_collider3.cpp:25-28) with
// Vector3::cross (include/jet/detail/vector3-inl.h:125-127).
__device__ __forceinline__ void bc_collider_velocity(const BcCollider& c, const double p[3],
                                                     double out[3]) {
#pragma clang fp contract(off)
    const double r0 = p[0] - c.tr[0], r1 = p[1] - c.tr[1], r2 = p[2] - c.tr[2];
    out[0] = c.lin[0] + (c.ang[1] * r2 - r1 * c.ang[2]);
    out[1] = c.lin[1] + (c.ang[2] * r0 - r2 * c.ang[0]);
    out[2] = c.lin[2] + (c.ang[0] * r1 - r0 * c.ang[1]);
}

// getBarycentric (include/jet/detail/math_utils-inl.h:147-175), iLow = 0.
__device__ __forceinline__ void bc_barycentric(double x, int ihigh, int& i, double& f) {
#pragma clang fp contract(off)
    const double s = floor(x);
    i = (int)s;
    if (ihigh == 0) {
        i = 0;
        f = 0;
    } else if (i < 0) {
        i = 0;
        f = 0;
    } else if (i > ihigh - 1) {
        i = ihigh - 1;
        f = 1;
    } else {
        f = x - s;
    }
}

__device__ __forceinline__ double bc_lerp(double a, double b, double f) {
#pragma clang fp contract(off)
    return (1 - f) * a + f * b;  // include/jet/detail/math_utils-inl.h:177-180
}

// LinearArraySampler3::operator() (include/jet/detail/array_samplers3-inl.h:
// 116-156): data of size s at data origin dorg, spacing h.
__device__ double bc_sample(const double* __restrict__ data, const int s[3], const double h[3],
                            const double dorg[3], const double p[3]) {
#pragma clang fp contract(off)
    int i, j, k;
    double fx, fy, fz;
    bc_barycentric((p[0] - dorg[0]) / h[0], s[0] - 1, i, fx);
    bc_barycentric((p[1] - dorg[1]) / h[1], s[1] - 1, j, fy);
    bc_barycentric((p[2] - dorg[2]) / h[2], s[2] - 1, k, fz);
    const int ip1 = min(i + 1, s[0] - 1), jp1 = min(j + 1, s[1] - 1), kp1 = min(k + 1, s[2] - 1);
    auto D = [&](int a, int b, int c) { return data[a + (idx_t)s[0] * (b + (idx_t)s[1] * c)]; };
    const double b0 = bc_lerp(bc_lerp(D(i, j, k), D(ip1, j, k), fx),
                              bc_lerp(D(i, jp1, k), D(ip1, jp1, k), fx), fy);
    const double b1 = bc_lerp(bc_lerp(D(i, j, kp1), D(ip1, j, kp1), fx),
                              bc_lerp(D(i, jp1, kp1), D(ip1, jp1, kp1), fx), fy);
    return bc_lerp(b0, b1, fz);
}

// The collider SDF (cell-centred, data origin o + h/2) at p.
__device__ __forceinline__ double bc_sdf(const double* __restrict__ sdf, const BcGrid& g,
                                         const double p[3]) {
#pragma clang fp contract(off)
    const int s[3] = {g.d.nx, g.d.ny, g.d.nz};
    const double dorg[3] = {g.o[0] + 0.5 * g.h[0], g.o[1] + 0.5 * g.h[1], g.o[2] + 0.5 * g.h[2]};
    return bc_sample(sdf, s, g.h, dorg, p);
}

// ScalarGrid3::gradient (src/jet/scalar_grid3.cpp:89-102): the weights of
// getCoordinatesAndWeights (include/jet/detail/array_samplers3-inl.h:152-194,
// which scales by 1/h) on gradient3 at the 8 data points (src/jet/fdm_utils.cpp:49-67).
__device__ void bc_sdf_gradient(const double* __restrict__ sdf, const BcGrid& g, const double p[3],
                                double out[3]) {
#pragma clang fp contract(off)
    const int s[3] = {g.d.nx, g.d.ny, g.d.nz};
    int lo[3], hi[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
        const double dorg = g.o[a] + 0.5 * g.h[a];
        const double inv = 1.0 / g.h[a];
        bc_barycentric((p[a] - dorg) * inv, s[a] - 1, lo[a], f[a]);
        hi[a] = min(lo[a] + 1, s[a] - 1);
    }
    const double fx = f[0], fy = f[1], fz = f[2];
    const double wt[8] = {(1 - fx) * (1 - fy) * (1 - fz), fx * (1 - fy) * (1 - fz),
                          (1 - fx) * fy * (1 - fz),       fx * fy * (1 - fz),
                          (1 - fx) * (1 - fy) * fz,       fx * (1 - fy) * fz,
                          (1 - fx) * fy * fz,             fx * fy * fz};
    auto D = [&](int a, int b, int c) { return sdf[a + (idx_t)s[0] * (b + (idx_t)s[1] * c)]; };
    out[0] = out[1] = out[2] = 0.0;
    for (int q = 0; q < 8; ++q) {
        const int i = (q & 1) ? hi[0] : lo[0];
        const int j = (q & 2) ? hi[1] : lo[1];
        const int k = (q & 4) ? hi[2] : lo[2];
        const double left = D(i > 0 ? i - 1 : i, j, k), right = D(i + 1 < s[0] ? i + 1 : i, j, k);
        const double down = D(i, j > 0 ? j - 1 : j, k), up = D(i, j + 1 < s[1] ? j + 1 : j, k);
        const double back = D(i, j, k > 0 ? k - 1 : k), front = D(i, j, k + 1 < s[2] ? k + 1 : k);
        out[0] += wt[q] * (0.5 * (right - left) / g.h[0]);
        out[1] += wt[q] * (0.5 * (up - down) / g.h[1]);
        out[2] += wt[q] * (0.5 * (front - back) / g.h[2]);
    }
}

// projectAndApplyFriction (src/jet/physics_helpers.h:34-45).
__device__ __forceinline__ void bc_friction(const double vel[3], const double n[3], double friction,
                                            double out[3]) {
#pragma clang fp contract(off)
    const double dn = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
    out[0] = vel[0] - n[0] * dn;
    out[1] = vel[1] - n[1] * dn;
    out[2] = vel[2] - n[2] * dn;
    if (out[0] * out[0] + out[1] * out[1] + out[2] * out[2] > 0) {
        const double nd = -(vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2]);
        const double veln = nd < 0.0 ? 0.0 : nd;
        double sc = 1.0 - friction * veln / sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
        sc = sc < 0.0 ? 0.0 : sc;
        out[0] *= sc;
        out[1] *= sc;
        out[2] *= sc;
    }
}

// fractionInsideSdf (include/jet/detail/level_set_utils-inl.h:51-62).
__device__ __forceinline__ double bc_fraction_inside_sdf(double phi0, double phi1) {
#pragma clang fp contract(off)
    if (phi0 < 0 && phi1 < 0) return 1;
    if (phi0 < 0 && !(phi1 < 0)) return phi0 / (phi0 - phi1);
    if (!(phi0 < 0) && phi1 < 0) return phi1 / (phi1 - phi0);
    return 0;
}

#define JG_BC_FACES(A, g, s, n)                                                      \
    int s[3];                                                                        \
    bc_face_dims<A>((g).d, s);                                                       \
    const idx_t n = (idx_t)s[0] * s[1] * s[2];                                       \
    for (idx_t f = blockIdx.x * (idx_t)blockDim.x + threadIdx.x; f < n;               \
         f += (idx_t)gridDim.x * blockDim.x)

// Stage 1 (grid_fractional_boundary_condition_solver3.cpp:54-102): the face
// is valid (1) unless its whole extent lies in the collider; there it takes
// the collider velocity.
template <int A>
__global__ __launch_bounds__(256) void k_bc_init(BcGrid g, const double* __restrict__ sdf,
                                                 BcCollider c, double* __restrict__ comp,
                                                 int8_t* __restrict__ valid) {
#pragma clang fp contract(off)
    double fo[3];
    bc_face_origin<A>(g, fo);
    JG_BC_FACES(A, g, s, n) {
        const int i = (int)(f % s[0]), j = (int)((f / s[0]) % s[1]), k = (int)(f / ((idx_t)s[0] * s[1]));
        double p[3], q0[3], q1[3];
        bc_pos(fo, g.h, i, j, k, p);
        for (int b = 0; b < 3; ++b) {
            const double off = b == A ? 0.5 * g.h[A] : 0.0;
            q0[b] = p[b] - off;
            q1[b] = p[b] + off;
        }
        double frac = bc_fraction_inside_sdf(bc_sdf(sdf, g, q0), bc_sdf(sdf, g, q1));
        frac = 1.0 - (frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac));  // clamp(frac, 0, 1)
        if (frac > 0.0) {
            valid[f] = 1;
        } else {
            double cv[3];
            bc_collider_velocity(c, p, cv);
            comp[f] = cv[A];
            valid[f] = 0;
        }
    }
}

// Stage 2: one pass of extrapolateToRegion, neighbour order +x, -x, +y, -y,
// +z, -z as Jet's.
template <int A>
__global__ __launch_bounds__(256) void k_bc_extrapolate(Dims d, double* __restrict__ comp,
                                                        const int8_t* __restrict__ v0,
                                                        int8_t* __restrict__ v1) {
#pragma clang fp contract(off)
    BcGrid g;
    g.d = d;
    JG_BC_FACES(A, g, s, n) {
        if (v0[f]) {
            v1[f] = 1;
            continue;
        }
        const int i = (int)(f % s[0]), j = (int)((f / s[0]) % s[1]), k = (int)(f / ((idx_t)s[0] * s[1]));
        const idx_t sy = s[0], sz = (idx_t)s[0] * s[1];
        double sum = 0.0;
        unsigned count = 0;
        if (i + 1 < s[0] && v0[f + 1]) { sum += comp[f + 1]; ++count; }
        if (i > 0 && v0[f - 1]) { sum += comp[f - 1]; ++count; }
        if (j + 1 < s[1] && v0[f + sy]) { sum += comp[f + sy]; ++count; }
        if (j > 0 && v0[f - sy]) { sum += comp[f - sy]; ++count; }
        if (k + 1 < s[2] && v0[f + sz]) { sum += comp[f + sz]; ++count; }
        if (k > 0 && v0[f - sz]) { sum += comp[f - sz]; ++count; }
        if (count > 0) {
            comp[f] = sum / (double)count;
            v1[f] = 1;
        } else {
            v1[f] = 0;
        }
    }
}

// Stage 3 (:104-176): inside the collider, the velocity relative to the
// collider loses its normal component and friction scales the tangential
// one; elsewhere the value is kept.  Reads u, v, w; writes out.
template <int A>
__global__ __launch_bounds__(256) void k_bc_project(BcGrid g, const double* __restrict__ sdf,
                                                    BcCollider c, const double* __restrict__ u,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
    double fo[3];
    bc_face_origin<A>(g, fo);
    const double* comp = A == 0 ? u : (A == 1 ? v : w);
    JG_BC_FACES(A, g, s, n) {
        const int i = (int)(f % s[0]), j = (int)((f / s[0]) % s[1]), k = (int)(f / ((idx_t)s[0] * s[1]));
        double p[3];
        bc_pos(fo, g.h, i, j, k, p);
        if (!(bc_sdf(sdf, g, p) < 0)) {
            out[f] = comp[f];
            continue;
        }
        double cv[3], vel[3], gr[3];
        bc_collider_velocity(c, p, cv);
        {
            int su[3], sv[3], sw[3];
            double ou[3], ov[3], ow[3];
            bc_face_dims<0>(g.d, su);
            bc_face_dims<1>(g.d, sv);
            bc_face_dims<2>(g.d, sw);
            bc_face_origin<0>(g, ou);
            bc_face_origin<1>(g, ov);
            bc_face_origin<2>(g, ow);
            vel[0] = bc_sample(u, su, g.h, ou, p);  // FaceCenteredGrid3::sample
            vel[1] = bc_sample(v, sv, g.h, ov, p);
            vel[2] = bc_sample(w, sw, g.h, ow, p);
        }
        bc_sdf_gradient(sdf, g, p, gr);
        if (gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2] > 0.0) {
            const double l = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
            const double nrm[3] = {gr[0] / l, gr[1] / l, gr[2] / l};
            const double velr[3] = {vel[0] - cv[0], vel[1] - cv[1], vel[2] - cv[2]};
            double velt[3];
            bc_friction(velr, nrm, c.friction, velt);
            out[f] = velt[A] + cv[A];
        } else {
            out[f] = cv[A];
        }
    }
}

// Stage 4 (:178-232): the projected values back into the component, zero on
// closed domain walls (closed: Jet's kDirection* bits, include/jet/
// constants.h:312-335).
template <int A>
__global__ __launch_bounds__(256) void k_bc_finish(Dims d, const double* __restrict__ tmp,
                                                   double* __restrict__ comp, int closed) {
    BcGrid g;
    g.d = d;
    const int lo = A == 0 ? 1 : (A == 1 ? 4 : 16), hi = 2 * lo;
    JG_BC_FACES(A, g, s, n) {
        const int i = (int)(f % s[0]), j = (int)((f / s[0]) % s[1]), k = (int)(f / ((idx_t)s[0] * s[1]));
        const int a = A == 0 ? i : (A == 1 ? j : k);
        const bool wall = ((closed & lo) && a == 0) || ((closed & hi) && a == s[A] - 1);
        comp[f] = wall ? 0.0 : tmp[f];
    }
}

// Blocked solver's pass (src/jet/grid_blocked_boundary_condition_solver3.cpp:
// 39-66, marker :86-94): an interior face whose two cells are one collider
// (SDF < 0 at the centre) and one fluid takes the collider velocity.
template <int A>
__global__ __launch_bounds__(256) void k_bc_blocked(BcGrid g, const double* __restrict__ sdf,
                                                    BcCollider c, double* __restrict__ comp) {
#pragma clang fp contract(off)
    double fo[3];
    bc_face_origin<A>(g, fo);
    JG_BC_FACES(A, g, s, n) {
        const int i = (int)(f % s[0]), j = (int)((f / s[0]) % s[1]), k = (int)(f / ((idx_t)s[0] * s[1]));
        const int a = A == 0 ? i : (A == 1 ? j : k);
        if (a == 0 || a == s[A] - 1) continue;
        const idx_t hi = i + (idx_t)g.d.nx * (j + (idx_t)g.d.ny * k);  // the cell at +A
        const idx_t lo = hi - (A == 0 ? 1 : (A == 1 ? (idx_t)g.d.nx : (idx_t)g.d.nx * g.d.ny));
        const bool chi = sdf[hi] < 0, clo = sdf[lo] < 0;
        if (chi == clo) continue;
        double p[3], cv[3];
        bc_pos(fo, g.h, i, j, k, p);
        bc_collider_velocity(c, p, cv);
        comp[f] = cv[A];
    }
}

#undef JG_BC_FACES

}  // namespace jg
