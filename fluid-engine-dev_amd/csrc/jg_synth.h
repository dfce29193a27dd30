// jg_synth.h -- synthetic pressure-projection inputs (host and device).
//
// One source for the CPU and the GPU generators, so a test can compare a GPU
// result with the CPU oracle on identical inputs without shipping data.
// Contraction is off in every numeric function so host and device evaluate
// every SDF and velocity bit-identically.
//
// RNG: counter-based SplitMix64 (Steele, Lea, Flood 2014, finaliser of
// Vigna's splitmix64.c): u = (mix(seed ^ (stream << 56) ^ idx) >> 11) * 2^-53.
// Streams: 1 = u faces, 2 = v faces, 3 = w faces, 4 = scene parameters,
// 5 = batch sizes.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace jg {
namespace synth {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double uniform01(uint64_t seed, uint64_t stream, uint64_t idx) {
    return (double)(mix64(seed ^ (stream << 56) ^ idx) >> 11) * 0x1.0p-53;
}

__host__ __device__ inline double uniform(uint64_t seed, uint64_t stream, uint64_t idx, double a,
                                          double b) {
#pragma clang fp contract(off)
    return a + (b - a) * uniform01(seed, stream, idx);
}

constexpr int kMaxSpheres = 8;

// Scene description after parameter derivation (plain data, host/device).
struct Scene {
    int kind;        // jg scene kind (see jetgpu.h)
    int n;           // cubic grid edge
    double h;        // spacing
    uint64_t seed;
    double gravity_dt;   // added to every v face before boundary zeroing
    int nspheres;
    double sc[kMaxSpheres][3];
    double sr[kMaxSpheres];
    double plane_y;      // fluid below y = plane_y (world units); huge = all fluid
    double dam_x;        // dam-break: fluid where x < dam_x and y < plane_y
    int torus;           // torus collider present
    double tc[3], tR, tr;
};

constexpr double kBig = 1.0e300;

__host__ __device__ inline double boundary_sdf(const Scene& s, double x, double y, double z) {
#pragma clang fp contract(off)
    double d = kBig;
    for (int q = 0; q < s.nspheres; ++q) {
        const double dx = x - s.sc[q][0], dy = y - s.sc[q][1], dz = z - s.sc[q][2];
        const double v = sqrt(dx * dx + dy * dy + dz * dz) - s.sr[q];
        d = v < d ? v : d;
    }
    if (s.torus) {
        const double dx = x - s.tc[0], dz = z - s.tc[2];
        const double a = sqrt(dx * dx + dz * dz) - s.tR;
        const double b = y - s.tc[1];
        const double v = sqrt(a * a + b * b) - s.tr;
        d = v < d ? v : d;
    }
    return d;
}

__host__ __device__ inline double fluid_sdf(const Scene& s, double x, double y, double z) {
#pragma clang fp contract(off)
    (void)z;
    const double a = y - s.plane_y;
    if (s.kind == 2) {
        const double b = x - s.dam_x;
        return a > b ? a : b;
    }
    return a;
}

// Cell-centre position (origin 0): Grid3::cellCenterPosition,
// src/jet/grid3.cpp:32-38.
__host__ __device__ inline double cc(const Scene& s, int i) {
#pragma clang fp contract(off)
    return s.h * ((double)i + 0.5);
}

// Is the cell (i,j,k) solid?  Out-of-range cells are not solid.
__host__ __device__ inline bool solid_cell(const Scene& s, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= s.n || j >= s.n || k >= s.n) return false;
    return boundary_sdf(s, cc(s, i), cc(s, j), cc(s, k)) < 0;
}

// Face velocity: U(-1,1) (+ gravity on v), then the blocked boundary
// condition: domain-wall faces and faces touching a solid cell are zero
// (src/jet/grid_blocked_boundary_condition_solver3.cpp:23-69,
// src/jet/grid_fractional_boundary_condition_solver3.cpp:191-230).
// axis 0/1/2 = u/v/w; (i,j,k) indexes the face array of that axis.
__host__ __device__ inline double face_velocity(const Scene& s, int axis, int i, int j, int k) {
#pragma clang fp contract(off)
    const int n = s.n;
    const int sx = n + (axis == 0), sy = n + (axis == 1);
    const uint64_t idx = (uint64_t)i + (uint64_t)sx * ((uint64_t)j + (uint64_t)sy * (uint64_t)k);
    double val = uniform(s.seed, (uint64_t)(axis + 1), idx, -1.0, 1.0);
    if (axis == 1) val = val + s.gravity_dt;
    const int f = axis == 0 ? i : (axis == 1 ? j : k);
    if (f == 0 || f == n) return 0.0;
    const int im = i - (axis == 0), jm = j - (axis == 1), km = k - (axis == 2);
    if (solid_cell(s, im, jm, km) || solid_cell(s, i, j, k)) return 0.0;
    return val;
}

}  // namespace synth
}  // namespace jg

