// jg_pressure.hip -- C ABI of the pressure-system assembly, the gradient
// projection and the synthetic-scene generators.
#include <thread>

#include <hipcub/hipcub.hpp>

#include "jg_internal.h"
#include "jg_pressure.h"
#include "jg_fractional.h"
#include "jg_boundary.h"

using namespace jg;

namespace {

// Scene parameters from the description (stream 4 of the RNG).  The plane
// of JG_SCENE_SPHERE_POOL is chosen afterwards from row counts.
int fill_scene(const jg_scene_desc* d, synth::Scene* s) {
    if (!d || d->n < 2 || d->n > 4096) return JG_ERR_INVALID_ARG;
    std::memset(s, 0, sizeof(*s));
    s->kind = d->kind;
    s->n = (int)d->n;
    s->h = d->h > 0 ? d->h : 1.0;
    s->seed = d->seed;
    s->plane_y = 2.0 * synth::kBig;  // all fluid
    const double n = (double)d->n;
    const double L = n * s->h;
    auto U = [&](uint64_t idx, double a, double b) { return synth::uniform(d->seed, 4, idx, a, b); };
    switch (d->kind) {
        case JG_SCENE_CLOSED_SPHERE:
            s->nspheres = 1;
            s->sc[0][0] = s->sc[0][1] = s->sc[0][2] = 0.5 * L;
            s->sr[0] = 0.25 * L;
            break;
        case JG_SCENE_DAM_BREAK:
            s->plane_y = 0.4 * L;
            s->dam_x = 0.5 * L;
            s->gravity_dt = -9.8 * (1.0 / 60.0);
            break;
        case JG_SCENE_SPHERE_POOL:
            s->nspheres = 8;
            for (int q = 0; q < 8; ++q) {
                const double r = U(4 * q, 8.0, 32.0) * n / 256.0;
                s->sr[q] = r * s->h;
                for (int a = 0; a < 3; ++a) s->sc[q][a] = U(4 * q + 1 + a, r + 1.0, n - 1.0 - r) * s->h;
            }
            break;
        case JG_SCENE_BATCH_MEMBER: {
            s->nspheres = 1 + (int)(4.0 * U(100, 0.0, 1.0));
            if (s->nspheres > 4) s->nspheres = 4;
            for (int q = 0; q < s->nspheres; ++q) {
                const double r = U(4 * q, n / 16.0, n / 6.0);
                s->sr[q] = r * s->h;
                for (int a = 0; a < 3; ++a) s->sc[q][a] = U(4 * q + 1 + a, r + 1.0, n - 1.0 - r) * s->h;
            }
            s->plane_y = U(101, 0.3, 0.8) * L;
            break;
        }
        case JG_SCENE_TORUS_POOL:
            s->torus = 1;
            s->tc[0] = s->tc[1] = s->tc[2] = 0.5 * L;
            s->tR = 150.0 / 512.0 * L;
            s->tr = 50.0 / 512.0 * L;
            s->plane_y = 0.35 * L;
            break;
        default:
            return JG_ERR_INVALID_ARG;
    }
    return JG_OK;
}

// Fluid-below-plane height giving the requested fluid fraction: the lowest
// whole row j* with sum_{j<j*} open(j) >= fraction * n^3.
void choose_plane(synth::Scene* s, const std::vector<unsigned>& open, double fraction) {
    const double goal = fraction * (double)s->n * s->n * s->n;
    double acc = 0.0;
    int j = 0;
    while (j < s->n && acc < goal) acc += open[j++];
    s->plane_y = s->h * (double)j;
}

template <typename F>
void parallel_planes(int nz, F f) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int k = t; k < nz; k += nt) f(k);
        });
    for (auto& x : th) x.join();
}

dim3 plane_grid(int nx, int ny, int nz, dim3* block) {
    const int bx = std::min(64, pow2ceil(std::max(1, nx)));
    *block = dim3(bx, 256 / bx, 1);
    return dim3((nx + bx - 1) / bx, (ny + block->y - 1) / block->y, nz);
}

}  // namespace

extern "C" {

int jg_scene_generate_host(const jg_scene_desc* d, double* u, double* v, double* w,
                           double* bsdf_cc, double* fsdf_cc, double* bsdf_vertex) {
    synth::Scene s;
    int st = fill_scene(d, &s);
    if (st != JG_OK) return st;
    const int n = s.n;
    if (d->kind == JG_SCENE_SPHERE_POOL) {
        std::vector<unsigned> open(n, 0);
        std::vector<std::vector<unsigned>> part(n, std::vector<unsigned>(n, 0));
        parallel_planes(n, [&](int k) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) part[k][j] += synth::solid_cell(s, i, j, k) ? 0u : 1u;
        });
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) open[j] += part[k][j];
        choose_plane(&s, open, d->fluid_fraction > 0 ? d->fluid_fraction : 0.6);
    }
    const long n1 = n + 1;
    parallel_planes(n + 1, [&](int k) {
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                if (u && k < n && j < n) u[i + n1 * (j + (long)n * k)] = synth::face_velocity(s, 0, i, j, k);
                if (v && k < n && i < n) v[i + (long)n * (j + n1 * k)] = synth::face_velocity(s, 1, i, j, k);
                if (w && j < n && i < n) w[i + (long)n * (j + (long)n * k)] = synth::face_velocity(s, 2, i, j, k);
                if (k < n && j < n && i < n) {
                    const long c = i + (long)n * (j + (long)n * k);
                    const double x = synth::cc(s, i), y = synth::cc(s, j), z = synth::cc(s, k);
                    if (bsdf_cc) bsdf_cc[c] = synth::boundary_sdf(s, x, y, z);
                    if (fsdf_cc) fsdf_cc[c] = synth::fluid_sdf(s, x, y, z);
                }
                if (bsdf_vertex)
                    bsdf_vertex[i + n1 * (j + n1 * k)] =
                        synth::boundary_sdf(s, s.h * (double)i, s.h * (double)j, s.h * (double)k);
            }
    });
    return JG_OK;
}

int jg_scene_generate_device(jg_solver* s, const jg_scene_desc* d, double* u, double* v,
                             double* w, double* bsdf_cc, double* fsdf_cc, double* bsdf_vertex) {
    if (!s) return JG_ERR_INVALID_ARG;
    synth::Scene sc;
    int st = fill_scene(d, &sc);
    if (st != JG_OK) return st;
    JG_HIP(hipSetDevice(s->device));
    const int n = sc.n;
    dim3 blk;
    if (d->kind == JG_SCENE_SPHERE_POOL) {
        unsigned* rows = nullptr;
        JG_HIP(hipMalloc(&rows, n * sizeof(unsigned)));
        std::vector<unsigned> open(n, 0);
        hipError_t e = hipMemsetAsync(rows, 0, n * sizeof(unsigned), s->stream);
        if (e == hipSuccess) {
            dim3 g = plane_grid(n, n, n, &blk);
            k_synth_row_open<<<g, blk, 0, s->stream>>>(sc, rows);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(open.data(), rows, n * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        (void)hipFree(rows);
        if (e != hipSuccess) {
            s->err = std::string("scene row count: ") + hipGetErrorString(e);
            return JG_ERR_HIP;
        }
        choose_plane(&sc, open, d->fluid_fraction > 0 ? d->fluid_fraction : 0.6);
    }
    if (u) k_synth_faces<<<plane_grid(n + 1, n, n, &blk), blk, 0, s->stream>>>(sc, 0, u);
    if (v) k_synth_faces<<<plane_grid(n, n + 1, n, &blk), blk, 0, s->stream>>>(sc, 1, v);
    if (w) k_synth_faces<<<plane_grid(n, n, n + 1, &blk), blk, 0, s->stream>>>(sc, 2, w);
    if (bsdf_cc && fsdf_cc)
        k_synth_cells<<<plane_grid(n, n, n, &blk), blk, 0, s->stream>>>(sc, bsdf_cc, fsdf_cc);
    else if (bsdf_cc || fsdf_cc) {
        s->err = "bsdf_cc and fsdf_cc are generated together";
        return JG_ERR_INVALID_ARG;
    }
    if (bsdf_vertex)
        k_synth_vertices<<<plane_grid(n + 1, n + 1, n + 1, &blk), blk, 0, s->stream>>>(sc, bsdf_vertex);
    JG_TRY(check_launch(s, "scene"));
    return sync(s);
}

int jg_sp_build(jg_solver* s, const double* u, const double* v, const double* w,
                const double* boundary_sdf, const double* fluid_sdf, double hx, double hy,
                double hz) {
    if (!s || s->lv.empty() || !u || !v || !w || !boundary_sdf || !fluid_sdf || !(hx > 0) ||
        !(hy > 0) || !(hz > 0))
        return JG_ERR_INVALID_ARG;
    if (s->twod) {
        // k_sp_coarsen reads fine planes 2k and 2k+1: a one-plane hierarchy has no plane 1
        s->err = "jg_sp_build: the resident hierarchy is a 2-D (one-plane) one; use jg_sp_build2";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const Dims& d0 = s->lv[0].d;
    k_sp_markers<<<nblocks_1d(d0.n()), 256, 0, s->stream>>>(boundary_sdf, fluid_sdf,
                                                            s->lv[0].markers, d0.n());
    for (size_t l = 1; l < s->lv.size(); ++l) {
        Geo g = geo_planes(s->lv[l].d);
        k_sp_coarsen<<<g.grid, g.block, 0, s->stream>>>(s->lv[l - 1].markers, s->lv[l].markers,
                                                        s->lv[l].d);
    }
    H3 h{hx, hy, hz};
    for (size_t l = 0; l < s->lv.size(); ++l) {
        Geo g = geo_planes(s->lv[l].d);
        k_sp_build<<<g.grid, g.block, 0, s->stream>>>(s->lv[l].markers, l == 0 ? u : nullptr, v, w,
                                                      s->lv[l].d, h, s->lv[l].A64,
                                                      l == 0 ? s->b : nullptr);
        h.x *= 2.0;
        h.y *= 2.0;
        h.z *= 2.0;
    }
    s->f32_valid = false;
    JG_TRY(check_launch(s, "sp_build"));
    return sync(s);
}

int jg_sp_apply_gradient(jg_solver* s, const double* u, const double* v, const double* w,
                         double* uo, double* vo, double* wo, double hx, double hy, double hz) {
    if (!s || s->lv.empty() || !u || !v || !w || !uo || !vo || !wo) return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    const Dims& d0 = s->lv[0].d;
    Geo g = geo_planes(d0);
    k_sp_gradient<<<g.grid, g.block, 0, s->stream>>>(s->lv[0].markers, s->x, u, v, w, uo, vo, wo,
                                                     d0, H3{hx, hy, hz});
    JG_TRY(check_launch(s, "sp_gradient"));
    return sync(s);
}

int jg_sp_build2(jg_solver* s, const double* u, const double* v, const double* boundary_sdf,
                 const double* fluid_sdf, double hx, double hy) {
    if (!s || s->lv.empty() || !u || !v || !boundary_sdf || !fluid_sdf || !(hx > 0) || !(hy > 0))
        return JG_ERR_INVALID_ARG;
    if (s->lv[0].d.nz != 1) {
        s->err = "jg_sp_build2: the resident hierarchy is not a 2-D (one-plane) one";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const Dims& d0 = s->lv[0].d;
    k_sp_markers<<<nblocks_1d(d0.n()), 256, 0, s->stream>>>(boundary_sdf, fluid_sdf,
                                                            s->lv[0].markers, d0.n());
    for (size_t l = 1; l < s->lv.size(); ++l) {
        Geo g = geo_planes(s->lv[l].d);
        k_sp_coarsen2<<<g.grid, g.block, 0, s->stream>>>(s->lv[l - 1].markers, s->lv[l].markers,
                                                         s->lv[l].d);
    }
    H3 h{hx, hy, 1.0};
    for (size_t l = 0; l < s->lv.size(); ++l) {
        Geo g = geo_planes(s->lv[l].d);
        k_sp_build<<<g.grid, g.block, 0, s->stream>>>(s->lv[l].markers, l == 0 ? u : nullptr, v,
                                                      nullptr, s->lv[l].d, h, s->lv[l].A64,
                                                      l == 0 ? s->b : nullptr);
        h.x *= 2.0;
        h.y *= 2.0;
    }
    s->f32_valid = false;
    JG_TRY(check_launch(s, "sp_build2"));
    return sync(s);
}

int jg_sp_apply_gradient2(jg_solver* s, const double* u, const double* v, double* uo, double* vo,
                          double hx, double hy) {
    if (!s || s->lv.empty() || !u || !v || !uo || !vo) return JG_ERR_INVALID_ARG;
    if (s->lv[0].d.nz != 1) {
        s->err = "jg_sp_apply_gradient2: the resident hierarchy is not a 2-D (one-plane) one";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const Dims& d0 = s->lv[0].d;
    Geo g = geo_planes(d0);
    // one plane: the kernel's +z branch never runs, so w / wo are not read
    k_sp_gradient<<<g.grid, g.block, 0, s->stream>>>(s->lv[0].markers, s->x, u, v, nullptr, uo, vo,
                                                     nullptr, d0, H3{hx, hy, 1.0});
    JG_TRY(check_launch(s, "sp_gradient2"));
    return sync(s);
}

const signed char* jg_solver_markers(jg_solver* s, size_t level) {
    if (!s || level >= s->lv.size()) return nullptr;
    return (const signed char*)s->lv[level].markers;
}

int jg_fr_build(jg_solver* s, const double* u, const double* v, const double* w,
                const double* boundary_sdf_vertex, const double* fluid_sdf, const double* bu,
                const double* bv, const double* bw, double hx, double hy, double hz) {
    if (!s || s->lv.empty() || !u || !v || !w || !boundary_sdf_vertex || !fluid_sdf ||
        !(hx > 0) || !(hy > 0) || !(hz > 0))
        return JG_ERR_INVALID_ARG;
    if (s->twod) {
        // the 3-D weight restriction reads fine planes 2k-1..2k+2: a one-plane hierarchy has one
        s->err = "jg_fr_build: the resident hierarchy is a 2-D (one-plane) one";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    for (auto& lv : s->lv) {
        if (lv.fs) continue;
        const Dims& d = lv.d;
        JG_TRY(dalloc(s, &lv.fs, d.n()));
        JG_TRY(dalloc(s, &lv.uw, (size_t)(d.nx + 1) * d.ny * d.nz));
        JG_TRY(dalloc(s, &lv.vw, (size_t)d.nx * (d.ny + 1) * d.nz));
        JG_TRY(dalloc(s, &lv.ww, (size_t)d.nx * d.ny * (d.nz + 1)));
    }
    const Dims& d0 = s->lv[0].d;
    auto faces = [](const Dims& d, int a) {
        return Dims{d.nx + (a == 0), d.ny + (a == 1), d.nz + (a == 2)};
    };
    k_fr_weights<0><<<nblocks_1d(faces(d0, 0).n()), 256, 0, s->stream>>>(boundary_sdf_vertex, d0,
                                                                         s->lv[0].uw);
    k_fr_weights<1><<<nblocks_1d(faces(d0, 1).n()), 256, 0, s->stream>>>(boundary_sdf_vertex, d0,
                                                                         s->lv[0].vw);
    k_fr_weights<2><<<nblocks_1d(faces(d0, 2).n()), 256, 0, s->stream>>>(boundary_sdf_vertex, d0,
                                                                         s->lv[0].ww);
    k_fr_sdf_to_float<<<nblocks_1d(d0.n()), 256, 0, s->stream>>>(fluid_sdf, s->lv[0].fs, d0.n());
    for (size_t l = 1; l < s->lv.size(); ++l) {
        const Dims& fd = s->lv[l - 1].d;
        const Dims& cd = s->lv[l].d;
        k_fr_restrict<<<nblocks_1d(cd.n()), 256, 0, s->stream>>>(s->lv[l - 1].fs, fd, s->lv[l].fs, cd);
        float* fw[3] = {s->lv[l - 1].uw, s->lv[l - 1].vw, s->lv[l - 1].ww};
        float* cw[3] = {s->lv[l].uw, s->lv[l].vw, s->lv[l].ww};
        for (int a = 0; a < 3; ++a)
            k_fr_restrict<<<nblocks_1d(faces(cd, a).n()), 256, 0, s->stream>>>(fw[a], faces(fd, a), cw[a],
                                                                             faces(cd, a));
    }
    H3 h{hx, hy, hz};
    for (size_t l = 0; l < s->lv.size(); ++l) {
        Level& lv = s->lv[l];
        const bool top = l == 0;
        k_fr_build<<<nblocks_1d(lv.d.n()), 256, 0, s->stream>>>(
            lv.fs, lv.uw, lv.vw, lv.ww, top ? u : nullptr, v, w, top ? bu : nullptr,
            top ? bv : nullptr, top ? bw : nullptr, lv.d, h, lv.A64, top ? s->b : nullptr);
        h.x *= 2.0;
        h.y *= 2.0;
        h.z *= 2.0;
    }
    s->f32_valid = false;
    JG_TRY(check_launch(s, "fr_build"));
    return sync(s);
}

int jg_fr_apply_gradient(jg_solver* s, const double* u, const double* v, const double* w,
                         double* uo, double* vo, double* wo, double hx, double hy, double hz) {
    if (!s || s->lv.empty() || !u || !v || !w || !uo || !vo || !wo) return JG_ERR_INVALID_ARG;
    if (!s->lv[0].fs) {
        s->err = "jg_fr_apply_gradient before jg_fr_build";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const Level& l0 = s->lv[0];
    k_fr_gradient<<<nblocks_1d(l0.d.n()), 256, 0, s->stream>>>(l0.fs, l0.uw, l0.vw, l0.ww, s->x, u, v,
                                                                w, uo, vo, wo, l0.d, H3{hx, hy, hz});
    JG_TRY(check_launch(s, "fr_gradient"));
    return sync(s);
}

int jg_fr_build2(jg_solver* s, const double* u, const double* v, const double* boundary_sdf,
                 const double* fluid_sdf, const double* bu, const double* bv, double hx, double hy) {
    if (!s || s->lv.empty() || !u || !v || !boundary_sdf || !fluid_sdf || !(hx > 0) || !(hy > 0))
        return JG_ERR_INVALID_ARG;
    if (s->lv[0].d.nz != 1 || (s->lv.size() > 1 && !s->twod)) {
        s->err = "jg_fr_build2: the resident hierarchy is not a 2-D (one-plane) one";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    for (auto& lv : s->lv) {
        if (lv.fs) continue;
        const Dims& d = lv.d;
        JG_TRY(dalloc(s, &lv.fs, d.n()));
        JG_TRY(dalloc(s, &lv.uw, (size_t)(d.nx + 1) * d.ny));
        JG_TRY(dalloc(s, &lv.vw, (size_t)d.nx * (d.ny + 1)));
        JG_TRY(dalloc(s, &lv.ww, (size_t)d.nx * d.ny * 2));
    }
    const Dims& d0 = s->lv[0].d;
    k_fr_weights2<0><<<nblocks_1d((idx_t)(d0.nx + 1) * d0.ny), 256, 0, s->stream>>>(boundary_sdf, d0,
                                                                                  s->lv[0].uw);
    k_fr_weights2<1><<<nblocks_1d((idx_t)d0.nx * (d0.ny + 1)), 256, 0, s->stream>>>(boundary_sdf, d0,
                                                                                  s->lv[0].vw);
    k_fr_sdf_to_float<<<nblocks_1d(d0.n()), 256, 0, s->stream>>>(fluid_sdf, s->lv[0].fs, d0.n());
    for (size_t l = 1; l < s->lv.size(); ++l) {
        const Dims& fd = s->lv[l - 1].d;
        const Dims& cd = s->lv[l].d;
        k_fr_restrict2<<<nblocks_1d(cd.n()), 256, 0, s->stream>>>(s->lv[l - 1].fs, fd.nx, fd.ny,
                                                                  s->lv[l].fs, cd.nx, cd.ny);
        k_fr_restrict2<<<nblocks_1d((idx_t)(cd.nx + 1) * cd.ny), 256, 0, s->stream>>>(
            s->lv[l - 1].uw, fd.nx + 1, fd.ny, s->lv[l].uw, cd.nx + 1, cd.ny);
        k_fr_restrict2<<<nblocks_1d((idx_t)cd.nx * (cd.ny + 1)), 256, 0, s->stream>>>(
            s->lv[l - 1].vw, fd.nx, fd.ny + 1, s->lv[l].vw, cd.nx, cd.ny + 1);
    }
    double h[2] = {hx, hy};
    for (size_t l = 0; l < s->lv.size(); ++l) {
        Level& lv = s->lv[l];
        const bool top = l == 0;
        k_fr_build2<<<nblocks_1d(lv.d.n()), 256, 0, s->stream>>>(
            lv.fs, lv.uw, lv.vw, top ? u : nullptr, v, top ? bu : nullptr, top ? bv : nullptr, lv.d,
            h[0], h[1], lv.A64, top ? s->b : nullptr);
        h[0] *= 2.0;
        h[1] *= 2.0;
    }
    s->f32_valid = false;
    JG_TRY(check_launch(s, "fr_build2"));
    return sync(s);
}

int jg_fr_apply_gradient2(jg_solver* s, const double* u, const double* v, double* uo, double* vo,
                          double hx, double hy) {
    if (!s || s->lv.empty() || !u || !v || !uo || !vo) return JG_ERR_INVALID_ARG;
    if (!s->lv[0].fs || s->lv[0].d.nz != 1) {
        s->err = "jg_fr_apply_gradient2 before jg_fr_build2";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const Level& l0 = s->lv[0];
    k_fr_gradient2<<<nblocks_1d(l0.d.n()), 256, 0, s->stream>>>(l0.fs, l0.uw, l0.vw, s->x, u, v, uo,
                                                                 vo, l0.d, hx, hy);
    JG_TRY(check_launch(s, "fr_gradient2"));
    return sync(s);
}

int jg_constrain_velocity(jg_solver* s, int blocked, const double h[3], const double o[3],
                          double* u, double* v, double* w, const double* collider_sdf,
                          const jg_rigid_collider* collider, unsigned extrapolation_depth,
                          int closed_walls) {
    if (!s || s->lv.empty() || !h || !o || !u || !v || !w || !collider_sdf ||
        !(h[0] > 0) || !(h[1] > 0) || !(h[2] > 0))
        return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    BcGrid g;
    g.d = s->lv[0].d;
    for (int a = 0; a < 3; ++a) {
        g.h[a] = h[a];
        g.o[a] = o[a];
    }
    BcCollider c{};
    if (collider) {
        for (int a = 0; a < 3; ++a) {
            c.lin[a] = collider->linear_velocity[a];
            c.ang[a] = collider->angular_velocity[a];
            c.tr[a] = collider->translation[a];
        }
        c.friction = collider->friction;
    }
    const Dims& d = g.d;
    const idx_t nf[3] = {(idx_t)(d.nx + 1) * d.ny * d.nz, (idx_t)d.nx * (d.ny + 1) * d.nz,
                         (idx_t)d.nx * d.ny * (d.nz + 1)};
    const idx_t total = nf[0] + nf[1] + nf[2];
    if (!s->bc_tmp) {
        JG_TRY(dalloc(s, &s->bc_tmp, total));
        JG_TRY(dalloc(s, &s->bc_valid, 2 * total));
    }
    double* comp[3] = {u, v, w};
    double* tmp[3] = {s->bc_tmp, s->bc_tmp + nf[0], s->bc_tmp + nf[0] + nf[1]};
    int8_t* va[3] = {s->bc_valid, s->bc_valid + nf[0], s->bc_valid + nf[0] + nf[1]};
    int8_t* vb[3] = {va[0] + total, va[1] + total, va[2] + total};
    hipStream_t st = s->stream;
    auto nb = [&](int a) { return nblocks_1d(nf[a]); };
    // markers and collider velocity on covered faces
    k_bc_init<0><<<nb(0), 256, 0, st>>>(g, collider_sdf, c, u, va[0]);
    k_bc_init<1><<<nb(1), 256, 0, st>>>(g, collider_sdf, c, v, va[1]);
    k_bc_init<2><<<nb(2), 256, 0, st>>>(g, collider_sdf, c, w, va[2]);
    // free slip: extrapolation into the collider, validity ping-pong
    for (unsigned it = 0; it < extrapolation_depth; ++it) {
        int8_t* const* in = (it & 1) ? vb : va;
        int8_t* const* out = (it & 1) ? va : vb;
        k_bc_extrapolate<0><<<nb(0), 256, 0, st>>>(d, u, in[0], out[0]);
        k_bc_extrapolate<1><<<nb(1), 256, 0, st>>>(d, v, in[1], out[1]);
        k_bc_extrapolate<2><<<nb(2), 256, 0, st>>>(d, w, in[2], out[2]);
    }
    // no flux inside the collider (reads all three components), then back
    k_bc_project<0><<<nb(0), 256, 0, st>>>(g, collider_sdf, c, u, v, w, tmp[0]);
    k_bc_project<1><<<nb(1), 256, 0, st>>>(g, collider_sdf, c, u, v, w, tmp[1]);
    k_bc_project<2><<<nb(2), 256, 0, st>>>(g, collider_sdf, c, u, v, w, tmp[2]);
    k_bc_finish<0><<<nb(0), 256, 0, st>>>(d, tmp[0], comp[0], closed_walls);
    k_bc_finish<1><<<nb(1), 256, 0, st>>>(d, tmp[1], comp[1], closed_walls);
    k_bc_finish<2><<<nb(2), 256, 0, st>>>(d, tmp[2], comp[2], closed_walls);
    if (blocked) {
        k_bc_blocked<0><<<nb(0), 256, 0, st>>>(g, collider_sdf, c, u);
        k_bc_blocked<1><<<nb(1), 256, 0, st>>>(g, collider_sdf, c, v);
        k_bc_blocked<2><<<nb(2), 256, 0, st>>>(g, collider_sdf, c, w);
    }
    JG_TRY(check_launch(s, "constrain_velocity"));
    return sync(s);
}

// The compressed build of the finest level; w == nullptr builds a one-plane
// 2-D system (no w term, hz unused).
static int sp_build_compressed(jg_solver* s, const double* u, const double* v, const double* w,
                               const double* boundary_sdf, const double* fluid_sdf, double hx,
                               double hy, double hz, size_t* rows_out, size_t* nnz_out) {
    JG_HIP(hipSetDevice(s->device));
    Level& l0 = s->lv[0];
    const Dims& d = l0.d;
    const idx_t n = d.n();
    if (n >= (idx_t(1) << 31) / 7) {
        s->err = "jg_sp_build_compressed: grid too large for 32-bit CSR indices";
        return JG_ERR_UNSUPPORTED;
    }
    jg_solver::Csr& c = s->csr;
    if (!c.ro) {  // per-cell counts and offsets of the finest level (freed with it)
        JG_TRY(dalloc(s, &c.ro, n));
        JG_TRY(dalloc(s, &c.eo, n));
        JG_TRY(dalloc(s, &c.cr, n));
        JG_TRY(dalloc(s, &c.ce, n));
        JG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, c.scan_bytes, c.cr, c.ro, (int)n, s->stream));
        char* tmp = nullptr;
        JG_TRY(dalloc(s, &tmp, c.scan_bytes));
        c.scan_tmp = tmp;
    }
    const int nb = nblocks_1d(n);
    k_sp_markers<<<nb, 256, 0, s->stream>>>(boundary_sdf, fluid_sdf, l0.markers, n);
    k_csr_count<<<nb, 256, 0, s->stream>>>(l0.markers, d, c.cr, c.ce);
    size_t bytes = c.scan_bytes;
    JG_HIP(hipcub::DeviceScan::ExclusiveSum(c.scan_tmp, bytes, c.cr, c.ro, (int)n, s->stream));
    JG_HIP(hipcub::DeviceScan::ExclusiveSum(c.scan_tmp, bytes, c.ce, c.eo, (int)n, s->stream));
    int last[4];
    JG_HIP(hipMemcpyAsync(&last[0], c.ro + n - 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    JG_HIP(hipMemcpyAsync(&last[1], c.cr + n - 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    JG_HIP(hipMemcpyAsync(&last[2], c.eo + n - 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    JG_HIP(hipMemcpyAsync(&last[3], c.ce + n - 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    JG_TRY(sync(s));
    const size_t rows = (size_t)last[0] + last[1], nnz = (size_t)last[2] + last[3];
    JG_TRY(csr_reserve(s, rows > 0 ? rows : 1, nnz));
    c.rows = rows;
    const int end = (int)nnz;  // the source of the copy below: alive until the sync
    if (rows > 0) {
        k_csr_fill<<<nb, 256, 0, s->stream>>>(l0.markers, u, v, w, d, hx, hy, hz, c.ro, c.eo, c.rp,
                                              c.ci, c.val, c.b);
        JG_HIP(hipMemcpyAsync(c.rp + rows, &end, sizeof(int), hipMemcpyHostToDevice, s->stream));
    }
    JG_TRY(check_launch(s, "sp_build_compressed"));
    c.from_grid = true;
    c.max_row = 7;  // the centre and at most six fluid neighbours
    c.ell_valid = false;
    if (rows_out) *rows_out = rows;
    if (nnz_out) *nnz_out = nnz;
    return sync(s);
}

int jg_sp_build_compressed(jg_solver* s, const double* u, const double* v, const double* w,
                           const double* boundary_sdf, const double* fluid_sdf, double hx,
                           double hy, double hz, size_t* rows_out, size_t* nnz_out) {
    if (!s || s->lv.empty() || !u || !v || !w || !boundary_sdf || !fluid_sdf || !(hx > 0) ||
        !(hy > 0) || !(hz > 0))
        return JG_ERR_INVALID_ARG;
    return sp_build_compressed(s, u, v, w, boundary_sdf, fluid_sdf, hx, hy, hz, rows_out, nnz_out);
}

int jg_sp_build_compressed2(jg_solver* s, const double* u, const double* v,
                            const double* boundary_sdf, const double* fluid_sdf, double hx,
                            double hy, size_t* rows_out, size_t* nnz_out) {
    if (!s || s->lv.empty() || !u || !v || !boundary_sdf || !fluid_sdf || !(hx > 0) || !(hy > 0))
        return JG_ERR_INVALID_ARG;
    if (s->lv[0].d.nz != 1) {
        s->err = "jg_sp_build_compressed2: the finest level is not one plane";
        return JG_ERR_INVALID_ARG;
    }
    return sp_build_compressed(s, u, v, nullptr, boundary_sdf, fluid_sdf, hx, hy, 1.0, rows_out,
                               nnz_out);
}

int jg_diffusion_build(jg_solver* s, const double* f, const double* boundary_sdf,
                       const double* fluid_sdf, double dt, double mu, double hx, double hy,
                       double hz, int dirichlet) {
    if (!s || s->lv.empty() || !f || !boundary_sdf || !fluid_sdf || !(hx > 0) || !(hy > 0) ||
        !(hz > 0))
        return JG_ERR_INVALID_ARG;
    JG_HIP(hipSetDevice(s->device));
    Level& l0 = s->lv[0];
    // c = dt * mu / (h * h), as GridBackwardEulerDiffusionSolver3::solve
    // evaluates it (src/jet/grid_backward_euler_diffusion_solver3.cpp:33)
    const double sc = dt * mu;
    const double cx = sc / (hx * hx), cy = sc / (hy * hy), cz = sc / (hz * hz);
    k_diffusion_build<<<nblocks_1d(l0.d.n()), 256, 0, s->stream>>>(
        f, boundary_sdf, fluid_sdf, l0.d, cx, cy, cz, dirichlet ? 1 : 0, l0.A64, s->b);
    JG_TRY(check_launch(s, "diffusion_build"));
    s->f32_valid = false;
    return sync(s);
}

int jg_csr_decompress(jg_solver* s) {
    if (!s || s->lv.empty() || !s->csr.from_grid || !s->csr.ro) {
        if (s) s->err = "jg_csr_decompress needs a system from jg_sp_build_compressed";
        return JG_ERR_INVALID_ARG;
    }
    JG_HIP(hipSetDevice(s->device));
    const idx_t n = s->lv[0].d.n();
    if (s->csr.rows == 0) {
        JG_HIP(hipMemsetAsync(s->x, 0, (size_t)n * sizeof(double), s->stream));
    } else {
        k_csr_scatter<<<nblocks_1d(n), 256, 0, s->stream>>>(s->lv[0].markers, s->csr.ro, s->csr.x,
                                                            s->x, n);
    }
    JG_TRY(check_launch(s, "csr_decompress"));
    return sync(s);
}

const double* jg_csr_array(jg_solver* s, int which) {
    if (!s) return nullptr;
    switch (which) {
        case 0: return s->csr.val;
        case 1: return s->csr.b;
        case 2: return s->csr.x;
        default: return nullptr;
    }
}

const int* jg_csr_index(jg_solver* s, int which) {
    if (!s) return nullptr;
    return which == 0 ? s->csr.rp : (which == 1 ? s->csr.ci : nullptr);
}

const float* jg_fr_array(jg_solver* s, size_t level, int which) {
    if (!s || level >= s->lv.size()) return nullptr;
    const Level& lv = s->lv[level];
    switch (which) {
        case 0: return lv.uw;
        case 1: return lv.vw;
        case 2: return lv.ww;
        case 3: return lv.fs;
        default: return nullptr;
    }
}

}  // extern "C"
