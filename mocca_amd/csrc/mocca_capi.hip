// C ABI of the mocca_amd library: solve handles, device memory, launches.
// The entry points and their reference counterparts are documented in
// include/mocca_amd.h.
#include "mocca_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

using namespace mocca;

namespace mocca {
namespace detail {

thread_local std::string g_err;

int fail(mocca_solve *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    g_err = buf;
    return code;
}

hipEvent_t take_event(mocca_solve *h) {
    if (!h->pool.empty()) {
        hipEvent_t e = h->pool.back();
        h->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

int drain_profile(mocca_solve *h) {
    if (h->pending.empty()) return 0;
    HIPCHK(h, hipStreamSynchronize(h->st));
    for (auto &p : h->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto &slot = h->acc[p.name];
        slot.first += ms;
        slot.second += 1;
        h->pool.push_back(p.a);
        h->pool.push_back(p.b);
    }
    h->pending.clear();
    return 0;
}

// Lagrange-mesh matrices on a line of n stored points whose full line has
// nfull points (IniLag + reduceLag, Derivatives.f90:899-1078): the
// Baye-Heenen first-derivative coefficients pi/(N dx) (-1)^k / sin(pi k/N)
// and second-derivative coefficients, folded onto the stored half line for
// even (+) and odd (-) reflection.  Output order: first(-), first(+),
// second(-), second(+), each n x n row-major.
void lagrange_axis(int n, int nfull, double dx, double *out) {
    const double pi = 4.0 * std::atan2(1.0, 1.0);
    std::vector<double> c(2 * nfull + 2, 0.0), l(2 * nfull + 2, 0.0);
    const double f = pi / nfull, df = f / dx;
    const double diag = -pi * pi / (3.0 * dx * dx) * (1.0 - 1.0 / double(nfull * nfull));
    for (int k = 1; k < nfull; ++k) {
        const double s = std::sin(f * k), co = std::cos(f * k);
        const double sg = (k & 1) ? -1.0 : 1.0;
        c[k] = df * sg / s;
        l[k] = -sg * 2.0 * df * df * co / (s * s);
    }
    double *f1m = out, *f1p = out + n * n, *f2m = out + 2 * n * n, *f2p = out + 3 * n * n;
    for (int i = 1; i <= n; ++i)
        for (int j = 1; j <= n; ++j) {
            const int r = (i - 1) * n + (j - 1);
            const int a = std::abs(i - j), e = i + j - 1;
            if (i == j) {
                f1m[r] = -c[e];
                f1p[r] = c[e];
                f2m[r] = diag - l[e];
                f2p[r] = diag + l[e];
            } else {
                const double ca = (j < i) ? c[a] : -c[a];
                f1p[r] = ca + c[e];
                f1m[r] = ca - c[e];
                f2p[r] = l[a] + l[e];
                f2m[r] = l[a] - l[e];
            }
        }
}

template <typename T>
void fill_mesh(mocca_solve *h, MeshK<T> &m, const T *lag) {
    const mocca_mesh &g = h->mesh;
    m.nx = g.nx; m.ny = g.ny; m.nz = g.nz; m.N = h->N; m.lag = g.lagrange;
    m.inv60dx = T(1.0 / (60.0 * g.dx));
    const double xf = -8064.0;
    double xd = 5040.0 / xf * (g.dx * g.dx);
    xd = 1.0 / xd;
    m.l2m = T(43050.0 / (xf * 3));
    m.l2c2 = T(1008.0 / xf);
    m.l2c3 = T(-128.0 / xf);
    m.l2c4 = T(9.0 / xf);
    m.l2scale = T(-xd);
    const int ns[3] = {g.nx, g.ny, g.nz};
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        const size_t nn = (size_t)ns[a] * ns[a];
        m.lag1[a][0] = lag + off;
        m.lag1[a][1] = lag + off + nn;
        m.lag2[a][0] = lag + off + 2 * nn;
        m.lag2[a][1] = lag + off + 3 * nn;
        off += 4 * nn;
    }
}

int sign_first_h(int axis, int P, int S, int TS, int c, int lag) {
    const int cx = (c == 0 || c == 3) ? 1 : -1;
    const int cy = (c == 0 || c == 2) ? 1 : -1;
    const int cz = (c < 2) ? 1 : -1;
    if (axis == 0) return cx * (lag ? S * TS : S);
    if (axis == 1) return cy * TS;
    return cz * S * P;
}

int sign_lap_h(int axis, int P, int S, int TS, int c) {
    const int cx = (c == 0 || c == 3) ? 1 : -1;
    const int cy = (c == 0 || c == 2) ? 1 : -1;
    const int cz = (c < 2) ? 1 : -1;
    if (axis == 0) return cx * S;
    if (axis == 1) return cy * TS;
    return cz * S * P;
}

// Jobs for the derived density fields (Densities.f90:541-550,1003-1101).
// Field "quantum numbers" (P, S, TS, component) follow the reference calls.
void build_density_jobs(mocca_solve *h, std::vector<FieldJob> &jobs) {
    const long N = h->N;
    const int lag = h->mesh.lagrange, trc = h->mesh.time_reversal;
    const int nf = 9;
    double *dens = h->b.dens;
    double *derrho = h->b.derived, *laprho = derrho + 6 * N, *rotj = laprho + 2 * N,
           *rots = rotj + 6 * N, *ders = rots + 6 * N;
    const int P = 1, S = 1, TS = 1;
    auto first = [&](FieldJob &j, int t, const double *src, int axis, int p, int s,
                     int ts, int comp, double coef) {
        j.src[t] = src;
        j.axis[t] = axis;
        j.sign[t][0] = j.sign[t][1] = j.sign[t][2] = sign_first_h(axis, p, s, ts, comp - 1, lag);
        j.coef[t] = coef;
        j.nterms = t + 1;
    };
    for (int it = 0; it < 2; ++it) {
        const double *rho = dens + ((long)it * nf + 0) * N;
        for (int a = 0; a < 3; ++a) {
            FieldJob j{};
            first(j, 0, rho, a, P, S, TS, 1, 1.0);
            j.dst = derrho + (it * 3 + a) * N;
            jobs.push_back(j);
        }
        FieldJob j{};
        j.src[0] = rho;
        j.axis[0] = 3;
        for (int a = 0; a < 3; ++a) j.sign[0][a] = sign_lap_h(a, P, S, TS, 0);
        j.coef[0] = 1.0;
        j.nterms = 1;
        j.dst = laprho + it * N;
        jobs.push_back(j);
        if (trc) continue;
        const double *jx = dens + ((long)it * nf + 3) * N, *jy = jx + N, *jz = jy + N;
        const double *sx = dens + ((long)it * nf + 6) * N, *sy = sx + N, *sz = sy + N;
        // rot j: (Densities.f90:1008-1025)
        FieldJob r{};
        first(r, 0, jz, 1, -P, S, TS, 2, 1.0);
        first(r, 1, jy, 2, -P, -S, TS, 1, -1.0);
        r.dst = rotj + (it * 3 + 0) * N;
        jobs.push_back(r);
        r = FieldJob{};
        first(r, 0, jx, 2, -P, -S, TS, 2, 1.0);
        first(r, 1, jz, 0, -P, S, TS, 2, -1.0);
        r.dst = rotj + (it * 3 + 1) * N;
        jobs.push_back(r);
        r = FieldJob{};
        first(r, 0, jy, 0, -P, -S, TS, 1, 1.0);
        first(r, 1, jx, 1, -P, -S, TS, 2, -1.0);
        r.dst = rotj + (it * 3 + 2) * N;
        jobs.push_back(r);
        // DerS and rot s (Densities.f90:1058-1101)
        const double *sc[3] = {sx, sy, sz};
        const int sS[3] = {-S, -S, S}, sC[3] = {1, 2, 1};
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) {
                FieldJob d{};
                first(d, 0, sc[c], a, P, sS[c], TS, sC[c], 1.0);
                d.dst = ders + ((it * 3 + c) * 3 + a) * N;
                jobs.push_back(d);
            }
        // rot s_x = d_y s_z - d_z s_y ; rot s_y = d_z s_x - d_x s_z ;
        // rot s_z = d_x s_y - d_y s_x
        const int rc[3][4] = {{2, 1, 1, 2}, {0, 2, 2, 0}, {1, 0, 0, 1}};
        for (int q = 0; q < 3; ++q) {
            FieldJob d{};
            first(d, 0, sc[rc[q][0]], rc[q][1], P, sS[rc[q][0]], TS, sC[rc[q][0]], 1.0);
            first(d, 1, sc[rc[q][2]], rc[q][3], P, sS[rc[q][2]], TS, sC[rc[q][2]], -1.0);
            d.dst = rots + (it * 3 + q) * N;
            jobs.push_back(d);
        }
    }
}

template <typename T>
const MeshK<T> &mesh_of(mocca_solve *h);
template <>
const MeshK<double> &mesh_of<double>(mocca_solve *h) { return h->md; }
template <>
const MeshK<float> &mesh_of<float>(mocca_solve *h) { return h->mf; }

int check_ready(mocca_solve *h) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    if (!h->meta_set) return fail(h, MOCCA_ERR_STATE, "mocca_set_states_meta not called");
    return 0;
}

int check_status(mocca_solve *h) {
    int s = 0;
    HIPCHK(h, hipMemcpyAsync(&s, h->b.status, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (s) {
        HIPCHK(h, hipMemsetAsync(h->b.status, 0, sizeof(int), h->st));
        return fail(h, MOCCA_ERR_NUMERIC,
                    "Gram-Schmidt met a non-positive norm (linearly dependent or NaN states)");
    }
    return 0;
}

template <typename T>
int gram_schmidt_impl(mocca_solve *h, const T *src, T *dst, T *upd, int update_mode) {
    const long K4 = 4L * h->N;
    int rc;
    if (h->mfma_gram) {
        rc = run(h, "gram", [&] {
            launch_gram_mfma<T>(src, K4, h->b.blk_states, h->b.blk_m, h->nblk, h->max_m,
                                h->mpad, h->b.gpart, h->nsplit_mfma, h->mesh.dv, h->b.G,
                                h->st);
        });
        if (rc) return rc;
    } else {
        rc = run(h, "gram", [&] {
            launch_gram<T>(src, K4, h->b.blk_states, h->b.blk_m, h->nblk, h->max_m, h->mmax,
                           h->gtile, h->b.gpart, h->nsplit, h->st);
        });
        if (rc) return rc;
        rc = run(h, "gram_reduce", [&] {
            launch_gram_reduce(h->b.gpart, h->b.blk_m, h->nblk, h->max_m, h->mmax, h->gtile,
                               h->nsplit, h->mesh.dv, h->b.G, h->st);
        });
        if (rc) return rc;
    }
    rc = run(h, "cholesky", [&] {
        launch_cholesky_ld(h->b.G, h->b.blk_states, h->b.blk_m, h->nblk, h->max_m, h->mmax,
                           h->b.E, h->b.Tm, h->b.perm, h->b.work, h->b.status,
                           h->chol_lds, h->st);
        launch_gs_norms(h->b.Tm, h->b.perm, h->b.blk_m, h->nblk, h->max_m, h->b.nrm, h->st);
    });
    if (rc) return rc;
    if (h->mfma_apply && !update_mode)
        return run(h, "gs_apply", [&] {
            launch_gs_apply_mfma<T>(src, dst, K4, h->b.perm, h->b.blk_m, h->nblk, h->max_m,
                                    h->mpad, h->b.Tm, h->st);
        });
    return run(h, "gs_apply", [&] {
        launch_gs_apply<T>(src, dst, upd, K4, h->b.perm, h->b.blk_m, h->nblk, h->max_m,
                           h->b.Tm, update_mode, h->st);
    });
}


// The matrix-core Lagrange derivatives of every state into the scratch:
// FP64 matrix cores, or for FP32 states on request the split FP32 ones.
template <typename T>
void lag_derivs(mocca_solve *h, const T *psi, T *dsc) {
    if constexpr (sizeof(T) == 4) {
        if (h->lag_split) {
            launch_lag_derivs_split(h->md, psi, h->b.qn, h->M, dsc, h->st);
            return;
        }
    }
    launch_lag_derivs<T>(h->md, psi, h->b.qn, h->M, dsc, h->st);
}

template <typename T>
int grad_desc_impl(mocca_solve *h, double dth, double mu) {
    const MeshK<T> &m = mesh_of<T>(h);
    T *psi = (T *)h->b.psi, *psi2 = (T *)h->b.psi2, *upd = (T *)h->b.upd;
    // After an AlternateStep the stencils read the states before it and the
    // values come from the current ones (the reference keeps the derivatives
    // of DeriveAll until the next one, ImaginaryTime.f90:306-437).
    const bool stale = h->scf.alt_pending;
    const T *dsrc = stale ? (const T *)h->scf.alt_old : psi;
    const T *vsrc = stale ? psi : nullptr;
    h->scf.alt_pending = false;
    // no previous states yet (start, upload, reset): u = 0
    if (!h->prev_valid) {
        HIPCHK(h, hipMemcpyAsync(upd, dsrc, (size_t)h->M * 4 * h->N * sizeof(T),
                                 hipMemcpyDeviceToDevice, h->st));
        h->prev_valid = true;
    }
    int rc, slots;
    if constexpr (sizeof(T) == 8) {
        if (h->march && h->dma_hpsi && !vsrc) {
            slots = h->ntile_y * h->nzs;
            rc = run(h, "hpsi_step", [&] {
                launch_hpsi_fd_dma(m, dsrc, psi2, upd, (const T *)h->b.fields, h->b.qn,
                                   h->b.tasks_dma, h->ntasks_dma, h->TY, h->ntile_y, h->nzs,
                                   h->mesh.time_reversal, h->kin[0], h->kin[1], dth, mu,
                                   h->b.eparts, h->st, h->w_antisym);
            });
            goto stepped;
        }
    }
    if (h->march) {
        slots = h->ntile_y;
        rc = run(h, "hpsi_step", [&] {
            launch_hpsi_fd<T>(m, dsrc, psi2, upd, nullptr, (const T *)h->b.fields, h->b.qn,
                              h->b.tasks, h->ntasks, h->TY, h->ntile_y,
                              h->mesh.time_reversal, T(h->kin[0]), T(h->kin[1]), T(dth),
                              T(mu), kStep, h->b.eparts, h->st, vsrc, h->w_antisym);
        });
    } else {
        slots = h->ntiles;
        T *dsc = (T *)h->b.lagscr;
        if (dsc) {
            rc = run(h, "lag_derivs", [&] { lag_derivs<T>(h, dsrc, dsc); });
            if (rc) return rc;
        }
        rc = run(h, "hpsi_step", [&] {
            launch_hpsi_step<T>(m, dsrc, psi2, upd, nullptr, (const T *)h->b.fields, h->b.qn,
                                h->b.order, h->b.chunks, h->nchunks, h->mesh.time_reversal,
                                T(h->kin[0]), T(h->kin[1]), T(dth), T(mu), kStep,
                                h->b.eparts, h->ntiles, h->st, vsrc, dsc);
        });
    }
stepped:
    if (rc) return rc;
    rc = run(h, "reduce_ed", [&] {
        launch_reduce_ed(h->b.eparts, slots, h->b.all_states, h->M, h->mesh.dv,
                         h->b.E, h->b.D, h->b.K, h->st);
    });
    if (rc) return rc;
    // Gram-Schmidt of the stepped states psi2 into the buffer of the previous
    // states, which then become the current ones; the states before the
    // step are the new previous ones (u = psi_new - psi_old next time).
    if ((rc = gram_schmidt_impl<T>(h, psi2, upd, nullptr, 0))) return rc;
    std::swap(h->b.psi, h->b.upd);
    return 0;
}

// hPsi of every state without outputs other than the per-state sums; only
// K = dv <psi|Lap psi> is kept (CompKinetic's inner products).
template <typename T>
int kinetic_impl(mocca_solve *h) {
    const MeshK<T> &m = mesh_of<T>(h);
    // the density basis: the canonical states under HFB
    const T *psi = (const T *)(h->hfb.on ? h->hfb.can_psi : h->b.psi);
    int rc, slots;
    if constexpr (sizeof(T) == 8) {
        if (h->march && h->dma_hpsi) {
            slots = h->ntile_y * h->nzs;
            rc = run(h, "kinetic", [&] {
                launch_hpsi_kinetic_dma(m, psi, h->b.qn, h->b.tasks_dma, h->ntasks_dma, h->TY,
                                        h->ntile_y, h->nzs, h->mesh.time_reversal,
                                        h->b.eparts, h->st);
            });
            if (rc) return rc;
            return run(h, "reduce_ed", [&] {
                launch_reduce_ed(h->b.eparts, slots, h->b.all_states, h->M, h->mesh.dv,
                                 nullptr, nullptr, h->b.K, h->st);
            });
        }
    }
    if (h->march) {
        slots = h->ntile_y;
        rc = run(h, "kinetic", [&] {
            launch_hpsi_fd<T>(m, psi, nullptr, nullptr, nullptr,
                              (const T *)h->b.fields, h->b.qn, h->b.tasks, h->ntasks, h->TY,
                              h->ntile_y, h->mesh.time_reversal, T(h->kin[0]), T(h->kin[1]),
                              T(0), T(0), kKinetic, h->b.eparts, h->st);
        });
    } else {
        slots = h->ntiles;
        T *dsc = (T *)h->b.lagscr;
        if (dsc) {
            rc = run(h, "lag_derivs", [&] { lag_derivs<T>(h, psi, dsc); });
            if (rc) return rc;
        }
        rc = run(h, "kinetic", [&] {
            launch_hpsi_step<T>(m, psi, nullptr, nullptr, nullptr,
                                (const T *)h->b.fields, h->b.qn, h->b.order, h->b.chunks,
                                h->nchunks, h->mesh.time_reversal, T(h->kin[0]),
                                T(h->kin[1]), T(0), T(0), 0, h->b.eparts, h->ntiles, h->st,
                                nullptr, dsc);
        });
    }
    if (rc) return rc;
    return run(h, "reduce_ed", [&] {
        launch_reduce_ed(h->b.eparts, slots, h->b.all_states, h->M, h->mesh.dv, nullptr,
                         nullptr, h->b.K, h->st);
    });
}

int up(mocca_solve *h, void *dst, const double *src, size_t n);

// hPsi of the estimation state into est.h, with its partial sums.
template <typename T>
int est_hpsi(mocca_solve *h) {
    const MeshK<T> &m = mesh_of<T>(h);
    Est &e = h->est;
    if (h->march)
        return run(h, "est_hpsi", [&] {
            launch_hpsi_fd<T>(m, (const T *)e.psi, nullptr, nullptr, (T *)e.h,
                              (const T *)h->b.fields, e.qn, e.tasks, h->ntile_y, h->TY,
                              h->ntile_y, h->mesh.time_reversal, T(h->kin[0]), T(h->kin[1]),
                              T(0), T(0), kWriteH, e.eparts, h->st);
        });
    return run(h, "est_hpsi", [&] {
        launch_hpsi_step<T>(m, (const T *)e.psi, nullptr, nullptr, (T *)e.h,
                            (const T *)h->b.fields, e.qn, e.order, e.chunk, 1,
                            h->mesh.time_reversal, T(h->kin[0]), T(h->kin[1]), T(0), T(0),
                            kWriteH, e.eparts, h->ntiles, h->st);
    });
}

int est_alloc(mocca_solve *h) {
    Est &e = h->est;
    const size_t n4 = 4 * (size_t)h->N, slots = (size_t)std::max(h->ntile_y, h->ntiles);
    e.nblk = (int)std::min<size_t>(512, (n4 + kBlock - 1) / kBlock);
    int rc;
    if ((rc = dalloc(h, &e.psi, n4 * h->tsize)) || (rc = dalloc(h, &e.h, n4 * h->tsize)) ||
        (rc = dalloc(h, &e.u, n4 * h->tsize)) ||
        (rc = dalloc(h, &e.eparts, 3 * slots * sizeof(double))) ||
        (rc = dalloc(h, &e.npart, e.nblk * sizeof(double))) ||
        (rc = dalloc(h, &e.rec, kEstSlots * sizeof(double))) ||
        (rc = dalloc(h, &e.qn, 4 * sizeof(int))) ||
        (rc = dalloc(h, &e.tasks, std::max(1, h->ntile_y) * sizeof(Task))) ||
        (rc = dalloc(h, &e.order, sizeof(int))) || (rc = dalloc(h, &e.chunk, sizeof(Chunk))))
        return rc;
    return 0;
}

// IterativeEstimation (ImaginaryTime.f90:28-211) on the device.
template <typename T>
int estimate_impl(mocca_solve *h, int iteration, double dt, double mu, double *out) {
    Est &e = h->est;
    const long n4 = 4L * h->N;
    int rc;
    if (!e.psi && (rc = est_alloc(h))) return rc;
    const bool init = iteration == 1 || !e.ready;
    int fill = 0;
    if (init) {
        // maxspwf copies the quantum numbers of state nwt, or of state nwn
        // when there are no protons (ImaginaryTime.f90:58-63).
        int s = h->M - 1;
        if (h->scf.set && h->scf.sk.protons < 0.000000001) {
            int nwn = 0;
            for (int i = 0; i < h->M; ++i) nwn += h->qn[4 * i] == -1;
            if (nwn == 0) return fail(h, MOCCA_ERR_STATE, "no neutron state to start from");
            s = nwn - 1;
        }
        HIPCHK(h, hipMemcpy(e.qn, &h->qn[4 * s], 4 * sizeof(int), hipMemcpyHostToDevice));
        const int it = (h->qn[4 * s] + 3) / 2 - 1;
        std::vector<Task> tk(std::max(1, h->ntile_y));
        for (int t = 0; t < (int)tk.size(); ++t) tk[t] = Task{0, t};
        const int zero = 0;
        const Chunk c{0, 1, it, 0};
        HIPCHK(h, hipMemcpy(e.tasks, tk.data(), tk.size() * sizeof(Task), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(e.order, &zero, sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(e.chunk, &c, sizeof c, hipMemcpyHostToDevice));
        if (!e.seed.empty()) {
            if ((rc = up(h, e.psi, e.seed.data(), (size_t)n4))) return rc;
            e.seed.clear();
        } else {
            fill = 1;
        }
    }
    const unsigned long long seed = 0x243F6A8885A308D3ull + e.calls++;
    rc = run(h, "est_reset", [&] {
        launch_est_reset<T>((T *)e.psi, (T *)e.u, n4, fill, seed, e.rec, e.nblk, h->st);
    });
    if (rc) return rc;
    if (init) {
        rc = run(h, "est_start", [&] {
            launch_est_start<T>((T *)e.psi, n4, e.npart, e.nblk, h->mesh.dv, e.rec, h->st);
        });
        if (rc) return rc;
        e.ready = true;
    }
    // Power iterations, enqueued in growing batches; the device record says
    // where |con| < 1e-2 was met.
    constexpr int kMaxIter = 500;
    const double dth = dt / kHbar, tol = 1e-2;
    const int slots = h->march ? h->ntile_y : h->ntiles;
    double rec[kEstSlots];
    int k = 1, batch = 4;
    while (k <= kMaxIter) {
        const int nb = std::min(batch, kMaxIter + 1 - k);
        for (int j = 0; j < nb; ++j) {
            if ((rc = est_hpsi<T>(h))) return rc;
            const int iter = k + j;
            rc = run(h, "est_step", [&] {
                launch_est_iteration<T>((T *)e.psi, (const T *)e.h, (T *)e.u, n4, e.eparts,
                                        slots, h->mesh.dv, dth, mu, iter, iter == 1, tol, e.rec,
                                        e.npart, e.nblk, h->st);
            });
            if (rc) return rc;
        }
        HIPCHK(h, hipMemcpyAsync(rec, e.rec, sizeof rec, hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
        k += nb;
        if (rec[3] != 0.0) break;
        batch = std::min(2 * batch, 64);
    }
    if (h->hfb.on && (rc = hfb_ensure_qpe(h))) return rc;
    rc = run(h, "est_finalize", [&] {
        const bool pair = h->hfb.on || h->hfb.bcs;   // (BCS: qpe holds Eqp of every state)
        launch_est_finalize(h->b.E, h->b.occ, h->b.qn, h->M, pair ? 2 : 0, h->hfb.qpe,
                            pair ? h->M : 0, kHbar, e.rec, h->st);
    });
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(rec, e.rec, sizeof rec, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (!std::isfinite(rec[0]))
        return fail(h, MOCCA_ERR_NUMERIC, "IterativeEstimation: something went wrong (maxE = %g)",
                    rec[0]);
    out[0] = rec[4];
    out[1] = rec[5];
    out[2] = rec[0];
    out[3] = rec[6];
    out[4] = rec[1];
    out[5] = rec[2];
    return 0;
}

int estimate_any(mocca_solve *h, int iteration, double dt, double mu, double *out) {
    return h->prec == MOCCA_FP64 ? estimate_impl<double>(h, iteration, dt, mu, out)
                                 : estimate_impl<float>(h, iteration, dt, mu, out);
}

// AlternateStep (ImaginaryTime.f90:306-437): psi <- psi - multipole psi,
// then Gramschmidt; the states before the step stay in alt_old for the
// derivatives of the next GradDesc.
int alternate_prepare(mocca_solve *h) {
    Scf &s = h->scf;
    if (s.alt_old) return 0;
    return dalloc(h, &s.alt_old, (size_t)h->M * 4 * h->N * h->tsize);
}

template <typename T>
int alternate_impl(mocca_solve *h) {
    Scf &s = h->scf;
    const long N = h->N;
    if (int rc = alternate_prepare(h)) return rc;
    int rc = run(h, "alternate", [&] {
        launch_alternate<T>(s.con, s.ncon, s.Y, s.calt, s.cut, N, s.mult, (const T *)h->b.psi,
                            (T *)h->b.psi2, h->b.qn, h->M, h->st);
    });
    if (rc) return rc;
    if ((rc = gram_schmidt_impl<T>(h, (const T *)h->b.psi2, (T *)s.alt_old, nullptr, 0)))
        return rc;
    std::swap(h->b.psi, s.alt_old);
    s.alt_pending = true;
    return 0;
}

int alternate_any(mocca_solve *h) {
    return h->prec == MOCCA_FP64 ? alternate_impl<double>(h) : alternate_impl<float>(h);
}

int grad_desc_any(mocca_solve *h, double dth, double mu) {
    return h->prec == MOCCA_FP64 ? grad_desc_impl<double>(h, dth, mu)
                                 : grad_desc_impl<float>(h, dth, mu);
}

int kinetic_any(mocca_solve *h) {
    return h->prec == MOCCA_FP64 ? kinetic_impl<double>(h) : kinetic_impl<float>(h);
}

template <typename T>
int hpsi_one_impl(mocca_solve *h, int state) {
    const MeshK<T> &m = mesh_of<T>(h);
    if (h->march) {
        std::vector<Task> tk(h->ntile_y);
        for (int t = 0; t < h->ntile_y; ++t) tk[t] = Task{state, t};
        HIPCHK(h, hipMemcpy(h->b.one_tasks, tk.data(), tk.size() * sizeof(Task),
                            hipMemcpyHostToDevice));
        return run(h, "hpsi", [&] {
            launch_hpsi_fd<T>(m, (const T *)h->b.psi, nullptr, nullptr, (T *)h->b.hbuf,
                              (const T *)h->b.fields, h->b.qn, h->b.one_tasks, h->ntile_y,
                              h->TY, h->ntile_y, h->mesh.time_reversal, T(h->kin[0]),
                              T(h->kin[1]), T(0), T(0), kWriteH, nullptr, h->st);
        });
    }
    Chunk c{0, 1, (h->qn[4 * state] + 3) / 2 - 1, 0};
    HIPCHK(h, hipMemcpy(h->b.one_chunk, &c, sizeof c, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->b.one_order, &state, sizeof(int), hipMemcpyHostToDevice));
    return run(h, "hpsi", [&] {
        launch_hpsi_step<T>(m, (const T *)h->b.psi, nullptr, nullptr, (T *)h->b.hbuf,
                            (const T *)h->b.fields, h->b.qn, h->b.one_order,
                            h->b.one_chunk, 1, h->mesh.time_reversal, T(h->kin[0]),
                            T(h->kin[1]), T(0), T(0), kWriteH, nullptr, h->ntiles, h->st);
    });
}

template <typename T>
int density_impl(mocca_solve *h) {
    const MeshK<T> &m = mesh_of<T>(h);
    // the density basis (DensityBasis): the canonical states under HFB
    const T *psi = (const T *)(h->hfb.on ? h->hfb.can_psi : h->b.psi);
    const double *occ = h->hfb.on ? h->hfb.can_occ : h->b.occ;
    const int nf = h->mesh.time_reversal ? 3 : 9;
    int rc;
    if (h->march) {
        rc = run(h, "density", [&] {
            launch_density_fd<T>(m, psi, h->b.qn, occ, h->b.order,
                                 h->b.dchunks, h->b.dtasks, h->ndtasks, h->TY,
                                 h->mesh.time_reversal, (T *)h->b.dpart, h->st);
        });
        if (rc) return rc;
        rc = run(h, "density_reduce", [&] {
            launch_density_reduce<T>((const T *)h->b.dpart, h->b.dchunks, h->ndchunks, nf,
                                     h->N, h->b.dens, h->st);
        });
    } else {
        rc = run(h, "density", [&] {
            launch_density<T>(m, psi, h->b.qn, occ, h->b.order,
                              h->b.chunks, h->nchunks, h->mesh.time_reversal,
                              (T *)h->b.dpart, h->st);
        });
        if (rc) return rc;
        rc = run(h, "density_reduce", [&] {
            launch_density_reduce<T>((const T *)h->b.dpart, h->b.chunks, h->nchunks, nf,
                                     h->N, h->b.dens, h->st);
        });
    }
    if (rc) return rc;
    return run(h, "field_jobs", [&] {
        launch_field_jobs(h->md, h->b.jobs, h->njobs, h->st);
    });
}

int density_any(mocca_solve *h) {
    return h->prec == MOCCA_FP64 ? density_impl<double>(h) : density_impl<float>(h);
}

template <typename T>
int upload_convert(mocca_solve *h, T *dst, const double *src, size_t n) {
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(h, hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    }
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
    HIPCHK(h, hipMemcpy(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int download_convert(mocca_solve *h, double *dst, const T *src, size_t n) {
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(h, hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<T> tmp(n);
    HIPCHK(h, hipMemcpy(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) dst[i] = (double)tmp[i];
    return 0;
}

int up(mocca_solve *h, void *dst, const double *src, size_t n) {
    return h->prec == MOCCA_FP64 ? upload_convert(h, (double *)dst, src, n)
                                 : upload_convert(h, (float *)dst, src, n);
}

int down(mocca_solve *h, double *dst, const void *src, size_t n) {
    return h->prec == MOCCA_FP64 ? download_convert(h, dst, (const double *)src, n)
                                 : download_convert(h, dst, (const float *)src, n);
}

int set_scheme(mocca_solve *h, int lagrange) {
    lagrange = lagrange ? 1 : 0;
    if (h->mesh.lagrange == lagrange) return 0;
    if (!lagrange && (h->mesh.nx < 6 || h->mesh.ny < 6 || h->mesh.nz < 6))
        return fail(h, MOCCA_ERR_ARG, "EV8 finite differences need at least 6 points per axis");
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->mesh.lagrange = lagrange;
    Buffers &b = h->b;
    fill_mesh<double>(h, h->md, b.lagD);
    fill_mesh<float>(h, h->mf, (const float *)b.lagT);
    if (h->prec == MOCCA_FP64) fill_mesh<double>(h, h->md, (const double *)b.lagT);
    // the z-marching kernels are finite-difference only; a mesh created with
    // Lagrange derivatives has no z-marching task lists
    h->march = !lagrange && h->mesh.nx <= 64 && h->ntasks > 0;
    h->dma_hpsi = h->march && h->prec == MOCCA_FP64 && h->ntasks_dma > 0 &&
                  hpsi_dma_ok(h->md, h->TY, h->mesh.time_reversal);
    std::vector<FieldJob> jobs;
    build_density_jobs(h, jobs);
    if ((int)jobs.size() != h->njobs)
        return fail(h, MOCCA_ERR_STATE, "derived-density jobs changed in number");
    HIPCHK(h, hipMemcpy(b.jobs, jobs.data(), jobs.size() * sizeof(FieldJob),
                        hipMemcpyHostToDevice));
    return 0;
}

void free_all(mocca_solve *h) {
    Buffers &b = h->b;
    void *ptrs[] = {b.psi, b.psi2, b.upd, b.hbuf, b.fields, b.dpart, b.lagT, b.lagD,
                    b.der, b.lap, b.dens, b.derived, b.qn, b.occ, b.E, b.D, b.nrm, b.K, b.eparts,
                    b.order, b.chunks, b.all_states, b.blk_states, b.blk_m, b.perm,
                    b.G, b.Tm, b.work, b.gpart, b.status, b.jobs, b.one_chunk,
                    b.one_order, b.tasks, b.one_tasks, b.dchunks, b.dtasks, b.tasks_dma,
                    b.lagscr};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    Est &e = h->est;
    void *eptrs[] = {e.psi, e.h, e.u, e.eparts, e.npart, e.rec, e.qn, e.tasks, e.order, e.chunk};
    for (void *p : eptrs)
        if (p) (void)hipFree(p);
    scf_free(h);
    hfb_free(h);
    for (auto &p : h->pending) { h->pool.push_back(p.a); h->pool.push_back(p.b); }
    for (auto e : h->pool) (void)hipEventDestroy(e);
    if (h->st) (void)hipStreamDestroy(h->st);
}

}  // namespace detail
}  // namespace mocca

namespace mocca {
namespace detail {

GsPlan gs_plan(int nblk, int mmax, long K4, int M, size_t tsize) {
    GsPlan p;
    p.nblk = nblk;
    p.mmax = mmax;
    p.max_m = (mmax + kGramTile - 1) / kGramTile * kGramTile;
    p.gtile = mmax <= 32 ? 32 : 64;
    // Tiles are counted from the largest block (the leading dimension max_m
    // may be larger); each block gets enough K splits to fill the GPU.
    const int nt = (mmax + p.gtile - 1) / p.gtile;
    p.ntile_max = nt * (nt + 1) / 2;
    int nsplit = std::max(1, 2048 / std::max(1, nblk * p.ntile_max));
    nsplit = (int)std::min<long>(nsplit, std::max<long>(1, K4 / (kGramKC * 8)));
    p.nsplit = nsplit;
    // matrix-core Gram for blocks of up to kMfmaMaxRows states
    p.mpad = (mmax + 15) / 16 * 16;
    p.mfma_gram = p.mpad <= kMfmaMaxRows;
    // (the apply keeps one output row of tiles per wave: up to 192 states)
    p.mfma_apply = p.mpad <= kMfmaApplyMaxRows;
    const int nt16 = p.mpad / 16;
    p.ntile16 = nt16 * (nt16 + 1) / 2;
    // one round of single-wave workgroups (1024 = 256 CUs x 4 SIMDs at one
    // wave per SIMD; four per SIMD for blocks of up to 64 states, whose
    // waves hold few registers), partial tiles under 15% of the states'
    // bytes, at least 4 chunks of 16 values per wave
    const double data = (double)M * K4 * tsize;
    long ns = std::max(1, (nt16 <= 4 ? 4096 : 1024) / std::max(1, nblk));
    ns = std::min<long>(ns, (long)(0.15 * data / ((double)std::max(1, nblk) * p.ntile16 * 256 * 8)));
    ns = std::min<long>(ns, std::max<long>(1, K4 / 64));
    p.nsplit_mfma = (int)std::max<long>(1, ns);
    return p;
}

}  // namespace detail
}  // namespace mocca

using namespace mocca::detail;

// ===========================================================================
extern "C" {

const char *mocca_last_error(const mocca_solve *h) {
    return h ? h->err.c_str() : g_err.c_str();
}

int mocca_solve_create(const mocca_mesh *mesh, int nwt, int precision,
                       mocca_solve **out) {
    if (!mesh || !out) return fail(nullptr, MOCCA_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->nx < 1 || mesh->ny < 1 || mesh->nz < 1 || !(mesh->dx > 0) || !(mesh->dv > 0))
        return fail(nullptr, MOCCA_ERR_ARG, "invalid mesh (nx=%d ny=%d nz=%d dx=%g dv=%g)",
                    mesh->nx, mesh->ny, mesh->nz, mesh->dx, mesh->dv);
    if (!mesh->lagrange && (mesh->nx < 6 || mesh->ny < 6 || mesh->nz < 6))
        return fail(nullptr, MOCCA_ERR_ARG,
                    "EV8 finite differences need at least 6 points per axis");
    if (nwt < 1) return fail(nullptr, MOCCA_ERR_ARG, "nwt must be positive (got %d)", nwt);
    if (precision != MOCCA_FP64 && precision != MOCCA_MIXED)
        return fail(nullptr, MOCCA_ERR_ARG, "unknown precision %d", precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, MOCCA_ERR_HIP, "no HIP device available");
    mocca_solve *h = new mocca_solve;
    h->mesh = *mesh;
    h->M = nwt;
    h->N = mesh->nx * mesh->ny * mesh->nz;
    h->prec = precision;
    h->tsize = precision == MOCCA_FP64 ? 8 : 4;
    int rc = 0;
#define TRY(x)                     \
    do {                           \
        if ((rc = (x)) != 0) {     \
            g_err = h->err;        \
            free_all(h);           \
            delete h;              \
            return rc;             \
        }                          \
    } while (0)
    if (hipStreamCreate(&h->st) != hipSuccess) {
        delete h;
        return fail(nullptr, MOCCA_ERR_HIP, "hipStreamCreate failed");
    }
    const size_t N = h->N, M = h->M, ts = h->tsize;
    Buffers &b = h->b;
    TRY(dalloc(h, &b.psi, M * 4 * N * ts));
    TRY(dalloc(h, &b.psi2, M * 4 * N * ts));
    TRY(dalloc(h, &b.upd, M * 4 * N * ts));
    TRY(dalloc(h, &b.hbuf, 4 * N * ts));
    TRY(dalloc(h, &b.fields, 48 * N * ts));
    TRY(dalloc(h, &b.dens, 2 * 9 * N * sizeof(double)));
    TRY(dalloc(h, &b.derived, 38 * N * sizeof(double)));
    // no densities yet: zero (the pairing factor of a first SolvePairing
    // reads them)
    HIPCHK(h, hipMemset(b.dens, 0, 2 * 9 * N * sizeof(double)));
    HIPCHK(h, hipMemset(b.derived, 0, 38 * N * sizeof(double)));
    TRY(dalloc(h, &b.qn, M * 4 * sizeof(int)));
    TRY(dalloc(h, &b.occ, M * sizeof(double)));
    TRY(dalloc(h, &b.E, M * sizeof(double)));
    TRY(dalloc(h, &b.D, M * sizeof(double)));
    TRY(dalloc(h, &b.nrm, M * sizeof(double)));
    TRY(dalloc(h, &b.status, sizeof(int)));
    TRY(dalloc(h, &b.one_chunk, sizeof(Chunk)));
    TRY(dalloc(h, &b.one_order, sizeof(int)));
    h->chol_lds = cholesky_lds_limit();
    h->ntiles = (int)((N + kBlock - 1) / kBlock);
    h->march = !mesh->lagrange && mesh->nx <= 64;
    h->TY = plane_tile_rows(mesh->nx, mesh->ny);
    h->ntile_y = (mesh->ny + h->TY - 1) / h->TY;
    TRY(dalloc(h, &b.one_tasks, h->ntile_y * sizeof(Task)));
    TRY(dalloc(h, &b.eparts, M * (size_t)h->ntiles * 3 * sizeof(double)));
    // Lagrange meshes with lines of 16..64 points take their derivatives from
    // the matrix cores (k_lag_axis) through a scratch of 24 values per state
    // and point; if that does not fit, the on-the-fly path is used.
    if (mesh->lagrange) {
        const int nmin = std::min(mesh->nx, std::min(mesh->ny, mesh->nz));
        const int nmax = std::max(mesh->nx, std::max(mesh->ny, mesh->nz));
        if (nmin >= 16 && nmax <= kLagMaxN) {
            if (hipMalloc(&b.lagscr, (size_t)M * 24 * N * ts) != hipSuccess) {
                b.lagscr = nullptr;
                (void)hipGetLastError();
            }
        }
    }
    TRY(dalloc(h, &b.K, M * sizeof(double)));
    HIPCHK(h, hipMemset(b.K, 0, M * sizeof(double)));
    HIPCHK(h, hipMemset(b.upd, 0, M * 4 * N * ts));
    HIPCHK(h, hipMemset(b.fields, 0, 48 * N * ts));
    HIPCHK(h, hipMemset(b.dens, 0, 2 * 9 * N * sizeof(double)));
    HIPCHK(h, hipMemset(b.derived, 0, 38 * N * sizeof(double)));
    HIPCHK(h, hipMemset(b.E, 0, M * sizeof(double)));
    HIPCHK(h, hipMemset(b.D, 0, M * sizeof(double)));
    HIPCHK(h, hipMemset(b.nrm, 0, M * sizeof(double)));
    HIPCHK(h, hipMemset(b.status, 0, sizeof(int)));
    // Lagrange matrices (EV8: the full line doubles every stored axis).
    const int ns[3] = {mesh->nx, mesh->ny, mesh->nz};
    size_t nl = 0;
    for (int a = 0; a < 3; ++a) nl += 4 * (size_t)ns[a] * ns[a];
    std::vector<double> lag(nl, 0.0);
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        lagrange_axis(ns[a], 2 * ns[a], mesh->dx, lag.data() + off);
        off += 4 * (size_t)ns[a] * ns[a];
    }
    TRY(dalloc(h, &b.lagD, nl * sizeof(double)));
    HIPCHK(h, hipMemcpy(b.lagD, lag.data(), nl * sizeof(double), hipMemcpyHostToDevice));
    TRY(dalloc(h, &b.lagT, nl * ts));
    TRY(up(h, b.lagT, lag.data(), nl));
    fill_mesh<double>(h, h->md, b.lagD);
    fill_mesh<float>(h, h->mf, (const float *)b.lagT);
    if (precision == MOCCA_FP64) fill_mesh<double>(h, h->md, (const double *)b.lagT);
    // Derived-density jobs.
    std::vector<FieldJob> jobs;
    build_density_jobs(h, jobs);
    h->njobs = (int)jobs.size();
    TRY(dalloc(h, &b.jobs, jobs.size() * sizeof(FieldJob)));
    HIPCHK(h, hipMemcpy(b.jobs, jobs.data(), jobs.size() * sizeof(FieldJob),
                        hipMemcpyHostToDevice));
#undef TRY
    *out = h;
    return 0;
}

int mocca_solve_destroy(mocca_solve *h) {
    if (!h) return 0;
    graphs_clear(h);
    if (h->graph_rows) (void)hipFree(h->graph_rows);
    free_all(h);
    delete h;
    return 0;
}

int mocca_synchronize(mocca_solve *h) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    HIPCHK(h, hipStreamSynchronize(h->st));
    return check_status(h);
}

int mocca_set_states_meta(mocca_solve *h, const int *qn, const double *occ) {
    if (h) ++h->graph_epoch;
    if (!h || !qn || !occ) return fail(h, MOCCA_ERR_ARG, "null argument");
    const int M = h->M;
    for (int i = 0; i < M; ++i) {
        const int *q = qn + 4 * i;
        if ((q[0] != -1 && q[0] != 1) || (q[1] != -1 && q[1] != 1) ||
            (q[2] != -1 && q[2] != 1) || (q[3] != -1 && q[3] != 1))
            return fail(h, MOCCA_ERR_ARG, "state %d: quantum numbers must be +-1", i);
        if (q[3] != 1)
            return fail(h, MOCCA_ERR_UNSUPPORTED,
                        "state %d: time simplex -1 is not stored in the EV8 octant", i);
    }
    h->qn.assign(qn, qn + 4 * M);
    h->occ.assign(occ, occ + M);
    Buffers &b = h->b;
    HIPCHK(h, hipMemcpy(b.qn, qn, M * 4 * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.occ, occ, M * sizeof(double), hipMemcpyHostToDevice));
    // Chunks: states grouped by isospin, kChunk per chunk.
    std::vector<int> order;
    std::vector<Chunk> chunks;
    for (int it = 0; it < 2; ++it) {
        std::vector<int> st;
        for (int i = 0; i < M; ++i)
            if ((qn[4 * i] + 3) / 2 - 1 == it) st.push_back(i);
        for (size_t k = 0; k < st.size(); k += kChunk) {
            Chunk c{(int)order.size(), (int)std::min<size_t>(kChunk, st.size() - k), it, 0};
            for (int q = 0; q < c.count; ++q) order.push_back(st[k + q]);
            chunks.push_back(c);
        }
    }
    // Density chunks (kDensChunk states of one isospin) over the same order.
    std::vector<Chunk> dchunks;
    for (const Chunk &c : chunks) {
        if (!dchunks.empty() && dchunks.back().it == c.it &&
            dchunks.back().count + c.count <= kDensChunk)
            dchunks.back().count += c.count;
        else
            dchunks.push_back(c);
    }
    // Density tasks: (tile, slab, chunk); all tasks of one chunk on one XCD.
    std::vector<DTask> dtasks;
    {
        const int slab = density_slab(h->mesh.time_reversal, h->tsize);
        const int nslab = (h->mesh.nz + slab - 1) / slab;
        std::vector<std::vector<DTask>> q(8);
        for (int c = 0; c < (int)dchunks.size(); ++c)
            for (int t = 0; t < h->ntile_y; ++t)
                for (int sl = 0; sl < nslab; ++sl) q[c % 8].push_back(DTask{t, sl, c, 0});
        std::vector<size_t> pos(8, 0);
        size_t total = 0;
        for (auto &v : q) total += v.size();
        for (size_t w = 0; dtasks.size() < total; ++w) {
            int k = (int)(w % 8);
            for (int tries = 0; tries < 8 && pos[k] >= q[k].size(); ++tries) k = (k + 1) % 8;
            dtasks.push_back(q[k][pos[k]++]);
        }
    }
    // Gram-Schmidt blocks: (isospin, parity, signature) (Spwf.f90:842-866).
    std::vector<std::vector<int>> blocks;
    for (int iso = -1; iso <= 1; iso += 2)
        for (int P = -1; P <= 1; P += 2)
            for (int S = -1; S <= 1; S += 2) {
                std::vector<int> st;
                for (int i = 0; i < M; ++i)
                    if (qn[4 * i] == iso && qn[4 * i + 1] == P && qn[4 * i + 2] == S)
                        st.push_back(i);
                if (!st.empty()) blocks.push_back(st);
            }
    int mmax = 0;
    for (auto &bl : blocks) mmax = std::max<int>(mmax, (int)bl.size());
    if (mmax > kMaxBlock)
        return fail(h, MOCCA_ERR_UNSUPPORTED,
                    "a (isospin, parity, signature) block holds %d states (max %d)", mmax,
                    kMaxBlock);
    const int max_m = (mmax + kGramTile - 1) / kGramTile * kGramTile;
    const int nblk = (int)blocks.size();
    std::vector<int> bs((size_t)nblk * max_m, 0), bm(nblk);
    for (int k = 0; k < nblk; ++k) {
        bm[k] = (int)blocks[k].size();
        for (int q = 0; q < bm[k]; ++q) bs[(size_t)k * max_m + q] = blocks[k][q];
    }
    std::vector<int> all(M);
    for (int i = 0; i < M; ++i) all[i] = i;
    // Task list of the z-marching kernels, XCD-aware: workgroup w runs on XCD
    // w % 8, so all tasks of one (isospin, y-tile) pair go to one XCD and share
    // that XCD's L2 copy of the tile's potentials.
    std::vector<Task> tasks;
    {
        std::vector<std::vector<Task>> q(8);
        for (int i = 0; i < M; ++i) {
            const int it = (qn[4 * i] + 3) / 2 - 1;
            for (int t = 0; t < h->ntile_y; ++t) q[(it * h->ntile_y + t) % 8].push_back(Task{i, t});
        }
        std::vector<size_t> pos(8, 0);
        const size_t total = (size_t)M * h->ntile_y;
        for (size_t w = 0; tasks.size() < total; ++w) {
            int k = (int)(w % 8);
            for (int tries = 0; tries < 8 && pos[k] >= q[k].size(); ++tries) k = (k + 1) % 8;
            tasks.push_back(q[k][pos[k]++]);
        }
    }
    // The pipelined FP64 step splits each (state, tile) into nzs z slabs when
    // that fills the CUs better: cost ~ rounds x (planes per slab + ~1.5
    // planes to start the register queue and the first loads).
    std::vector<Task> tasks_dma;
    h->dma_hpsi = h->march && h->prec == MOCCA_FP64 &&
                  hpsi_dma_ok(h->md, h->TY, h->mesh.time_reversal);
    h->nzs = 1;
    if (h->dma_hpsi) {
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        const long T1 = (long)tasks.size();
        const int nz = h->mesh.nz;
        double best = 0;
        for (int k = 1; k <= 4; ++k) {
            if (k > 1 && (nz / k < 6 || h->ntile_y * k > h->ntiles)) break;
            const double rounds = (double)((T1 * k + ncu - 1) / std::max(ncu, 1));
            const double cost = rounds * ((double)nz / k + 1.5);
            if (k == 1 || cost < best) { best = cost; h->nzs = k; }
        }
        // slab-major, so the first round holds every (state, tile) once
        for (int sl = 0; sl < h->nzs; ++sl)
            for (const Task &t : tasks) tasks_dma.push_back(Task{t.state, t.tile + h->ntile_y * sl});
    }
    // (Re)allocate index and GS buffers.
    void *old[] = {b.order, b.chunks, b.all_states, b.blk_states, b.blk_m, b.perm,
                   b.G, b.Tm, b.work, b.gpart, b.dpart, b.tasks, b.dchunks, b.dtasks,
                   b.tasks_dma};
    for (void *p : old)
        if (p) (void)hipFree(p);
    b.order = nullptr; b.chunks = nullptr; b.all_states = nullptr;
    b.blk_states = nullptr; b.blk_m = nullptr; b.perm = nullptr;
    b.G = b.Tm = b.work = b.gpart = nullptr; b.dpart = nullptr; b.tasks = nullptr;
    b.dchunks = nullptr; b.dtasks = nullptr; b.tasks_dma = nullptr;
    int rc;
    if ((rc = dalloc(h, &b.order, order.size() * sizeof(int)))) return rc;
    if ((rc = dalloc(h, &b.chunks, chunks.size() * sizeof(Chunk)))) return rc;
    if ((rc = dalloc(h, &b.all_states, M * sizeof(int)))) return rc;
    if ((rc = dalloc(h, &b.blk_states, bs.size() * sizeof(int)))) return rc;
    if ((rc = dalloc(h, &b.blk_m, nblk * sizeof(int)))) return rc;
    if ((rc = dalloc(h, &b.perm, bs.size() * sizeof(int)))) return rc;
    const GsPlan gp = gs_plan(nblk, mmax, 4L * h->N, h->M, h->tsize);
    const size_t gsz = (size_t)nblk * max_m * max_m * sizeof(double);
    if ((rc = dalloc(h, &b.G, gsz))) return rc;
    if ((rc = dalloc(h, &b.Tm, gsz))) return rc;
    if ((rc = dalloc(h, &b.work, gsz))) return rc;
    h->gtile = gp.gtile;
    h->mmax = mmax;
    h->nsplit = gp.nsplit;
    h->mpad = gp.mpad;
    h->mfma_gram = gp.mfma_gram;
    h->mfma_apply = gp.mfma_apply;
    h->nsplit_mfma = gp.nsplit_mfma;
    if ((rc = dalloc(h, &b.gpart, gp.gpart_doubles() * sizeof(double)))) return rc;
    const size_t npart = std::max(chunks.size(), dchunks.size());
    if ((rc = dalloc(h, &b.dpart, npart * 9 * (size_t)h->N * h->tsize))) return rc;
    if ((rc = dalloc(h, &b.dchunks, dchunks.size() * sizeof(Chunk)))) return rc;
    if ((rc = dalloc(h, &b.dtasks, dtasks.size() * sizeof(DTask)))) return rc;
    HIPCHK(h, hipMemcpy(b.dchunks, dchunks.data(), dchunks.size() * sizeof(Chunk),
                        hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.dtasks, dtasks.data(), dtasks.size() * sizeof(DTask),
                        hipMemcpyHostToDevice));
    h->ndchunks = (int)dchunks.size();
    h->ndtasks = (int)dtasks.size();
    if ((rc = dalloc(h, &b.tasks, tasks.size() * sizeof(Task)))) return rc;
    HIPCHK(h, hipMemcpy(b.tasks, tasks.data(), tasks.size() * sizeof(Task),
                        hipMemcpyHostToDevice));
    h->ntasks = (int)tasks.size();
    if ((rc = dalloc(h, &b.tasks_dma, std::max<size_t>(tasks_dma.size(), 1) * sizeof(Task))))
        return rc;
    if (!tasks_dma.empty())
        HIPCHK(h, hipMemcpy(b.tasks_dma, tasks_dma.data(), tasks_dma.size() * sizeof(Task),
                            hipMemcpyHostToDevice));
    h->ntasks_dma = (int)tasks_dma.size();
    HIPCHK(h, hipMemcpy(b.order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.chunks, chunks.data(), chunks.size() * sizeof(Chunk),
                        hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.all_states, all.data(), M * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.blk_states, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(b.blk_m, bm.data(), nblk * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(b.G, 0, gsz));
    HIPCHK(h, hipMemset(b.Tm, 0, gsz));
    HIPCHK(h, hipMemset(b.work, 0, gsz));
    h->nchunks = (int)chunks.size();
    h->nblk = nblk;
    h->max_m = max_m;
    h->meta_set = true;
    return 0;
}

int mocca_set_occupations(mocca_solve *h, const double *occ) {
    if (int rc = check_ready(h)) return rc;
    if (!occ) return fail(h, MOCCA_ERR_ARG, "null occupations");
    h->occ.assign(occ, occ + h->M);
    HIPCHK(h, hipMemcpy(h->b.occ, occ, h->M * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int mocca_set_energies(mocca_solve *h, const double *energy) {
    if (!h || !energy) return fail(h, MOCCA_ERR_ARG, "null argument");
    HIPCHK(h, hipMemcpy(h->b.E, energy, h->M * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int mocca_get_energies(mocca_solve *h, double *energy, double *disp) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (energy)
        HIPCHK(h, hipMemcpy(energy, h->b.E, h->M * sizeof(double), hipMemcpyDeviceToHost));
    if (disp)
        HIPCHK(h, hipMemcpy(disp, h->b.D, h->M * sizeof(double), hipMemcpyDeviceToHost));
    return check_status(h);
}

int mocca_upload_states(mocca_solve *h, int first, int count, const double *grid) {
    if (h) ++h->graph_epoch;
    if (!h || !grid) return fail(h, MOCCA_ERR_ARG, "null argument");
    if (first < 0 || count < 0 || first + count > h->M)
        return fail(h, MOCCA_ERR_ARG, "states [%d, %d) out of range (nwt=%d)", first,
                    first + count, h->M);
    HIPCHK(h, hipStreamSynchronize(h->st));
    const size_t n4 = 4 * (size_t)h->N;
    h->scf.alt_pending = false;
    h->prev_valid = false;
    return up(h, (char *)h->b.psi + first * n4 * h->tsize, grid, count * n4);
}

int mocca_download_states(mocca_solve *h, int first, int count, double *grid) {
    if (!h || !grid) return fail(h, MOCCA_ERR_ARG, "null argument");
    if (first < 0 || count < 0 || first + count > h->M)
        return fail(h, MOCCA_ERR_ARG, "states [%d, %d) out of range (nwt=%d)", first,
                    first + count, h->M);
    const size_t n4 = 4 * (size_t)h->N;
    if (int rc = down(h, grid, (char *)h->b.psi + first * n4 * h->tsize, count * n4))
        return rc;
    return check_status(h);
}

int mocca_reset_updates(mocca_solve *h) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    h->prev_valid = false;   // the next GradDesc starts with u = 0
    return 0;
}

int mocca_set_potentials(mocca_solve *h, const mocca_fields *f) {
    if (h) ++h->graph_epoch;
    if (!h || !f) return fail(h, MOCCA_ERR_ARG, "null argument");
    if (!f->U || !f->B || !f->NablaB || !f->W || !f->DivW)
        return fail(h, MOCCA_ERR_ARG, "U, B, NablaB, W and DivW are required");
    if (!h->mesh.time_reversal && (!f->S || !f->A || !f->DivA))
        return fail(h, MOCCA_ERR_ARG, "S, A and DivA are required when time reversal is broken");
    HIPCHK(h, hipStreamSynchronize(h->st));
    const size_t N = h->N, ts = h->tsize;
    char *F = (char *)h->b.fields;
    struct { const double *p; int off, n; } parts[] = {
        {f->U, 0, 2}, {f->B, 2, 2}, {f->NablaB, 4, 6}, {f->W, 10, 18}, {f->DivW, 28, 6},
        {f->S, 34, 6}, {f->A, 40, 6}, {f->DivA, 46, 2}};
    for (auto &q : parts) {
        if (!q.p) {
            HIPCHK(h, hipMemset(F + q.off * N * ts, 0, q.n * N * ts));
            continue;
        }
        if (int rc = up(h, F + q.off * N * ts, q.p, q.n * N)) return rc;
    }
    // W[it][nu][mu]: antisymmetric with a zero diagonal, as without tensor
    // terms, lets the step kernel move 3 of its 9 components
    bool anti = true;
    for (int it = 0; it < 2 && anti; ++it)
        for (int nu = 0; nu < 3 && anti; ++nu)
            for (int mu = nu; mu < 3 && anti; ++mu) {
                const double *a = f->W + ((size_t)(it * 3 + nu) * 3 + mu) * N;
                const double *b = f->W + ((size_t)(it * 3 + mu) * 3 + nu) * N;
                for (size_t p = 0; p < N && anti; ++p)
                    anti = mu == nu ? a[p] == 0.0 : a[p] == -b[p];
            }
    h->w_antisym = anti;
    h->kin[0] = f->kin[0];
    h->kin[1] = f->kin[1];
    h->fields_set = true;
    return 0;
}

int mocca_derive_all(mocca_solve *h) {
    if (int rc = check_ready(h)) return rc;
    const size_t N = h->N, M = h->M;
    h->scf.alt_pending = false;   // DeriveAll refreshes the derivatives
    if (!h->b.der) {
        if (int rc = dalloc(h, &h->b.der, M * 12 * N * sizeof(double))) return rc;
        if (int rc = dalloc(h, &h->b.lap, M * 4 * N * sizeof(double))) return rc;
    }
    return run(h, "derive_all", [&] {
        if (h->prec == MOCCA_FP64)
            launch_derive_all<double>(h->md, (const double *)h->b.psi, h->b.qn, h->M,
                                      h->b.der, h->b.lap, h->st);
        else
            launch_derive_all<float>(h->mf, (const float *)h->b.psi, h->b.qn, h->M,
                                     h->b.der, h->b.lap, h->st);
    });
}

int mocca_download_derivatives(mocca_solve *h, int state, double *der, double *lap) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    if (!h->b.der) return fail(h, MOCCA_ERR_STATE, "mocca_derive_all not called");
    if (state < 0 || state >= h->M) return fail(h, MOCCA_ERR_ARG, "state %d out of range", state);
    const size_t N = h->N;
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (der)
        HIPCHK(h, hipMemcpy(der, h->b.der + state * 12 * N, 12 * N * sizeof(double),
                            hipMemcpyDeviceToHost));
    if (lap)
        HIPCHK(h, hipMemcpy(lap, h->b.lap + state * 4 * N, 4 * N * sizeof(double),
                            hipMemcpyDeviceToHost));
    return 0;
}

int mocca_hpsi(mocca_solve *h, int state, double *out) {
    if (int rc = check_ready(h)) return rc;
    if (!h->fields_set) return fail(h, MOCCA_ERR_STATE, "mocca_set_potentials not called");
    if (state < 0 || state >= h->M || !out)
        return fail(h, MOCCA_ERR_ARG, "invalid state %d or null output", state);
    int rc = h->prec == MOCCA_FP64 ? hpsi_one_impl<double>(h, state)
                                   : hpsi_one_impl<float>(h, state);
    if (rc) return rc;
    return down(h, out, h->b.hbuf, 4 * (size_t)h->N);
}

int mocca_grad_desc(mocca_solve *h, double dt_over_hbar, double momentum) {
    if (int rc = check_ready(h)) return rc;
    if (!h->fields_set) return fail(h, MOCCA_ERR_STATE, "mocca_set_potentials not called");
    return grad_desc_any(h, dt_over_hbar, momentum);
}

int mocca_gram_schmidt(mocca_solve *h) {
    if (int rc = check_ready(h)) return rc;
    int rc = h->prec == MOCCA_FP64
                 ? gram_schmidt_impl<double>(h, (const double *)h->b.psi,
                                             (double *)h->b.psi2, nullptr, 0)
                 : gram_schmidt_impl<float>(h, (const float *)h->b.psi,
                                            (float *)h->b.psi2, nullptr, 0);
    if (rc) return rc;
    std::swap(h->b.psi, h->b.psi2);
    return 0;
}

int mocca_set_lagrange_precision(mocca_solve *h, int mode) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    if (mode != 0 && mode != 1) return fail(h, MOCCA_ERR_ARG, "unknown mode %d (0 or 1)", mode);
    if (mode == 1 && h->prec != MOCCA_MIXED)
        return fail(h, MOCCA_ERR_UNSUPPORTED,
                    "FP32 matrix-core Lagrange derivatives need the mixed-precision states");
    ++h->graph_epoch;
    h->lag_split = mode == 1;
    return 0;
}

int mocca_angular_momentum(mocca_solve *h, double *states, double *canonical,
                           double *totals) {
    if (int rc = check_ready(h)) return rc;
    if (!states && !canonical && !totals) return fail(h, MOCCA_ERR_ARG, "null arguments");
    if (canonical && !h->hfb.on)
        return fail(h, MOCCA_ERR_STATE, "no canonical basis: HFB pairing is not active");
    const int M = h->M;
    const long N = h->N;
    const int nblk = (int)((N + kBlock - 1) / kBlock);
    const bool can = h->hfb.on;
    const int nb = can ? 2 : 1;
    double *part = nullptr, *out = nullptr;
    int rc;
    if ((rc = dalloc(h, &part, std::max<size_t>((size_t)M * nblk * kAmSums, nblk) *
                                   sizeof(double))) ||
        (rc = dalloc(h, &out, (size_t)nb * M * 21 * sizeof(double)))) {
        if (part) (void)hipFree(part);
        return rc;
    }
    std::vector<double> res((size_t)nb * M * 21), dpart(nblk);
    auto body = [&]() -> int {
        HIPCHK(h, hipMemsetAsync(out, 0, res.size() * sizeof(double), h->st));
        int r = run(h, "angmom", [&] {
            for (int b = 0; b < nb; ++b) {
                const void *src = b == 0 ? h->b.psi : h->hfb.can_psi;
                if (h->prec == MOCCA_FP64)
                    launch_angmom<double>(h->md, (const double *)src, h->b.qn, M, h->mesh.dx,
                                          h->mesh.dv, part, out + (size_t)b * M * 21, h->st);
                else
                    launch_angmom<float>(h->mf, (const float *)src, h->b.qn, M, h->mesh.dx,
                                         h->mesh.dv, part, out + (size_t)b * M * 21, h->st);
            }
        });
        if (r) return r;
        HIPCHK(h, hipStreamSynchronize(h->st));
        HIPCHK(h, hipMemcpy(res.data(), out, res.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (totals && !h->mesh.time_reversal) {
            r = run(h, "angmom", [&] {
                launch_angmom_dens(h->b.dens, h->mesh.nx, h->mesh.ny, N, h->mesh.dx, part, h->st);
            });
            if (r) return r;
            HIPCHK(h, hipStreamSynchronize(h->st));
            HIPCHK(h, hipMemcpy(dpart.data(), part, nblk * sizeof(double), hipMemcpyDeviceToHost));
        }
        return 0;
    };
    rc = body();
    (void)hipStreamSynchronize(h->st);
    (void)hipFree(part);
    (void)hipFree(out);
    if (rc) return rc;
    if (states) std::copy(res.begin(), res.begin() + (size_t)M * 21, states);
    if (canonical) std::copy(res.begin() + (size_t)M * 21, res.end(), canonical);
    if (!totals) return 0;
    // UpdateAM's sums (Cranking.f90:135-232) over the density basis, only
    // without time reversal: TotalAngMom, JTR, JTI, J2Total,
    // TotalAngMom_dens, AMIsoblock(S, P, it, i)
    std::fill(totals, totals + 39, 0.0);
    if (h->mesh.time_reversal) return 0;
    std::vector<double> occ(M);
    std::vector<int> qn(4 * (size_t)M);
    HIPCHK(h, hipMemcpy(occ.data(), can ? h->hfb.can_occ : h->b.occ, M * sizeof(double),
                        hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(qn.data(), h->b.qn, qn.size() * sizeof(int), hipMemcpyDeviceToHost));
    const double *db = res.data() + (can ? (size_t)M * 21 : 0);
    for (int w = 0; w < M; ++w) {
        const double o = occ[w];
        const double *r = db + (size_t)w * 21;
        const int it = (qn[4 * w] + 3) / 2 - 1, P = (qn[4 * w + 1] + 3) / 2 - 1,
                  S = (qn[4 * w + 2] + 3) / 2 - 1;
        for (int i = 0; i < 3; ++i) {
            totals[i] += o * r[i];
            totals[3 + i] += o * r[3 + i];
            totals[6 + i] += o * r[6 + i];
            totals[9 + i] += o * r[9 + i];
            totals[15 + S + 2 * P + 4 * it + 8 * i] += o * r[i];
        }
    }
    double dz = 0.0;
    for (int b = 0; b < nblk; ++b) dz += dpart[b];
    totals[14] = dz * h->mesh.dv;
    return 0;
}

int mocca_compute_density(mocca_solve *h, mocca_density *out) {
    if (int rc = check_ready(h)) return rc;
    int rc = density_any(h);
    if (rc) return rc;
    if (!out) return 0;
    HIPCHK(h, hipStreamSynchronize(h->st));
    const size_t N = h->N;
    auto cp = [&](double *dst, const double *src, size_t n) -> int {
        if (!dst) return 0;
        HIPCHK(h, hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    };
    // dens is [2][9][N] (slots beyond nf unused when time reversal holds)
    std::vector<double> d(2 * 9 * N);
    HIPCHK(h, hipMemcpy(d.data(), h->b.dens, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    auto scatter = [&](double *dst, int f, int ncomp) {
        if (!dst) return;
        for (int it = 0; it < 2; ++it)
            for (int c = 0; c < ncomp; ++c)
                std::memcpy(dst + ((size_t)it * ncomp + c) * N,
                            d.data() + ((size_t)it * 9 + f + c) * N, N * sizeof(double));
    };
    scatter(out->rho, 0, 1);
    scatter(out->tau, 1, 1);
    scatter(out->nablaJ, 2, 1);
    const double *der = h->b.derived;
    int rc2 = 0;
    if ((rc2 = cp(out->derrho, der, 6 * N))) return rc2;
    if ((rc2 = cp(out->laprho, der + 6 * N, 2 * N))) return rc2;
    if (!h->mesh.time_reversal) {
        scatter(out->vecj, 3, 3);
        scatter(out->vecs, 6, 3);
        if ((rc2 = cp(out->rotvecj, der + 8 * N, 6 * N))) return rc2;
        if ((rc2 = cp(out->rots, der + 14 * N, 6 * N))) return rc2;
        if ((rc2 = cp(out->ders, der + 20 * N, 18 * N))) return rc2;
    }
    return 0;
}

int mocca_profile_enable(mocca_solve *h, int enable) {
    if (!h) return fail(nullptr, MOCCA_ERR_ARG, "null solve handle");
    if (int rc = drain_profile(h)) return rc;
    h->prof = enable != 0;
    if (enable) h->acc.clear();
    return 0;
}

int mocca_profile_get(mocca_solve *h, const char *name, double *ms, long *count) {
    if (!h || !name) return fail(h, MOCCA_ERR_ARG, "null argument");
    if (int rc = drain_profile(h)) return rc;
    auto it = h->acc.find(name);
    if (ms) *ms = it == h->acc.end() ? 0.0 : it->second.first;
    if (count) *count = it == h->acc.end() ? 0 : it->second.second;
    return 0;
}

}  // extern "C"
