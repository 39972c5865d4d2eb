// C ABI of the self-consistent part: functional set-up, HF filling,
// density mixing, moments, Coulomb, potentials, energy and the fused
// iteration loop.  Entry points and their reference counterparts are listed
// in include/mocca_amd.h.
#include "mocca_internal.h"

#include <algorithm>
#include <functional>
#include <thread>
#include <cmath>
#include <cstring>

using namespace mocca;
using namespace mocca::detail;

namespace {

constexpr double kPi = 3.14159265358979323846;

// Associated Legendre functions P_n^m(x), n = m..nmax, with the Condon-Shortley
// phase, by the recurrence legendre_pnm uses (Mesh.f90:92-268):
// P_m^m = (-1)^m (2m-1)!! (1-x^2)^(m/2), P_{m+1}^m = x (2m+1) P_m^m,
// P_n^m = ((2n-1) x P_{n-1}^m - (n+m-1) P_{n-2}^m) / (n-m).
void legendre(int nmax, int m, double x, double *p) {
    for (int i = 0; i < m; ++i) p[i] = 0.0;
    p[m] = 1.0;
    const double s = std::sqrt(1.0 - x * x);
    double odd = 1.0;
    for (int i = 1; i <= m; ++i) {
        p[m] = -p[m] * odd * s;
        odd += 2.0;
    }
    if (m == nmax) return;
    p[m + 1] = x * double(2 * m + 1) * p[m];
    for (int n = m + 2; n <= nmax; ++n)
        p[n] = (double(2 * n - 1) * x * p[n - 1] - double(n + m - 1) * p[n - 2]) / double(n - m);
}

// r^l Re Y_lm(X, Y, Z) with z the quantisation axis and phi = atan2(Y, X)
// (SphericalHarmonics, Moments.f90:584-691).
double rlylm(int l, int m, double X, double Y, double Z) {
    double p[32];
    const double r = std::sqrt(X * X + Y * Y + Z * Z);
    legendre(l, m, Z / r, p);
    double q = 1.0;
    for (int k = 1; k <= 2 * m; ++k) q = q / double(l + m - k + 1);
    double fac = ((m & 1) ? -1.0 : 1.0) * std::sqrt(q) * std::pow(r, l);
    fac = fac * std::sqrt(double(2 * l + 1) / (4 * kPi));
    return fac * p[l] * std::cos(m * std::atan2(Y, X));
}

// Eigen-decomposition of a symmetric n x n matrix by cyclic Jacobi rotations
// (Golub & Van Loan, Matrix Computations, sec. 8.5): a = Q diag(w) Q^T with
// Q[i][k] the i-th component of eigenvector k.
void jacobi_eigen(int n, std::vector<double> a, std::vector<double> &Q, std::vector<double> &w) {
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A(i, j) * A(i, j);
                if (i != j) off += A(i, j) * A(i, j);
            }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - sn * akq;
                    A(k, q) = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - sn * aqk;
                    A(q, k) = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double &qkp = Q[(size_t)k * n + p], &qkq = Q[(size_t)k * n + q];
                    const double x = qkp, y = qkq;
                    qkp = c * x - sn * y;
                    qkq = sn * x + c * y;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A(i, i);
}

// The 1-D Coulomb operator on n interior points: 5-point second derivative
// (-1/12, 4/3, -5/2, 4/3, -1/12)/dx^2, mirrored (+) below the symmetry plane
// (Coulomb_1D, CoulombDerivatives.f90:228-262); the two boundary planes
// beyond the far edge are known values, not unknowns.
std::vector<double> coulomb_1d(int n, double dx) {
    const double cf[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
    std::vector<double> t((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int q = -2; q <= 2; ++q) {
            int a = i + q;
            if (a < 0) a = -1 - a;
            if (a < n) t[(size_t)i * n + a] += cf[q + 2] / (dx * dx);
        }
    return t;
}

void free_ptr(void *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int check_scf(mocca_solve *h) {
    if (int rc = check_ready(h)) return rc;
    if (!h->scf.set) return fail(h, MOCCA_ERR_STATE, "mocca_scf_setup not called");
    return 0;
}

// DivWPot and DivAPot terms (MeanFields.f90:463-468, 580-592): for each
// output, the field derivative and the (P, S, TS, component) it is taken with.
void build_potential_jobs(mocca_solve *h, std::vector<FieldJob> &jobs) {
    const long N = h->N;
    const int lag = h->mesh.lagrange, trc = h->mesh.time_reversal;
    double *pot = h->scf.pot;
    auto term = [&](FieldJob &j, int t, const double *src, int axis, int P, int S, int TS,
                    int comp) {
        j.src[t] = src;
        j.axis[t] = axis;
        j.sign[t][0] = j.sign[t][1] = j.sign[t][2] = sign_first_h(axis, P, S, TS, comp - 1, lag);
        j.coef[t] = 1.0;
        j.nterms = t + 1;
    };
    for (int it = 0; it < 2; ++it)
        for (int a = 0; a < 3; ++a) {
            FieldJob j{};
            term(j, 0, pot + (kOffB + it) * N, a, 1, 1, 1, 1);
            j.dst = pot + (kOffNB + it * 3 + a) * N;
            jobs.push_back(j);
        }
    // DivW_nu = sum_mu d_mu W_{mu nu}; (P, S, TS, comp) per (nu, mu)
    static const int q[3][3][4] = {{{-1, 1, 1, 2}, {-1, 1, 1, 1}, {-1, -1, 1, 2}},
                                   {{-1, 1, 1, 1}, {-1, 1, 1, 2}, {-1, -1, 1, 1}},
                                   {{-1, -1, 1, 2}, {-1, -1, 1, 1}, {-1, 1, 1, 2}}};
    for (int it = 0; it < 2; ++it)
        for (int nu = 0; nu < 3; ++nu) {
            FieldJob j{};
            for (int mu = 0; mu < 3; ++mu)
                term(j, mu, pot + (kOffW + (it * 3 + nu) * 3 + mu) * N, mu, q[nu][mu][0],
                     q[nu][mu][1], q[nu][mu][2], q[nu][mu][3]);
            j.dst = pot + (kOffDW + it * 3 + nu) * N;
            jobs.push_back(j);
        }
    if (trc) return;
    for (int it = 0; it < 2; ++it) {
        FieldJob j{};
        term(j, 0, pot + (kOffA + it * 3 + 0) * N, 0, -1, -1, 1, 2);
        term(j, 1, pot + (kOffA + it * 3 + 1) * N, 1, -1, -1, 1, 1);
        term(j, 2, pot + (kOffA + it * 3 + 2) * N, 2, -1, 1, 1, 2);
        j.dst = pot + (kOffDA + it) * N;
        jobs.push_back(j);
    }
}

// CalculateAllMoments + ConstructPotentials.
int construct_impl(mocca_solve *h, int passes = 1) {
    Scf &s = h->scf;
    const long N = h->N;
    double *dens = h->b.dens;
    int rc = run(h, "moments", [&] {
        launch_moments(dens, dens + 9 * N, s.Y, s.nmom, N, h->mesh.dv, s.mpart, s.mom_nblk, s.Q,
                       h->st);
    });
    if (rc) return rc;
    if (s.ncon) {
        // the constraint part of CalculateAllMoments (Moments.f90:2250-2334):
        // cutoff, <Q^2>, targets and the constraint field; the start-up
        // calls it three times (Main.f90:213-215)
        rc = run(h, "cutoff", [&] {
            launch_rutz_cutoff(dens, h->mesh.nx, h->mesh.ny, h->mesh.nz, h->mesh.dx, h->mesh.dv,
                               s.radd, s.acut, s.cstat, s.ccnt, s.csurf, s.cut, h->st);
        });
        if (rc) return rc;
        rc = run(h, "constraints", [&] {
            launch_con_squared(s.con, s.ncon, s.Y, dens, N, h->mesh.dv, s.csq, h->st);
            for (int k = 0; k < passes; ++k) {
                launch_con_update(s.con, s.ncon, s.cctl, s.Q, s.csq, s.cstat, kConCalc, 0, 0.0,
                                  s.ccoef, s.calt, h->st);
                launch_con_field(s.con, s.ncon, s.Y, s.ccoef, s.cctl, s.cut, N, s.ce, h->st);
            }
            launch_con_history(s.con, s.ncon, s.Q, s.chist, passes, h->st);
        });
        if (rc) return rc;
    }
    if (s.sk.coulomb == 2) {
        rc = run(h, "coulomb_bc", [&] {
            launch_coulomb_bc(s.Q, s.nmom, s.lfac, s.bidx, s.bsh, s.brp, s.nb, s.sk.e2, s.cx,
                              h->st);
        });
        if (rc) return rc;
        rc = run(h, "coulomb_direct", [&] {
            const double *Qa[3] = {s.cq[0], s.cq[1], s.cq[2]};
            const double *La[3] = {s.clam[0], s.clam[1], s.clam[2]};
            launch_coulomb_direct(s.ck, s.cx, dens + 9 * N, Qa, La, s.cw0, s.cw1, s.vcoul,
                                  s.vcoulx, h->st);
        });
        if (rc) return rc;
    } else if (s.sk.coulomb == 1) {
        rc = run(h, "coulomb_bc", [&] {
            launch_coulomb_bc(s.Q, s.nmom, s.lfac, s.bidx, s.bsh, s.brp, s.nb, s.sk.e2, s.cx,
                              h->st);
        });
        if (rc) return rc;
        int herr = 0;
        rc = run(h, "coulomb_cg", [&] {
            herr = launch_coulomb_cg(s.ck, s.cx, dens + 9 * N, s.cr, s.cpa, s.cpb, s.cap, s.cpart,
                                     s.cg_nwg, s.vcoul, s.vcoulx, s.cg_iters, h->st);
        });
        if (rc) return rc;
        if (herr != (int)hipSuccess)
            return fail(h, MOCCA_ERR_HIP, "cooperative launch of the Coulomb solver failed: %s",
                        hipGetErrorString((hipError_t)herr));
    }
    // MixingScheme 2 updates U and S from their previous values (UpotOld,
    // SpotOld, PotentialMixing.f90:114-124), skipped on the first
    // construction, when those are all zero (:143)
    const bool precond = s.mixing_scheme == 2 && s.pot_built;
    if (precond) {
        HIPCHK(h, hipMemcpyAsync(s.pold, s.pot + kOffU * N, 2 * N * sizeof(double),
                                 hipMemcpyDeviceToDevice, h->st));
        if (!h->mesh.time_reversal)
            HIPCHK(h, hipMemcpyAsync(s.pold + 2 * N, s.pot + kOffS * N, 6 * N * sizeof(double),
                                     hipMemcpyDeviceToDevice, h->st));
    }
    rc = run(h, "potentials", [&] {
        launch_potentials(s.sk, N, h->mesh.time_reversal, dens, h->b.derived, s.vcoul, s.vcoulx,
                          s.ncon ? s.ce : nullptr, s.pot, h->st);
    });
    if (rc) return rc;
    rc = run(h, "pot_jobs", [&] { launch_field_jobs(h->md, s.pjobs, s.npjobs, h->st); });
    if (rc) return rc;
    if (precond) {
        rc = run(h, "precondition", [&] {
            launch_precondition(h->md, s.pcjobs, s.npc, s.pcwork, -s.preconu, 1.0, h->mesh.dv,
                                s.pc_iters, h->st);
        });
        if (rc) return rc;
    }
    s.pot_built = true;
    if (h->prec != MOCCA_FP64) {
        rc = run(h, "pot_convert", [&] {
            launch_convert_f32(s.pot, (float *)h->b.fields, (long)kNumFieldSlots * N, h->st);
        });
        if (rc) return rc;
    }
    h->fields_set = true;
    h->w_antisym = true;   // W_{mu nu} = -eps_{k mu nu} w_k (k_potentials), mixed linearly
    return 0;
}

int energy_terms_impl(mocca_solve *h, double *row) {
    Scf &s = h->scf;
    return run(h, "energy", [&] {
        launch_energy_terms(s.sk, h->N, h->mesh.time_reversal, h->b.dens, h->b.derived, s.vcoul,
                            s.epart, s.e_nblk, h->mesh.dv, row, h->st);
    });
}

int energy_kinetic_impl(mocca_solve *h, double *row) {
    const double *occ = h->hfb.on ? h->hfb.can_occ : h->b.occ;
    return run(h, "energy_kinetic", [&] {
        launch_energy_kinetic(h->scf.sk, h->b.K, occ, h->b.qn, h->M, row, h->st);
    });
}

int fill_impl(mocca_solve *h) {
    Scf &s = h->scf;
    if (h->hfb.bcs) return bcs_solve(h);
    if (h->hfb.on) {
        const bool first = !h->hfb.started;
        h->hfb.started = true;
        return hfb_solve_any(h, first);
    }
    return run(h, "hf_fill", [&] {
        launch_hf_fill(h->b.E, h->b.qn, h->M, s.ub_n, s.ub_p, s.fill_fac, h->b.occ, h->st);
    });
}

int save_history(mocca_solve *h) {
    const size_t N = h->N;
    HIPCHK(h, hipMemcpyAsync(h->scf.hist, h->b.dens, 18 * N * sizeof(double),
                             hipMemcpyDeviceToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(h->scf.hist + 18 * N, h->b.derived, 20 * N * sizeof(double),
                             hipMemcpyDeviceToDevice, h->st));
    return 0;
}

int mix_impl(mocca_solve *h, int keep) {
    const long N = h->N;
    Scf &s = h->scf;
    if (s.mixing_scheme == 2) return 0;   // MixDensities: potentials are mixed instead
    return run(h, "mix", [&] {
        launch_mix(h->b.dens, s.hist, 18 * N, s.damping, keep, h->st);
        launch_mix(h->b.derived, s.hist + 18 * N, 20 * N, s.damping, keep, h->st);
    });
}

// ReadjustAllMoments + TurnOffConstraints(iteration) (Moments.f90:2336-2366,
// 3516-3549), and the AlternateStep coefficients for the next iteration.
int readjust_impl(mocca_solve *h, int iteration, double momentum) {
    Scf &s = h->scf;
    if (!s.ncon) return 0;
    for (size_t c = 0; c < s.htype.size(); ++c)
        if (s.htype[c] != 0 && s.hiter[c] > 0 && iteration > s.hiter[c]) s.htype[c] = 0;
    return run(h, "constraints", [&] {
        launch_con_update(s.con, s.ncon, s.cctl, s.Q, s.csq, s.cstat,
                          kConReadjust | kConTurnOff, iteration, momentum, s.ccoef, s.calt,
                          h->st);
    });
}

bool alternating(const mocca_solve *h) {
    for (int t : h->scf.htype)
        if (t == 2) return true;
    return false;
}

int check_gs_status(mocca_solve *h) {
    int st = 0;
    HIPCHK(h, hipMemcpyAsync(&st, h->b.status, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (st) {
        HIPCHK(h, hipMemsetAsync(h->b.status, 0, sizeof(int), h->st));
        return fail(h, MOCCA_ERR_NUMERIC,
                    "Gram-Schmidt met a non-positive norm (linearly dependent or NaN states)");
    }
    return 0;
}

// CompEnergy's OldEnergy shift (Energy.f90:93-98) with the total energy of
// a completed energy row, and the Fermi energies of that row.
void push_energy(mocca_solve *h, const double *row) {
    Scf &s = h->scf;
    for (int i = 7; i > 0; --i) s.ehist[i] = s.ehist[i - 1];
    s.ehist[0] = row[0];
    s.efermi[0] = row[45];
    s.efermi[1] = row[46];
}

// The Evolve loop (Main.f90:255-328) for niter iterations.  With estimate,
// IterativeEstimation runs before each step (ParameterEstimation = 1,
// ImaginaryTime.f90:464) and dt, momentum carry over from one iteration to
// the next; params (when not null) receives the pair used at each.
// The Evolve loop in three parts, so that the iterations of several solves
// can be issued in turn on their own streams (mocca_batch_run): rows for
// niter iterations, one iteration, and the final kinetic energy with the
// status checks.  Only IterativeEstimation reads results back inside an
// iteration.
int iter_begin(mocca_solve *h, int niter) {
    if (int rc = check_scf(h)) return rc;
    if (!h->scf.started) return fail(h, MOCCA_ERR_STATE, "mocca_scf_start not called");
    if (niter < 0) return fail(h, MOCCA_ERR_ARG, "niter must be >= 0");
    Scf &s = h->scf;
    if (s.erow_cap < niter + 1) {
        if (s.erows) {
            HIPCHK(h, hipStreamSynchronize(h->st));
            (void)hipFree(s.erows);
            s.erows = nullptr;
        }
        if (int rc = dalloc(h, &s.erows, (size_t)(niter + 1) * kRowLen * sizeof(double))) return rc;
        HIPCHK(h, hipMemsetAsync(s.erows, 0, (size_t)(niter + 1) * kRowLen * sizeof(double),
                                 h->st));
        s.erow_cap = niter + 1;
    }
    HIPCHK(h, hipMemcpyAsync(s.erows, s.ecur, kRowLen * sizeof(double), hipMemcpyDeviceToDevice,
                             h->st));
    return 0;
}

// Iteration k (1-based within the call) with Evolve's iteration number iter,
// writing its energy row to row and completing the previous one in prev.
int iter_step(mocca_solve *h, int k, int iter, double &dt, double &dth, double &momentum,
              bool estimate, double *params, double *prev, double *row) {
    Scf &s = h->scf;
    int rc;
    // HF: a row's kinetic energy comes from the next GradDesc's hPsi pass
    const bool deferred = !h->hfb.on;
    // AlternateStep when alternating constraints were on after the last
    // SolvePairing (Main.f90:262-264); the previous row's kinetic energy
    // then comes from the states before it
    const bool alt = s.alt_next;
    if (alt) {
        if (deferred) {
            if ((rc = kinetic_any(h))) return rc;
            if ((rc = energy_kinetic_impl(h, prev))) return rc;
        }
        if ((rc = alternate_any(h))) return rc;
    }
    if (estimate) {
        double out[6];
        if ((rc = estimate_any(h, iter, dt, momentum, out))) return rc;
        dt = out[0];
        dth = dt / kHbar;
        momentum = out[1];
    }
    if (params) {
        params[2 * (k - 1)] = dt;
        params[2 * (k - 1) + 1] = momentum;
    }
    if ((rc = grad_desc_any(h, dth, momentum))) return rc;
    // the hPsi pass of GradDesc also gave <psi|Lap psi> of the states the
    // previous row's densities came from: that row's kinetic energy
    if (deferred && !alt && (rc = energy_kinetic_impl(h, prev))) return rc;
    if ((rc = fill_impl(h))) return rc;
    s.alt_next = alternating(h);
    if ((rc = density_any(h))) return rc;
    if ((rc = mix_impl(h, 1))) return rc;
    if ((rc = construct_impl(h))) return rc;
    if ((rc = readjust_impl(h, iter, momentum))) return rc;
    if ((rc = energy_terms_impl(h, row))) return rc;
    if (h->hfb.bcs && (rc = bcs_energy_impl(h, row))) return rc;
    if (!deferred) {
        // HFB: the density basis is the canonical one, whose kinetic sums
        // the GradDesc pass does not give
        if ((rc = hfb_energy_impl(h, row))) return rc;
        if ((rc = kinetic_any(h))) return rc;
        if ((rc = energy_kinetic_impl(h, row))) return rc;
    }
    return 0;
}

int iter_end(mocca_solve *h, int niter, double *energies) {
    Scf &s = h->scf;
    int rc;
    // kinetic energy of the final states: one hPsi-type pass with no outputs
    double *last = s.erows + (size_t)niter * kRowLen;
    if (!h->hfb.on) {
        if ((rc = kinetic_any(h))) return rc;
        if ((rc = energy_kinetic_impl(h, last))) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(s.ecur, last, kRowLen * sizeof(double), hipMemcpyDeviceToDevice,
                             h->st));
    // (the rows also feed the energy history of the convergence check)
    std::vector<double> rows((size_t)(niter + 1) * kRowLen);
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(rows.data(), s.erows, rows.size() * sizeof(double),
                        hipMemcpyDeviceToHost));
    if (energies) std::copy(rows.begin(), rows.end(), energies);
    if (s.start_pending && niter > 0) push_energy(h, rows.data());
    if (niter > 0) s.start_pending = false;
    for (int k = 1; k <= niter; ++k) push_energy(h, rows.data() + (size_t)k * kRowLen);
    if ((rc = check_gs_status(h))) return rc;
    return h->hfb.on ? hfb_check_status(h) : 0;
}

int iterate_impl(mocca_solve *h, int niter, int first_iteration, double dt, double dth,
                 double momentum, bool estimate, double *energies, double *params) {
    int rc;
    if ((rc = iter_begin(h, niter))) return rc;
    for (int k = 1; k <= niter; ++k) {
        if ((rc = iter_step(h, k, first_iteration + k - 1, dt, dth, momentum, estimate, params,
                            h->scf.erows + (size_t)(k - 1) * kRowLen,
                            h->scf.erows + (size_t)k * kRowLen)))
            return rc;
        h->scf.iteration = first_iteration + k - 1;
    }
    return iter_end(h, niter, energies);
}

// Replayed iterations of mocca_batch_run.  An iteration's kernels depend on
// the host only through the buffer pointers that rotate from one iteration
// to the next (states, stepped states, previous states, the states before an
// AlternateStep) and the energy rows; the rotation has period 2 (3 with
// AlternateStep).  So one graph is captured per pointer configuration, with
// the rows in two fixed slots that are copied out after each launch, and a
// replay restores the configuration the captured iteration left behind.
void ptr_state(mocca_solve *h, void *(&k)[4]) {
    k[0] = h->b.psi; k[1] = h->b.psi2; k[2] = h->b.upd; k[3] = h->scf.alt_old;
}

// Whether the iterations of a solve may be replayed: HF (the HFB step and
// its canonical basis keep host-side state), no profiling, and constraints
// that are never switched off (TurnOffConstraints reads the iteration
// number).
bool graph_ok(const mocca_solve *h) {
    if (h->hfb.on || h->hfb.bcs || h->prof) return false;
    for (size_t c = 0; c < h->scf.htype.size(); ++c)
        if (h->scf.htype[c] != 0 && h->scf.hiter[c] > 0) return false;
    return true;
}

}  // namespace

namespace mocca {
namespace detail {
void graphs_clear(mocca_solve *h) {
    if (!h->graphs.empty()) (void)hipStreamSynchronize(h->st);
    for (IterGraph &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}
}  // namespace detail
}  // namespace mocca

namespace {

// One replayed (or, for a new pointer configuration, captured) iteration k
// of a batch; rows: two slots [2][kRowLen] (previous, current).
int graph_step(mocca_solve *h, std::vector<IterGraph> &cache, int k, double dt, double dth,
               double momentum, double *rows, bool &ok, bool capture) {
    Scf &s = h->scf;
    void *cur[4];
    ptr_state(h, cur);
    IterGraph *g = nullptr;
    for (IterGraph &c : cache)
        if (std::equal(cur, cur + 4, c.pre)) g = &c;
    if (!g && !capture) {
        // a configuration not seen in the captured iterations: run eagerly
        ok = false;
        return 0;
    }
    if (!g) {
        IterGraph n{};
        std::copy(cur, cur + 4, n.pre);
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(h->st, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            ++h->acc["graph_capture_failed"].second;
            ok = false;
            return 0;
        }
        // (the iteration number only reaches the graph through
        // TurnOffConstraints, which graph_ok excludes)
        const int rc = iter_step(h, k, h->scf.iteration + 1, dt, dth, momentum, false, nullptr,
                                 rows, rows + kRowLen);
        const hipError_t ec = hipStreamEndCapture(h->st, &graph);
        if (rc || ec != hipSuccess || !graph ||
            hipGraphInstantiate(&n.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            h->err.clear();
            // the host state moved on as if the iteration had run: undo the
            // pointer rotation and run it eagerly
            h->b.psi = cur[0]; h->b.psi2 = cur[1]; h->b.upd = cur[2]; s.alt_old = cur[3];
            ++h->acc["graph_capture_failed"].second;
            ok = false;
            return 0;
        }
        (void)hipGraphDestroy(graph);
        ++h->acc["graph_capture"].second;
        ptr_state(h, n.post);
        cache.push_back(n);
        g = &cache.back();
    } else {
        h->b.psi = g->post[0]; h->b.psi2 = g->post[1]; h->b.upd = g->post[2];
        s.alt_old = g->post[3];
    }
    HIPCHK(h, hipGraphLaunch(g->exec, h->st));
    ++h->acc["graph_replay"].second;
    ++s.iteration;
    double *prev = s.erows + (size_t)(k - 1) * kRowLen, *row = s.erows + (size_t)k * kRowLen;
    HIPCHK(h, hipMemcpyAsync(prev, rows, kRowLen * sizeof(double), hipMemcpyDeviceToDevice,
                             h->st));
    HIPCHK(h, hipMemcpyAsync(row, rows + kRowLen, kRowLen * sizeof(double),
                             hipMemcpyDeviceToDevice, h->st));
    HIPCHK(h, hipMemcpyAsync(rows, rows + kRowLen, kRowLen * sizeof(double),
                             hipMemcpyDeviceToDevice, h->st));
    return 0;
}


}  // namespace

namespace mocca {
namespace detail {

void scf_free(mocca_solve *h) {
    Scf &s = h->scf;
    void *ptrs[] = {s.hist, s.Y, s.mpart, s.Q, s.lfac, s.bidx, s.bsh, s.brp, s.cx, s.cr, s.cpa,
                    s.cpb, s.cap, s.cpart, s.vcoul, s.vcoulx, s.cg_iters, s.pjobs, s.epart,
                    s.ecur, s.erows, s.cq[0], s.cq[1], s.cq[2], s.clam[0], s.clam[1],
                    s.clam[2], s.cw0, s.cw1, s.pold, s.pcwork, s.pcjobs, s.pc_iters, s.con,
                    s.cctl, s.csq, s.ccoef, s.calt, s.cut, s.ce, s.mult, s.cstat, s.ccnt,
                    s.csurf, s.alt_old, s.chist};
    for (void *p : ptrs) free_ptr(p);
    if (s.pot && s.pot != (double *)h->b.fields) (void)hipFree(s.pot);
    s = Scf{};
}

}  // namespace detail
}  // namespace mocca

extern "C" {

int mocca_scf_setup(mocca_solve *h, const mocca_functional *f) {
    if (h) ++h->graph_epoch;
    if (!h || !f) return fail(h, MOCCA_ERR_ARG, "null argument");
    for (int k = 20; k < 28; ++k)
        if (f->B[k] != 0.0)
            return fail(h, MOCCA_ERR_UNSUPPORTED,
                        "tensor terms (B14..B21) are not supported (B[%d] = %g)", k, f->B[k]);
    if (f->max_moment < 0 || f->max_moment > 12 || (f->max_moment & 1))
        return fail(h, MOCCA_ERR_ARG, "max_moment must be even and <= 12 (got %d)",
                    f->max_moment);
    if (f->mixing_scheme != 0 && f->mixing_scheme != 2)
        return fail(h, MOCCA_ERR_UNSUPPORTED,
                    "MixingScheme %d is not supported (0: linear mixing, 2: potential "
                    "preconditioning)", f->mixing_scheme);
    if (f->mixing_scheme == 2 && !(f->preconu > 0.0))
        return fail(h, MOCCA_ERR_ARG, "preconu must be positive");
    if (f->coulomb < 0 || f->coulomb > 2)
        return fail(h, MOCCA_ERR_UNSUPPORTED,
                    "coulomb: 0 none, 1 conjugate gradients, 2 direct (got %d)", f->coulomb);
    if (f->coulomb && f->coulomb_maxit < 1)
        return fail(h, MOCCA_ERR_ARG, "coulomb_maxit must be positive");
    if (f->cexchange < 0 || f->cexchange > 2) return fail(h, MOCCA_ERR_ARG, "cexchange 0..2");
    if (!(f->neutrons >= 0) || !(f->protons >= 0))
        return fail(h, MOCCA_ERR_ARG, "particle numbers must be non-negative");
    HIPCHK(h, hipStreamSynchronize(h->st));
    scf_free(h);
    hfb_free(h);
    Scf &s = h->scf;
    const mocca_mesh &g = h->mesh;
    const long N = h->N;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // functional
    SkyrmeK &k = s.sk;
    const double *B = f->B;
    k.B1 = B[0]; k.B2 = B[1]; k.B3 = B[2]; k.B4 = B[3]; k.B5 = B[4]; k.B6 = B[5];
    k.B7a = B[6]; k.B7b = B[7]; k.B8a = B[8]; k.B8b = B[9]; k.byt3a = B[10]; k.byt3b = B[11];
    k.B9 = B[12]; k.B9q = B[13]; k.B10 = B[14]; k.B11 = B[15]; k.B12a = B[16];
    k.B12b = B[17]; k.B13a = B[18]; k.B13b = B[19];
    for (int it = 0; it < 2; ++it) { k.hbm[it] = f->hbm[it]; k.mass[it] = f->mass[it]; }
    k.e2 = f->e2; k.neutrons = f->neutrons; k.protons = f->protons;
    k.com1body = f->com1body; k.cexchange = f->cexchange; k.coulomb = f->coulomb;
    s.damping = f->damping;
    s.mixing_scheme = f->mixing_scheme;
    s.preconu = f->preconu;
    // kinetic coefficients hbm/2 (1 - m_q / (N m_n + Z m_p)) with COM1Body = 2
    for (int it = 0; it < 2; ++it) {
        const double red = f->com1body == 2
                               ? 1.0 - f->mass[it] / (f->neutrons * f->mass[0] +
                                                      f->protons * f->mass[1])
                               : 1.0;
        h->kin[it] = f->hbm[it] / 2.0 * red;
    }
    // filling bounds (NaiveFill): pairs under time reversal
    const int trc = g.time_reversal;
    s.ub_p = trc ? (int)std::floor(f->protons / 2) : (int)std::floor(f->protons);
    s.ub_n = trc ? (int)std::floor(f->neutrons / 2) : (int)std::floor(f->neutrons);
    s.fill_fac = trc ? 2.0 : 1.0;
    int rc;
#define SCF_ALLOC(p, bytes)                                   \
    do {                                                      \
        if ((rc = dalloc(h, &(p), (bytes))) != 0) {           \
            scf_free(h);                                      \
            return rc;                                        \
        }                                                     \
    } while (0)
    // moments: even l <= max_moment, even m <= l, in IniMoments order
    s.nmom = 0;
    for (int l = 0; l <= f->max_moment; l += 2)
        for (int m = 0; m <= l; m += 2) {
            s.lm[2 * s.nmom] = l;
            s.lm[2 * s.nmom + 1] = m;
            ++s.nmom;
        }
    {
        std::vector<double> Y((size_t)s.nmom * N);
        for (int q = 0; q < s.nmom; ++q) {
            const int l = s.lm[2 * q], m = s.lm[2 * q + 1];
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        const long p = ((long)z * ny + y) * nx + x;
                        // the monopole form factor is 1/sqrt(4 pi) (IniMoments)
                        Y[q * N + p] = l == 0 ? 1.0 / std::sqrt(4.0 * kPi)
                                              : rlylm(l, m, (x + 0.5) * g.dx, (y + 0.5) * g.dx,
                                                      (z + 0.5) * g.dx);
                    }
        }
        SCF_ALLOC(s.Y, Y.size() * sizeof(double));
        HIPCHK(h, hipMemcpy(s.Y, Y.data(), Y.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    s.mom_nblk = (int)std::min<long>(256, (N + kBlock - 1) / kBlock);
    SCF_ALLOC(s.mpart, (size_t)s.mom_nblk * kMaxMom * 2 * sizeof(double));
    SCF_ALLOC(s.Q, kMaxMom * 2 * sizeof(double));
    HIPCHK(h, hipMemset(s.Q, 0, kMaxMom * 2 * sizeof(double)));
    // Coulomb box: (nx+2)(ny+2)(nz+2), boundary planes on the far faces
    const int mx = nx + 2, my = ny + 2, mz = nz + 2;
    const size_t M3 = (size_t)mx * my * mz;
    s.ck.nx = nx; s.ck.ny = ny; s.ck.nz = nz; s.ck.mx = mx; s.ck.my = my; s.ck.mz = mz;
    s.ck.f = 1.0 / (g.dx * g.dx);
    s.ck.prec = 1.0e-9 / (g.dx * g.dx * g.dx * nx * ny * nz);
    s.ck.e2 = f->e2;
    s.ck.maxit = f->coulomb_maxit;
    {
        std::vector<int> bidx;
        std::vector<double> bx;
        for (int z = 0; z < mz; ++z)
            for (int y = 0; y < my; ++y)
                for (int x = 0; x < mx; ++x)
                    if (x >= nx || y >= ny || z >= nz) bidx.push_back((z * my + y) * mx + x);
        s.nb = (int)bidx.size();
        std::vector<double> bsh((size_t)s.nmom * s.nb), brp((size_t)s.nmom * s.nb), lf(s.nmom);
        for (int q = 0; q < s.nmom; ++q) {
            const int l = s.lm[2 * q], m = s.lm[2 * q + 1];
            lf[q] = 4 * kPi / (2 * l + 1);
            for (int b = 0; b < s.nb; ++b) {
                const int i = bidx[b] % mx, j = (bidx[b] / mx) % my, kk = bidx[b] / (mx * my);
                const double X = (i + 0.5) * g.dx, Yc = (j + 0.5) * g.dx, Z = (kk + 0.5) * g.dx;
                const double r = std::sqrt(X * X + Yc * Yc + Z * Z);
                bsh[(size_t)q * s.nb + b] = rlylm(l, m, X, Yc, Z);
                brp[(size_t)q * s.nb + b] = std::pow(r, 2 * l + 1);
            }
        }
        SCF_ALLOC(s.bidx, bidx.size() * sizeof(int));
        SCF_ALLOC(s.bsh, bsh.size() * sizeof(double));
        SCF_ALLOC(s.brp, brp.size() * sizeof(double));
        SCF_ALLOC(s.lfac, lf.size() * sizeof(double));
        HIPCHK(h, hipMemcpy(s.bidx, bidx.data(), bidx.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(s.bsh, bsh.data(), bsh.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(s.brp, brp.data(), brp.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(s.lfac, lf.data(), lf.size() * sizeof(double), hipMemcpyHostToDevice));
        // InitialGuess: potential of a uniformly charged sphere of radius
        // 1.25 A^(1/3) holding the protons, on the interior of the box.
        std::vector<double> x0(M3, 0.0);
        const double R = 1.25 * std::pow(f->neutrons + f->protons, 1.0 / 3.0);
        const double Qc = f->protons * f->e2, F = Qc / (2 * R);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const double X = (x + 0.5) * g.dx, Yc = (y + 0.5) * g.dx, Z = (z + 0.5) * g.dx;
                    const double r = std::sqrt(X * X + Yc * Yc + Z * Z);
                    x0[((size_t)z * my + y) * mx + x] =
                        r <= R ? F * (3 - r * r / (R * R)) : Qc / r;
                }
        SCF_ALLOC(s.cx, M3 * sizeof(double));
        HIPCHK(h, hipMemcpy(s.cx, x0.data(), M3 * sizeof(double), hipMemcpyHostToDevice));
    }
    SCF_ALLOC(s.cr, M3 * sizeof(double));
    SCF_ALLOC(s.cpa, M3 * sizeof(double));
    SCF_ALLOC(s.cpb, M3 * sizeof(double));
    SCF_ALLOC(s.cap, M3 * sizeof(double));
    HIPCHK(h, hipMemset(s.cr, 0, M3 * sizeof(double)));
    HIPCHK(h, hipMemset(s.cpa, 0, M3 * sizeof(double)));
    HIPCHK(h, hipMemset(s.cpb, 0, M3 * sizeof(double)));
    HIPCHK(h, hipMemset(s.cap, 0, M3 * sizeof(double)));
    {
        // boxes up to kCgSingleMax interior points: one ordinary workgroup;
        // larger ones: one cooperative launch over up to 256 workgroups
        const long nint = (long)nx * ny * nz;
        if (nint <= kCgSingleMax) {
            s.cg_nwg = 1;
        } else {
            const int cap = coulomb_cg_max_groups();
            if (cap < 1) {
                scf_free(h);
                return fail(h, MOCCA_ERR_HIP,
                            "cooperative launch not available for the Coulomb solver");
            }
            s.cg_nwg = (int)std::min<long>(std::min(cap, 256), (nint + kBlock - 1) / kBlock);
        }
    }
    SCF_ALLOC(s.cpart, 2 * (size_t)s.cg_nwg * sizeof(double));
    // direct solve: eigen-decomposition of the 1-D operators
    {
        const int ns[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            std::vector<double> Qh, wh;
            jacobi_eigen(ns[a], coulomb_1d(ns[a], g.dx), Qh, wh);
            SCF_ALLOC(s.cq[a], Qh.size() * sizeof(double));
            SCF_ALLOC(s.clam[a], wh.size() * sizeof(double));
            HIPCHK(h, hipMemcpy(s.cq[a], Qh.data(), Qh.size() * sizeof(double),
                                hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(s.clam[a], wh.data(), wh.size() * sizeof(double),
                                hipMemcpyHostToDevice));
        }
        SCF_ALLOC(s.cw0, N * sizeof(double));
        SCF_ALLOC(s.cw1, N * sizeof(double));
    }
    SCF_ALLOC(s.vcoul, N * sizeof(double));
    SCF_ALLOC(s.vcoulx, N * sizeof(double));
    SCF_ALLOC(s.cg_iters, sizeof(int));
    HIPCHK(h, hipMemset(s.vcoul, 0, N * sizeof(double)));
    HIPCHK(h, hipMemset(s.vcoulx, 0, N * sizeof(double)));
    HIPCHK(h, hipMemset(s.cg_iters, 0, sizeof(int)));
    // potentials (FP64 master copy; the FP64 path uses it directly)
    if (h->prec == MOCCA_FP64) {
        s.pot = (double *)h->b.fields;
    } else {
        SCF_ALLOC(s.pot, (size_t)kNumFieldSlots * N * sizeof(double));
        HIPCHK(h, hipMemset(s.pot, 0, (size_t)kNumFieldSlots * N * sizeof(double)));
    }
    {
        std::vector<FieldJob> jobs;
        build_potential_jobs(h, jobs);
        s.npjobs = (int)jobs.size();
        SCF_ALLOC(s.pjobs, jobs.size() * sizeof(FieldJob));
        HIPCHK(h, hipMemcpy(s.pjobs, jobs.data(), jobs.size() * sizeof(FieldJob),
                            hipMemcpyHostToDevice));
    }
    if (s.mixing_scheme == 2) {
        // U (P, S, TS, comp 1) and, without time reversal, S_x (P, -S, TS, 1),
        // S_y (P, -S, TS, 2), S_z (P, S, TS, 1) for P = S = TS = 1
        // (PotentialMixing.f90:146-170); ghost signs from the component table
        // of Lapla_EV8 (OptimizedDerivatives.f90:533-542).
        static const int su[3] = {1, 1, 1};
        static const int ss[3][3] = {{-1, 1, -1}, {1, -1, -1}, {1, 1, 1}};
        SCF_ALLOC(s.pold, 8 * (size_t)N * sizeof(double));
        std::vector<PrecJob> pj;
        for (int it = 0; it < 2; ++it) {
            PrecJob j{};
            j.pot = s.pot + (kOffU + it) * N;
            j.old = s.pold + it * N;
            for (int a = 0; a < 3; ++a) j.sign[a] = su[a];
            pj.push_back(j);
        }
        if (!trc)
            for (int it = 0; it < 2; ++it)
                for (int c = 0; c < 3; ++c) {
                    PrecJob j{};
                    j.pot = s.pot + (kOffS + 3 * it + c) * N;
                    j.old = s.pold + (2 + 3 * it + c) * N;
                    for (int a = 0; a < 3; ++a) j.sign[a] = ss[c][a];
                    pj.push_back(j);
                }
        s.npc = (int)pj.size();
        SCF_ALLOC(s.pcwork, (size_t)s.npc * 4 * N * sizeof(double));
        SCF_ALLOC(s.pcjobs, pj.size() * sizeof(PrecJob));
        SCF_ALLOC(s.pc_iters, pj.size() * sizeof(int));
        HIPCHK(h, hipMemcpy(s.pcjobs, pj.data(), pj.size() * sizeof(PrecJob),
                            hipMemcpyHostToDevice));
        HIPCHK(h, hipMemset(s.pc_iters, 0, pj.size() * sizeof(int)));
    }
    s.e_nblk = (int)std::min<long>(512, (N + kBlock - 1) / kBlock);
    SCF_ALLOC(s.epart, (size_t)s.e_nblk * kEnergySlots * sizeof(double));
    SCF_ALLOC(s.ecur, kRowLen * sizeof(double));
    HIPCHK(h, hipMemset(s.ecur, 0, kRowLen * sizeof(double)));
    SCF_ALLOC(s.hist, 38 * (size_t)N * sizeof(double));
    HIPCHK(h, hipMemset(s.hist, 0, 38 * (size_t)N * sizeof(double)));
#undef SCF_ALLOC
    s.set = true;
    return 0;
}

int mocca_hf_fill(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    return fill_impl(h);
}

int mocca_get_occupations(mocca_solve *h, double *occ) {
    if (int rc = check_ready(h)) return rc;
    if (!occ) return fail(h, MOCCA_ERR_ARG, "null argument");
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(occ, h->b.occ, h->M * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int mocca_update_densities(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    if (int rc = save_history(h)) return rc;
    return density_any(h);
}

int mocca_mix_densities(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    return mix_impl(h, 0);
}

int mocca_construct_potentials(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    return construct_impl(h);
}

int mocca_comp_energy(mocca_solve *h, double *e) {
    if (int rc = check_scf(h)) return rc;
    if (!h->fields_set) return fail(h, MOCCA_ERR_STATE, "no potentials yet");
    int rc;
    if ((rc = kinetic_any(h))) return rc;
    if ((rc = energy_terms_impl(h, h->scf.ecur))) return rc;
    if ((rc = pairing_energy_impl(h, h->scf.ecur))) return rc;
    if ((rc = energy_kinetic_impl(h, h->scf.ecur))) return rc;
    double row[kRowLen];
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(row, h->scf.ecur, kRowLen * sizeof(double), hipMemcpyDeviceToHost));
    push_energy(h, row);
    if (e) std::copy(row, row + kRowLen, e);
    return 0;
}

int mocca_convergence_check(mocca_solve *h, const double *prec, int *out) {
    if (int rc = check_scf(h)) return rc;
    if (!prec || !out) return fail(h, MOCCA_ERR_ARG, "null argument");
    Scf &s = h->scf;
    // ConverMultipoleAll / ConverMultipole (Moments.f90:2628-2701): each
    // constrained moment's Value and OldValue(:,1:7), summed over isospins,
    // against sum(Constraint), relative unless sum(TrueConstraint) is 0;
    // moments without a constraint (or switched off) pass
    bool mom = true;
    if (s.ncon) {
        std::vector<ConstraintK> ck(s.ncon);
        std::vector<double> hist(16 * (size_t)s.ncon);
        HIPCHK(h, hipStreamSynchronize(h->st));
        HIPCHK(h, hipMemcpy(ck.data(), s.con, ck.size() * sizeof(ConstraintK),
                            hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(hist.data(), s.chist, hist.size() * sizeof(double),
                            hipMemcpyDeviceToHost));
        for (int c = 0; c < s.ncon && mom; ++c) {
            const ConstraintK &k = ck[c];
            if (k.type == 0) continue;
            // Isoswitch 1 keeps one target (the sum over isospins), 2 one each
            const double con = k.iso == 1 ? k.con[0] : k.con[0] + k.con[1];
            const double tru = k.iso == 1 ? k.truec[0] : k.truec[0] + k.truec[1];
            for (int i = 0; i < 8 && mom; ++i) {
                const double q = hist[16 * (size_t)c + 2 * i] + hist[16 * (size_t)c + 2 * i + 1];
                const double d = tru != 0.0 ? std::fabs((con - q) / con) : std::fabs(con - q);
                if (!(d < prec[0])) mom = false;
            }
        }
    }
    // ConverEnergy (Energy.f90:1458-1481)
    bool en = true;
    for (int i = 1; i <= 7 && en; ++i)
        en = std::fabs((s.ehist[i] - s.ehist[0]) / s.ehist[0]) <= prec[1];
    // ConverFermi (Pairing.f90:1048-1070): HF passes; with pairing each
    // Fermi energy is compared with FermiHistory(:,1:7), which the reference
    // declares zero and never updates
    bool fe = true;
    if (h->hfb.on || h->hfb.bcs)
        for (int it = 0; it < 2; ++it)
            if (std::fabs(s.efermi[it]) > prec[2]) fe = false;
    // ConverCranking (Cranking.f90:602-619): no cranking constraints
    const bool cr = true;
    out[0] = mom && en && fe && cr;
    out[1] = s.ehist[0] >= 0.0;   // Main.f90:389-396, no trap potential
    out[2] = mom;
    out[3] = en;
    out[4] = fe;
    out[5] = cr;
    return 0;
}

int mocca_final_iteration(mocca_solve *h, double *row) {
    if (int rc = check_scf(h)) return rc;
    if (!h->scf.started) return fail(h, MOCCA_ERR_STATE, "mocca_scf_start not called");
    if (!h->fields_set) return fail(h, MOCCA_ERR_STATE, "no potentials yet");
    const bool fd = !h->mesh.lagrange;
    // AssignFDCoefs(-1, -1, CoulombLapOrder): Lagrange first and second
    // derivatives for the states and the density fields (the Coulomb solve
    // is not repeated)
    int rc = set_scheme(h, 1);
    if (rc) return rc;
    auto lagrange_pass = [&]() -> int {
        int r;
        if ((r = density_any(h))) return r;
        if ((r = kinetic_any(h))) return r;
        if ((r = energy_terms_impl(h, h->scf.ecur))) return r;
        if ((r = pairing_energy_impl(h, h->scf.ecur))) return r;
        return energy_kinetic_impl(h, h->scf.ecur);
    };
    rc = lagrange_pass();
    if (fd) {
        const int r2 = set_scheme(h, 0);
        if (!rc) rc = r2;
    }
    if (rc) return rc;
    double e[kRowLen];
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(e, h->scf.ecur, kRowLen * sizeof(double), hipMemcpyDeviceToHost));
    push_energy(h, e);
    if (row) std::copy(e, e + kRowLen, row);
    // the densities of the final states with the run's own derivatives,
    // unmixed (UpdateDensities(0) after AssignFDCoefs(MaxFDOrder, ...))
    return fd ? density_any(h) : 0;
}

int mocca_scf_start(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    int rc;
    if ((rc = fill_impl(h))) return rc;
    // densities of the states, or those of a wavefunction file read without
    // Recalc (Main.f90:206-215)
    if (!h->scf.dens_loaded && (rc = density_any(h))) return rc;
    h->scf.dens_loaded = false;
    if ((rc = save_history(h))) return rc;
    if ((rc = construct_impl(h, 3))) return rc;
    if ((rc = energy_terms_impl(h, h->scf.ecur))) return rc;
    if (h->hfb.bcs && (rc = bcs_energy_impl(h, h->scf.ecur))) return rc;
    if (h->hfb.on) {
        if ((rc = hfb_energy_impl(h, h->scf.ecur))) return rc;
        if ((rc = kinetic_any(h))) return rc;
        if ((rc = energy_kinetic_impl(h, h->scf.ecur))) return rc;
    }
    h->scf.alt_next = alternating(h);
    h->scf.started = true;
    h->scf.iteration = 0;
    // the start-up row enters the energy history once it is complete: at
    // once with HFB, after the first iteration's hPsi pass with HF (the
    // kinetic part of an HF row comes from the next GradDesc)
    h->scf.start_pending = !h->hfb.on;
    if (h->hfb.on) {
        double row[kRowLen];
        HIPCHK(h, hipStreamSynchronize(h->st));
        HIPCHK(h, hipMemcpy(row, h->scf.ecur, kRowLen * sizeof(double), hipMemcpyDeviceToHost));
        push_energy(h, row);
    }
    return 0;
}

int mocca_iterate(mocca_solve *h, int niter, double dt_over_hbar, double momentum,
                  double *energies) {
    if (!h) return MOCCA_ERR_ARG;
    return iterate_impl(h, niter, h->scf.iteration + 1, dt_over_hbar * kHbar, dt_over_hbar,
                        momentum, false, energies, nullptr);
}

int mocca_iterate_estimated(mocca_solve *h, int first_iteration, int niter, double dt,
                            double momentum, double *energies, double *params) {
    if (first_iteration < 1) return fail(h, MOCCA_ERR_ARG, "first_iteration must be >= 1");
    return iterate_impl(h, niter, first_iteration, dt, dt / kHbar, momentum, true, energies,
                        params);
}

int mocca_batch_run(mocca_solve *const *solves, int nsolves, int niter,
                    const double *dt_over_hbar, const double *momentum, double *const *energies) {
    if (!solves || nsolves < 1 || !dt_over_hbar || !momentum) return MOCCA_ERR_ARG;
    for (int i = 0; i < nsolves; ++i)
        if (!solves[i]) return MOCCA_ERR_ARG;
    std::vector<double> dt(nsolves), dth(nsolves), mu(nsolves);
    for (int i = 0; i < nsolves; ++i) {
        dth[i] = dt_over_hbar[i];
        dt[i] = dth[i] * kHbar;
        mu[i] = momentum[i];
    }
    // every solve's rows first, then iteration k of every solve in turn (each
    // on its own stream, so the device overlaps them), then the final passes
    // and checks; the first failure stops the batch and is reported on the
    // failing solve
    // (a failure is also reported as mocca_last_error(NULL), naming the solve)
    auto failed = [&](int i, int rc) {
        return fail(nullptr, rc, "batch solve %d: %s", i, solves[i]->err.c_str());
    };
    for (int i = 0; i < nsolves; ++i)
        if (int rc = iter_begin(solves[i], niter)) return failed(i, rc);
    // Iteration 1 of every solve runs eagerly (it sets up what later
    // iterations reuse); later ones are graph replays where graph_ok holds.
    // The captured iterations stay with the solve for later calls.
    //
    // Every allocation and every stream capture happens on this thread,
    // before any worker starts: the buffers an iteration may need (the
    // row slots of the replays, the states before an AlternateStep) are
    // allocated here, and the iterations in which a solve's graphs are
    // captured (one per configuration of the rotating buffers, so the first
    // kWarm) are issued here too, as is the first iteration of a solve that
    // has not run in a batch before (whatever it sets up lazily).  The
    // workers then only replay captured graphs or run iterations eagerly,
    // while no capture is open anywhere.
    constexpr int kWarm = 4;     // 1 eager + one capture per buffer configuration (period <= 3)
    std::vector<char> use_graph(nsolves, 0);
    for (int i = 0; i < nsolves; ++i) {
        mocca_solve *h = solves[i];
        if (alternating(h))
            if (int rc = alternate_prepare(h)) return failed(i, rc);
        if (niter >= 3 && graph_ok(h)) {
            if (h->graphs_epoch != h->graph_epoch || h->graphs_dth != dth[i] ||
                h->graphs_mu != mu[i])
                graphs_clear(h);
            if (!h->graph_rows) {
                if (hipMalloc(&h->graph_rows, 2 * kRowLen * sizeof(double)) != hipSuccess) {
                    h->graph_rows = nullptr;
                    (void)hipGetLastError();
                }
                ++h->graph_epoch;
            }
            use_graph[i] = h->graph_rows != nullptr;
            if (use_graph[i]) {
                h->graphs_epoch = h->graph_epoch;
                h->graphs_dth = dth[i];
                h->graphs_mu = mu[i];
            }
        }
    }
    // one iteration k of solve i; capture = whether a graph may be captured
    auto step = [&](int i, int k, bool capture) -> int {
        mocca_solve *h = solves[i];
        if (use_graph[i] && k == 2) {
            // both slots start as row 1 (entries an iteration does not
            // write, such as the HF pairing terms, keep their values)
            if (hipMemcpyAsync(h->graph_rows, h->scf.erows + kRowLen, kRowLen * sizeof(double),
                               hipMemcpyDeviceToDevice, h->st) != hipSuccess ||
                hipMemcpyAsync(h->graph_rows + kRowLen, h->scf.erows + kRowLen,
                               kRowLen * sizeof(double), hipMemcpyDeviceToDevice,
                               h->st) != hipSuccess) {
                (void)hipGetLastError();
                use_graph[i] = 0;
            }
        }
        if (use_graph[i] && k >= 2) {
            bool ok = true;
            const int rc = graph_step(h, h->graphs, k, dt[i], dth[i], mu[i], h->graph_rows, ok,
                                      capture);
            if (rc || ok) return rc;
            use_graph[i] = 0;   // capture refused or not allowed here: run eagerly
            if (capture) graphs_clear(h);
        }
        const int rc = iter_step(h, k, h->scf.iteration + 1, dt[i], dth[i], mu[i], false,
                                 nullptr, h->scf.erows + (size_t)(k - 1) * kRowLen,
                                 h->scf.erows + (size_t)k * kRowLen);
        if (!rc) ++h->scf.iteration;
        return rc;
    };
    // whether the graphs a solve captured in earlier calls cover every
    // pointer configuration from its current one on (a closed cycle), so
    // that its later iterations need no capture
    auto covered = [&](int i) {
        const mocca_solve *h = solves[i];
        void *cur[4], *c[4];
        ptr_state(const_cast<mocca_solve *>(h), cur);
        std::copy(cur, cur + 4, c);
        for (size_t n = 0; n < h->graphs.size(); ++n) {
            const IterGraph *g = nullptr;
            for (const IterGraph &x : h->graphs)
                if (std::equal(c, c + 4, x.pre)) g = &x;
            if (!g) return false;
            std::copy(g->post, g->post + 4, c);
            if (std::equal(c, c + 4, cur)) return true;
        }
        return false;
    };
    const int kw = std::min(niter, kWarm);
    // first iteration a worker issues: 1 for solves whose earlier calls
    // captured every graph they need (nothing left to capture or
    // allocate); otherwise this thread runs iteration 1 and, for graph
    // solves, the capturing iterations up to kw
    std::vector<int> k0(nsolves, kw + 1);
    for (int i = 0; i < nsolves; ++i)
        if (use_graph[i] && covered(i)) k0[i] = 1;
    for (int i = 0; i < nsolves; ++i)
        if (k0[i] > 1)
            if (int rc = step(i, 1, true)) return failed(i, rc);
    for (int i = 0; i < nsolves; ++i)
        if (k0[i] > 1 && (!use_graph[i] || covered(i))) k0[i] = 2;
    for (int k = 2; k <= kw; ++k)
        for (int i = 0; i < nsolves; ++i)
            if (k0[i] > kw)
                if (int rc = step(i, k, true)) return failed(i, rc);
    // The remaining iterations, split over a few host threads, each issuing
    // its share of the solves in turn: one thread cannot issue the kernels
    // of many small solves as fast as the device runs them.
    auto run_range = [&](int i0, int i1, int &rc, int &bad) {
        for (int k = 1; k <= niter && !rc; ++k)
            for (int i = i0; i < i1 && !rc; ++i)
                if (k >= k0[i] && (rc = step(i, k, false))) bad = i;
    };
    const int nthr = std::max(1, std::min({nsolves, 8,
                                           (int)std::thread::hardware_concurrency() / 2}));
    std::vector<int> trc(nthr, 0), tbad(nthr, -1);
    std::vector<std::thread> pool;
    if (niter >= 1) {
        for (int t = 1; t < nthr; ++t)
            pool.emplace_back(run_range, nsolves * t / nthr, nsolves * (t + 1) / nthr,
                              std::ref(trc[t]), std::ref(tbad[t]));
        run_range(0, nsolves / nthr, trc[0], tbad[0]);
        for (std::thread &th : pool) th.join();
    }
    int rc = 0, bad = -1;
    for (int t = 0; t < nthr && !rc; ++t)
        if (trc[t]) { rc = trc[t]; bad = tbad[t]; }
    if (rc) return failed(bad, rc);
    int first = 0;
    for (int i = 0; i < nsolves; ++i) {
        const int r = iter_end(solves[i], niter, energies ? energies[i] : nullptr);
        if (r && !first) first = failed(i, r);
    }
    return first;
}

int mocca_estimate_parameters(mocca_solve *h, int iteration, double dt, double momentum,
                              double *out) {
    if (int rc = check_ready(h)) return rc;
    if (!h->fields_set) return fail(h, MOCCA_ERR_STATE, "no potentials set");
    if (!out) return fail(h, MOCCA_ERR_ARG, "null output");
    if (iteration < 1) return fail(h, MOCCA_ERR_ARG, "iteration must be >= 1");
    return estimate_any(h, iteration, dt, momentum, out);
}

int mocca_estimation_seed(mocca_solve *h, const double *grid) {
    if (int rc = check_ready(h)) return rc;
    if (!grid) {
        h->est.seed.clear();
        return 0;
    }
    h->est.seed.assign(grid, grid + 4 * (size_t)h->N);
    return 0;
}



int mocca_get_moments(mocca_solve *h, int *lm, double *q, int *count) {
    if (int rc = check_scf(h)) return rc;
    const Scf &s = h->scf;
    if (count) *count = s.nmom;
    if (lm) std::memcpy(lm, s.lm, 2 * s.nmom * sizeof(int));
    if (q) {
        HIPCHK(h, hipStreamSynchronize(h->st));
        HIPCHK(h, hipMemcpy(q, s.Q, 2 * s.nmom * sizeof(double), hipMemcpyDeviceToHost));
    }
    return 0;
}

int mocca_get_coulomb(mocca_solve *h, double *vcoul, double *vcoulx, double *cpot,
                      int *cg_iterations) {
    if (int rc = check_scf(h)) return rc;
    const Scf &s = h->scf;
    const size_t N = h->N, M3 = (size_t)s.ck.mx * s.ck.my * s.ck.mz;
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (vcoul) HIPCHK(h, hipMemcpy(vcoul, s.vcoul, N * sizeof(double), hipMemcpyDeviceToHost));
    if (vcoulx) HIPCHK(h, hipMemcpy(vcoulx, s.vcoulx, N * sizeof(double), hipMemcpyDeviceToHost));
    if (cpot) HIPCHK(h, hipMemcpy(cpot, s.cx, M3 * sizeof(double), hipMemcpyDeviceToHost));
    if (cg_iterations)
        HIPCHK(h, hipMemcpy(cg_iterations, s.cg_iters, sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int mocca_set_constraints(mocca_solve *h, const mocca_constraint *c, int count, double radd,
                          double acut, double damping, double readjust_slowdown) {
    if (h) ++h->graph_epoch;
    if (int rc = check_scf(h)) return rc;
    if (count < 0 || (count > 0 && !c)) return fail(h, MOCCA_ERR_ARG, "bad constraint list");
    if (!(acut > 0.0)) return fail(h, MOCCA_ERR_ARG, "acut must be positive");
    Scf &s = h->scf;
    const long N = h->N;
    std::vector<ConstraintK> ck;
    for (int k = 0; k < count; ++k) {
        const mocca_constraint &q = c[k];
        int idx = -1;
        for (int j = 0; j < s.nmom; ++j)
            if (s.lm[2 * j] == q.l && s.lm[2 * j + 1] == q.m) idx = j;
        if (idx < 0)
            return fail(h, MOCCA_ERR_UNSUPPORTED,
                        "no moment (l=%d, m=%d) under the symmetries (even l, even m <= l <= "
                        "max_moment)", q.l, q.m);
        if (q.type != 1 && q.type != 2)
            return fail(h, MOCCA_ERR_UNSUPPORTED, "ConstraintType %d (1 or 2)", q.type);
        if (q.isoswitch != 1 && q.isoswitch != 2)
            return fail(h, MOCCA_ERR_UNSUPPORTED, "Isoswitch %d (1 or 2)", q.isoswitch);
        for (const ConstraintK &o : ck)
            if (o.idx == idx)
                return fail(h, MOCCA_ERR_ARG, "two constraints on (l=%d, m=%d)", q.l, q.m);
        ConstraintK r{};
        r.idx = idx;
        r.type = q.type;
        r.iso = q.isoswitch;
        r.iteration = q.iteration;
        // Constraint / sqrt(4 pi) for the monopole (Moments.f90:3719-3720)
        const double f = q.l == 0 ? 1.0 / std::sqrt(4.0 * kPi) : 1.0;
        if (q.isoswitch == 1) {
            r.con[0] = r.con[1] = q.value[0] * f;
        } else {
            r.con[0] = q.value[0];
            r.con[1] = q.value[1];
        }
        r.truec[0] = r.con[0];
        r.truec[1] = r.con[1];
        r.inten[0] = r.inten[1] = q.intensity;
        ck.push_back(r);
    }
    // the order of the moment list (IniMoments, Moments.f90:372-485)
    std::sort(ck.begin(), ck.end(),
              [](const ConstraintK &a, const ConstraintK &b) { return a.idx < b.idx; });
    HIPCHK(h, hipStreamSynchronize(h->st));
    void *old[] = {s.con, s.cctl, s.csq, s.ccoef, s.calt, s.cut, s.ce, s.mult, s.cstat, s.ccnt,
                   s.csurf, s.chist};
    for (void *p : old) free_ptr(p);
    s.con = nullptr; s.cctl = nullptr; s.csq = s.ccoef = s.calt = s.chist = nullptr;
    s.cut = s.ce = s.mult = s.cstat = s.csurf = nullptr; s.ccnt = nullptr;
    s.ncon = (int)ck.size();
    s.htype.clear();
    s.hiter.clear();
    if (!s.ncon) return 0;
    for (const ConstraintK &r : ck) {
        s.htype.push_back(r.type);
        s.hiter.push_back(r.iteration);
    }
    const ConstraintCtl ctl{damping, readjust_slowdown, 1, 0};
    s.radd = radd;
    s.acut = acut;
    int rc;
    if ((rc = dalloc(h, &s.con, ck.size() * sizeof(ConstraintK))) ||
        (rc = dalloc(h, &s.cctl, sizeof(ConstraintCtl))) ||
        (rc = dalloc(h, &s.csq, 2 * ck.size() * sizeof(double))) ||
        (rc = dalloc(h, &s.ccoef, 2 * ck.size() * sizeof(double))) ||
        (rc = dalloc(h, &s.calt, 2 * ck.size() * sizeof(double))) ||
        (rc = dalloc(h, &s.cut, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &s.ce, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &s.mult, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &s.cstat, 4 * sizeof(double))) ||
        (rc = dalloc(h, &s.ccnt, 2 * sizeof(int))) ||
        (rc = dalloc(h, &s.csurf, 2 * 3 * (size_t)N * 3 * sizeof(double))) ||
        (rc = dalloc(h, &s.chist, 16 * ck.size() * sizeof(double))))
        return rc;
    HIPCHK(h, hipMemset(s.chist, 0, 16 * ck.size() * sizeof(double)));
    HIPCHK(h, hipMemcpy(s.con, ck.data(), ck.size() * sizeof(ConstraintK), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(s.cctl, &ctl, sizeof ctl, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(s.ce, 0, 2 * N * sizeof(double)));
    HIPCHK(h, hipMemset(s.cut, 0, 2 * N * sizeof(double)));
    HIPCHK(h, hipMemset(s.csq, 0, 2 * ck.size() * sizeof(double)));
    HIPCHK(h, hipMemset(s.calt, 0, 2 * ck.size() * sizeof(double)));
    HIPCHK(h, hipMemset(s.ccoef, 0, 2 * ck.size() * sizeof(double)));
    return 0;
}

int mocca_get_constraints(mocca_solve *h, double *out) {
    if (int rc = check_scf(h)) return rc;
    Scf &s = h->scf;
    if (!s.ncon) return 0;
    if (!out) return fail(h, MOCCA_ERR_ARG, "null output");
    std::vector<ConstraintK> ck(s.ncon);
    std::vector<double> sq(2 * s.ncon);
    HIPCHK(h, hipStreamSynchronize(h->st));
    HIPCHK(h, hipMemcpy(ck.data(), s.con, ck.size() * sizeof(ConstraintK), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(sq.data(), s.csq, sq.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < s.ncon; ++c) {
        double *o = out + 10 * c;
        o[0] = s.lm[2 * ck[c].idx];
        o[1] = s.lm[2 * ck[c].idx + 1];
        o[2] = ck[c].type;
        o[3] = ck[c].con[0];
        o[4] = ck[c].con[1];
        o[5] = ck[c].inten[0];
        o[6] = ck[c].inten[1];
        o[7] = sq[2 * c];
        o[8] = sq[2 * c + 1];
        o[9] = ck[c].iso;
    }
    return 0;
}

int mocca_alternate_step(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    if (!h->scf.ncon) return 0;
    return alternate_any(h);
}

int mocca_readjust_constraints(mocca_solve *h, int iteration, double momentum) {
    if (int rc = check_scf(h)) return rc;
    return readjust_impl(h, iteration, momentum);
}

int mocca_set_pairing(mocca_solve *h, const mocca_pairing *p) {
    if (h) ++h->graph_epoch;
    if (int rc = check_scf(h)) return rc;
    if (!p) return fail(h, MOCCA_ERR_ARG, "null pairing parameters");
    HIPCHK(h, hipStreamSynchronize(h->st));
    return hfb_setup(h, p);
}

int mocca_solve_pairing(mocca_solve *h) {
    if (int rc = check_scf(h)) return rc;
    int rc = fill_impl(h);
    if (rc) return rc;
    return h->hfb.on ? hfb_check_status(h) : 0;
}

int mocca_get_pairing(mocca_solve *h, double *out, double *canonical_occ) {
    if (int rc = check_scf(h)) return rc;
    Hfb &p = h->hfb;
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (out) {
        std::vector<double> row(kRowLen, 0.0);
        if (p.on || p.bcs) {
            double *tmp = nullptr;
            if (int rc = dalloc(h, &tmp, kRowLen * sizeof(double))) return rc;
            int rc = pairing_energy_impl(h, tmp);
            if (!rc) {
                (void)hipMemcpyAsync(row.data(), tmp, kRowLen * sizeof(double),
                                     hipMemcpyDeviceToHost, h->st);
                (void)hipStreamSynchronize(h->st);
            }
            (void)hipFree(tmp);
            if (rc || (rc = hfb_ensure_qpe(h))) return rc;
            HIPCHK(h, hipMemcpy(out + 4, p.qpe, h->M * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            for (int i = 0; i < h->M; ++i) out[4 + i] = 0.0;
        }
        out[0] = row[45];
        out[1] = row[46];
        out[2] = row[43];
        out[3] = row[44];
    }
    if (canonical_occ) {
        HIPCHK(h, hipMemcpy(canonical_occ, p.on ? p.can_occ : h->b.occ, h->M * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    return 0;
}

int mocca_get_potentials(mocca_solve *h, double *U, double *B, double *NablaB, double *W,
                         double *DivW, double *S, double *A, double *DivA) {
    if (int rc = check_scf(h)) return rc;
    const size_t N = h->N;
    HIPCHK(h, hipStreamSynchronize(h->st));
    struct { double *dst; int off, n; } parts[] = {
        {U, kOffU, 2}, {B, kOffB, 2}, {NablaB, kOffNB, 6}, {W, kOffW, 18}, {DivW, kOffDW, 6},
        {S, kOffS, 6}, {A, kOffA, 6}, {DivA, kOffDA, 2}};
    for (auto &q : parts)
        if (q.dst)
            HIPCHK(h, hipMemcpy(q.dst, h->scf.pot + q.off * N, q.n * N * sizeof(double),
                                hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
