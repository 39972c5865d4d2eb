/*
 * CPU restatement ("oracle") of the self-consistent part of the MOCCa
 * iteration: multipole moments, Coulomb Poisson solve, mean-field potentials,
 * density mixing, HF filling and the energy.  TEST INFRASTRUCTURE ONLY (see
 * mocca_oracle.h).  Scalar C, one thread, reference operation order where
 * cheap.  Pinned by the golden vectors of the reference's own modules
 * (tests/test_oracle_scf.py).
 */
#include "mocca_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Force coefficients in the order Force.f90:38-41 declares them. */
enum { B1, B2, B3, B4, B5, B6, B7a, B7b, B8a, B8b, Byt3a, Byt3b, B9, B9q, B10,
       B11, B12a, B12b, B13a, B13b, B14, B15, B16, B17, B18, B19, B20, B21 };

static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* Associated Legendre functions P_n^m(x), n = m..nmax, with the recurrence of
 * legendre_pnm (Mesh.f90:92-268), including its (-1)^m phase.               */
static void legendre_pnm(int nmax, int m, double x, double *cx)
{
    for (int i = 0; i < m; ++i) cx[i] = 0.0;
    cx[m] = 1.0;
    const double somx2 = sqrt(1.0 - x * x);
    double fact = 1.0;
    for (int i = 1; i <= m; ++i) {
        cx[m] = -cx[m] * fact * somx2;
        fact = fact + 2.0;
    }
    if (m == nmax) return;
    cx[m + 1] = x * (double)(2 * m + 1) * cx[m];
    for (int i = m + 2; i <= nmax; ++i)
        cx[i] = ((double)(2 * i - 1) * x * cx[i - 1] - (double)(i + m - 1) * cx[i - 2]) /
                (double)(i - m);
}

/* r^l Re Y_lm at (X, Y, Z), quantisation axis z, phi = atan2(Y, X)
 * (SphericalHarmonics, Moments.f90:584-691). */
double or_rlylm(int l, int m, double X, double Y, double Z)
{
    double cx[32];
    const double r = sqrt(X * X + Y * Y + Z * Z);
    const double cost = Z / r;
    const double phi = atan2(Y, X);
    legendre_pnm(l, m, cost, cx);
    double fq = 1.0;
    for (int q = 1; q <= 2 * m; ++q) fq = fq / (double)(l + m - q + 1);
    double fac = ((m % 2) ? -1.0 : 1.0) * sqrt(fq) * pow(r, l);
    fac = fac * sqrt((double)(2 * l + 1) / (4 * PI));
    return fac * cx[l] * cos(m * phi);
}

/* Electric multipole moments present in the EV8 octant (all symmetries
 * conserved, quantisation axis z): even l <= lmax, even m <= l, real parts,
 * in the linked-list order of IniMoments (Moments.f90:372-485).  Returns the
 * count; lm[2k] = l, lm[2k+1] = m; val[2k+it] = dv sum rho_it r^l Y_lm
 * (Calculate_Multipole, Moments.f90:1907-2004). */
int or_moments(const double *rho, int nx, int ny, int nz, double dx, double dv,
               int lmax, int *lm, double *val)
{
    int cnt = 0;
    for (int l = 0; l <= lmax; l += 2)
        for (int m = 0; m <= l; m += 2) {
            lm[2 * cnt] = l;
            lm[2 * cnt + 1] = m;
            for (int it = 0; it < 2; ++it) {
                double s = 0.0;
                const double *r = rho + (size_t)it * nx * ny * nz;
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            const double X = (i + 0.5) * dx, Y = (j + 0.5) * dx,
                                         Z = (k + 0.5) * dx;
                            const double y = (l == 0) ? 1.0 / sqrt(4.0 * PI)
                                                      : or_rlylm(l, m, X, Y, Z);
                            s = s + y * r[((size_t)k * ny + j) * nx + i];
                        }
                val[2 * cnt + it] = s * dv;
            }
            ++cnt;
        }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Coulomb: box of (nx+2)(ny+2)(nz+2) points (BCNumber = 2 boundary planes on
 * the far faces, Coulomb.f90:348-540), multipole Dirichlet values on the
 * boundary (CoulombBound :583-674), conjugate gradients with the 5-point
 * second derivative per axis (ConjugGrad :755-900, CoulombLaplacian and
 * Coulomb_1D, CoulombDerivatives.f90:144-281).                              */
static void coul_lap(const double *g, double *out, int mx, int my, int mz, int nx,
                     int ny, int nz, double dx)
{
    static const double c[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
    const double f = 1.0 / (dx * dx);
#define G(i, j, k) g[((size_t)(k) * my + (j)) * mx + (i)]
    memset(out, 0, sizeof(double) * (size_t)mx * my * mz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double dxx = 0, dyy = 0, dzz = 0;
                for (int q = -2; q <= 2; ++q) {
                    int a = i + q;  /* mirror with sign +1 below the plane */
                    dxx = dxx + c[q + 2] * (a < 0 ? G(-1 - a, j, k) : G(a, j, k));
                    a = j + q;
                    dyy = dyy + c[q + 2] * (a < 0 ? G(i, -1 - a, k) : G(i, a, k));
                    a = k + q;
                    dzz = dzz + c[q + 2] * (a < 0 ? G(i, j, -1 - a) : G(i, j, a));
                }
                out[((size_t)k * my + j) * mx + i] = (dxx * f + dyy * f) + dzz * f;
            }
#undef G
}

/* One SolveCoulomb.  cpot [mz][my][mx] is the persistent complete potential
 * (initial guess on the first call when *first != 0).  rho_p [N]. */
int or_coulomb(const double *rho_p, int nx, int ny, int nz, double dx, double e2,
               double nucleons, double protons, int nmom, const int *lm,
               const double *qp, double *cpot, int first, double *vcoul,
               double *vcoulx, int maxit)
{
    const int mx = nx + 2, my = ny + 2, mz = nz + 2;
    const size_t M3 = (size_t)mx * my * mz;
#define IX(i, j, k) (((size_t)(k) * my + (j)) * mx + (i))
    if (first) {
        memset(cpot, 0, sizeof(double) * M3);
        const double R = 1.25 * pow(nucleons, 1.0 / 3.0);
        const double Q = protons * e2, F = Q / (2 * R);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double X = (i + 0.5) * dx, Y = (j + 0.5) * dx, Z = (k + 0.5) * dx;
                    const double r = sqrt(X * X + Y * Y + Z * Z);
                    cpot[IX(i, j, k)] = r <= R ? F * (3 - r * r / (R * R)) : Q / r;
                }
    }
    double *bc = calloc(M3, sizeof(double));
    for (int q = 0; q < nmom; ++q) {
        const int l = lm[2 * q], m = lm[2 * q + 1];
        const double value = qp[q] * (4 * PI / (2 * l + 1));
        for (int k = 0; k < mz; ++k)
            for (int j = 0; j < my; ++j)
                for (int i = 0; i < mx; ++i) {
                    if (i < nx && j < ny && k < nz) continue;
                    const double X = (i + 0.5) * dx, Y = (j + 0.5) * dx, Z = (k + 0.5) * dx;
                    const double r = sqrt(X * X + Y * Y + Z * Z);
                    const double y = (l == 0) ? or_rlylm(0, 0, X, Y, Z) : or_rlylm(l, m, X, Y, Z);
                    bc[IX(i, j, k)] = bc[IX(i, j, k)] + value * y * e2 / pow(r, 2 * l + 1);
                }
    }
    for (size_t p = 0; p < M3; ++p)
        if (bc[p] != 0.0) cpot[p] = bc[p];
    free(bc);
    double *src = calloc(M3, sizeof(double)), *res = malloc(sizeof(double) * M3),
           *pk = malloc(sizeof(double) * M3), *tmp = malloc(sizeof(double) * M3);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                src[IX(i, j, k)] = -4 * PI * e2 * rho_p[((size_t)k * ny + j) * nx + i];
    const double prec = 1.0e-9 / (dx * dx * dx * nx * ny * nz);
    coul_lap(cpot, tmp, mx, my, mz, nx, ny, nz, dx);
    double norm = 0;
    for (size_t p = 0; p < M3; ++p) {
        res[p] = -1.0 * tmp[p];
        res[p] = res[p] + src[p];
        pk[p] = res[p];
        norm = norm + res[p] * res[p];
    }
    int it;
    for (it = 1; it <= maxit; ++it) {
        coul_lap(pk, tmp, mx, my, mz, nx, ny, nz, dx);
        double integral = 0;
        for (size_t p = 0; p < M3; ++p) integral = integral + res[p] * tmp[p];
        const double a = norm / integral;
        double nn = 0;
        for (size_t p = 0; p < M3; ++p) {
            cpot[p] = cpot[p] + a * pk[p];
            res[p] = res[p] - a * tmp[p];
            nn = nn + res[p] * res[p];
        }
        if (nn <= prec) break;
        const double c = nn / norm;
        norm = nn;
        for (size_t p = 0; p < M3; ++p) pk[p] = res[p] + c * pk[p];
    }
    const double fx = -pow(3 / PI, 1.0 / 3.0) * e2;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t p = ((size_t)k * ny + j) * nx + i;
                vcoul[p] = cpot[IX(i, j, k)];
                vcoulx[p] = fx * pow(rho_p[p], 1.0 / 3.0);
            }
    free(src); free(res); free(pk); free(tmp);
#undef IX
    return it;
}

/* ------------------------------------------------------------------------ */
/* Mean-field potentials of an SLy4-type functional (no tensor, N2LO or
 * constraint terms): CalcUPot (MeanFields.f90:269-409), CalcWPot (:494-597),
 * CalcBPot (:115-150), CalcSPot (:618-712), CalcAPot (:411-492).
 * dens: rho, tau, nablaJ, laprho [2][N]; derrho [2][3][N]; when trc == 0 also
 * vecj, vecs, rotvecj, rots [2][3][N].  Outputs in the layouts of or_hpsi. */
void or_potentials(const double *coef, int nx, int ny, int nz, double dx, int lag,
                   const double *lagx, const double *lagy, const double *lagz, int trc,
                   const double *rho, const double *tau, const double *nablaJ,
                   const double *derrho, const double *laprho, const double *vecj,
                   const double *vecs, const double *rotvecj, const double *rots,
                   const double *vcoul, const double *vcoulx, int cexchange,
                   double *U, double *Bp, double *NB, double *W, double *DW,
                   double *S, double *A, double *DA)
{
    const size_t n = (size_t)nx * ny * nz;
    const double *c = coef;
    const double eps = 1.0e-20, a3 = c[Byt3a];
    const double *lm[3] = {lagx, lagy, lagz};
    /* CalcUPot */
    for (size_t p = 0; p < n; ++p) {
        const double rt = rho[p] + rho[n + p];
        const double njt = nablaJ[p] + nablaJ[n + p];
        U[p] = (2.0 * c[B1] + 2.0 * c[B2]) * rho[p] + 2.0 * c[B1] * rho[n + p];
        U[n + p] = (2.0 * c[B1] + 2.0 * c[B2]) * rho[n + p] + 2.0 * c[B1] * rho[p];
        for (int it = 0; it < 2; ++it) {
            const int at = 1 - it;
            const double ri = rho[it * n + p], ra = rho[at * n + p];
            U[it * n + p] = U[it * n + p] + (c[B3] + c[B4]) * tau[it * n + p] +
                            c[B3] * tau[at * n + p] +
                            2.0 * ((c[B5] + c[B6]) * laprho[it * n + p] + c[B5] * laprho[at * n + p]) +
                            (2 + a3) * c[B7a] * pow(rt + eps, 1 + a3) +
                            c[B8a] * (a3 * pow(rt, a3) / (rt + eps) * (ri * ri + ra * ra) +
                                      2.0 * pow(rt, a3) * ri) +
                            c[B9] * njt + c[B9q] * nablaJ[it * n + p];
            if (c[B7b] != 0.0 || c[B8b] != 0.0) {
                const double a3b = c[Byt3b];
                U[it * n + p] = U[it * n + p] + (2 + a3b) * c[B7b] * pow(rt + eps, 1 + a3b) +
                                c[B8b] * (a3b * pow(rt, a3b) / (rt + eps) * (ri * ri + ra * ra) +
                                          2.0 * pow(rt, a3b) * ri);
            }
        }
        if (!trc) {
            double st2 = 0, s2[2] = {0, 0};
            for (int k = 0; k < 3; ++k) {
                const double s0 = vecs[k * n + p], s1 = vecs[(3 + k) * n + p];
                st2 = st2 + (s0 + s1) * (s0 + s1);
                s2[0] = s2[0] + s0 * s0;
                s2[1] = s2[1] + s1 * s1;
            }
            for (int it = 0; it < 2; ++it)
                U[it * n + p] = U[it * n + p] + a3 * pow(rt, a3) / (rt + eps) *
                                (c[B12a] * st2 + c[B13a] * s2[it] + c[B13a] * s2[1 - it]);
        }
        U[n + p] = U[n + p] + vcoul[p];
        if (cexchange == 2) U[n + p] = U[n + p] + vcoulx[p];
    }
    /* CalcWPot: W_{mu nu} = -sum_k eps_{k mu nu} ((B9+B9q) d_k rho_q + B9 d_k rho_q') */
    static const int lc[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}},
                                    {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
                                    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
    for (int it = 0; it < 2; ++it)
        for (int mu = 0; mu < 3; ++mu)
            for (int nu = 0; nu < 3; ++nu) {
                double *w = W + ((size_t)(it * 3 + nu) * 3 + mu) * n;
                for (size_t p = 0; p < n; ++p) {
                    double acc = 0.0;
                    for (int k = 0; k < 3; ++k)
                        acc = acc - lc[k][mu][nu] * ((c[B9] + c[B9q]) * derrho[(it * 3 + k) * n + p] +
                                                     c[B9] * derrho[((1 - it) * 3 + k) * n + p]);
                    w[p] = acc;
                }
            }
    /* DivWPot (MeanFields.f90:580-592): per nu, the (P,S,TS,comp) of each term */
    static const int dwq[3][3][4] = {
        {{-1, 1, 1, 2}, {-1, 1, 1, 1}, {-1, -1, 1, 2}},
        {{-1, 1, 1, 1}, {-1, 1, 1, 2}, {-1, -1, 1, 1}},
        {{-1, -1, 1, 2}, {-1, -1, 1, 1}, {-1, 1, 1, 2}}};
    double *t = malloc(sizeof(double) * n);
    for (int it = 0; it < 2; ++it)
        for (int nu = 0; nu < 3; ++nu) {
            double *dw = DW + (size_t)(it * 3 + nu) * n;
            memset(dw, 0, sizeof(double) * n);
            for (int mu = 0; mu < 3; ++mu) {
                const int *q = dwq[nu][mu];
                or_derive_field(W + ((size_t)(it * 3 + nu) * 3 + mu) * n, t, nx, ny, nz, dx, mu,
                                or_sign_first(mu, q[0], q[1], q[2], q[3] - 1, lag), lag, lm[mu]);
                for (size_t p = 0; p < n; ++p) dw[p] = dw[p] + t[p];
            }
        }
    /* CalcBPot */
    for (int it = 0; it < 2; ++it) {
        for (size_t p = 0; p < n; ++p)
            Bp[it * n + p] = c[B3] * rho[(1 - it) * n + p] + (c[B3] + c[B4]) * rho[it * n + p];
        for (int a = 0; a < 3; ++a)
            or_derive_field(Bp + it * n, NB + (size_t)(it * 3 + a) * n, nx, ny, nz, dx, a,
                            or_sign_first(a, 1, 1, 1, 0, lag), lag, lm[a]);
    }
    if (!trc) {
        /* CalcSPot */
        for (int it = 0; it < 2; ++it) {
            const int at = 1 - it;
            for (int k = 0; k < 3; ++k)
                for (size_t p = 0; p < n; ++p) {
                    const double rt = pow(rho[p] + rho[n + p], a3);
                    S[(it * 3 + k) * n + p] =
                        +(c[B9] + c[B9q]) * rotvecj[(it * 3 + k) * n + p] +
                        c[B9] * rotvecj[(at * 3 + k) * n + p] +
                        2.0 * ((c[B10] + c[B11]) * vecs[(it * 3 + k) * n + p] +
                               c[B10] * vecs[(at * 3 + k) * n + p]) +
                        2.0 * ((c[B12a] + c[B13a]) * (rt * vecs[(it * 3 + k) * n + p]) +
                               c[B12a] * (rt * vecs[(at * 3 + k) * n + p]));
                }
        }
        /* CalcAPot + DivAPot (:463-468) */
        for (int it = 0; it < 2; ++it) {
            const int at = 1 - it;
            for (int k = 0; k < 3; ++k)
                for (size_t p = 0; p < n; ++p) {
                    double a = -2.0 * ((c[B3] + c[B4]) * vecj[(it * 3 + k) * n + p] +
                                       c[B3] * vecj[(at * 3 + k) * n + p]);
                    a = a + ((c[B9] + c[B9q]) * rots[(it * 3 + k) * n + p] +
                             c[B9] * rots[(at * 3 + k) * n + p]);
                    A[(it * 3 + k) * n + p] = a;
                }
            double *da = DA + (size_t)it * n;
            or_derive_field(A + (size_t)(it * 3 + 0) * n, da, nx, ny, nz, dx, 0,
                            or_sign_first(0, -1, -1, 1, 1, lag), lag, lm[0]);
            or_derive_field(A + (size_t)(it * 3 + 1) * n, t, nx, ny, nz, dx, 1,
                            or_sign_first(1, -1, -1, 1, 0, lag), lag, lm[1]);
            for (size_t p = 0; p < n; ++p) da[p] = da[p] + t[p];
            or_derive_field(A + (size_t)(it * 3 + 2) * n, t, nx, ny, nz, dx, 2,
                            or_sign_first(2, -1, 1, 1, 1, lag), lag, lm[2]);
            for (size_t p = 0; p < n; ++p) da[p] = da[p] + t[p];
        }
    }
    free(t);
}

/* ------------------------------------------------------------------------ */
/* Energy of an SLy4-type functional (CompEnergy, Energy.f90:62-179;
 * compSkyrme :623-859; CompKinetic :1426-1456; CompCOMCorrection
 * :1608-1660; CompCoulombEnergy/Exchange, Coulomb.f90:215-253).
 * kin_sp[it] = sum_i occ_i <psi_i|Lap psi_i> (dv included).
 * consts = hbm(2), mass(2), e2, hbar, dv, N, Z.  out[0] = total, out[1..2]
 * kinetic, out[3] Coulomb, out[4] exchange, out[5..6] COM (1-body),
 * out[7..42] Skyrme terms 1..36. */
void or_energy(const double *coef, const double *consts, int com1body, int trc,
               int nx, int ny, int nz, const double *rho, const double *tau,
               const double *nablaJ, const double *laprho, const double *vecj,
               const double *vecs, const double *rots, const double *vcoul,
               const double *kin_sp, int cexchange, double *out)
{
    const size_t n = (size_t)nx * ny * nz;
    const double *c = coef, dv = consts[6];
    const double a3 = c[Byt3a];
    double T[37];
    memset(T, 0, sizeof T);
    double s;
    s = 0; for (size_t p = 0; p < n; ++p) { const double rt = rho[p] + rho[n + p]; s = s + rt * rt; }
    T[1] = c[B1] * s;
    for (int it = 0; it < 2; ++it) { s = 0; for (size_t p = 0; p < n; ++p) s = s + rho[it * n + p] * rho[it * n + p]; T[2] = T[2] + s; }
    T[2] = c[B2] * T[2];
    s = 0; for (size_t p = 0; p < n; ++p) s = s + (rho[p] + rho[n + p]) * (tau[p] + tau[n + p]);
    T[3] = c[B3] * s;
    for (int it = 0; it < 2; ++it) { s = 0; for (size_t p = 0; p < n; ++p) s = s + rho[it * n + p] * tau[it * n + p]; T[4] = T[4] + c[B4] * s; }
    if (!trc) {
        s = 0;
        for (int k = 0; k < 3; ++k) for (size_t p = 0; p < n; ++p) { const double jt = vecj[k * n + p] + vecj[(3 + k) * n + p]; s = s + jt * jt; }
        T[5] = -c[B3] * s;
        for (int it = 0; it < 2; ++it) { s = 0; for (int k = 0; k < 3; ++k) for (size_t p = 0; p < n; ++p) s = s + vecj[(it * 3 + k) * n + p] * vecj[(it * 3 + k) * n + p]; T[6] = T[6] - c[B4] * s; }
    }
    s = 0; for (size_t p = 0; p < n; ++p) s = s + (rho[p] + rho[n + p]) * (laprho[p] + laprho[n + p]);
    T[7] = c[B5] * s;
    for (int it = 0; it < 2; ++it) { s = 0; for (size_t p = 0; p < n; ++p) s = s + rho[it * n + p] * laprho[it * n + p]; T[8] = T[8] + c[B6] * s; }
    s = 0; for (size_t p = 0; p < n; ++p) s = s + pow(rho[p] + rho[n + p], 2 + a3);
    T[9] = c[B7a] * s;
    for (int it = 0; it < 2; ++it) { s = 0; for (size_t p = 0; p < n; ++p) s = s + rho[it * n + p] * rho[it * n + p] * pow(rho[p] + rho[n + p], a3); T[10] = T[10] + c[B8a] * s; }
    s = 0; for (size_t p = 0; p < n; ++p) s = s + (rho[p] + rho[n + p]) * (nablaJ[p] + nablaJ[n + p]);
    T[11] = c[B9] * s;
    for (int it = 0; it < 2; ++it) { s = 0; for (size_t p = 0; p < n; ++p) s = s + rho[it * n + p] * nablaJ[it * n + p]; T[12] = T[12] + c[B9q] * s; }
    if (!trc) {
        /* sum(sum(a*b, 4)): the vector components are summed per point first */
        s = 0;
        for (size_t p = 0; p < n; ++p) {
            double d = 0;
            for (int k = 0; k < 3; ++k)
                d = d + (vecj[k * n + p] + vecj[(3 + k) * n + p]) * (rots[k * n + p] + rots[(3 + k) * n + p]);
            s = s + d;
        }
        T[13] = c[B9] * s;
        for (int it = 0; it < 2; ++it) {
            s = 0;
            for (size_t p = 0; p < n; ++p) {
                double d = 0;
                for (int k = 0; k < 3; ++k) d = d + vecj[(it * 3 + k) * n + p] * rots[(it * 3 + k) * n + p];
                s = s + d;
            }
            T[14] = T[14] + c[B9q] * s;
        }
        s = 0; for (int k = 0; k < 3; ++k) for (size_t p = 0; p < n; ++p) { const double st = vecs[k * n + p] + vecs[(3 + k) * n + p]; s = s + st * st; }
        T[15] = c[B10] * s;
        for (int it = 0; it < 2; ++it) { s = 0; for (int k = 0; k < 3; ++k) for (size_t p = 0; p < n; ++p) s = s + vecs[(it * 3 + k) * n + p] * vecs[(it * 3 + k) * n + p]; T[16] = T[16] + c[B11] * s; }
        /* T17/T18: rho_t^a3 times s_t^2 and s_q^2 (vanish for SLy4-type sets) */
        s = 0;
        for (size_t p = 0; p < n; ++p) {
            double d = 0;
            for (int k = 0; k < 3; ++k) { const double st = vecs[k * n + p] + vecs[(3 + k) * n + p]; d = d + st * st; }
            s = s + pow(rho[p] + rho[n + p], a3) * d;
        }
        T[17] = c[B12a] * s;
        for (int it = 0; it < 2; ++it) {
            s = 0;
            for (size_t p = 0; p < n; ++p) {
                double d = 0;
                for (int k = 0; k < 3; ++k) d = d + vecs[(it * 3 + k) * n + p] * vecs[(it * 3 + k) * n + p];
                s = s + pow(rho[p] + rho[n + p], a3) * d;
            }
            T[18] = T[18] + c[B13a] * s;
        }
    }
    double sky = 0;
    for (int k = 1; k <= 36; ++k) { T[k] = T[k] * dv; sky = sky + T[k]; }
    double kin[2], com[2];
    for (int it = 0; it < 2; ++it) kin[it] = -kin_sp[it] / 2.0 * consts[it];
    const double mtot = consts[7] * consts[2] + consts[8] * consts[3];
    for (int it = 0; it < 2; ++it) com[it] = com1body > 0 ? -kin[it] * consts[2 + it] / mtot : 0.0;
    double ec = 0; for (size_t p = 0; p < n; ++p) ec = ec + rho[n + p] * vcoul[p];
    ec = ec * dv * 0.5;
    double ex = 0;
    if (cexchange >= 1) {
        const double fac = -pow(3 / PI, 1 / 3.0) * consts[4] * 0.75 * dv;
        for (size_t p = 0; p < n; ++p) ex = ex + rho[n + p] * pow(rho[n + p], 1 / 3.0);
        ex = fac * ex;
    }
    out[0] = sky + (kin[0] + kin[1]) + ec + (com[0] + com[1]) + ex;
    out[1] = kin[0]; out[2] = kin[1]; out[3] = ec; out[4] = ex; out[5] = com[0]; out[6] = com[1];
    for (int k = 1; k <= 36; ++k) out[6 + k] = T[k];
}

/* Kinetic sums of CompKinetic (Energy.f90:1426-1456) before the -hbm/2
 * factor: kin_sp[it] = sum_l occ_l * (dv * sum psi_l . lap_l), with the inner
 * product of InproductSpinorReal (Spinor.f90:715-731). */
void or_kinetic_sums(const double *psi, const double *lap, int M, int N,
                     const int *qn, const double *occ, double dv, double *kin_sp)
{
    kin_sp[0] = kin_sp[1] = 0.0;
    for (int l = 0; l < M; ++l) {
        const int it = (qn[4 * l] + 3) / 2 - 1;
        const double *a = psi + (size_t)l * 4 * N, *b = lap + (size_t)l * 4 * N;
        double ip = 0.0;
        for (size_t i = 0; i < (size_t)4 * N; ++i) ip = ip + a[i] * b[i];
        ip = ip * dv;
        kin_sp[it] = kin_sp[it] + occ[l] * ip;
    }
}

/* Linear density mixing, MixingScheme 0 (DensityMixing.f90:39-45 with the
 * DensityVector operators Densities.f90:211-344): d = (1-a) d + a h. */
void or_mix(double *d, const double *h, size_t count, double a)
{
    const double b = 1 - a;
    for (size_t i = 0; i < count; ++i) d[i] = b * d[i] + a * h[i];
}

/* NaiveFill (HartreeFock.f90:177-262): the lowest floor(Z/2) protons and
 * floor(N/2) neutrons (floor(Z), floor(N) when time reversal is broken) by
 * global ascending energy get occupation 1, doubled under time reversal. */
void or_hf_fill(const double *energy, const int *qn, int M, double neutrons,
                double protons, int trc, double *occ)
{
    int *ord = malloc(sizeof(int) * M);
    for (int i = 0; i < M; ++i) ord[i] = i;
    for (int i = 1; i < M; ++i) {  /* OrderSpwfs: stable insertion sort */
        int v = ord[i], h = i;
        while (h > 0 && energy[v] < energy[ord[h - 1]]) { ord[h] = ord[h - 1]; --h; }
        ord[h] = v;
    }
    const int pub = trc ? (int)floor(protons / 2) : (int)floor(protons);
    const int nub = trc ? (int)floor(neutrons / 2) : (int)floor(neutrons);
    for (int i = 0; i < M; ++i) occ[i] = 0.0;
    int cnt = 0;
    for (int i = 0; i < M && cnt < pub; ++i)
        if (qn[4 * ord[i]] == 1) { occ[ord[i]] = 1.0; ++cnt; }
    cnt = 0;
    for (int i = 0; i < M && cnt < nub; ++i)
        if (qn[4 * ord[i]] == -1) { occ[ord[i]] = 1.0; ++cnt; }
    if (trc)
        for (int i = 0; i < M; ++i) occ[i] = 2.0 * occ[i];
    free(ord);
}

/* ------------------------------------------------------------------------ */
/* IterativeEstimation (ImaginaryTime.f90:28-211).  psi[4][N] is maxspwf;
 * when start != 0 it is first normalised as the reference does after
 * random_number (:63-65).  Then up to 500 power iterations
 * (:73-100): h = hPsi, maxE = <psi|h>, con = previous maxE - maxE (the
 * first iteration of a call never converges: the reference's previous maxE
 * is an undefined local there), u = dt/hbar (h - maxE psi) + momentum u
 * with u = 0 on entry, psi += u, psi /= sqrt(<psi|psi>), exit at
 * |con| < 1e-2.  relE for HF (:109-128) or, when rel_hfb > 0, the given
 * HFB value (the smallest positive quasiparticle energy, :136-152); minh
 * (:189-195), then out[6] = { dt, momentum, maxE, relE, con, iterations }
 * (:199-208). */
static double dotdv(const double *a, const double *b, size_t n, double dv)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s = s + a[i] * b[i];
    return s * dv;
}

void or_estimate(double *psi, int start, int nx, int ny, int nz, double dx, int lag,
                 const double *lagx, const double *lagy, const double *lagz,
                 const int *qn_est, int trc, double kin, const double *U, const double *B,
                 const double *NablaB, const double *W, const double *DivW,
                 const double *S, const double *A, const double *DivA, double dt,
                 double momentum, double dv, double hbar, int M, const int *qn,
                 const double *occ, const double *energy, double rel_hfb, double *out)
{
    const int N = nx * ny * nz;
    const size_t n4 = (size_t)4 * N;
    double *der = malloc(sizeof(double) * 3 * n4), *lap = malloc(sizeof(double) * n4);
    double *h = malloc(sizeof(double) * n4), *u = calloc(n4, sizeof(double));
    const int it = (qn_est[0] + 3) / 2 - 1;
    if (start) {
        const double sc = 1.0 / sqrt(dotdv(psi, psi, n4, dv));
        for (size_t k = 0; k < n4; ++k) psi[k] = sc * psi[k];
    }
    double maxE = 0.0, con = 1e300;
    int iter;
    for (iter = 1; iter <= 500; ++iter) {
        or_derive_spinor(psi, der, lap, nx, ny, nz, dx, qn_est[1], qn_est[2], qn_est[3], lag,
                         lagx, lagy, lagz);
        or_hpsi(psi, der, lap, N, it, trc, kin, U, B, NablaB, W, DivW, S, A, DivA, h);
        const double prev = maxE;
        maxE = dotdv(psi, h, n4, dv);
        con = iter == 1 ? 1e300 : prev - maxE;
        const double f = dt / hbar;
        for (size_t k = 0; k < n4; ++k) {
            const double d = h[k] - maxE * psi[k];
            u[k] = f * d + momentum * u[k];
            psi[k] = psi[k] + u[k];
        }
        const double sc = 1.0 / sqrt(dotdv(psi, psi, n4, dv));
        for (size_t k = 0; k < n4; ++k) psi[k] = sc * psi[k];
        if (fabs(con) < 1e-2) break;
    }
    if (iter > 500) iter = 500;
    double relE = 10000.0, minh = 10000.0;
    if (rel_hfb > 0.0) {
        relE = rel_hfb;
    } else {
        for (int i = 0; i < M; ++i) {
            if (fabs(occ[i]) < 0.5) continue;
            for (int ii = 0; ii < M; ++ii) {
                if (fabs(occ[ii]) > 0.5) continue;
                if (qn[4 * ii] != qn[4 * i]) continue;
                const double c = -energy[i] + energy[ii];
                if (c > 0.0 && c < relE) relE = c;
            }
        }
        if (relE < 0.1) relE = 0.1;
    }
    for (int i = 0; i < M; ++i)
        if (energy[i] < minh) minh = energy[i];
    const double mE = maxE - minh;
    const double kappa = relE / mE;
    const double r = (sqrt(kappa) - 1) / (sqrt(kappa) + 1);
    out[0] = 4.0 / (mE + relE + 2 * sqrt(mE * relE)) * hbar * (double)0.90f;
    out[1] = r * r;
    out[2] = maxE;
    out[3] = relE;
    out[4] = con;
    out[5] = iter;
    free(der); free(lap); free(h); free(u);
}

/* ------------------------------------------------------------------------ */
/* PreconditionPotential (Damping.f90:158-198) with preconoperator
 * (:200-228): per isospin, conjugate gradients on
 * (beta + alpha Lap) x = pot[it] from x = 0; Lap is the wavefunction
 * Laplacian (FD or Lagrange) with ghost signs (sx, sy, sz); at most 300
 * iterations; the loop ends after the direction update once
 * dv |r|^2 < 1e-8.  pot[2][N] in, x[2][N] out, iters[2] iterations. */
void or_precondition(const double *pot, double *x, int nx, int ny, int nz, double dx, int lag,
                     const double *lagx, const double *lagy, const double *lagz, int sx, int sy,
                     int sz, double alpha, double beta, double dv, int *iters)
{
    const size_t N = (size_t)nx * ny * nz;
    double *r = malloc(sizeof(double) * N), *d = malloc(sizeof(double) * N);
    double *q = malloc(sizeof(double) * N);
    for (int it = 0; it < 2; ++it) {
        double *xi = x + it * N;
        for (size_t k = 0; k < N; ++k) {
            xi[k] = 0.0;
            r[k] = pot[it * N + k];
            d[k] = r[k];
        }
        double newres = dotdv(d, d, N, dv);
        int iter;
        for (iter = 1; iter <= 300; ++iter) {
            or_laplace_field(d, q, nx, ny, nz, dx, sx, sy, sz, lag, lagx, lagy, lagz);
            for (size_t k = 0; k < N; ++k) q[k] = beta * d[k] + alpha * q[k];
            const double a = newres / dotdv(d, q, N, dv);
            for (size_t k = 0; k < N; ++k) {
                xi[k] = xi[k] + a * d[k];
                r[k] = r[k] - a * q[k];
            }
            const double oldres = newres;
            newres = dotdv(r, r, N, dv);
            const double b = newres / oldres;
            for (size_t k = 0; k < N; ++k) d[k] = r[k] + b * d[k];
            if (newres < 1e-8) break;
        }
        iters[it] = iter > 300 ? 300 : iter;
    }
    free(r); free(d); free(q);
}

/* ------------------------------------------------------------------------ */
/* RutzCutOff (Moments.f90:2703-2849): per isospin, threshold = max(rho)/10;
 * surface points by linear interpolation between neighbours (i, i+1),
 * (j, j+1), (k, k+1) of the cells i < nx-1, j < ny-1, k < nz-1 of the
 * octant whose densities straddle the threshold; DeltaR = -+ distance to
 * the nearest surface point (minus inside, rho > threshold);
 * cutoff = 1 / (1 + exp((DeltaR - radd) / acut)).  An isospin with
 * dv sum rho < 1e-10 gets cutoff 0.  Mesh points at (i + 1/2) dx. */
void or_rutz_cutoff(const double *rho, int nx, int ny, int nz, double dx, double dv,
                    double radd, double acut, double *cutoff)
{
    const size_t N = (size_t)nx * ny * nz;
    double *surf = malloc(sizeof(double) * 3 * 3 * N);
    for (int it = 0; it < 2; ++it) {
        const double *r = rho + it * N;
        double *cut = cutoff + it * N;
        double tot = 0.0, mx = r[0];
        for (size_t p = 0; p < N; ++p) {
            tot += r[p];
            if (r[p] > mx) mx = r[p];
        }
        if (dv * tot < 0.0000000001) {
            for (size_t p = 0; p < N; ++p) cut[p] = 0.0;
            continue;
        }
        const double tr = mx / 10.0;
        size_t T = 0;
#define RHO(i, j, k) r[((size_t)(k) * ny + (j)) * nx + (i)]
        for (int k = 0; k < nz - 1; ++k) {
            const double Z = dx / 2.0 + dx * k;
            for (int j = 0; j < ny - 1; ++j) {
                const double Y = dx / 2.0 + dx * j;
                for (int i = 0; i < nx - 1; ++i) {
                    const double X = dx / 2.0 + dx * i;
                    const double a = RHO(i, j, k);
                    const double nbr[3] = {RHO(i + 1, j, k), RHO(i, j + 1, k), RHO(i, j, k + 1)};
                    for (int ax = 0; ax < 3; ++ax) {
                        const double b = nbr[ax];
                        if ((a >= tr && b <= tr) || (a <= tr && b >= tr)) {
                            double P[3] = {X, Y, Z};
                            P[ax] = P[ax] + dx * (tr - a) / (b - a);
                            surf[3 * T] = P[0];
                            surf[3 * T + 1] = P[1];
                            surf[3 * T + 2] = P[2];
                            ++T;
                        }
                    }
                }
            }
        }
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double X = dx / 2.0 + dx * i, Y = dx / 2.0 + dx * j,
                                 Z = dx / 2.0 + dx * k;
                    double dmin = 1e12;
                    for (size_t l = 0; l < T; ++l) {
                        const double ex = X - surf[3 * l], ey = Y - surf[3 * l + 1],
                                     ez = Z - surf[3 * l + 2];
                        const double d = ex * ex + ey * ey + ez * ez;
                        if (d < dmin) dmin = d;
                    }
                    const double sig = RHO(i, j, k) > tr ? -1.0 : 1.0;
                    const double dr = sig * sqrt(dmin);
                    cut[((size_t)k * ny + j) * nx + i] = 1.0 / (1.0 + exp((dr - radd) / acut));
                }
#undef RHO
    }
    free(surf);
}
