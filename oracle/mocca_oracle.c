/*
 * CPU restatement ("oracle") of the MOCCa imaginary-time hot path.
 * TEST INFRASTRUCTURE ONLY -- see mocca_oracle.h for the contract, the layout
 * and how each function is pinned.  Plain scalar C, one thread.
 */
#include "mocca_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(x, y, z) (((size_t)(z) * ny + (y)) * nx + (x))

/* ------------------------------------------------------------------------ */
/* Lagrange matrices: IniLag + reduceLag (Derivatives.f90:899-1078).        */
static void lag_axis(int n, int nfull, double dx, double *m)
{
    const double pi = 4.0 * atan2(1.0, 1.0);
    double *c = calloc((size_t)2 * nfull + 2, sizeof(double));
    double *l = calloc((size_t)2 * nfull + 2, sizeof(double));
    const double factor = pi / nfull;
    const double dxf = factor / dx;
    const double diag = -pi * pi / (3.0 * dx * dx) *
                        (1.0 - 1.0 / (double)(nfull * nfull));
    for (int k = 1; k <= nfull - 1; ++k) {
        const double s = sin(factor * k), co = cos(factor * k);
        const double sg = (k % 2 == 0) ? 1.0 : -1.0; /* (-1)**k */
        c[k] = dxf * sg / s;
        l[k] = -sg * 2.0 * (dxf * dxf) * co / (s * s); /* (-1)**(k+1) */
    }
    double *f1m = m, *f1p = m + n * n, *f2m = m + 2 * n * n, *f2p = m + 3 * n * n;
    for (int i = 1; i <= n; ++i) {
        for (int j = 1; j <= n; ++j) {
            const int r = (i - 1) * n + (j - 1);
            if (j < i) {
                f1p[r] = c[i - j] + c[i + j - 1];
                f1m[r] = c[i - j] - c[i + j - 1];
                f2p[r] = l[i - j] + l[i + j - 1];
                f2m[r] = l[i - j] - l[i + j - 1];
            } else if (j == i) {
                f2m[r] = diag - l[2 * i - 1];
                f2p[r] = diag + l[2 * i - 1];
                f1m[r] = -c[2 * i - 1];
                f1p[r] = c[2 * i - 1];
            } else {
                f1p[r] = -c[j - i] + c[i + j - 1];
                f1m[r] = -c[j - i] - c[i + j - 1];
                f2p[r] = l[j - i] + l[i + j - 1];
                f2m[r] = l[j - i] - l[i + j - 1];
            }
        }
    }
    free(c);
    free(l);
}

void or_lag_build(int nx, int ny, int nz, double dx, int nfx, int nfy, int nfz,
                  double *lagx, double *lagy, double *lagz)
{
    lag_axis(nx, nfx, dx, lagx);
    lag_axis(ny, nfy, dx, lagy);
    lag_axis(nz, nfz, dx, lagz);
}

/* ------------------------------------------------------------------------ */
/* EV8 sign rules.  First derivatives: Opt_X_EV8 (:24-44) uses S only, the
 * Lagrange variant Opt_LAGX_EV8 (:115-132) S*TS; y: TS rule (:208-225,
 * :306-320); z: S*P rule (:340-357, :432-446).  Laplacian: Lapla_EV8 and
 * Lapla_LAG_EV8 share one table (:522-549, :551-571).                      */
int or_sign_first(int axis, int P, int S, int TS, int comp, int lag)
{
    static const int cx[4] = {1, -1, -1, 1}, cy[4] = {1, -1, 1, -1},
                     cz[4] = {1, 1, -1, -1};
    if (axis == 0) return cx[comp] * (lag ? S * TS : S);
    if (axis == 1) return cy[comp] * TS;
    return cz[comp] * S * P;
}

int or_sign_lap(int axis, int P, int S, int TS, int comp)
{
    static const int cx[4] = {1, -1, -1, 1}, cy[4] = {1, -1, 1, -1},
                     cz[4] = {1, 1, -1, -1};
    if (axis == 0) return cx[comp] * S;
    if (axis == 1) return cy[comp] * TS;
    return cz[comp] * S * P;
}

/* Gather one line along `axis` through (a, b) -- the other two indices. */
static void get_line(const double *f, int nx, int ny, int nz, int axis, int a,
                     int b, double *line)
{
    if (axis == 0) for (int i = 0; i < nx; ++i) line[i] = f[IDX(i, a, b)];
    else if (axis == 1) for (int i = 0; i < ny; ++i) line[i] = f[IDX(a, i, b)];
    else for (int i = 0; i < nz; ++i) line[i] = f[IDX(a, b, i)];
}

static void put_line(double *f, int nx, int ny, int nz, int axis, int a, int b,
                     const double *line)
{
    if (axis == 0) for (int i = 0; i < nx; ++i) f[IDX(i, a, b)] = line[i];
    else if (axis == 1) for (int i = 0; i < ny; ++i) f[IDX(a, i, b)] = line[i];
    else for (int i = 0; i < nz; ++i) f[IDX(a, b, i)] = line[i];
}

/* Value at line index j (0-based) with the reflection ghost w(-1-j)*s below
 * the symmetry plane and zero beyond the far edge of the box. */
static inline double ghost(const double *w, int n, int j, int s)
{
    if (j < 0) return s * w[-1 - j];
    if (j >= n) return 0.0;
    return w[j];
}

/* 7-point first derivative on one line: derx (OptimizedDerivatives.f90:46-113). */
static void fd1_line(const double *w, double *d, int n, int s, double dx)
{
    const double ca = 45.0, cb = 9.0, cc = 60.0;
    const double ccdx = 1.0 / (cc * dx);
    const int sg = s > 0 ? 1 : -1;
    for (int i = 0; i < n; ++i) {
        double v = ca * (ghost(w, n, i + 1, sg) - ghost(w, n, i - 1, sg))
                 - cb * (ghost(w, n, i + 2, sg) - ghost(w, n, i - 2, sg))
                 + (ghost(w, n, i + 3, sg) - ghost(w, n, i - 3, sg));
        d[i] = v * ccdx;
    }
}

/* 9-point second derivative on one line, term order of lapla_X_ev8
 * (OptimizedDerivatives.f90:617-676). */
static void fd2_line(const double *w, double *d, int n, int s, double dx)
{
    const double xf = -8064.0;
    const double xm = 43050.0 / (xf * 3), x2 = 1008.0 / xf, x3 = -128.0 / xf,
                 x4 = 9.0 / xf;
    double xd = 5040.0 / xf * (dx * dx);
    xd = 1.0 / xd;
    const double xp = s;
    for (int i = 1; i <= n; ++i) {
        double acc = xm * w[i - 1];
        /* explicit edge rows near the symmetry plane (ghosts) */
        if (i == 1) acc = acc + w[0] * xp + x2 * w[1] * xp + x3 * w[2] * xp + x4 * w[3] * xp;
        if (i == 2) acc = acc + w[0] + x2 * w[0] * xp + x3 * w[1] * xp + x4 * w[2] * xp;
        if (i == 3) acc = acc + w[1] + x2 * w[0] + x3 * w[0] * xp + x4 * w[1] * xp;
        if (i == 4) acc = acc + w[2] + x2 * w[1] + x3 * w[0] + x4 * w[0] * xp;
        /* explicit edge rows at the far edge (zeros beyond) */
        if (i == n - 1) acc = acc + w[n - 1];
        if (i == n - 2) acc = acc + w[n - 2] + x2 * w[n - 1];
        if (i == n - 3) acc = acc + w[n - 3] + x2 * w[n - 2] + x3 * w[n - 1];
        if (i <= n - 4)
            acc = acc + w[i] + x2 * w[i + 1] + x3 * w[i + 2] + x4 * w[i + 3];
        if (i >= 5)
            acc = acc + w[i - 2] + x2 * w[i - 3] + x3 * w[i - 4] + x4 * w[i - 5];
        d[i - 1] = -acc * xd;
    }
}

/* Dense line product der(i) = sum_j M(i,j) w(j) (matmul in Opt_LAG*_EV8). */
static void mat_line(const double *w, double *d, int n, const double *m)
{
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += m[i * n + j] * w[j];
        d[i] = acc;
    }
}

void or_derive_field(const double *f, double *out, int nx, int ny, int nz,
                     double dx, int axis, int sign, int lag, const double *lagmat)
{
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int na = axis == 0 ? ny : nx;
    const int nb = axis == 2 ? ny : nz;
    double *w = malloc(sizeof(double) * n), *d = malloc(sizeof(double) * n);
    /* Opt_LAG*: matrix 2 (+) when the sign is +1, else matrix 1 (-). */
    const double *m = lagmat ? lagmat + (sign > 0 ? n * n : 0) : NULL;
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
            get_line(f, nx, ny, nz, axis, a, b, w);
            if (lag) mat_line(w, d, n, m);
            else fd1_line(w, d, n, sign, dx);
            put_line(out, nx, ny, nz, axis, a, b, d);
        }
    free(w);
    free(d);
}

static void second_axis(const double *f, double *out, int nx, int ny, int nz,
                        double dx, int axis, int sign, int lag,
                        const double *lagmat)
{
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int na = axis == 0 ? ny : nx;
    const int nb = axis == 2 ? ny : nz;
    double *w = malloc(sizeof(double) * n), *d = malloc(sizeof(double) * n);
    const double *m = lag ? lagmat + 2 * n * n + (sign == 1 ? n * n : 0) : NULL;
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
            get_line(f, nx, ny, nz, axis, a, b, w);
            if (lag) mat_line(w, d, n, m);
            else fd2_line(w, d, n, sign, dx);
            put_line(out, nx, ny, nz, axis, a, b, d);
        }
    free(w);
    free(d);
}

void or_laplace_field(const double *f, double *out, int nx, int ny, int nz,
                      double dx, int sx, int sy, int sz, int lag,
                      const double *lagx, const double *lagy, const double *lagz)
{
    const size_t N = (size_t)nx * ny * nz;
    double *tx = malloc(sizeof(double) * N), *ty = malloc(sizeof(double) * N),
           *tz = malloc(sizeof(double) * N);
    second_axis(f, tx, nx, ny, nz, dx, 0, sx, lag, lagx);
    second_axis(f, ty, nx, ny, nz, dx, 1, sy, lag, lagy);
    second_axis(f, tz, nx, ny, nz, dx, 2, sz, lag, lagz);
    for (size_t i = 0; i < N; ++i) out[i] = tx[i] + ty[i] + tz[i];
    free(tx);
    free(ty);
    free(tz);
}

void or_derive_spinor(const double *psi, double *der, double *lap,
                      int nx, int ny, int nz, double dx, int P, int S, int TS,
                      int lag, const double *lagx, const double *lagy,
                      const double *lagz)
{
    const size_t N = (size_t)nx * ny * nz;
    const double *lm[3] = {lagx, lagy, lagz};
    for (int c = 0; c < 4; ++c) {
        for (int a = 0; a < 3; ++a)
            or_derive_field(psi + c * N, der + (a * 4 + c) * N, nx, ny, nz, dx,
                            a, or_sign_first(a, P, S, TS, c, lag), lag, lm[a]);
        or_laplace_field(psi + c * N, lap + c * N, nx, ny, nz, dx,
                         or_sign_lap(0, P, S, TS, c), or_sign_lap(1, P, S, TS, c),
                         or_sign_lap(2, P, S, TS, c), lag, lagx, lagy, lagz);
    }
}

void or_derive_states(const double *psi, double *der, double *lap, int M, const int *qn,
                      int nx, int ny, int nz, double dx, int lag, const double *lagx,
                      const double *lagy, const double *lagz)
{
    const size_t N = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < M; ++i)
        or_derive_spinor(psi + (size_t)i * 4 * N, der + (size_t)i * 12 * N,
                         lap + (size_t)i * 4 * N, nx, ny, nz, dx, qn[4 * i + 1],
                         qn[4 * i + 2], qn[4 * i + 3], lag, lagx, lagy, lagz);
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Spinor algebra (Spinor.f90:113-157,264-281,353-369).                     */
static inline void pauli(const double in[4], int dir, double out[4])
{
    if (dir == 0) { out[0] = in[2]; out[1] = in[3]; out[2] = in[0]; out[3] = in[1]; }
    else if (dir == 1) { out[0] = in[3]; out[1] = -in[2]; out[2] = -in[1]; out[3] = in[0]; }
    else { out[0] = in[0]; out[1] = in[1]; out[2] = -in[2]; out[3] = -in[3]; }
}

static inline double re_dot(const double a[4], const double b[4])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

static inline double im_dot(const double a[4], const double b[4])
{
    return a[0] * b[1] - a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
}

void or_hpsi(const double *psi, const double *der, const double *lap, int N,
             int it, int trc, double kin, const double *U, const double *B,
             const double *NablaB, const double *W, const double *DivW,
             const double *S, const double *A, const double *DivA,
             double *out)
{
    const size_t n = (size_t)N;
    for (size_t p = 0; p < n; ++p) {
        double v[4], d[3][4], l[4], h[4];
        for (int c = 0; c < 4; ++c) {
            v[c] = psi[c * n + p];
            l[c] = lap[c * n + p];
            for (int m = 0; m < 3; ++m) d[m][c] = der[(m * 4 + c) * n + p];
        }
        /* ActionOfBNew (MeanFields.f90:811-851) */
        const double bq = B[it * n + p];
        double ab[4], sd[4];
        for (int c = 0; c < 4; ++c) {
            ab[c] = -bq * l[c];
            ab[c] = ab[c] - kin * l[c];
            double d1 = NablaB[(it * 3 + 0) * n + p] * d[0][c];
            double d2 = NablaB[(it * 3 + 1) * n + p] * d[1][c];
            double d3 = NablaB[(it * 3 + 2) * n + p] * d[2][c];
            sd[c] = d1 + d2 + d3;
            ab[c] = ab[c] - sd[c];
        }
        /* ActionOfU (MeanFields.f90:792-809) */
        const double uq = U[it * n + p];
        /* ActionOfW (MeanFields.f90:1029-1065) */
        double aw[4] = {0, 0, 0, 0}, t[4];
        for (int mu = 0; mu < 3; ++mu)
            for (int nu = 0; nu < 3; ++nu) {
                const double w = W[((it * 3 + nu) * 3 + mu) * n + p];
                pauli(d[mu], nu, t);
                for (int c = 0; c < 4; ++c) aw[c] = aw[c] + w * t[c];
            }
        for (int nu = 0; nu < 3; ++nu) {
            const double dw = DivW[(it * 3 + nu) * n + p];
            pauli(v, nu, t);
            for (int c = 0; c < 4; ++c) aw[c] = aw[c] + 0.5 * dw * t[c];
        }
        /* -i * aw : (1,2,3,4) <- (a2, -a1, a4, -a3) */
        double w4[4] = {aw[1], -aw[0], aw[3], -aw[2]};
        for (int c = 0; c < 4; ++c) h[c] = (ab[c] + uq * v[c]) + w4[c];
        if (!trc) {
            /* ActionOfS (MeanFields.f90:1131-1147) */
            double as[4] = {0, 0, 0, 0};
            for (int m = 0; m < 3; ++m) {
                pauli(v, m, t);
                const double sm = S[(it * 3 + m) * n + p];
                for (int c = 0; c < 4; ++c) as[c] = as[c] + sm * t[c];
            }
            /* ActionOfA (MeanFields.f90:1107-1129) */
            double aa[4] = {0, 0, 0, 0};
            for (int m = 0; m < 3; ++m) {
                const double am = A[(it * 3 + m) * n + p];
                for (int c = 0; c < 4; ++c) aa[c] = aa[c] + am * d[m][c];
            }
            const double da = DivA[it * n + p];
            for (int c = 0; c < 4; ++c) aa[c] = aa[c] + 0.5 * da * v[c];
            /* -(i*aa) */
            double a4[4] = {aa[1], -aa[0], aa[3], -aa[2]};
            for (int c = 0; c < 4; ++c) h[c] = (h[c] + as[c]) + a4[c];
        }
        for (int c = 0; c < 4; ++c) out[c * n + p] = h[c];
    }
}

/* ------------------------------------------------------------------------ */
void or_density(const double *psi, const double *der, int M, int N,
                const int *qn, const double *occ, int trc,
                double *rho, double *tau, double *nablaJ,
                double *vecj, double *vecs)
{
    const size_t n = (size_t)N;
    memset(rho, 0, sizeof(double) * 2 * n);
    memset(tau, 0, sizeof(double) * 2 * n);
    memset(nablaJ, 0, sizeof(double) * 2 * n);
    if (!trc) {
        memset(vecj, 0, sizeof(double) * 6 * n);
        memset(vecs, 0, sizeof(double) * 6 * n);
    }
    /* (i,j,k) order of GetNablaJ (Spwf.f90:2206-2243) */
    static const int tri[6][3] = {{0, 1, 2}, {0, 2, 1}, {2, 1, 0},
                                  {2, 0, 1}, {1, 0, 2}, {1, 2, 0}};
    static const double lc[6] = {1, -1, -1, 1, -1, 1};
    /* point blocks over threads; every point sums its states in index order */
    const size_t pblk = 4096;
#pragma omp parallel for schedule(static)
    for (size_t p0 = 0; p0 < n; p0 += pblk)
    for (int s = 0; s < M; ++s) {
        const double o = occ[s];
        if (o == 0.0) continue;
        const int it = (qn[4 * s] + 3) / 2 - 1;
        const double *ps = psi + (size_t)s * 4 * n;
        const double *ds = der + (size_t)s * 12 * n;
        const size_t p1 = p0 + pblk < n ? p0 + pblk : n;
        for (size_t p = p0; p < p1; ++p) {
            double v[4], d[3][4], t[4];
            for (int c = 0; c < 4; ++c) {
                v[c] = ps[c * n + p];
                for (int m = 0; m < 3; ++m) d[m][c] = ds[(m * 4 + c) * n + p];
            }
            rho[it * n + p] += o * re_dot(v, v);
            double ta = 0.0;
            for (int m = 0; m < 3; ++m) ta = ta + re_dot(d[m], d[m]);
            tau[it * n + p] += o * ta;
            double nj = 0.0;
            for (int q = 0; q < 6; ++q) {
                pauli(d[tri[q][1]], tri[q][2], t);
                nj = nj + lc[q] * im_dot(d[tri[q][0]], t);
            }
            nablaJ[it * n + p] += o * nj;
            if (!trc) {
                for (int m = 0; m < 3; ++m)
                    vecj[(it * 3 + m) * n + p] += o * im_dot(v, d[m]);
                for (int l = 0; l < 3; ++l) {
                    pauli(v, l, t);
                    vecs[(it * 3 + l) * n + p] += o * re_dot(v, t);
                }
            }
        }
    }
}

void or_density_derived(const double *rho, const double *vecj,
                        const double *vecs, int nx, int ny, int nz, double dx,
                        int lag, const double *lagx, const double *lagy,
                        const double *lagz, int trc, double *derrho,
                        double *laprho, double *rotvecj, double *rots,
                        double *ders)
{
    const size_t n = (size_t)nx * ny * nz;
    const double *lm[3] = {lagx, lagy, lagz};
    const int P = 1, S = 1, TS = 1; /* ParityInt, SignatureInt, TimeSimplexInt */
#define DER(src, dst, ax, p_, s_, ts_, c_)                                      \
    or_derive_field(src, dst, nx, ny, nz, dx, ax,                              \
                    or_sign_first(ax, p_, s_, ts_, (c_) - 1, lag), lag, lm[ax])
    for (int it = 0; it < 2; ++it) {
        for (int a = 0; a < 3; ++a)
            DER(rho + it * n, derrho + (it * 3 + a) * n, a, P, S, TS, 1);
        or_laplace_field(rho + it * n, laprho + it * n, nx, ny, nz, dx,
                         or_sign_lap(0, P, S, TS, 0), or_sign_lap(1, P, S, TS, 0),
                         or_sign_lap(2, P, S, TS, 0), lag, lagx, lagy, lagz);
    }
    if (trc) return;
    double *d = malloc(sizeof(double) * 9 * n); /* d[a*3+m]: d_a of j_m */
    for (int it = 0; it < 2; ++it) {
        const double *jx = vecj + (it * 3 + 0) * n, *jy = vecj + (it * 3 + 1) * n,
                     *jz = vecj + (it * 3 + 2) * n;
        /* ComputeDensity (Densities.f90:1008-1025) */
        DER(jx, d + (1 * 3 + 0) * n, 1, -P, -S, TS, 2);
        DER(jx, d + (2 * 3 + 0) * n, 2, -P, -S, TS, 2);
        DER(jy, d + (0 * 3 + 1) * n, 0, -P, -S, TS, 1);
        DER(jy, d + (2 * 3 + 1) * n, 2, -P, -S, TS, 1);
        DER(jz, d + (0 * 3 + 2) * n, 0, -P, S, TS, 2);
        DER(jz, d + (1 * 3 + 2) * n, 1, -P, S, TS, 2);
        for (size_t p = 0; p < n; ++p) {
            rotvecj[(it * 3 + 0) * n + p] = d[(1 * 3 + 2) * n + p] - d[(2 * 3 + 1) * n + p];
            rotvecj[(it * 3 + 1) * n + p] = d[(2 * 3 + 0) * n + p] - d[(0 * 3 + 2) * n + p];
            rotvecj[(it * 3 + 2) * n + p] = d[(0 * 3 + 1) * n + p] - d[(1 * 3 + 0) * n + p];
        }
        /* Densities.f90:1058-1101: s_x (P,-S,TS,1), s_y (P,-S,TS,2), s_z (P,S,TS,1) */
        double *ds = ders + (size_t)it * 9 * n; /* [comp][dir] */
        const double *sx = vecs + (it * 3 + 0) * n, *sy = vecs + (it * 3 + 1) * n,
                     *sz = vecs + (it * 3 + 2) * n;
        for (int a = 0; a < 3; ++a) {
            DER(sx, ds + (0 * 3 + a) * n, a, P, -S, TS, 1);
            DER(sy, ds + (1 * 3 + a) * n, a, P, -S, TS, 2);
            DER(sz, ds + (2 * 3 + a) * n, a, P, S, TS, 1);
        }
        for (size_t p = 0; p < n; ++p) {
            /* RotS_x = d_y s_z - d_z s_y, etc. */
            rots[(it * 3 + 0) * n + p] = ds[(2 * 3 + 1) * n + p] - ds[(1 * 3 + 2) * n + p];
            rots[(it * 3 + 1) * n + p] = ds[(0 * 3 + 2) * n + p] - ds[(2 * 3 + 0) * n + p];
            rots[(it * 3 + 2) * n + p] = ds[(1 * 3 + 0) * n + p] - ds[(0 * 3 + 1) * n + p];
        }
    }
    free(d);
#undef DER
}

/* ------------------------------------------------------------------------ */
static double inprod(const double *a, const double *b, size_t n4, double dv)
{
    double s = 0.0;
    for (size_t i = 0; i < n4; ++i) s = s + a[i] * b[i];
    return s * dv;
}

void or_gram_schmidt(double *psi, int M, int N, const int *qn,
                     const double *energy, double dv)
{
    const size_t n4 = (size_t)4 * N;
    int *idx = malloc(sizeof(int) * M);
    for (int iso = -1; iso <= 1; iso += 2) {
        /* OrderSpwfsISO: stable insertion sort by energy within the isospin */
        int nw = 0;
        for (int i = 0; i < M; ++i)
            if (qn[4 * i] == iso) idx[nw++] = i;
        for (int i = 1; i < nw; ++i) {
            int v = idx[i], h = i;
            while (h > 0 && energy[v] < energy[idx[h - 1]]) {
                idx[h] = idx[h - 1];
                --h;
            }
            idx[h] = v;
        }
        for (int k = 0; k < nw; ++k) {
            double *pk = psi + (size_t)idx[k] * n4;
            const double norm = inprod(pk, pk, n4, dv);
            const double sc = 1.0 / sqrt(norm);
            for (size_t i = 0; i < n4; ++i) pk[i] = sc * pk[i];
            const int *q = qn + 4 * idx[k];
#pragma omp parallel for schedule(dynamic)
            for (int l = k + 1; l < nw; ++l) {
                const int *r = qn + 4 * idx[l];
                if (r[0] != q[0] || r[1] != q[1] || r[2] != q[2]) continue;
                double *pl = psi + (size_t)idx[l] * n4;
                const double me = inprod(pl, pk, n4, dv);
                for (size_t i = 0; i < n4; ++i) pl[i] = pl[i] - me * pk[i];
            }
        }
    }
    free(idx);
}

void or_graddesc(double *psi, const double *der, const double *lap,
                 double *upd, int M, int N, const int *qn, int trc,
                 const double *kin2, const double *U, const double *B,
                 const double *NablaB, const double *W, const double *DivW,
                 const double *S, const double *A, const double *DivA,
                 double dt_over_hbar, double momentum, double dv,
                 double *energy, double *disp)
{
    const size_t n4 = (size_t)4 * N;
#pragma omp parallel
    {
    double *h = malloc(sizeof(double) * n4);
#pragma omp for schedule(dynamic)
    for (int i = 0; i < M; ++i) {
        double *p = psi + i * n4, *u = upd + i * n4;
        const int it = (qn[4 * i] + 3) / 2 - 1;
        or_hpsi(p, der + (size_t)i * 3 * n4, lap + i * n4, N, it, trc, kin2[it],
                U, B, NablaB, W, DivW, S, A, DivA, h);
        const double e = inprod(p, h, n4, dv);
        disp[i] = inprod(h, h, n4, dv) - e * e;
        energy[i] = e;
        for (size_t k = 0; k < n4; ++k) {
            const double step = -(dt_over_hbar * h[k]) + momentum * u[k];
            u[k] = p[k];
            p[k] = p[k] + step;
        }
    }
    free(h);
    }
    or_gram_schmidt(psi, M, N, qn, energy, dv);
    for (size_t k = 0; k < (size_t)M * n4; ++k) upd[k] = psi[k] - upd[k];
}
