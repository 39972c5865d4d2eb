/*
 * CPU restatement ("oracle") of the MOCCa imaginary-time hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker of the HIP library (or as the timed CPU baseline).  The product
 * (mocca_amd/) never links or calls it.
 *
 * Every function follows a named reference routine (file:line in the
 * reference MOCCa tree) and keeps its operation order where that is cheap,
 * so that results agree with the reference to rounding.
 *
 * Layout (identical to the HIP library's):
 *   field     f[(z*ny + y)*nx + x]                      N = nx*ny*nz
 *   spinor    psi[c][N], c = 0..3 = (Re up, Im up, Re down, Im down)
 *   batch     psi[i][c][N] for states i = 0..M-1
 *   Lagrange  lag[a] for axis a: 4 matrices n_a x n_a (row-major) in the
 *             order first(-), first(+), second(-), second(+)
 *
 * Parity pinning (see tests/test_oracle.py): every function below is
 * compared with golden vectors written by the reference's own modules
 * (the .npz fixtures in tests/golden, produced by oracle/ref_driver.f90), and the
 * derivative operators additionally against closed forms.
 */
#ifndef MOCCA_ORACLE_H
#define MOCCA_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Lagrange-mesh matrices (Derivatives.f90:899-1078). nfull[a] = points on the
 * full line (n_a * 2 when the plane symmetry of that axis is conserved). */
void or_lag_build(int nx, int ny, int nz, double dx, int nfx, int nfy, int nfz,
                  double *lagx, double *lagy, double *lagz);

/* Reflection sign of component c (0..3) along axis (0..2) for quantum
 * numbers P, S, TS (OptimizedDerivatives.f90:24-44,115-132,208-225,340-357). */
int or_sign_first(int axis, int P, int S, int TS, int comp, int lag);
int or_sign_lap(int axis, int P, int S, int TS, int comp);

/* First derivative of one field along axis with ghost sign s.
 * FD: derx/dery/derz (OptimizedDerivatives.f90:46-113,230-304,362-430).
 * Lagrange: Opt_LAG{X,Y,Z}_EV8 (OptimizedDerivatives.f90:115-148,306-338,432-464). */
void or_derive_field(const double *f, double *out, int nx, int ny, int nz,
                     double dx, int axis, int sign, int lag, const double *lagmat);

/* Laplacian of one field with per-axis ghost signs.
 * FD: lapla_{X,Y,Z}_ev8 (OptimizedDerivatives.f90:617-795).
 * Lagrange: Lapla_LAG_EV8 (OptimizedDerivatives.f90:551-615). */
void or_laplace_field(const double *f, double *out, int nx, int ny, int nz,
                      double dx, int sx, int sy, int sz, int lag,
                      const double *lagx, const double *lagy, const double *lagz);

/* CompDer for one state: der[3][4][N], lap[4][N] (Spwf.f90:778-792). */
void or_derive_spinor(const double *psi, double *der, double *lap,
                      int nx, int ny, int nz, double dx, int P, int S, int TS,
                      int lag, const double *lagx, const double *lagy,
                      const double *lagz);

/* CompDer for states 0..M-1 (DeriveAll, SpwfStorage.f90:586-604), over
 * OpenMP threads; qn[4*i+1..3] = parity, signature, time simplex. */
void or_derive_states(const double *psi, double *der, double *lap, int M, const int *qn,
                      int nx, int ny, int nz, double dx, int lag, const double *lagx,
                      const double *lagy, const double *lagz);

/* OpenMP threads of the loops over states and points (results do not depend
 * on the count: every point and every state keeps its operation order). */
void or_set_threads(int n);
int or_max_threads(void);

/* hPsi for one state (ImaginaryTime.f90:213-304): ActionOfBNew + ActionOfU
 * + ActionOfW (+ ActionOfS + ActionOfA when trc == 0)
 * (MeanFields.f90:792-851,1029-1065,1107-1147).
 * Fields: U[2][N] B[2][N] NablaB[2][3][N] W[2][nu][mu][N] DivW[2][3][N]
 *         S[2][3][N] A[2][3][N] DivA[2][N]; it = 0 neutron, 1 proton.
 * kin = hbm(it)/2 * reducedmass(it). */
void or_hpsi(const double *psi, const double *der, const double *lap, int N,
             int it, int trc, double kin, const double *U, const double *B,
             const double *NablaB, const double *W, const double *DivW,
             const double *S, const double *A, const double *DivA,
             double *out);

/* ComputeDensity accumulation over states (Densities.f90:431-534):
 * rho[2][N] tau[2][N] nablaJ[2][N] and, when trc == 0, vecj[2][3][N],
 * vecs[2][3][N].  Outputs are overwritten. */
void or_density(const double *psi, const double *der, int M, int N,
                const int *qn, const double *occ, int trc,
                double *rho, double *tau, double *nablaJ,
                double *vecj, double *vecs);

/* Derived density fields (Densities.f90:536-550,1003-1101):
 * derrho[2][3][N], laprho[2][N]; when trc == 0: rotvecj[2][3][N],
 * rots[2][3][N], ders[2][3 comp][3 dir][N]. */
void or_density_derived(const double *rho, const double *vecj,
                        const double *vecs, int nx, int ny, int nz, double dx,
                        int lag, const double *lagx, const double *lagy,
                        const double *lagz, int trc, double *derrho,
                        double *laprho, double *rotvecj, double *rots,
                        double *ders);

/* Energy-ordered modified Gram-Schmidt per (isospin, parity, signature)
 * block (SpwfStorage.f90:372-480, OrderSpwfsISO :795-847).  Time simplex
 * conserved: real inner products only. */
void or_gram_schmidt(double *psi, int M, int N, const int *qn,
                     const double *energy, double dv);

/* One GradDesc call with fixed dt and momentum (ImaginaryTime.f90:439-500).
 * psi, der, lap: current states and derivatives; upd[M][4][N]: heavy-ball
 * memory (zero on the first call).  On return psi holds the orthonormalised
 * states, upd the new differences, energy/disp the <h> and dispersions. */
void or_graddesc(double *psi, const double *der, const double *lap,
                 double *upd, int M, int N, const int *qn, int trc,
                 const double *kin2, const double *U, const double *B,
                 const double *NablaB, const double *W, const double *DivW,
                 const double *S, const double *A, const double *DivA,
                 double dt_over_hbar, double momentum, double dv,
                 double *energy, double *disp);

/* ---- self-consistent part (mocca_oracle_scf.c) --------------------------- */

/* r^l Re Y_lm at (X,Y,Z) (SphericalHarmonics, Moments.f90:584-691). */
double or_rlylm(int l, int m, double X, double Y, double Z);

/* EV8 electric multipole moments (even l <= lmax, even m, real part) in the
 * list order of IniMoments (Moments.f90:372-485); lm[2k], lm[2k+1] = l, m;
 * val[2k+it] = dv sum rho_it r^l Y_lm (Calculate_Multipole :1907-2004).
 * Returns the number of moments. */
int or_moments(const double *rho, int nx, int ny, int nz, double dx, double dv,
               int lmax, int *lm, double *val);

/* SolveCoulomb with the conjugate-gradient solver (Coulomb.f90:255-340,
 * 583-674, 755-900).  cpot[(nz+2)][(ny+2)][(nx+2)] is the persistent complete
 * potential; first != 0 fills it with InitialGuess (:542-581) first.
 * qp[k] = proton moment k of the list of or_moments.  Returns the number of
 * CG iterations taken. */
int or_coulomb(const double *rho_p, int nx, int ny, int nz, double dx, double e2,
               double nucleons, double protons, int nmom, const int *lm,
               const double *qp, double *cpot, int first, double *vcoul,
               double *vcoulx, int maxit);

/* Mean-field potentials of an SLy4-type functional (MeanFields.f90:115-150,
 * 269-409, 411-492, 494-597, 618-712).  coef = B1..B21 in Force.f90 order
 * (28 values).  Layouts as in or_hpsi. */
void or_potentials(const double *coef, int nx, int ny, int nz, double dx, int lag,
                   const double *lagx, const double *lagy, const double *lagz, int trc,
                   const double *rho, const double *tau, const double *nablaJ,
                   const double *derrho, const double *laprho, const double *vecj,
                   const double *vecs, const double *rotvecj, const double *rots,
                   const double *vcoul, const double *vcoulx, int cexchange,
                   double *U, double *B, double *NablaB, double *W, double *DivW,
                   double *S, double *A, double *DivA);

/* Energy (CompEnergy, Energy.f90:62-179 with compSkyrme :623-859 and
 * CompCOMCorrection :1608-1660).  consts = hbm(2), mass(2), e2, hbar, dv, N,
 * Z.  out[0] total, [1..2] kinetic, [3] Coulomb, [4] exchange, [5..6]
 * one-body c.o.m., [7..42] the 36 Skyrme terms. */
void or_energy(const double *coef, const double *consts, int com1body, int trc,
               int nx, int ny, int nz, const double *rho, const double *tau,
               const double *nablaJ, const double *laprho, const double *vecj,
               const double *vecs, const double *rots, const double *vcoul,
               const double *kin_sp, int cexchange, double *out);

/* sum_l occ_l <psi_l|lap_l> per isospin (CompKinetic, Energy.f90:1426-1456). */
void or_kinetic_sums(const double *psi, const double *lap, int M, int N,
                     const int *qn, const double *occ, double dv, double *kin_sp);

/* d = (1-a) d + a h (MixDensities scheme 0, DensityMixing.f90:39-45). */
void or_mix(double *d, const double *h, size_t count, double a);

/* HF occupations by NaiveFill (HartreeFock.f90:177-262). */
void or_hf_fill(const double *energy, const int *qn, int M, double neutrons,
                double protons, int trc, double *occ);

/* IterativeEstimation for HF (ImaginaryTime.f90:28-211) on maxspwf psi[4][N]
 * (normalised first when start != 0), quantum numbers qn_est = iso, P, S,
 * TS; out[6] = { dt, momentum, maxE, relE, con, iterations }. */
void or_estimate(double *psi, int start, int nx, int ny, int nz, double dx, int lag,
                 const double *lagx, const double *lagy, const double *lagz,
                 const int *qn_est, int trc, double kin, const double *U, const double *B,
                 const double *NablaB, const double *W, const double *DivW,
                 const double *S, const double *A, const double *DivA, double dt,
                 double momentum, double dv, double hbar, int M, const int *qn,
                 const double *occ, const double *energy, double rel_hfb, double *out);

/* PreconditionPotential (Damping.f90:158-228): per isospin, CG on
 * (beta + alpha Lap) x = pot[it] (Lap with ghost signs sx, sy, sz), x = 0
 * start, <= 300 iterations, stop at dv |r|^2 < 1e-8.  x[2][N], iters[2]. */
void or_precondition(const double *pot, double *x, int nx, int ny, int nz, double dx, int lag,
                     const double *lagx, const double *lagy, const double *lagz, int sx, int sy,
                     int sz, double alpha, double beta, double dv, int *iters);

/* RutzCutOff (Moments.f90:2703-2849): density cutoff function of the
 * multipole constraints, rho[2][N] -> cutoff[2][N]. */
void or_rutz_cutoff(const double *rho, int nx, int ny, int nz, double dx, double dv,
                    double radd, double acut, double *cutoff);

#ifdef __cplusplus
}
#endif
#endif
