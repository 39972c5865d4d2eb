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

#ifdef __cplusplus
}
#endif
#endif
