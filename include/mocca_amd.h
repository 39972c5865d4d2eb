/*
 * mocca_amd -- MI355X (gfx950) library for the imaginary-time iteration hot
 * path of the MOCCa Skyrme-HF(B) mesh solver.  C ABI only: plain pointers,
 * sizes and status codes, so Fortran (iso_c_binding), C and Python (ctypes)
 * bind it directly.
 *
 * Each entry point replaces one module procedure of the reference (cited as
 * path:line in the reference tree).  A solve handle owns all device memory:
 * the states of HFBasis, the heavy-ball memory, mean-field potentials and
 * densities.  Host arrays stay owned by the caller.
 *
 * Layouts (x fastest, identical to the reference's Fortran arrays):
 *   field     f[(z*ny + y)*nx + x],  N = nx*ny*nz stored points
 *   spinor    grid[c][N], c = 0..3 = (Re up, Im up, Re down, Im down)
 *             (Spinor.f90:6-19: Grid(nx,ny,nz,4,1))
 *   states    grid[i][c][N] for i = first .. first+count-1
 *   qn        qn[i][4] = (isospin -1/+1, parity, signature, time simplex)
 *   U, B      [2][N]        (isospin 0 = neutron, 1 = proton)
 *   NablaB    [2][3][N]     DivW [2][3][N]    S, A [2][3][N]   DivA [2][N]
 *   W         [2][nu][mu][N]  (WPot(:,:,:,mu,nu,it), MeanFields.f90:80-95)
 *
 * Errors: every call returns 0 on success or a negative MOCCA_ERR_* code;
 * mocca_last_error() then describes it.  The reference aborts through stp
 * (GenInfo.f90:91-177); a Fortran shim maps a non-zero status to
 * stp(mocca_last_error(h)).  No call returns partial results silently.
 *
 * Threading: a handle is not thread-safe; its work is ordered on one HIP
 * stream that the handle creates.  Calls that return host data synchronise.
 */
#ifndef MOCCA_AMD_H
#define MOCCA_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

#define MOCCA_OK 0
#define MOCCA_ERR_ARG -1      /* invalid argument or size */
#define MOCCA_ERR_UNSUPPORTED -2
#define MOCCA_ERR_HIP -3      /* HIP runtime error (message has details) */
#define MOCCA_ERR_STATE -4    /* call order: something was not set yet */
#define MOCCA_ERR_NUMERIC -5  /* non-positive norm in Gram-Schmidt, NaN */

#define MOCCA_FP64 0          /* everything in FP64 */
#define MOCCA_MIXED 1         /* FP32 states/stencils/hPsi/densities,
                                 FP64 Gram-Schmidt and energy sums */

typedef struct mocca_solve mocca_solve;

/* Mesh and symmetry set-up: GenInfo (nx, ny, nz, dx, dv after
 * AssignSymmetries, GenInfo.f90:220-297), the derivative scheme of
 * AssignFDCoefs (Derivatives.f90:206-346): lagrange = 1 for
 * MaxFDOrder = MaxFDLapOrder = -1 (Opt_LAG*_EV8, Lapla_LAG_EV8), 0 for the
 * EV8 finite differences (MaxFDOrder = 3, MaxFDLapOrder = 4).  Parity,
 * signature and time simplex must be conserved (the EV8 octant). */
typedef struct mocca_mesh {
    int nx, ny, nz;
    double dx;
    double dv;
    int time_reversal;   /* 1 = TRC conserved (no time-odd fields) */
    int lagrange;
} mocca_mesh;

/* Mean-field potentials of one iteration (MeanFields.f90:80-95), host
 * pointers in the layouts above.  S, A, DivA are read only when
 * time_reversal == 0.  kin[it] = hbm(it)/2 * reduced-mass factor(it)
 * (ActionOfBNew, MeanFields.f90:826-838). */
typedef struct mocca_fields {
    const double *U, *B, *NablaB, *W, *DivW, *S, *A, *DivA;
    double kin[2];
} mocca_fields;

/* Densities (DensityVector, Densities.f90:39-94), host pointers; any may be
 * NULL.  rho, tau, nablaJ, laprho [2][N]; derrho, vecj, vecs, rotvecj,
 * rots [2][3][N]; ders [2][3 comp][3 dir][N].  Time-odd members are only
 * written when time_reversal == 0. */
typedef struct mocca_density {
    double *rho, *tau, *nablaJ, *derrho, *laprho;
    double *vecj, *vecs, *rotvecj, *rots, *ders;
} mocca_density;

/* Create / destroy a solve on the current HIP device.  nwt = number of
 * states (SpwfStorage nwt).  No reference counterpart beyond the
 * allocations of ReadSpwfStorageInfo/NewSpinor (SpwfStorage.f90:77-97). */
int mocca_solve_create(const mocca_mesh *mesh, int nwt, int precision,
                       mocca_solve **out);
int mocca_solve_destroy(mocca_solve *h);
const char *mocca_last_error(const mocca_solve *h);
int mocca_synchronize(mocca_solve *h);

/* Quantum numbers and occupations of HFBasis(i) (Spwf.f90:14-194,
 * SetOcc :624).  Must be called before any compute call. */
int mocca_set_states_meta(mocca_solve *h, const int *qn, const double *occ);
int mocca_set_occupations(mocca_solve *h, const double *occ);
/* HFBasis(i)%Energy as used by the Gram-Schmidt ordering (SetEnergy). */
int mocca_set_energies(mocca_solve *h, const double *energy);
int mocca_get_energies(mocca_solve *h, double *energy, double *disp);

/* HFBasis(first+k)%Value%Grid <-> host (SetGrid / GetValue). */
int mocca_upload_states(mocca_solve *h, int first, int count,
                        const double *grid);
int mocca_download_states(mocca_solve *h, int first, int count, double *grid);
/* Heavy-ball memory `updates` of GradDesc (ImaginaryTime.f90:24,456-461). */
int mocca_reset_updates(mocca_solve *h);

int mocca_set_potentials(mocca_solve *h, const mocca_fields *f);

/* DeriveAll (SpwfStorage.f90:586-604) -> CompDer of every state: first
 * derivatives and Laplacian kept on the device; download one state's
 * der[3][4][N] and lap[4][N]. */
int mocca_derive_all(mocca_solve *h);
int mocca_download_derivatives(mocca_solve *h, int state, double *der,
                               double *lap);

/* hPsi(HFBasis(state)) (ImaginaryTime.f90:213-304) into host out[4][N]. */
int mocca_hpsi(mocca_solve *h, int state, double *out);

/* GradDesc (ImaginaryTime.f90:439-500) with given dt/hbar and momentum:
 * hPsi of every state, energies and dispersions, heavy-ball step, then
 * Gramschmidt and the update of the heavy-ball memory.  Derivatives are
 * formed on the fly from the current states. */
int mocca_grad_desc(mocca_solve *h, double dt_over_hbar, double momentum);

/* Gramschmidt (SpwfStorage.f90:372-480): ascending-energy orthonormalisation
 * per (isospin, parity, signature) block, done as Cholesky QR in FP64. */
int mocca_gram_schmidt(mocca_solve *h);

/* The orthonormalisation of mocca_gram_schmidt without any allocation, on
 * FP64 states the caller keeps in device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).
 *   psi, out   device, double [nwt][4][n]; out must not overlap psi
 *   qn         host, int [nwt][4] as in mocca_set_states_meta (blocks by
 *              isospin, parity, signature)
 *   energy     device, double [nwt]: the ordering energies (mocca_set_energies)
 *   dv         volume element of the inner products
 *   status     device int or NULL: set to 0, then non-zero if a block had a
 *              non-positive norm (the call does not synchronise, so it
 *              cannot report that in its return code)
 *   stream     hipStream_t to order the work on (NULL: the default stream)
 * Its workspace is one static device array of 32 MiB per GPU, held until
 * the process ends; nwt <= 512, blocks of at most 512 states, and sizes
 * whose workspace exceeds the array fail with MOCCA_ERR_UNSUPPORTED (when
 * only the Gram partials do not fit, fewer K splits are taken).  Every call
 * on one GPU shares that array: calls that overlap on one GPU race.
 * Errors are described by mocca_last_error(NULL). */
int mocca_gram_schmidt_noalloc(const double *psi, double *out, int nwt, long n, const int *qn,
                               const double *energy, double dv, int *status, void *stream);

/* ComputeDensity(Density, .false.) (Densities.f90:431-1187, SLy4-type
 * functional: no tensor or N2LO densities). */
int mocca_compute_density(mocca_solve *h, mocca_density *out);

/* ---- self-consistent part of the iteration --------------------------------
 * The loop of Evolve (Main.f90:255-328) for Hartree-Fock with fixed dt and
 * momentum (ParameterEstimation = 0) and linear density mixing
 * (MixingScheme = 0): EvolveSpwf (GradDesc), SolvePairing (NaiveFill),
 * UpdateDensities, MixDensities, CalculateAllMoments, ConstructPotentials
 * (SolveCoulomb + Calc*Pot), CompEnergy.  Everything stays on the device;
 * only the energies come back. */

/* Functional and constants (Force.f90:38-41, ReadForceInfo): B[] = B1, B2,
 * B3, B4, B5, B6, B7a, B7b, B8a, B8b, byt3a, byt3b, B9, B9q, B10, B11, B12a,
 * B12b, B13a, B13b, B14 .. B21 (28 values, Force.f90 order).  Tensor and
 * N2LO terms (B14 .. B21, t1n2, t2n2) are not supported: non-zero B14 .. B21
 * returns MOCCA_ERR_UNSUPPORTED. */
typedef struct mocca_functional {
    double B[28];
    double hbm[2];        /* Force.f90 hbm after ReadForceInfo doubles it: hbar^2/m,
                             so the kinetic coefficient is hbm/2 */
    double mass[2];       /* nucleonmass */
    double e2;            /* e^2 (Force.f90 e2) */
    double neutrons, protons;
    int com1body;         /* COM1Body: 2 = one-body c.o.m. in the kinetic term */
    int cexchange;        /* Cexchange: 0 none, 1 perturbative, 2 self-consistent */
    int coulomb;          /* 0 = no Coulomb (CoulombSolver 0);
                             1 = conjugate gradients, the reference's iterate
                                 (CoulombSolver 1, ConjugGrad);
                             2 = direct solve of the same discrete problem
                                 (exact to rounding; no iterations) */
    int coulomb_maxit;    /* CoulombMax (500), for coulomb = 1 */
    int max_moment;       /* MaxMoment = MaxLCoul (10, even, <= 12) */
    double damping;       /* DampingParam of MixingScheme 0 (0.75) */
    int mixing_scheme;    /* MixingScheme (Densities.f90:27): 0 = linear density mixing
                             with damping; 2 = no density mixing, U and S (without
                             time reversal) preconditioned in ConstructPotentials
                             (PotentialMixing.f90:143-171) */
    int pad_;
    double preconu;       /* preconu (Densities.f90:34, 1.0): the preconditioner is
                             (1 - preconu Lap)^-1 */
} mocca_functional;

/* Set the functional, the kinetic coefficients hbm/2 * reduced-mass factor
 * (MeanFields.f90:826-838), the multipole tables (IniMoments,
 * Moments.f90:372-485) and the Coulomb box with its initial guess
 * (SetupCoulomb + InitialGuess, Coulomb.f90:342-581). */
int mocca_scf_setup(mocca_solve *h, const mocca_functional *f);

/* SolvePairing for HF: NaiveFill (HartreeFock.f90:177-262) from the
 * current single-particle energies. */
int mocca_hf_fill(mocca_solve *h);
int mocca_get_occupations(mocca_solve *h, double *occ);

/* UpdateDensities(0) (Densities.f90:406-429): the current densities become
 * DensityHistory(1), then ComputeDensity. */
int mocca_update_densities(mocca_solve *h);

/* MixDensities, MixingScheme 0 (DensityMixing.f90:22-95):
 * density = (1 - damping) density + damping DensityHistory(1). */
int mocca_mix_densities(mocca_solve *h);

/* CalculateAllMoments (Moments.f90:1907-2004, electric moments) and
 * ConstructPotentials (PotentialMixing.f90:17-201): SolveCoulomb with the
 * multipole boundary values (Coulomb.f90:255-340, 583-674) and the
 * conjugate-gradient (ConjugGrad :755-900) or direct solve, then
 * CalcUPot, CalcWPot, CalcBPot, CalcSPot, CalcAPot (MeanFields.f90:115-712)
 * into the potentials used by hPsi / GradDesc. */
int mocca_construct_potentials(mocca_solve *h);

/* CompEnergy (Energy.f90:62-179): e[47] = total, kinetic(2), Coulomb,
 * Coulomb exchange, one-body c.o.m.(2), Skyrme terms 1..36
 * (compSkyrme, Energy.f90:623-859), pairing energy(2) and Fermi energy(2)
 * (zero for HF; HFBEnergy, HFB.f90:296-348). */
int mocca_comp_energy(mocca_solve *h, double *e);

/* Start-up of Evolve (Main.f90:172-225): SolvePairing, UpdateDensities,
 * CalculateAllMoments, ConstructPotentials, CompEnergy. */
int mocca_scf_start(mocca_solve *h);

/* Arithmetic of the Lagrange derivatives (Opt_LAG{X,Y,Z}_EV8,
 * Lapla_LAG_EV8, OptimizedDerivatives.f90:115-148,306-338,432-464,551-615)
 * on lines of 16-64 points, which run as small GEMMs on the matrix cores:
 * mode 0 (default) the FP64 matrix cores (mixed precision: FP32 states
 * widened, FP64 sums); mode 1, mixed precision only, the FP32 matrix cores
 * with each FP64 Lagrange matrix split into two FP32 parts (high and low)
 * and FP32 sums. */
int mocca_set_lagrange_precision(mocca_solve *h, int mode);

/* UpdateAM (Cranking.f90:135-232) with DiagSpin and DiagAng
 * (Spwf.f90:977-1489) of every state, derivatives formed on the fly from the
 * current states.  states[M][21] (may be NULL) for the HF basis and
 * canonical[M][21] (may be NULL; HFB only) for the canonical basis: J(3)
 * JTR(3) JTI(3) J2(3) Sp(3) SpTR(3) SpTI(3) in units of hbar (hbar^2 for
 * J2), with the components that signature and time-simplex symmetry
 * remove from the EV8 octant at zero, as the reference sets them.
 * totals[39] (may be NULL) over the density basis with its occupations,
 * zero with time reversal as in the reference: TotalAngMom(3), JTR(3),
 * JTI(3), J2Total(3), TotalAngMom_dens(3) (the integrals of the spin and
 * current densities of the handle's last densities), AMIsoblock(2,2,2,3)
 * in Fortran order (signature, parity, isospin, direction). */
int mocca_angular_momentum(mocca_solve *h, double *states, double *canonical,
                           double *totals);

/* ConvergenceCheck (Main.f90:368-397) after the last CompEnergy of the
 * handle (mocca_scf_start, each iteration of mocca_iterate and its variants,
 * mocca_comp_energy, mocca_final_iteration).  prec[4] = MomentPrec,
 * EnergyPrec, PairingPrec, CrankPrec (the reference's defaults: 1e-4,
 * 1e-9, 1e-4, 1e-4).  out[6] = converged (all four tests), the
 * "Positive total energy!" condition on which the reference stops (total
 * energy >= 0), then the tests one by one: ConverMultipoleAll
 * (Moments.f90:2628-2701: each constrained moment's last 8 values within
 * prec[0] of its target, relative unless the target is 0), ConverEnergy
 * (Energy.f90:1458-1481: the total energy within prec[1] relative of each of
 * the 7 before it), ConverFermi (Pairing.f90:1048-1070: with HFB, |Fermi|
 * <= prec[2] for both isospins, the reference's FermiHistory staying zero;
 * always true for HF) and ConverCranking (Cranking.f90:602-619: true, no
 * cranking constraints are offered). */
int mocca_convergence_check(mocca_solve *h, const double *prec, int *out);

/* FinalIteration (Main.f90:482-529): the derivatives of the density basis
 * with the Lagrange scheme, UpdateDensities(0), CompEnergy into row[47]
 * (the layout of mocca_comp_energy; may be NULL), then the densities again
 * with the handle's own scheme (finite differences), unmixed.  The Coulomb
 * potential is the last one solved, as in the reference. */
int mocca_final_iteration(mocca_solve *h, double *row);

/* niter iterations of the Evolve loop (Main.f90:255-328) on the device.
 * energies (may be NULL): [niter+1][47] in the layout of mocca_comp_energy,
 * row 0 for the states on entry, row k after iteration k.
 * Iterations are numbered as Evolve numbers them (the number that
 * TurnOffConstraints compares with a constraint's switch-off iteration):
 * the handle counts the iterations run since mocca_scf_start, so a run split
 * over several calls (of mocca_iterate, mocca_iterate_estimated or
 * mocca_batch_run) continues the count where the previous call ended. */
int mocca_iterate(mocca_solve *h, int niter, double dt_over_hbar, double momentum,
                  double *energies);

/* IterativeEstimation (ImaginaryTime.f90:28-211), called by GradDesc when
 * ParameterEstimation = 1 (:464): up to 500 power iterations of an extra
 * state towards the largest eigenvalue of h, heavy-ball steps with the
 * given dt and momentum, stopping at |change of <psi|h psi>| < 1e-2 MeV; then
 * relE (HF: lowest particle-hole energy of one isospin, at least 0.1 MeV),
 * and out[6] = { dt, momentum, maxE, relE, last change, iterations }.
 * At iteration 1 (or the first call on the handle) the extra state starts
 * from uniform [0,1) numbers (random_number in the reference) with the
 * quantum numbers of the last state (of the last neutron state when
 * protons = 0); later calls continue from where the previous one ended. */
int mocca_estimate_parameters(mocca_solve *h, int iteration, double dt, double momentum,
                              double *out);

/* Start vector grid[4][N] (the layout of mocca_upload_states) for the next
 * call that starts the estimation state, instead of random numbers; NULL
 * clears it.  Lets a caller reproduce a run whose random start it knows. */
int mocca_estimation_seed(mocca_solve *h, const double *grid);

/* mocca_iterate with ParameterEstimation = 1: before each GradDesc,
 * mocca_estimate_parameters with the current dt, momentum (dt in units of
 * 1e-22 s, not divided by hbar), whose results carry over to the next
 * iteration.  first_iteration numbers the first iteration (1 at the start
 * of a run); params (may be NULL) receives [niter][2] = the dt and momentum
 * used by each GradDesc. */
int mocca_iterate_estimated(mocca_solve *h, int first_iteration, int niter, double dt,
                            double momentum, double *energies, double *params);

/* niter iterations of the Evolve loop (as mocca_iterate) for each of nsolves
 * independent solves, each after its own mocca_scf_start: iteration k of
 * every solve is issued in turn, each on its solve's stream, so the device
 * runs the solves concurrently (the batched constraint scans that the
 * reference runs as separate jobs, one process each).  dt_over_hbar[i] and
 * momentum[i] per solve; energies (may be NULL, and any energies[i] may be
 * NULL) receives [niter+1][47] rows per solve.  The first failing solve stops
 * the batch; its mocca_last_error has the message, and mocca_last_error(NULL)
 * names it. */
int mocca_batch_run(mocca_solve *const *solves, int nsolves, int niter,
                    const double *dt_over_hbar, const double *momentum,
                    double *const *energies);

/* One multipole constraint (namelist MomentConstraint, Moments.f90:3576-3800):
 * the moment (l, m) of the density (real part, even l and m under the EV8
 * symmetries), ConstraintType 1 (augmented Lagrangian) or 2 (alternating
 * step, the default), Isoswitch 1 (neutrons + protons: value[0]) or 2
 * (value[0] neutrons, value[1] protons), the iteration after which it is
 * switched off (<= 0: never) and Intensity (0: estimated at the first
 * readjustment). */
typedef struct mocca_constraint {
    int l, m;
    int type;
    int isoswitch;
    int iteration;
    int pad_;
    double value[2];
    double intensity;
} mocca_constraint;

/* Multipole constraints for the run (ReadMomentData, Moments.f90:3558-3800),
 * with the Rutz cutoff parameters radd, acut (4.0, 0.4), the Damping of the
 * constraint field (0.0) and ReadjustSlowDown (0.01).  After
 * mocca_scf_setup and before mocca_scf_start; count = 0 removes them.  They
 * act through CalculateAllMoments in mocca_construct_potentials (cutoff,
 * <Q^2>, the constraint field added to U, MeanFields.f90:402-403),
 * mocca_readjust_constraints and mocca_alternate_step; mocca_iterate and
 * mocca_iterate_estimated run all three in the reference's order. */
int mocca_set_constraints(mocca_solve *h, const mocca_constraint *c, int count, double radd,
                          double acut, double damping, double readjust_slowdown);

/* ReadjustAllMoments + TurnOffConstraints(iteration) (Moments.f90:2336-2366,
 * 3516-3549); momentum is the current heavy-ball momentum (it halves the
 * estimated intensity of alternating constraints). */
int mocca_readjust_constraints(mocca_solve *h, int iteration, double momentum);

/* AlternateStep (ImaginaryTime.f90:306-437): psi <- psi - sum_c
 * 0.5 (<Q_c> - Q_c) / (<Q_c^2> + d0) cutoff Y_c psi, then Gramschmidt.  As in
 * the reference, the next mocca_grad_desc takes the derivatives of the
 * states before this step (their values after it). */
int mocca_alternate_step(mocca_solve *h);

/* State of the constraints: out[count][10] = l, m, type (0 once switched
 * off), target(2), intensity(2), <Q^2>(2), isoswitch. */
int mocca_get_constraints(mocca_solve *h, double *out);

/* Pairing (namelist Pairing, Pairing.f90:143-560).  type 0 = HF (NaiveFill,
 * the default), 1 = BCS (BCS.f90: gaps, Fermi energies and occupations of
 * the HF states, the loop of SolvePairing over at most `iterations`
 * (PairingIter, 50 for BCS) steps to FermiPrec; no Lipkin-Nogami), 2 = HFB
 * with a density-dependent contact interaction
 * -G (1 - alpha rho / rho_sat) (CompDensityFactor,
 * PairingInteraction.f90:265-354), SymmetricFermi cutoff factors, the
 * Broyden Fermi solver and HFBMix mixing of rho and kappa.  HFB here needs
 * time reversal and states of signature +1.  fermi: the start Fermi
 * energies (-10, -10 in the reference); guess_kappa: GuessKappa. */
typedef struct mocca_pairing {
    int type;
    int guess_kappa;
    int iterations;       /* HFBIter (50 for the Broyden solver); with BCS
                             PairingIter (50) */
    int solver;           /* Fermi-energy solver: 0 = automatic (1 while every
                             (isospin, parity) block holds at most 32 states,
                             else 4); 1 = the reference's Broyden iteration
                             (HFBFindFermiEnergyBroyden) in one workgroup per
                             isospin with the matrices in LDS, blocks of at most
                             32 states; 2 = the same iteration on the
                             large-block path (blocks of up to 128 states),
                             each evaluation a Jacobi diagonalisation;
                             3 = Newton's method with the exact derivative
                             dN/dlambda on the large-block path: fewer
                             evaluations to the same FermiPrec, but where the
                             reference's iteration stops unconverged it ends
                             elsewhere; 4 = the iteration of 2 with rho, kappa
                             and N(lambda) from the spectral projector
                             (1 + sign T) / 2 of each evaluation (matrix-core
                             products, no diagonalisation; the quasiparticle
                             energies are formed when asked for) */
    double strength[2];   /* PairingNeutron, PairingProton (MeV fm^3) */
    double alpha[2];      /* AlphaNeutron, AlphaProton */
    double cut[2];        /* CutNeutron, CutProton (5 MeV) */
    double mu[2];         /* PairingMu (0.5 MeV) */
    double rhosat;        /* RhoSat (0.16 fm^-3) */
    double hfb_mix;       /* HFBMix (0.1) */
    double fermi_prec;    /* FermiPrec (1e-12) */
    double fermi[2];
} mocca_pairing;

/* Select the pairing treatment of SolvePairing (after mocca_scf_setup,
 * before mocca_scf_start).  With HFB the canonical basis is the density
 * basis (Pairing.f90:812) and energy rows carry the pairing energies and
 * Fermi energies in elements 43..46. */
int mocca_set_pairing(mocca_solve *h, const mocca_pairing *p);

/* SolvePairing (Pairing.f90:743-890): HF filling, or the HFB step (pairing
 * field, gaps, Fermi energies, canonical basis, mixing).  Replaces
 * mocca_hf_fill when HFB is selected (mocca_hf_fill calls it then). */
int mocca_solve_pairing(mocca_solve *h);

/* Fermi energies, pairing energies and selected quasiparticle energies of
 * the last pairing step: out[4 + nwt] = Fermi(2), E_pair(2), qp energies
 * (the h largest of each block, block by block), and canonical occupations
 * occ[nwt] (may be NULL). */
int mocca_get_pairing(mocca_solve *h, double *out, double *canonical_occ);

/* Results of the last ConstructPotentials.  lm[k][2] = (l, m) and
 * q[k][2] = (neutron, proton) moments; *count moments (<= 28).
 * vcoul, vcoulx [N]; cpot [(nz+2)][(ny+2)][(nx+2)] (CPotComplete);
 * potentials in the layouts of mocca_fields.  Any pointer may be NULL. */
int mocca_get_moments(mocca_solve *h, int *lm, double *q, int *count);
int mocca_get_coulomb(mocca_solve *h, double *vcoul, double *vcoulx, double *cpot,
                      int *cg_iterations);
int mocca_get_potentials(mocca_solve *h, double *U, double *B, double *NablaB,
                         double *W, double *DivW, double *S, double *A, double *DivA);

/* ---- MOCCa v1 wavefunction files ------------------------------------------
 * The file WriteMOCCa_v1 writes and ReadMOCCa_v1 reads (InOut.f90:905-1416):
 * Fortran sequential unformatted records, each framed by its byte count
 * (int32, little endian) before and after; int32 integers and logicals
 * (1 = true), FP64 reals, complex as (re, im) FP64 pairs.  Records in order:
 * version (1); TotalEnergy, TotalEnergy - OldEnergy(1:7); nx, ny, nz, dx;
 * TRC, PC, SC, TSC, IC; nwt, neutrons, protons (reals); per state the grid
 * (nx,ny,nz,4) and Occupation, Energy, Dispersion, Norm, J(3), Isospin,
 * TimeSimplex, Parity, Signature, TimeReversal (WriteSpwf, Spwf.f90:2246-
 * 2262); the densities (WriteDensity, Densities.f90:1189-1240: Rho DerRho
 * LapRho | Tau | NablaJ | JMuNu (empty) | 5 time-odd records, empty with
 * time reversal); MaxFDOrder, MaxFDLapOrder, CoulombLapOrder; the force name;
 * t0 .. wsoq; name, hbar, hbm(2), COM1Body, COM2Body; B1 .. B21 (28);
 * PairingType, PairingStrength(2), alpha(2), PairingCut(2); PCutoffs (empty
 * without pairing); Fermi(2), LNLambda(2); blocksizes; U, V; HFBColumns;
 * RhoHFB, KappaHFB; CanTransfo; Omega(3), CrankValues(3); per constrained
 * moment two records (WriteMoment, Moments.f90:3330-3363).  Records of 2 GiB
 * or more are not supported. */
#define MOCCA_V1_MAX_MOMENTS 32

/* One constrained moment (WriteMoment / ReadMoment, Moments.f90:3330-3420).
 * With Isoswitch 1 the file repeats Intensity, Constraint and
 * TrueConstraint; the values here are the file's pairs. */
typedef struct mocca_v1_moment {
    int constraint_type, isoswitch, l, m, impart;
    double intensity[2], constraint[2], true_constraint[2];
    double value[2];          /* Value(2): neutrons, protons */
    double old_value[14];     /* OldValue(2,7), Fortran order */
} mocca_v1_moment;

/* Every record of a v1 file except the per-state and array ones. */
typedef struct mocca_v1_info {
    double energy;            /* TotalEnergy */
    double denergy[7];        /* TotalEnergy - OldEnergy(1:7) */
    int nx, ny, nz;
    double dx;
    int trc, pc, sc, tsc, ic; /* symmetries conserved (logicals) */
    int nwt;
    double neutrons, protons;
    int fd_order[3];          /* MaxFDOrder, MaxFDLapOrder, CoulombLapOrder */
    char force[200];          /* afor, blank padded, no terminator */
    double skyrme[14];        /* t0 x0 t1 x1 t2 x2 t3a x3a yt3a t3b x3b yt3b wso wsoq */
    double hbar, hbm[2];
    int com1body, com2body;
    double B[28];             /* B1 .. B21 in the order of mocca_functional.B */
    int pairing_type;
    double pairing_strength[2], pairing_alpha[2], pairing_cut[2];
    int has_cutoffs;          /* the PCutoffs record holds nwt values */
    double fermi[2], lnlambda[2];
    int hfb_size;             /* S = HFBSize: 2 nwt with time reversal, else nwt */
    int npar, niso;           /* Pindex, Iindex: 2 with parity / isospin conserved */
    int blocksizes[4];        /* blocksizes(npar, niso), Fortran order */
    double omega[3], crank_values[3];
    int nmoments;
    mocca_v1_moment moments[MOCCA_V1_MAX_MOMENTS];
} mocca_v1_info;

/* The per-state and array records, host memory (N = nx ny nz, S =
 * hfb_size).  On read any pointer may be NULL (that part is skipped); on
 * write grid, state and qn are required, NULL density members write zeros,
 * NULL cutoffs write an empty record, NULL HFB matrices write zeros. */
typedef struct mocca_v1_arrays {
    double *grid;             /* [nwt][4][N] HFBasis(i)%Value%Grid */
    double *state;            /* [nwt][7] Occupation, Energy, Dispersion, Norm, J(3) */
    int *qn;                  /* [nwt][5] Isospin, TimeSimplex, Parity, Signature,
                                 TimeReversal */
    double *cutoffs;          /* [nwt] PCutoffs */
    mocca_density dens;       /* rho, derrho, laprho, tau, nablaJ; without time
                                 reversal also vecj, rotvecj, vecs, rots, ders */
    double *u, *v;            /* complex U(S, 2S, 2, 2), V(S, 2S, 2, 2) */
    int *hfb_columns;         /* HFBColumns(S, npar, niso) */
    double *rho_hfb, *kappa_hfb;   /* complex (S, S, 2, 2) */
    double *can_transfo;      /* complex (S, S, npar, niso) */
} mocca_v1_arrays;

/* Read a v1 file (ReadMOCCa_v1, InOut.f90:1040-1416): every header record
 * into info, and the arrays whose pointers are set.  Host only (no device
 * use).  Returns MOCCA_ERR_ARG for a file that is not a v1 file or does not
 * hold the records its header announces. */
int mocca_v1_read(const char *path, mocca_v1_info *info, const mocca_v1_arrays *arrays);

/* Write a v1 file from info and host arrays (WriteMOCCa_v1,
 * InOut.f90:905-1038).  Host only.  A file read by mocca_v1_read and
 * written back is identical byte for byte. */
int mocca_v1_write(const char *path, const mocca_v1_info *info,
                   const mocca_v1_arrays *arrays);

/* Checkpoint of a solve as Evolve writes one (Main.f90:318-325 calling
 * WriteMOCCa_v1): the states with their occupations, energies, dispersions,
 * Gram-Schmidt norms and angular momenta (UpdateAM), the current (mixed)
 * densities, the energy history, the functional, the pairing (HFB: the
 * cutoff factors, Fermi energies, RhoHFB and KappaHFB; HF: KappaHFB as
 * GuessHFBMatrices(0) fills it, HFB.f90:3686-3715) and the constrained
 * moments.  U, V and CanTransfo are written as zeros (a restart rebuilds
 * them in its first SolvePairing).  extra gives what the handle does not
 * hold: fd_order[2], force, skyrme, hbar, com2body, omega, crank_values,
 * and for HF the Fermi energies; NULL takes CoulombLapOrder 2, a blank
 * name, zeros, hbar = 6.58211928, Fermi = -10. */
int mocca_v1_save(mocca_solve *h, const char *path, const mocca_v1_info *extra);

/* Restart a solve from a v1 file as Input/ReadMOCCa_v1 and Evolve's
 * start-up take one (InOut.f90:122-187,1040-1416; Main.f90:206-215): the
 * mesh and nwt must equal the handle's, and the file may not break a
 * symmetry the handle conserves.  Loads the states, quantum numbers,
 * occupations, single-particle energies and dispersions; with HFB selected
 * (mocca_set_pairing before this call) the Fermi energies and RhoHFB /
 * KappaHFB; unless recalc (the reference's Recalc) the file's densities,
 * which the next mocca_scf_start then uses instead of recomputing them.
 * info (may be NULL) receives the header, including the constrained
 * moments (ContinueMoment is left to the caller: mocca_set_constraints). */
int mocca_v1_load(mocca_solve *h, const char *path, int recalc, mocca_v1_info *info);

/* Kernel timing over the handle's stream with HIP events (for benchmarks):
 * enable = 1 starts recording each launch; mocca_profile_get returns the
 * summed milliseconds and launch count for a kernel family name
 * ("hpsi_step", "gram", "cholesky", "gs_apply", "density",
 * "density_reduce", "field_jobs", "derive_all", "reduce_ed"). */
int mocca_profile_enable(mocca_solve *h, int enable);
int mocca_profile_get(mocca_solve *h, const char *name, double *ms,
                      long *count);

#ifdef __cplusplus
}
#endif
#endif
