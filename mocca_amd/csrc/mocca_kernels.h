// Device kernels of the mocca_amd library (gfx950, wave64).
//
// Storage: states psi[i][c][N] (x fastest), fields [it][...][N].  All kernels
// are templated on the storage/stencil type T (double for the FP64 path,
// float for the mixed path); reductions that feed Gram-Schmidt and energies
// accumulate in double.
#pragma once
#include <hip/hip_runtime.h>

namespace mocca {

constexpr int kBlock = 256;      // threads per workgroup for point kernels
constexpr int kChunk = 8;        // states per chunk (one isospin per chunk)
constexpr int kGramTile = 64;    // Gram tile edge (16x16 threads x 4x4)
constexpr int kGramKC = 16;      // K sub-chunk staged in LDS by the Gram kernel
constexpr int kApplyL = 32;      // output states per workgroup in the apply
constexpr int kApplyChunks = 1;  // chunks of kBlock values per apply workgroup
constexpr int kMaxBlock = 512;   // largest (isospin, parity, signature) block

// Raises the dynamic-LDS limit of kernel f to bytes.  A failure is recorded
// under the kernel's name and reported by the next take_launch_failure()
// (the C-ABI layer checks it after every launch), so no kernel is launched
// with a smaller LDS allocation than it was written for.
bool require_lds(const void *f, size_t bytes, const char *kernel);
// The kernel whose LDS limit could not be raised since the last call, or
// nullptr; clears the record.
const char *take_launch_failure();

// Mesh description passed by value to kernels.
template <typename T>
struct MeshK {
    int nx, ny, nz, N;
    int lag;                  // 1 = Lagrange matrices, 0 = EV8 finite differences
    T inv60dx;                // 1/(60 dx) of the 7-point first derivative
    T l2m, l2c2, l2c3, l2c4;  // 9-point second-derivative weights
    T l2scale;                // -1/(5040/(-8064) dx^2)
    // Lagrange matrices (row-major n x n): first[axis][sign(0:-,1:+)],
    // second[axis][sign].
    const T *lag1[3][2];
    const T *lag2[3][2];
};

// 9-point second derivative of v[0..8] centred on v[4] (lapla_*_ev8,
// OptimizedDerivatives.f90:617-795).  FP64 keeps the reference's sum with
// the centre weight; FP32 sums the differences from the centre instead, which
// is the same operator (l2m = -2 (1 + l2c2 + l2c3 + l2c4)) but maps a constant
// to exactly 0: the rounded FP32 centre weight would leave an offset of
// ~1e-7 |v| / dx^2, a systematic bias of the kinetic energy (~1 keV for 200
// nucleons).
template <typename T>
__device__ __forceinline__ T lap9(const MeshK<T> &m, const T v[9]) {
    if constexpr (sizeof(T) == 4) {
        const T c = v[4] + v[4];
        return m.l2scale * (((v[5] + v[3]) - c) + m.l2c2 * ((v[6] + v[2]) - c) +
                            m.l2c3 * ((v[7] + v[1]) - c) + m.l2c4 * ((v[8] + v[0]) - c));
    } else {
        return m.l2scale * (m.l2m * v[4] + (v[5] + v[3]) + m.l2c2 * (v[6] + v[2]) +
                            m.l2c3 * (v[7] + v[1]) + m.l2c4 * (v[8] + v[0]));
    }
}

// Chunk of states of one isospin: states order[start .. start+count-1].
struct Chunk {
    int start, count, it, pad;
};

// One job of the field-derivative kernel: dst = sum_t coef[t] * D_t(src[t]),
// with D_t the first derivative along axis[t] (0..2) or the Laplacian
// (axis 3) with per-axis ghost signs sign[t][0..2].
struct FieldJob {
    const double *src[3];
    double *dst;
    int axis[3];
    int sign[3][3];
    double coef[3];
    int nterms;
    int pad;
};

// One MixingScheme 2 solve: pot <- old + (beta + alpha Lap)^-1 (pot - old),
// Laplacian ghost signs sign[axis].
struct PrecJob {
    double *pot;
    const double *old;
    int sign[3];
    int pad;
};

enum HpsiMode { kWriteH = 1, kStep = 2, kKinetic = 4 };   // kKinetic: only <psi|Lap psi> partials

// Field buffer offsets, in units of N (one contiguous buffer per handle):
// U[2] B[2] NablaB[2][3] W[2][nu][mu] DivW[2][3] S[2][3] A[2][3] DivA[2]
constexpr int kOffU = 0, kOffB = 2, kOffNB = 4, kOffW = 10, kOffDW = 28,
              kOffS = 34, kOffA = 40, kOffDA = 46, kNumFieldSlots = 48;

// One workgroup task of the z-marching stencil kernels: a y-tile of one
// state (all x, rows [tile*TY, tile*TY+TY), all z).
struct Task {
    int state, tile;
};

// Density task of the z-marching density kernel.
struct DTask {
    int tile, slab, chunk, pad;
};
constexpr int kDensSlab = 4;     // z planes per density task
// FP32 states with the 6 time-odd fields: 2 planes per task (twice the
// tasks, fewer accumulators per thread; in FP64 the longer slab is faster)
constexpr int kDensSlabTodd32 = 2;
inline int density_slab(int trc, int tsize) {
    return !trc && tsize == 4 ? kDensSlabTodd32 : kDensSlab;
}
constexpr int kDensChunk = 16;   // states per density chunk

// Rows per y-tile of the z-marching kernels (x spans the full line).
inline int plane_tile_rows(int nx, int ny) {
    int ty = kBlock / nx;
    return ty < ny ? ty : ny;
}

// Finite-difference hPsi (+ heavy-ball step) marching along z: register
// queue of 9 planes per component, LDS tile of the current plane with y
// halos.  Requires nx <= kBlock.  wa: W antisymmetric with a zero diagonal
// (3 of its 9 components read).
template <typename T>
void launch_hpsi_fd(const MeshK<T> &m, const T *psi, T *psi_out, T *upd, T *hout,
                    const T *fields, const int *qn, const Task *tasks, int ntasks,
                    int TY, int ntile_y, int trc, T kin0, T kin1, T dth, T mu,
                    int mode, double *eparts, hipStream_t st, const T *vsrc = nullptr,
                    bool wa = false);
// The FP64 GradDesc step (no stale derivatives) with the next plane's halo
// rows and potentials moved to LDS by direct-to-LDS loads while the current
// plane is computed; tasks carry tile + ntile_y * slab for nzs z slabs, and
// eparts gets ntile_y * nzs slots per state.  w_antisym: the spin-orbit
// field W is antisymmetric with a zero diagonal (3 components moved, not 9).
bool hpsi_dma_ok(const MeshK<double> &m, int TY, int trc);
// <psi|Lap psi> partial sums of every state by the pipelined FP64 kernel
// (tasks, eparts as launch_hpsi_fd_dma; element 2 of each slot).
void launch_hpsi_kinetic_dma(const MeshK<double> &m, const double *psi, const int *qn,
                             const Task *tasks, int ntasks, int TY, int ntile_y, int nzs,
                             int trc, double *eparts, hipStream_t st);
void launch_hpsi_fd_dma(const MeshK<double> &m, const double *psi, double *psi_out,
                        const double *prev, const double *fields, const int *qn,
                        const Task *tasks, int ntasks, int TY, int ntile_y, int nzs, int trc,
                        double kin0, double kin1, double dth, double mu, double *eparts,
                        hipStream_t st, bool w_antisym);

// DiagSpin / DiagAng of M states (Spwf.f90:977-1489): out[M][21] = J(3)
// JTR(3) JTI(3) J2(3) Sp(3) SpTR(3) SpTI(3) (the entries the EV8 symmetries
// zero are left untouched: clear out first); part: [nblk][M][kAmSums] with
// nblk = ceil(N / kBlock).
constexpr int kAmSums = 9;
template <typename T>
void launch_angmom(const MeshK<T> &m, const T *psi, const int *qn, int M, double dx, double dv,
                   double *part, double *out, hipStream_t st);
// Per-workgroup sums (ceil(N / kBlock) of them) of the z components of the
// spin and current density integrals of UpdateAM.
void launch_angmom_dens(const double *dens, int nx, int ny, long N, double dx, double *part,
                        hipStream_t st);
template <typename T>
void launch_derive_all(const MeshK<T> &m, const T *psi, const int *qn, int M,
                       double *der, double *lap, hipStream_t st);

template <typename T>
void launch_hpsi_step(const MeshK<T> &m, const T *psi, T *psi_out, T *upd,
                      T *hout, const T *fields, const int *qn, const int *order,
                      const Chunk *chunks, int nchunks, int trc, T kin0, T kin1,
                      T dth, T mu, int mode, double *eparts, int ntiles,
                      hipStream_t st, const T *vsrc = nullptr, const T *dsc = nullptr);
constexpr int kLagMaxN = 64;   // longest line of the matrix-core Lagrange derivatives
// Lagrange first and second derivatives of every state along x, y, z on the
// FP64 matrix cores (lines of up to kLagMaxN = 64 points):
// dsc[state][comp][6][N] = d/dx, d/dy, d/dz, d2/dx2, d2/dy2, d2/dz2.
template <typename T>
void launch_lag_derivs(const MeshK<double> &md, const T *psi, const int *qn, int M, T *dsc,
                       hipStream_t st);
// The same for FP32 states on the FP32 matrix cores, the Lagrange matrices
// split into a high and a low FP32 part (k_lag_axis_split).
void launch_lag_derivs_split(const MeshK<double> &md, const float *psi, const int *qn, int M,
                             float *dsc, hipStream_t st);

// E = dv sum psi.h, D = dv sum h.h - E^2, K = dv sum psi.Lap psi per state
// (any of E, D, K may be null).
void launch_reduce_ed(const double *eparts, int ntiles, const int *states,
                      int nstates, double dv, double *E, double *D, double *K,
                      hipStream_t st);

template <typename T>
void launch_density(const MeshK<T> &m, const T *psi, const int *qn,
                    const double *occ, const int *order, const Chunk *chunks,
                    int nchunks, int trc, T *part, hipStream_t st);

template <typename T>
void launch_density_reduce(const T *part, const Chunk *chunks, int nchunks,
                           int nf, int N, double *dens, hipStream_t st);

void launch_field_jobs(const MeshK<double> &m, const FieldJob *jobs, int njobs,
                       hipStream_t st);

// Gram matrices with TILE x TILE output tiles (tile = 32 or 64).
// max_m: leading dimension of the per-block arrays; mmax: largest block.
template <typename T>
void launch_gram(const T *psi, long K4, const int *blk_states, const int *blk_m,
                 int nblk, int max_m, int mmax, int tile, double *partial,
                 int nsplit, hipStream_t st);

void launch_gram_reduce(const double *partial, const int *blk_m, int nblk,
                        int max_m, int mmax, int tile, int nsplit, double dv,
                        double *G, hipStream_t st);

// Energy order, Cholesky and T = L^{-T} per block: in registers up to 63
// states, in LDS when the largest block fits in lds_limit bytes
// (cholesky_lds_limit()), else with its packed lower triangle in LDS (up to
// 192 states), else in global memory.
void launch_cholesky_ld(const double *G, const int *blk_states,
                        const int *blk_m, int nblk, int max_m, int mmax, const double *E,
                        double *Tm, int *perm, double *work, int *status,
                        size_t lds_limit, hipStream_t st);
size_t cholesky_lds_limit();
// Norm of each state before its normalisation in the Gram-Schmidt whose
// factor Tm = L^-T, perm holds: nrm[perm[b][r]] = L[r][r]^2 (the squared
// norm of the state less its projections on the lower-energy states of its
// block, Gramschmidt's CompNorm, SpwfStorage.f90:388-389).
void launch_gs_norms(const double *Tm, const int *perm, const int *blk_m, int nblk,
                     int max_m, double *nrm, hipStream_t st);

template <typename T>
void launch_gs_apply(const T *src, T *dst, T *upd, long K4, const int *perm,
                     const int *blk_m, int nblk, int max_m, const double *Tm,
                     int update_mode, hipStream_t st);
// The apply on the FP64 matrix cores for blocks of 64..128 states
// (mpad_max = the largest block padded to 16, a multiple of 16 in 64..128).
template <typename T>
void launch_gs_apply_mfma(const T *src, T *dst, long K4, const int *perm, const int *blk_m,
                          int nblk, int max_m, int mpad_max, const double *Tm,
                          hipStream_t st);

// out[list[b][l]] = sum_k src[list[b][k]] Tm[b][k][l] for a full (dense) Tm,
// on the matrix cores as the apply above; blocks of up to 192 states.
template <typename T>
void launch_dense_apply_mfma(const T *src, T *dst, long K4, const int *list, const int *blk_m,
                             int nblk, int max_m, int mpad_max, const double *Tm,
                             hipStream_t st);

// Gram matrices on the FP64 matrix cores (v_mfma_f64_16x16x4f64) for blocks
// of up to kMfmaMaxRows states (one wave holds all upper 16x16 tiles of up
// to 128 states; 129-160 states take three waves per K slice).
constexpr int kMfmaMaxRows = 160;
constexpr int kMfmaApplyMaxRows = 192;

template <typename T>
void launch_gram_mfma(const T *psi, long K4, const int *blk_states, const int *blk_m, int nblk,
                      int max_m, int mpad_max, double *part, int nsplit, double dv, double *G,
                      hipStream_t st);

template <typename T>
void launch_density_fd(const MeshK<T> &m, const T *psi, const int *qn, const double *occ,
                       const int *order, const Chunk *dchunks, const DTask *tasks,
                       int ntasks, int TY, int trc, T *part, hipStream_t st);

// ---- self-consistent part (mocca_scf.hip) ----------------------------------

constexpr int kMaxMom = 28;      // multipole moments (even l <= 12, even m)

// Skyrme-type functional: the B coefficients of Force.f90:38-41 and the
// constants the potentials and the energy need.
struct SkyrmeK {
    double B1, B2, B3, B4, B5, B6, B7a, B7b, B8a, B8b, byt3a, byt3b, B9, B9q, B10, B11,
        B12a, B12b, B13a, B13b;
    double hbm[2], mass[2], e2, neutrons, protons;
    int com1body, cexchange, coulomb, pad;
};

// Coulomb box: interior nx*ny*nz, box strides mx = nx+2, my = ny+2.
struct CoulombK {
    int nx, ny, nz, mx, my, mz;
    double f;      // 1/dx^2
    double prec;   // stop when sum r^2 <= prec
    double e2;
    int maxit, pad;
};

void launch_hf_fill(const double *E, const int *qn, int M, int ub_n, int ub_p, double fac,
                    double *occ, hipStream_t st);
void launch_mix(double *d, double *h, long n, double a, int keep, hipStream_t st);
void launch_moments(const double *rho_n, const double *rho_p, const double *Y, int nmom, long N,
                    double dv, double *part, int nblk, double *Q, hipStream_t st);
void launch_coulomb_bc(const double *Q, int nmom, const double *lfac, const int *bidx,
                       const double *bsh, const double *brp, int nb, double e2, double *x,
                       hipStream_t st);
// Cooperative launch of the whole CG solve; returns the hipError_t.
int launch_coulomb_cg(const CoulombK &c, double *x, const double *rho_p, double *r, double *pa,
                      double *pb, double *ap, double *part, int nwg, double *vcoul,
                      double *vcoulx, int *iters, hipStream_t st);
int coulomb_cg_max_groups();
// interior points of the Coulomb box up to which the CG runs as one workgroup
constexpr long kCgSingleMax = 4096;
// Direct solve: Q[a] (n_a x n_a, row-major, columns = eigenvectors) and
// lam[a] of the 1-D operators; w0, w1 scratch of nx*ny*nz doubles.
void launch_coulomb_direct(const CoulombK &c, double *x, const double *rho_p,
                           const double *const Q[3], const double *const lam[3], double *w0,
                           double *w1, double *vcoul, double *vcoulx, hipStream_t st);
void launch_potentials(const SkyrmeK &f, long N, int trc, const double *dens,
                       const double *derived, const double *vcoul, const double *vcoulx,
                       const double *ce, double *pot, hipStream_t st);
void launch_convert_f32(const double *a, float *b, long n, hipStream_t st);
// e[43] = [total, kin(2), E_C, E_Cx, com(2), terms 1..36]
void launch_energy_terms(const SkyrmeK &f, long N, int trc, const double *dens,
                         const double *derived, const double *vcoul, double *part, int nblk,
                         double dv, double *e, hipStream_t st);
void launch_energy_kinetic(const SkyrmeK &f, const double *K, const double *occ, const int *qn,
                           int M, double *e, hipStream_t st);
constexpr int kEnergySlots = 35;   // partial sums per workgroup of the energy pass

void launch_precondition(const MeshK<double> &m, const PrecJob *jobs, int njobs, double *work,
                         double alpha, double beta, double dv, int *iters, hipStream_t st);

// Multipole constraints (Moments.f90): one record per constraint, in the
// order of the moment list.
struct ConstraintK {
    int idx;             // moment slot (row of the Y table, of Q)
    int type;            // ConstraintType: 0 off, 1 augmented Lagrangian, 2 alternating
    int iso;             // Isoswitch: 1 neutrons + protons, 2 each isospin
    int iteration;       // switched off after this iteration (<= 0: never)
    double con[2], truec[2], inten[2];   // Constraint, TrueConstraint, Intensity
};
struct ConstraintCtl {
    double damping;      // Damping of the constraint field
    double slowdown;     // ReadjustSlowDown
    int first;           // next CalculateAllMoments is the first call
    int ce_nz;           // the constraint field is not all zero
};
enum ConMode { kConCalc = 1, kConReadjust = 2, kConTurnOff = 4 };
void launch_rutz_cutoff(const double *dens, int nx, int ny, int nz, double dx, double dv,
                        double radd, double acut, double *stat, int *cnt, double *surf,
                        double *cut, hipStream_t st);
void launch_con_squared(const ConstraintK *con, int nc, const double *Y, const double *dens,
                        long N, double dv, double *sq, hipStream_t st);
void launch_con_update(ConstraintK *con, int nc, ConstraintCtl *ctl, const double *Q,
                       const double *sq, const double *cstat, int mode, int iteration,
                       double momentum, double *ce_coef, double *alt, hipStream_t st);
// Value / OldValue bookkeeping of CalculateAllMoments (ConvergenceCheck, the
// wavefunction file): passes times, hist[c][1:8][:] <- hist[c][0:7][:] and
// hist[c][0][it] = the constraint's moment of isospin it.
void launch_con_history(const ConstraintK *con, int nc, const double *Q, double *hist,
                        int passes, hipStream_t st);
void launch_con_field(const ConstraintK *con, int nc, const double *Y, const double *coef,
                      const ConstraintCtl *ctl, const double *cut, long N, double *ce,
                      hipStream_t st);
template <typename T>
void launch_alternate(const ConstraintK *con, int nc, const double *Y, const double *alt,
                      const double *cut, long N, double *mult, const T *in, T *out,
                      const int *qn, int M, hipStream_t st);

// HFB pairing (time reversal, conserved signature): one block per
// (isospin, parity) with h states and their h partners.
constexpr double kHfbNumCut = 1e-12;   // HFBNumCut (HFB.f90)
constexpr int kHfbMaxN = 64;           // block size 2h of the in-LDS solvers
constexpr int kHlNewton = 12;
constexpr int kHfbEnergyParts = 16;   // workgroups per block of the pairing-energy sum          // Fermi-solver record of the large-block path
struct HfbBlockK {
    int it, P, h;
    int first;      // index of the block's first state in the state table
    int off;        // offset of its (2h)^2 matrices in the packed arrays
    int toff;       // offset of its h^2 canonical transform
};
struct HfbK {
    int nblk;
    HfbBlockK blk[4];
    double G[2], alpha[2], cut[2], mu[2];   // PairingStrength, Alpha, PairingCut, PairingMu
    double rhosat, mix, prec, particles[2];  // RhoSat, HFBMix, FermiPrec, N and Z
    int iters;                               // HFBIter
    const int *state;        // [M] states block by block
    const int *slot_block;   // [M] block of each state
    const int *slot_index;   // [M] position of each state in its block
};

// Pairing field of the large-block HFB step without phi in HBM: the dense
// apply phi_l = sum_k psi_list[k] Tm[k][l] of every block contracted with
// the states in the same kernel (per-block partials in fpart, nblk x 2 x 4N
// doubles), then field[2][2][N] = dfac x their sums over the blocks of each
// isospin.  False (nothing launched) for blocks above 128 states.
template <typename T>
bool launch_field_apply(const HfbK &k, const T *psi, long N, const int *list,
                        const int *blk_m, int max_m, int mpad_max, const double *Tm,
                        const double *dfac, double *fpart, double *field, hipStream_t st);

// IterativeEstimation (power iteration for dt and momentum).
constexpr int kEstSlots = 8;
template <typename T>
void launch_est_reset(T *psi, T *u, long n, int fill, unsigned long long seed, double *est,
                      int nblk, hipStream_t st);
template <typename T>
void launch_est_start(T *psi, long n, double *npart, int nblk, double dv, double *est,
                      hipStream_t st);
template <typename T>
void launch_est_iteration(T *psi, const T *hpsi, T *u, long n, const double *eparts, int slots,
                          double dv, double dth, double mu, int iter, int first, double tol,
                          double *est, double *npart, int nblk, hipStream_t st);
void launch_est_finalize(const double *E, const double *occ, const int *qn, int M, int pairing,
                         const double *qpe, int nqp, double hbar, double *est, hipStream_t st);

}  // namespace mocca
