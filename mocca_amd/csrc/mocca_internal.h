// Internal state of a solve handle, shared by the C-ABI translation units
// (mocca_capi.hip: states, hPsi/GradDesc, Gram-Schmidt, densities;
// mocca_scf_capi.hip: filling, mixing, moments, Coulomb, potentials, energy).
#pragma once

#include "../../include/mocca_amd.h"
#include "mocca_kernels.h"

#include <map>
#include <string>
#include <vector>

namespace mocca {
namespace detail {

struct Pending {
    const char *name;
    hipEvent_t a, b;
};

struct Buffers {
    // psi: current states; psi2: stepped states; upd: the states before the
    // last GradDesc step (the heavy-ball memory is u = psi - upd)
    void *psi = nullptr, *psi2 = nullptr, *upd = nullptr, *hbuf = nullptr;
    void *fields = nullptr;        // T [48][N]
    void *dpart = nullptr;         // T [nchunks][9][N]
    void *lagT = nullptr;          // T lagrange matrices
    double *lagD = nullptr;        // double lagrange matrices
    double *der = nullptr, *lap = nullptr;
    double *dens = nullptr;        // [2][9][N]: rho tau nablaJ j(3) s(3)
    double *derived = nullptr;     // derrho[2][3] laprho[2] rotvecj[2][3] rots[2][3] ders[2][3][3]
    int *qn = nullptr;
    double *occ = nullptr, *E = nullptr, *D = nullptr, *K = nullptr;
    double *nrm = nullptr;         // [M] norm of each state in the last Gram-Schmidt
                                   // before its normalisation (the Spwf's Norm)
    double *eparts = nullptr;      // [M][tiles][3]: <psi|h psi>, <h psi|h psi>, <psi|Lap psi>
    int *order = nullptr;
    Chunk *chunks = nullptr;
    int *all_states = nullptr;
    int *blk_states = nullptr, *blk_m = nullptr, *perm = nullptr;
    double *G = nullptr, *Tm = nullptr, *work = nullptr, *gpart = nullptr;
    int *status = nullptr;
    FieldJob *jobs = nullptr;
    Chunk *one_chunk = nullptr;
    int *one_order = nullptr;
    Task *tasks = nullptr, *one_tasks = nullptr;
    Task *tasks_dma = nullptr;     // pipelined FP64 step: (state, tile + ntile_y * slab)
    void *lagscr = nullptr;        // T [M][4][6][N]: matrix-core Lagrange derivatives
    Chunk *dchunks = nullptr;
    DTask *dtasks = nullptr;
};

// Device buffers and settings of the self-consistent part.
struct Scf {
    bool set = false, started = false;
    SkyrmeK sk{};
    CoulombK ck{};
    double damping = 0.75;
    int ub_n = 0, ub_p = 0;
    double fill_fac = 1.0;
    int nmom = 0, lm[2 * kMaxMom] = {};
    int nb = 0, cg_nwg = 0, mom_nblk = 0, e_nblk = 0, npjobs = 0;
    bool coulomb_first = true;
    double *hist = nullptr;        // [18N] dens + [20N] derived (derrho laprho rotvecj rots)
    double *Y = nullptr;           // [nmom][N] r^l Y_lm on the mesh
    double *mpart = nullptr;       // [mom_nblk][kMaxMom][2]
    double *Q = nullptr;           // [kMaxMom][2]
    double *lfac = nullptr;        // [nmom] 4 pi / (2l+1)
    int *bidx = nullptr;           // [nb] boundary points of the Coulomb box
    double *bsh = nullptr, *brp = nullptr;  // [nmom][nb] r^l Y_lm and r^(2l+1) there
    double *cx = nullptr, *cr = nullptr, *cpa = nullptr, *cpb = nullptr, *cap = nullptr;
    double *cpart = nullptr;       // [2 cg_nwg]
    double *vcoul = nullptr, *vcoulx = nullptr;
    int *cg_iters = nullptr;
    double *cq[3] = {nullptr, nullptr, nullptr};    // direct solve: eigenvectors per axis
    double *clam[3] = {nullptr, nullptr, nullptr};  // and eigenvalues
    double *cw0 = nullptr, *cw1 = nullptr;          // [N] scratch
    double *pot = nullptr;         // FP64 potentials [48][N] (aliases fields in FP64)
    FieldJob *pjobs = nullptr;
    double *epart = nullptr;       // [e_nblk][kEnergySlots]
    double *ecur = nullptr;        // [kRowLen] energy of the current state
    double *erows = nullptr;       // [erow_cap][kRowLen]
    int erow_cap = 0;
    // ConvergenceCheck (Main.f90:368-397) state: on the host, TotalEnergy of
    // the last CompEnergy and OldEnergy(1:7) before it (Energy.f90:93-98),
    // and the Fermi energies of the last energy row; on the device, the last
    // 8 values of sum(Value) of each constrained moment (Value, then
    // OldValue(:,1:7), Moments.f90:1920-1932, per isospin)
    double ehist[8] = {};
    double efermi[2] = {};
    bool start_pending = false;    // the start-up row is not yet in ehist (HF)
    bool dens_loaded = false;      // dens/derived hold a wavefunction file's densities,
                                   // which the next mocca_scf_start keeps
    double *chist = nullptr;       // [ncon][8][2]
    // MixingScheme 2 (potential preconditioning)
    int mixing_scheme = 0;
    double preconu = 1.0;
    bool pot_built = false;        // a ConstructPotentials has run on the handle
    double *pold = nullptr;        // [8N] U(2), S(2x3) before the update
    double *pcwork = nullptr;      // [npc][4N] CG vectors
    PrecJob *pcjobs = nullptr;
    int *pc_iters = nullptr;       // [npc]
    int npc = 0;
    // multipole constraints
    int ncon = 0;
    ConstraintK *con = nullptr;    // [ncon] device records
    ConstraintCtl *cctl = nullptr;
    double *csq = nullptr, *ccoef = nullptr, *calt = nullptr;  // [ncon][2]
    double *cut = nullptr, *ce = nullptr, *mult = nullptr;     // [2][N]
    double *cstat = nullptr;       // [2][2] threshold, particles present
    int *ccnt = nullptr;           // [2] surface points
    double *csurf = nullptr;       // [2][3N][3]
    double radd = 4.0, acut = 0.4;
    std::vector<int> htype, hiter;  // host copy of the types and switch-off iterations
    void *alt_old = nullptr;       // states before the last AlternateStep (T [M][4][N])
    bool alt_pending = false;      // the next GradDesc takes its derivatives from alt_old
    bool alt_next = false;         // Evolve's AlternateCheck for the next iteration
    int iteration = 0;             // Evolve iterations run since mocca_scf_start
};

// The extra state of IterativeEstimation (maxspwf) and its records.
struct Est {
    bool ready = false;            // psi holds the state of an earlier call
    void *psi = nullptr, *h = nullptr, *u = nullptr;  // T [4N]
    double *eparts = nullptr;      // [slots][3] hPsi partial sums
    double *npart = nullptr;       // [nblk] norm partial sums
    double *rec = nullptr;         // [kEstSlots]
    int *qn = nullptr;             // [4]
    Task *tasks = nullptr;         // z-marching hPsi: state 0, every y tile
    int *order = nullptr;          // chunked hPsi: the one state
    Chunk *chunk = nullptr;
    int nblk = 0;
    std::vector<double> seed;      // start vector for the next first call
    unsigned long long calls = 0;
};

// HFB pairing state (PairingType 2).
struct Hfb {
    bool on = false, started = false, unconverged = false;
    // BCS pairing (PairingType 1): per state the gap delta[i], the
    // quasiparticle energy qpe[i] (Eqp), the cutoff factor pc[i]; the
    // pairing field in field[2][N] (real); no HFB blocks
    bool bcs = false;
    mocca_pairing in{};            // the parameters of the last mocca_set_pairing
    bool in_set = false;
    HfbK k{};
    int total = 0, hmax = 0;
    size_t lds_broyden = 0, lds_canonical = 0;
    double *kappa = nullptr, *rho = nullptr, *orho = nullptr, *okappa = nullptr;
    double *delta = nullptr;       // packed (2h)^2 per block
    double *fermi = nullptr;       // [2]
    double *pc = nullptr;          // [M] cutoff factors
    double *dfac = nullptr;        // [2][N] density factor
    double *field = nullptr;       // [2][2][N] pairing field (re, im)
    double *qpe = nullptr;         // [M] selected quasiparticle energies
    double *tcan = nullptr;        // packed h^2 canonical transforms
    void *can_psi = nullptr;       // T [M][4][N] canonical states (the density basis)
    double *can_occ = nullptr;     // [M] canonical occupations
    double *work = nullptr;
    int *status = nullptr;
    int *tables = nullptr;         // [3][M] state, slot_block, slot_index
    int *mflags = nullptr;         // [2] whether old rho / kappa or delta are non-zero
    double *epart = nullptr;       // [nblk][kHfbEnergyParts] pairing-energy partials
    // blocks above the in-LDS solver (mocca_hfb_large.hip): per block, n x n
    // matrices with leading dimension nmax (the largest 2h)
    bool large = false;
    bool newton = false;           // Newton's method instead of the Broyden iteration
    bool sign = false;             // particle numbers from the matrix sign function
    bool qpe_stale = false;        // qpe not formed by the last pairing step (sign path)
    bool can_warm = false;         // lVc holds the canonical basis of the last step
    double *lsig = nullptr;        // [nblk][parts] sign iteration: residual of each wave
    double *lhost = nullptr;       // pinned host copy of the Fermi-solve records
    int nmax = 0, lsplit = 1;
    long lmax = 0;                 // rotation-log entries per block
    size_t lds_limit = 0;
    bool lwarm = false;            // lV holds the eigenbasis of the last call
    double *lT = nullptr, *lV = nullptr, *lA = nullptr, *lW = nullptr, *lC = nullptr;
    double *lVc = nullptr;         // canonical-basis eigenvectors
    double *lw = nullptr, *lwc = nullptr;   // eigenvalues [nblk][nmax]
    int *lrank = nullptr;          // [nblk][nmax] eigenvalue ranks (descending)
    double2 *llog = nullptr;       // [nblk][lmax] Jacobi rotations
    int *llogij = nullptr;         // [nblk][lmax] their pairs, i | j << 16
    double *lAf = nullptr;         // [nblk][ld][ld] the matrix after the last Jacobi
    int *lrounds = nullptr;        // [2][nblk] rounds logged (HFB, canonical)
    double *lnum = nullptr;        // [nblk][2] particle number, derivative
    double *loff = nullptr;        // [nblk][2] off-diagonal square sums of T and C
    int *lcold = nullptr;          // [nblk] evaluation started in the original basis
    double *lnst = nullptr;        // [2][kHlNewton] Newton records
    int *lactive = nullptr;        // [nblk] block still evaluated
    int *lnb = nullptr;            // [2][nblk] 2h, h
    int *lhb = nullptr;            // [nblk] h (apply block sizes)
    int *llist = nullptr;          // [nblk][hmax] the block's states (apply lists)
    double *lTap = nullptr;        // [nblk][hmax][hmax] apply matrices
    double *lgpart = nullptr;      // pair-Gram partial sums; pairing-field partials before
};

}  // namespace detail
}  // namespace mocca

namespace mocca {
namespace detail {
// A captured GradDesc-to-energy iteration of mocca_batch_run for one
// configuration of the rotating state buffers (pre), which it leaves in post.
struct IterGraph {
    void *pre[4], *post[4];
    hipGraphExec_t exec;
};
}  // namespace detail
}  // namespace mocca

struct mocca_solve {
    mocca_mesh mesh{};
    int M = 0, N = 0, prec = MOCCA_FP64, tsize = 8;
    hipStream_t st = nullptr;
    std::string err;
    mocca::detail::Buffers b;
    mocca::detail::Scf scf;
    mocca::detail::Est est;
    mocca::detail::Hfb hfb;
    mocca::MeshK<double> md{};
    mocca::MeshK<float> mf{};
    bool meta_set = false, fields_set = false;
    bool w_antisym = false;   // W[nu][mu] = -W[mu][nu], zero diagonal (no tensor terms)
    bool prev_valid = false;   // b.upd holds the states before the last GradDesc step
    // mocca_batch_run's captured iterations, valid while no buffer is
    // (re)allocated and no setting changes (graph_epoch counts both) and for
    // the dt, momentum they were captured with
    unsigned long graph_epoch = 0, graphs_epoch = 0;
    double graphs_dth = 0, graphs_mu = 0;
    std::vector<mocca::detail::IterGraph> graphs;
    double *graph_rows = nullptr;   // [2][row]: previous, current
    double kin[2] = {0, 0};
    std::vector<int> qn;
    std::vector<double> occ;
    int nchunks = 0, ntiles = 0;
    int nblk = 0, max_m = 0, nsplit = 1;
    int njobs = 0;
    // z-marching FD kernels: y-tile rows, tiles per plane, task count
    bool march = false;
    // mixed precision: Lagrange derivatives on the FP32 matrix cores with
    // split matrices (mocca_set_lagrange_precision) instead of the FP64 ones
    bool lag_split = false;
    int TY = 0, ntile_y = 0, ntasks = 0;
    // the pipelined FP64 step kernel: usable, z slabs per state and tile, tasks
    bool dma_hpsi = false;
    int nzs = 1, ntasks_dma = 0;
    int ndchunks = 0, ndtasks = 0, gtile = 64, mmax = 0;
    size_t chol_lds = 0;   // dynamic LDS bytes available to the in-LDS Cholesky
    bool mfma_gram = false;  // Gram on the matrix cores (blocks <= kMfmaMaxRows)
    bool mfma_apply = false; // apply on the matrix cores (blocks <= kMfmaApplyMaxRows)
    int mpad = 0, nsplit_mfma = 1;
    // profiling
    bool prof = false;
    std::vector<mocca::detail::Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, long>> acc;
};

namespace mocca {
namespace detail {

int fail(mocca_solve *h, int code, const char *fmt, ...);

// Launch plan of the Cholesky-QR Gram-Schmidt for nblk blocks of at most
// mmax states over K4 = 4N values (M states of tsize bytes in all): tile
// sizes, K splits and which kernels run.  The partial Gram matrices need
// gpart_doubles() doubles.
struct GsPlan {
    int nblk = 0, mmax = 0, max_m = 0, gtile = 32, ntile_max = 0, nsplit = 1;
    int mpad = 0, ntile16 = 0, nsplit_mfma = 1;
    bool mfma_gram = false, mfma_apply = false;
    size_t gpart_doubles() const {
        const size_t a = (size_t)nblk * ntile_max * nsplit * gtile * gtile;
        const size_t b = (size_t)nblk * nsplit_mfma * ntile16 * 256;
        return a > b ? a : b;
    }
};
GsPlan gs_plan(int nblk, int mmax, long K4, int M, size_t tsize);

#define HIPCHK(h, call)                                                           \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::mocca::detail::fail(h, MOCCA_ERR_HIP, "%s failed: %s", #call, \
                                         hipGetErrorString(e_));                  \
    } while (0)

template <typename P>
int dalloc(mocca_solve *h, P **p, size_t bytes) {
    ++h->graph_epoch;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 8);
    if (e != hipSuccess)
        return fail(h, MOCCA_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes,
                    hipGetErrorString(e));
    *p = static_cast<P *>(q);
    return 0;
}

hipEvent_t take_event(mocca_solve *h);

// Launch f on the handle's stream, timed with events when profiling is on.
template <typename F>
int run(mocca_solve *h, const char *name, F &&f) {
    if (h->prof) {
        hipEvent_t a = take_event(h), b = take_event(h);
        (void)hipEventRecord(a, h->st);
        f();
        (void)hipEventRecord(b, h->st);
        h->pending.push_back({name, a, b});
    } else {
        f();
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(h, MOCCA_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e));
    if (const char *k = take_launch_failure())
        return fail(h, MOCCA_ERR_HIP, "launch of %s: the dynamic LDS limit of %s could not be raised",
                    name, k);
    return 0;
}

int check_ready(mocca_solve *h);
int sign_first_h(int axis, int P, int S, int TS, int c, int lag);
// GradDesc on the device (hPsi + step + energies + Gram-Schmidt); also
// leaves K[i] = dv <psi_i|Lap psi_i> of the states before the step.
int grad_desc_any(mocca_solve *h, double dth, double mu);
// Densities of the current states into dens/derived.
int density_any(mocca_solve *h);
// K[i] = dv <psi_i|Lap psi_i> of the current states (no other output).
int kinetic_any(mocca_solve *h);
// HFB pairing: setup from the C-ABI parameters, one SolvePairing step, the
// pairing energy and Fermi energies into an energy row, the status check.
// BCS pairing (PairingType 1): setup, one SolvePairing step, the energy
// row's pairing part.  pairing_energy_impl: the pairing part of a row for
// either (nothing for HF).
int bcs_solve(mocca_solve *h);
int bcs_energy_impl(mocca_solve *h, double *row);
int pairing_energy_impl(mocca_solve *h, double *row);
int hfb_setup(mocca_solve *h, const mocca_pairing *q);
int hfb_free(mocca_solve *h);
int hfb_solve_any(mocca_solve *h, bool first);
int hfb_energy_impl(mocca_solve *h, double *row);
int hfb_check_status(mocca_solve *h);
// The quasiparticle energies of the last pairing step where that step did
// not form them (the sign-function path): one diagonalisation at its Fermi
// energies.
int hfb_ensure_qpe(mocca_solve *h);
// Kernels of mocca_hfb.hip launched by the large-block path: the rho/kappa
// history step of SolvePairing (first = 1) or the HFBMix mixing (first = 0),
// and the cutoff factors with the density factor.
void hfb_launch_mix(mocca_solve *h, int first);
void hfb_launch_cutoffs(mocca_solve *h);
// The large-block path (mocca_hfb_large.hip): whether hmax states per block
// fit it, its buffers, one SolvePairing step.
bool hfb_large_fits(int hmax);
int hfb_large_setup(mocca_solve *h);
int hfb_large_free(mocca_solve *h);
int hfb_large_solve_any(mocca_solve *h);
constexpr int kRowLen = 47;   // energy row: 43 of CompEnergy + pairing energy(2) + Fermi(2)
// AlternateStep (constraints of type 2) on the device; alternate_prepare
// allocates the buffer of the states before the step ahead of time.
int alternate_any(mocca_solve *h);
int alternate_prepare(mocca_solve *h);
// IterativeEstimation: out = dt, momentum, maxE, relE, con, iterations.
int estimate_any(mocca_solve *h, int iteration, double dt, double mu, double *out);
constexpr double kHbar = 6.58211928;   // Force.f90:103
// Destroys mocca_batch_run's captured iterations of a solve.
void graphs_clear(mocca_solve *h);
// Releases the buffers of the self-consistent part.
void scf_free(mocca_solve *h);
// Switches the derivative scheme of the handle between the EV8 finite
// differences and the Lagrange mesh (AssignFDCoefs, Derivatives.f90:206-346):
// the mesh descriptors, the kernel paths and the derived-density jobs.
int set_scheme(mocca_solve *h, int lagrange);
// Densities of the density basis without mixing (UpdateDensities(0)).
int density_any(mocca_solve *h);
// dv <psi|Lap psi> of every state of the density basis.
int kinetic_any(mocca_solve *h);

}  // namespace detail
}  // namespace mocca
