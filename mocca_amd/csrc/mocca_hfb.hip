// HFB pairing on the device (PairingType 2 of SolvePairing, Pairing.f90:743-890)
// for runs with time reversal and conserved signature: the states and their
// time-reversed partners form one (parity, isospin) block each
// (PrepareHFBModule, HFB.f90:248-295).  Everything is FP64.
//
//   pairing field  Delta(r) from kappa over the block pairs (HFBPairingField,
//                  HFB.f90:350-518), gaps Delta_ij (HFBGaps :520-632):
//                  point kernels over the states;
//   Fermi energies Broyden iteration (HFBFindFermiEnergyBroyden :1849-2161)
//                  over HFBNumberofParticles (:810-1073): one workgroup per
//                  isospin runs the whole iteration; every evaluation
//                  diagonalises the signature-reduced HFB matrix of each
//                  parity block (DiagonaliseHFBHamiltonian_Signature
//                  :2695-2801) by parallel cyclic Jacobi in LDS, warm-started
//                  from the previous evaluation's eigenvectors;
//   canonical basis eigenvectors of rho (DiagonaliseRhoHFB :3156-3285,
//                  HFBOccupations :3361-3561), the canonical states as
//                  combinations of the HF states, rho/kappa mixing.
//
// Block layout: the first h entries of a block are its states in state
// order, the next h their time-reversed partners (HFB size 2 nwt).  With all
// signatures +1 the reduced HFB matrix of a block is
//     T = [[E - lambda, D], [D^T, -(E - lambda)]]   (2h x 2h, real),
// E the states' single-particle energies and D_ab = Delta(a, h+b).  Its
// eigenvectors x_k of the h largest eigenvalues give
//     rho(a, b) = rho(h+a, h+b) = sum_k x_k(h+a) x_k(h+b),
//     kappa(h+a, b) = -kappa(b, h+a) = sum_k x_k(h+a) x_k(b),
// which is what constructRhoHFB / ConstructKappaHFB (:3128-3360) build from
// the U, V columns the reference selects.
#include "mocca_internal.h"

#include <algorithm>
#include <cmath>

namespace mocca {

namespace {

constexpr int kHfbThreads = 256;

__device__ __forceinline__ double hfb_wave_sum(double a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

// Sum over the workgroup; every thread gets the result.
__device__ double hfb_block_sum(double a, double *sh) {
    a = hfb_wave_sum(a);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64, nw = blockDim.x / 64;
    __syncthreads();
    if (l == 0) sh[w] = a;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < nw; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// Pair amplitude of GetPairDensity (PairingInteraction.f90:101-180).
__device__ __forceinline__ void pair_amp(const double a[4], const double b[4], bool tr,
                                         double &re, double &im) {
    if (!tr) {
        re = a[0] * b[2] - a[1] * b[3] - a[2] * b[0] + a[3] * b[1];
        im = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] - a[3] * b[0];
    } else {
        re = -a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
        im = a[0] * b[1] - a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Cutoff factors (SymmetricFermi, PairingInteraction.f90:198-210) and the
// density factor -G (1 - alpha rho0 / rho_sat) (CompDensityFactor :265-354).
__global__ void k_hfb_cutoffs(HfbK k, const double *__restrict__ E, const int *__restrict__ qn,
                              const double *__restrict__ fermi, int M, double *__restrict__ pc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int it = (qn[4 * i] + 3) / 2 - 1;
    const double up = (E[i] - fermi[it] - k.cut[it]) / k.mu[it];
    const double down = -(E[i] - fermi[it] + k.cut[it]) / k.mu[it];
    double c = sqrt(sqrt(1.0 / (1.0 + exp(up))));
    c = c * sqrt(sqrt(1.0 / (1.0 + exp(down))));
    pc[i] = c;
}

__global__ void k_hfb_dfac(HfbK k, const double *__restrict__ dens, long N,
                           double *__restrict__ dfac) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const double rho0 = dens[p] + dens[9 * N + p];
    for (int it = 0; it < 2; ++it)
        dfac[it * N + p] = -k.G[it] * (1.0 - k.alpha[it] / k.rhosat * rho0);
}

// Pairing field of each isospin: -dfac sum_{a', b} c_b c_a' kappa(h+a', b)
// P(psi_b, psi_a', TR), written as -dfac sum_a' P(phi_a', psi_a') with
// phi_a' = sum_b w_a'b psi_b (P is linear in its first argument).  Pairs
// below the cut are dropped as in the reference.
template <typename T>
__global__ __launch_bounds__(kHfbThreads) void k_hfb_field(HfbK k, const T *__restrict__ psi,
                                                          long N,
                                                          const double *__restrict__ kappa,
                                                          const double *__restrict__ pc,
                                                          const double *__restrict__ dfac,
                                                          double *__restrict__ field) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    if (p >= N) return;
    double fr = 0, fi = 0;
    for (int b = 0; b < k.nblk; ++b) {
        const HfbBlockK &bk = k.blk[b];
        if (bk.it != it) continue;
        const int h = bk.h, n = 2 * h;
        const double *kap = kappa + bk.off;
        for (int a2 = 0; a2 < h; ++a2) {
            const int s2 = k.state[bk.first + a2];
            const double c2 = pc[s2];
            double phi[4] = {0, 0, 0, 0};
            bool any = false;
            for (int b1 = 0; b1 < h; ++b1) {
                const int s1 = k.state[bk.first + b1];
                const double c1 = pc[s1];
                if (c1 < kHfbNumCut) continue;
                const double kv = kap[(h + a2) * n + b1];
                if (c1 * c2 * fabs(kv) < kHfbNumCut) continue;
                const double w = c1 * c2 * kv;
                const T *q = psi + (long)s1 * 4 * N + p;
#pragma unroll
                for (int c = 0; c < 4; ++c) phi[c] += w * (double)q[c * N];
                any = true;
            }
            if (!any) continue;
            double v[4];
            const T *q = psi + (long)s2 * 4 * N + p;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = (double)q[c * N];
            double re, im;
            pair_amp(phi, v, true, re, im);
            fr = fr - re;
            fi = fi - im;
        }
    }
    const double d = dfac[it * N + p];
    field[(2 * it) * N + p] = fr * d;
    field[(2 * it + 1) * N + p] = fi * d;
}

// Gaps Delta(a, h+a') = dv c_a c_a' sum_r (Re P Re F + Im P Im F), P the
// pair amplitude of (psi_a, psi_a') with the time-reversal form, and
// Delta(h+a', a) = -Delta(a, h+a'): one workgroup per (block, a, a').
template <typename T>
__global__ __launch_bounds__(kHfbThreads) void k_hfb_gaps(HfbK k, const T *__restrict__ psi,
                                                         long N, const double *__restrict__ pc,
                                                         const double *__restrict__ field,
                                                         double dv, double *__restrict__ delta) {
    __shared__ double sh[kHfbThreads / 64];
    const int b = blockIdx.y;
    if (b >= k.nblk) return;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int a = blockIdx.x / h, a2 = blockIdx.x % h;
    if (a >= h) return;
    const int s1 = k.state[bk.first + a], s2 = k.state[bk.first + a2];
    const double cc = pc[s1] * pc[s2];
    double *dl = delta + bk.off;
    if (cc < kHfbNumCut) {
        if (threadIdx.x == 0) {
            dl[a * n + h + a2] = 0.0;
            dl[(h + a2) * n + a] = 0.0;
        }
        return;
    }
    const T *q1 = psi + (long)s1 * 4 * N, *q2 = psi + (long)s2 * 4 * N;
    const double *fr = field + (2 * bk.it) * N, *fi = fr + N;
    double sr = 0, si = 0;
    for (long p = threadIdx.x; p < N; p += blockDim.x) {
        double x[4], y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = (double)q1[c * N + p];
            y[c] = (double)q2[c * N + p];
        }
        double re, im;
        pair_amp(x, y, true, re, im);
        sr += re * fr[p];
        si += im * fi[p];
    }
    sr = hfb_block_sum(sr, sh);
    si = hfb_block_sum(si, sh);
    if (threadIdx.x == 0) {
        const double d = dv * pc[s1] * pc[s2] * (sr + si);
        dl[a * n + h + a2] = d;
        dl[(h + a2) * n + a] = -d;
    }
}

// ---------------------------------------------------------------------------
// Parallel cyclic Jacobi on a symmetric n x n matrix A (row-major, n <= kHfbMaxN),
// eigenvectors accumulated into the columns of V (which holds the starting
// basis: A must already be expressed in it).  Rotations of one round act on
// disjoint index pairs (round-robin schedule), so a round is two parallel
// passes (rows, then columns) separated by barriers.
__device__ void jacobi_sym(double *A, double *V, int n, double *rc, double *rs, int *rp,
                           int *rq, double *sh) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int m = n + (n & 1);          // even number of players
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0, dia = 0;
        for (int e = tid; e < n * n; e += nt) {
            const int i = e / n, j = e % n;
            const double v = A[e] * A[e];
            if (i == j) dia += v; else off += v;
        }
        off = hfb_block_sum(off, sh);
        dia = hfb_block_sum(dia, sh);
        if (off <= 1e-32 * dia || off == 0.0) break;
        for (int r = 0; r < m - 1; ++r) {
            // pairs of this round and their rotations (Golub & Van Loan 8.4.2)
            for (int pp = tid; pp < m / 2; pp += nt) {
                auto pos = [&](int x) { return x == 0 ? 0 : ((x - 1 + r) % (m - 1)) + 1; };
                int i = pos(pp), j = pos(m - 1 - pp);
                if (i > j) { const int t = i; i = j; j = t; }
                double c = 1.0, s = 0.0;
                if (j < n) {
                    const double apq = A[i * n + j];
                    if (apq != 0.0) {
                        const double tau = (A[j * n + j] - A[i * n + i]) / (2.0 * apq);
                        const double t = tau >= 0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                                  : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                rp[pp] = i;
                rq[pp] = j;
                rc[pp] = c;
                rs[pp] = s;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = tid; e < (m / 2) * n; e += nt) {
                const int pp = e / n, col = e % n;
                const int i = rp[pp], j = rq[pp];
                if (j >= n || rs[pp] == 0.0) continue;
                const double c = rc[pp], s = rs[pp];
                const double ai = A[i * n + col], aj = A[j * n + col];
                A[i * n + col] = c * ai - s * aj;
                A[j * n + col] = s * ai + c * aj;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J
            for (int e = tid; e < (m / 2) * n; e += nt) {
                const int pp = e / n, row = e % n;
                const int i = rp[pp], j = rq[pp];
                if (j >= n || rs[pp] == 0.0) continue;
                const double c = rc[pp], s = rs[pp];
                const double ai = A[row * n + i], aj = A[row * n + j];
                A[row * n + i] = c * ai - s * aj;
                A[row * n + j] = s * ai + c * aj;
                const double vi = V[row * n + i], vj = V[row * n + j];
                V[row * n + i] = c * vi - s * vj;
                V[row * n + j] = s * vi + c * vj;
            }
            __syncthreads();
        }
    }
}

// One evaluation of HFBNumberofParticles for one block at lambda: reduced
// HFB matrix, eigen-decomposition (warm start from V), the h largest
// eigenvalues' vectors -> rho, kappa of the block; returns 2 tr(rho_h).
__device__ double hfb_block_eval(const HfbK &k, const HfbBlockK &bk, const double *E,
                                 const double *delta, double lam, double *A, double *V,
                                 double *W, double *rho, double *kappa, double *qpe, double *rc,
                                 double *rs, int *rp, int *rq, int *rank, double *sh) {
    const int h = bk.h, n = 2 * h, tid = threadIdx.x, nt = blockDim.x;
    const double *dl = delta + bk.off;
    // W = T(lambda)
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e % n;
        double v = 0.0;
        if (i < h && j < h) v = i == j ? E[k.state[bk.first + i]] - lam : 0.0;
        else if (i >= h && j >= h) v = i == j ? -(E[k.state[bk.first + i - h]] - lam) : 0.0;
        else if (i < h) v = dl[i * n + h + (j - h)];
        else v = dl[j * n + h + (i - h)];
        W[e] = v;
    }
    __syncthreads();
    // A = V^T W V (the warm start's basis) in two products
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e % n;
        double acc = 0;
        for (int r = 0; r < n; ++r) acc += W[i * n + r] * V[r * n + j];
        A[e] = acc;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e % n;
        double acc = 0;
        for (int q = 0; q < n; ++q) acc += V[q * n + i] * A[q * n + j];
        W[e] = acc;
    }
    __syncthreads();
    // symmetrise against the rounding of the products
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e % n;
        A[e] = 0.5 * (W[i * n + j] + W[j * n + i]);
    }
    __syncthreads();
    jacobi_sym(A, V, n, rc, rs, rp, rq, sh);
    // ranks of the eigenvalues (descending; ties by index)
    for (int i = tid; i < n; i += nt) {
        const double wi = A[i * n + i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double wj = A[j * n + j];
            if (wj > wi || (wj == wi && j < i)) ++r;
        }
        rank[i] = r;
    }
    __syncthreads();
    double tr = 0;
    double *rb = rho + bk.off, *kb = kappa + bk.off;
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e % n;
        double v = 0.0;
        if (i < h && j < h) {
            for (int q = 0; q < n; ++q)
                if (rank[q] < h) v += V[(h + i) * n + q] * V[(h + j) * n + q];
            if (i == j) tr += v;
        } else if (i >= h && j >= h) {
            for (int q = 0; q < n; ++q)
                if (rank[q] < h) v += V[i * n + q] * V[j * n + q];
        }
        rb[e] = v;
        double kv = 0.0;
        if (i >= h && j < h) {
            for (int q = 0; q < n; ++q)
                if (rank[q] < h) kv += V[i * n + q] * V[j * n + q];
        } else if (i < h && j >= h) {
            for (int q = 0; q < n; ++q)
                if (rank[q] < h) kv -= V[j * n + q] * V[i * n + q];
        }
        kb[e] = kv;
    }
    for (int i = tid; i < n; i += nt)
        if (rank[i] < h) qpe[bk.first + rank[i]] = A[i * n + i];
    tr = hfb_block_sum(tr, sh);
    return 2.0 * tr;
}

// The Broyden Fermi solver of one isospin (blockIdx.x), HFBIter iterations
// at most, stop at |N - Z| < prec.  Every evaluation recomputes rho and
// kappa of the isospin's blocks; the last one is kept, as in the reference.
__global__ __launch_bounds__(kHfbThreads) void k_hfb_broyden(
    HfbK k, const double *__restrict__ E, const double *__restrict__ delta,
    double *__restrict__ fermi, double *__restrict__ rho, double *__restrict__ kappa,
    double *__restrict__ qpe, double *__restrict__ work, int *__restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double sh[kHfbThreads / 64];
    __shared__ double rc[kHfbMaxN / 2], rs[kHfbMaxN / 2];
    __shared__ int rp[kHfbMaxN / 2], rq[kHfbMaxN / 2], rank[kHfbMaxN];
    const int it = blockIdx.x, tid = threadIdx.x;
    // the blocks of this isospin, matrices in LDS: A, V, W per block
    int bl[2], nb = 0;
    for (int b = 0; b < k.nblk; ++b)
        if (k.blk[b].it == it) bl[nb++] = b;
    double *A[2], *Vm[2], *Wm[2];
    {
        double *base = lds;
        for (int q = 0; q < nb; ++q) {
            const int n = 2 * k.blk[bl[q]].h;
            A[q] = base; base += n * n;
            Vm[q] = base; base += n * n;
            Wm[q] = base; base += n * n;
            for (int e = tid; e < n * n; e += blockDim.x) Vm[q][e] = (e / n == e % n) ? 1.0 : 0.0;
        }
    }
    (void)work;
    __syncthreads();
    const double particles = k.particles[it];
    auto eval = [&](double lam) {
        double nn = 0;
        for (int q = 0; q < nb; ++q)
            nn += hfb_block_eval(k, k.blk[bl[q]], E, delta, lam, A[q], Vm[q], Wm[q], rho, kappa,
                                 qpe, rc, rs, rp, rq, rank, sh);
        return nn - particles;
    };
    const double prec = k.prec;
    double F = fermi[it];
    double FH = F + 0.01;
    double Nv = eval(F);
    bool conv = fabs(Nv) < prec;
    const double NX = eval(FH);
    double J = (Nv - NX) / (F - FH);
    double upd = 0, norm = 0;
    int iter;
    for (iter = 1; iter <= k.iters; ++iter) {
        const double det = J;
        const double inv = fabs(det) < 0.001 ? copysign(0.99, det) : 1.0 / det;
        if (!conv) {
            upd = -inv * Nv;
            norm = upd * upd;
        }
        upd = 0.5 * upd;
        if (!conv) {
            FH = F;
            F = F + upd;
        }
        Nv = eval(F);
        if (fabs(Nv) < prec) conv = true;
        if (conv) break;      // the other isospin's evaluations leave this one unchanged
        if (iter > 1) J = J + Nv * upd / norm;
        if (Nv == Nv + 1.0 || Nv != Nv) {
            if (tid == 0) atomicOr(status, 2);
            break;
        }
    }
    if (tid == 0) {
        fermi[it] = F;
        if (!conv) atomicOr(status, 1);
    }
}

// Canonical basis of one block (blockIdx.x): eigenvectors of rho_h (the
// signature +1 part of rho, DiagonaliseRhoHFB with FindCorrectColumns),
// entries below 1e-11 set to zero, occupations 2 x eigenvalue clamped to
// [0, 2]; tcan[block] (h x h, column c = canonical state c in ascending
// occupation), canonical occupations into the slots of the block's states;
// the HF states' occupations 2 rho_ii (HFBOccupations :3470-3480).
__global__ __launch_bounds__(kHfbThreads) void k_hfb_canonical(HfbK k,
                                                              const double *__restrict__ rho,
                                                              double *__restrict__ tcan,
                                                              double *__restrict__ can_occ,
                                                              double *__restrict__ hf_occ,
                                                              int *__restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double sh[kHfbThreads / 64];
    __shared__ double rc[kHfbMaxN / 2], rs[kHfbMaxN / 2];
    __shared__ int rp[kHfbMaxN / 2], rq[kHfbMaxN / 2], rank[kHfbMaxN];
    const HfbBlockK &bk = k.blk[blockIdx.x];
    const int h = bk.h, n = 2 * h, tid = threadIdx.x;
    double *A = lds, *V = lds + h * h;
    const double *rb = rho + bk.off;
    for (int e = tid; e < h * h; e += blockDim.x) {
        const int i = e / h, j = e % h;
        A[e] = rb[i * n + j];
        V[e] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    jacobi_sym(A, V, h, rc, rs, rp, rq, sh);
    for (int i = tid; i < h; i += blockDim.x) {
        const double wi = A[i * h + i];
        int r = 0;
        for (int j = 0; j < h; ++j) {
            const double wj = A[j * h + j];
            if (wj < wi || (wj == wi && j < i)) ++r;
        }
        rank[i] = r;
    }
    __syncthreads();
    double *tb = tcan + bk.toff;
    for (int e = tid; e < h * h; e += blockDim.x) {
        const int j = e / h, c = e % h;            // row j (HF state), column = source c
        const double v = V[j * h + c];
        tb[j * h + rank[c]] = fabs(v) < 1e-11 ? 0.0 : v;
    }
    for (int c = tid; c < h; c += blockDim.x) {
        double w = A[c * h + c];
        if (w - 1.0 > 1e-5) atomicOr(status, 4);
        if (w < 0.0) w = 0.0;
        if (w > 1.0) w = 1.0;
        can_occ[k.state[bk.first + rank[c]]] = 2.0 * w;
        hf_occ[k.state[bk.first + c]] = 2.0 * rb[c * n + c];
    }
}

// Canonical states: can[slot c] = sum_j tcan(j, c) psi_j within each block.
template <typename T>
__global__ __launch_bounds__(kHfbThreads) void k_hfb_can_states(HfbK k, const T *__restrict__ psi,
                                                               long N,
                                                               const double *__restrict__ tcan,
                                                               T *__restrict__ can) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (p >= N) return;
    const int b = k.slot_block[slot], c = k.slot_index[slot];
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h;
    const double *tb = tcan + bk.toff;
    double acc[4] = {0, 0, 0, 0};
    for (int j = 0; j < h; ++j) {
        const double t = tb[j * h + c];
        if (t == 0.0) continue;
        const T *q = psi + (long)k.state[bk.first + j] * 4 * N + p;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc[cc] += t * (double)q[cc * N];
    }
    T *o = can + (long)k.state[bk.first + c] * 4 * N + p;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) o[cc * N] = T(acc[cc]);
}

// rho <- mix rho + (1 - mix) rho_old when rho_old is not all zero, kappa
// likewise when kappa is not all zero (HFBOccupations :3492-3510); then the
// old matrices take the mixed ones.  The reference's new kappa is exactly
// zero only when every gap is (its Householder-QL eigenvectors of a matrix
// with any non-zero coupling carry rounding-level pair components), so the
// test is made on the gaps as well: Jacobi leaves couplings below its
// stopping threshold untouched.  first = 1 instead sets rho_old, kappa_old
// to the current ones while rho_old is all zero (SolvePairing :800-806).
// HFBMix (SolvePairing :800-806, :820-840) over every element of the packed
// blocks, in three launches (the flags are reduced over the whole grid):
// zero the flags; flags[0] = any old rho != 0, flags[1] = any kappa or
// delta != 0; then with first = 1 the old matrices take the current ones
// while the old rho is all zero, and with first = 0 rho and kappa are mixed
// with the old ones (each where its flag is set) and become the old ones.
__global__ void k_hfb_mix_zero(int *__restrict__ flags) {
    if (threadIdx.x < 2) flags[threadIdx.x] = 0;
}

__global__ void k_hfb_mix_flags(int total, const double *__restrict__ kappa,
                                const double *__restrict__ orho,
                                const double *__restrict__ delta, int *__restrict__ flags) {
    int f0 = 0, f1 = 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        f0 |= orho[e] != 0.0;
        f1 |= kappa[e] != 0.0 || delta[e] != 0.0;
    }
    f0 = __syncthreads_or(f0);
    f1 = __syncthreads_or(f1);
    if (threadIdx.x == 0) {
        if (f0) atomicOr(flags, 1);
        if (f1) atomicOr(flags + 1, 1);
    }
}

__global__ void k_hfb_mix(int total, double mix, int first, double *__restrict__ rho,
                          double *__restrict__ kappa, double *__restrict__ orho,
                          double *__restrict__ okappa, const int *__restrict__ flags) {
    const bool f0 = flags[0] != 0, f1 = flags[1] != 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        if (first) {
            if (!f0) {
                orho[e] = rho[e];
                okappa[e] = kappa[e];
            }
            continue;
        }
        if (f0) rho[e] = mix * rho[e] + (1.0 - mix) * orho[e];
        if (f1) kappa[e] = mix * kappa[e] + (1.0 - mix) * okappa[e];
        orho[e] = rho[e];
        okappa[e] = kappa[e];
    }
}

// Pairing energy E_q = 1/2 sum_blocks sum_ij Delta_ij kappa_ij (HFBEnergy
// :296-348): partial sums of kHfbEnergyParts workgroups per block, then one
// workgroup adds them in a fixed order into the energy row: row[43..44] =
// E_pair, row[45..46] = Fermi.
__global__ void k_hfb_energy_part(HfbK k, const double *__restrict__ delta,
                                  const double *__restrict__ kappa, double *__restrict__ part) {
    __shared__ double sh[kHfbThreads / 64];
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int n = 2 * bk.h;
    double s = 0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n * n; q += gridDim.x * blockDim.x)
        s += delta[bk.off + q] * kappa[bk.off + q];
    s = hfb_block_sum(s, sh);
    if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = s;
}

__global__ void k_hfb_energy(HfbK k, const double *__restrict__ part, int nparts,
                             const double *__restrict__ fermi, double *__restrict__ row) {
    if (threadIdx.x != 0) return;
    double e[2] = {0, 0};
    for (int b = 0; b < k.nblk; ++b) {
        double s = 0;
        for (int x = 0; x < nparts; ++x) s += part[b * nparts + x];
        e[k.blk[b].it] += s;
    }
    row[43] = -0.5 * (-e[0]);
    row[44] = -0.5 * (-e[1]);
    row[45] = fermi[0];
    row[46] = fermi[1];
}

// ---------------------------------------------------------------------------
// Guess of kappa (GuessHFBMatrices, HFB.f90:3645-3836, PairingType 2): every
// state paired with its partner, kappa(h+a, a) = 0.25, kappa(a, h+a) = -0.25.
void hfb_guess(const HfbK &k, std::vector<double> &kappa) {
    for (int b = 0; b < k.nblk; ++b) {
        const HfbBlockK &bk = k.blk[b];
        const int h = bk.h, n = 2 * h;
        for (int a = 0; a < h; ++a) {
            kappa[bk.off + (h + a) * n + a] = 0.25;
            kappa[bk.off + a * n + h + a] = -0.25;
        }
    }
}

}  // namespace mocca

namespace mocca {

// ---------------------------------------------------------------------------
// BCS pairing (PairingType 1: BCS.f90, the loop of SolvePairing,
// Pairing.f90:826-858).  Per state i of isospin it: cutoff factor c_i,
// gap Delta_i, quasiparticle energy E_i = sqrt((e_i - lambda)^2 + Delta_i^2)
// (BCSQPEnergies :151-188), occupation 1 - (e_i - lambda) / E_i clamped at
// 0 (BCSOccupations :190-226); everything real (the pair density of a state
// with its own time reverse is real: -rho_i).

// Pairing field F_it(r) = dfac_it(r) sum_{i in it} (1/2) c_i^2 Delta_i / E_i
// (-rho_i(r)), summed state by state as BCSPairingField (:65-101) does.
template <typename T>
__global__ __launch_bounds__(256) void k_bcs_field(const T *__restrict__ psi, long N, int M,
                                                   const int *__restrict__ qn,
                                                   const double *__restrict__ pc,
                                                   const double *__restrict__ delta,
                                                   const double *__restrict__ eqp,
                                                   const double *__restrict__ dfac,
                                                   double *__restrict__ field) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    if (p >= N) return;
    double f = 0.0;
    for (int i = 0; i < M; ++i) {
        if ((qn[4 * i] + 3) / 2 - 1 != it) continue;
        const double w = 0.5 * pc[i] * pc[i] * delta[i] / eqp[i];
        const T *ps = psi + (long)i * 4 * N + p;
        double r = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double v = (double)ps[c * N];
            r += v * v;
        }
        f += w * (-r);
    }
    field[it * N + p] = dfac[it * N + p] * f;
}

// Gaps Delta_i = c_i^2 dv sum_r rho_i(r) F_it(r) (BCSGaps :16-63), one
// workgroup per state.
template <typename T>
__global__ __launch_bounds__(kHfbThreads) void k_bcs_gaps(const T *__restrict__ psi, long N,
                                                         const int *__restrict__ qn,
                                                         const double *__restrict__ pc,
                                                         const double *__restrict__ field,
                                                         double dv, double *__restrict__ delta) {
    __shared__ double sh[kHfbThreads / 64];
    const int i = blockIdx.x;
    const int it = (qn[4 * i] + 3) / 2 - 1;
    const T *ps = psi + (long)i * 4 * N;
    const double *F = field + it * N;
    double s = 0.0;
    for (long p = threadIdx.x; p < N; p += blockDim.x) {
        double r = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double v = (double)ps[c * N + p];
            r += v * v;
        }
        s += r * F[p];
    }
    s = hfb_block_sum(s, sh);
    if (threadIdx.x == 0) delta[i] = pc[i] * pc[i] * dv * s;
}

// BCSFindFermiEnergy (:103-149) for isospin threadIdx.x: the quasiparticle
// energies at the current lambda, then lambda = (N - sum (1 - e/E)) /
// sum 1/E; rec[it] keeps the old lambda.  With first set, gaps all below
// 1e-5 are first set to 1 (SolvePairing :812) -- that happens in
// k_bcs_start instead.
__global__ void k_bcs_fermi(int M, const int *__restrict__ qn, const double *__restrict__ E,
                            const double *__restrict__ delta, double n0, double n1,
                            double *__restrict__ fermi, double *__restrict__ eqp,
                            double *__restrict__ rec) {
    const int it = threadIdx.x;
    if (it >= 2) return;
    const double lam = fermi[it];
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < M; ++i) {
        if ((qn[4 * i] + 3) / 2 - 1 != it) continue;
        const double d = E[i] - lam;
        const double e = sqrt(d * d + delta[i] * delta[i]);
        eqp[i] = e;
        s1 += 1.0 - E[i] / e;
        s2 += 1.0 / e;
    }
    rec[it] = lam;
    fermi[it] = ((it == 0 ? n0 : n1) - s1) / s2;
}

// BCSOccupations (:190-226) at the new lambda, and the convergence test of
// the loop (|lambda - old| < FermiPrec for both isospins, :854) into rec[2].
__global__ void k_bcs_occ(int M, const int *__restrict__ qn, const double *__restrict__ E,
                          const double *__restrict__ delta, const double *__restrict__ fermi,
                          double *__restrict__ eqp, double *__restrict__ occ, double prec,
                          double *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        rec[2] = (fabs(fermi[0] - rec[0]) < prec && fabs(fermi[1] - rec[1]) < prec) ? 1.0 : 0.0;
    if (i >= M) return;
    const int it = (qn[4 * i] + 3) / 2 - 1;
    const double d = E[i] - fermi[it];
    const double e = sqrt(d * d + delta[i] * delta[i]);
    eqp[i] = e;
    const double o = 1.0 - d / e;
    occ[i] = o < 0.0 ? 0.0 : o;
}

// Gaps below 1e-5 in every state are set to 1 (SolvePairing :812).
__global__ void k_bcs_start(int M, double *__restrict__ delta) {
    __shared__ int big;
    if (threadIdx.x == 0) big = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < M; i += blockDim.x)
        if (fabs(delta[i]) >= 1e-5) big = 1;
    __syncthreads();
    if (!big)
        for (int i = threadIdx.x; i < M; i += blockDim.x) delta[i] = 1.0;
}

// BCSEnergy (:228-247): E_pair(it) = -1/2 sum |Delta_i|^2 / E_i, and the
// Fermi energies, into the row (elements 43..46).
__global__ void k_bcs_energy(int M, const int *__restrict__ qn, const double *__restrict__ delta,
                             const double *__restrict__ eqp, const double *__restrict__ fermi,
                             double *__restrict__ row) {
    if (threadIdx.x != 0) return;
    double e[2] = {0.0, 0.0};
    for (int i = 0; i < M; ++i) {
        const int it = (qn[4 * i] + 3) / 2 - 1;
        e[it] += delta[i] * delta[i] / eqp[i];
    }
    row[43] = -0.5 * e[0];
    row[44] = -0.5 * e[1];
    row[45] = fermi[0];
    row[46] = fermi[1];
}

}  // namespace mocca

// ===========================================================================
// Host side
namespace mocca {
namespace detail {

int hfb_free(mocca_solve *h) {
    Hfb &p = h->hfb;
    hfb_large_free(h);
    void *ptrs[] = {p.kappa, p.rho, p.orho, p.okappa, p.delta, p.fermi, p.pc, p.dfac, p.field,
                    p.qpe, p.tcan, p.can_psi, p.can_occ, p.work, p.status, p.tables,
                    p.mflags, p.epart};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p = Hfb{};
    return 0;
}

template <typename T>
int hfb_solve_impl(mocca_solve *h, bool first) {
    Hfb &p = h->hfb;
    const long N = h->N;
    const int M = h->M;
    (void)first;
    // SolvePairing :800-806: old rho, kappa take the current ones while the
    // old rho is all zero (the first call)
    int rc = run(h, "hfb_mix", [&] { hfb_launch_mix(h, 1); });
    if (rc) return rc;
    rc = run(h, "hfb_field", [&] {
        hipLaunchKernelGGL(k_hfb_cutoffs, dim3((M + 63) / 64), dim3(64), 0, h->st, p.k, h->b.E,
                           h->b.qn, p.fermi, M, p.pc);
        hipLaunchKernelGGL(k_hfb_dfac, dim3((N + 255) / 256), dim3(256), 0, h->st, p.k,
                           h->b.dens, N, p.dfac);
        hipLaunchKernelGGL(k_hfb_field<T>, dim3((N + kHfbThreads - 1) / kHfbThreads, 2),
                           dim3(kHfbThreads), 0, h->st, p.k, (const T *)h->b.psi, N, p.kappa,
                           p.pc, p.dfac, p.field);
    });
    if (rc) return rc;
    rc = run(h, "hfb_gaps", [&] {
        hipLaunchKernelGGL(k_hfb_gaps<T>, dim3(p.hmax * p.hmax, p.k.nblk), dim3(kHfbThreads), 0,
                           h->st, p.k, (const T *)h->b.psi, N, p.pc, p.field, h->mesh.dv,
                           p.delta);
    });
    if (rc) return rc;
    rc = run(h, "hfb_fermi", [&] {
        hipLaunchKernelGGL(k_hfb_broyden, dim3(2), dim3(kHfbThreads), p.lds_broyden, h->st, p.k,
                           h->b.E, p.delta, p.fermi, p.rho, p.kappa, p.qpe, p.work, p.status);
    });
    if (rc) return rc;
    rc = run(h, "hfb_canonical", [&] {
        hipLaunchKernelGGL(k_hfb_canonical, dim3(p.k.nblk), dim3(kHfbThreads), p.lds_canonical,
                           h->st, p.k, p.rho, p.tcan, p.can_occ, h->b.occ, p.status);
        hipLaunchKernelGGL(k_hfb_can_states<T>, dim3((N + kHfbThreads - 1) / kHfbThreads, M),
                           dim3(kHfbThreads), 0, h->st, p.k, (const T *)h->b.psi, N, p.tcan,
                           (T *)p.can_psi);
        hfb_launch_mix(h, 0);
    });
    return rc;
}

int hfb_solve_any(mocca_solve *h, bool first) {
    if (h->hfb.large) return hfb_large_solve_any(h);
    return h->prec == MOCCA_FP64 ? hfb_solve_impl<double>(h, first)
                                 : hfb_solve_impl<float>(h, first);
}

void hfb_launch_mix(mocca_solve *h, int first) {
    Hfb &p = h->hfb;
    const int grid = std::max(1, std::min(1024, (p.total + 255) / 256));
    hipLaunchKernelGGL(k_hfb_mix_zero, dim3(1), dim3(64), 0, h->st, p.mflags);
    hipLaunchKernelGGL(k_hfb_mix_flags, dim3(grid), dim3(256), 0, h->st, p.total, p.kappa,
                       p.orho, p.delta, p.mflags);
    hipLaunchKernelGGL(k_hfb_mix, dim3(grid), dim3(256), 0, h->st, p.total, p.k.mix, first, p.rho,
                       p.kappa, p.orho, p.okappa, p.mflags);
}

void hfb_launch_cutoffs(mocca_solve *h) {
    Hfb &p = h->hfb;
    const long N = h->N;
    hipLaunchKernelGGL(k_hfb_cutoffs, dim3((h->M + 63) / 64), dim3(64), 0, h->st, p.k, h->b.E,
                       h->b.qn, p.fermi, h->M, p.pc);
    hipLaunchKernelGGL(k_hfb_dfac, dim3((N + 255) / 256), dim3(256), 0, h->st, p.k, h->b.dens,
                       N, p.dfac);
}

int hfb_energy_impl(mocca_solve *h, double *row) {
    Hfb &p = h->hfb;
    return run(h, "hfb_energy", [&] {
        hipLaunchKernelGGL(k_hfb_energy_part, dim3(kHfbEnergyParts, p.k.nblk),
                           dim3(kHfbThreads), 0, h->st, p.k, p.delta, p.kappa, p.epart);
        hipLaunchKernelGGL(k_hfb_energy, dim3(1), dim3(64), 0, h->st, p.k, p.epart,
                           kHfbEnergyParts, p.fermi, row);
    });
}

int hfb_check_status(mocca_solve *h) {
    int st = 0;
    HIPCHK(h, hipMemcpyAsync(&st, h->hfb.status, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (st & 14) {
        HIPCHK(h, hipMemsetAsync(h->hfb.status, 0, sizeof(int), h->st));
        return fail(h, MOCCA_ERR_NUMERIC,
                    (st & 2) ? "NaN in the calculation of Fermi energies"
                             : (st & 4) ? "canonical occupations above one"
                                        : "the HFB eigen-solver did not converge");
    }
    h->hfb.unconverged = (st & 1) != 0;
    return 0;
}

// BCS: the parameters in p.k (cutoffs, density factor), Fermi energies,
// gaps 0 (set to 1 by the first SolvePairing), Eqp 1 (the Spwf default).
int bcs_setup(mocca_solve *h, const mocca_pairing *q) {
    Hfb &p = h->hfb;
    if (!h->scf.set) return fail(h, MOCCA_ERR_STATE, "mocca_scf_setup not called");
    if (q->iterations < 1) return fail(h, MOCCA_ERR_ARG, "BCS: iterations (PairingIter) < 1");
    const int M = h->M;
    const long N = h->N;
    HfbK &k = p.k;
    k = HfbK{};
    for (int it = 0; it < 2; ++it) {
        k.G[it] = q->strength[it];
        k.alpha[it] = q->alpha[it];
        k.cut[it] = q->cut[it];
        k.mu[it] = q->mu[it];
    }
    k.rhosat = q->rhosat;
    k.prec = q->fermi_prec;
    k.iters = q->iterations;
    k.particles[0] = h->scf.sk.neutrons;
    k.particles[1] = h->scf.sk.protons;
    int rc;
    if ((rc = dalloc(h, &p.delta, M * sizeof(double))) ||
        (rc = dalloc(h, &p.qpe, M * sizeof(double))) ||
        (rc = dalloc(h, &p.pc, M * sizeof(double))) ||
        (rc = dalloc(h, &p.fermi, 2 * sizeof(double))) ||
        (rc = dalloc(h, &p.dfac, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &p.field, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &p.work, 4 * sizeof(double))))
        return rc;
    std::vector<double> ones(M, 1.0);
    HIPCHK(h, hipMemset(p.delta, 0, M * sizeof(double)));
    HIPCHK(h, hipMemcpy(p.qpe, ones.data(), M * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(p.pc, 0, M * sizeof(double)));
    HIPCHK(h, hipMemcpy(p.fermi, q->fermi, 2 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(p.work, 0, 4 * sizeof(double)));
    p.bcs = true;
    return 0;
}

template <typename T>
int bcs_solve_impl(mocca_solve *h) {
    Hfb &p = h->hfb;
    const int M = h->M;
    const long N = h->N;
    const T *psi = (const T *)h->b.psi;
    double *rec = p.work;   // old lambda (2), converged flag
    int rc = run(h, "bcs", [&] {
        hipLaunchKernelGGL(k_bcs_start, dim3(1), dim3(256), 0, h->st, M, p.delta);
        hipLaunchKernelGGL(k_hfb_dfac, dim3((N + 255) / 256), dim3(256), 0, h->st, p.k,
                           h->b.dens, N, p.dfac);
    });
    if (rc) return rc;
    for (int outer = 0; outer < p.k.iters; ++outer) {
        rc = run(h, "bcs", [&] {
            hipLaunchKernelGGL(k_hfb_cutoffs, dim3((M + 63) / 64), dim3(64), 0, h->st, p.k,
                               h->b.E, h->b.qn, p.fermi, M, p.pc);
            hipLaunchKernelGGL(k_bcs_field<T>, dim3((N + 255) / 256, 2), dim3(256), 0, h->st,
                               psi, N, M, h->b.qn, p.pc, p.delta, p.qpe, p.dfac, p.field);
            hipLaunchKernelGGL(k_bcs_gaps<T>, dim3(M), dim3(kHfbThreads), 0, h->st, psi, N,
                               h->b.qn, p.pc, p.field, h->mesh.dv, p.delta);
            hipLaunchKernelGGL(k_bcs_fermi, dim3(1), dim3(64), 0, h->st, M, h->b.qn, h->b.E,
                               p.delta, p.k.particles[0], p.k.particles[1], p.fermi, p.qpe,
                               rec);
            hipLaunchKernelGGL(k_hfb_cutoffs, dim3((M + 63) / 64), dim3(64), 0, h->st, p.k,
                               h->b.E, h->b.qn, p.fermi, M, p.pc);
            hipLaunchKernelGGL(k_bcs_occ, dim3((M + 255) / 256), dim3(256), 0, h->st, M,
                               h->b.qn, h->b.E, p.delta, p.fermi, p.qpe, h->b.occ, p.k.prec,
                               rec);
        });
        if (rc) return rc;
        double conv = 0.0;
        HIPCHK(h, hipMemcpyAsync(&conv, rec + 2, sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
        if (conv != 0.0) break;
    }
    return 0;
}

int bcs_solve(mocca_solve *h) {
    return h->prec == MOCCA_FP64 ? bcs_solve_impl<double>(h) : bcs_solve_impl<float>(h);
}

int bcs_energy_impl(mocca_solve *h, double *row) {
    Hfb &p = h->hfb;
    return run(h, "hfb_energy", [&] {
        hipLaunchKernelGGL(k_bcs_energy, dim3(1), dim3(64), 0, h->st, h->M, h->b.qn, p.delta,
                           p.qpe, p.fermi, row);
    });
}

int pairing_energy_impl(mocca_solve *h, double *row) {
    if (h->hfb.on) return hfb_energy_impl(h, row);
    if (h->hfb.bcs) return bcs_energy_impl(h, row);
    return 0;
}

int hfb_setup(mocca_solve *h, const mocca_pairing *q) {
    hfb_free(h);
    Hfb &p = h->hfb;
    if (q) {
        p.in = *q;
        p.in_set = true;
    }
    if (!q || q->type == 0) return 0;
    if (q->type == 1) return bcs_setup(h, q);
    if (q->type != 2)
        return fail(h, MOCCA_ERR_UNSUPPORTED, "pairing type %d (0 HF, 1 BCS, 2 HFB)", q->type);
    if (!h->mesh.time_reversal)
        return fail(h, MOCCA_ERR_UNSUPPORTED, "HFB needs time-reversal symmetry here");
    if (!h->scf.set) return fail(h, MOCCA_ERR_STATE, "mocca_scf_setup not called");
    const int M = h->M;
    for (int i = 0; i < M; ++i)
        if (h->qn[4 * i + 2] != 1)
            return fail(h, MOCCA_ERR_UNSUPPORTED, "HFB with signature -1 states (state %d)", i);
    HfbK &k = p.k;
    k = HfbK{};
    // blocks in the reference's order: isospin outer, parity inner
    std::vector<int> st_all, slot_block(M, 0), slot_index(M, 0);
    int off = 0, toff = 0, nb = 0;
    for (int it = 0; it < 2; ++it)
        for (int P = 0; P < 2; ++P) {
            std::vector<int> st;
            for (int i = 0; i < M; ++i)
                if ((h->qn[4 * i] + 3) / 2 - 1 == it && (h->qn[4 * i + 1] + 3) / 2 - 1 == P)
                    st.push_back(i);
            if (st.empty()) continue;
            HfbBlockK &bk = k.blk[nb];
            bk.it = it;
            bk.P = P;
            bk.h = (int)st.size();
            bk.first = (int)st_all.size();
            bk.off = off;
            bk.toff = toff;
            off += 4 * bk.h * bk.h;
            toff += bk.h * bk.h;
            for (size_t a = 0; a < st.size(); ++a) {
                slot_block[st[a]] = nb;
                slot_index[st[a]] = (int)a;
            }
            st_all.insert(st_all.end(), st.begin(), st.end());
            p.hmax = std::max(p.hmax, bk.h);
            ++nb;
        }
    k.nblk = nb;
    std::vector<int> sblock(M), sindex(M);
    for (int i = 0; i < M; ++i) {
        sblock[i] = slot_block[i];
        sindex[i] = slot_index[i];
    }
    for (int it = 0; it < 2; ++it) {
        k.G[it] = q->strength[it];
        k.alpha[it] = q->alpha[it];
        k.cut[it] = q->cut[it];
        k.mu[it] = q->mu[it];
    }
    k.rhosat = q->rhosat;
    k.mix = q->hfb_mix;
    k.iters = q->iterations;
    k.prec = q->fermi_prec;
    k.particles[0] = h->scf.sk.neutrons;
    k.particles[1] = h->scf.sk.protons;
    p.total = off;
    // LDS: three n x n matrices per block of an isospin (Broyden), two h x h
    // (canonical)
    size_t lb = 0;
    for (int it = 0; it < 2; ++it) {
        size_t s = 0;
        for (int b = 0; b < nb; ++b)
            if (k.blk[b].it == it) s += 3 * (size_t)(2 * k.blk[b].h) * (2 * k.blk[b].h);
        lb = std::max(lb, s);
    }
    p.lds_broyden = lb * sizeof(double);
    p.lds_canonical = 2 * (size_t)p.hmax * p.hmax * sizeof(double);
    // the in-LDS Broyden solver while its matrices fit, else the large-block
    // Newton solver
    const bool small_ok = 2 * p.hmax <= kHfbMaxN && p.lds_broyden <= 60 * 1024;
    if (q->solver < 0 || q->solver > 4)
        return fail(h, MOCCA_ERR_ARG,
                    "pairing solver %d (0 automatic, 1 in-LDS Broyden, 2 large-block Broyden, "
                    "3 large-block Newton, 4 large-block Broyden with the sign function)",
                    q->solver);
    if (q->solver == 1 && !small_ok)
        return fail(h, MOCCA_ERR_UNSUPPORTED,
                    "HFB block of %d states: the in-LDS solver takes at most %d", p.hmax,
                    kHfbMaxN / 2);
    p.large = q->solver >= 2 || (q->solver == 0 && !small_ok);
    p.newton = q->solver == 3;
    p.sign = q->solver == 4 || (q->solver == 0 && p.large);
    if (p.large && !hfb_large_fits(p.hmax))
        return fail(h, MOCCA_ERR_UNSUPPORTED, "HFB block of %d states (at most 128)", p.hmax);
    const long N = h->N;
    int rc;
    if ((rc = dalloc(h, &p.kappa, off * sizeof(double))) ||
        (rc = dalloc(h, &p.rho, off * sizeof(double))) ||
        (rc = dalloc(h, &p.orho, off * sizeof(double))) ||
        (rc = dalloc(h, &p.okappa, off * sizeof(double))) ||
        (rc = dalloc(h, &p.delta, off * sizeof(double))) ||
        (rc = dalloc(h, &p.fermi, 2 * sizeof(double))) ||
        (rc = dalloc(h, &p.pc, M * sizeof(double))) ||
        (rc = dalloc(h, &p.dfac, 2 * N * sizeof(double))) ||
        (rc = dalloc(h, &p.field, 4 * N * sizeof(double))) ||
        (rc = dalloc(h, &p.qpe, M * sizeof(double))) ||
        (rc = dalloc(h, &p.tcan, std::max(1, toff) * sizeof(double))) ||
        (rc = dalloc(h, &p.can_psi, (size_t)M * 4 * N * h->tsize)) ||
        (rc = dalloc(h, &p.can_occ, M * sizeof(double))) ||
        (rc = dalloc(h, &p.work, 8)) || (rc = dalloc(h, &p.status, sizeof(int))) ||
        (rc = dalloc(h, &p.tables, 3 * (size_t)M * sizeof(int))) ||
        (rc = dalloc(h, &p.mflags, 2 * sizeof(int))) ||
        (rc = dalloc(h, &p.epart, (size_t)std::max(1, p.k.nblk) * kHfbEnergyParts *
                                      sizeof(double))))
        return rc;
    std::vector<int> tab(3 * (size_t)M);
    for (int i = 0; i < M; ++i) {
        tab[i] = st_all[i];
        tab[M + i] = sblock[i];
        tab[2 * M + i] = sindex[i];
    }
    HIPCHK(h, hipMemcpy(p.tables, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    k.state = p.tables;
    k.slot_block = p.tables + M;
    k.slot_index = p.tables + 2 * M;
    std::vector<double> zero(off, 0.0), kap(off, 0.0);
    if (q->guess_kappa) hfb_guess(k, kap);
    HIPCHK(h, hipMemcpy(p.kappa, kap.data(), off * sizeof(double), hipMemcpyHostToDevice));
    for (double *d : {p.rho, p.orho, p.okappa, p.delta})
        HIPCHK(h, hipMemcpy(d, zero.data(), off * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(p.fermi, q->fermi, 2 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(p.status, 0, sizeof(int)));
    HIPCHK(h, hipMemset(p.qpe, 0, M * sizeof(double)));
    HIPCHK(h, hipMemset(p.can_occ, 0, M * sizeof(double)));
    if (p.large)
        if (int rc2 = hfb_large_setup(h)) return rc2;
    p.on = true;
    p.started = false;
    return 0;
}

}  // namespace detail
}  // namespace mocca
