// HFB pairing step for blocks too large for the in-LDS solvers of
// mocca_hfb.hip (PairingType 2 of SolvePairing, Pairing.f90:743-890; time
// reversal and signature +1 states, as there).  Everything is FP64 except
// where the states are FP32 (mixed precision: they are widened).
//
//   gaps           Delta(a, h+a') = dv c_a c_a' sum_r [Re P Re F + Im P Im F]
//                  (HFBGaps, HFB.f90:520-632) as a matrix product
//                  D = Psi Ytilde^T over K = 4N, Ytilde a pointwise transform
//                  of the states by the pairing field: FP64 matrix cores;
//   pairing field  Delta_q(r) = -g(r) sum_a P(phi_a, psi_a),
//                  phi_a = sum_b c_a c_b kappa(h+a, b) psi_b (HFBPairingField,
//                  HFB.f90:350-518): the phi by the dense matrix-core apply
//                  with T in LDS, contracted with the states in its epilogue
//                  (launch_field_apply: phi never in HBM);
//   Fermi energies N(lambda) = particle number (HFBNumberofParticles,
//                  HFB.f90:810-1073) solved by the reference's Broyden
//                  iteration (HFBFindFermiEnergyBroyden :1849-2161: the same
//                  evaluation points and stopping rule, so also the same
//                  result where it stops unconverged), or on request by
//                  Newton's method with the exact derivative dN/dlambda =
//                  2 sum_{k+, l-} C_kl^2 / (E_k - E_l), C = V^T Sigma_z V
//                  (first-order perturbation of the eigenvectors by
//                  dT = -Sigma_z dlambda), safeguarded by a bracket, which
//                  reaches FermiPrec in fewer evaluations.  Each
//                  evaluation diagonalises the signature-reduced HFB matrix
//                  T = [[E - lambda, D], [D^T, -(E - lambda)]] of every block
//                  (DiagonaliseHFBHamiltonian_Signature :2695-2801) by cyclic
//                  Jacobi in the eigenbasis of the previous evaluation;
//   canonical basis eigenvectors of rho_h (DiagonaliseRhoHFB :3156-3285,
//                  HFBOccupations :3361-3561) by the same Jacobi, the
//                  canonical states by the dense matrix-core apply.
//
// The Jacobi kernel keeps the packed upper triangle of the matrix in LDS (up
// to 200 rows; in global memory above), runs whole sweeps of the round-robin
// ordering (n/2 disjoint rotations per round, every 2x2 block of the matrix
// updated from its old values in one pass) and logs each round's rotations;
// the eigenvectors are then rotated by the log in a separate kernel whose
// workgroups own disjoint rows of the basis, so the one-workgroup solver only
// streams the matrix.
#include "mocca_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mocca {

namespace {

constexpr int kJacThreads = 1024;
constexpr int kJacLdsMaxN = 200;     // packed upper triangle + rotation tables in 160 KiB
constexpr int kJacMaxSweeps = 24;
constexpr int kJacBatch = 3;         // 2x2 blocks a thread reads before it updates them
constexpr int kApplyRows = 4;        // basis rows per workgroup of the rotation apply (a wave each)
constexpr int kGramWaves = 1024;     // waves of one pair-Gram launch

typedef double hl_d4 __attribute__((ext_vector_type(4)));   // FP64 MFMA accumulator

__device__ __forceinline__ double hl_wave_sum(double a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

__device__ double hl_block_sum(double a, double *sh) {
    a = hl_wave_sum(a);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64, nw = blockDim.x / 64;
    __syncthreads();
    if (l == 0) sh[w] = a;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < nw; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// packed upper triangle, row-major: element (i, j), i <= j
__device__ __forceinline__ int pk(int i, int j, int n) { return i * n - (i * (i - 1)) / 2 + (j - i); }

// pair pp of round r of the round-robin ordering of m (even) players
__device__ __forceinline__ void rr_pair(int pp, int r, int m, int &i, int &j) {
    auto pos = [&](int x) { return x == 0 ? 0 : ((x - 1 + r) % (m - 1)) + 1; };
    i = pos(pp);
    j = pos(m - 1 - pp);
    if (i > j) { const int t = i; i = j; j = t; }
}

}  // namespace

// LDS of the Jacobi kernel in doubles before its reduction slots: (c, s),
// pair and the two packed row offsets of each of the M2 = ceil(n / 2) pairs
__host__ __device__ inline int jac_sh_offset(int n) {
    const int M2 = (n + (n & 1)) / 2;
    return (16 * M2 + 12 * M2 + 7) / 8;
}

// ---------------------------------------------------------------------------
// Cyclic Jacobi on the symmetric n x n matrices A[b] (dense, leading
// dimension ld, stride mstride; symmetrised on load), one workgroup each:
// diagonal into w[b][.], the final matrix into Afin[b] (dense, when given),
// the rotations of every round into log[b][round][pair] = (c, s) with their
// pairs in logij[b][round][pair] = i | j << 16, the number of rounds into
// rounds[b].  The rotation of pair (i, j) is that of Golub & Van Loan 8.4.2,
// applied as A <- J^T A J with J = [[c, s], [-s, c]] in rows/columns (i, j).
//
// The round-robin ordering pairs all n (+1) indices; sweeps stop when the
// off-diagonal square sum is below 1e-32 of the diagonal one.
// GA: the packed matrix in global memory (Ag[b]) instead of LDS.
template <bool GA, bool ODD>
__global__ __launch_bounds__(kJacThreads) void k_lj_jacobi(
    const double *__restrict__ Ain, int ld, long mstride, const int *__restrict__ nb,
    const int *__restrict__ active, double *__restrict__ Ag,
    double *__restrict__ w, double *__restrict__ Afin, double2 *__restrict__ log,
    int *__restrict__ logij, long logstride, int *__restrict__ rounds, int maxsweeps,
    int *__restrict__ status) {
    extern __shared__ __align__(16) double lds[];
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (active && !active[b]) return;
    const int n = nb[b], m = n + (n & 1), M2 = m / 2;
    // per round: rotation (c, s), pair i | j << 16 and the packed offsets
    // ro(i), ro(j) of each pair
    double2 *cs = reinterpret_cast<double2 *>(lds);                      // [M2]
    int *pij = reinterpret_cast<int *>(cs + M2);                         // [M2]
    int *pri = pij + M2, *prj = pri + M2;                                // [M2]
    double *sh = reinterpret_cast<double *>(lds) + jac_sh_offset(n);     // [16]
    double *A = GA ? Ag + b * mstride : sh + 16;                         // packed upper
    // packed offset of row x less x: element (x, y), x <= y, at ro(x) + y
    auto ro = [&](int x) { return x * n - (x * (x - 1)) / 2 - x; };
    const double *Ab = Ain + b * mstride;
    for (int q = tid; q < n * n; q += nt) {
        const int i = q / n, j = q % n;
        if (j >= i) A[ro(i) + j] = 0.5 * (Ab[(long)i * ld + j] + Ab[(long)j * ld + i]);
    }
    __syncthreads();
    double2 *lb = log + b * logstride;
    int *li = logij + b * logstride;
    // 2x2 blocks (k, l), k <= l, over threads: column pairs u and M2-1-u are
    // folded into one list of M2+1 blocks (M2/2+... lists of equal length),
    // GT threads per list
    const int L2 = (M2 + 1) / 2;
    const int GT = nt / L2 > 0 ? nt / L2 : 1;
    const int u = tid / GT, g = tid % GT;
    const bool worker = u < L2;
    const int l1 = u, l2 = M2 - 1 - u;
    const int cnt = worker ? (l1 == l2 ? l1 + 1 : M2 + 1) : 0;
    int done = 0;
    bool conv = false;
    for (int sweep = 0; sweep <= maxsweeps; ++sweep) {
        double off = 0, dia = 0;
        for (int i = tid; i < n; i += nt) {
            const double *row = A + ro(i);
            dia += row[i] * row[i];
            for (int j = i + 1; j < n; ++j) off += 2.0 * row[j] * row[j];
        }
        off = hl_block_sum(off, sh);
        dia = hl_block_sum(dia, sh);
        if (off <= 1e-32 * dia || off == 0.0) { conv = true; break; }
        if (sweep == maxsweeps) break;
        for (int r = 0; r < m - 1; ++r) {
            if (tid < M2) {
                const int pp = tid;
                int i, j;
                rr_pair(pp, r, m, i, j);
                double c = 1.0, s = 0.0;
                const int ri = ro(i), rj = j < n ? ro(j) : 0;
                if (j < n) {
                    const double apq = A[ri + j];
                    // a rotation by an element below 1e-18 of the geometric
                    // mean of its diagonal pair moves those diagonals by
                    // ~1e-36 relative: it is skipped, so that the sweeps
                    // that only confirm convergence cost little
                    const double aii = A[ri + i], ajj = A[rj + j];
                    if (apq != 0.0 && fabs(apq) > 1e-18 * sqrt(fabs(aii * ajj))) {
                        const double tau = (ajj - aii) / (2.0 * apq);
                        const double t = tau >= 0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                                  : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                cs[pp] = make_double2(c, s);
                pij[pp] = i | (j << 16);
                pri[pp] = ri;
                prj[pp] = rj;
                lb[(long)done * M2 + pp] = make_double2(c, s);
                li[(long)done * M2 + pp] = i | (j << 16);
            }
            __syncthreads();
            // the column pair of each of this thread's two lists
            const double2 cl1 = worker ? cs[l1] : make_double2(1.0, 0.0);
            const double2 cl2 = worker ? cs[l2] : make_double2(1.0, 0.0);
            const int4 R1 = worker ? make_int4(pij[l1] & 0xffff, pij[l1] >> 16, pri[l1], prj[l1])
                                   : make_int4(0, 0, 0, 0);
            const int4 R2 = worker ? make_int4(pij[l2] & 0xffff, pij[l2] >> 16, pri[l2], prj[l2])
                                   : make_int4(0, 0, 0, 0);
            // kJacBatch blocks at a time: their tables, then their elements,
            // are read back to back so that the LDS latencies overlap; every
            // element belongs to one block, so the reads of a batch see the
            // old values and its writes touch nothing another block reads
            for (int q0 = g; q0 < cnt; q0 += kJacBatch * GT) {
                int e[kJacBatch][4];
                double x[kJacBatch][4];
                double2 ck[kJacBatch], cl[kJacBatch];
                int kind[kJacBatch];   // 0 none, 1 off-diagonal block, 2 diagonal block
#pragma unroll
                for (int t = 0; t < kJacBatch; ++t) {
                    const int q = q0 + t * GT;
                    kind[t] = 0;
                    if (q >= cnt) continue;
                    const bool first = q <= l1;
                    const int k = first ? q : q - l1 - 1, l = first ? l1 : l2;
                    const int4 Q = first ? R1 : R2;
                    cl[t] = first ? cl1 : cl2;
                    ck[t] = cs[k];
                    // a block whose two rotations are both the identity
                    // keeps its values (the sweeps that only confirm
                    // convergence read nothing more)
                    if (ck[t].y == 0.0 && cl[t].y == 0.0) continue;
                    const int w1 = pij[k];
                    const int4 P = make_int4(w1 & 0xffff, w1 >> 16, pri[k], prj[k]);
                    if (k == l) {
                        kind[t] = (ODD && P.y >= n) || ck[t].y == 0.0 ? 0 : 2;
                        e[t][0] = P.z + P.x;   // (i, i)
                        e[t][1] = P.z + P.y;   // (i, j)
                        e[t][3] = P.w + P.y;   // (j, j)
                        e[t][2] = 0;
                        continue;
                    }
                    kind[t] = 1;
                    // rows (P.x, P.y) x columns (Q.x, Q.y); with an odd
                    // size, entries with the index n (the dummy player) stay
                    // out (their offset is set to -1)
                    e[t][0] = P.x <= Q.x ? P.z + Q.x : Q.z + P.x;
                    e[t][1] = P.x <= Q.y ? P.z + Q.y : Q.w + P.x;
                    e[t][2] = P.y <= Q.x ? P.w + Q.x : Q.z + P.y;
                    e[t][3] = P.y <= Q.y ? P.w + Q.y : Q.w + P.y;
                    if (ODD) {
                        if (Q.y >= n) e[t][1] = e[t][3] = -1;
                        if (P.y >= n) e[t][2] = e[t][3] = -1;
                    }
                }
#pragma unroll
                for (int t = 0; t < kJacBatch; ++t) {
                    if (kind[t] == 0) continue;
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        x[t][c] = (kind[t] == 2 && c == 2) || (ODD && e[t][c] < 0) ? 0.0
                                                                                   : A[e[t][c]];
                }
#pragma unroll
                for (int t = 0; t < kJacBatch; ++t) {
                    if (kind[t] == 0) continue;
                    const double2 a = ck[t], bq = cl[t];
                    if (kind[t] == 2) {
                        // J^T [[aii, aij], [aij, ajj]] J, off-diagonal exactly 0
                        const double aii = x[t][0], aij = x[t][1], ajj = x[t][3];
                        const double c = a.x, s = a.y;
                        A[e[t][0]] = c * c * aii - 2.0 * c * s * aij + s * s * ajj;
                        A[e[t][3]] = s * s * aii + 2.0 * c * s * aij + c * c * ajj;
                        A[e[t][1]] = 0.0;
                        continue;
                    }
                    // Y = Jk^T X Jl, Jk = [[ck, sk], [-sk, ck]]
                    const double a0 = a.x * x[t][0] - a.y * x[t][2];
                    const double a1 = a.x * x[t][1] - a.y * x[t][3];
                    const double b0 = a.y * x[t][0] + a.x * x[t][2];
                    const double b1 = a.y * x[t][1] + a.x * x[t][3];
                    A[e[t][0]] = a0 * bq.x - a1 * bq.y;
                    if (!ODD || e[t][1] >= 0) A[e[t][1]] = a0 * bq.y + a1 * bq.x;
                    if (!ODD || e[t][2] >= 0) A[e[t][2]] = b0 * bq.x - b1 * bq.y;
                    if (!ODD || e[t][3] >= 0) A[e[t][3]] = b0 * bq.y + b1 * bq.x;
                }
            }
            __syncthreads();
            ++done;
        }
    }
    double *wb = w + (long)b * ld;
    for (int i = tid; i < n; i += nt) wb[i] = A[ro(i) + i];
    if (Afin) {
        double *Fb = Afin + b * mstride;
        for (int q = tid; q < n * n; q += nt) {
            const int i = q / n, j = q % n;
            Fb[(long)i * ld + j] = i <= j ? A[ro(i) + j] : A[ro(j) + i];
        }
    }
    if (tid == 0) {
        rounds[b] = done;
        if (!conv) atomicOr(status, 8);
    }
}

// ---------------------------------------------------------------------------
// The same cyclic Jacobi for even n with the rows and columns kept in slot
// order: the round-robin ordering of the circle method seats the players
// (matrix indices) in slots, pair k of a round being slots (2k, 2k+1); after
// each round every player but the one in slot 0 moves one seat along the
// circle 2 -> 4 -> ... -> m-2 -> m-1 -> m-3 -> ... -> 1 -> 2.  The matrix is
// stored by 2x2 blocks of slot pairs (K, L), K <= L (packed upper triangle
// of blocks; a diagonal block holds all four elements), element (a, c) of
// block (K, L) in plane 2a + c, so the lanes of a wave read consecutive
// blocks of a plane.  A round reads every block into registers, then writes
// the rotated elements to the seats their players move to: every element is
// read and written once per round, as in place, with the moves for free and
// no index tables.  slot_player[s] is the player in seat s; the log holds
// each rotation with its players (i < j), the rotation being that of
// Golub & Van Loan 8.4.2 for (i, j) as in k_lj_jacobi.
__device__ __forceinline__ int seat_next(int u, int m) {
    if (u == 0) return 0;
    if (!(u & 1)) return u == m - 2 ? m - 1 : u + 2;
    return u == 1 ? 2 : u - 2;
}
__device__ __forceinline__ int blk_index(int K, int L, int M2) {
    return K * M2 - (K * (K - 1)) / 2 + (L - K);
}
// block row K of packed block index t: the largest K with
// blk_index(K, K) <= t, K = floor((2 M2 + 1 - sqrt((2 M2 + 1)^2 - 8 t)) / 2)
// (the radicand is an integer below 2^24, its square root exact enough in
// FP64 that the floor is right, exact where it is a perfect square)
__device__ __forceinline__ int blk_row(int t, int M2) {
    const int a = 2 * M2 + 1;
    return (int)((a - sqrt((double)(a * a - 8 * t))) * 0.5);
}
size_t slots_jacobi_lds(int n) {
    const int M2 = n / 2, nb = M2 * (M2 + 1) / 2;
    return sizeof(double) * (4 * (size_t)nb + 2 * (size_t)M2 + 16) +
           sizeof(short) * (size_t)((n + 3) / 4 * 4);
}
constexpr int kSlotThreads = 1024;   // the rotated blocks of a round stay in registers
__global__ __launch_bounds__(kSlotThreads) void k_lj_jacobi_slots(
    const double *__restrict__ Ain, int ld, long mstride, const int *__restrict__ nb,
    const int *__restrict__ active, double *__restrict__ w, double *__restrict__ Afin,
    double2 *__restrict__ log, int *__restrict__ logij, long logstride,
    int *__restrict__ rounds, int maxsweeps, int *__restrict__ status) {
    extern __shared__ __align__(16) double lds_s[];
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (active && !active[b]) return;
    const int n = nb[b], m = n, M2 = n / 2, NB = M2 * (M2 + 1) / 2;
    double *E = lds_s;                                              // [4][NB]
    double2 *cs = reinterpret_cast<double2 *>(E + 4 * NB);          // [M2] slot rotations
    double *sh = reinterpret_cast<double *>(cs + M2);               // [16]
    short *sp = reinterpret_cast<short *>(sh + 16);                 // [n] slot -> player
    // location of element (seat u, seat v): plane and block index, with
    // *mir set when it sits in a diagonal block off its diagonal (the
    // mirror position holds the same value)
    auto loc = [&](int u, int v, int &mir) {
        int K = u >> 1, a = u & 1, L = v >> 1, c = v & 1;
        mir = 0;
        if (K > L) {
            const int t = K; K = L; L = t;
            const int q = a; a = c; c = q;
        } else if (K == L && a != c) {
            mir = 1;
        }
        return (2 * a + c) * NB + blk_index(K, L, M2);
    };
    const double *Ab = Ain + b * mstride;
    for (int q = tid; q < n * n; q += nt) {
        const int i = q / n, j = q % n;
        if ((i >> 1) <= (j >> 1)) {
            int mir;
            E[loc(i, j, mir)] = 0.5 * (Ab[(long)i * ld + j] + Ab[(long)j * ld + i]);
        }
    }
    for (int i = tid; i < n; i += nt) sp[i] = (short)i;
    __syncthreads();
    double2 *lb = log + b * logstride;
    int *li = logij + b * logstride;
    constexpr int kMaxBlk = (kJacLdsMaxN / 2 * (kJacLdsMaxN / 2 + 1) / 2 + kSlotThreads - 1) /
                            kSlotThreads;
    // this thread's blocks t = tid + q nt: block (K, L) and, for each of its
    // elements, where the seat move sends it (the same every round): 16-bit
    // offsets, two per word, 0xffff for none (element (1, 0) of a diagonal
    // block, which is its (0, 1)); bit 4q + e of mirb set when the target
    // needs its mirror too
    int kl[kMaxBlk], tg[kMaxBlk][2], mirb = 0;
#pragma unroll
    for (int q = 0; q < kMaxBlk; ++q) {
        const int t = tid + q * nt;
        kl[q] = -1;
        tg[q][0] = tg[q][1] = -1;
        if (t >= NB) continue;
        const int K = blk_row(t, M2), L = K + (t - blk_index(K, K, M2));
        kl[q] = K | (L << 16);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int o = 0xffff, mir = 0;
            if (!(K == L && e == 2))
                o = loc(seat_next(2 * K + (e >> 1), m), seat_next(2 * L + (e & 1), m), mir);
            const int sh16 = 16 * (e & 1);
            tg[q][e >> 1] = (tg[q][e >> 1] & ~(0xffff << sh16)) | (o << sh16);
            if (mir) mirb |= 1 << (4 * q + e);
        }
    }
    int done = 0;
    bool conv = false;
    for (int sweep = 0; sweep <= maxsweeps; ++sweep) {
        double off = 0, dia = 0;
        for (int t = tid; t < NB; t += nt) {
            // block (K, L) of index t
            const int K = blk_row(t, M2), L = K + (t - blk_index(K, K, M2));
            const double x0 = E[t], x1 = E[NB + t], x2 = E[2 * NB + t], x3 = E[3 * NB + t];
            if (K == L) {
                dia += x0 * x0 + x3 * x3;
                off += 2.0 * x1 * x1;
            } else {
                off += 2.0 * (x0 * x0 + x1 * x1 + x2 * x2 + x3 * x3);
            }
        }
        off = hl_block_sum(off, sh);
        dia = hl_block_sum(dia, sh);
        if (off <= 1e-32 * dia || off == 0.0) { conv = true; break; }
        if (sweep == maxsweeps) break;
        for (int r = 0; r < m - 1; ++r) {
            // rotations of the pairs (players of seats 2k, 2k+1)
            if (tid < M2) {
                const int k = tid, dk = blk_index(k, k, M2);
                const int pu = sp[2 * k], pv = sp[2 * k + 1];
                const bool sw = pu > pv;          // player i of (i < j) in seat 2k+1
                const double xuu = E[dk], xuv = E[NB + dk], xvv = E[3 * NB + dk];
                const double aii = sw ? xvv : xuu, ajj = sw ? xuu : xvv, apq = xuv;
                double c = 1.0, s = 0.0;
                // (a rotation by an element below 1e-18 of the geometric mean
                // of its diagonal pair is skipped, as in k_lj_jacobi)
                if (apq != 0.0 && fabs(apq) > 1e-18 * sqrt(fabs(aii * ajj))) {
                    const double tau = (ajj - aii) / (2.0 * apq);
                    const double t = tau >= 0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                              : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                }
                cs[k] = make_double2(c, sw ? -s : s);
                lb[(long)done * M2 + k] = make_double2(c, s);
                li[(long)done * M2 + k] = sw ? (pv | (pu << 16)) : (pu | (pv << 16));
            }
            __syncthreads();
            // every block into registers (and the players of the seats)
            double x[kMaxBlk][4];
#pragma unroll
            for (int q = 0; q < kMaxBlk; ++q) {
                if (kl[q] < 0) continue;
                const int t = tid + q * nt;
#pragma unroll
                for (int e = 0; e < 4; ++e) x[q][e] = E[e * NB + t];
            }
            short pl = 0;
            if (tid < n) pl = sp[tid];
            __syncthreads();
            if (tid < n) sp[seat_next(tid, m)] = pl;
#pragma unroll
            for (int q = 0; q < kMaxBlk; ++q) {
                if (kl[q] < 0) continue;
                const int K = kl[q] & 0xffff, L = kl[q] >> 16;
                const double2 a = cs[K], bq = cs[L];
                double y[4];
                if (K == L) {
                    // J^T [[x00, x01], [x01, x11]] J, off-diagonal exactly 0
                    const double c = a.x, sn = a.y;
                    y[0] = c * c * x[q][0] - 2.0 * c * sn * x[q][1] + sn * sn * x[q][3];
                    y[3] = sn * sn * x[q][0] + 2.0 * c * sn * x[q][1] + c * c * x[q][3];
                    y[1] = 0.0;
                    y[2] = 0.0;
                } else {
                    // Y = JK^T X JL
                    const double a0 = a.x * x[q][0] - a.y * x[q][2];
                    const double a1 = a.x * x[q][1] - a.y * x[q][3];
                    const double b0 = a.y * x[q][0] + a.x * x[q][2];
                    const double b1 = a.y * x[q][1] + a.x * x[q][3];
                    y[0] = a0 * bq.x - a1 * bq.y;
                    y[1] = a0 * bq.y + a1 * bq.x;
                    y[2] = b0 * bq.x - b1 * bq.y;
                    y[3] = b0 * bq.y + b1 * bq.x;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int o = (tg[q][e >> 1] >> (16 * (e & 1))) & 0xffff;
                    if (o == 0xffff) continue;
                    E[o] = y[e];
                    // the mirror of an element (0, 1) of a diagonal block is
                    // its (1, 0), one plane further, and back
                    if (mirb & (1 << (4 * q + e))) E[o < 2 * NB ? o + NB : o - NB] = y[e];
                }
            }
            __syncthreads();
            ++done;
        }
    }
    // diagonal and the final matrix in player order
    double *wb = w + (long)b * ld;
    for (int u = tid; u < n; u += nt) {
        int mir;
        wb[sp[u]] = E[loc(u, u, mir)];
    }
    if (Afin) {
        double *Fb = Afin + b * mstride;
        for (int q = tid; q < n * n; q += nt) {
            const int u = q / n, v = q % n;
            int mir;
            Fb[(long)sp[u] * ld + sp[v]] = E[loc(u, v, mir)];
        }
    }
    if (tid == 0) {
        rounds[b] = done;
        if (!conv) atomicOr(status, 8);
    }
}

// The logged rotations applied to the basis: V[b] <- V[b] J_1 J_2 ... (in
// place; V = identity first when ident or cold[b]).  kApplyLanes lanes own
// one row of the basis in LDS and take disjoint rotations of each round
// (pairs q, q + kApplyLanes, ...), so a round touches disjoint columns of a
// row and the rounds follow one another without a barrier: the lanes of a
// row are in one wave, whose LDS accesses complete in order.  The pairs of
// each round come from the pair log.
constexpr int kApplyLanes = 64;
constexpr int kApplyMaxPairs = 128 / kApplyLanes;
constexpr int kApplyChunk = 16;        // rounds of the log staged in LDS at a time
size_t apply_lds(int n) {
    const int M2 = (n + (n & 1)) / 2;
    return sizeof(double) * kApplyRows * n + (sizeof(double2) + sizeof(int)) * kApplyChunk * M2;
}
__global__ __launch_bounds__(256) void k_lj_apply(double *__restrict__ V, int ld, long mstride,
                                                  const int *__restrict__ nb,
                                                  const int *__restrict__ active,
                                                  const double2 *__restrict__ log,
                                                  const int *__restrict__ logij,
                                                  long logstride, const int *__restrict__ rounds,
                                                  int ident, const int *__restrict__ cold) {
    extern __shared__ __align__(16) double lds_apply[];
    const int b = blockIdx.y, tid = threadIdx.x;
    if (active && !active[b]) return;
    const int n = nb[b], m = n + (n & 1), M2 = m / 2;
    const int r0 = blockIdx.x * kApplyRows;
    if (r0 >= n) return;
    const int nr = min(kApplyRows, n - r0);
    double2 *lgc = reinterpret_cast<double2 *>(lds_apply);   // [kApplyChunk][M2]
    double *rows = reinterpret_cast<double *>(lgc + kApplyChunk * M2);   // [kApplyRows][n]
    int *lij = reinterpret_cast<int *>(rows + kApplyRows * n);          // [kApplyChunk][M2]
    double *Vb = V + b * mstride;
    const bool id = ident || (cold && cold[b]);
    for (int e = tid; e < nr * n; e += blockDim.x) {
        const int r = e / n, c = e % n;
        rows[e] = id ? (r0 + r == c ? 1.0 : 0.0) : Vb[(long)(r0 + r) * ld + c];
    }
    const int row = tid / kApplyLanes, q0 = tid % kApplyLanes;
    double *rw = rows + row * n;
    const bool mine = row < nr;
    const double2 *lb = log + b * logstride;
    const int *lb2 = logij + b * logstride;
    const int R = rounds[b];
    for (int g0 = 0; g0 < R; g0 += kApplyChunk) {
        const int nc = min(kApplyChunk, R - g0);
        __syncthreads();
        for (int e = tid; e < nc * M2; e += blockDim.x) {
            lgc[e] = lb[(long)g0 * M2 + e];
            lij[e] = lb2[(long)g0 * M2 + e];
        }
        __syncthreads();
        if (!mine) continue;
        for (int gc = 0; gc < nc; ++gc) {
            const double2 *lg = lgc + gc * M2;
            const int *lp = lij + gc * M2;
#pragma unroll
            for (int t = 0; t < kApplyMaxPairs; ++t) {
                const int pp = q0 + kApplyLanes * t;
                if (pp < M2) {
                    const double2 cs = lg[pp];
                    const int ij = lp[pp];
                    const int i = ij & 0xffff, j = ij >> 16;
                    if (j < n && cs.y != 0.0) {
                        const double vi = rw[i], vj = rw[j];
                        rw[i] = cs.x * vi - cs.y * vj;
                        rw[j] = cs.y * vi + cs.x * vj;
                    }
                }
            }
            // the lanes of a row are in one wave, whose LDS accesses
            // complete in order: the next round may read what this one wrote
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    for (int e = tid; e < nr * n; e += blockDim.x) {
        const int r = e / n, c = e % n;
        Vb[(long)(r0 + r) * ld + c] = rows[e];
    }
}

// C[b] = op(A[b]) B'[b] for n x n matrices (leading dimension ld), where
// op = transpose when TA and B' = B with its rows k >= h[b] negated when SZ
// (Sigma_z B); one 16 x 16 tile of C per workgroup.
template <bool TA, bool SZ>
__global__ __launch_bounds__(256) void k_lj_gemm(const double *__restrict__ A,
                                                 const double *__restrict__ B,
                                                 double *__restrict__ C, int ld, long mstride,
                                                 const int *__restrict__ nb,
                                                 const int *__restrict__ hb,
                                                 const int *__restrict__ active) {
    __shared__ double As[16][17], Bs[16][17];
    const int b = blockIdx.z;
    if (active && !active[b]) return;
    const int n = nb[b], h = SZ ? hb[b] : 0;
    const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
    if (i0 >= n || j0 >= n) return;
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    const double *Ab = A + b * mstride, *Bb = B + b * mstride;
    double acc = 0;
    for (int k0 = 0; k0 < n; k0 += 16) {
        const int ia = i0 + ty, ka = k0 + tx;
        As[ty][tx] = (ia < n && ka < n) ? (TA ? Ab[(long)ka * ld + ia] : Ab[(long)ia * ld + ka])
                                        : 0.0;
        const int kb = k0 + ty, jb = j0 + tx;
        double bv = (kb < n && jb < n) ? Bb[(long)kb * ld + jb] : 0.0;
        if (SZ && kb >= h) bv = -bv;
        Bs[ty][tx] = bv;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += As[ty][k] * Bs[k][tx];
        __syncthreads();
    }
    if (i0 + ty < n && j0 + tx < n) C[b * mstride + (long)(i0 + ty) * ld + j0 + tx] = acc;
}

// Four consecutive values (16- or 32-byte aligned) by vector loads.
__device__ __forceinline__ void ld4(const double *p, double (&v)[4]) {
    const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void ld4(const float *p, float (&v)[4]) {
    const float4 a = *(const float4 *)p;
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}

// ---------------------------------------------------------------------------
// Gaps: partial sums S[a][a'] = sum_K psi_a[K] Ytilde_a'[K] over a K slice
// per wave, all tiles of row tiles 0..NI-1 (NI = the block's tile count)
// against the column tiles J0 = Z NJ .. J0+NJ-1 (group Z = blockIdx.z).
// With F = Fr + i Fi the pairing field, the transformed partner state is
//   Ytilde_c = -Fr y_c + Fi y_{c+1} (c even),  -Fr y_c - Fi y_{c-1} (c odd),
// so that sum_c x_c Ytilde_c = Re P(x, y) Fr + Im P(x, y) Fi for the
// time-reversal pair amplitude P (GetPairDensity, PairingInteraction.f90:
// 101-180).  K is walked in sub-chunks: component c of 16 consecutive
// points, lane group g = lane / 16 holding points 4g .. 4g+3 of the state
// 16 I + lane % 16, so that one load instruction reads whole 128-byte lines
// of a row (with one component per lane group, the rows' lines were
// fetched from HBM several times over).  The partner tiles are row tiles
// of the same block: component c comes from the rows already in registers,
// component c^1 from a second read of the partner rows.  Sub-chunk s+1 is
// loaded in the storage type, without branches (points past N and states
// past h read point / state 0), while the MFMAs of sub-chunk s run, and is
// widened to FP64 with those entries zeroed afterwards; with N % 4 == 0
// (V4) a lane's four points are one aligned vector load.  Result lane L,
// register r: S[16 I + L/16 + 4 r][16 J + L%16].
template <typename T, int NI, int NJ, int Z, bool V4>
__device__ __forceinline__ void pair_gram_group(const HfbK &k, const T *__restrict__ psi, long N,
                                                const double *__restrict__ field,
                                                double *__restrict__ part, int nsplit) {
    constexpr int J0 = Z * NJ;
    constexpr int NJE = NI - J0 < NJ ? NI - J0 : NJ;   // column tiles of this group
    const int w = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, g = lane / 16;
    const double *Fr = field + (2 * bk.it) * N, *Fi = Fr + N;
    const T *xr[NI];
    const T *yr[NJE];
    bool xok[NI];
#pragma unroll
    for (int I = 0; I < NI; ++I) {
        const int a = 16 * I + lane % 16;
        xok[I] = a < h;
        xr[I] = psi + (long)k.state[bk.first + (xok[I] ? a : 0)] * 4 * N;
    }
#pragma unroll
    for (int J = 0; J < NJE; ++J) {
        const int a = 16 * (J0 + J) + lane % 16;
        yr[J] = psi + (long)k.state[bk.first + (a < h ? a : 0)] * 4 * N;
    }
    hl_d4 acc[NI][NJE];
#pragma unroll
    for (int I = 0; I < NI; ++I)
#pragma unroll
        for (int J = 0; J < NJE; ++J) acc[I][J] = hl_d4{0, 0, 0, 0};
    // sub-chunk s: component s % 4 of the 16 points of chunk s / 4
    const long nch = (N + 15) / 16;
    const long s0 = 4 * (nch * w / nsplit), s1 = 4 * (nch * (w + 1) / nsplit);
    T raw[NI][4], rawy[NJE][4];
    double fraw[2][4];
    auto load = [&](long sc) {
        const long pb = 16 * (sc >> 2) + 4 * g;
        const long cx = (long)(sc & 3) * N, cy = (long)((sc & 3) ^ 1) * N;
        if constexpr (V4) {
            // N % 4 == 0: a lane's 4 points are in range together, aligned
            const long p = pb < N ? pb : 0;
            ld4(Fr + p, fraw[0]);
            ld4(Fi + p, fraw[1]);
#pragma unroll
            for (int I = 0; I < NI; ++I) ld4(xr[I] + cx + p, raw[I]);
#pragma unroll
            for (int J = 0; J < NJE; ++J) ld4(yr[J] + cy + p, rawy[J]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long p = pb + q < N ? pb + q : 0;
                fraw[0][q] = Fr[p];
                fraw[1][q] = Fi[p];
#pragma unroll
                for (int I = 0; I < NI; ++I) raw[I][q] = xr[I][cx + p];
#pragma unroll
                for (int J = 0; J < NJE; ++J) rawy[J][q] = yr[J][cy + p];
            }
        }
    };
    double x[NI][4], y1[NJE][4], f[2][4];
    auto widen = [&](long sc) {
        const long pb = 16 * (sc >> 2) + 4 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool pin = pb + q < N;
            f[0][q] = fraw[0][q];
            f[1][q] = fraw[1][q];
#pragma unroll
            for (int I = 0; I < NI; ++I) x[I][q] = (pin && xok[I]) ? (double)raw[I][q] : 0.0;
#pragma unroll
            for (int J = 0; J < NJE; ++J)
                y1[J][q] = (pin && xok[J0 + J]) ? (double)rawy[J][q] : 0.0;
        }
    };
    if (s0 < s1) {
        load(s0);
        widen(s0);
    }
    for (long sc = s0; sc < s1; ++sc) {
        if (sc + 1 < s1) load(sc + 1);
        // keep the widening of sub-chunk sc+1 behind the MFMAs of sc (the
        // scheduler would otherwise pull it up and wait for the loads there)
        __builtin_amdgcn_sched_barrier(0);
        const double sg = (sc & 1) ? -1.0 : 1.0;   // odd components take -Fi
        double yt[NJE][4];
#pragma unroll
        for (int J = 0; J < NJE; ++J)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                yt[J][q] = -f[0][q] * x[J0 + J][q] + sg * f[1][q] * y1[J][q];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int I = 0; I < NI; ++I)
#pragma unroll
                for (int J = 0; J < NJE; ++J)
                    acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[I][q], yt[J][q],
                                                                     acc[I][J], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        widen(sc + 1);   // after the last sub-chunk: unused
    }
    double *P = part + ((long)b * nsplit + w) * NI * NI * 256 + lane * 4;
#pragma unroll
    for (int I = 0; I < NI; ++I)
#pragma unroll
        for (int J = 0; J < NJE; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) P[(I * NI + J0 + J) * 256 + r] = acc[I][J][r];
}

// One wave per (K slice, block, column-tile group); the group index selects
// a compile-time tile range, so every register array is indexed statically.
template <typename T, int NI, int NJ, bool V4>
__global__ __launch_bounds__(64) void k_hl_pair_gram(HfbK k, const T *__restrict__ psi, long N,
                                                     const double *__restrict__ field,
                                                     double *__restrict__ part, int nsplit) {
    constexpr int NZ = (NI + NJ - 1) / NJ;
    static_assert(NZ <= 3, "at most three column-tile groups");
    const int z = blockIdx.z;
    if (z == 0) {
        pair_gram_group<T, NI, NJ, 0, V4>(k, psi, N, field, part, nsplit);
    } else if constexpr (NZ > 1) {
        if (z == 1) {
            pair_gram_group<T, NI, NJ, 1, V4>(k, psi, N, field, part, nsplit);
        } else if constexpr (NZ > 2) {
            if (z == 2) pair_gram_group<T, NI, NJ, 2, V4>(k, psi, N, field, part, nsplit);
        }
    }
}
// Delta(a, h+a') = dv c_a c_a' sum_slices S, Delta(h+a', a) = -Delta(a, h+a'),
// zero below the pair cut (HFBGaps :520-632): one thread per (a, a').
__global__ void k_hl_gaps_reduce(HfbK k, const double *__restrict__ part, int nsplit, int nt,
                                 const double *__restrict__ pc, double dv,
                                 double *__restrict__ delta) {
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= h * h) return;
    const int a = e / h, a2 = e % h;
    const int s1 = k.state[bk.first + a], s2 = k.state[bk.first + a2];
    const double cc = pc[s1] * pc[s2];
    double *dl = delta + bk.off;
    if (cc < kHfbNumCut) {
        dl[a * n + h + a2] = 0.0;
        dl[(h + a2) * n + a] = 0.0;
        return;
    }
    const int I = a / 16, J = a2 / 16, ri = a % 16, cj = a2 % 16;
    // lane L = 16 (ri % 4) + cj, register r = ri / 4
    const int idx = (16 * (ri % 4) + cj) * 4 + ri / 4;
    double s = 0;
#pragma unroll 8
    for (int w = 0; w < nsplit; ++w)
        s += part[(((long)b * nsplit + w) * nt * nt + I * nt + J) * 256 + idx];
    const double d = dv * pc[s1] * pc[s2] * s;
    dl[a * n + h + a2] = d;
    dl[(h + a2) * n + a] = -d;
}

// The apply matrices of the pairing field: Tap[b][b1][a2] = c_b1 c_a2
// kappa(h+a2, b1), zero for c_b1 or the product below the pair cut (the
// pairs HFBPairingField drops); the block lists and sizes for the apply.
__global__ void k_hl_field_w(HfbK k, const double *__restrict__ kappa,
                             const double *__restrict__ pc, int hmax,
                             double *__restrict__ Tap, int *__restrict__ list,
                             int *__restrict__ hlist) {
    const int b = blockIdx.x;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const double *kap = kappa + bk.off;
    double *Tb = Tap + (long)b * hmax * hmax;
    for (int e = threadIdx.x; e < hmax * hmax; e += blockDim.x) {
        const int b1 = e / hmax, a2 = e % hmax;
        double v = 0.0;
        if (b1 < h && a2 < h) {
            const double c1 = pc[k.state[bk.first + b1]], c2 = pc[k.state[bk.first + a2]];
            const double kv = kap[(h + a2) * n + b1];
            if (c1 >= kHfbNumCut && c1 * c2 * fabs(kv) >= kHfbNumCut) v = c1 * c2 * kv;
        }
        Tb[e] = v;
    }
    for (int i = threadIdx.x; i < hmax; i += blockDim.x)
        list[b * hmax + i] = k.state[bk.first + (i < h ? i : 0)];
    if (threadIdx.x == 0) hlist[b] = h;
}

// T(lambda) of every block (dense, leading dimension ld): the HFB matrix of
// hfb_block_eval (mocca_hfb.hip) without the warm-start products.
__global__ void k_hl_build_T(HfbK k, const double *__restrict__ E,
                             const double *__restrict__ delta, const double *__restrict__ lam,
                             int ld, double *__restrict__ Tm) {
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int i = e / n, j = e % n;
    const double *dl = delta + bk.off;
    const double l = lam[bk.it];
    double v = 0.0;
    if (i < h && j < h) v = i == j ? E[k.state[bk.first + i]] - l : 0.0;
    else if (i >= h && j >= h) v = i == j ? -(E[k.state[bk.first + i - h]] - l) : 0.0;
    else if (i < h) v = dl[i * n + h + (j - h)];
    else v = dl[j * n + h + (i - h)];
    Tm[(long)b * ld * ld + (long)i * ld + j] = v;
}

// The next evaluation's start matrix.  In the current basis it is
// A = Afin - dlambda C (Afin the last Jacobi's matrix, diagonal to its
// convergence threshold; T changes by -Sigma_z dlambda between evaluations
// of one Fermi solve); when dlambda is so large that its
// off-diagonal part exceeds that of T(lambda') itself (the gaps), the
// evaluation starts from T(lambda') in the original basis instead (cold[b]
// = 1: the rotations are then applied to the identity).
__global__ void k_hl_shift_A(HfbK k, const double *__restrict__ Af,
                             const double *__restrict__ C, const double *__restrict__ nst,
                             const int *__restrict__ active, int ld, double *__restrict__ A,
                             const double *__restrict__ E, const double *__restrict__ delta,
                             const double *__restrict__ off, int *__restrict__ cold) {
    const int b = blockIdx.y;
    if (!active[b]) return;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const double dl = nst[bk.it * kHlNewton + 6];
    const bool c0 = dl * dl * off[2 * b + 1] > off[2 * b];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) cold[b] = c0 ? 1 : 0;
    if (e >= n * n) return;
    const int i = e / n, j = e % n;
    const long o = (long)b * ld * ld + (long)i * ld + j;
    if (!c0) {
        A[o] = Af[o] - dl * C[o];
        return;
    }
    const double *dd = delta + bk.off;
    const double l = nst[bk.it * kHlNewton + 0];   // the next evaluation's lambda
    double v;
    if (i < h && j < h) v = i == j ? E[k.state[bk.first + i]] - l : 0.0;
    else if (i >= h && j >= h) v = i == j ? -(E[k.state[bk.first + i - h]] - l) : 0.0;
    else if (i < h) v = dd[i * n + h + (j - h)];
    else v = dd[j * n + h + (i - h)];
    A[o] = v;
}

constexpr int kNumThreads = 1024;
constexpr int kHfbLargeMaxH = 128;   // states per (isospin, parity) block

// Particle number of each block and its derivative: ranks of the
// eigenvalues (descending, ties by index; the h largest are the
// quasiparticles), N_b = 2 sum_{k+} sum_a V[h+a][k]^2,
// dN_b = 2 sum_{k+} sum_{l-} C_kl^2 / (w_k - w_l).
__global__ __launch_bounds__(kNumThreads) void k_hl_number(
    HfbK k, const double *__restrict__ V, const double *__restrict__ C,
    const double *__restrict__ w, int ld, const int *__restrict__ active,
    int *__restrict__ rank, double *__restrict__ num, const double *__restrict__ delta,
    double *__restrict__ off) {
    __shared__ double sh[kNumThreads / 64];
    __shared__ int rk[2 * kHfbLargeMaxH];
    __shared__ double ws[2 * kHfbLargeMaxH];
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (!active[b]) return;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const double *wb = w + (long)b * ld;
    const double *Vb = V + (long)b * ld * ld, *Cb = C + (long)b * ld * ld;
    int *rb = rank + (long)b * ld;
    for (int i = tid; i < n; i += nt) ws[i] = wb[i];
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        const double wi = ws[i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double wj = ws[j];
            if (wj > wi || (wj == wi && j < i)) ++r;
        }
        rk[i] = r;
        rb[i] = r;
    }
    __syncthreads();
    // particle number, its derivative and the off-diagonal square sum of C
    // in one pass over the n x n elements
    double nn = 0, dn = 0, oc = 0;
    for (int e = tid; e < n * n; e += nt) {
        const int a = e / n, q = e % n;
        const double caq = Cb[(long)a * ld + q];
        if (a != q) oc += caq * caq;
        if (rk[q] >= h) continue;
        if (a < h) {
            const double v = Vb[(long)(h + a) * ld + q];
            nn += v * v;
        }
        const int l = a;   // C[q][l], q positive, l negative
        if (rk[l] < h) continue;
        const double c = Cb[(long)q * ld + l];
        dn += c * c / (ws[q] - ws[l]);
    }
    // off-diagonal square sum of T(lambda') in the original basis (the gaps,
    // both triangles; independent of lambda), against dlambda^2 times that
    // of C for Afin - dlambda C in the current basis
    double od = 0;
    const double *dl = delta + bk.off;
    for (int e = tid; e < h * h; e += nt) {
        const double d = dl[(e / h) * n + h + e % h];
        od += 2.0 * d * d;
    }
    nn = hl_block_sum(nn, sh);
    dn = hl_block_sum(dn, sh);
    od = hl_block_sum(od, sh);
    oc = hl_block_sum(oc, sh);
    if (tid == 0) {
        num[2 * b] = 2.0 * nn;
        num[2 * b + 1] = 2.0 * dn;
        off[2 * b] = od;
        off[2 * b + 1] = oc;
    }
}

// Newton step on the Fermi energy of each isospin (thread it).  Record nst
// [it][kHlNewton] = lambda, lo, hi, has_lo, has_hi, converged, delta, iter.
__global__ void k_hl_newton(HfbK k, const double *__restrict__ num, double *__restrict__ nst,
                            int *__restrict__ active, double *__restrict__ fermi,
                            int *__restrict__ status) {
    const int it = threadIdx.x;
    if (it >= 2) return;
    double *s = nst + it * kHlNewton;
    if (s[5] != 0.0) return;
    double nn = 0, dn = 0;
    bool any = false;
    for (int b = 0; b < k.nblk; ++b)
        if (k.blk[b].it == it) {
            nn += num[2 * b];
            dn += num[2 * b + 1];
            any = true;
        }
    if (!any) { s[5] = 1.0; return; }
    const double f = nn - k.particles[it], lam = s[0];
    if (f != f || dn != dn) {
        atomicOr(status, 2);
        s[5] = 1.0;
        return;
    }
    if (fabs(f) < k.prec) {
        s[5] = 1.0;
        fermi[it] = lam;
        for (int b = 0; b < k.nblk; ++b)
            if (k.blk[b].it == it) active[b] = 0;
        return;
    }
    if (f < 0) { s[1] = lam; s[3] = 1.0; } else { s[2] = lam; s[4] = 1.0; }
    double x = dn > 1e-14 ? lam - f / dn : lam + (f < 0 ? 1.0 : -1.0);
    const bool blo = s[3] != 0.0, bhi = s[4] != 0.0;
    if (blo && bhi) {
        if (!(x > s[1] && x < s[2])) x = 0.5 * (s[1] + s[2]);
    } else if (f < 0) {
        if (!(x > lam)) x = lam + 1.0;
        x = fmin(x, lam + 5.0);
    } else {
        if (!(x < lam)) x = lam - 1.0;
        x = fmax(x, lam - 5.0);
    }
    s[6] = x - lam;
    s[0] = x;
    s[7] += 1.0;
    fermi[it] = x;
    if (s[7] >= k.iters) {
        // HFBIter evaluations without reaching FermiPrec: keep the last
        // evaluation's rho and kappa, as the reference does
        s[5] = 1.0;
        s[0] = lam;
        fermi[it] = lam;
        atomicOr(status, 1);
        for (int b = 0; b < k.nblk; ++b)
            if (k.blk[b].it == it) active[b] = 0;
    }
}

// The reference's Broyden (secant) iteration on the Fermi energy of each
// isospin (thread it), HFBFindFermiEnergyBroyden (HFB.f90:1849-2161) without
// Lipkin-Nogami, as an event-driven machine: called after each evaluation
// with the particle number there, it sets the next evaluation point.  The
// evaluation points, their order and the stopping rule are the reference's:
// F, F + 0.01, then F += 0.5 * (-N / J) with the Broyden update of J after
// iteration 1 and the 0.99 ceiling of 1/J, until |N - Z| < FermiPrec or
// HFBIter iterations; the last evaluation's rho and kappa are kept.
// Record nst[it][kHlNewton]: 0 lambda of the current evaluation, 1 FH,
// 2 J, 3 upd, 4 norm, 5 done, 6 next lambda - lambda, 7 iteration,
// 8 N - Z at F, 9 phase (0: F evaluated, 1: FH evaluated, 2: loop),
// 10 converged, 11 F.
__global__ void k_hl_broyden(HfbK k, const double *__restrict__ num, double *__restrict__ nst,
                             int *__restrict__ active, double *__restrict__ fermi,
                             int *__restrict__ status, const double *__restrict__ skip) {
    const int it = threadIdx.x;
    if (it >= 2 || (skip && *skip != 0.0)) return;
    double *s = nst + it * kHlNewton;
    if (s[5] != 0.0) return;
    double nn = 0;
    bool any = false;
    for (int b = 0; b < k.nblk; ++b)
        if (k.blk[b].it == it) {
            nn += num[2 * b];
            any = true;
        }
    auto finish = [&](bool unconverged) {
        s[5] = 1.0;
        fermi[it] = s[11];
        if (unconverged) atomicOr(status, 1);
        for (int b = 0; b < k.nblk; ++b)
            if (k.blk[b].it == it) active[b] = 0;
    };
    if (!any) { s[5] = 1.0; return; }
    const double Nv = nn - k.particles[it];
    const double lam = s[0];
    double F = s[11], FH = s[1], J = s[2], upd = s[3], norm = s[4];
    int iter = (int)s[7];
    bool conv = s[10] != 0.0;
    const int phase = (int)s[9];
    // the part of a loop iteration before its evaluation
    auto pre_eval = [&]() {
        const double det = J;
        const double inv = fabs(det) < 0.001 ? copysign(0.99, det) : 1.0 / det;
        if (!conv) {
            upd = -inv * s[8];
            norm = upd * upd;
        }
        upd = 0.5 * upd;
        if (!conv) {
            FH = F;
            F = F + upd;
        }
    };
    double next;
    if (phase == 0) {
        s[8] = Nv;
        conv = fabs(Nv) < k.prec;
        next = FH;                      // F + 0.01
        s[9] = 1.0;
    } else if (phase == 1) {
        J = (s[8] - Nv) / (F - FH);
        iter = 1;
        pre_eval();
        next = F;
        s[9] = 2.0;
    } else {
        s[8] = Nv;
        if (fabs(Nv) < k.prec) conv = true;
        if (conv) {
            s[10] = 1.0;
            finish(false);
            return;
        }
        if (iter > 1) J = J + Nv * upd / norm;
        if (Nv == Nv + 1.0 || Nv != Nv) {
            atomicOr(status, 2);
            finish(false);
            return;
        }
        if (iter >= k.iters) {
            finish(true);
            return;
        }
        ++iter;
        pre_eval();
        next = F;
    }
    s[1] = FH;
    s[2] = J;
    s[3] = upd;
    s[4] = norm;
    s[7] = iter;
    s[10] = conv ? 1.0 : 0.0;
    s[11] = F;
    s[6] = next - lam;
    s[0] = next;
}

// ---------------------------------------------------------------------------
// Particle numbers without an eigen-decomposition.  The quasiparticles that
// HFBColumns selects are the eigenvectors of T(lambda) with positive
// eigenvalue (the spectrum of the signature-reduced T comes in +-pairs), so
// rho, kappa and N(lambda) are blocks of the spectral projector
// P+ = (1 + sign T) / 2:
//   rho = [[R, 0], [0, R]],  R = P+[h:, h:];
//   kappa = [[0, -Q^T], [Q, 0]],  Q = P+[h:, :h];   N = 2 tr R.
// sign T is the limit of the scaled Newton-Schulz iteration
//   X0 = T / s,  X <- (a/2) X (3 - a^2 X^2),
// s the largest absolute row sum of T (a bound of its spectral radius), a
// the scaling for the eigenvalue interval [l, 1] of X (l from a bound of the
// smallest |eigenvalue| relative to s, raised along the iteration), two
// symmetric matrix products per step on the FP64 matrix cores.  Every X is
// a polynomial in T, so the products commute and stay symmetric: each is
// formed for the upper 32 x 32 tiles and mirrored.
constexpr int kSignTile = 16;
constexpr int kSignSteps = 16;           // k steps of 4 per wave (blocks up to 256 rows)
// residual slots per block: four waves of each of the at most 16 x 17 / 2
// upper tiles
constexpr int kSignParts = 4 * 16 * 17 / 2;
constexpr double kSignResidual = 1e-8;   // max |X^2 - 1| before the last step

__device__ __forceinline__ double hl_wave_max(double a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    return a;
}

// Off-diagonal absolute row sums of T (the gaps: independent of lambda),
// once per Fermi solve: rs[b][i] = sum_j |T_ij|, j != i (one wave per row).
__global__ __launch_bounds__(256) void k_hs_rowsum(HfbK k, const double *__restrict__ delta,
                                                   int ld, double *__restrict__ rs) {
    const int b = blockIdx.y, lane = threadIdx.x % 64;
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    if (i >= n) return;
    const double *dl = delta + bk.off;
    double r = 0.0;
    for (int j = lane; j < h; j += 64) r += fabs(i < h ? dl[i * n + h + j] : dl[j * n + i]);
    r = hl_wave_sum(r);
    if (lane == 0) rs[(long)b * ld + i] = r;
}

// X = T(lambda) / s for the active blocks (lambda of the evaluation from the
// record nst[it][0]), s = the largest absolute row sum of T (a bound of its
// spectral radius; every workgroup of a block forms it from rs and the
// diagonal); the redo flag nst[2][0] cleared.  T as k_hl_build_T forms it.
__global__ __launch_bounds__(256) void k_hs_build(HfbK k, const double *__restrict__ E,
                                                  const double *__restrict__ delta,
                                                  double *__restrict__ nst,
                                                  const double *__restrict__ rs,
                                                  const int *__restrict__ active, int ld,
                                                  double *__restrict__ X) {
    __shared__ double sinv;
    const int b = blockIdx.y, tid = threadIdx.x;
    if (b == 0 && blockIdx.x == 0 && tid == 0) nst[2 * kHlNewton] = 0.0;   // redo flag
    if (!active[b]) return;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const double l = nst[bk.it * kHlNewton];   // lambda of this evaluation
    if (tid < 64) {
        double mx = 0.0;
        for (int i = tid; i < n; i += 64) {
            const double d = E[k.state[bk.first + (i < h ? i : i - h)]] - l;
            mx = fmax(mx, fabs(d) + rs[(long)b * ld + i]);
        }
        mx = hl_wave_max(mx);
        if (tid == 0) sinv = mx > 0.0 ? 1.0 / mx : 1.0;
    }
    __syncthreads();
    const double inv = sinv;
    const int e = blockIdx.x * blockDim.x + tid;
    if (e >= n * n) return;
    const int i = e / n, j = e % n;
    const double *dl = delta + bk.off;
    double v = 0.0;
    if (i < h && j < h) v = i == j ? E[k.state[bk.first + i]] - l : 0.0;
    else if (i >= h && j >= h) v = i == j ? -(E[k.state[bk.first + i - h]] - l) : 0.0;
    else if (i < h) v = dl[i * n + h + (j - h)];
    else v = dl[j * n + h + (i - h)];
    X[(long)b * ld * ld + (long)i * ld + j] = v * inv;
}

// Matrix sizes of the blocks, passed by value so that a kernel's first
// loads do not wait for a table read.
struct SignDims {
    int n[4];
};

// C = c1 D + c2 A B for symmetric A, B whose product is symmetric; one
// workgroup of four waves per upper 16 x 16 tile (I <= J) of a block,
// mirrored into (J, I).  The waves split the k range in four (every operand
// of a wave loaded before its first product) and add their partial tiles
// through LDS.  The block's active flag is read while the operands load
// (an inactive block's tiles are read and dropped).  With RESID each wave's
// largest |C - 1| goes into its slot of sig[b][.] (reduced by k_hs_number).
template <bool RESID>
__global__ __launch_bounds__(256) void k_hs_gemm(const double *__restrict__ A,
                                                 const double *__restrict__ B,
                                                 const double *__restrict__ D, double c1,
                                                 double c2, double *__restrict__ C, int ld,
                                                 long mstride, SignDims dims,
                                                 const int *__restrict__ active,
                                                 double *__restrict__ sig) {
    __shared__ double part[4][4][64];   // [wave][register][lane]
    const int b = blockIdx.y, tid = threadIdx.x, wv = tid / 64, lane = tid % 64;
    const int n = dims.n[b], nt = (n + kSignTile - 1) / kSignTile;
    int t = blockIdx.x, I = 0;
    while (I < nt && t >= nt - I) {
        t -= nt - I;
        ++I;
    }
    if (I >= nt) return;
    const int J = I + t;
    const int i0 = I * kSignTile, j0 = J * kSignTile;
    const double *Ab = A + b * mstride, *Bb = B + b * mstride;
    const int lr = lane % 16, lk = lane / 16;
    hl_d4 acc = hl_d4{0, 0, 0, 0};
    // this wave's quarter of k, in steps of 4 (at most kSignSteps: n <= 256)
    const int kq = ((n + 15) / 16) * 4;
    const int kb = wv * kq, ke = min(n, kb + kq);
    // A is symmetric: its operand A[i][k] is read as A[k][i] (rows, coalesced)
    const bool ia = i0 + lr < n, jb = j0 + lr < n;
    double av[kSignSteps], bv[kSignSteps];
#pragma unroll
    for (int q = 0; q < kSignSteps; ++q) {
        const int kk = kb + 4 * q + lk;
        const bool kin = kk < ke;
        const long row = (long)(kin ? kk : 0) * ld;
        av[q] = kin && ia ? Ab[row + i0 + lr] : 0.0;
        bv[q] = kin && jb ? Bb[row + j0 + lr] : 0.0;
    }
    if (!active[b]) return;
#pragma unroll
    for (int q = 0; q < kSignSteps; ++q) {
        if (kb + 4 * q >= ke) break;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) part[wv][r][lane] = acc[r];
    __syncthreads();
    // wave wv finishes register wv of the tile
    const int r = wv;
    const double sum = part[0][r][lane] + part[1][r][lane] + part[2][r][lane] + part[3][r][lane];
    // accumulator lane L, register r: row L/16 + 4 r, column L%16
    const int row = i0 + lk + 4 * r, col = j0 + lr;
    double res = 0.0;
    if (row < n && col < n && !(I == J && row > col)) {
        double v = c2 * sum;
        if (D) v += c1 * D[b * mstride + (long)row * ld + col];
        double *Cb = C + b * mstride;
        Cb[(long)row * ld + col] = v;
        Cb[(long)col * ld + row] = v;
        if (RESID) res = fabs(v - (row == col ? 1.0 : 0.0));
    }
    if (RESID) {
        res = hl_wave_max(res);
        if (lane == 0) sig[(long)b * kSignParts + 4 * blockIdx.x + wv] = res;
    }
}

// N of each active block from the converged X: num[2b] = 2 tr R =
// sum_{i >= h} (1 + X_ii) (one wave per block); *redo = 1 when the
// block's residual max |X^2 - 1| before the last step (the slots sig[b][.])
// is above kSignResidual (the evaluation then continues before the Fermi
// step).
__global__ __launch_bounds__(64) void k_hs_number(HfbK k, const double *__restrict__ X, int ld,
                                                  const int *__restrict__ active,
                                                  const double *__restrict__ sig,
                                                  double *__restrict__ num,
                                                  double *__restrict__ redo) {
    const int b = blockIdx.x, lane = threadIdx.x;
    if (!active[b]) return;
    const int h = k.blk[b].h, n = 2 * h;
    const int nt = (n + kSignTile - 1) / kSignTile, nparts = 4 * nt * (nt + 1) / 2;
    double res = 0.0;
    for (int q = lane; q < nparts; q += 64) res = fmax(res, sig[(long)b * kSignParts + q]);
    res = hl_wave_max(res);
    if (!(res <= kSignResidual)) {
        if (lane == 0) *redo = 1.0;
        return;
    }
    const double *Xb = X + (long)b * ld * ld;
    double s = 0.0;
    for (int i = h + lane; i < n; i += 64) s += 1.0 + Xb[(long)i * ld + i];
    s = hl_wave_sum(s);
    if (lane == 0) {
        num[2 * b] = s;
        num[2 * b + 1] = 0.0;
    }
}

__global__ void k_hs_clear(double *__restrict__ flag) { *flag = 0.0; }


__global__ void k_hs_activate(int *__restrict__ active, int nblk) {
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) active[b] = 1;
}


// rho and kappa from X (the layout of k_hl_rho_kappa).
__global__ void k_hs_rho_kappa(HfbK k, const double *__restrict__ X, int ld,
                               double *__restrict__ rho, double *__restrict__ kappa) {
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int i = e / n, j = e % n;
    const double *Xb = X + (long)b * ld * ld;
    double v = 0.0, kv = 0.0;
    if (i < h && j < h)
        v = 0.5 * ((i == j ? 1.0 : 0.0) + Xb[(long)(h + i) * ld + h + j]);
    else if (i >= h && j >= h)
        v = 0.5 * ((i == j ? 1.0 : 0.0) + Xb[(long)i * ld + j]);
    else if (i >= h)
        kv = 0.5 * Xb[(long)i * ld + j];
    else
        kv = -0.5 * Xb[(long)j * ld + i];
    rho[bk.off + e] = v;
    kappa[bk.off + e] = kv;
}

// The h largest eigenvalues of each block, in descending order, as
// k_hl_rho_kappa stores them.
__global__ void k_hl_qpe(HfbK k, const double *__restrict__ w, int ld,
                         double *__restrict__ qpe) {
    const int b = blockIdx.x;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const double *wb = w + (long)b * ld;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double wi = wb[i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double wj = wb[j];
            if (wj > wi || (wj == wi && j < i)) ++r;
        }
        if (r < h) qpe[bk.first + r] = wi;
    }
}

// rho and kappa of the blocks from the eigenvectors of the h largest
// eigenvalues (constructRhoHFB, ConstructKappaHFB :3128-3360), in the layout
// of hfb_block_eval; the quasiparticle energies.
__global__ void k_hl_rho_kappa(HfbK k, const double *__restrict__ V,
                               const double *__restrict__ w, const int *__restrict__ rank,
                               int ld, double *__restrict__ rho, double *__restrict__ kappa,
                               double *__restrict__ qpe) {
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const double *Vb = V + (long)b * ld * ld, *wb = w + (long)b * ld;
    const int *rb = rank + (long)b * ld;
    if (e < n && rb[e] < h) qpe[bk.first + rb[e]] = wb[e];
    if (e >= n * n) return;
    const int i = e / n, j = e % n;
    double v = 0.0, kv = 0.0;
    if (i < h && j < h) {
        for (int q = 0; q < n; ++q)
            if (rb[q] < h) v += Vb[(long)(h + i) * ld + q] * Vb[(long)(h + j) * ld + q];
    } else if (i >= h && j >= h) {
        for (int q = 0; q < n; ++q)
            if (rb[q] < h) v += Vb[(long)i * ld + q] * Vb[(long)j * ld + q];
    } else if (i >= h) {
        for (int q = 0; q < n; ++q)
            if (rb[q] < h) kv += Vb[(long)i * ld + q] * Vb[(long)j * ld + q];
    } else {
        for (int q = 0; q < n; ++q)
            if (rb[q] < h) kv -= Vb[(long)j * ld + q] * Vb[(long)i * ld + q];
    }
    rho[bk.off + e] = v;
    kappa[bk.off + e] = kv;
}

// rho_h (the signature +1 part of rho) of every block into A (h x h,
// leading dimension ld) for the canonical-basis diagonalisation.
__global__ void k_hl_can_prep(HfbK k, const double *__restrict__ rho, int ld,
                              double *__restrict__ A, int *__restrict__ hlist) {
    const int b = blockIdx.y;
    const HfbBlockK &bk = k.blk[b];
    const int h = bk.h, n = 2 * h;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0 && e < k.nblk) hlist[e] = k.blk[e].h;
    if (e >= h * h) return;
    const int i = e / h, j = e % h;
    A[(long)b * ld * ld + (long)i * ld + j] = rho[bk.off + i * n + j];
}

// Canonical transforms and occupations from the eigenvectors of rho_h
// (k_hfb_canonical's rules): columns in ascending occupation, entries below
// 1e-11 set to zero, occupations 2 x eigenvalue clamped to [0, 2]; the HF
// states' occupations 2 rho_ii; the transforms also into the apply matrices
// Tap (leading dimension hmax) with the block lists.
__global__ __launch_bounds__(256) void k_hl_can_finish(
    HfbK k, const double *__restrict__ V, const double *__restrict__ w, int ld,
    const double *__restrict__ rho, int hmax, double *__restrict__ tcan,
    double *__restrict__ can_occ, double *__restrict__ hf_occ, double *__restrict__ Tap,
    int *__restrict__ list, int *__restrict__ status) {
    __shared__ int rank[256];
    const HfbBlockK &bk = k.blk[blockIdx.x];
    const int h = bk.h, n = 2 * h, tid = threadIdx.x, nt = blockDim.x;
    const double *Vb = V + (long)blockIdx.x * ld * ld, *wb = w + (long)blockIdx.x * ld;
    for (int i = tid; i < h; i += nt) {
        const double wi = wb[i];
        int r = 0;
        for (int j = 0; j < h; ++j) {
            const double wj = wb[j];
            if (wj < wi || (wj == wi && j < i)) ++r;
        }
        rank[i] = r;
    }
    __syncthreads();
    double *tb = tcan + bk.toff;
    double *Tb = Tap + (long)blockIdx.x * hmax * hmax;
    for (int e = tid; e < hmax * hmax; e += nt) Tb[e] = 0.0;
    __syncthreads();
    for (int e = tid; e < h * h; e += nt) {
        const int j = e / h, c = e % h;
        const double v = Vb[(long)j * ld + c];
        const double t = fabs(v) < 1e-11 ? 0.0 : v;
        tb[j * h + rank[c]] = t;
        Tb[j * hmax + rank[c]] = t;
    }
    for (int c = tid; c < h; c += nt) {
        double x = wb[c];
        if (x - 1.0 > 1e-5) atomicOr(status, 4);
        if (x < 0.0) x = 0.0;
        if (x > 1.0) x = 1.0;
        can_occ[k.state[bk.first + rank[c]]] = 2.0 * x;
        hf_occ[k.state[bk.first + c]] = 2.0 * rho[bk.off + c * n + c];
    }
    for (int i = tid; i < hmax; i += nt)
        list[blockIdx.x * hmax + i] = k.state[bk.first + (i < h ? i : 0)];
}

__global__ void k_hl_newton_init(const double *__restrict__ fermi, double *__restrict__ nst,
                                 int nblk, int *__restrict__ active) {
    const int t = threadIdx.x;
    if (t < 2) {
        double *s = nst + t * kHlNewton;
        for (int q = 0; q < kHlNewton; ++q) s[q] = 0.0;
        s[0] = fermi[t];          // first evaluation at the last Fermi energy
        s[11] = fermi[t];         // F of the Broyden iteration
        s[1] = fermi[t] + 0.01;   // FH: FermiHistory = Fermi + 0.01
    }
    if (t < nblk) active[t] = 1;
}

// ===========================================================================
// Host side
namespace detail {

namespace {

template <typename T>
void launch_pair_gram(const HfbK &k, const T *psi, long N, const double *field, double *part,
                      int nsplit, int nt, int nblk, hipStream_t st) {
    auto grid = [&](int nj) { return dim3(nsplit, nblk, (nt + nj - 1) / nj); };
#define MOCCA_PG(NI, NJ)                                                                    \
    case NI:                                                                                \
        if (N % 4 == 0)                                                                     \
            hipLaunchKernelGGL((k_hl_pair_gram<T, NI, NJ, true>), grid(NJ), dim3(64), 0, st, k, \
                               psi, N, field, part, nsplit);                                \
        else                                                                                \
            hipLaunchKernelGGL((k_hl_pair_gram<T, NI, NJ, false>), grid(NJ), dim3(64), 0, st, \
                               k, psi, N, field, part, nsplit);                             \
        break;
    switch (nt) {
        MOCCA_PG(1, 1) MOCCA_PG(2, 2) MOCCA_PG(3, 3) MOCCA_PG(4, 4)
        MOCCA_PG(5, 5) MOCCA_PG(6, 4) MOCCA_PG(7, 4) MOCCA_PG(8, 3)
        default: break;
    }
#undef MOCCA_PG
}

int pair_gram_groups(int nt) {
    static const int nj[9] = {1, 1, 2, 3, 4, 5, 4, 4, 3};
    return (nt + nj[nt] - 1) / nj[nt];
}

}  // namespace

bool hfb_large_fits(int hmax) { return hmax >= 1 && (hmax + 15) / 16 <= 8; }

size_t hfb_large_jacobi_lds(int n) {
    return sizeof(double) * ((size_t)jac_sh_offset(n) + 16 + (size_t)n * (n + 1) / 2);
}

int hfb_large_setup(mocca_solve *h) {
    Hfb &p = h->hfb;
    const HfbK &k = p.k;
    int nmax = 0;
    for (int b = 0; b < k.nblk; ++b) nmax = std::max(nmax, 2 * k.blk[b].h);
    p.nmax = nmax;
    const int ld = nmax, hmax = p.hmax;
    const size_t mat = (size_t)k.nblk * ld * ld;
    const int m = nmax + (nmax & 1);
    p.lmax = (long)kJacMaxSweeps * (m - 1) * (m / 2);
    const int ntile = (hmax + 15) / 16;
    p.lsplit = std::max(1, kGramWaves / std::max(1, k.nblk * pair_gram_groups(ntile)));
    p.lsplit = (int)std::min<long>(p.lsplit, std::max<long>(1, ((long)h->N + 3) / 4 / 8));
    int rc;
    if ((rc = dalloc(h, &p.lT, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lV, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lA, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lW, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lC, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lVc, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lw, (size_t)k.nblk * ld * sizeof(double))) ||
        (rc = dalloc(h, &p.lwc, (size_t)k.nblk * ld * sizeof(double))) ||
        (rc = dalloc(h, &p.lrank, (size_t)k.nblk * ld * sizeof(int))) ||
        (rc = dalloc(h, &p.llog, (size_t)k.nblk * p.lmax * sizeof(double2))) ||
        (rc = dalloc(h, &p.llogij, (size_t)k.nblk * p.lmax * sizeof(int))) ||
        (rc = dalloc(h, &p.lAf, mat * sizeof(double))) ||
        (rc = dalloc(h, &p.lrounds, 2 * (size_t)k.nblk * sizeof(int))) ||
        (rc = dalloc(h, &p.lnum, 2 * (size_t)k.nblk * sizeof(double))) ||
        (rc = dalloc(h, &p.loff, 2 * (size_t)k.nblk * sizeof(double))) ||
        (rc = dalloc(h, &p.lcold, (size_t)k.nblk * sizeof(int))) ||
        (rc = dalloc(h, &p.lnst, (2 * kHlNewton + 2) * sizeof(double))) ||
        (rc = dalloc(h, &p.lactive, (size_t)k.nblk * sizeof(int))) ||
        (rc = dalloc(h, &p.lnb, 2 * (size_t)k.nblk * sizeof(int))) ||
        (rc = dalloc(h, &p.lhb, (size_t)k.nblk * sizeof(int))) ||
        (rc = dalloc(h, &p.llist, (size_t)k.nblk * hmax * sizeof(int))) ||
        (rc = dalloc(h, &p.lTap, (size_t)k.nblk * hmax * hmax * sizeof(double))) ||
        (rc = dalloc(h, &p.lgpart, sizeof(double) *
                                       std::max((size_t)k.nblk * p.lsplit * ntile * ntile * 256,
                                                (size_t)k.nblk * 8 * h->N))) ||
        (rc = dalloc(h, &p.lsig, (size_t)k.nblk * kSignParts * sizeof(double))))
        return rc;
    std::vector<int> nb(2 * k.nblk), hb(k.nblk);
    for (int b = 0; b < k.nblk; ++b) {
        nb[b] = 2 * k.blk[b].h;
        nb[k.nblk + b] = k.blk[b].h;
        hb[b] = k.blk[b].h;
    }
    if (hipHostMalloc((void **)&p.lhost, (2 * kHlNewton + 2) * sizeof(double)) != hipSuccess) {
        p.lhost = nullptr;
        return fail(h, MOCCA_ERR_HIP, "hipHostMalloc of the Fermi-solve records failed");
    }
    HIPCHK(h, hipMemcpy(p.lnb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(p.lhb, hb.data(), hb.size() * sizeof(int), hipMemcpyHostToDevice));
    p.lwarm = false;
    // the in-LDS Jacobi wants the whole device limit
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
        p.lds_limit = (size_t)lim;
    return 0;
}

int hfb_large_free(mocca_solve *h) {
    Hfb &p = h->hfb;
    void *ptrs[] = {p.lT, p.lV, p.lA, p.lW, p.lC, p.lVc, p.lw, p.lwc, p.lrank, p.llog,
                    p.lrounds, p.lnum, p.lnst, p.lactive, p.lnb, p.lhb, p.llist, p.lTap,
                    p.lgpart, p.loff, p.lcold, p.llogij, p.lAf, p.lsig};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (p.lhost) (void)hipHostFree(p.lhost);
    p.lhost = nullptr;
    return 0;
}

namespace {

// Jacobi + rotation apply on nblk matrices of sizes nb[] (A, leading
// dimension ld) with the basis V (identity first when ident).
int jacobi_run(mocca_solve *h, const double *A, double *V, double *w, const int *nb, int nmax,
               const int *active, int *rounds, bool ident, bool odd, const int *cold,
               double *Afin) {
    Hfb &p = h->hfb;
    const int nblk = p.k.nblk, ld = p.nmax;
    const long ms = (long)ld * ld;
    int rc;
    if (!odd && nmax <= kJacLdsMaxN && slots_jacobi_lds(nmax) <= p.lds_limit) {
        const size_t shm = slots_jacobi_lds(nmax);
        const size_t top = std::min(slots_jacobi_lds(kJacLdsMaxN), p.lds_limit);
        rc = run(h, "hfb_jacobi", [&] {
            require_lds((const void *)k_lj_jacobi_slots, top, "k_lj_jacobi_slots");
            hipLaunchKernelGGL(k_lj_jacobi_slots, dim3(nblk), dim3(kSlotThreads), shm, h->st, A,
                               ld, ms, nb, active, w, Afin, p.llog, p.llogij, p.lmax, rounds,
                               kJacMaxSweeps, p.status);
        });
    } else if (nmax <= kJacLdsMaxN && hfb_large_jacobi_lds(nmax) <= p.lds_limit) {
        const size_t shm = hfb_large_jacobi_lds(nmax);
        // (the limit for the largest in-LDS matrix, once)
        const size_t top = std::min(hfb_large_jacobi_lds(kJacLdsMaxN), p.lds_limit);
        rc = run(h, "hfb_jacobi", [&] {
            auto kern = odd ? k_lj_jacobi<false, true> : k_lj_jacobi<false, false>;
            require_lds((const void *)kern, top, "k_lj_jacobi");
            hipLaunchKernelGGL(kern, dim3(nblk), dim3(kJacThreads), shm, h->st, A, ld, ms, nb,
                               active, nullptr, w, Afin, p.llog, p.llogij, p.lmax, rounds,
                               kJacMaxSweeps, p.status);
        });
    } else {
        const size_t shm = sizeof(double) * ((size_t)jac_sh_offset(nmax) + 16);
        rc = run(h, "hfb_jacobi", [&] {
            auto kern = odd ? k_lj_jacobi<true, true> : k_lj_jacobi<true, false>;
            hipLaunchKernelGGL(kern, dim3(nblk), dim3(kJacThreads), shm, h->st, A, ld, ms, nb,
                               active, p.lW, w, Afin, p.llog, p.llogij, p.lmax, rounds,
                               kJacMaxSweeps, p.status);
        });
    }
    if (rc) return rc;
    return run(h, "hfb_jacobi_apply", [&] {
        const size_t shm = apply_lds(nmax);
        require_lds((const void *)k_lj_apply, shm, "k_lj_apply");
        hipLaunchKernelGGL(k_lj_apply, dim3((nmax + kApplyRows - 1) / kApplyRows, nblk),
                           dim3(256), shm, h->st, V, ld, ms, nb, active, p.llog, p.llogij,
                           p.lmax, rounds, ident ? 1 : 0, cold);
    });
}

}  // namespace

// The canonical basis from rho (in lT: rho_h of every block): eigenvectors
// by the Jacobi kernel, the canonical states by the dense matrix-core apply,
// then the rho/kappa mixing.
template <typename T>
int hfb_large_canonical(mocca_solve *h) {
    Hfb &p = h->hfb;
    const HfbK &k = p.k;
    const long N = h->N;
    const int nblk = k.nblk, ld = p.nmax, hmax = p.hmax;
    const int mpad = (hmax + 15) / 16 * 16;
    // (the block sizes h are the second half of lnb)
    bool odd = false;
    for (int b = 0; b < nblk; ++b) odd = odd || (k.blk[b].h & 1);
    // after the first step the Jacobi starts in the last step's canonical
    // basis: A = Vc^T rho_h Vc, the rotations applied to Vc
    const bool warm = p.can_warm;
    if (warm) {
        const long ms = (long)ld * ld;
        const dim3 gm((hmax + 15) / 16, (hmax + 15) / 16, nblk);
        if (int rc = run(h, "hfb_canonical", [&] {
                hipLaunchKernelGGL((k_lj_gemm<false, false>), gm, dim3(256), 0, h->st, p.lT,
                                   p.lVc, p.lW, ld, ms, p.lnb + nblk, nullptr, nullptr);
                hipLaunchKernelGGL((k_lj_gemm<true, false>), gm, dim3(256), 0, h->st, p.lVc,
                                   p.lW, p.lC, ld, ms, p.lnb + nblk, nullptr, nullptr);
            }))
            return rc;
    }
    if (int rc = jacobi_run(h, warm ? p.lC : p.lT, p.lVc, p.lwc, p.lnb + nblk, hmax, nullptr,
                            p.lrounds + nblk, !warm, odd, nullptr, nullptr))
        return rc;
    p.can_warm = true;
    return run(h, "hfb_canonical", [&] {
        hipLaunchKernelGGL(k_hl_can_finish, dim3(nblk), dim3(256), 0, h->st, p.k, p.lVc, p.lwc,
                           ld, p.rho, hmax, p.tcan, p.can_occ, h->b.occ, p.lTap, p.llist,
                           p.status);
        launch_dense_apply_mfma<T>((const T *)h->b.psi, (T *)p.can_psi, 4 * N, p.llist, p.lhb,
                                   nblk, hmax, mpad, p.lTap, h->st);
        hfb_launch_mix(h, 0);
    });
}

// The scaled Newton-Schulz schedule: the scalings a_k for an interval
// [l0, 1] of |eigenvalues| / s, padded with a plain step (a = 1) to an even
// count (the iterate then ends in the buffer it starts in).
const std::vector<double> &sign_schedule() {
    static const std::vector<double> sched = [] {
        std::vector<double> a;
        double l = 1e-2;
        while (1.0 - l > 1e-15 && a.size() < 40) {
            const double al = std::sqrt(3.0 / (1.0 + l + l * l));
            a.push_back(al);
            l = 0.5 * al * l * (3.0 - al * al * l * l);
        }
        if (a.size() & 1) a.push_back(1.0);
        return a;
    }();
    return sched;
}

// Newton-Schulz steps on X (in X0, ending in X0 for an even count; X1 and Y
// scratch); the first products of the last step record the residual.
void sign_steps(mocca_solve *h, const double *alphas, int nsteps) {
    Hfb &p = h->hfb;
    const int nblk = p.k.nblk, ld = p.nmax;
    const long ms = (long)ld * ld;
    const int nt = (p.nmax + kSignTile - 1) / kSignTile;
    const dim3 grid(nt * (nt + 1) / 2, nblk);
    double *X = p.lA, *Xn = p.lC, *Y = p.lW;
    SignDims dims{};
    for (int b = 0; b < nblk; ++b) dims.n[b] = 2 * p.k.blk[b].h;
    for (int q = 0; q < nsteps; ++q) {
        const double a = alphas[q];
        if (q == nsteps - 1)
            hipLaunchKernelGGL(k_hs_gemm<true>, grid, dim3(256), 0, h->st, X, X, nullptr, 0.0,
                               1.0, Y, ld, ms, dims, p.lactive, p.lsig);
        else
            hipLaunchKernelGGL(k_hs_gemm<false>, grid, dim3(256), 0, h->st, X, X, nullptr, 0.0,
                               1.0, Y, ld, ms, dims, p.lactive, p.lsig);
        hipLaunchKernelGGL(k_hs_gemm<false>, grid, dim3(256), 0, h->st, X, Y, X, 1.5 * a,
                           -0.5 * a * a * a, Xn, ld, ms, dims, p.lactive, p.lsig);
        std::swap(X, Xn);
    }
}

// The Fermi solve of the sign path: the reference's Broyden iteration with
// N(lambda) from sign T(lambda) at every evaluation (lambda taken from the
// iteration's record), until both isospins meet FermiPrec or HFBIter
// evaluations; X of every block's last evaluation stays in lA.
int sign_fermi_solve(mocca_solve *h) {
    Hfb &p = h->hfb;
    const HfbK &k = p.k;
    const int nblk = k.nblk, ld = p.nmax;
    const int nel = p.nmax * p.nmax;
    double *redo = p.lnst + 2 * kHlNewton;
    const std::vector<double> &sched = sign_schedule();
    const double polish[2] = {1.0, 1.0};
    double *rs = p.lw;   // (the Jacobi's eigenvalue buffer, unused on this path)
    int rc = run(h, "hfb_fermi", [&] {
        hipLaunchKernelGGL(k_hl_newton_init, dim3(1), dim3(64), 0, h->st, p.fermi, p.lnst,
                           nblk, p.lactive);
        hipLaunchKernelGGL(k_hs_rowsum, dim3((ld + 3) / 4, nblk), dim3(256), 0, h->st, k,
                           p.delta, ld, rs);
    });
    if (rc) return rc;
    auto number = [&]() {
        return run(h, "hfb_fermi", [&] {
            hipLaunchKernelGGL(k_hs_number, dim3(nblk), dim3(64), 0, h->st, k, p.lA, ld,
                               p.lactive, p.lsig, p.lnum, redo);
            hipLaunchKernelGGL(k_hl_broyden, dim3(1), dim3(64), 0, h->st, k, p.lnum, p.lnst,
                               p.lactive, p.fermi, p.status, redo);
        });
    };
    const size_t nrec = (2 * kHlNewton + 1) * sizeof(double);
    for (int ev = 0; ev < std::max(1, k.iters) + 3; ++ev) {
        rc = run(h, "hfb_sign", [&] {
            hipLaunchKernelGGL(k_hs_build, dim3((nel + 255) / 256, nblk), dim3(256), 0, h->st,
                               k, h->b.E, p.delta, p.lnst, rs, p.lactive, ld, p.lA);
            sign_steps(h, sched.data(), (int)sched.size());
        });
        if (rc || (rc = number())) return rc;
        int extra = 0;
        for (;;) {
            const double *st = p.lhost;
            HIPCHK(h, hipMemcpyAsync(p.lhost, p.lnst, nrec, hipMemcpyDeviceToHost, h->st));
            HIPCHK(h, hipStreamSynchronize(h->st));
            if (st[2 * kHlNewton] == 0.0) {
                if (st[5] != 0.0 && st[kHlNewton + 5] != 0.0) return 0;
                break;
            }
            // an eigenvalue closer to 0 than the schedule assumed: more
            // plain steps (pairs, so that X stays in lA)
            if ((extra += 2) > 64)
                return fail(h, MOCCA_ERR_NUMERIC,
                            "HFB: the sign iteration of a Fermi-solve evaluation did not "
                            "converge (an eigenvalue of the HFB matrix at 0?); pairing "
                            "solver 2 diagonalises instead");
            rc = run(h, "hfb_sign", [&] {
                hipLaunchKernelGGL(k_hs_clear, dim3(1), dim3(1), 0, h->st, redo);
                sign_steps(h, polish, 2);
            });
            if (rc || (rc = number())) return rc;
        }
    }
    return 0;
}

template <typename T>
int hfb_large_solve_impl(mocca_solve *h) {
    Hfb &p = h->hfb;
    const HfbK &k = p.k;
    const long N = h->N;
    const int M = h->M, nblk = k.nblk, ld = p.nmax, hmax = p.hmax;
    const long ms = (long)ld * ld;
    const int mpad = (hmax + 15) / 16 * 16, ntile = (hmax + 15) / 16;
    int rc;
    // SolvePairing :800-806: old rho, kappa take the current ones while the
    // old rho is all zero (the first call)
    if ((rc = run(h, "hfb_mix", [&] { hfb_launch_mix(h, 1); }))) return rc;
    // cutoffs, density factor, pairing field
    rc = run(h, "hfb_field", [&] {
        hfb_launch_cutoffs(h);
        hipLaunchKernelGGL(k_hl_field_w, dim3(nblk), dim3(256), 0, h->st, p.k, p.kappa, p.pc,
                           hmax, p.lTap, p.llist, p.lhb);
        // (this path's blocks have at most 128 states: hfb_large_fits)
        launch_field_apply<T>(p.k, (const T *)h->b.psi, N, p.llist, p.lhb, hmax, mpad, p.lTap,
                              p.dfac, p.lgpart, p.field, h->st);
    });
    if (rc) return rc;
    // gaps
    rc = run(h, "hfb_gaps", [&] {
        launch_pair_gram<T>(p.k, (const T *)h->b.psi, N, p.field, p.lgpart, p.lsplit, ntile,
                            nblk, h->st);
        hipLaunchKernelGGL(k_hl_gaps_reduce, dim3((hmax * hmax + 255) / 256, nblk), dim3(256),
                           0, h->st, p.k, p.lgpart, p.lsplit, ntile, p.pc, h->mesh.dv, p.delta);
    });
    if (rc) return rc;
    // Fermi energies: Newton evaluations until both isospins meet FermiPrec
    const dim3 gmat((p.nmax + 15) / 16, (p.nmax + 15) / 16, nblk);
    const int nel = p.nmax * p.nmax;
    if (p.sign) {
        if ((rc = sign_fermi_solve(h))) return rc;
        p.lwarm = false;
        p.qpe_stale = true;
        rc = run(h, "hfb_canonical", [&] {
            hipLaunchKernelGGL(k_hs_rho_kappa, dim3((nel + 255) / 256, nblk), dim3(256), 0, h->st,
                               p.k, p.lA, ld, p.rho, p.kappa);
            hipLaunchKernelGGL(k_hl_can_prep, dim3((hmax * hmax + 255) / 256, nblk), dim3(256),
                               0, h->st, p.k, p.rho, ld, p.lT, p.lhb);
        });
        if (rc) return rc;
        return hfb_large_canonical<T>(h);
    }
    p.qpe_stale = false;
    rc = run(h, "hfb_fermi", [&] {
        hipLaunchKernelGGL(k_hl_newton_init, dim3(1), dim3(64), 0, h->st, p.fermi, p.lnst,
                           nblk, p.lactive);
        hipLaunchKernelGGL(k_hl_build_T, dim3((nel + 255) / 256, nblk), dim3(256), 0, h->st,
                           p.k, h->b.E, p.delta, p.fermi, ld, p.lT);
        if (p.lwarm) {
            // A = V^T T V in the eigenbasis of the last call
            hipLaunchKernelGGL((k_lj_gemm<false, false>), gmat, dim3(256), 0, h->st, p.lT, p.lV,
                               p.lW, ld, ms, p.lnb, nullptr, nullptr);
            hipLaunchKernelGGL((k_lj_gemm<true, false>), gmat, dim3(256), 0, h->st, p.lV, p.lW,
                               p.lA, ld, ms, p.lnb, nullptr, nullptr);
        }
    });
    if (rc) return rc;
    const double *A0 = p.lwarm ? p.lA : p.lT;
    int conv[2] = {0, 0};
    // (Broyden: two evaluations before its HFBIter iterations)
    for (int ev = 0; ev < std::max(1, k.iters) + 3; ++ev) {
        if ((rc = jacobi_run(h, ev == 0 ? A0 : p.lA, p.lV, p.lw, p.lnb, p.nmax, p.lactive,
                             p.lrounds, ev == 0 && !p.lwarm, false,
                             ev == 0 ? nullptr : p.lcold, p.lAf)))
            return rc;
        rc = run(h, "hfb_fermi", [&] {
            hipLaunchKernelGGL((k_lj_gemm<true, true>), gmat, dim3(256), 0, h->st, p.lV, p.lV,
                               p.lC, ld, ms, p.lnb, p.lhb, p.lactive);
            hipLaunchKernelGGL(k_hl_number, dim3(nblk), dim3(kNumThreads), 0, h->st, p.k, p.lV,
                               p.lC, p.lw, ld, p.lactive, p.lrank, p.lnum, p.delta, p.loff);
            if (p.newton)
                hipLaunchKernelGGL(k_hl_newton, dim3(1), dim3(64), 0, h->st, p.k, p.lnum,
                                   p.lnst, p.lactive, p.fermi, p.status);
            else
                hipLaunchKernelGGL(k_hl_broyden, dim3(1), dim3(64), 0, h->st, p.k, p.lnum,
                                   p.lnst, p.lactive, p.fermi, p.status, nullptr);
            hipLaunchKernelGGL(k_hl_shift_A, dim3((nel + 255) / 256, nblk), dim3(256), 0,
                               h->st, p.k, p.lAf, p.lC, p.lnst, p.lactive, ld, p.lA, h->b.E,
                               p.delta, p.loff, p.lcold);
        });
        if (rc) return rc;
        double st[2 * kHlNewton];
        HIPCHK(h, hipMemcpyAsync(st, p.lnst, sizeof st, hipMemcpyDeviceToHost, h->st));
        HIPCHK(h, hipStreamSynchronize(h->st));
        conv[0] = st[5] != 0.0;
        conv[1] = st[kHlNewton + 5] != 0.0;
        if (conv[0] && conv[1]) break;
    }
    p.lwarm = true;
    // rho, kappa, quasiparticle energies; canonical basis and states; mixing
    rc = run(h, "hfb_canonical", [&] {
        hipLaunchKernelGGL(k_hl_rho_kappa, dim3((nel + 255) / 256, nblk), dim3(256), 0, h->st,
                           p.k, p.lV, p.lw, p.lrank, ld, p.rho, p.kappa, p.qpe);
        hipLaunchKernelGGL(k_hl_can_prep, dim3((hmax * hmax + 255) / 256, nblk), dim3(256), 0,
                           h->st, p.k, p.rho, ld, p.lT, p.lhb);
    });
    if (rc) return rc;
    return hfb_large_canonical<T>(h);
}

int hfb_ensure_qpe(mocca_solve *h) {
    Hfb &p = h->hfb;
    if (!p.on || !p.large || !p.qpe_stale) return 0;
    const HfbK &k = p.k;
    const int nblk = k.nblk, ld = p.nmax, nel = p.nmax * p.nmax;
    int rc = run(h, "hfb_fermi", [&] {
        hipLaunchKernelGGL(k_hs_activate, dim3(1), dim3(64), 0, h->st, p.lactive, nblk);
        hipLaunchKernelGGL(k_hl_build_T, dim3((nel + 255) / 256, nblk), dim3(256), 0, h->st, k,
                           h->b.E, p.delta, p.fermi, ld, p.lT);
    });
    if (rc || (rc = jacobi_run(h, p.lT, p.lV, p.lw, p.lnb, p.nmax, p.lactive, p.lrounds, true,
                               false, nullptr, nullptr)))
        return rc;
    rc = run(h, "hfb_fermi", [&] {
        hipLaunchKernelGGL(k_hl_qpe, dim3(nblk), dim3(256), 0, h->st, k, p.lw, ld, p.qpe);
    });
    if (!rc) p.qpe_stale = false;
    return rc;
}

int hfb_large_solve_any(mocca_solve *h) {
    return h->prec == MOCCA_FP64 ? hfb_large_solve_impl<double>(h)
                                 : hfb_large_solve_impl<float>(h);
}

}  // namespace detail
}  // namespace mocca
