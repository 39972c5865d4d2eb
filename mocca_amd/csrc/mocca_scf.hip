// Self-consistent part of the imaginary-time iteration on the device
// (gfx950, wave64): HF filling, density mixing, multipole moments, the
// Coulomb Poisson solve, the mean-field potentials and the energy.
//
// All of it is FP64 (it is O(N) per iteration against the O(M N) of the
// spinor kernels).  Reductions are deterministic: per-workgroup partial sums
// in a fixed layout, combined in a fixed order.
#include <hip/hip_cooperative_groups.h>

#include "mocca_kernels.h"

namespace mocca {

namespace {

__device__ __forceinline__ double wave_sum(double a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    return a;
}

// Sum of one value over the workgroup (kBlock threads); result on every thread.
__device__ __forceinline__ double block_sum(double a, double *sh) {
    a = wave_sum(a);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = a;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) s += sh[k];
    return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// NaiveFill (HartreeFock.f90:177-262): the states are ordered by a stable
// ascending sort of the energies (OrderSpwfs, SpwfStorage.f90:754-793); the
// first ub[it] states of each isospin in that order get occupation fac
// (2 under time reversal, which doubles every occupation), the rest 0.
// A state's place in the order is the number of states before it, so one
// thread per state decides its own occupation.
__global__ __launch_bounds__(kBlock) void k_hf_fill(const double *__restrict__ E,
                                                     const int *__restrict__ qn, int M,
                                                     int ub_n, int ub_p, double fac,
                                                     double *__restrict__ occ) {
    extern __shared__ double se[];          // [M] energies, then [M] isospins
    int *si = reinterpret_cast<int *>(se + M);
    for (int j = threadIdx.x; j < M; j += kBlock) {
        se[j] = E[j];
        si[j] = qn[4 * j];
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M) return;
    const int iso = si[i];
    const double e = se[i];
    int before = 0;
    for (int j = 0; j < M; ++j) {
        const double f = se[j];
        before += (si[j] == iso && (f < e || (f == e && j < i))) ? 1 : 0;
    }
    const int ub = iso == -1 ? ub_n : ub_p;
    occ[i] = before < ub ? fac : 0.0;
}

void launch_hf_fill(const double *E, const int *qn, int M, int ub_n, int ub_p, double fac,
                    double *occ, hipStream_t st) {
    const size_t shm = (size_t)M * (sizeof(double) + sizeof(int));
    hipLaunchKernelGGL(k_hf_fill, dim3((M + kBlock - 1) / kBlock), dim3(kBlock), shm, st, E, qn,
                       M, ub_n, ub_p, fac, occ);
}

// ---------------------------------------------------------------------------
// MixDensities, scheme 0 (DensityMixing.f90:39-45): d = (1-a) d + a h over
// every mixed density; with keep != 0 the history takes the mixed value too
// (what the next UpdateDensities stores as DensityHistory(1)).
__global__ __launch_bounds__(kBlock) void k_mix(double *__restrict__ d, double *__restrict__ h,
                                                 long n, double a, int keep) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double v = (1.0 - a) * d[i] + a * h[i];
    d[i] = v;
    if (keep) h[i] = v;
}

void launch_mix(double *d, double *h, long n, double a, int keep, hipStream_t st) {
    hipLaunchKernelGGL(k_mix, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d, h, n, a,
                       keep);
}

// ---------------------------------------------------------------------------
// Multipole moments (Calculate_Multipole, Moments.f90:1907-2004):
// Q[k][it] = dv * sum_p Y[k][p] rho_it(p).  Pass 1: per-workgroup partial
// sums part[wg][k][it]; pass 2: one thread per (k, it) sums the partials in
// workgroup order.
__global__ __launch_bounds__(kBlock) void k_moments(const double *__restrict__ rho_n,
                                                     const double *__restrict__ rho_p,
                                                     const double *__restrict__ Y, int nmom,
                                                     long N, double *__restrict__ part) {
    __shared__ double sh[kBlock / 64];
    double an[kMaxMom], ap[kMaxMom];
#pragma unroll
    for (int k = 0; k < kMaxMom; ++k) { an[k] = 0; ap[k] = 0; }
    for (long p = (long)blockIdx.x * kBlock + threadIdx.x; p < N; p += (long)gridDim.x * kBlock) {
        const double rn = rho_n[p], rp = rho_p[p];
#pragma unroll
        for (int k = 0; k < kMaxMom; ++k)
            if (k < nmom) {
                const double y = Y[k * N + p];
                an[k] += y * rn;
                ap[k] += y * rp;
            }
    }
    for (int k = 0; k < nmom; ++k) {
        const double sn = block_sum(an[k], sh);
        const double sp = block_sum(ap[k], sh);
        if (threadIdx.x == 0) {
            part[((long)blockIdx.x * kMaxMom + k) * 2] = sn;
            part[((long)blockIdx.x * kMaxMom + k) * 2 + 1] = sp;
        }
    }
}

// One wave per output (moment, isospin): lanes sum strided partials, then a
// fixed shuffle tree.
__global__ __launch_bounds__(64) void k_moments_reduce(const double *__restrict__ part, int nblk,
                                                      int nmom, double dv,
                                                      double *__restrict__ Q) {
    const int t = blockIdx.x;
    if (t >= 2 * nmom) return;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 64) s += part[(long)b * kMaxMom * 2 + t];
    s = wave_sum(s);
    if (threadIdx.x == 0) Q[t] = s * dv;
}

void launch_moments(const double *rho_n, const double *rho_p, const double *Y, int nmom, long N,
                    double dv, double *part, int nblk, double *Q, hipStream_t st) {
    hipLaunchKernelGGL(k_moments, dim3(nblk), dim3(kBlock), 0, st, rho_n, rho_p, Y, nmom, N,
                       part);
    hipLaunchKernelGGL(k_moments_reduce, dim3(2 * nmom), dim3(64), 0, st, part, nblk, nmom, dv,
                       Q);
}

// ---------------------------------------------------------------------------
// Coulomb boundary values (CoulombBound, Coulomb.f90:583-674): on the
// boundary planes of the Coulomb box, V = sum_lm Q_lm(p) 4pi/(2l+1) e2
// r^l Y_lm / r^(2l+1); a non-zero value replaces the potential there.
__global__ __launch_bounds__(kBlock) void k_coulomb_bc(const double *__restrict__ Q, int nmom,
                                                        const double *__restrict__ lfac,
                                                        const int *__restrict__ bidx,
                                                        const double *__restrict__ bsh,
                                                        const double *__restrict__ brp, int nb,
                                                        double e2, double *__restrict__ x) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= nb) return;
    double bc = 0.0;
    for (int k = 0; k < nmom; ++k) {
        const double value = Q[2 * k + 1] * lfac[k];
        bc = bc + value * bsh[(long)k * nb + b] * e2 / brp[(long)k * nb + b];
    }
    if (bc != 0.0) x[bidx[b]] = bc;
}

void launch_coulomb_bc(const double *Q, int nmom, const double *lfac, const int *bidx,
                       const double *bsh, const double *brp, int nb, double e2, double *x,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_coulomb_bc, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, Q,
                       nmom, lfac, bidx, bsh, brp, nb, e2, x);
}

// ---------------------------------------------------------------------------
// Conjugate gradients for Lap V = -4 pi e2 rho_p on the Coulomb box
// (ConjugGrad, Coulomb.f90:755-900; CoulombLaplacian / Coulomb_1D,
// CoulombDerivatives.f90:144-281): 5-point second derivative per axis,
// mirror (sign +1) below the symmetry planes, boundary planes held fixed.
// The reference step length a = r.r / r.Ap and stop sum r^2 <= prec.
//
// One cooperative launch runs the whole solve.  Per iteration: phase A
// forms p = r + c p_old on the fly (double-buffered p), Ap and the partial
// r.Ap; grid barrier; phase B updates V and r and the partial r.r; grid
// barrier.  Every workgroup sums the partials in the same order, so all
// take the same decisions and leave the loop together.  Small boxes run the
// same code as one ordinary workgroup (COOP = false, workgroup barriers):
// a cooperative launch in a process costs the later launches of other
// streams their concurrency on this runtime (measured: 64 graph-replayed
// solves on their own streams ran 1.6x slower after one), which the
// single-workgroup form avoids where it is fast enough.
namespace {

__device__ __forceinline__ int mirror(int a) { return a < 0 ? -1 - a : a; }

// Fixed-order sum of nwg partials; identical on every workgroup.
__device__ __forceinline__ double grid_total(const double *part, int nwg, double *sh) {
    double s = 0;
    if (threadIdx.x < 64) {
        for (int b = threadIdx.x; b < nwg; b += 64) s += part[b];
        s = wave_sum(s);
        if (threadIdx.x == 0) sh[0] = s;
    }
    __syncthreads();
    s = sh[0];
    __syncthreads();
    return s;
}

}  // namespace

template <typename PF>
__device__ __forceinline__ double coul_lap(const CoulombK &c, int i, int j, int k, PF &&val) {
    constexpr double c2 = -1.0 / 12.0, c1 = 4.0 / 3.0, c0 = -5.0 / 2.0;
    const double cf[5] = {c2, c1, c0, c1, c2};
    auto at = [&](int a, int b, int d) { return val(((long)d * c.my + b) * c.mx + a); };
    double dx = 0, dy = 0, dz = 0;
#pragma unroll
    for (int q = -2; q <= 2; ++q) {
        dx = dx + cf[q + 2] * at(mirror(i + q), j, k);
        dy = dy + cf[q + 2] * at(i, mirror(j + q), k);
        dz = dz + cf[q + 2] * at(i, j, mirror(k + q));
    }
    return (dx * c.f + dy * c.f) + dz * c.f;
}

template <bool COOP>
__global__ __launch_bounds__(kBlock) void k_coulomb_cg(CoulombK c, double *__restrict__ x,
                                                        const double *__restrict__ rho_p,
                                                        double *__restrict__ r,
                                                        double *__restrict__ pa,
                                                        double *__restrict__ pb,
                                                        double *__restrict__ ap,
                                                        double *__restrict__ part,
                                                        double *__restrict__ vcoul,
                                                        double *__restrict__ vcoulx,
                                                        int *__restrict__ iters) {
    namespace cg = cooperative_groups;
    auto gsync = [&] {
        if constexpr (COOP) cg::this_grid().sync();
        else __syncthreads();
    };
    __shared__ double sh[kBlock / 64 + 1];
    const int nwg = gridDim.x;
    const long stride = (long)nwg * kBlock;
    const long nint = (long)c.nx * c.ny * c.nz;
    const double src_f = -4.0 * 3.14159265358979323846 * c.e2;
    auto coords = [&](long q, int &i, int &j, int &k) {
        i = q % c.nx;
        j = (q / c.nx) % c.ny;
        k = q / ((long)c.nx * c.ny);
    };
    auto bidx = [&](int i, int j, int k) { return ((long)k * c.my + j) * c.mx + i; };
    // r = -Lap V + source, p = r, norm = r.r
    double loc = 0;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nint; q += stride) {
        int i, j, k;
        coords(q, i, j, k);
        const double lv = coul_lap(c, i, j, k, [&](long t) { return x[t]; });
        double rv = -1.0 * lv;
        rv = rv + src_f * rho_p[q];
        r[bidx(i, j, k)] = rv;
        pa[bidx(i, j, k)] = rv;
        loc += rv * rv;
    }
    // (partials of r.r live in part[nwg..2nwg), those of r.Ap in part[0..nwg):
    // a workgroup may only overwrite a region after the barrier that follows
    // every other workgroup's read of it)
    double s = block_sum(loc, sh);
    if (threadIdx.x == 0) part[nwg + blockIdx.x] = s;
    gsync();
    double norm = grid_total(part + nwg, nwg, sh);
    double cc = 0;
    double *pcur = pa, *pnext = pb;
    int it;
    for (it = 1; it <= c.maxit; ++it) {
        // phase A: p (on the fly for it > 1), Ap, r.Ap
        loc = 0;
        const bool first = it == 1;
        for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nint; q += stride) {
            int i, j, k;
            coords(q, i, j, k);
            double lv;
            if (first)
                lv = coul_lap(c, i, j, k, [&](long t) { return pcur[t]; });
            else
                lv = coul_lap(c, i, j, k, [&](long t) { return r[t] + cc * pcur[t]; });
            const long b = bidx(i, j, k);
            if (!first) pnext[b] = r[b] + cc * pcur[b];
            ap[b] = lv;
            loc += r[b] * lv;
        }
        s = block_sum(loc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
        gsync();
        if (!first) { double *t = pcur; pcur = pnext; pnext = t; }
        const double integral = grid_total(part, nwg, sh);
        const double a = norm / integral;
        // phase B: V += a p, r -= a Ap, r.r
        loc = 0;
        for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nint; q += stride) {
            int i, j, k;
            coords(q, i, j, k);
            const long b = bidx(i, j, k);
            x[b] = x[b] + a * pcur[b];
            const double rv = r[b] - a * ap[b];
            r[b] = rv;
            loc += rv * rv;
        }
        s = block_sum(loc, sh);
        if (threadIdx.x == 0) part[nwg + blockIdx.x] = s;
        gsync();
        const double nn = grid_total(part + nwg, nwg, sh);
        if (nn <= c.prec) break;
        cc = nn / norm;
        norm = nn;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *iters = it > c.maxit ? c.maxit : it;
    // Coulomb potential on the mesh and the Slater exchange potential
    // (SolveCoulomb, Coulomb.f90:329-339).
    const double fx = -pow(3.0 / 3.14159265358979323846, 1.0 / 3.0) * c.e2;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < nint; q += stride) {
        int i, j, k;
        coords(q, i, j, k);
        vcoul[q] = x[bidx(i, j, k)];
        vcoulx[q] = fx * pow(rho_p[q], 1.0 / 3.0);
    }
}

int launch_coulomb_cg(const CoulombK &c, double *x, const double *rho_p, double *r, double *pa,
                      double *pb, double *ap, double *part, int nwg, double *vcoul,
                      double *vcoulx, int *iters, hipStream_t st) {
    CoulombK cc = c;
    if (nwg == 1) {
        hipLaunchKernelGGL(k_coulomb_cg<false>, dim3(1), dim3(kBlock), 0, st, cc, x, rho_p, r, pa,
                           pb, ap, part, vcoul, vcoulx, iters);
        return (int)hipGetLastError();
    }
    void *args[] = {&cc, &x, &rho_p, &r, &pa, &pb, &ap, &part, &vcoul, &vcoulx, &iters};
    return (int)hipLaunchCooperativeKernel((const void *)k_coulomb_cg<true>, dim3(nwg),
                                           dim3(kBlock), args, 0, st);
}

int coulomb_cg_max_groups() {
    int dev = 0, per_cu = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_coulomb_cg<true>,
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return per_cu * ncu;
}

// ---------------------------------------------------------------------------
// Direct solve of the same discrete Poisson problem.  The box operator is a
// sum of three 1-D operators, A = T_x + T_y + T_z, each the 5-point second
// derivative with the mirror rows at the symmetry plane; with T_a = Q_a
// diag(lambda_a) Q_a^T (symmetric, Q_a orthogonal, computed once on the
// host) the interior solution is V = (Q_z Q_y Q_x) D^-1 (Q_x^T Q_y^T Q_z^T) b,
// D = lambda_x + lambda_y + lambda_z, with b the source minus what the fixed
// boundary planes contribute through the stencil.  Six small transforms
// along the axes replace the conjugate-gradient iterations; the result is
// the exact discrete solution (to rounding), which the reference's CG
// approaches to its stopping residual.
__global__ __launch_bounds__(kBlock) void k_coulomb_rhs(CoulombK c, const double *__restrict__ x,
                                                         const double *__restrict__ rho_p,
                                                         double *__restrict__ b) {
    const long q = (long)blockIdx.x * kBlock + threadIdx.x;
    const long nint = (long)c.nx * c.ny * c.nz;
    if (q >= nint) return;
    const int i = q % c.nx, j = (q / c.nx) % c.ny, k = q / ((long)c.nx * c.ny);
    const double cf[3] = {-5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
    auto X = [&](int a, int bb, int d) { return x[((long)d * c.my + bb) * c.mx + a]; };
    double acc = -4.0 * 3.14159265358979323846 * c.e2 * rho_p[q];
    for (int s = 1; s <= 2; ++s) {
        if (i + s >= c.nx) acc -= cf[s] * c.f * X(i + s, j, k);
        if (j + s >= c.ny) acc -= cf[s] * c.f * X(i, j + s, k);
        if (k + s >= c.nz) acc -= cf[s] * c.f * X(i, j, k + s);
    }
    b[q] = acc;
}

// out = Q^T line (forward) or Q line (inverse) along one axis of length n
// with element stride `inner`; with lam != null the forward result is
// divided by lam_x + lam_y + lam_z.
__global__ __launch_bounds__(kBlock) void k_axis_transform(const double *__restrict__ in,
                                                            double *__restrict__ out,
                                                            const double *__restrict__ Q, int n,
                                                            long inner, long total, int inverse,
                                                            const double *__restrict__ lam_x,
                                                            const double *__restrict__ lam_y,
                                                            const double *__restrict__ lam_z,
                                                            int nx, int ny) {
    extern __shared__ double qs[];
    const bool lds = n <= 64;
    if (lds) {
        for (int t = threadIdx.x; t < n * n; t += kBlock) qs[t] = Q[t];
        __syncthreads();
    }
    const double *M = lds ? qs : Q;
    const long e = (long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= total) return;
    const long line = e / n;           // lines are enumerated (outer, inner-offset)
    const int a = e % n;
    const long base = (line / inner) * n * inner + (line % inner);
    double acc = 0;
    if (!inverse) {
        for (int i = 0; i < n; ++i) acc += M[i * n + a] * in[base + i * inner];
    } else {
        for (int i = 0; i < n; ++i) acc += M[a * n + i] * in[base + i * inner];
    }
    const long o = base + a * inner;
    if (lam_x) {
        const int xi = o % nx, yi = (o / nx) % ny, zi = o / ((long)nx * ny);
        acc = acc / (lam_x[xi] + lam_y[yi] + lam_z[zi]);
    }
    out[o] = acc;
}

__global__ __launch_bounds__(kBlock) void k_coulomb_finish(CoulombK c, double *__restrict__ x,
                                                            const double *__restrict__ v,
                                                            const double *__restrict__ rho_p,
                                                            double *__restrict__ vcoul,
                                                            double *__restrict__ vcoulx) {
    const long q = (long)blockIdx.x * kBlock + threadIdx.x;
    const long nint = (long)c.nx * c.ny * c.nz;
    if (q >= nint) return;
    const int i = q % c.nx, j = (q / c.nx) % c.ny, k = q / ((long)c.nx * c.ny);
    x[((long)k * c.my + j) * c.mx + i] = v[q];
    vcoul[q] = v[q];
    vcoulx[q] = -pow(3.0 / 3.14159265358979323846, 1.0 / 3.0) * c.e2 * pow(rho_p[q], 1.0 / 3.0);
}

void launch_coulomb_direct(const CoulombK &c, double *x, const double *rho_p,
                           const double *const Q[3], const double *const lam[3], double *w0,
                           double *w1, double *vcoul, double *vcoulx, hipStream_t st) {
    const long nint = (long)c.nx * c.ny * c.nz;
    const dim3 g((nint + kBlock - 1) / kBlock), b(kBlock);
    const int n[3] = {c.nx, c.ny, c.nz};
    const long inner[3] = {1, c.nx, (long)c.nx * c.ny};
    auto shm = [&](int a) { return n[a] <= 64 ? sizeof(double) * n[a] * n[a] : 0; };
    hipLaunchKernelGGL(k_coulomb_rhs, g, b, 0, st, c, x, rho_p, w0);
    // forward x, y, z (scaled), inverse z, y, x: w0 -> w1 -> w0 -> w1 -> w0 -> w1 -> w0
    double *buf[2] = {w0, w1};
    int cur = 0;
    for (int a = 0; a < 3; ++a, cur ^= 1)
        hipLaunchKernelGGL(k_axis_transform, g, b, shm(a), st, buf[cur], buf[cur ^ 1], Q[a], n[a],
                           inner[a], nint, 0, a == 2 ? lam[0] : nullptr, lam[1], lam[2], c.nx,
                           c.ny);
    for (int a = 2; a >= 0; --a, cur ^= 1)
        hipLaunchKernelGGL(k_axis_transform, g, b, shm(a), st, buf[cur], buf[cur ^ 1], Q[a], n[a],
                           inner[a], nint, 1, nullptr, nullptr, nullptr, c.nx, c.ny);
    hipLaunchKernelGGL(k_coulomb_finish, g, b, 0, st, c, x, buf[cur], rho_p, vcoul, vcoulx);
}

// ---------------------------------------------------------------------------
// Mean-field potentials, pointwise part: U (CalcUPot, MeanFields.f90:269-409),
// B (CalcBPot :115-150), W (CalcWPot :494-541), S (CalcSPot :618-712) and
// A (CalcAPot :411-446) into pot[48][N] at the field offsets of the hPsi
// kernels.  Their derivative fields (NablaB, DivW, DivA) are field jobs.
__global__ __launch_bounds__(kBlock) void k_potentials(SkyrmeK f, long N, int trc,
                                                        const double *__restrict__ dens,
                                                        const double *__restrict__ derived,
                                                        const double *__restrict__ vcoul,
                                                        const double *__restrict__ vcoulx,
                                                        const double *__restrict__ ce,
                                                        double *__restrict__ pot) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    const double eps = 1.0e-20;
    auto D = [&](int it, int slot) { return dens[((long)it * 9 + slot) * N + p]; };
    const double *derrho = derived, *laprho = derived + 6 * N, *rotj = derived + 8 * N,
                 *rots = derived + 14 * N;
    const double rn = D(0, 0), rp = D(1, 0);
    const double rt = rn + rp;
    const double njt = D(0, 2) + D(1, 2);
    const double a3 = f.byt3a, a3b = f.byt3b;
    double U[2];
    U[0] = (2.0 * f.B1 + 2.0 * f.B2) * rn + 2.0 * f.B1 * rp;
    U[1] = (2.0 * f.B1 + 2.0 * f.B2) * rp + 2.0 * f.B1 * rn;
    const double rta = pow(rt, a3), rt1 = pow(rt + eps, 1 + a3);
    const double rtb = pow(rt, a3b), rt1b = pow(rt + eps, 1 + a3b);
    for (int it = 0; it < 2; ++it) {
        const int at = 1 - it;
        const double ri = D(it, 0), ra = D(at, 0);
        U[it] = U[it] + (f.B3 + f.B4) * D(it, 1) + f.B3 * D(at, 1) +
                2.0 * ((f.B5 + f.B6) * laprho[it * N + p] + f.B5 * laprho[at * N + p]) +
                (2 + a3) * f.B7a * rt1 +
                f.B8a * (a3 * rta / (rt + eps) * (ri * ri + ra * ra) + 2.0 * rta * ri) +
                f.B9 * njt + f.B9q * D(it, 2);
        if (f.B7b != 0.0 || f.B8b != 0.0)
            U[it] = U[it] + (2 + a3b) * f.B7b * rt1b +
                    f.B8b * (a3b * rtb / (rt + eps) * (ri * ri + ra * ra) + 2.0 * rtb * ri);
    }
    double s[2][3] = {{0, 0, 0}, {0, 0, 0}};
    if (!trc) {
        double st2 = 0, s2[2] = {0, 0};
        for (int k = 0; k < 3; ++k) {
            s[0][k] = D(0, 6 + k);
            s[1][k] = D(1, 6 + k);
            st2 = st2 + (s[0][k] + s[1][k]) * (s[0][k] + s[1][k]);
            s2[0] = s2[0] + s[0][k] * s[0][k];
            s2[1] = s2[1] + s[1][k] * s[1][k];
        }
        for (int it = 0; it < 2; ++it)
            U[it] = U[it] + a3 * rta / (rt + eps) *
                                (f.B12a * st2 + f.B13a * s2[it] + f.B13a * s2[1 - it]);
        if (f.B12b != 0.0 || f.B13b != 0.0)
            for (int it = 0; it < 2; ++it)
                U[it] = U[it] + a3b * rtb / (rt + eps) *
                                    (f.B12b * st2 + f.B13b * s2[it] + f.B13b * s2[1 - it]);
    }
    U[1] = U[1] + vcoul[p];
    if (f.cexchange == 2) U[1] = U[1] + vcoulx[p];
    // multipole constraints (MeanFields.f90:402-403)
    if (ce) {
        U[0] = U[0] + ce[p];
        U[1] = U[1] + ce[N + p];
    }
    pot[(kOffU + 0) * N + p] = U[0];
    pot[(kOffU + 1) * N + p] = U[1];
    // B (CalcBPot)
    for (int it = 0; it < 2; ++it)
        pot[(kOffB + it) * N + p] = f.B3 * D(1 - it, 0) + (f.B3 + f.B4) * D(it, 0);
    // W_{mu nu} = -sum_k eps_{k mu nu} ((B9+B9q) d_k rho_q + B9 d_k rho_q'),
    // stored W[it][nu][mu].
    for (int it = 0; it < 2; ++it) {
        double g[3];
        for (int k = 0; k < 3; ++k)
            g[k] = (f.B9 + f.B9q) * derrho[(it * 3 + k) * N + p] +
                   f.B9 * derrho[((1 - it) * 3 + k) * N + p];
        for (int nu = 0; nu < 3; ++nu)
            for (int mu = 0; mu < 3; ++mu) {
                double w = 0.0;
                if (mu != nu) {
                    const int k = 3 - mu - nu;
                    // eps_{k mu nu} = +1 for (k, mu, nu) an even permutation of (0,1,2)
                    const int e = ((mu - k + 3) % 3 == 1) ? 1 : -1;
                    w = e > 0 ? 0.0 - g[k] : 0.0 + g[k];
                }
                pot[(kOffW + (it * 3 + nu) * 3 + mu) * N + p] = w;
            }
    }
    if (trc) return;
    // S (CalcSPot) and A (CalcAPot)
    for (int it = 0; it < 2; ++it) {
        const int at = 1 - it;
        for (int k = 0; k < 3; ++k) {
            double sv = +(f.B9 + f.B9q) * rotj[(it * 3 + k) * N + p] +
                        f.B9 * rotj[(at * 3 + k) * N + p] +
                        2.0 * ((f.B10 + f.B11) * s[it][k] + f.B10 * s[at][k]) +
                        2.0 * ((f.B12a + f.B13a) * (rta * s[it][k]) + f.B12a * (rta * s[at][k]));
            if (f.B12b != 0.0 || f.B13b != 0.0)
                sv = sv + 2.0 * ((f.B12b + f.B13b) * (rtb * s[it][k]) + f.B12b * (rtb * s[at][k]));
            pot[(kOffS + it * 3 + k) * N + p] = sv;
            double av = -2.0 * ((f.B3 + f.B4) * D(it, 3 + k) + f.B3 * D(at, 3 + k));
            av = av + ((f.B9 + f.B9q) * rots[(it * 3 + k) * N + p] +
                       f.B9 * rots[(at * 3 + k) * N + p]);
            pot[(kOffA + it * 3 + k) * N + p] = av;
        }
    }
}

void launch_potentials(const SkyrmeK &f, long N, int trc, const double *dens,
                       const double *derived, const double *vcoul, const double *vcoulx,
                       const double *ce, double *pot, hipStream_t st) {
    hipLaunchKernelGGL(k_potentials, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, f, N,
                       trc, dens, derived, vcoul, vcoulx, ce, pot);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert(const double *__restrict__ a,
                                                     T *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

void launch_convert_f32(const double *a, float *b, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_convert<float>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, a,
                       b, n);
}

// ---------------------------------------------------------------------------
// Energy (compSkyrme, Energy.f90:623-859; Coulomb energies Coulomb.f90:
// 215-253).  Pass 1: per-workgroup partial sums of the integrands, slot
// layout kE*; pass 2 (one workgroup): fixed-order sums, coefficients, dv.
enum {
    kE_rt2, kE_rn2, kE_rp2, kE_rttt, kE_rntn, kE_rptp, kE_jt2, kE_jn2, kE_jp2, kE_rtlt,
    kE_rnln, kE_rplp, kE_rta2, kE_rn2a, kE_rp2a, kE_rtnj, kE_rnnj, kE_rpnj, kE_jrst, kE_jrsn,
    kE_jrsp, kE_st2, kE_sn2, kE_sp2, kE_ast2, kE_asn2, kE_asp2, kE_rtb2, kE_rn2b, kE_rp2b,
    kE_bst2, kE_bsn2, kE_bsp2, kE_coul, kE_cx, kE_count
};
static_assert(kE_count == kEnergySlots, "energy partial-sum slots");

__global__ __launch_bounds__(kBlock) void k_energy_partial(SkyrmeK f, long N, int trc,
                                                            const double *__restrict__ dens,
                                                            const double *__restrict__ derived,
                                                            const double *__restrict__ vcoul,
                                                            double *__restrict__ part) {
    __shared__ double sh[kBlock / 64];
    double acc[kE_count];
#pragma unroll
    for (int k = 0; k < kE_count; ++k) acc[k] = 0;
    const double *laprho = derived + 6 * N, *rots = derived + 14 * N;
    for (long p = (long)blockIdx.x * kBlock + threadIdx.x; p < N; p += (long)gridDim.x * kBlock) {
        auto D = [&](int it, int slot) { return dens[((long)it * 9 + slot) * N + p]; };
        const double rn = D(0, 0), rp = D(1, 0), rt = rn + rp;
        const double tn = D(0, 1), tp = D(1, 1);
        const double ln = laprho[p], lp = laprho[N + p];
        const double nn = D(0, 2), np_ = D(1, 2);
        const double rta = pow(rt, f.byt3a);
        acc[kE_rt2] += rt * rt;
        acc[kE_rn2] += rn * rn;
        acc[kE_rp2] += rp * rp;
        acc[kE_rttt] += rt * (tn + tp);
        acc[kE_rntn] += rn * tn;
        acc[kE_rptp] += rp * tp;
        acc[kE_rtlt] += rt * (ln + lp);
        acc[kE_rnln] += rn * ln;
        acc[kE_rplp] += rp * lp;
        acc[kE_rta2] += pow(rt, 2 + f.byt3a);
        acc[kE_rn2a] += rn * rn * rta;
        acc[kE_rp2a] += rp * rp * rta;
        acc[kE_rtnj] += rt * (nn + np_);
        acc[kE_rnnj] += rn * nn;
        acc[kE_rpnj] += rp * np_;
        if (f.B7b != 0.0) acc[kE_rtb2] += pow(rt, 2 + f.byt3b);
        if (f.B8b != 0.0) {
            const double rtb = pow(rt, f.byt3b);
            acc[kE_rn2b] += rn * rn * rtb;
            acc[kE_rp2b] += rp * rp * rtb;
        }
        if (!trc) {
            double jt2 = 0, jn2 = 0, jp2 = 0, jrst = 0, jrsn = 0, jrsp = 0, st2 = 0, sn2 = 0,
                   sp2 = 0;
            for (int k = 0; k < 3; ++k) {
                const double jn = D(0, 3 + k), jp = D(1, 3 + k);
                const double sn = D(0, 6 + k), sp = D(1, 6 + k);
                const double rsn = rots[k * N + p], rsp = rots[(3 + k) * N + p];
                jt2 += (jn + jp) * (jn + jp);
                jn2 += jn * jn;
                jp2 += jp * jp;
                jrst += (jn + jp) * (rsn + rsp);
                jrsn += jn * rsn;
                jrsp += jp * rsp;
                st2 += (sn + sp) * (sn + sp);
                sn2 += sn * sn;
                sp2 += sp * sp;
            }
            acc[kE_jt2] += jt2;
            acc[kE_jn2] += jn2;
            acc[kE_jp2] += jp2;
            acc[kE_jrst] += jrst;
            acc[kE_jrsn] += jrsn;
            acc[kE_jrsp] += jrsp;
            acc[kE_st2] += st2;
            acc[kE_sn2] += sn2;
            acc[kE_sp2] += sp2;
            acc[kE_ast2] += rta * st2;
            acc[kE_asn2] += rta * sn2;
            acc[kE_asp2] += rta * sp2;
            if (f.B12b != 0.0 || f.B13b != 0.0) {
                const double rtb = pow(rt, f.byt3b);
                acc[kE_bst2] += rtb * st2;
                acc[kE_bsn2] += rtb * sn2;
                acc[kE_bsp2] += rtb * sp2;
            }
        }
        acc[kE_coul] += rp * vcoul[p];
        acc[kE_cx] += rp * pow(rp, 1.0 / 3.0);
    }
    for (int k = 0; k < kE_count; ++k) {
        const double s = block_sum(acc[k], sh);
        if (threadIdx.x == 0) part[(long)blockIdx.x * kE_count + k] = s;
    }
}

// Terms, Coulomb energies into e[0..42] = [total, kin(2), E_C, E_Cx, com(2),
// terms 1..36]; the kinetic and c.o.m. slots are filled by k_energy_kinetic.
__global__ void k_energy_terms(SkyrmeK f, const double *__restrict__ part, int nblk, int trc,
                               double dv, double *__restrict__ e) {
    // one wave per slot: lanes sum strided partials, fixed shuffle tree
    __shared__ double S[kE_count];
    const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
    for (int slot = wv; slot < kE_count; slot += blockDim.x / 64) {
        double s = 0;
        for (int b = ln; b < nblk; b += 64) s += part[(long)b * kE_count + slot];
        s = wave_sum(s);
        if (ln == 0) S[slot] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double T[37];
    for (int k = 0; k < 37; ++k) T[k] = 0;
    T[1] = f.B1 * S[kE_rt2];
    T[2] = f.B2 * ((0.0 + S[kE_rn2]) + S[kE_rp2]);
    T[3] = f.B3 * S[kE_rttt];
    T[4] = (0.0 + f.B4 * S[kE_rntn]) + f.B4 * S[kE_rptp];
    T[7] = f.B5 * S[kE_rtlt];
    T[8] = (0.0 + f.B6 * S[kE_rnln]) + f.B6 * S[kE_rplp];
    T[9] = f.B7a * S[kE_rta2];
    T[10] = (0.0 + f.B8a * S[kE_rn2a]) + f.B8a * S[kE_rp2a];
    if (f.B7b != 0.0) T[33] = f.B7b * S[kE_rtb2];
    if (f.B8b != 0.0) T[34] = (0.0 + f.B8b * S[kE_rn2b]) + f.B8b * S[kE_rp2b];
    T[11] = f.B9 * S[kE_rtnj];
    T[12] = (0.0 + f.B9q * S[kE_rnnj]) + f.B9q * S[kE_rpnj];
    if (!trc) {
        T[5] = -f.B3 * S[kE_jt2];
        T[6] = (0.0 - f.B4 * S[kE_jn2]) - f.B4 * S[kE_jp2];
        T[13] = f.B9 * S[kE_jrst];
        T[14] = (0.0 + f.B9q * S[kE_jrsn]) + f.B9q * S[kE_jrsp];
        T[15] = f.B10 * S[kE_st2];
        T[16] = (0.0 + f.B11 * S[kE_sn2]) + f.B11 * S[kE_sp2];
        T[17] = f.B12a * S[kE_ast2];
        T[18] = (0.0 + f.B13a * S[kE_asn2]) + f.B13a * S[kE_asp2];
        if (f.B12b != 0.0) T[35] = f.B12b * S[kE_bst2];
        if (f.B13b != 0.0) T[36] = (0.0 + f.B13b * S[kE_bsn2]) + f.B13b * S[kE_bsp2];
    }
    for (int k = 1; k <= 36; ++k) e[6 + k] = T[k] * dv;
    e[3] = f.coulomb ? S[kE_coul] * dv * 0.5 : 0.0;
    e[4] = (f.coulomb && f.cexchange >= 1)
               ? -pow(3.0 / 3.14159265358979323846, 1.0 / 3.0) * f.e2 * 0.75 * dv * S[kE_cx]
               : 0.0;
}

// Kinetic energy (CompKinetic, Energy.f90:1426-1456) from the per-state sums
// K_i = dv <psi_i|Lap psi_i>, the one-body c.o.m. correction
// (CompCOMCorrection :1608-1660) and the total (CompEnergy :62-179).
__global__ __launch_bounds__(kBlock) void k_energy_kinetic(SkyrmeK f, const double *__restrict__ K,
                                                          const double *__restrict__ occ,
                                                          const int *__restrict__ qn, int M,
                                                          double *__restrict__ e) {
    __shared__ double sh[kBlock / 64];
    double kn = 0, kp = 0;
    for (int l = threadIdx.x; l < M; l += kBlock) {
        const double v = occ[l] * K[l];
        if (qn[4 * l] == -1) kn += v; else kp += v;
    }
    const double sn = block_sum(kn, sh);
    const double sp = block_sum(kp, sh);
    if (threadIdx.x != 0) return;
    double kin[2] = {sn, sp};
    double com[2] = {0, 0};
    for (int it = 0; it < 2; ++it) {
        kin[it] = -kin[it] / 2.0 * f.hbm[it];
        if (f.com1body > 0)
            com[it] = -kin[it] * f.mass[it] / (f.neutrons * f.mass[0] + f.protons * f.mass[1]);
    }
    double sky = 0;
    for (int k = 1; k <= 36; ++k) sky = sky + e[6 + k];
    e[1] = kin[0];
    e[2] = kin[1];
    e[5] = com[0];
    e[6] = com[1];
    e[0] = sky + (kin[0] + kin[1]) + e[3] + (com[0] + com[1]) + e[4] + (e[43] + e[44]);
}

void launch_energy_terms(const SkyrmeK &f, long N, int trc, const double *dens,
                         const double *derived, const double *vcoul, double *part, int nblk,
                         double dv, double *e, hipStream_t st) {
    hipLaunchKernelGGL(k_energy_partial, dim3(nblk), dim3(kBlock), 0, st, f, N, trc, dens,
                       derived, vcoul, part);
    hipLaunchKernelGGL(k_energy_terms, dim3(1), dim3(1024), 0, st, f, part, nblk, trc, dv, e);
}

void launch_energy_kinetic(const SkyrmeK &f, const double *K, const double *occ, const int *qn,
                           int M, double *e, hipStream_t st) {
    hipLaunchKernelGGL(k_energy_kinetic, dim3(1), dim3(kBlock), 0, st, f, K, occ, qn, M, e);
}

// ---------------------------------------------------------------------------
// Multipole constraints (Moments.f90).  Scalars of every constraint live in
// device records updated by one thread in the reference's order, so an
// iteration needs no host round trip; the fields (cutoff, constraint field
// in U, AlternateStep multipole) are point kernels.

// RutzCutOff (Moments.f90:2703-2849), step 1: per isospin threshold
// max(rho)/10 and the particle-number test; clears the surface counters.
__global__ __launch_bounds__(1024) void k_cut_stats(const double *__restrict__ dens, long N,
                                                   double dv, double *__restrict__ stat,
                                                   int *__restrict__ cnt) {
    __shared__ double smax[1024], ssum[16];
    const int it = blockIdx.x;
    const double *r = dens + (long)it * 9 * N;
    double mx = r[0], sm = 0;
    for (long p = threadIdx.x; p < N; p += 1024) {
        mx = fmax(mx, r[p]);
        sm += r[p];
    }
    smax[threadIdx.x] = mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sm += __shfl_xor(sm, o, 64);
    if (threadIdx.x % 64 == 0) ssum[threadIdx.x / 64] = sm;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 16; ++k) t += ssum[k];
        stat[2 * it] = smax[0] / 10.0;
        stat[2 * it + 1] = dv * t < 0.0000000001 ? 0.0 : 1.0;
        cnt[it] = 0;
    }
}

// Step 2: surface points between neighbours straddling the threshold, for
// the cells i < nx-1, j < ny-1, k < nz-1 (the reference's loop bounds).  The
// order of the points does not matter: only their minimum distance is used.
__global__ __launch_bounds__(kBlock) void k_cut_surface(const double *__restrict__ dens, int nx,
                                                       int ny, int nz, double dx,
                                                       const double *__restrict__ stat,
                                                       int *__restrict__ cnt,
                                                       double *__restrict__ surf) {
    const long N = (long)nx * ny * nz;
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int it = blockIdx.y;
    if (p >= N || stat[2 * it + 1] == 0.0) return;
    const int i = p % nx, j = (p / nx) % ny, k = p / ((long)nx * ny);
    if (i >= nx - 1 || j >= ny - 1 || k >= nz - 1) return;
    const double *r = dens + (long)it * 9 * N;
    const double tr = stat[2 * it];
    const double a = r[p];
    const long nb[3] = {p + 1, p + nx, p + (long)nx * ny};
    for (int ax = 0; ax < 3; ++ax) {
        const double b = r[nb[ax]];
        if ((a >= tr && b <= tr) || (a <= tr && b >= tr)) {
            double P[3] = {dx / 2.0 + dx * i, dx / 2.0 + dx * j, dx / 2.0 + dx * k};
            P[ax] = P[ax] + dx * (tr - a) / (b - a);
            const int slot = atomicAdd(&cnt[it], 1);
            double *o = surf + ((long)it * 3 * N + slot) * 3;
            o[0] = P[0];
            o[1] = P[1];
            o[2] = P[2];
        }
    }
}

// Step 3: signed distance to the nearest surface point and the Fermi
// function 1 / (1 + exp((DeltaR - radd) / acut)).
__global__ __launch_bounds__(kBlock) void k_cut_apply(const double *__restrict__ dens, int nx,
                                                     int ny, int nz, double dx, double radd,
                                                     double acut,
                                                     const double *__restrict__ stat,
                                                     const int *__restrict__ cnt,
                                                     const double *__restrict__ surf,
                                                     double *__restrict__ cut) {
    __shared__ double sp[3 * kBlock];
    const long N = (long)nx * ny * nz;
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int it = blockIdx.y;
    if (stat[2 * it + 1] == 0.0) {
        if (p < N) cut[(long)it * N + p] = 0.0;
        return;
    }
    const int T = cnt[it];
    const long pp = p < N ? p : 0;
    const int i = pp % nx, j = (pp / nx) % ny, k = pp / ((long)nx * ny);
    const double X = dx / 2.0 + dx * i, Y = dx / 2.0 + dx * j, Z = dx / 2.0 + dx * k;
    const double *su = surf + (long)it * 3 * N * 3;
    double dmin = 1e12;
    for (int t0 = 0; t0 < T; t0 += kBlock) {
        const int nt = min(kBlock, T - t0);
        __syncthreads();
        for (int q = threadIdx.x; q < 3 * nt; q += kBlock) sp[q] = su[(long)t0 * 3 + q];
        __syncthreads();
        for (int l = 0; l < nt; ++l) {
            const double ex = X - sp[3 * l], ey = Y - sp[3 * l + 1], ez = Z - sp[3 * l + 2];
            const double d = ex * ex + ey * ey + ez * ez;
            dmin = fmin(dmin, d);
        }
    }
    if (p >= N) return;
    const double sig = dens[(long)it * 9 * N + p] > stat[2 * it] ? -1.0 : 1.0;
    cut[(long)it * N + p] = 1.0 / (1.0 + exp((sig * sqrt(dmin) - radd) / acut));
}

void launch_rutz_cutoff(const double *dens, int nx, int ny, int nz, double dx, double dv,
                        double radd, double acut, double *stat, int *cnt, double *surf,
                        double *cut, hipStream_t st) {
    const long N = (long)nx * ny * nz;
    const dim3 g((N + kBlock - 1) / kBlock, 2);
    hipLaunchKernelGGL(k_cut_stats, dim3(2), dim3(1024), 0, st, dens, N, dv, stat, cnt);
    hipLaunchKernelGGL(k_cut_surface, g, dim3(kBlock), 0, st, dens, nx, ny, nz, dx, stat, cnt,
                       surf);
    hipLaunchKernelGGL(k_cut_apply, g, dim3(kBlock), 0, st, dens, nx, ny, nz, dx, radd, acut,
                       stat, cnt, surf, cut);
}

// <Q^2> = dv sum Y^2 rho per constraint and isospin (Calculate_Multipole,
// Moments.f90:1943-1947): one workgroup per (constraint, isospin).
__global__ __launch_bounds__(1024) void k_con_squared(const ConstraintK *__restrict__ con,
                                                     const double *__restrict__ Y,
                                                     const double *__restrict__ dens, long N,
                                                     double dv, double *__restrict__ sq) {
    __shared__ double sh[16];
    const int c = blockIdx.x, it = blockIdx.y;
    const double *y = Y + (long)con[c].idx * N, *r = dens + (long)it * 9 * N;
    double a = 0;
    for (long p = threadIdx.x; p < N; p += 1024) a += y[p] * y[p] * r[p];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 16; ++k) t += sh[k];
        sq[2 * c + it] = t * dv;
    }
}

// The scalar part, one thread, constraints in the moment list's order:
// kConCalc: first-call targets (CalculateAllMoments :2278-2301) and the
// coefficients of the constraint field (CalcConstraintEnergy :2368-2531);
// kConReadjust: intensity estimate and target update (Readjust
// :2929-3058); kConTurnOff: switch-off after the given iteration
// (TurnOffConstraints :3516-3549).  Always: the AlternateStep coefficients
// 0.5 (<Q> - Q_0) / (<Q^2> + d0) (ImaginaryTime.f90:338-377).
__global__ void k_con_update(ConstraintK *__restrict__ con, int nc, ConstraintCtl *ctl,
                             const double *__restrict__ Q, const double *__restrict__ sq,
                             const double *__restrict__ cstat, int mode, int iteration,
                             double momentum, double *__restrict__ ce_coef,
                             double *__restrict__ alt) {
    if (threadIdx.x != 0) return;
    const double d0 = (double)0.1f;      // a default-real literal (Moments.f90:258)
    if (mode & kConCalc) {
        if (ctl->first) {
            for (int c = 0; c < nc; ++c) {
                ConstraintK &k = con[c];
                if (k.type == 0) continue;
                const double v0 = Q[2 * k.idx], v1 = Q[2 * k.idx + 1];
                if (k.iso == 1) { k.con[0] = k.con[1] = v0 + v1; }
                else { k.con[0] = v0; k.con[1] = v1; }
            }
            ctl->first = 0;
        }
        const double damp = ctl->damping;
        ctl->ce_nz = ctl->ce_nz && damp != 0.0;
        for (int c = 0; c < nc; ++c) {
            ConstraintK &k = con[c];
            double fac[2], val[2], des[2];
            const double v0 = Q[2 * k.idx], v1 = Q[2 * k.idx + 1];
            if (k.type == 0) {
                ce_coef[2 * c] = ce_coef[2 * c + 1] = 0.0;
                continue;
            }
            if (k.iso == 1) {
                fac[0] = fac[1] = 2.0 * k.inten[0];
                val[0] = val[1] = v0 + v1;
                des[0] = des[1] = k.con[0];
            } else {
                fac[0] = 2.0 * k.inten[0]; fac[1] = 2.0 * k.inten[1];
                val[0] = v0; val[1] = v1;
                des[0] = k.con[0]; des[1] = k.con[1];
            }
            if (ctl->ce_nz) { fac[0] = fac[0] * (1.0 - damp); fac[1] = fac[1] * (1.0 - damp); }
            for (int it = 0; it < 2; ++it) {
                const double cf = fac[it] * (val[it] - des[it]);
                ce_coef[2 * c + it] = cf;
                if (cf != 0.0 && cstat[2 * it + 1] != 0.0) ctl->ce_nz = 1;
            }
        }
    }
    if (mode & kConReadjust) {
        for (int c = 0; c < nc; ++c) {
            ConstraintK &k = con[c];
            const double v0 = Q[2 * k.idx], v1 = Q[2 * k.idx + 1];
            const double s0 = sq[2 * c], s1 = sq[2 * c + 1];
            if (k.type == 1) {
                if (k.iso == 1) {
                    if (k.inten[0] == 0.0) {
                        const double o2 = s0 + s1;         // O2(1:2) = sum(Squared)
                        k.inten[0] = k.inten[1] = 2 / (o2 + o2);
                        ctl->slowdown = 2 * k.inten[0];
                    }
                    k.con[0] = k.con[1] = k.con[0] - ctl->slowdown * ((v0 + v1) - k.truec[0]);
                } else {
                    k.con[0] = k.con[0] - ctl->slowdown * (v0 - k.truec[0]);
                    k.con[1] = k.con[1] - ctl->slowdown * (v1 - k.truec[1]);
                }
            } else if (k.type == 2) {
                if (k.inten[0] == 0.0) {
                    if (k.iso == 1) {
                        k.inten[0] = k.inten[1] = 2 / (s0 + s1);
                        if (momentum != 0.0) k.inten[0] = k.inten[1] = 0.5 * k.inten[0];
                    } else {
                        k.inten[0] = 2 / s0;
                        k.inten[1] = 2 / s1;
                        if (momentum != 0.0) {
                            k.inten[0] = 0.5 * k.inten[0];
                            k.inten[1] = 0.5 * k.inten[1];
                        }
                    }
                }
                if (k.iso == 1) {
                    k.con[0] = k.con[1] = k.con[0] - ((v0 + v1) - k.truec[0]);
                } else {
                    k.con[0] = k.con[0] - (v0 - k.truec[0]);
                    k.con[1] = k.con[1] - (v1 - k.truec[1]);
                }
            }
        }
    }
    if (mode & kConTurnOff)
        for (int c = 0; c < nc; ++c) {
            ConstraintK &k = con[c];
            if (k.type != 0 && k.iteration > 0 && iteration > k.iteration) k.type = 0;
        }
    for (int c = 0; c < nc; ++c) {
        const ConstraintK &k = con[c];
        const double v0 = Q[2 * k.idx], v1 = Q[2 * k.idx + 1];
        const double o2 = sq[2 * c] + sq[2 * c + 1];
        for (int it = 0; it < 2; ++it) {
            double a = 0.0;
            if (k.type == 2) {
                const double v = k.iso == 1 ? v0 + v1 : (it == 0 ? v0 : v1);
                const double d = k.iso == 1 ? k.truec[0] : k.truec[it];
                a = 0.5 * (v - d) / (o2 + d0);
            }
            alt[2 * c + it] = a;
        }
    }
}

// ConstraintEnergy = Damping * ConstraintEnergy + sum_c coef_c Y_c cutoff.
__global__ __launch_bounds__(kBlock) void k_con_field(const ConstraintK *__restrict__ con,
                                                     int nc, const double *__restrict__ Y,
                                                     const double *__restrict__ coef,
                                                     const ConstraintCtl *__restrict__ ctl,
                                                     const double *__restrict__ cut, long N,
                                                     double *__restrict__ ce) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    const double damp = ctl->damping;
    for (int it = 0; it < 2; ++it) {
        double v = damp * ce[(long)it * N + p];
        const double cu = cut[(long)it * N + p];
        for (int c = 0; c < nc; ++c) v = v + coef[2 * c + it] * Y[(long)con[c].idx * N + p] * cu;
        ce[(long)it * N + p] = v;
    }
}

// AlternateStep's multipole field sum_c alt_c cutoff Y_c.
__global__ __launch_bounds__(kBlock) void k_alt_field(const ConstraintK *__restrict__ con,
                                                     int nc, const double *__restrict__ Y,
                                                     const double *__restrict__ alt,
                                                     const double *__restrict__ cut, long N,
                                                     double *__restrict__ mult) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    for (int it = 0; it < 2; ++it) {
        double v = 0.0;
        const double cu = cut[(long)it * N + p];
        for (int c = 0; c < nc; ++c) v = v + alt[2 * c + it] * cu * Y[(long)con[c].idx * N + p];
        mult[(long)it * N + p] = v;
    }
}

// psi <- psi - multipole(isospin) psi (ImaginaryTime.f90:415-426).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_alt_apply(const T *__restrict__ in,
                                                     T *__restrict__ out,
                                                     const int *__restrict__ qn,
                                                     const double *__restrict__ mult, long N) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int s = blockIdx.y;
    if (p >= N) return;
    const int it = (qn[4 * s] + 3) / 2 - 1;
    const double mv = mult[(long)it * N + p];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const long q = ((long)s * 4 + c) * N + p;
        const double v = (double)in[q];
        out[q] = T(v - mv * v);
    }
}

void launch_con_squared(const ConstraintK *con, int nc, const double *Y, const double *dens,
                        long N, double dv, double *sq, hipStream_t st) {
    hipLaunchKernelGGL(k_con_squared, dim3(nc, 2), dim3(1024), 0, st, con, Y, dens, N, dv, sq);
}

void launch_con_update(ConstraintK *con, int nc, ConstraintCtl *ctl, const double *Q,
                       const double *sq, const double *cstat, int mode, int iteration,
                       double momentum, double *ce_coef, double *alt, hipStream_t st) {
    hipLaunchKernelGGL(k_con_update, dim3(1), dim3(64), 0, st, con, nc, ctl, Q, sq, cstat, mode,
                       iteration, momentum, ce_coef, alt);
}

// The value history of each constrained moment (CalculateAllMoments shifts
// OldValue(:,1:7) and stores the new Value at every call, Moments.f90:
// 1920-1932; ConverMultipole compares sum(Value) and sum(OldValue(:,1:7))
// with the target, :2628-2682).  One thread per constraint.
__global__ void k_con_history(const ConstraintK *__restrict__ con, int nc,
                              const double *__restrict__ Q, double *__restrict__ hist,
                              int passes) {
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        const double v0 = Q[2 * con[c].idx], v1 = Q[2 * con[c].idx + 1];
        double *hc = hist + 16 * c;
        for (int p = 0; p < passes; ++p) {
            for (int i = 7; i > 0; --i) {
                hc[2 * i] = hc[2 * i - 2];
                hc[2 * i + 1] = hc[2 * i - 1];
            }
            hc[0] = v0;
            hc[1] = v1;
        }
    }
}

void launch_con_history(const ConstraintK *con, int nc, const double *Q, double *hist,
                        int passes, hipStream_t st) {
    hipLaunchKernelGGL(k_con_history, dim3(1), dim3(64), 0, st, con, nc, Q, hist, passes);
}

void launch_con_field(const ConstraintK *con, int nc, const double *Y, const double *coef,
                      const ConstraintCtl *ctl, const double *cut, long N, double *ce,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_con_field, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, st, con,
                       nc, Y, coef, ctl, cut, N, ce);
}

template <typename T>
void launch_alternate(const ConstraintK *con, int nc, const double *Y, const double *alt,
                      const double *cut, long N, double *mult, const T *in, T *out,
                      const int *qn, int M, hipStream_t st) {
    const unsigned gx = (unsigned)((N + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_alt_field, dim3(gx), dim3(kBlock), 0, st, con, nc, Y, alt, cut, N, mult);
    hipLaunchKernelGGL(k_alt_apply<T>, dim3(gx, M), dim3(kBlock), 0, st, in, out, qn, mult, N);
}
template void launch_alternate<double>(const ConstraintK *, int, const double *, const double *,
                                       const double *, long, double *, const double *,
                                       double *, const int *, int, hipStream_t);
template void launch_alternate<float>(const ConstraintK *, int, const double *, const double *,
                                      const double *, long, double *, const float *, float *,
                                      const int *, int, hipStream_t);

// ---------------------------------------------------------------------------
// IterativeEstimation (ImaginaryTime.f90:28-211): power iteration of one
// extra state (maxspwf) towards the largest eigenvalue of h, then dt and the
// heavy-ball momentum from that eigenvalue and the lowest excitation.
//
// Device record est[kEstSlots]: est[0] = maxE of the last executed power
// iteration, est[1] = con = previous maxE - maxE, est[2] = iterations
// executed, est[3] = iteration at which |con| < tol was met (0 while
// running), est[4..7] = dt, momentum, relE, maxE - minh from k_est_finalize.
// The host enqueues power iterations in batches; every kernel of iteration
// k returns at once when est[3] holds an earlier iteration, so a batch that
// runs past convergence leaves the state as it was at convergence.

// Counter-based uniform [0,1) numbers (splitmix64 of seed + index), the
// start vector of the first call (random_number in the reference).
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long i) {
    unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_est_reset(T *__restrict__ psi, T *__restrict__ u,
                                                     long n, int fill, unsigned long long seed,
                                                     double *__restrict__ est) {
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        if (fill) psi[i] = T(uniform01(seed, (unsigned long long)i));
        u[i] = T(0);
    }
    if (blockIdx.x == 0 && threadIdx.x < kEstSlots)
        est[threadIdx.x] = threadIdx.x == 1 ? 1e300 : 0.0;
}

// One power iteration, first half: maxE = dv <psi|h psi> from the hPsi
// partial sums, u = dth (h psi - maxE psi) + mu u, psi += u, and the
// workgroup's part of <psi|psi>.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_est_update(
    T *__restrict__ psi, const T *__restrict__ hpsi, T *__restrict__ u, long n,
    const double *__restrict__ eparts, int slots, double dv, double dth, double mu, int iter,
    int first, double *__restrict__ est, double *__restrict__ npart) {
    __shared__ double sh[kBlock / 64];
    const double stop = est[3];
    if (stop != 0.0 && stop < iter) return;
    double a = 0;
    for (int t = threadIdx.x; t < slots; t += kBlock) a += eparts[3 * t];
    const double maxE = block_sum(a, sh) * dv;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        est[1] = first ? 1e300 : est[0] - maxE;
        est[0] = maxE;
    }
    double nn = 0;
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        const double v = (double)psi[i];
        const double w = dth * ((double)hpsi[i] - maxE * v) + mu * (double)u[i];
        const T nv = T(v + w);
        u[i] = T(w);
        psi[i] = nv;
        nn += (double)nv * (double)nv;
    }
    nn = block_sum(nn, sh);
    if (threadIdx.x == 0) npart[blockIdx.x] = nn;
}

// Second half: psi /= sqrt(dv <psi|psi>); records the iteration and, when
// |con| < tol, the stop.  With n4 partials from k_est_sumsq (iter 0) it is
// the normalisation of the start vector.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_est_normalize(T *__restrict__ psi, long n,
                                                         const double *__restrict__ npart,
                                                         int nblk, double dv, double tol,
                                                         int iter, double *__restrict__ est) {
    __shared__ double sh[kBlock / 64];
    const double stop = est[3];
    if (stop != 0.0 && stop < iter) return;
    double a = 0;
    for (int t = threadIdx.x; t < nblk; t += kBlock) a += npart[t];
    const double scale = 1.0 / sqrt(block_sum(a, sh) * dv);
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        psi[i] = T((double)psi[i] * scale);
    if (blockIdx.x == 0 && threadIdx.x == 0 && iter > 0) {
        est[2] = iter;
        if (fabs(est[1]) < tol) est[3] = iter;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_est_sumsq(const T *__restrict__ psi, long n,
                                                     double *__restrict__ npart) {
    __shared__ double sh[kBlock / 64];
    double nn = 0;
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        nn += (double)psi[i] * (double)psi[i];
    nn = block_sum(nn, sh);
    if (threadIdx.x == 0) npart[blockIdx.x] = nn;
}

// relE, minh and the parameters (ImaginaryTime.f90:151-209).  HF (pairing
// 0): relE = min over occupied i, empty ii of the same isospin of
// E_ii - E_i > 0, at least 0.1; HFB (pairing 2): relE = min of the positive
// quasiparticle energies qpe[0..nqp).  dt uses the single-precision 0.90 of
// the reference (0.899999976158142...).
__global__ __launch_bounds__(kBlock) void k_est_finalize(const double *__restrict__ E,
                                                        const double *__restrict__ occ,
                                                        const int *__restrict__ qn, int M,
                                                        int pairing, const double *__restrict__ qpe,
                                                        int nqp, double hbar,
                                                        double *__restrict__ est) {
    __shared__ double sr[kBlock], sm[kBlock];
    double rel = 10000.0, mh = 10000.0;
    if (pairing == 0) {
        for (int i = 0; i < M; ++i) {
            if (fabs(occ[i]) < 0.5) continue;
            const int iso = qn[4 * i];
            for (int ii = threadIdx.x; ii < M; ii += kBlock) {
                if (fabs(occ[ii]) > 0.5 || qn[4 * ii] != iso) continue;
                const double c = -E[i] + E[ii];
                if (c > 0.0) rel = fmin(rel, c);
            }
        }
    } else {
        for (int i = threadIdx.x; i < nqp; i += kBlock)
            if (qpe[i] > 0.0) rel = fmin(rel, fabs(qpe[i]));
    }
    for (int i = threadIdx.x; i < M; i += kBlock) mh = fmin(mh, E[i]);
    sr[threadIdx.x] = rel;
    sm[threadIdx.x] = mh;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmin(sr[threadIdx.x], sr[threadIdx.x + o]);
            sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double relE = sr[0];
    if (pairing == 0 && relE < 0.1) relE = 0.1;
    const double maxE = est[0] - sm[0];
    const double kappa = relE / maxE;
    const double sk = sqrt(kappa);
    const double m = ((sk - 1) / (sk + 1)) * ((sk - 1) / (sk + 1));
    est[4] = 4.0 / (maxE + relE + 2 * sqrt(maxE * relE)) * hbar * (double)0.90f;
    est[5] = m;
    est[6] = relE;
    est[7] = maxE;
}

template <typename T>
void launch_est_reset(T *psi, T *u, long n, int fill, unsigned long long seed, double *est,
                      int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_est_reset<T>, dim3(nblk), dim3(kBlock), 0, st, psi, u, n, fill, seed,
                       est);
}

template <typename T>
void launch_est_start(T *psi, long n, double *npart, int nblk, double dv, double *est,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_est_sumsq<T>, dim3(nblk), dim3(kBlock), 0, st, psi, n, npart);
    hipLaunchKernelGGL(k_est_normalize<T>, dim3(nblk), dim3(kBlock), 0, st, psi, n, npart, nblk,
                       dv, 0.0, 0, est);
}

template <typename T>
void launch_est_iteration(T *psi, const T *hpsi, T *u, long n, const double *eparts, int slots,
                          double dv, double dth, double mu, int iter, int first, double tol,
                          double *est, double *npart, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_est_update<T>, dim3(nblk), dim3(kBlock), 0, st, psi, hpsi, u, n, eparts,
                       slots, dv, dth, mu, iter, first, est, npart);
    hipLaunchKernelGGL(k_est_normalize<T>, dim3(nblk), dim3(kBlock), 0, st, psi, n, npart, nblk,
                       dv, tol, iter, est);
}

void launch_est_finalize(const double *E, const double *occ, const int *qn, int M, int pairing,
                         const double *qpe, int nqp, double hbar, double *est, hipStream_t st) {
    hipLaunchKernelGGL(k_est_finalize, dim3(1), dim3(kBlock), 0, st, E, occ, qn, M, pairing, qpe,
                       nqp, hbar, est);
}

#define MOCCA_EST_INST(T)                                                                      \
    template void launch_est_reset<T>(T *, T *, long, int, unsigned long long, double *, int,  \
                                      hipStream_t);                                            \
    template void launch_est_start<T>(T *, long, double *, int, double, double *, hipStream_t); \
    template void launch_est_iteration<T>(T *, const T *, T *, long, const double *, int,      \
                                          double, double, double, int, int, double, double *,  \
                                          double *, int, hipStream_t);
MOCCA_EST_INST(double)
MOCCA_EST_INST(float)
#undef MOCCA_EST_INST

}  // namespace mocca
