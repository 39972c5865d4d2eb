// HIP kernels of the mocca_amd library (gfx950 / CDNA4, wave64).
//
// Reference routines each kernel replaces are cited as path:line in the
// reference MOCCa tree.
#include "mocca_kernels.h"

#include <algorithm>
#include <type_traits>

namespace mocca {

namespace {
thread_local const char *g_launch_failure = nullptr;
}

bool require_lds(const void *f, size_t bytes, const char *kernel) {
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
        hipSuccess)
        return true;
    g_launch_failure = kernel;
    return false;
}

const char *take_launch_failure() {
    const char *k = g_launch_failure;
    g_launch_failure = nullptr;
    return k;
}

// ---------------------------------------------------------------------------
// Sign tables of the EV8 octant (OptimizedDerivatives.f90:24-44,115-132,
// 208-225,340-357 for first derivatives; :522-571 for the Laplacian).
__device__ __forceinline__ int sign_first(int axis, int P, int S, int TS,
                                          int c, int lag) {
    const int cx = (c == 0 || c == 3) ? 1 : -1;
    const int cy = (c == 0 || c == 2) ? 1 : -1;
    const int cz = (c < 2) ? 1 : -1;
    if (axis == 0) return cx * (lag ? S * TS : S);
    if (axis == 1) return cy * TS;
    return cz * S * P;
}

__device__ __forceinline__ int sign_lap(int axis, int P, int S, int TS, int c) {
    const int cx = (c == 0 || c == 3) ? 1 : -1;
    const int cy = (c == 0 || c == 2) ? 1 : -1;
    const int cz = (c < 2) ? 1 : -1;
    if (axis == 0) return cx * S;
    if (axis == 1) return cy * TS;
    return cz * S * P;
}

// Line value with the reflection ghost below the symmetry plane and zero
// beyond the far box edge (derx/lapla_X_ev8 boundary rows).
template <typename T>
__device__ __forceinline__ T line_at(const T *__restrict__ line, int j, int n,
                                     int stride, T s) {
    if (j < 0) return s * line[(-1 - j) * stride];
    if (j >= n) return T(0);
    return line[j * stride];
}

// First derivative d1 (sign s1) and second derivative d2 (sign s2) along one
// axis at the point (x, y, z) of field f.
template <typename T>
__device__ __forceinline__ void axis_derivs(const MeshK<T> &m,
                                            const T *__restrict__ f, int x,
                                            int y, int z, int axis, int s1,
                                            int s2, T &d1, T &d2, bool need2) {
    int n, i, stride;
    long base;
    if (axis == 0) {
        n = m.nx; i = x; stride = 1; base = ((long)z * m.ny + y) * m.nx;
    } else if (axis == 1) {
        n = m.ny; i = y; stride = m.nx; base = (long)z * m.ny * m.nx + x;
    } else {
        n = m.nz; i = z; stride = m.nx * m.ny; base = (long)y * m.nx + x;
    }
    const T *line = f + base;
    if (!m.lag) {
        const T s = s2 > 0 ? T(1) : T(-1);
        T v[9];
#pragma unroll
        for (int k = -4; k <= 4; ++k) v[k + 4] = line_at(line, i + k, n, stride, s);
        if (need2)
            d2 = lap9(m, v);
        if (s1 != s2) {
            const T s1t = s1 > 0 ? T(1) : T(-1);
#pragma unroll
            for (int k = -3; k <= 3; ++k) v[k + 4] = line_at(line, i + k, n, stride, s1t);
        }
        d1 = (T(45) * (v[5] - v[3]) - T(9) * (v[6] - v[2]) + (v[7] - v[1])) * m.inv60dx;
    } else {
        const T *M1 = m.lag1[axis][s1 > 0 ? 1 : 0] + (long)i * n;
        const T *M2 = m.lag2[axis][s2 > 0 ? 1 : 0] + (long)i * n;
        T a1 = 0, a2 = 0;
        for (int j = 0; j < n; ++j) {
            const T w = line[j * stride];
            a1 += M1[j] * w;
            if (need2) a2 += M2[j] * w;
        }
        d1 = a1;
        d2 = a2;
    }
}

// Value, gradient and Laplacian of all four components of state `s` at a
// point (CompDer, Spwf.f90:778-792).
template <typename T>
__device__ __forceinline__ void state_derivs(const MeshK<T> &m,
                                             const T *__restrict__ ps, int P,
                                             int S, int TS, int x, int y, int z,
                                             long p, T v[4], T d[3][4], T l[4],
                                             bool need_lap) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const T *f = ps + (long)c * m.N;
        v[c] = f[p];
        T lap = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T d1, d2 = 0;
            axis_derivs(m, f, x, y, z, a, sign_first(a, P, S, TS, c, m.lag),
                        sign_lap(a, P, S, TS, c), d1, d2, need_lap);
            d[a][c] = d1;
            lap += d2;
        }
        l[c] = lap;
    }
}

// ---------------------------------------------------------------------------
// DeriveAll (SpwfStorage.f90:586-604): materialised derivatives.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_derive_all(MeshK<T> m,
                                                       const T *__restrict__ psi,
                                                       const int *__restrict__ qn,
                                                       double *__restrict__ der,
                                                       double *__restrict__ lap) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int s = blockIdx.y;
    if (p >= m.N) return;
    const int x = p % m.nx, y = (p / m.nx) % m.ny, z = p / ((long)m.nx * m.ny);
    T v[4], d[3][4], l[4];
    state_derivs(m, psi + (long)s * 4 * m.N, qn[4 * s + 1], qn[4 * s + 2],
                 qn[4 * s + 3], x, y, z, p, v, d, l, true);
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 4; ++c)
            der[(((long)s * 3 + a) * 4 + c) * m.N + p] = (double)d[a][c];
    for (int c = 0; c < 4; ++c) lap[((long)s * 4 + c) * m.N + p] = (double)l[c];
}

template <typename T>
void launch_derive_all(const MeshK<T> &m, const T *psi, const int *qn, int M,
                       double *der, double *lap, hipStream_t st) {
    dim3 g((m.N + kBlock - 1) / kBlock, M);
    hipLaunchKernelGGL(k_derive_all<T>, g, dim3(kBlock), 0, st, m, psi, qn, der, lap);
}

// ---------------------------------------------------------------------------
// DiagSpin and DiagAng (Spwf.f90:977-1489) of every state: <S>, <S T>,
// <J>, <J T> and <J_i^2> from the values, first derivatives and coordinates
// (x, y, z) = ((i + 1/2) dx, ...) of the EV8 octant.  Signature and time
// simplex are conserved there, which zeroes J_x, J_y, (J T)_x imaginary,
// (J T)_y real and the z parts of J T; the nonzero sums are, per state,
// kAmSums partial sums per workgroup: J_z, Re (J T)_x, Im (J T)_y,
// <J_x^2>, <J_y^2>, <J_z^2>, S_z, Re (S T)_x, Im (S T)_y.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_angmom(MeshK<T> m, const T *__restrict__ psi,
                                                   const int *__restrict__ qn, double dx,
                                                   double *__restrict__ part) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int s = blockIdx.y;
    double acc[kAmSums];
#pragma unroll
    for (int k = 0; k < kAmSums; ++k) acc[k] = 0.0;
    if (p < m.N) {
        const int x = p % m.nx, y = (p / m.nx) % m.ny, z = p / ((long)m.nx * m.ny);
        T vt[4], dt[3][4], lt[4];
        state_derivs(m, psi + (long)s * 4 * m.N, qn[4 * s + 1], qn[4 * s + 2], qn[4 * s + 3], x,
                     y, z, p, vt, dt, lt, false);
        double v[4], dX[4], dY[4], dZ[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            v[c] = (double)vt[c];
            dX[c] = (double)dt[0][c];
            dY[c] = (double)dt[1][c];
            dZ[c] = (double)dt[2][c];
        }
        const double X = (x + 0.5) * dx, Y = (y + 0.5) * dx, Z = (z + 0.5) * dx;
        // spin parts
        const double sz = 0.5 * (v[0] * v[0] + v[1] * v[1] - v[2] * v[2] - v[3] * v[3]);
        const double stx = 0.5 * (-v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - v[3] * v[3]);
        const double sty = 0.5 * (-v[0] * v[0] + v[1] * v[1] - v[2] * v[2] + v[3] * v[3]);
        // orbital parts: l = -i r x grad acting on the spinor, with the
        // component pairing of the reference's expressions
        const double lz = v[1] * Y * dX[0] - v[1] * X * dY[0] - v[0] * Y * dX[1] +
                          v[0] * X * dY[1] + v[3] * Y * dX[2] - v[3] * X * dY[2] -
                          v[2] * Y * dX[3] + v[2] * X * dY[3];
        const double ltx = v[1] * Z * dY[2] - v[1] * Y * dZ[2] + v[0] * Z * dY[3] -
                           v[0] * Y * dZ[3] - v[3] * Z * dY[0] + v[3] * Y * dZ[0] -
                           v[2] * Z * dY[1] + v[2] * Y * dZ[1];
        const double lty = v[0] * Z * dX[2] - v[0] * X * dZ[2] - v[1] * Z * dX[3] +
                           v[1] * X * dZ[3] - v[2] * Z * dX[0] + v[2] * X * dZ[0] +
                           v[3] * Z * dX[1] - v[3] * X * dZ[1];
        // <J_i^2> = |J_i psi|^2, J_i psi component by component
        const double a1 = -Z * dY[1] + Y * dZ[1] + 0.5 * v[2];
        const double a2 = Z * dY[0] - Y * dZ[0] + 0.5 * v[3];
        const double a3 = -Z * dY[3] + Y * dZ[3] + 0.5 * v[0];
        const double a4 = Z * dY[2] - Y * dZ[2] + 0.5 * v[1];
        const double b1 = -X * dZ[1] + Z * dX[1] + 0.5 * v[3];
        const double b2 = X * dZ[0] - Z * dX[0] - 0.5 * v[2];
        const double b3 = -X * dZ[3] + Z * dX[3] - 0.5 * v[1];
        const double b4 = X * dZ[2] - Z * dX[2] + 0.5 * v[0];
        const double c1 = -Y * dX[1] + X * dY[1] + 0.5 * v[0];
        const double c2 = Y * dX[0] - X * dY[0] + 0.5 * v[1];
        const double c3 = -Y * dX[3] + X * dY[3] - 0.5 * v[2];
        const double c4 = Y * dX[2] - X * dY[2] - 0.5 * v[3];
        acc[0] = sz + lz;
        acc[1] = stx + ltx;
        acc[2] = sty + lty;
        acc[3] = a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4;
        acc[4] = b1 * b1 + b2 * b2 + b3 * b3 + b4 * b4;
        acc[5] = c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4;
        acc[6] = sz;
        acc[7] = stx;
        acc[8] = sty;
    }
    __shared__ double sh[kBlock / 64][kAmSums];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < kAmSums; ++k) {
        double a = acc[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if (l == 0) sh[w][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < kAmSums) {
        double t = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) t += sh[q][threadIdx.x];
        part[((long)s * gridDim.x + blockIdx.x) * kAmSums + threadIdx.x] = t;
    }
}

// The sums over the workgroups, in their order, times dv, into the layout
// J(3) JTR(3) JTI(3) J2(3) Sp(3) SpTR(3) SpTI(3) of each state.
__global__ void k_angmom_reduce(const double *__restrict__ part, int nblk, int M, double dv,
                                double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * kAmSums) return;
    const int s = t / kAmSums, k = t % kAmSums;
    double a = 0.0;
    for (int b = 0; b < nblk; ++b) a += part[((long)s * nblk + b) * kAmSums + k];
    // slot of each sum in the 21-value record
    const int slot[kAmSums] = {2, 3, 7, 9, 10, 11, 14, 15, 19};
    out[(long)s * 21 + slot[k]] = a * dv;
}

// Integrals of the spin and current densities (UpdateAM, Cranking.f90:
// 186-212), z component: sum over isospins of 1/2 s_z - y j_x + x j_y, per
// workgroup; dens = [2][9][N] with j at fields 3..5 and s at 6..8.
__global__ __launch_bounds__(kBlock) void k_angmom_dens(const double *__restrict__ dens,
                                                        int nx, int ny, long N, double dx,
                                                        double *__restrict__ part) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    double a = 0.0;
    if (p < N) {
        const int x = p % nx, y = (p / nx) % ny;
        const double X = (x + 0.5) * dx, Y = (y + 0.5) * dx;
        for (int it = 0; it < 2; ++it) {
            const double *d = dens + (long)it * 9 * N;
            a += 0.5 * d[8 * N + p];
            a += -Y * d[3 * N + p] + X * d[4 * N + p];
        }
    }
    __shared__ double sh[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) t += sh[q];
        part[blockIdx.x] = t;
    }
}

template <typename T>
void launch_angmom(const MeshK<T> &m, const T *psi, const int *qn, int M, double dx, double dv,
                   double *part, double *out, hipStream_t st) {
    const int nblk = (int)((m.N + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_angmom<T>, dim3(nblk, M), dim3(kBlock), 0, st, m, psi, qn, dx, part);
    const int nt = M * kAmSums;
    hipLaunchKernelGGL(k_angmom_reduce, dim3((nt + 255) / 256), dim3(256), 0, st, part, nblk, M,
                       dv, out);
}
template void launch_angmom<double>(const MeshK<double> &, const double *, const int *, int,
                                    double, double, double *, double *, hipStream_t);
template void launch_angmom<float>(const MeshK<float> &, const float *, const int *, int,
                                   double, double, double *, double *, hipStream_t);

void launch_angmom_dens(const double *dens, int nx, int ny, long N, double dx, double *part,
                        hipStream_t st) {
    const int nblk = (int)((N + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_angmom_dens, dim3(nblk), dim3(kBlock), 0, st, dens, nx, ny, N, dx, part);
}

// ---------------------------------------------------------------------------
// Block-wide sum of two values (wave64 shuffles, then LDS over the waves).
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { sh[3 * w] = a; sh[3 * w + 1] = b; sh[3 * w + 2] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0, sc = 0;
        for (int k = 0; k < kBlock / 64; ++k) { sa += sh[3 * k]; sb += sh[3 * k + 1]; sc += sh[3 * k + 2]; }
        a = sa; b = sb; c = sc;
    }
    __syncthreads();
}

// FP64 matrix-core tile: D[16x16] += A[16x4] B[4x16].
typedef double mfma_d4 __attribute__((ext_vector_type(4)));


__device__ __forceinline__ mfma_d4 mfma16x16x4(double a, double b, mfma_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Lagrange derivatives on the FP64 matrix cores (Opt_LAG{X,Y,Z}_EV8,
// Lapla_LAG_EV8: OptimizedDerivatives.f90:115-148,306-338,432-464,551-615).
// Along each axis the derivative of every line is a small GEMM with the
// folded n x n matrix of the component's reflection sign: D = L Lambda, with
// Lambda[k][p] the value at position k of line ("pencil") p.  A workgroup
// (4 waves) owns one (state, component, axis): it keeps L1 and L2 (first and
// second derivative) transposed in LDS and walks blocks of 16 pencils, each
// staged in the wave's own LDS slice.  v_mfma_f64_16x16x4f64: lane l holds
// A[l%16][l/16] and B[l/16][l%16]; the result lane l, register r holds
// D[l/16 + 4r][l%16].  For y and z a pencil block is 16 consecutive x (so the
// results are stored 16 consecutive values at a time) and D = L Lambda; for x
// a block is 16 rows and D^T = Lambda^T L^T, which stores along x as well.
// Output: dsc[state][comp][6][N] = d/dx, d/dy, d/dz, d2/dx2, d2/dy2, d2/dz2.
template <typename T, int NPAD, int AXIS>
__global__ __launch_bounds__(256) void k_lag_axis(MeshK<double> md, const T *__restrict__ psi,
                                                  const int *__restrict__ qn, int M,
                                                  T *__restrict__ dsc) {
    constexpr int NT = NPAD / 16, NKS = NPAD / 4, NEL = NPAD / 4;
    const int s = blockIdx.x / 4, c = blockIdx.x % 4;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int nx = md.nx, ny = md.ny, nz = md.nz;
    const long N = md.N;
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    extern __shared__ double lag_lds[];
    double *Lt = lag_lds;                                 // [2][NPAD][NPAD]: L^T (k rows)
    double *Lam = lag_lds + 2 * NPAD * NPAD + wave * NPAD * 16;   // [NPAD][16]
    const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
    const int s1 = sign_first(AXIS, P, S, TS, c, 1), s2 = sign_lap(AXIS, P, S, TS, c);
    const double *L1 = md.lag1[AXIS][s1 > 0 ? 1 : 0], *L2 = md.lag2[AXIS][s2 > 0 ? 1 : 0];
    for (int q = threadIdx.x; q < 2 * NPAD * NPAD; q += blockDim.x) {
        const int mt = q / (NPAD * NPAD), r = q % (NPAD * NPAD);
        const int k = r / NPAD, i = r % NPAD;
        Lt[q] = (i < n && k < n) ? (mt ? L2 : L1)[i * n + k] : 0.0;
    }
    __syncthreads();
    const T *ps = psi + ((long)s * 4 + c) * N;
    T *d1 = dsc + (((long)s * 4 + c) * 6 + AXIS) * N;
    T *d2 = dsc + (((long)s * 4 + c) * 6 + 3 + AXIS) * N;
    const int nxb = (nx + 15) / 16;
    const int nblk = AXIS == 0 ? (ny * nz + 15) / 16 : nxb * (AXIS == 1 ? nz : ny);
    const int kr = lane / 16, col = lane % 16;
    // geometry of pencil block b: x pencils are rows 16 b .. 16 b + 15; y and
    // z pencils are 16 consecutive x at (z, or y) = outer
    auto geom = [&](int b, int &outer, int &x0, int &npen) {
        outer = 0; x0 = 0;
        if (AXIS == 0) {
            npen = min(16, ny * nz - 16 * b);
        } else {
            outer = b / nxb;
            x0 = (b % nxb) * 16;
            npen = min(16, nx - x0);
        }
    };
    // element j of the lane in a block: Lam index e = lane + 64 j, position
    // k and pencil p; the next block is loaded into registers while the
    // MFMAs of the current one run
    auto gather = [&](int b, double (&pf)[NEL]) {
        int outer, x0, npen;
        geom(b, outer, x0, npen);
#pragma unroll
        for (int j = 0; j < NEL; ++j) {
            const int e = lane + 64 * j;
            int k, p;
            long a;
            if (AXIS == 0) { p = e / NPAD; k = e % NPAD; a = (long)(16 * b + p) * nx + k; }
            else {
                k = e / 16; p = e % 16;
                a = AXIS == 1 ? ((long)outer * ny + k) * nx + x0 + p
                              : ((long)k * ny + outer) * nx + x0 + p;
            }
            const bool ok = k < n && p < npen;
            const T v = ps[ok ? a : 0];
            pf[j] = ok ? (double)v : 0.0;
        }
    };
    auto put = [&](const double (&pf)[NEL]) {
#pragma unroll
        for (int j = 0; j < NEL; ++j) {
            const int e = lane + 64 * j;
            const int k = AXIS == 0 ? e % NPAD : e / 16, p = AXIS == 0 ? e / NPAD : e % 16;
            Lam[k * 16 + p] = pf[j];
        }
    };
    double pf[NEL];
    if (wave < nblk) {
        gather(wave, pf);
        put(pf);
    }
    for (int blk = wave; blk < nblk; blk += 4) {
        const int nb = blk + 4;
        if (nb < nblk) gather(nb, pf);
        int outer, x0, npen;
        geom(blk, outer, x0, npen);
        // (the wave reads only its own slice: its LDS writes complete in order)
        __builtin_amdgcn_s_waitcnt(0xc07f);
        mfma_d4 acc[2][NT];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[mt][t] = mfma_d4{0, 0, 0, 0};
#pragma unroll 4
        for (int ks = 0; ks < NKS; ++ks) {
            const int k = 4 * ks + kr;
            const double lam = Lam[k * 16 + col];
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const double ell = Lt[(mt * NPAD + k) * NPAD + 16 * t + col];
                    acc[mt][t] = AXIS == 0 ? mfma16x16x4(lam, ell, acc[mt][t])
                                           : mfma16x16x4(ell, lam, acc[mt][t]);
                }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i, p;
                if (AXIS == 0) { p = kr + 4 * r; i = 16 * t + col; }
                else { i = 16 * t + kr + 4 * r; p = col; }
                if (i < n && p < npen) {
                    long a;
                    if (AXIS == 0) a = (long)(16 * blk + p) * nx + i;
                    else if (AXIS == 1) a = ((long)outer * ny + i) * nx + x0 + p;
                    else a = ((long)i * ny + outer) * nx + x0 + p;
                    d1[a] = (T)acc[0][t][r];
                    d2[a] = (T)acc[1][t][r];
                }
            }
        if (nb < nblk) put(pf);
    }
}

// The same derivatives for FP32 states on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32: operands as above, result lane l register r =
// D[4 (l/16) + r][l%16]) with the FP64 Lagrange matrices split into two FP32
// parts, L = L_hi + L_lo (L_hi = fl32(L), L_lo = fl32(L - L_hi)), each
// product accumulated in its own FP32 accumulator and the two added at the
// end: the matrices keep ~48 bits, the sums are FP32.
typedef float lag_f4 __attribute__((ext_vector_type(4)));
template <int NPAD, int AXIS>
__global__ __launch_bounds__(256) void k_lag_axis_split(MeshK<double> md,
                                                        const float *__restrict__ psi,
                                                        const int *__restrict__ qn, int M,
                                                        float *__restrict__ dsc) {
    constexpr int NT = NPAD / 16, NKS = NPAD / 4, NEL = NPAD / 4;
    const int s = blockIdx.x / 4, c = blockIdx.x % 4;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int nx = md.nx, ny = md.ny, nz = md.nz;
    const long N = md.N;
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    extern __shared__ float lags_lds[];
    float *Lt = lags_lds;                                  // [4][NPAD][NPAD]: L1 hi, lo, L2 hi, lo
    float *Lam = lags_lds + 4 * NPAD * NPAD + wave * NPAD * 16;   // [NPAD][16]
    const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
    const int s1 = sign_first(AXIS, P, S, TS, c, 1), s2 = sign_lap(AXIS, P, S, TS, c);
    const double *L1 = md.lag1[AXIS][s1 > 0 ? 1 : 0], *L2 = md.lag2[AXIS][s2 > 0 ? 1 : 0];
    for (int q = threadIdx.x; q < 2 * NPAD * NPAD; q += blockDim.x) {
        const int mt = q / (NPAD * NPAD), r = q % (NPAD * NPAD);
        const int k = r / NPAD, i = r % NPAD;
        const double v = (i < n && k < n) ? (mt ? L2 : L1)[i * n + k] : 0.0;
        const float hi = (float)v;
        Lt[(2 * mt) * NPAD * NPAD + r] = hi;
        Lt[(2 * mt + 1) * NPAD * NPAD + r] = (float)(v - (double)hi);
    }
    __syncthreads();
    const float *ps = psi + ((long)s * 4 + c) * N;
    float *d1 = dsc + (((long)s * 4 + c) * 6 + AXIS) * N;
    float *d2 = dsc + (((long)s * 4 + c) * 6 + 3 + AXIS) * N;
    const int nxb = (nx + 15) / 16;
    const int nblk = AXIS == 0 ? (ny * nz + 15) / 16 : nxb * (AXIS == 1 ? nz : ny);
    const int kr = lane / 16, col = lane % 16;
    auto geom = [&](int b, int &outer, int &x0, int &npen) {
        outer = 0; x0 = 0;
        if (AXIS == 0) {
            npen = min(16, ny * nz - 16 * b);
        } else {
            outer = b / nxb;
            x0 = (b % nxb) * 16;
            npen = min(16, nx - x0);
        }
    };
    auto gather = [&](int b, float (&pf)[NEL]) {
        int outer, x0, npen;
        geom(b, outer, x0, npen);
#pragma unroll
        for (int j = 0; j < NEL; ++j) {
            const int e = lane + 64 * j;
            int k, p;
            long a;
            if (AXIS == 0) { p = e / NPAD; k = e % NPAD; a = (long)(16 * b + p) * nx + k; }
            else {
                k = e / 16; p = e % 16;
                a = AXIS == 1 ? ((long)outer * ny + k) * nx + x0 + p
                              : ((long)k * ny + outer) * nx + x0 + p;
            }
            const bool ok = k < n && p < npen;
            const float v = ps[ok ? a : 0];
            pf[j] = ok ? v : 0.0f;
        }
    };
    auto put = [&](const float (&pf)[NEL]) {
#pragma unroll
        for (int j = 0; j < NEL; ++j) {
            const int e = lane + 64 * j;
            const int k = AXIS == 0 ? e % NPAD : e / 16, p = AXIS == 0 ? e / NPAD : e % 16;
            Lam[k * 16 + p] = pf[j];
        }
    };
    float pf[NEL];
    if (wave < nblk) {
        gather(wave, pf);
        put(pf);
    }
    for (int blk = wave; blk < nblk; blk += 4) {
        const int nb = blk + 4;
        if (nb < nblk) gather(nb, pf);
        int outer, x0, npen;
        geom(blk, outer, x0, npen);
        // (the wave reads only its own slice: its LDS writes complete in order)
        __builtin_amdgcn_s_waitcnt(0xc07f);
        lag_f4 acc[2][2][NT];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[mt][h][t] = lag_f4{0, 0, 0, 0};
#pragma unroll 4
        for (int ks = 0; ks < NKS; ++ks) {
            const int k = 4 * ks + kr;
            const float lam = Lam[k * 16 + col];
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const float ell = Lt[((2 * mt + h) * NPAD + k) * NPAD + 16 * t + col];
                        acc[mt][h][t] =
                            AXIS == 0
                                ? __builtin_amdgcn_mfma_f32_16x16x4f32(lam, ell, acc[mt][h][t], 0, 0, 0)
                                : __builtin_amdgcn_mfma_f32_16x16x4f32(ell, lam, acc[mt][h][t], 0, 0, 0);
                    }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int i, p;
                if (AXIS == 0) { p = 4 * kr + r; i = 16 * t + col; }
                else { i = 16 * t + 4 * kr + r; p = col; }
                if (i < n && p < npen) {
                    long a;
                    if (AXIS == 0) a = (long)(16 * blk + p) * nx + i;
                    else if (AXIS == 1) a = ((long)outer * ny + i) * nx + x0 + p;
                    else a = ((long)i * ny + outer) * nx + x0 + p;
                    d1[a] = acc[0][0][t][r] + acc[0][1][t][r];
                    d2[a] = acc[1][0][t][r] + acc[1][1][t][r];
                }
            }
        if (nb < nblk) put(pf);
    }
}

void launch_lag_derivs_split(const MeshK<double> &md, const float *psi, const int *qn, int M,
                             float *dsc, hipStream_t st) {
    const int ns[3] = {md.nx, md.ny, md.nz};
#define MOCCA_LAG_AXIS(NP, AX)                                                                  \
    {                                                                                           \
        constexpr size_t shm = sizeof(float) * (4 * NP * NP + 4 * NP * 16);                    \
        static const bool lds_set =                                                             \
            require_lds((const void *)k_lag_axis_split<NP, AX>, shm, "k_lag_axis_split");     \
        if (!lds_set)                                                                           \
            require_lds((const void *)k_lag_axis_split<NP, AX>, shm, "k_lag_axis_split");     \
        hipLaunchKernelGGL((k_lag_axis_split<NP, AX>), dim3(4 * M), dim3(256), shm, st, md,     \
                           psi, qn, M, dsc);                                                    \
    }
#define MOCCA_LAG_NP(AX)                                                                        \
    switch ((ns[AX] + 15) / 16) {                                                               \
        case 1: MOCCA_LAG_AXIS(16, AX) break;                                                   \
        case 2: MOCCA_LAG_AXIS(32, AX) break;                                                   \
        case 3: MOCCA_LAG_AXIS(48, AX) break;                                                   \
        case 4: MOCCA_LAG_AXIS(64, AX) break;                                                   \
        default: break;                                                                         \
    }
    MOCCA_LAG_NP(0)
    MOCCA_LAG_NP(1)
    MOCCA_LAG_NP(2)
#undef MOCCA_LAG_NP
#undef MOCCA_LAG_AXIS
}

template <typename T>
void launch_lag_derivs(const MeshK<double> &md, const T *psi, const int *qn, int M, T *dsc,
                       hipStream_t st) {
    const int ns[3] = {md.nx, md.ny, md.nz};
#define MOCCA_LAG_AXIS(NP, AX)                                                               \
    {                                                                                        \
        constexpr size_t shm = sizeof(double) * (2 * NP * NP + 4 * NP * 16);               \
        static const bool lds_set =                                                          \
            require_lds((const void *)k_lag_axis<T, NP, AX>, shm, "k_lag_axis");           \
        if (!lds_set) require_lds((const void *)k_lag_axis<T, NP, AX>, shm, "k_lag_axis"); \
        hipLaunchKernelGGL((k_lag_axis<T, NP, AX>), dim3(4 * M), dim3(256), shm, st, md, psi, \
                           qn, M, dsc);                                                      \
    }
#define MOCCA_LAG_NP(AX)                                                                     \
    switch ((ns[AX] + 15) / 16) {                                                            \
        case 1: MOCCA_LAG_AXIS(16, AX) break;                                                \
        case 2: MOCCA_LAG_AXIS(32, AX) break;                                                \
        case 3: MOCCA_LAG_AXIS(48, AX) break;                                                \
        case 4: MOCCA_LAG_AXIS(64, AX) break;                                                \
        default: break;                                                                      \
    }
    MOCCA_LAG_NP(0)
    MOCCA_LAG_NP(1)
    MOCCA_LAG_NP(2)
#undef MOCCA_LAG_NP
#undef MOCCA_LAG_AXIS
}

// hPsi (ImaginaryTime.f90:213-304) fused with the heavy-ball step of GradDesc
// (ImaginaryTime.f90:467-491).  One thread per mesh point, looping over the
// states of one chunk (one isospin), so the isospin's potentials are read
// once per point and reused.  Partial sums of <psi|h psi>, <h psi|h psi> and
// <psi|Lap psi> (the kinetic sum of CompKinetic, Energy.f90:1426-1456) go to
// eparts[state][tile][3].
// STALE: values from vsrc, derivatives from psi (see k_hpsi_fd).
template <typename T, bool STALE>
__global__ __launch_bounds__(kBlock) void k_hpsi_step(
    MeshK<T> m, const T *__restrict__ psi, T *__restrict__ psi_out,
    T *__restrict__ upd, T *__restrict__ hout, const T *__restrict__ F,
    const int *__restrict__ qn, const int *__restrict__ order,
    const Chunk *__restrict__ chunks, int trc, T kin0, T kin1, T dth, T mu,
    int mode, double *__restrict__ eparts, int ntiles, const T *__restrict__ vsrc,
    const T *__restrict__ dsc) {
    __shared__ double sh[3 * kBlock / 64];
    const long N = m.N;
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const Chunk ch = chunks[blockIdx.y];
    const bool valid = p < N;
    const long pp = valid ? p : 0;
    const int x = pp % m.nx, y = (pp / m.nx) % m.ny, z = pp / ((long)m.nx * m.ny);
    const int it = ch.it;
    const T kin = it == 0 ? kin0 : kin1;
    // Potentials of this isospin at the point (MeanFields.f90:80-95).
    const T u = F[(kOffU + it) * N + pp];
    const T b = F[(kOffB + it) * N + pp];
    T nb[3], w[3][3], dw[3], sv[3] = {0, 0, 0}, av[3] = {0, 0, 0}, da = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        nb[k] = F[(kOffNB + it * 3 + k) * N + pp];
        dw[k] = F[(kOffDW + it * 3 + k) * N + pp];
#pragma unroll
        for (int mu_ = 0; mu_ < 3; ++mu_) w[k][mu_] = F[(kOffW + it * 9 + k * 3 + mu_) * N + pp];
    }
    if (!trc) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sv[k] = F[(kOffS + it * 3 + k) * N + pp];
            av[k] = F[(kOffA + it * 3 + k) * N + pp];
        }
        da = F[(kOffDA + it) * N + pp];
    }
    for (int q = 0; q < ch.count; ++q) {
        const int s = order[ch.start + q];
        const T *ps = psi + (long)s * 4 * N;
        T v[4] = {0, 0, 0, 0}, d[3][4] = {{0}}, l[4] = {0, 0, 0, 0}, h[4] = {0, 0, 0, 0};
        T vd[4] = {0, 0, 0, 0};   // values of the derivative source
        if (valid) {
            if (dsc) {
                // derivatives from the matrix-core Lagrange pass (k_lag_axis)
                const T *dd = dsc + (long)s * 24 * N + p;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    v[c] = ps[(long)c * N + p];
#pragma unroll
                    for (int a = 0; a < 3; ++a) d[a][c] = dd[(c * 6 + a) * N];
                    l[c] = dd[(c * 6 + 3) * N] + dd[(c * 6 + 4) * N] + dd[(c * 6 + 5) * N];
                }
            } else {
                state_derivs(m, ps, qn[4 * s + 1], qn[4 * s + 2], qn[4 * s + 3], x, y, z,
                             p, v, d, l, true);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) vd[c] = v[c];
            if constexpr (STALE) {
#pragma unroll
                for (int c = 0; c < 4; ++c) v[c] = vsrc[(long)s * 4 * N + (long)c * N + p];
            }
            // ActionOfBNew (MeanFields.f90:811-851)
            // ActionOfU    (MeanFields.f90:792-809)
            // ActionOfW    (MeanFields.f90:1029-1065): -i[W_mn s_n d_m + 1/2 divW_n s_n] psi
            T aw[4] = {0, 0, 0, 0};
#pragma unroll
            for (int mu_ = 0; mu_ < 3; ++mu_) {
                const T *e = d[mu_];
                const T wx = w[0][mu_], wy = w[1][mu_], wz = w[2][mu_];
                aw[0] += wx * e[2] + wy * e[3] + wz * e[0];
                aw[1] += wx * e[3] - wy * e[2] + wz * e[1];
                aw[2] += wx * e[0] - wy * e[1] - wz * e[2];
                aw[3] += wx * e[1] + wy * e[0] - wz * e[3];
            }
            {
                const T hx = T(0.5) * dw[0], hy = T(0.5) * dw[1], hz = T(0.5) * dw[2];
                aw[0] += hx * v[2] + hy * v[3] + hz * v[0];
                aw[1] += hx * v[3] - hy * v[2] + hz * v[1];
                aw[2] += hx * v[0] - hy * v[1] - hz * v[2];
                aw[3] += hx * v[1] + hy * v[0] - hz * v[3];
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T ab = -b * l[c] - kin * l[c] -
                             (nb[0] * d[0][c] + nb[1] * d[1][c] + nb[2] * d[2][c]);
                h[c] = ab + u * v[c];
            }
            h[0] += aw[1]; h[1] -= aw[0]; h[2] += aw[3]; h[3] -= aw[2];
            if (!trc) {
                // ActionOfS (MeanFields.f90:1131-1147): S . sigma psi
                h[0] += sv[0] * v[2] + sv[1] * v[3] + sv[2] * v[0];
                h[1] += sv[0] * v[3] - sv[1] * v[2] + sv[2] * v[1];
                h[2] += sv[0] * v[0] - sv[1] * v[1] - sv[2] * v[2];
                h[3] += sv[0] * v[1] + sv[1] * v[0] - sv[2] * v[3];
                // ActionOfA (MeanFields.f90:1107-1129): -i[A.grad + 1/2 divA] psi
                T aa[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    aa[c] = av[0] * d[0][c] + av[1] * d[1][c] + av[2] * d[2][c] +
                            T(0.5) * da * v[c];
                h[0] += aa[1]; h[1] -= aa[0]; h[2] += aa[3]; h[3] -= aa[2];
            }
        }
        double e = 0, dd = 0, kk = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            e += (double)v[c] * (double)h[c];
            dd += (double)h[c] * (double)h[c];
            kk += (double)v[c] * (double)l[c];
        }
        if (valid) {
            if (mode & kWriteH) {
#pragma unroll
                for (int c = 0; c < 4; ++c) hout[(long)c * N + p] = h[c];
            }
            if (mode & kStep) {
                T *us = upd + (long)s * 4 * N;
                T *po = psi_out + (long)s * 4 * N;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    // u = psi - psi_prev (ImaginaryTime.f90:497-499)
                    const T u = vd[c] - us[(long)c * N + p];
                    po[(long)c * N + p] = v[c] + (-(dth * h[c]) + mu * u);
                }
            }
        }
        if (eparts) {
            block_sum3(e, dd, kk, sh);
            if (threadIdx.x == 0) {
                double *ep = eparts + ((long)s * ntiles + blockIdx.x) * 3;
                ep[0] = e;
                ep[1] = dd;
                ep[2] = kk;
            }
        }
    }
}

template <typename T>
void launch_hpsi_step(const MeshK<T> &m, const T *psi, T *psi_out, T *upd,
                      T *hout, const T *fields, const int *qn, const int *order,
                      const Chunk *chunks, int nchunks, int trc, T kin0, T kin1,
                      T dth, T mu, int mode, double *eparts, int ntiles,
                      hipStream_t st, const T *vsrc, const T *dsc) {
    dim3 g(ntiles, nchunks);
    if (vsrc)
        hipLaunchKernelGGL((k_hpsi_step<T, true>), g, dim3(kBlock), 0, st, m, psi, psi_out,
                           upd, hout, fields, qn, order, chunks, trc, kin0, kin1, dth, mu,
                           mode, eparts, ntiles, vsrc, dsc);
    else
        hipLaunchKernelGGL((k_hpsi_step<T, false>), g, dim3(kBlock), 0, st, m, psi, psi_out,
                           upd, hout, fields, qn, order, chunks, trc, kin0, kin1, dth, mu,
                           mode, eparts, ntiles, vsrc, dsc);
}

// Energies and dispersions (ImaginaryTime.f90:470-471): E = dv sum psi.h,
// disp = dv sum h.h - E^2.
__global__ __launch_bounds__(kBlock) void k_reduce_ed(const double *__restrict__ eparts,
                                                      int ntiles,
                                                      const int *__restrict__ states,
                                                      double dv, double *E, double *D,
                                                      double *K) {
    __shared__ double sh[3 * kBlock / 64];
    const int s = states[blockIdx.x];
    double a = 0, b = 0, c = 0;
    for (int t = threadIdx.x; t < ntiles; t += kBlock) {
        const double *ep = eparts + ((long)s * ntiles + t) * 3;
        a += ep[0];
        b += ep[1];
        c += ep[2];
    }
    block_sum3(a, b, c, sh);
    if (threadIdx.x == 0) {
        const double e = a * dv;
        if (E) E[s] = e;
        if (D) D[s] = b * dv - e * e;
        if (K) K[s] = c * dv;
    }
}

void launch_reduce_ed(const double *eparts, int ntiles, const int *states,
                      int nstates, double dv, double *E, double *D, double *K,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_ed, dim3(nstates), dim3(kBlock), 0, st, eparts,
                       ntiles, states, dv, E, D, K);
}

// ---------------------------------------------------------------------------
// Density accumulation (ComputeDensity, Densities.f90:464-534; getters
// Spwf.f90:536,1897,2034,2071,2206): per chunk partial fields
// part[chunk][f][N], f = rho, tau, nablaJ, [j_x, j_y, j_z, s_x, s_y, s_z].
template <typename T>
__device__ __forceinline__ T im_dot(const T a[4], const T b[4]) {
    return a[0] * b[1] - a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
}

template <typename T>
__device__ __forceinline__ void pauli(const T a[4], int dir, T o[4]) {
    if (dir == 0) { o[0] = a[2]; o[1] = a[3]; o[2] = a[0]; o[3] = a[1]; }
    else if (dir == 1) { o[0] = a[3]; o[1] = -a[2]; o[2] = -a[1]; o[3] = a[0]; }
    else { o[0] = a[0]; o[1] = a[1]; o[2] = -a[2]; o[3] = -a[3]; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_density(
    MeshK<T> m, const T *__restrict__ psi, const int *__restrict__ qn,
    const double *__restrict__ occ, const int *__restrict__ order,
    const Chunk *__restrict__ chunks, int trc, T *__restrict__ part) {
    const long N = m.N;
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= N) return;
    const Chunk ch = chunks[blockIdx.y];
    const int x = p % m.nx, y = (p / m.nx) % m.ny, z = p / ((long)m.nx * m.ny);
    T acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < ch.count; ++q) {
        const int s = order[ch.start + q];
        const T o = (T)occ[s];
        if (occ[s] == 0.0) continue;
        T v[4], d[3][4], l[4];
        state_derivs(m, psi + (long)s * 4 * N, qn[4 * s + 1], qn[4 * s + 2],
                     qn[4 * s + 3], x, y, z, p, v, d, l, false);
        T t[4];
        acc[0] += o * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
        T ta = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            ta += d[a][0] * d[a][0] + d[a][1] * d[a][1] + d[a][2] * d[a][2] + d[a][3] * d[a][3];
        acc[1] += o * ta;
        // nabla.J = sum eps_ijk Im(d_i psi^+ sigma_k d_j psi) (Spwf.f90:2206-2243)
        T nj = 0;
        pauli(d[1], 2, t); nj += im_dot(d[0], t);
        pauli(d[2], 1, t); nj -= im_dot(d[0], t);
        pauli(d[1], 0, t); nj -= im_dot(d[2], t);
        pauli(d[0], 1, t); nj += im_dot(d[2], t);
        pauli(d[0], 2, t); nj -= im_dot(d[1], t);
        pauli(d[2], 0, t); nj += im_dot(d[1], t);
        acc[2] += o * nj;
        if (!trc) {
#pragma unroll
            for (int a = 0; a < 3; ++a) acc[3 + a] += o * im_dot(v, d[a]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                pauli(v, a, t);
                acc[6 + a] += o * (v[0] * t[0] + v[1] * t[1] + v[2] * t[2] + v[3] * t[3]);
            }
        }
    }
    const int nf = trc ? 3 : 9;
    for (int f = 0; f < nf; ++f) part[((long)blockIdx.y * nf + f) * N + p] = acc[f];
}

template <typename T>
void launch_density(const MeshK<T> &m, const T *psi, const int *qn,
                    const double *occ, const int *order, const Chunk *chunks,
                    int nchunks, int trc, T *part, hipStream_t st) {
    dim3 g((m.N + kBlock - 1) / kBlock, nchunks);
    hipLaunchKernelGGL(k_density<T>, g, dim3(kBlock), 0, st, m, psi, qn, occ, order,
                       chunks, trc, part);
}

// Sum of chunk partials per isospin in chunk order: dens[it][f][N] (double,
// nine field slots per isospin).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_density_reduce(const T *__restrict__ part,
                                                           const Chunk *__restrict__ chunks,
                                                           int nchunks, int nf, long N,
                                                           double *__restrict__ dens) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    const int f = blockIdx.y;
    if (p >= N) return;
    double a0 = 0, a1 = 0;
    for (int c = 0; c < nchunks; ++c) {
        const double v = (double)part[((long)c * nf + f) * N + p];
        if (chunks[c].it == 0) a0 += v; else a1 += v;
    }
    dens[((long)0 * 9 + f) * N + p] = a0;
    dens[((long)1 * 9 + f) * N + p] = a1;
}

template <typename T>
void launch_density_reduce(const T *part, const Chunk *chunks, int nchunks,
                           int nf, int N, double *dens, hipStream_t st) {
    dim3 g((N + kBlock - 1) / kBlock, nf);
    hipLaunchKernelGGL(k_density_reduce<T>, g, dim3(kBlock), 0, st, part, chunks,
                       nchunks, nf, (long)N, dens);
}

// ---------------------------------------------------------------------------
// Derivatives of fields (DerRho, LapRho, rot j, rot s, DerS ...:
// Densities.f90:541-550,1003-1101), batched as a job list.
__global__ __launch_bounds__(kBlock) void k_field_jobs(MeshK<double> m,
                                                       const FieldJob *__restrict__ jobs) {
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= m.N) return;
    const FieldJob &jb = jobs[blockIdx.y];
    const int x = p % m.nx, y = (p / m.nx) % m.ny, z = p / ((long)m.nx * m.ny);
    double acc = 0;
    for (int t = 0; t < jb.nterms; ++t) {
        double val;
        if (jb.axis[t] < 3) {
            double d1, d2 = 0;
            axis_derivs(m, jb.src[t], x, y, z, jb.axis[t], jb.sign[t][0],
                        jb.sign[t][0], d1, d2, false);
            val = d1;
        } else {
            double lap = 0;
            for (int a = 0; a < 3; ++a) {
                double d1, d2;
                axis_derivs(m, jb.src[t], x, y, z, a, jb.sign[t][a], jb.sign[t][a],
                            d1, d2, true);
                lap += d2;
            }
            val = lap;
        }
        acc += jb.coef[t] * val;
    }
    jb.dst[p] = acc;
}

void launch_field_jobs(const MeshK<double> &m, const FieldJob *jobs, int njobs,
                       hipStream_t st) {
    if (njobs <= 0) return;
    dim3 g((m.N + kBlock - 1) / kBlock, njobs);
    hipLaunchKernelGGL(k_field_jobs, g, dim3(kBlock), 0, st, m, jobs);
}

// ---------------------------------------------------------------------------
// PreconditionPotential (Damping.f90:158-198) with preconoperator
// (:200-228), the MixingScheme 2 step of ConstructPotentials
// (PotentialMixing.f90:143-171): for each job, du = pot - old, then the
// reference's conjugate gradients on (beta + alpha Lap) x = du from x = 0
// (at most 300 iterations, ending after the direction update once
// dv |r|^2 < 1e-8), and pot = old + x.  Lap is the wavefunction Laplacian
// (FD or Lagrange, axis_derivs) with the job's ghost signs.  One workgroup
// per job keeps every reduction inside the workgroup, so the whole solve is
// one launch with no host round trip; work holds x, r, d, q per job.
constexpr int kPrecBlock = 1024;

__device__ __forceinline__ double prec_sum(double a, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = a;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int k = 0; k < kPrecBlock / 64; ++k) s += sh[k];
    return s;
}

__global__ __launch_bounds__(kPrecBlock) void k_precondition(MeshK<double> m,
                                                            const PrecJob *__restrict__ jobs,
                                                            double *__restrict__ work,
                                                            double alpha, double beta,
                                                            double dv, int *__restrict__ iters) {
    __shared__ double sh[kPrecBlock / 64];
    const PrecJob jb = jobs[blockIdx.x];
    const long N = m.N;
    double *x = work + (long)blockIdx.x * 4 * N, *r = x + N, *d = r + N, *q = d + N;
    double rr = 0;
    for (long p = threadIdx.x; p < N; p += kPrecBlock) {
        const double v = jb.pot[p] - jb.old[p];
        x[p] = 0.0;
        r[p] = v;
        d[p] = v;
        rr += v * v;
    }
    double newres = prec_sum(rr, sh) * dv;
    int iter = 0;
    // du = 0 exactly (a field that vanishes by symmetry) leaves pot as it
    // is; the reference divides 0 by 0 there.
    if (newres > 0.0) {
        for (iter = 1; iter <= 300; ++iter) {
            double dq = 0;
            for (long p = threadIdx.x; p < N; p += kPrecBlock) {
                const int xx = p % m.nx, yy = (p / m.nx) % m.ny, zz = p / ((long)m.nx * m.ny);
                double lap = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    double d1, d2;
                    axis_derivs(m, (const double *)d, xx, yy, zz, a, jb.sign[a], jb.sign[a], d1,
                                d2, true);
                    lap += d2;
                }
                const double qv = beta * d[p] + alpha * lap;
                q[p] = qv;
                dq += d[p] * qv;
            }
            const double a = newres / (prec_sum(dq, sh) * dv);
            double r2 = 0;
            for (long p = threadIdx.x; p < N; p += kPrecBlock) {
                x[p] = x[p] + a * d[p];
                const double rv = r[p] - a * q[p];
                r[p] = rv;
                r2 += rv * rv;
            }
            const double oldres = newres;
            newres = prec_sum(r2, sh) * dv;
            const double b = newres / oldres;
            for (long p = threadIdx.x; p < N; p += kPrecBlock) d[p] = r[p] + b * d[p];
            __syncthreads();
            if (newres < 1e-8) break;
        }
        if (iter > 300) iter = 300;
        for (long p = threadIdx.x; p < N; p += kPrecBlock) jb.pot[p] = jb.old[p] + x[p];
    }
    if (threadIdx.x == 0 && iters) iters[blockIdx.x] = iter;
}

void launch_precondition(const MeshK<double> &m, const PrecJob *jobs, int njobs, double *work,
                         double alpha, double beta, double dv, int *iters, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_precondition, dim3(njobs), dim3(kPrecBlock), 0, st, m, jobs, work,
                       alpha, beta, dv, iters);
}

// ---------------------------------------------------------------------------
// Gram-Schmidt as Cholesky QR (SpwfStorage.f90:372-480).  Gram matrices of
// each (isospin, parity, signature) block, G = dv * Psi^T Psi, in FP64.
// Grid: x = K split, y = upper 64x64 tile of G, z = block.
__device__ __forceinline__ bool tile_of(int tix, int nt, int &ti, int &tj) {
    int cnt = 0;
    for (int a = 0; a < nt; ++a)
        for (int c = a; c < nt; ++c) {
            if (cnt == tix) { ti = a; tj = c; return true; }
            ++cnt;
        }
    return false;
}

template <typename T, int TILE>
__global__ __launch_bounds__(kBlock) void k_gram(const T *__restrict__ psi, long K4,
                                                 const int *__restrict__ blk_states,
                                                 const int *__restrict__ blk_m,
                                                 int max_m, double *__restrict__ partial,
                                                 int nsplit, int ntile_max) {
    constexpr int R = TILE / 16;  // register tile edge per thread
    const int w = blockIdx.x, tix = blockIdx.y, b = blockIdx.z;
    const int m = blk_m[b];
    int ti, tj;
    if (!tile_of(tix, (m + TILE - 1) / TILE, ti, tj)) return;
    __shared__ double As[kGramKC][TILE + 2];
    __shared__ double Bs[kGramKC][TILE + 2];
    const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
    double acc[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = 0;
    const long kw = ((K4 + nsplit - 1) / nsplit + kGramKC - 1) / kGramKC * kGramKC;
    const long k0 = (long)w * kw, k1 = min(K4, k0 + kw);
    const int *st = blk_states + (long)b * max_m;
    for (long kb = k0; kb < k1; kb += kGramKC) {
        for (int q = tid; q < TILE * kGramKC; q += kBlock) {
            const int r = q / kGramKC, kk = q % kGramKC;
            const long k = kb + kk;
            const int ra = ti * TILE + r, rb = tj * TILE + r;
            As[kk][r] = (ra < m && k < k1) ? (double)psi[(long)st[ra] * K4 + k] : 0.0;
            Bs[kk][r] = (rb < m && k < k1) ? (double)psi[(long)st[rb] * K4 + k] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGramKC; ++kk) {
            double a[R], bb[R];
#pragma unroll
            for (int i = 0; i < R; ++i) { a[i] = As[kk][tr * R + i]; bb[i] = Bs[kk][tc * R + i]; }
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int j = 0; j < R; ++j) acc[i][j] += a[i] * bb[j];
        }
        __syncthreads();
    }
    double *out = partial + (((long)b * ntile_max + tix) * nsplit + w) * TILE * TILE;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) out[(tr * R + i) * TILE + tc * R + j] = acc[i][j];
}

template <typename T>
void launch_gram(const T *psi, long K4, const int *blk_states, const int *blk_m,
                 int nblk, int max_m, int mmax, int tile, double *partial,
                 int nsplit, hipStream_t st) {
    const int nt = (mmax + tile - 1) / tile;
    const int ntile_max = nt * (nt + 1) / 2;
    dim3 g(nsplit, ntile_max, nblk);
    if (tile == 32)
        hipLaunchKernelGGL((k_gram<T, 32>), g, dim3(kBlock), 0, st, psi, K4, blk_states,
                           blk_m, max_m, partial, nsplit, ntile_max);
    else
        hipLaunchKernelGGL((k_gram<T, 64>), g, dim3(kBlock), 0, st, psi, K4, blk_states,
                           blk_m, max_m, partial, nsplit, ntile_max);
}

template <int TILE>
__global__ __launch_bounds__(kBlock) void k_gram_reduce(const double *__restrict__ partial,
                                                        const int *__restrict__ blk_m,
                                                        int max_m, int nsplit,
                                                        int ntile_max, double dv,
                                                        double *__restrict__ G) {
    const int tix = blockIdx.y, b = blockIdx.z;
    const int m = blk_m[b];
    int ti, tj;
    if (!tile_of(tix, (m + TILE - 1) / TILE, ti, tj)) return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= TILE * TILE) return;
    const int r = e / TILE, c = e % TILE;
    const int i = ti * TILE + r, j = tj * TILE + c;
    if (i >= m || j >= m) return;
    const double *src = partial + ((long)b * ntile_max + tix) * nsplit * TILE * TILE + e;
    double s = 0;
    for (int w = 0; w < nsplit; ++w) s += src[(long)w * TILE * TILE];
    s *= dv;
    double *g = G + (long)b * max_m * max_m;
    g[(long)i * max_m + j] = s;
    g[(long)j * max_m + i] = s;
}

void launch_gram_reduce(const double *partial, const int *blk_m, int nblk,
                        int max_m, int mmax, int tile, int nsplit, double dv,
                        double *G, hipStream_t st) {
    const int nt = (mmax + tile - 1) / tile;
    const int ntile_max = nt * (nt + 1) / 2;
    dim3 g((tile * tile + kBlock - 1) / kBlock, ntile_max, nblk);
    if (tile == 32)
        hipLaunchKernelGGL(k_gram_reduce<32>, g, dim3(kBlock), 0, st, partial, blk_m, max_m,
                           nsplit, ntile_max, dv, G);
    else
        hipLaunchKernelGGL(k_gram_reduce<64>, g, dim3(kBlock), 0, st, partial, blk_m, max_m,
                           nsplit, ntile_max, dv, G);
}

// Per block: ascending-energy order (the stable insertion sort of
// OrderSpwfsISO, SpwfStorage.f90:795-847), Cholesky G = L L^T of the
// reordered Gram matrix and T = (L^T)^{-1}, so that the orthonormal states
// are Q = Psi_sorted T -- the same Q as the reference's modified
// Gram-Schmidt in exact arithmetic.
__global__ __launch_bounds__(kBlock) void k_cholesky(const double *__restrict__ G,
                                                     const int *__restrict__ blk_states,
                                                     const int *__restrict__ blk_m,
                                                     int max_m, const double *__restrict__ E,
                                                     double *__restrict__ Tm,
                                                     int *__restrict__ perm,
                                                     double *__restrict__ work,
                                                     int *__restrict__ status) {
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = blk_m[b];
    const int *st = blk_states + (long)b * max_m;
    const double *g = G + (long)b * max_m * max_m;
    double *A = work + (long)b * max_m * max_m;
    double *Tb = Tm + (long)b * max_m * max_m;
    int *pb = perm + (long)b * max_m;
    __shared__ int rank[kMaxBlock];
    for (int i = tid; i < m; i += kBlock) {
        const double ei = E[st[i]];
        int r = 0;
        for (int j = 0; j < m; ++j) {
            const double ej = E[st[j]];
            if (ej < ei || (ej == ei && j < i)) ++r;
        }
        rank[i] = r;
        pb[r] = st[i];
    }
    __syncthreads();
    for (long q = tid; q < (long)m * m; q += kBlock) {
        const int i = q / m, j = q % m;
        A[(long)rank[i] * max_m + rank[j]] = g[(long)i * max_m + j];
    }
    __syncthreads();
    __shared__ double diag;
    for (int k = 0; k < m; ++k) {
        if (tid == 0) {
            const double d = A[(long)k * max_m + k];
            if (!(d > 0.0)) atomicOr(status, 1);
            diag = sqrt(d > 0.0 ? d : 1.0);
            A[(long)k * max_m + k] = diag;
        }
        __syncthreads();
        const double dk = diag;
        for (int i = k + 1 + tid; i < m; i += kBlock) A[(long)i * max_m + k] /= dk;
        __syncthreads();
        const int r = m - k - 1;
        for (long q = tid; q < (long)r * r; q += kBlock) {
            const int i = k + 1 + q / r, j = k + 1 + q % r;
            if (j <= i) A[(long)i * max_m + j] -= A[(long)i * max_m + k] * A[(long)j * max_m + k];
        }
        __syncthreads();
    }
    // Row j of T = column j of L^{-1} (forward substitution).
    for (int j = tid; j < m; j += kBlock) {
        double *X = Tb + (long)j * max_m;
        for (int i = 0; i < j; ++i) X[i] = 0.0;
        for (int i = j; i < m; ++i) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = j; k < i; ++k) s -= A[(long)i * max_m + k] * X[k];
            X[i] = s / A[(long)i * max_m + i];
        }
    }
}

// The same factorisation with the block resident in LDS (leading dimension
// m+1): right-looking Cholesky with one barrier per column (the division of
// column k by its pivot is deferred to step k+1, when nothing reads it),
// then L^{-1} by rows into the free upper triangle, so that T = L^{-T} is
// that upper triangle with the reciprocal pivots on the diagonal.
__global__ __launch_bounds__(1024) void k_cholesky_lds(const double *__restrict__ G,
                                                       const int *__restrict__ blk_states,
                                                       const int *__restrict__ blk_m,
                                                       int max_m, const double *__restrict__ E,
                                                       double *__restrict__ Tm,
                                                       int *__restrict__ perm,
                                                       int *__restrict__ status) {
    extern __shared__ double A[];           // [m][m+1]
    __shared__ int rank[kMaxBlock];
    __shared__ double dinv[kMaxBlock];      // first the block's energies, then 1/L[i][i]
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int m = blk_m[b], ld = m + 1;
    const int *st = blk_states + (long)b * max_m;
    const double *g = G + (long)b * max_m * max_m;
    double *Tb = Tm + (long)b * max_m * max_m;
    int *pb = perm + (long)b * max_m;
    for (int i = tid; i < m; i += nt) dinv[i] = E[st[i]];
    __syncthreads();
    for (int i = tid; i < m; i += nt) {
        const double ei = dinv[i];
        int r = 0;
        for (int j = 0; j < m; ++j) {
            const double ej = dinv[j];
            if (ej < ei || (ej == ei && j < i)) ++r;
        }
        rank[i] = r;
        pb[r] = st[i];
    }
    __syncthreads();
    for (int q = tid; q < m * m; q += nt) {
        const int i = q / m, j = q % m;
        A[rank[i] * ld + rank[j]] = g[(long)i * max_m + j];
    }
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const double akk = A[k * ld + k];
        if (k > 0) {
            // finish column k-1: L[i][k-1] = A[i][k-1] / d_{k-1}
            const double d = sqrt(A[(k - 1) * ld + k - 1] > 0 ? A[(k - 1) * ld + k - 1] : 1.0);
            for (int i = k + tid; i < m; i += nt) A[i * ld + k - 1] /= d;
        }
        if (tid == 0 && !(akk > 0.0)) atomicOr(status, 1);
        const double inv = 1.0 / (akk > 0.0 ? akk : 1.0);
        // trailing update on a 32 x 32 thread grid (blockDim = 1024)
        for (int i = k + 1 + tid / 32; i < m; i += 32) {
            const double aik = A[i * ld + k] * inv;
            for (int j = k + 1 + tid % 32; j <= i; j += 32) A[i * ld + j] -= aik * A[j * ld + k];
        }
        __syncthreads();
        if (k > 0 && tid == 0) {
            const double pv = A[(k - 1) * ld + k - 1];
            A[(k - 1) * ld + k - 1] = sqrt(pv > 0 ? pv : 1.0);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const double pv = A[(m - 1) * ld + m - 1];
        A[(m - 1) * ld + m - 1] = sqrt(pv > 0 ? pv : 1.0);
    }
    __syncthreads();
    for (int i = tid; i < m; i += nt) dinv[i] = 1.0 / A[i * ld + i];
    // X = L^{-1} by right-looking substitution: R = I - (sum over finished
    // rows), row k of X = R[k][:] / L[k][k], then R[i][j] -= L[i][k] X[k][j]
    // for i > k, j <= k.  R and X (j < i) live in the free upper triangle at
    // A[j][i]; the diagonal of R stays 1, so X[k][k] = 1/L[k][k].
    for (int q = tid; q < m * m; q += nt) {
        const int j = q / m, i = q % m;
        if (j < i) A[j * ld + i] = 0.0;
    }
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const double dk = dinv[k];
        for (int j = tid; j < k; j += nt) A[j * ld + k] *= dk;   // X[k][j]
        __syncthreads();
        for (int i = k + 1 + tid / 32; i < m; i += 32) {
            const double lik = A[i * ld + k];
            for (int j = tid % 32; j <= k; j += 32) {
                const double xkj = j == k ? dk : A[j * ld + k];
                A[j * ld + i] -= lik * xkj;
            }
        }
        __syncthreads();
    }
    // T = L^{-T}: row j of T holds X[i][j] for i >= j
    for (int q = tid; q < m * m; q += nt) {
        const int j = q / m, i = q % m;
        Tb[(long)j * max_m + i] = i < j ? 0.0 : (i == j ? dinv[j] : A[j * ld + i]);
    }
}

// The same factorisation for blocks of up to kCholPackedMax states, with only
// the lower triangle resident in LDS (row i at i(i+1)/2), so 150-190 states
// fit.  The factorisation is the right-looking one above, operation for
// operation; X = L^{-1} then goes column by column, one wave per column with
// the column in registers (lane l holds rows l, l+64, l+128), in the same
// order of updates as the in-LDS substitution, and is stored as T = X^T.
constexpr int kCholPackedMax = 192;

__device__ __forceinline__ int tri_row(int i) { return i * (i + 1) / 2; }

__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(1024) void k_cholesky_packed(const double *__restrict__ G,
                                                          const int *__restrict__ blk_states,
                                                          const int *__restrict__ blk_m,
                                                          int max_m,
                                                          const double *__restrict__ E,
                                                          double *__restrict__ Tm,
                                                          int *__restrict__ perm,
                                                          int *__restrict__ status) {
    extern __shared__ double A[];           // packed lower triangle
    __shared__ int rank[kMaxBlock];
    __shared__ double dinv[kMaxBlock];      // first the block's energies, then 1/L[i][i]
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid % 64, wave = tid / 64, nwave = nt / 64;
    const int m = blk_m[b];
    const int *st = blk_states + (long)b * max_m;
    const double *g = G + (long)b * max_m * max_m;
    double *Tb = Tm + (long)b * max_m * max_m;
    int *pb = perm + (long)b * max_m;
    for (int i = tid; i < m; i += nt) dinv[i] = E[st[i]];
    __syncthreads();
    for (int i = tid; i < m; i += nt) {
        const double ei = dinv[i];
        int r = 0;
        for (int j = 0; j < m; ++j) {
            const double ej = dinv[j];
            if (ej < ei || (ej == ei && j < i)) ++r;
        }
        rank[i] = r;
        pb[r] = st[i];
    }
    __syncthreads();
    for (int q = tid; q < m * m; q += nt) {
        const int i = q / m, j = q % m;
        const int ri = rank[i], rj = rank[j];
        if (rj <= ri) A[tri_row(ri) + rj] = g[(long)i * max_m + j];
    }
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const double akk = A[tri_row(k) + k];
        if (k > 0) {
            const double pv = A[tri_row(k - 1) + k - 1];
            const double d = sqrt(pv > 0 ? pv : 1.0);
            for (int i = k + tid; i < m; i += nt) A[tri_row(i) + k - 1] /= d;
        }
        if (tid == 0 && !(akk > 0.0)) atomicOr(status, 1);
        const double inv = 1.0 / (akk > 0.0 ? akk : 1.0);
        for (int i = k + 1 + tid / 32; i < m; i += 32) {
            const int ri = tri_row(i);
            const double aik = A[ri + k] * inv;
            for (int j = k + 1 + tid % 32; j <= i; j += 32) A[ri + j] -= aik * A[tri_row(j) + k];
        }
        __syncthreads();
        if (k > 0 && tid == 0) {
            const double pv = A[tri_row(k - 1) + k - 1];
            A[tri_row(k - 1) + k - 1] = sqrt(pv > 0 ? pv : 1.0);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const double pv = A[tri_row(m - 1) + m - 1];
        A[tri_row(m - 1) + m - 1] = sqrt(pv > 0 ? pv : 1.0);
    }
    __syncthreads();
    for (int i = tid; i < m; i += nt) dinv[i] = 1.0 / A[tri_row(i) + i];
    __syncthreads();
    // column j of X: R = e_j; for k = j..m-1: X[k][j] = R[k] / L[k][k],
    // R[i] -= L[i][k] X[k][j] for i > k.
    const int nr = (m + 63) / 64;
    for (int j = wave; j < m; j += nwave) {
        double x[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) x[r] = (lane + 64 * r == j) ? 1.0 : 0.0;
        for (int k = j; k < m; ++k) {
            const int rk = k >> 6, lk = k & 63;
            const double own = rk == 0 ? x[0] : (rk == 1 ? x[1] : x[2]);
            const double dk = dinv[k];
            const double xk = k == j ? dk : readlane_d(own, lk) * dk;
            if (lane == lk) {
                if (rk == 0) x[0] = xk; else if (rk == 1) x[1] = xk; else x[2] = xk;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int i = lane + 64 * r;
                if (r < nr && i > k && i < m) x[r] -= A[tri_row(i) + k] * xk;
            }
        }
        // T = L^{-T}: row j of T holds X[i][j] for i >= j
        double *trow = Tb + (long)j * max_m;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int i = lane + 64 * r;
            if (r < nr && i < m) trow[i] = i < j ? 0.0 : x[r];
        }
    }
}

// The same factorisation and inverse, operation for operation, for blocks
// of up to kCholRegMax states with the lower triangle held in registers:
// thread t owns the packed elements t and t + 1024 (EE slots).  Step k of
// the factorisation reads column k (its raw values, published to LDS by
// their owners at the end of step k-1) and updates the owned elements right
// of it: one barrier per step, no LDS read-modify-write.  L then goes to
// LDS (packed), and X = L^{-1} runs the same way on register-owned
// R = I - ..., with row k of X published per step.  (Larger blocks keep
// the in-LDS kernels: with 5-6 slots per thread, blocks of 100 states, this
// form took 0.215 ms against 0.165.)
constexpr int kCholRegMax = 63;   // 63 * 64 / 2 <= 2 * 1024 elements

template <int EE>
__global__ __launch_bounds__(1024) void k_cholesky_reg(const double *__restrict__ G,
                                                       const int *__restrict__ blk_states,
                                                       const int *__restrict__ blk_m,
                                                       int max_m, const double *__restrict__ E,
                                                       double *__restrict__ Tm,
                                                       int *__restrict__ perm,
                                                       int *__restrict__ status) {
    extern __shared__ double Lp[];          // packed L at the end
    __shared__ int irank[kMaxBlock];   // sorted position -> block index
    __shared__ double dinv[kMaxBlock];      // first the block's energies, then 1/L[i][i]
    __shared__ double vec[2][kCholRegMax];
    const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int m = blk_m[b];
    const int *st = blk_states + (long)b * max_m;
    const double *g = G + (long)b * max_m * max_m;
    double *Tb = Tm + (long)b * max_m * max_m;
    int *pb = perm + (long)b * max_m;
    for (int i = tid; i < m; i += nt) dinv[i] = E[st[i]];
    __syncthreads();
    for (int i = tid; i < m; i += nt) {
        const double ei = dinv[i];
        int r = 0;
        for (int j = 0; j < m; ++j) {
            const double ej = dinv[j];
            if (ej < ei || (ej == ei && j < i)) ++r;
        }
        irank[r] = i;
        pb[r] = st[i];
    }
    __syncthreads();
    const int ne = m * (m + 1) / 2;
    double a[EE];
    int oi[EE], oj[EE];   // owned element (i, j), j <= i; oi = -1: none
#pragma unroll
    for (int r = 0; r < EE; ++r) {
        const int e = tid + nt * r;
        oi[r] = -1;
        oj[r] = 0;
        a[r] = 0.0;
        if (e < ne) {
            int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while (tri_row(i) > e) --i;
            while (tri_row(i + 1) <= e) ++i;
            const int j = e - tri_row(i);
            oi[r] = i;
            oj[r] = j;
            a[r] = g[(long)irank[i] * max_m + irank[j]];
        }
    }
    // column 0 for step 0
#pragma unroll
    for (int r = 0; r < EE; ++r)
        if (oi[r] >= 0 && oj[r] == 0) vec[0][oi[r]] = a[r];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const double *col = vec[k & 1];
        double *nxt = vec[(k + 1) & 1];
        const double akk = col[k];
        if (tid == 0 && !(akk > 0.0)) atomicOr(status, 1);
        const double inv = 1.0 / (akk > 0.0 ? akk : 1.0);
        const double d = sqrt(akk > 0 ? akk : 1.0);
#pragma unroll
        for (int r = 0; r < EE; ++r) {
            const int i = oi[r], j = oj[r];
            if (i < 0) continue;
            if (j > k) {
                const double aik = col[i] * inv;
                a[r] -= aik * col[j];
                if (j == k + 1) nxt[i] = a[r];
            } else if (j == k) {
                a[r] = i == k ? d : a[r] / d;   // column k of L
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < EE; ++r)
        if (oi[r] >= 0) Lp[tri_row(oi[r]) + oj[r]] = a[r];
    __syncthreads();
    for (int i = tid; i < m; i += nt) dinv[i] = 1.0 / Lp[tri_row(i) + i];
    __syncthreads();
    // X = L^{-1}: R = I - (sum over finished rows); row k of X = R[k][:] / L[k][k]
    // (published), then R[i][j] -= L[i][k] X[k][j] for i > k, j <= k.  Owned
    // elements (i, j), j < i, hold R and end as X[i][j].
#pragma unroll
    for (int r = 0; r < EE; ++r) a[r] = 0.0;
    for (int k = 0; k < m; ++k) {
        double *xrow = vec[k & 1];
        const double dk = dinv[k];
#pragma unroll
        for (int r = 0; r < EE; ++r)
            if (oi[r] == k && oj[r] < k) {
                a[r] *= dk;            // X[k][j]
                xrow[oj[r]] = a[r];
            }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < EE; ++r) {
            const int i = oi[r], j = oj[r];
            if (i > k && j <= k && j < i) {
                const double xkj = j == k ? dk : xrow[j];
                a[r] -= Lp[tri_row(i) + k] * xkj;
            }
        }
    }
    // T = L^{-T}: row j of T holds X[i][j] for i >= j
#pragma unroll
    for (int r = 0; r < EE; ++r) {
        const int i = oi[r], j = oj[r];
        if (i >= 0) Tb[(long)j * max_m + i] = i == j ? dinv[j] : a[r];
    }
    for (int q = tid; q < m * m; q += nt) {
        const int j = q / m, i = q % m;
        if (i < j) Tb[(long)j * max_m + i] = 0.0;
    }
}

// Q = Psi_sorted T for each block, fused with the heavy-ball memory update
// u = psi_new - u of GradDesc (ImaginaryTime.f90:497-499) when update_mode.
// The T entries are the same for every lane: read from memory with
// workgroup-uniform addresses they become scalar loads, so the products
// take T from scalar registers instead of staging it through LDS.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gs_apply(const T *__restrict__ src,
                                                     T *__restrict__ dst,
                                                     T *__restrict__ upd, long K4,
                                                     const int *__restrict__ perm,
                                                     const int *__restrict__ blk_m,
                                                     int max_m,
                                                     const double *__restrict__ Tm,
                                                     int update_mode) {
    extern __shared__ double Ts[];  // [kend][kApplyL]
    const int b = blockIdx.z;
    const int m = blk_m[b];
    const int l0 = blockIdx.y * kApplyL;
    if (l0 >= m) return;
    const int kend = min(m, l0 + kApplyL);
    const double *Tb = Tm + (long)b * max_m * max_m;
    const int *pb = perm + (long)b * max_m;
    for (int q = threadIdx.x; q < kend * kApplyL; q += kBlock) {
        const int k = q / kApplyL, l = q % kApplyL;
        Ts[q] = (l0 + l < m) ? Tb[(long)k * max_m + l0 + l] : 0.0;
    }
    __syncthreads();
    // the staged T columns serve kApplyChunks consecutive chunks of kBlock values
    for (int ch = 0; ch < kApplyChunks; ++ch) {
        const long e = ((long)blockIdx.x * kApplyChunks + ch) * kBlock + threadIdx.x;
        if (e >= K4) return;
        double acc[kApplyL];
#pragma unroll
        for (int l = 0; l < kApplyL; ++l) acc[l] = 0;
        // the values of 8 states are loaded before their products
        for (int k0 = 0; k0 < kend; k0 += 8) {
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                v[q] = k0 + q < kend ? (double)src[(long)pb[k0 + q] * K4 + e] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (k0 + q >= kend) break;
                const double *tr = Ts + (k0 + q) * kApplyL;
#pragma unroll
                for (int l = 0; l < kApplyL; ++l) acc[l] += v[q] * tr[l];
            }
        }
#pragma unroll
        for (int l = 0; l < kApplyL; ++l) {
            if (l0 + l < m) {
                const long o = (long)pb[l0 + l] * K4 + e;
                const T r = (T)acc[l];
                dst[o] = r;
                if (update_mode) upd[o] = r - upd[o];
            }
        }
    }
}

template <typename T>
void launch_gs_apply(const T *src, T *dst, T *upd, long K4, const int *perm,
                     const int *blk_m, int nblk, int max_m, const double *Tm,
                     int update_mode, hipStream_t st) {
    const long per = (long)kBlock * kApplyChunks;
    dim3 g((K4 + per - 1) / per, (max_m + kApplyL - 1) / kApplyL, nblk);
    const size_t shm = sizeof(double) * (size_t)max_m * kApplyL;
    hipLaunchKernelGGL(k_gs_apply<T>, g, dim3(kBlock), shm, st, src, dst, upd, K4,
                       perm, blk_m, max_m, Tm, update_mode);
}

// ---------------------------------------------------------------------------
// Gram matrices and the Gram-Schmidt apply on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: D[16x16] += A[16x4] B[4x16]; operand lane l holds
// A[l%16][l/16] and B[l/16][l%16]; result lane l holds D[l/16 + 4r][l%16] in
// register r -- measured on gfx950).

// Gram partials: one wave per (K slice, block, job) holds a set of upper
// 16x16 tiles of the block.  A symmetric job (SYM) owns the tiles I <= J of
// NI row tiles starting at tile I0 = job * NI; the rectangular job owns the
// NI x NJ tiles of row tiles [0, NI) against [NI, NI + NJ).  Blocks of up to
// 8 row tiles are one symmetric job; 9-10 row tiles are two symmetric jobs
// and one rectangular one (each state read twice).  Per 16 mesh values each
// lane loads 4 consecutive values of its state row (16 lanes = 16 states, 4
// lane groups = 16 consecutive k): the k order inside a 16-chunk is permuted
// the same way for both operands, which leaves the sum over k unchanged.
// Output: the raw accumulator layout part[b][split][tile][lane][reg], tiles
// numbered over the upper triangle of all nt row tiles; tiles past nt (an
// odd nt split in two) are skipped.
template <typename T, int NI, int NJ, bool SYM>
__global__ __launch_bounds__(64) void k_gram_mfma(const T *__restrict__ psi, long K4,
                                                  const int *__restrict__ blk_states,
                                                  const int *__restrict__ blk_m, int max_m,
                                                  double *__restrict__ part, int nsplit,
                                                  int nt) {
    constexpr int NR = SYM ? NI : NI + NJ;                   // row tiles loaded
    constexpr int NTILE = SYM ? NI * (NI + 1) / 2 : NI * NJ;  // tiles accumulated
    const int w = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const int I0 = SYM ? blockIdx.z * NI : 0;
    const int m = blk_m[b];
    const int *st = blk_states + (long)b * max_m;
    const long nch = (K4 + 15) / 16;
    const long c0 = nch * w / nsplit, c1 = nch * (w + 1) / nsplit;
    const T *rows[NR];
    bool ok[NR];
#pragma unroll
    for (int I = 0; I < NR; ++I) {
        const int r = (I0 + I) * 16 + lane % 16;
        ok[I] = r < m;
        rows[I] = psi + (long)(ok[I] ? st[r] : 0) * K4;
    }
    mfma_d4 acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t) acc[t] = mfma_d4{0, 0, 0, 0};
    // Chunk c+1 is loaded in the storage type, without branches (rows past
    // the block and k past the end read row 0 / k 0), while the MFMAs of
    // chunk c run; it is widened to FP64 afterwards, with those entries
    // zeroed.  K4 is a multiple of 4, so a lane's 4 values are all in range
    // or all out.
    auto load = [&](long c, T (&raw)[NR][4]) {
        const long kb = c * 16 + 4 * (lane / 16);
        const long ks = kb < K4 ? kb : 0;
#pragma unroll
        for (int I = 0; I < NR; ++I)
#pragma unroll
            for (int q = 0; q < 4; ++q) raw[I][q] = rows[I][ks + q];
    };
    auto widen = [&](long c, const T (&raw)[NR][4], double (&v)[NR][4]) {
        const bool kin = c * 16 + 4 * (lane / 16) < K4;
#pragma unroll
        for (int I = 0; I < NR; ++I)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[I][q] = (kin && ok[I]) ? (double)raw[I][q] : 0.0;
    };
    double v[NR][4];
    T raw[NR][4];
    if (c0 < c1) {
        load(c0, raw);
        widen(c0, raw, v);
    }
    for (long c = c0; c < c1; ++c) {
        if (c + 1 < c1) load(c + 1, raw);
        int t = 0;
#pragma unroll
        for (int I = 0; I < NI; ++I)
#pragma unroll
            for (int J = SYM ? I : NI; J < (SYM ? NI : NR); ++J, ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[t] = mfma16x16x4(v[I][q], v[J][q], acc[t]);
        widen(c + 1, raw, v);   // after the last chunk: unused
    }
    const int ntile = nt * (nt + 1) / 2;
    double *P = part + (((long)b * nsplit + w) * ntile) * 256 + lane * 4;
    int t = 0;
#pragma unroll
    for (int I = 0; I < NI; ++I)
#pragma unroll
        for (int J = SYM ? I : NI; J < (SYM ? NI : NR); ++J, ++t) {
            const int gi = I0 + I, gj = I0 + J;   // row tiles of the whole block
            if (gj >= nt) continue;
            const int tg = gi * nt - gi * (gi - 1) / 2 + (gj - gi);
#pragma unroll
            for (int r = 0; r < 4; ++r) P[tg * 256 + r] = acc[t][r];
        }
}

// Q = Psi_sorted T on the FP64 matrix cores, for blocks of 129-192 states
// (k_apply_direct below takes the smaller ones, whose T fits LDS whole),
// as a GEMM [K4 x m] = [K4 x m] [m x m] tiled for LDS: a workgroup of NW
// waves owns 32 NW consecutive values of K and all m outputs; the sorted
// states and T pass through LDS in chunks of 16 states, double-buffered
// (chunk c+1 is loaded into registers while the MFMAs of chunk c run).
//   D[16 l][16 e] += A[16 l][4 k] B[4 k][16 e]: A = T^T (lane L holds
//   T[k0 + L/16][l0 + L%16]), B = the states (lane L holds psi_sorted
//   [k0 + L/16][e0 + L%16]); result lane L, register r: output state
//   l0 + L/16 + 4r at e0 + L%16, so 16 lanes store 16 consecutive values.
// T is upper triangular (T[k][l] = 0 for k > l): chunk c skips the l tiles
// below c.  Each wave accumulates 32 values of K x NT*16 outputs.  Every
// workgroup reads all of T, so wider workgroups (NW = 8) halve that traffic
// for the larger blocks.
// FP32 states (mixed precision) take the FP32 matrix cores
// (v_mfma_f32_16x16x4f32, twice the FP64 rate) with T rounded to FP32 and
// FP32 accumulation; their results are stored in FP32 anyway.  Its result
// lane l, register r holds D[4 (l/16) + r][l%16].
constexpr int kApplyNE = 2;               // 16-value K tiles per wave
template <int NW>
constexpr int apply_ew() { return 16 * kApplyNE * NW; }   // K values per workgroup
template <typename T, int NT, int NW>
constexpr size_t gs_apply_mfma_lds() {
    return sizeof(T) * 2 * 16 * (apply_ew<NW>() + 8 + NT * 16 + 8);
}
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mfma_d4 mma_apply(double a, double b, mfma_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mfma_f4 mma_apply(float a, float b, mfma_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// DENSE: T is a full matrix (the canonical transformation and the pairing
// field's kappa combinations of the HFB step), no chunk is skipped.
template <typename T, int NT, int NW, bool DENSE = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(
    NW == 4 && NT <= 8 ? 2 : (NW == 8 ? 2 : 1)))) void k_gs_apply_mfma(const T *__restrict__ src,
                                                       T *__restrict__ dst, long K4,
                                                       const int *__restrict__ perm,
                                                       const int *__restrict__ blk_m, int max_m,
                                                       const double *__restrict__ Tm) {
    constexpr int NTH = 64 * NW, EW = apply_ew<NW>(), PS = EW + 8;
    constexpr int LT = NT * 16, PT = LT + 8;
    constexpr int NTT = (16 * LT + NTH - 1) / NTH;   // T values staged per thread
    typedef std::conditional_t<sizeof(T) == 8, mfma_d4, mfma_f4> Acc;
    extern __shared__ double lds_raw[];
    T *const lds = reinterpret_cast<T *>(lds_raw);   // psi [2][16][PS], T [2][16][PT]
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    const int m = blk_m[b];
    const long eb = (long)blockIdx.x * EW;
    const double *Tb = Tm + (long)b * max_m * max_m;
    const int *pb = perm + (long)b * max_m;
    const int kr = lane / 16, col = lane % 16;
    const int nchunk = (m + 15) / 16;
    // staging: thread tid moves K value eb + tid % EW of the chunk rows
    // tid / EW + RS j
    constexpr int RS = NTH / EW, RPT = 16 / RS;
    const int my_col = tid % EW, my_row = tid / EW;
    const long my_e = eb + my_col;
    const bool my_in = my_e < K4;
    const long my_es = my_in ? my_e : 0;
    // global -> registers (addresses clamped in range; masked when stored)
    T ps[RPT];
    T tt[NTT];
    auto gload = [&](int c) {
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int k = c * 16 + my_row + RS * j;
            ps[j] = src[(long)pb[k < m ? k : 0] * K4 + my_es];
        }
#pragma unroll
        for (int q = 0; q < NTT; ++q) {
            const int idx = tid + q * NTH, r = idx / LT, l = idx % LT, k = c * 16 + r;
            const bool in = idx < 16 * LT && k < m && l < m;
            const double v = Tb[in ? (long)k * max_m + l : 0];
            tt[q] = in ? (T)v : T(0);
        }
    };
    auto lstore = [&](int c, int buf) {
        T *P = lds + buf * 16 * PS;
        T *Tl = lds + 2 * 16 * PS + buf * 16 * PT;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = my_row + RS * j;
            P[r * PS + my_col] = (my_in && c * 16 + r < m) ? ps[j] : T(0);
        }
#pragma unroll
        for (int q = 0; q < NTT; ++q) {
            const int idx = tid + q * NTH;
            if (idx < 16 * LT) Tl[(idx / LT) * PT + idx % LT] = tt[q];
        }
    };
    Acc acc[NT][kApplyNE];
#pragma unroll
    for (int L = 0; L < NT; ++L)
#pragma unroll
        for (int E = 0; E < kApplyNE; ++E) acc[L][E] = Acc{0, 0, 0, 0};
    gload(0);
    lstore(0, 0);
    __syncthreads();
    for (int c = 0; c < nchunk; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunk) gload(c + 1);
        const T *P = lds + buf * 16 * PS + wv * (16 * kApplyNE) + col;
        const T *Tl = lds + 2 * 16 * PS + buf * 16 * PT + col;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int row = 4 * kk + kr;
            T bv[kApplyNE];
#pragma unroll
            for (int E = 0; E < kApplyNE; ++E) bv[E] = P[row * PS + 16 * E];
#pragma unroll
            for (int L = 0; L < NT; ++L) {
                if (!DENSE && L < c) continue;
                const T av = Tl[row * PT + 16 * L];
#pragma unroll
                for (int E = 0; E < kApplyNE; ++E) acc[L][E] = mma_apply(av, bv[E], acc[L][E]);
            }
        }
        if (c + 1 < nchunk) lstore(c + 1, buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int L = 0; L < NT; ++L)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int l = sizeof(T) == 8 ? 16 * L + kr + 4 * r : 16 * L + 4 * kr + r;
            if (l < m) {
                T *out = dst + (long)pb[l] * K4;
#pragma unroll
                for (int E = 0; E < kApplyNE; ++E) {
                    const long e = eb + wv * (16 * kApplyNE) + 16 * E + col;
                    if (e < K4) out[e] = (T)acc[L][E][r];
                }
            }
        }
}

// The same product for blocks of up to 128 states with the whole T (rounded
// to the storage type) in LDS for the launch: after the one barrier of its
// staging, every wave streams the states of its own 16 NE K values from HBM
// (16 lanes = 16 consecutive K values of one state, the four lane groups
// four states) and reads its A operands (T rows) from LDS; the next chunk of
// 16 states is loaded while the MFMAs of this one run.  No barrier in the
// loop, and no LDS traffic for the states.
template <typename T>
constexpr int apply_t_pitch(int nt) {
    return nt * 16 + (sizeof(T) == 8 ? 2 : 4);
}
template <typename T, int NT>
constexpr size_t apply_direct_lds() {
    return sizeof(T) * (size_t)(NT * 16) * apply_t_pitch<T>(NT);
}
// FIELD (the pairing field of the HFB step, src = the states, K4 = 4N):
// instead of storing phi_l = (Q)_l, the wave contracts it with the states
// at its K values e = c N + p: fpart[b][0][e] = sum_l phi_l[e] psi_l[e],
// fpart[b][1][e] = s_c sum_l phi_l[e] psi_l[(c^1) N + p] (s_c = -1 for odd
// c), summed over the lanes of the four output classes; phi never reaches
// HBM.
template <typename T, int NT, int NE, int NW, bool DENSE, bool FIELD = false>
__global__ __launch_bounds__(64 * NW) void k_apply_direct(const T *__restrict__ src,
                                                          T *__restrict__ dst, long K4,
                                                          const int *__restrict__ perm,
                                                          const int *__restrict__ blk_m,
                                                          int max_m,
                                                          const double *__restrict__ Tm,
                                                          double *__restrict__ fpart) {
    constexpr int LT = NT * 16, PT = apply_t_pitch<T>(NT);
    typedef std::conditional_t<sizeof(T) == 8, mfma_d4, mfma_f4> Acc;
    extern __shared__ double lds_raw[];
    T *const Tl = reinterpret_cast<T *>(lds_raw);   // T[k][l], k, l < LT (zero past m)
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    const int m = blk_m[b];
    const double *Tb = Tm + (long)b * max_m * max_m;
    const int *pb = perm + (long)b * max_m;
    for (int idx = tid; idx < LT * LT; idx += 64 * NW) {
        const int k = idx / LT, l = idx % LT;
        const bool in = k < m && l < m;
        const double v = Tb[in ? (long)k * max_m + l : 0];
        Tl[k * PT + l] = in ? (T)v : T(0);
    }
    const int kr = lane / 16, col = lane % 16;
    const long e0 = ((long)blockIdx.x * NW + wv) * 16 * NE + col;
    bool ein[NE];
    long es[NE];
#pragma unroll
    for (int E = 0; E < NE; ++E) {
        ein[E] = e0 + 16 * E < K4;
        es[E] = ein[E] ? e0 + 16 * E : 0;
    }
    const int nchunk = (m + 15) / 16;
    // raw[kk][E]: state 16 c + 4 kk + kr at K value es[E] (state 0 past m,
    // zeroed when widened; T's row is zero there anyway)
    T raw[4][NE], bv[4][NE];
    auto gload = [&](int c) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = 16 * c + 4 * kk + kr;
            const T *row = src + (long)pb[k < m ? k : 0] * K4;
#pragma unroll
            for (int E = 0; E < NE; ++E) raw[kk][E] = row[es[E]];
        }
    };
    auto take = [&](int c) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const bool kin = 16 * c + 4 * kk + kr < m;
#pragma unroll
            for (int E = 0; E < NE; ++E) bv[kk][E] = (kin && ein[E]) ? raw[kk][E] : T(0);
        }
    };
    Acc acc[NT][NE];
#pragma unroll
    for (int L = 0; L < NT; ++L)
#pragma unroll
        for (int E = 0; E < NE; ++E) acc[L][E] = Acc{0, 0, 0, 0};
    gload(0);
    __syncthreads();   // T staged
    take(0);
    for (int c = 0; c < nchunk; ++c) {
        if (c + 1 < nchunk) gload(c + 1);
        __builtin_amdgcn_sched_barrier(0);
        const T *Tr = Tl + (16 * c + kr) * PT + col;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
            for (int L = 0; L < NT; ++L) {
                if (!DENSE && L < c) continue;
                const T av = Tr[4 * kk * PT + 16 * L];
#pragma unroll
                for (int E = 0; E < NE; ++E) acc[L][E] = mma_apply(av, bv[kk][E], acc[L][E]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        take(c + 1);   // after the last chunk: unused
    }
    if constexpr (FIELD) {
        const long N = K4 / 4;
        double re[NE], im[NE];
        long ex[NE];
#pragma unroll
        for (int E = 0; E < NE; ++E) {
            re[E] = im[E] = 0.0;
            const long e = es[E], cc = e / N;
            ex[E] = (cc ^ 1) * N + (e - cc * N);
        }
#pragma unroll
        for (int L = 0; L < NT; ++L)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int l = sizeof(T) == 8 ? 16 * L + kr + 4 * r : 16 * L + 4 * kr + r;
                // (past m: phi is 0, T's columns being 0 there; state 0 read)
                const T *row = src + (long)pb[l < m ? l : 0] * K4;
#pragma unroll
                for (int E = 0; E < NE; ++E) {
                    const double ph = (double)acc[L][E][r];
                    re[E] += ph * (double)row[es[E]];
                    im[E] += ph * (double)row[ex[E]];
                }
            }
#pragma unroll
        for (int E = 0; E < NE; ++E) {
            re[E] += __shfl_xor(re[E], 16, 64);
            im[E] += __shfl_xor(im[E], 16, 64);
            re[E] += __shfl_xor(re[E], 32, 64);
            im[E] += __shfl_xor(im[E], 32, 64);
            if (kr == 0 && ein[E]) {
                const long e = es[E];
                fpart[(2L * b) * K4 + e] = re[E];
                fpart[(2L * b + 1) * K4 + e] = ((e / N) & 1) ? -im[E] : im[E];
            }
        }
        return;
    }
#pragma unroll
    for (int L = 0; L < NT; ++L)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int l = sizeof(T) == 8 ? 16 * L + kr + 4 * r : 16 * L + 4 * kr + r;
            if (l < m) {
                T *out = dst + (long)pb[l] * K4;
#pragma unroll
                for (int E = 0; E < NE; ++E)
                    if (ein[E]) out[e0 + 16 * E] = (T)acc[L][E][r];
            }
        }
}

// Pairing field from the partials of k_apply_direct<FIELD>: field[2 it][p] =
// dfac sum_{blocks of it} sum_c fpart[b][0][c N + p], field[2 it + 1][p] =
// -dfac sum_b sum_c fpart[b][1][c N + p] (block order, then components).
__global__ void k_field_from_parts(HfbK k, const double *__restrict__ fpart, long N,
                                   const double *__restrict__ dfac,
                                   double *__restrict__ field) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    if (p >= N) return;
    double fr = 0.0, fi = 0.0;
    for (int b = 0; b < k.nblk; ++b) {
        if (k.blk[b].it != it) continue;
        const double *f = fpart + 2 * b * 4 * N;
        fr += f[p] + f[N + p] + f[2 * N + p] + f[3 * N + p];
        f += 4 * N;
        fi += f[p] + f[N + p] + f[2 * N + p] + f[3 * N + p];
    }
    const double d = dfac[it * N + p];
    field[(2 * it) * N + p] = fr * d;
    field[(2 * it + 1) * N + p] = -fi * d;
}

template <typename T, bool DENSE>
static void apply_mfma_launch(const T *src, T *dst, long K4, const int *perm, const int *blk_m,
                              int nblk, int max_m, int mpad_max, const double *Tm,
                              hipStream_t st) {
    // (the double-buffered LDS tiles exceed the default 64 KiB)
#define MOCCA_APPLY_NT(K, W)                                                                \
    case K: {                                                                               \
        constexpr size_t shm = gs_apply_mfma_lds<T, K, W>();                                \
        const void *fn = (const void *)k_gs_apply_mfma<T, K, W, DENSE>;                     \
        static const bool lds_set = require_lds(fn, shm, "k_gs_apply_mfma");                \
        if (!lds_set) require_lds(fn, shm, "k_gs_apply_mfma");                              \
        const dim3 g((K4 + apply_ew<W>() - 1) / apply_ew<W>(), nblk);                       \
        hipLaunchKernelGGL((k_gs_apply_mfma<T, K, W, DENSE>), g, dim3(64 * W), shm, st, src, \
                           dst, K4, perm, blk_m, max_m, Tm);                                \
        break;                                                                              \
    }
    // up to 8 row tiles: T resident in LDS (k_apply_direct)
#define MOCCA_APPLY_DIRECT(K, NE, W)                                                        \
    case K: {                                                                               \
        constexpr size_t shm = apply_direct_lds<T, K>();                                    \
        const void *fn = (const void *)k_apply_direct<T, K, NE, W, DENSE>;                  \
        static const bool lds_set = require_lds(fn, shm, "k_apply_direct");                 \
        if (!lds_set) require_lds(fn, shm, "k_apply_direct");                               \
        const dim3 g((K4 + 16 * NE * W - 1) / (16 * NE * W), nblk);                         \
        hipLaunchKernelGGL((k_apply_direct<T, K, NE, W, DENSE>), g, dim3(64 * W), shm, st,   \
                           src, dst, K4, perm, blk_m, max_m, Tm, nullptr);                  \
        return;                                                                             \
    }
    constexpr int DNE = sizeof(T) == 8 ? 1 : 2, DNW = sizeof(T) == 8 ? 16 : 8;
    switch (mpad_max / 16) {
        MOCCA_APPLY_DIRECT(1, DNE, DNW) MOCCA_APPLY_DIRECT(2, DNE, DNW)
        MOCCA_APPLY_DIRECT(3, DNE, DNW) MOCCA_APPLY_DIRECT(4, DNE, DNW)
        MOCCA_APPLY_DIRECT(5, DNE, DNW) MOCCA_APPLY_DIRECT(6, DNE, DNW)
        MOCCA_APPLY_DIRECT(7, DNE, DNW) MOCCA_APPLY_DIRECT(8, DNE, DNW)
        default: break;
    }
#undef MOCCA_APPLY_DIRECT
    // 9-12 row tiles: T staged in LDS chunk by chunk; 11-12 row tiles do
    // not fit two waves per SIMD, and keep four
    switch (mpad_max / 16) {
        MOCCA_APPLY_NT(9, 8) MOCCA_APPLY_NT(10, 8) MOCCA_APPLY_NT(11, 4) MOCCA_APPLY_NT(12, 4)
        default: return;
    }
#undef MOCCA_APPLY_NT
}

template <typename T>
void launch_gs_apply_mfma(const T *src, T *dst, long K4, const int *perm, const int *blk_m,
                          int nblk, int max_m, int mpad_max, const double *Tm,
                          hipStream_t st) {
    apply_mfma_launch<T, false>(src, dst, K4, perm, blk_m, nblk, max_m, mpad_max, Tm, st);
}

template <typename T>
bool launch_field_apply(const HfbK &k, const T *psi, long N, const int *list,
                        const int *blk_m, int max_m, int mpad_max, const double *Tm,
                        const double *dfac, double *fpart, double *field, hipStream_t st) {
    const long K4 = 4 * N;
    const int nblk = k.nblk;
    constexpr int NE = sizeof(T) == 8 ? 1 : 2, W = sizeof(T) == 8 ? 16 : 8;
#define MOCCA_FIELD_APPLY(K)                                                                \
    case K: {                                                                               \
        constexpr size_t shm = apply_direct_lds<T, K>();                                    \
        const void *fn = (const void *)k_apply_direct<T, K, NE, W, true, true>;             \
        static const bool lds_set = require_lds(fn, shm, "k_apply_direct");                 \
        if (!lds_set) require_lds(fn, shm, "k_apply_direct");                               \
        const dim3 g((K4 + 16 * NE * W - 1) / (16 * NE * W), nblk);                         \
        hipLaunchKernelGGL((k_apply_direct<T, K, NE, W, true, true>), g, dim3(64 * W), shm,  \
                           st, psi, nullptr, K4, list, blk_m, max_m, Tm, fpart);            \
        break;                                                                              \
    }
    switch (mpad_max / 16) {
        MOCCA_FIELD_APPLY(1) MOCCA_FIELD_APPLY(2) MOCCA_FIELD_APPLY(3) MOCCA_FIELD_APPLY(4)
        MOCCA_FIELD_APPLY(5) MOCCA_FIELD_APPLY(6) MOCCA_FIELD_APPLY(7) MOCCA_FIELD_APPLY(8)
        default: return false;
    }
#undef MOCCA_FIELD_APPLY
    hipLaunchKernelGGL(k_field_from_parts, dim3((N + 255) / 256, 2), dim3(256), 0, st, k,
                       fpart, N, dfac, field);
    return true;
}

template <typename T>
void launch_dense_apply_mfma(const T *src, T *dst, long K4, const int *list, const int *blk_m,
                             int nblk, int max_m, int mpad_max, const double *Tm,
                             hipStream_t st) {
    apply_mfma_launch<T, true>(src, dst, K4, list, blk_m, nblk, max_m, mpad_max, Tm, st);
}

// G[b][i][j] = G[b][j][i] = dv sum_split of the tile element holding
// (i, j), i <= j: one wave per element, lanes over the splits, then a fixed
// shuffle tree.
__global__ __launch_bounds__(kBlock) void k_gram_mfma_reduce(const double *__restrict__ part,
                                                             const int *__restrict__ blk_m,
                                                             int max_m, int nt, int nsplit,
                                                             double dv, double *__restrict__ G) {
    const int b = blockIdx.y;
    const int m = blk_m[b];
    const long pidx = (long)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (pidx >= (long)m * (m + 1) / 2) return;
    // pidx -> (i, j), i <= j, row-major over the upper triangle
    int i = 0;
    long rem = pidx;
    while (rem >= m - i) { rem -= m - i; ++i; }
    const int j = i + (int)rem;
    const int I = i / 16, J = j / 16, li = i % 16, lj = j % 16;
    const int t = I * nt - I * (I - 1) / 2 + (J - I);
    const int ntile = nt * (nt + 1) / 2;
    // lane l, register r of a tile hold row l/16 + 4r, column l%16
    const long off = (long)t * 256 + (lj + 16 * (li % 4)) * 4 + li / 4;
    const double *P = part + (long)b * nsplit * ntile * 256 + off;
    double s = 0;
    for (int w = lane; w < nsplit; w += 64) s += P[(long)w * ntile * 256];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        G[((long)b * max_m + i) * max_m + j] = s * dv;
        G[((long)b * max_m + j) * max_m + i] = s * dv;
    }
}

template <typename T>
void launch_gram_mfma(const T *psi, long K4, const int *blk_states, const int *blk_m, int nblk,
                      int max_m, int mpad_max, double *part, int nsplit, double dv, double *G,
                      hipStream_t st) {
    const int nt = mpad_max / 16;
    const dim3 g(nsplit, nblk), g2(nsplit, nblk, 2), blk(64);
#define MOCCA_GRAM_NT(K)                                                                    \
    case K:                                                                                 \
        hipLaunchKernelGGL((k_gram_mfma<T, K, K, true>), g, blk, 0, st, psi, K4, blk_states, \
                           blk_m, max_m, part, nsplit, nt);                                \
        break;
#define MOCCA_GRAM_NT2(K, H)                                                                \
    case K:                                                                                 \
        hipLaunchKernelGGL((k_gram_mfma<T, H, H, true>), g2, blk, 0, st, psi, K4,          \
                           blk_states, blk_m, max_m, part, nsplit, nt);                    \
        hipLaunchKernelGGL((k_gram_mfma<T, H, H, false>), g, blk, 0, st, psi, K4,          \
                           blk_states, blk_m, max_m, part, nsplit, nt);                    \
        break;
    switch (nt) {
        MOCCA_GRAM_NT(1) MOCCA_GRAM_NT(2) MOCCA_GRAM_NT(3) MOCCA_GRAM_NT(4)
        MOCCA_GRAM_NT(5) MOCCA_GRAM_NT(6) MOCCA_GRAM_NT(7) MOCCA_GRAM_NT(8)
        MOCCA_GRAM_NT2(9, 5) MOCCA_GRAM_NT2(10, 5)
        default: return;
    }
#undef MOCCA_GRAM_NT2
#undef MOCCA_GRAM_NT
    const long pairs = (long)max_m * (max_m + 1) / 2;
    const long per = kBlock / 64;
    hipLaunchKernelGGL(k_gram_mfma_reduce, dim3((pairs + per - 1) / per, nblk), dim3(kBlock), 0,
                       st, part, blk_m, max_m, nt, nsplit, dv, G);
}

// Explicit instantiations.
#define MOCCA_INST(T)                                                                   \
    template void launch_derive_all<T>(const MeshK<T> &, const T *, const int *, int,    \
                                       double *, double *, hipStream_t);                \
    template void launch_hpsi_step<T>(const MeshK<T> &, const T *, T *, T *, T *,        \
                                      const T *, const int *, const int *,               \
                                      const Chunk *, int, int, T, T, T, T, int,          \
                                      double *, int, hipStream_t, const T *, const T *); \
    template void launch_lag_derivs<T>(const MeshK<double> &, const T *, const int *,    \
                                       int, T *, hipStream_t);                           \
    template void launch_density<T>(const MeshK<T> &, const T *, const int *,            \
                                    const double *, const int *, const Chunk *, int,     \
                                    int, T *, hipStream_t);                              \
    template void launch_density_reduce<T>(const T *, const Chunk *, int, int, int,      \
                                           double *, hipStream_t);                       \
    template void launch_gram<T>(const T *, long, const int *, const int *, int, int,    \
                                 int, int, double *, int, hipStream_t);                  \
    template void launch_gs_apply_mfma<T>(const T *, T *, long, const int *, const int *, \
                                          int, int, int, const double *, hipStream_t);    \
    template void launch_dense_apply_mfma<T>(const T *, T *, long, const int *,           \
                                             const int *, int, int, int, const double *,  \
                                             hipStream_t);                                \
    template bool launch_field_apply<T>(const HfbK &, const T *, long, const int *,       \
                                        const int *, int, int, const double *,           \
                                        const double *, double *, double *, hipStream_t); \
    template void launch_gs_apply<T>(const T *, T *, T *, long, const int *,             \
                                     const int *, int, int, const double *, int,         \
                                     hipStream_t);                                       \
    template void launch_gram_mfma<T>(const T *, long, const int *, const int *, int,    \
                                      int, int, double *, int, double, double *,         \
                                      hipStream_t);

MOCCA_INST(double)
MOCCA_INST(float)

// Dynamic LDS available to k_cholesky_reg (the device limit less its static
// arrays), set as the kernel's limit once.
static size_t cholesky_reg_lds_limit() {
    static const size_t lim = [] {
        int dev = 0, l = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&l, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess)
            return (size_t)0;
        const size_t stat = kMaxBlock * (sizeof(int) + sizeof(double)) +
                            2 * kCholRegMax * sizeof(double);
        if ((size_t)l <= stat) return (size_t)0;
        const size_t dyn = (size_t)l - stat;
        const void *fns[] = {(const void *)k_cholesky_reg<1>, (const void *)k_cholesky_reg<2>};
        for (const void *f : fns)
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) !=
                hipSuccess)
                return (size_t)0;
        return dyn;
    }();
    return lim;
}

void launch_cholesky_ld(const double *G, const int *blk_states, const int *blk_m,
                        int nblk, int max_m, int mmax, const double *E, double *Tm,
                        int *perm, double *work, int *status, size_t lds_limit,
                        hipStream_t st) {
    // up to kCholRegMax states: the lower triangle in registers and L packed
    // in LDS
    const size_t lreg = sizeof(double) * (size_t)mmax * (mmax + 1) / 2;
    if (mmax <= kCholRegMax && lreg <= cholesky_reg_lds_limit()) {
        if (mmax * (mmax + 1) / 2 <= 1024)
            hipLaunchKernelGGL(k_cholesky_reg<1>, dim3(nblk), dim3(1024), lreg, st, G,
                               blk_states, blk_m, max_m, E, Tm, perm, status);
        else
            hipLaunchKernelGGL(k_cholesky_reg<2>, dim3(nblk), dim3(1024), lreg, st, G,
                               blk_states, blk_m, max_m, E, Tm, perm, status);
        return;
    }
    // the whole block in LDS while it fits (up to ~137 states), else the
    // packed lower triangle (up to kCholPackedMax), else global memory
    const size_t lds = sizeof(double) * (size_t)mmax * (mmax + 1);
    if (lds <= lds_limit) {
        hipLaunchKernelGGL(k_cholesky_lds, dim3(nblk), dim3(1024), lds, st, G, blk_states,
                           blk_m, max_m, E, Tm, perm, status);
        return;
    }
    const size_t packed = sizeof(double) * (size_t)mmax * (mmax + 1) / 2;
    if (mmax <= kCholPackedMax && packed <= lds_limit) {
        hipLaunchKernelGGL(k_cholesky_packed, dim3(nblk), dim3(1024), packed, st, G,
                           blk_states, blk_m, max_m, E, Tm, perm, status);
        return;
    }
    hipLaunchKernelGGL(k_cholesky, dim3(nblk), dim3(kBlock), 0, st, G, blk_states,
                       blk_m, max_m, E, Tm, perm, work, status);
}

__global__ void k_gs_norms(const double *__restrict__ Tm, const int *__restrict__ perm,
                           const int *__restrict__ blk_m, int max_m, double *__restrict__ nrm) {
    const int b = blockIdx.x;
    const double *Tb = Tm + (long)b * max_m * max_m;
    for (int r = threadIdx.x; r < blk_m[b]; r += blockDim.x) {
        const double d = Tb[(long)r * max_m + r];   // 1 / L[r][r]
        nrm[perm[(long)b * max_m + r]] = 1.0 / (d * d);
    }
}

void launch_gs_norms(const double *Tm, const int *perm, const int *blk_m, int nblk,
                     int max_m, double *nrm, hipStream_t st) {
    hipLaunchKernelGGL(k_gs_norms, dim3(nblk), dim3(64), 0, st, Tm, perm, blk_m, max_m, nrm);
}

size_t cholesky_lds_limit() {
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
        hipSuccess)
        return 0;
    // static LDS of the kernel (rank, dinv)
    const size_t stat = kMaxBlock * (sizeof(int) + sizeof(double));
    if ((size_t)lim <= stat) return 0;
    const size_t dyn = (size_t)lim - stat;
    if (hipFuncSetAttribute((const void *)k_cholesky_lds,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_cholesky_packed,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
        return 0;
    return dyn;
}

}  // namespace mocca
