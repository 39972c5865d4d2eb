// Z-marching finite-difference kernels for the EV8 octant (gfx950, wave64).
//
// One workgroup owns a y-tile (full x lines, TY rows) of one state and walks
// along z.  The 9-point z stencil comes from a per-thread register queue of
// 9 planes per spinor component; the x and y stencils read the current plane
// from LDS, where the tile is stored with 4-row y halos.  Reflection ghosts
// (below the symmetry plane) and zeros (beyond the box) follow the EV8 rules
// of derx/dery/derz and lapla_{X,Y,Z}_ev8 (OptimizedDerivatives.f90:46-113,
// 230-304, 362-430, 617-795).  Each state is read from HBM once per launch.
#include "mocca_kernels.h"

namespace mocca {

namespace {

__device__ __forceinline__ int comp_sign_x(int c, int S) { return (c == 0 || c == 3) ? S : -S; }
__device__ __forceinline__ int comp_sign_y(int c, int TS) { return (c == 0 || c == 2) ? TS : -TS; }
__device__ __forceinline__ int comp_sign_z(int c, int S, int P) { return c < 2 ? S * P : -S * P; }

// 7-point first and 9-point second derivative from the 9 values v[0..8]
// centred on v[4].
template <typename T>
__device__ __forceinline__ void fd_pair(const MeshK<T> &m, const T v[9], T &d1, T &d2) {
    d1 = (T(45) * (v[5] - v[3]) - T(9) * (v[6] - v[2]) + (v[7] - v[1])) * m.inv60dx;
    d2 = lap9(m, v);
}

__device__ __forceinline__ void block_sum3_fd(double &a, double &b, double &c, double *sh) {
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

}  // namespace

// LDS row layout of the z-marching kernels: one row holds x = -G .. nx+G-1
// at offsets 0 .. nx+2G-1: the G left cells are the mirrored reflection
// ghosts (with the component's sign), the G right cells are zeros, so the
// x stencils read without branches.
//
// STALE: the values come from vsrc instead of psi, whose stencils give the
// derivatives (the first GradDesc after an AlternateStep, which changes the
// values but not the stored derivatives, ImaginaryTime.f90:306-437).
// KIN: only the partial sums of <psi|Lap psi> (the kinetic sums of a basis
// that no GradDesc produced, such as the canonical states under HFB); the
// potentials and the first derivatives are not touched.
template <typename T, int HMAX, bool TRC, bool STALE, bool WA, bool KIN = false>
__global__ __launch_bounds__(kBlock) void k_hpsi_fd(
    MeshK<T> m, const T *__restrict__ psi, T *__restrict__ psi_out,
    T *__restrict__ upd, T *__restrict__ hout, const T *__restrict__ F,
    const int *__restrict__ qn, const Task *__restrict__ tasks, int TY,
    int ntile_y, T kin0, T kin1, T dth, T mu, int mode,
    double *__restrict__ eparts, const T *__restrict__ vsrc) {
    constexpr int G = 4;  // ghost width; HMAX = halo loads per thread (32 nx / 256)
    constexpr bool trc = TRC;
    extern __shared__ double smem_d[];
    T *sp = reinterpret_cast<T *>(smem_d);  // [4][TY+8][nx+8]
    __shared__ double red[3 * kBlock / 64];
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const long N = m.N, plane = (long)nx * ny;
    const int W = nx + 2 * G;
    const Task tk = tasks[blockIdx.x];
    const int s = tk.state, tile = tk.tile;
    const int it = (qn[4 * s] + 3) / 2 - 1;
    const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
    const T kin = it == 0 ? kin0 : kin1;
    const int tid = threadIdx.x, tx = tid % nx, tyl = tid / nx;
    const int y0 = tile * TY, y = y0 + tyl;
    const bool active = (tyl < TY) && (y < ny);
    const int rows = TY + 8;
    T sgx[4], sgy[4], sgz[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        sgx[c] = T(comp_sign_x(c, S));
        sgy[c] = T(comp_sign_y(c, TS));
        sgz[c] = T(comp_sign_z(c, S, P));
    }
    const T *ps = psi + (long)s * 4 * N;
    const long col = (long)(active ? y : 0) * nx + tx;
    T q[4][9];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int zz = k - 4;
            const int zi = zz < 0 ? -1 - zz : (zz < nz ? zz : nz - 1);
            const T val = ps[c * N + zi * plane + col];
            q[c][k] = !active || zz >= nz ? T(0) : (zz < 0 ? sgz[c] * val : val);
        }
    const T *Fu = F + (long)(kOffU + it) * N, *Fb = F + (long)(kOffB + it) * N;
    const T *Fnb = F + (long)(kOffNB + 3 * it) * N, *Fw = F + (long)(kOffW + 9 * it) * N;
    const T *Fdw = F + (long)(kOffDW + 3 * it) * N, *Fs = F + (long)(kOffS + 3 * it) * N;
    const T *Fa = F + (long)(kOffA + 3 * it) * N, *Fda = F + (long)(kOffDA + it) * N;
    T *po = psi_out ? psi_out + (long)s * 4 * N : nullptr;
    T *us = upd ? upd + (long)s * 4 * N : nullptr;
    double eacc = 0, dacc = 0, kacc = 0;
    const int nhalo = 4 * 8 * nx;
    for (int z = 0; z < nz; ++z) {
        // ---- all global loads of this plane first ----
        T hv[HMAX];
#pragma unroll
        for (int r = 0; r < HMAX; ++r) {
            const int idx = tid + r * kBlock;
            const int c = idx / (8 * nx), rem = idx % (8 * nx);
            const int hr = rem / nx, xx = rem % nx;
            const int yy = y0 - 4 + (hr < 4 ? hr : TY + hr);
            const int yi = yy < 0 ? -1 - yy : (yy < ny ? yy : ny - 1);
            const int cc = idx < nhalo ? c : 0;
            hv[r] = ps[cc * N + z * plane + yi * nx + xx];
        }
        const long p = (long)z * plane + col;
        const int zn = z + 5 < nz ? z + 5 : nz - 1;
        T nxt[4], uo[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 4; ++c) nxt[c] = ps[c * N + zn * plane + col];
        const T fu = Fu[p], fb = Fb[p] + kin;
        const T nb0 = Fnb[p], nb1 = Fnb[N + p], nb2 = Fnb[2 * N + p];
        // W[nu][mu]; WA: antisymmetric with a zero diagonal (no tensor
        // terms), only W[0][1], W[0][2], W[1][2] are read
        T fw[3][3];
        if constexpr (WA) {
            fw[0][1] = Fw[1 * N + p];
            fw[0][2] = Fw[2 * N + p];
            fw[1][2] = Fw[5 * N + p];
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int mu_ = 0; mu_ < 3; ++mu_) fw[a][mu_] = Fw[(a * 3 + mu_) * N + p];
        }
        const T hx = T(0.5) * Fdw[p], hy = T(0.5) * Fdw[N + p], hz = T(0.5) * Fdw[2 * N + p];
        T s0 = 0, s1 = 0, s2 = 0, a0 = 0, a1 = 0, a2 = 0, hda = 0;
        if (!trc) {
            s0 = Fs[p]; s1 = Fs[N + p]; s2 = Fs[2 * N + p];
            a0 = Fa[p]; a1 = Fa[N + p]; a2 = Fa[2 * N + p];
            hda = T(0.5) * Fda[p];
        }
        if (mode & kStep) {
            // heavy-ball memory u = psi - psi_prev (ImaginaryTime.f90:497-499):
            // the previous states are kept instead of u
#pragma unroll
            for (int c = 0; c < 4; ++c) uo[c] = us[c * N + p];
        }
        // ---- current plane -> LDS with ghosts and pads ----
        if (tyl < TY) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                T *row = sp + (c * rows + tyl + 4) * W;
                const T v = active ? q[c][4] : T(0);
                row[tx + G] = v;
                if (tx < G) row[G - 1 - tx] = sgx[c] * v;
                if (tx >= nx - G) row[tx + 2 * G] = T(0);
            }
        }
#pragma unroll
        for (int r = 0; r < HMAX; ++r) {
            const int idx = tid + r * kBlock;
            if (idx < nhalo) {
                const int c = idx / (8 * nx), rem = idx % (8 * nx);
                const int hr = rem / nx, xx = rem % nx;
                const int rr = hr < 4 ? hr : TY + hr;
                const int yy = y0 - 4 + rr;
                const T v = yy < 0 ? sgy[c] * hv[r] : (yy < ny ? hv[r] : T(0));
                T *row = sp + (c * rows + rr) * W;
                row[xx + G] = v;
                if (xx < G) row[G - 1 - xx] = sgx[c] * v;
                if (xx >= nx - G) row[xx + 2 * G] = T(0);
            }
        }
        __syncthreads();
        if (active) {
            // Component by component: derivatives, then their contributions to
            // h (ActionOfBNew, ActionOfU: MeanFields.f90:792-851), to the
            // spin-orbit sum aw (ActionOfW, :1029-1065) and to the current
            // term aa (ActionOfA, :1107-1129).
            T h[4] = {0, 0, 0, 0}, aw[4] = {0, 0, 0, 0}, aa[4] = {0, 0, 0, 0};
            T val[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if constexpr (STALE) val[c] = vsrc[(long)s * 4 * N + c * N + p];
                else val[c] = q[c][4];
            }
            if constexpr (KIN) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T *row = sp + (c * rows + tyl + 4) * W + tx + G;
                    T w[9];
#pragma unroll
                    for (int k = -4; k <= 4; ++k) w[k + 4] = row[k];
                    const T lx = lap9(m, w);
#pragma unroll
                    for (int k = -4; k <= 4; ++k) w[k + 4] = row[k * W];
                    const T ly = lap9(m, w);
                    const T lc = lx + ly + lap9(m, q[c]);
                    kacc += (double)val[c] * (double)lc;
                }
            } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T *row = sp + (c * rows + tyl + 4) * W + tx + G;
                T w[9], dxc, dyc, dzc, lx, ly, lz;
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = row[k];
                fd_pair(m, w, dxc, lx);
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = row[k * W];
                fd_pair(m, w, dyc, ly);
                fd_pair(m, q[c], dzc, lz);
                const T lc = lx + ly + lz, vc = val[c];
                kacc += (double)vc * (double)lc;
                h[c] += -fb * lc - (nb0 * dxc + nb1 * dyc + nb2 * dzc) + fu * vc;
                T ex, ey, ez;
                if constexpr (WA) {
                    ex = fw[0][1] * dyc + fw[0][2] * dzc + hx * vc;
                    ey = -fw[0][1] * dxc + fw[1][2] * dzc + hy * vc;
                    ez = -fw[0][2] * dxc - fw[1][2] * dyc + hz * vc;
                } else {
                    ex = fw[0][0] * dxc + fw[0][1] * dyc + fw[0][2] * dzc + hx * vc;
                    ey = fw[1][0] * dxc + fw[1][1] * dyc + fw[1][2] * dzc + hy * vc;
                    ez = fw[2][0] * dxc + fw[2][1] * dyc + fw[2][2] * dzc + hz * vc;
                }
                if (c == 0) { aw[0] += ez; aw[2] += ex; aw[3] += ey; }
                if (c == 1) { aw[1] += ez; aw[2] -= ey; aw[3] += ex; }
                if (c == 2) { aw[0] += ex; aw[1] -= ey; aw[2] -= ez; }
                if (c == 3) { aw[0] += ey; aw[1] += ex; aw[3] -= ez; }
                if (!trc) {
                    // S . sigma psi (ActionOfS, :1131-1147)
                    if (c == 0) { h[0] += s2 * vc; h[2] += s0 * vc; h[3] += s1 * vc; }
                    if (c == 1) { h[1] += s2 * vc; h[2] -= s1 * vc; h[3] += s0 * vc; }
                    if (c == 2) { h[0] += s0 * vc; h[1] -= s1 * vc; h[2] -= s2 * vc; }
                    if (c == 3) { h[0] += s1 * vc; h[1] += s0 * vc; h[3] -= s2 * vc; }
                    aa[c] = a0 * dxc + a1 * dyc + a2 * dzc + hda * vc;
                }
            }
            h[0] += aw[1]; h[1] -= aw[0]; h[2] += aw[3]; h[3] -= aw[2];
            if (!trc) { h[0] += aa[1]; h[1] -= aa[0]; h[2] += aa[3]; h[3] -= aa[2]; }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T vc = val[c];
                eacc += (double)vc * (double)h[c];
                dacc += (double)h[c] * (double)h[c];
            }
            }   // !KIN
            if (mode & kWriteH) {
#pragma unroll
                for (int c = 0; c < 4; ++c) hout[c * N + p] = h[c];
            }
            if (mode & kStep) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T vc = val[c];
                    const T u = q[c][4] - uo[c];
                    po[c * N + p] = vc + (-(dth * h[c]) + mu * u);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int k = 0; k < 8; ++k) q[c][k] = q[c][k + 1];
            q[c][8] = (active && z + 5 < nz) ? nxt[c] : T(0);
        }
    }
    if (eparts) {
        block_sum3_fd(eacc, dacc, kacc, red);
        if (tid == 0) {
            double *ep = eparts + ((long)s * ntile_y + tile) * 3;
            ep[0] = eacc;
            ep[1] = dacc;
            ep[2] = kacc;
        }
    }
}

// ---------------------------------------------------------------------------
// The GradDesc step (kStep, no stale derivatives) with a two-stage pipeline:
// while plane z is computed, the y-halo rows and the potentials of plane z+1
// travel to LDS by direct-to-LDS loads (global_load_lds_dwordx4: lane L of a
// wave moves 16 bytes to LDS base + 16 L), so no registers hold data in
// flight and the loads overlap the arithmetic.  One barrier per plane: the
// compiler waits for the outstanding direct-to-LDS loads there.
//   LDS (elements of T), each double-buffered by plane parity:
//     X[4][TY][nx+8]   the tile's own rows with x ghosts (reflection, zeros),
//                      written from the register queue (x stencil);
//     Y[4][TY+8][nx]   the same rows plus 4 halo rows above and below (y
//                      stencil); halo rows below the symmetry plane arrive
//                      unsigned from the mirrored rows and rows past the box
//                      from the last row: the y stencil of the edge tiles
//                      applies the reflection sign / zero when it reads them;
//     P[NF][TNP]       the potentials of the tile (TNP: TY nx rounded up to
//                      whole 1 KiB loads).
// The new values of plane z are stored at the start of plane z+1, so their
// completion is not waited for at the next barrier.  Requires nx * sizeof(T)
// to be a multiple of 16 (whole 16-byte lane loads inside a row).
typedef __attribute__((address_space(3))) void lds_raw_t;

// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt = 0 only:
// vmcnt bits [3:0] and [15:14] zero, expcnt [6:4] and lgkmcnt [11:8] at
// their "no wait" maxima.
constexpr unsigned kWaitVmcnt0 = 0x0F70;

template <typename T>
__device__ __forceinline__ void dma16(const T *g, T *lds_base) {
    __builtin_amdgcn_global_load_lds((const void *)g, (lds_raw_t *)lds_base, 16, 0, 0);
}

// Potential j of the pipelined step's LDS list: U, B, grad B (3), W (9:
// W[nu][mu] at 5 + 3 nu + mu), div W (3), S (3), A (3), div A.  WA: W is
// antisymmetric with a zero diagonal (no tensor terms: W_{mu nu} =
// -eps_{k mu nu} w_k), so only W[0][1], W[0][2], W[1][2] are moved.
template <bool WA>
__device__ __forceinline__ int pot_slot(int j, int it) {
    if (j == 0) return kOffU + it;
    if (j == 1) return kOffB + it;
    if (j < 5) return kOffNB + 3 * it + (j - 2);
    if constexpr (WA) {
        if (j < 8) return kOffW + 9 * it + (j == 5 ? 1 : (j == 6 ? 2 : 5));
        j += 6;
    }
    if (j < 14) return kOffW + 9 * it + (j - 5);
    if (j < 17) return kOffDW + 3 * it + (j - 14);
    if (j < 20) return kOffS + 3 * it + (j - 17);
    if (j < 23) return kOffA + 3 * it + (j - 20);
    return kOffDA + it;
}

template <typename T, bool TRC, bool WA>
constexpr int hpsi_dma_fields() { return (TRC ? 17 : 24) - (WA ? 6 : 0); }

template <typename T, bool TRC, bool WA>
size_t hpsi_dma_lds(int nx, int TY) {
    constexpr int EPI = 1024 / (int)sizeof(T);
    const int TNP = (TY * nx + EPI - 1) / EPI * EPI;
    return sizeof(T) * ((size_t)2 * 4 * TY * (nx + 8) + (size_t)2 * 4 * (TY + 8) * nx +
                        (size_t)2 * hpsi_dma_fields<T, TRC, WA>() * TNP);
}

// Two threads per point: lanes 0-31 of a wave own spinor components 0 and
// 1 of 32 points, lanes 32-63 components 2 and 3 of the same points.  Each
// thread forms the derivatives of its two components and their
// contributions to all four components of h psi; the partners then swap the
// partial sums of each other's components (a lane-distance-32 shuffle).  This
// halves the register queue per thread, so two waves share a SIMD.
constexpr int kDmaThreads = 2 * kBlock;

// KIN: only the sums <psi|Lap psi> of every state (CompKinetic's inner
// products), no potentials, no step, no stores of states.
template <typename T, bool TRC, bool WA, bool KIN = false>
__global__ __launch_bounds__(kDmaThreads) void k_hpsi_fd_dma(
    MeshK<T> m, const T *__restrict__ psi, T *__restrict__ psi_out,
    const T *__restrict__ prev, const T *__restrict__ F, const int *__restrict__ qn,
    const Task *__restrict__ tasks, int TY, int ntile_y, int nzs, T kin0, T kin1, T dth,
    T mu, double *__restrict__ eparts) {
    constexpr int G = 4;
    constexpr bool trc = TRC;
    constexpr int NF = KIN ? 0 : hpsi_dma_fields<T, TRC, WA>();
    constexpr int FD = WA ? 8 : 14;   // first div W slot; S, A, div A follow
    constexpr int NW = kDmaThreads / 64;       // waves
    constexpr int VPL = 16 / (int)sizeof(T);   // values per lane of one 16-byte load
    constexpr int EPI = 64 * VPL;              // values per wave load (1 KiB)
    extern __shared__ double smem_d[];
    __shared__ double red[3 * NW];
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const long N = m.N, plane = (long)nx * ny;
    const int W = nx + 2 * G, YR = TY + 8;
    const int TNP = (TY * nx + EPI - 1) / EPI * EPI;
    T *const Xb = reinterpret_cast<T *>(smem_d);
    T *const Yb = Xb + 2 * 4 * TY * W;
    T *const Pb = Yb + 2 * 4 * YR * nx;
    // task: state, and tile + ntile_y * slab (z slab slab of nzs)
    const Task tk = tasks[blockIdx.x];
    const int s = tk.state, tile = tk.tile % ntile_y, slab = tk.tile / ntile_y;
    const int zb = nz * slab / nzs, ze = nz * (slab + 1) / nzs;
    const int it = (qn[4 * s] + 3) / 2 - 1;
    const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
    const T kin = it == 0 ? kin0 : kin1;
    const int tid = threadIdx.x, lane = tid % 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int half = lane / 32;                // components 2 half, 2 half + 1
    const int pt = wave * 32 + lane % 32;      // point of the tile plane
    const int tyl = pt / nx, tx = pt % nx;
    const int y0 = tile * TY, y = y0 + tyl;
    const bool inplane = tyl < TY;
    const bool active = inplane && (y < ny);
    const bool edge = y0 < 4 || y0 + TY + 4 > ny;
    const int rows_valid = min(TY, ny - y0);
    T sgx[2], sgy[2], sgz[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int c = 2 * half + a;
        sgx[a] = T(comp_sign_x(c, S));
        sgy[a] = T(comp_sign_y(c, TS));
        sgz[a] = T(comp_sign_z(c, S, P));
    }
    const long cN = (long)(2 * half) * N;      // offset of this thread's first component
    const T *ps = psi + (long)s * 4 * N;
    const T *pv = prev + (long)s * 4 * N + cN;
    T *po = psi_out + (long)s * 4 * N + cN;
    const long col = (long)(active ? y : 0) * nx + tx;
    // the register queue of the z stencil: planes zb-4 .. zb+4 of both components
    T q[2][9];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int zz = zb + k - 4;
            const int zi = zz < 0 ? -1 - zz : (zz < nz ? zz : nz - 1);
            const T val = ps[cN + a * N + zi * plane + col];
            q[a][k] = !active || zz >= nz ? T(0) : (zz < 0 ? sgz[a] * val : val);
        }
    // Direct-to-LDS loads of plane zz into buffer b.  Wave w < 8 moves one
    // halo block (component w/2, 4 rows above (w even) or below the tile:
    // nhb <= 2 loads) and the potentials j = w, w+NW, ... (nfc <= 2 loads
    // each).  The lane offsets inside a load are fixed, so they are set up
    // once here.
    const int nhb = (4 * nx + EPI - 1) / EPI, nfc = TNP / EPI;
    int hrow[2], hx[2];
    bool hval[2], fval[2];
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        const int e = ch * EPI + lane * VPL;
        hval[ch] = ch < nhb && e < 4 * nx;
        hrow[ch] = e / nx;
        hx[ch] = e % nx;
        fval[ch] = ch < nfc && e < rows_valid * nx;
    }
    const int hc = wave >> 1, hrr0 = (wave & 1) ? TY + 4 : 0;
    const T *psw = ps + (long)hc * N;
    auto issue = [&](int zz, int b) {
        if (wave < 8) {
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
                T *dst = Yb + ((b * 4 + hc) * YR + hrr0) * nx + ch * EPI;
                if (hval[ch]) {
                    const int yy = y0 - 4 + hrr0 + hrow[ch];
                    const int yi = yy < 0 ? -1 - yy : (yy < ny ? yy : ny - 1);
                    dma16(psw + (long)zz * plane + (long)yi * nx + hx[ch], dst);
                }
            }
        }
        for (int j = wave; j < NF; j += NW) {
            const T *src = F + (long)pot_slot<WA>(j, it) * N + (long)zz * plane + (long)y0 * nx +
                           lane * VPL;
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
                T *dst = Pb + (b * NF + j) * TNP + ch * EPI;
                if (fval[ch]) dma16(src + ch * EPI, dst);
            }
        }
    };
    issue(zb, 0);
    double eacc = 0, dacc = 0, kacc = 0;
    T pst[2] = {0, 0};
    long pst_p = -1;
    for (int z = zb; z < ze; ++z) {
        const int b = (z - zb) & 1;
        T *const Xc = Xb + b * 4 * TY * W;
        T *const Yc = Yb + b * 4 * YR * nx;
        const T *const Pc = Pb + b * NF * TNP;
        // ---- the tile's own rows of plane z from the queue ----
        if (inplane) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int c = 2 * half + a;
                const T v = active ? q[a][4] : T(0);
                T *xr = Xc + (c * TY + tyl) * W;
                xr[tx + G] = v;
                if (tx < G) xr[G - 1 - tx] = sgx[a] * v;
                if (tx >= nx - G) xr[tx + 2 * G] = T(0);
                Yc[(c * YR + tyl + 4) * nx + tx] = v;
            }
        }
        // The direct-to-LDS loads of plane z (issued during plane z-1) must
        // have landed before the barrier publishes buffer b to the other
        // waves: a workgroup barrier does not wait on vmcnt by itself.
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();
        // ---- stores of plane z-1, loads of plane z+1 and of the queue ----
        if (!KIN && pst_p >= 0) {
            po[pst_p] = pst[0];
            po[N + pst_p] = pst[1];
        }
        if (z + 1 < ze) issue(z + 1, b ^ 1);
        const long p = (long)z * plane + col;
        const int zn = z + 5 < nz ? z + 5 : nz - 1;
        T nxt[2], uo[2] = {T(0), T(0)};
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            nxt[a] = ps[cN + a * N + zn * plane + col];
            if constexpr (!KIN) uo[a] = pv[a * N + p];
        }
        __builtin_amdgcn_sched_barrier(0);
        // Contributions of this thread's two components (a = 0, 1: c = 2 half
        // + a) to h psi (ActionOfBNew, ActionOfU: MeanFields.f90:792-851;
        // ActionOfW :1029-1065; ActionOfS, ActionOfA :1107-1147).  With
        // sg = +1 for components (0, 1) and -1 for (2, 3) the couplings have
        // the same form in both halves:
        //   own0  = hown_0 + sg ez_1 + sg s2 v_0 + aa_1
        //   own1  = hown_1 - sg ez_0 + sg s2 v_1 - aa_0
        //   send0 = ex_1 + sg ey_0 + s0 v_0 - sg s1 v_1   (partner's first)
        //   send1 = -ex_0 + sg ey_1 + sg s1 v_0 + s0 v_1  (partner's second)
        // Lanes outside the tile compute on zeros.
        const T sg = half ? T(-1) : T(1);
        T own0 = 0, own1 = 0, send0 = 0, send1 = 0;
        if constexpr (KIN) {
            const int rowi = inplane ? tyl : 0;
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int c = 2 * half + a;
                const T *row = Xc + (c * TY + rowi) * W + tx + G;
                T w[9], dxc, dyc, dzc, lx, ly, lz;
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = row[k];
                fd_pair(m, w, dxc, lx);
                const T *colp = Yc + (c * YR + rowi + 4) * nx + tx;
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = colp[k * nx];
                if (edge) {
#pragma unroll
                    for (int k = -4; k <= 4; ++k) {
                        const int yy = y + k;
                        w[k + 4] *= yy < 0 ? sgy[a] : (yy < ny ? T(1) : T(0));
                    }
                }
                fd_pair(m, w, dyc, ly);
                fd_pair(m, q[a], dzc, lz);
                kacc += (double)q[a][4] * (double)(lx + ly + lz);
            }
        } else {
            const T *pp = Pc + (inplane ? pt : 0);
            const T fu = pp[0], fb = pp[TNP] + kin;
            const T nb0 = pp[2 * TNP], nb1 = pp[3 * TNP], nb2 = pp[4 * TNP];
            const int rowi = inplane ? tyl : 0;
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int c = 2 * half + a;
                const T *row = Xc + (c * TY + rowi) * W + tx + G;
                T w[9], dxc, dyc, dzc, lx, ly, lz;
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = row[k];
                fd_pair(m, w, dxc, lx);
                const T *colp = Yc + (c * YR + rowi + 4) * nx + tx;
#pragma unroll
                for (int k = -4; k <= 4; ++k) w[k + 4] = colp[k * nx];
                if (edge) {
#pragma unroll
                    for (int k = -4; k <= 4; ++k) {
                        const int yy = y + k;
                        w[k + 4] *= yy < 0 ? sgy[a] : (yy < ny ? T(1) : T(0));
                    }
                }
                fd_pair(m, w, dyc, ly);
                fd_pair(m, q[a], dzc, lz);
                const T lc = lx + ly + lz, vc = q[a][4];
                kacc += (double)vc * (double)lc;
                const T hown = -fb * lc - (nb0 * dxc + nb1 * dyc + nb2 * dzc) + fu * vc;
                T ex, ey, ez;
                if constexpr (WA) {
                    const T w01 = pp[5 * TNP], w02 = pp[6 * TNP], w12 = pp[7 * TNP];
                    ex = w01 * dyc + w02 * dzc + T(0.5) * pp[FD * TNP] * vc;
                    ey = -w01 * dxc + w12 * dzc + T(0.5) * pp[(FD + 1) * TNP] * vc;
                    ez = -w02 * dxc - w12 * dyc + T(0.5) * pp[(FD + 2) * TNP] * vc;
                } else {
                    ex = pp[5 * TNP] * dxc + pp[6 * TNP] * dyc + pp[7 * TNP] * dzc +
                         T(0.5) * pp[FD * TNP] * vc;
                    ey = pp[8 * TNP] * dxc + pp[9 * TNP] * dyc + pp[10 * TNP] * dzc +
                         T(0.5) * pp[(FD + 1) * TNP] * vc;
                    ez = pp[11 * TNP] * dxc + pp[12 * TNP] * dyc + pp[13 * TNP] * dzc +
                         T(0.5) * pp[(FD + 2) * TNP] * vc;
                }
                if (a == 0) {
                    own0 += hown;
                    own1 -= sg * ez;
                    send0 += sg * ey;
                    send1 -= ex;
                } else {
                    own1 += hown;
                    own0 += sg * ez;
                    send0 += ex;
                    send1 += sg * ey;
                }
                if (!trc) {
                    const T s0 = pp[(FD + 3) * TNP], s1 = pp[(FD + 4) * TNP],
                            s2 = pp[(FD + 5) * TNP];
                    const T aac = pp[(FD + 6) * TNP] * dxc + pp[(FD + 7) * TNP] * dyc +
                                  pp[(FD + 8) * TNP] * dzc + T(0.5) * pp[(FD + 9) * TNP] * vc;
                    if (a == 0) {
                        own0 += sg * s2 * vc;
                        send0 += s0 * vc;
                        send1 += sg * s1 * vc;
                        own1 -= aac;
                    } else {
                        own1 += sg * s2 * vc;
                        send0 -= sg * s1 * vc;
                        send1 += s0 * vc;
                        own0 += aac;
                    }
                }
            }
        }
        // the partner's partial sums of this thread's components
        const T h0 = own0 + __shfl_xor(send0, 32, 64);
        const T h1 = own1 + __shfl_xor(send1, 32, 64);
        if (!KIN && active) {
            const T hh[2] = {h0, h1};
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const T vc = q[a][4];
                eacc += (double)vc * (double)hh[a];
                dacc += (double)hh[a] * (double)hh[a];
                // heavy-ball memory u = psi - psi_prev (ImaginaryTime.f90:497-499)
                const T u = vc - uo[a];
                pst[a] = vc + (-(dth * hh[a]) + mu * u);
            }
            pst_p = p;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int k = 0; k < 8; ++k) q[a][k] = q[a][k + 1];
            q[a][8] = (active && z + 5 < nz) ? nxt[a] : T(0);
        }
    }
    if (!KIN && pst_p >= 0) {
        po[pst_p] = pst[0];
        po[N + pst_p] = pst[1];
    }
    if (eparts) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            eacc += __shfl_xor(eacc, o, 64);
            dacc += __shfl_xor(dacc, o, 64);
            kacc += __shfl_xor(kacc, o, 64);
        }
        if (lane == 0) { red[3 * wave] = eacc; red[3 * wave + 1] = dacc; red[3 * wave + 2] = kacc; }
        __syncthreads();
        if (tid == 0) {
            double sa = 0, sb = 0, sc = 0;
            for (int k = 0; k < NW; ++k) { sa += red[3 * k]; sb += red[3 * k + 1]; sc += red[3 * k + 2]; }
            double *ep = eparts + ((long)s * ntile_y * nzs + tk.tile) * 3;
            ep[0] = sa;
            ep[1] = sb;
            ep[2] = sc;
        }
    }
}

template <typename T>
void launch_hpsi_fd(const MeshK<T> &m, const T *psi, T *psi_out, T *upd, T *hout,
                    const T *fields, const int *qn, const Task *tasks, int ntasks,
                    int TY, int ntile_y, int trc, T kin0, T kin1, T dth, T mu,
                    int mode, double *eparts, hipStream_t st, const T *vsrc, bool wa) {
    // (whole 16-byte lane loads inside a row; the halo rows of one component
    // in at most two 1 KiB loads; the LDS within the device's per-workgroup
    // limit less the kernel's static reduction array)
    const size_t shm = sizeof(T) * 4 * (size_t)(TY + 8) * (m.nx + 8);
#define MOCCA_HPSI_LAUNCH(HM, TR, ST)                                                         \
    if (wa)                                                                                   \
        hipLaunchKernelGGL((k_hpsi_fd<T, HM, TR, ST, true>), dim3(ntasks), dim3(kBlock), shm, \
                           st, m, psi, psi_out, upd, hout, fields, qn, tasks, TY, ntile_y,    \
                           kin0, kin1, dth, mu, mode, eparts, vsrc);                          \
    else                                                                                      \
        hipLaunchKernelGGL((k_hpsi_fd<T, HM, TR, ST, false>), dim3(ntasks), dim3(kBlock),     \
                           shm, st, m, psi, psi_out, upd, hout, fields, qn, tasks, TY,        \
                           ntile_y, kin0, kin1, dth, mu, mode, eparts, vsrc)
#define MOCCA_HPSI_TRC(HM, ST)                                                                \
    if (trc) { MOCCA_HPSI_LAUNCH(HM, true, ST); } else { MOCCA_HPSI_LAUNCH(HM, false, ST); }
#define MOCCA_HPSI_KIN(HM, TR)                                                                \
    hipLaunchKernelGGL((k_hpsi_fd<T, HM, TR, false, false, true>), dim3(ntasks), dim3(kBlock), \
                       shm, st, m, psi, psi_out, upd, hout, fields, qn, tasks, TY, ntile_y,   \
                       kin0, kin1, dth, mu, 0, eparts, vsrc)
    if (mode & kKinetic) {
        if (m.nx <= 32) {
            if (trc) MOCCA_HPSI_KIN(4, true); else MOCCA_HPSI_KIN(4, false);
        } else {
            if (trc) MOCCA_HPSI_KIN(8, true); else MOCCA_HPSI_KIN(8, false);
        }
    } else if (vsrc) {
        if (m.nx <= 32) { MOCCA_HPSI_TRC(4, true); } else { MOCCA_HPSI_TRC(8, true); }
    } else {
        if (m.nx <= 32) { MOCCA_HPSI_TRC(4, false); } else { MOCCA_HPSI_TRC(8, false); }
    }
#undef MOCCA_HPSI_KIN
#undef MOCCA_HPSI_TRC
#undef MOCCA_HPSI_LAUNCH
}

// The pipelined FP64 step kernel is usable when a row splits into whole
// 16-byte lane loads, the halo rows of one component fit two 1 KiB loads, and
// its LDS fits the device's per-workgroup limit (less its static array).
static size_t hpsi_dma_lds_limit() {
    static const size_t lim = [] {
        int dev = 0, l = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&l, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess)
            return (size_t)0;
        const size_t stat = 3 * kDmaThreads / 64 * sizeof(double);
        return (size_t)l > stat ? (size_t)l - stat : (size_t)0;
    }();
    return lim;
}

bool hpsi_dma_ok(const MeshK<double> &m, int TY, int trc) {
    if ((m.nx * sizeof(double)) % 16 != 0 || 4 * m.nx * sizeof(double) > 2048) return false;
    // (the general-W variant needs the most LDS)
    const size_t dshm = trc ? hpsi_dma_lds<double, true, false>(m.nx, TY)
                            : hpsi_dma_lds<double, false, false>(m.nx, TY);
    return dshm <= hpsi_dma_lds_limit();
}

void launch_hpsi_fd_dma(const MeshK<double> &m, const double *psi, double *psi_out,
                        const double *prev, const double *fields, const int *qn,
                        const Task *tasks, int ntasks, int TY, int ntile_y, int nzs, int trc,
                        double kin0, double kin1, double dth, double mu, double *eparts,
                        hipStream_t st, bool w_antisym) {
#define MOCCA_HPSI_DMA(TR, WA)                                                              \
    {                                                                                       \
        const size_t dshm = hpsi_dma_lds<double, TR, WA>(m.nx, TY);                         \
        static const bool lds_set = require_lds((const void *)k_hpsi_fd_dma<double, TR, WA>, \
                                                hpsi_dma_lds_limit(), "k_hpsi_fd_dma");     \
        if (!lds_set)                                                                       \
            require_lds((const void *)k_hpsi_fd_dma<double, TR, WA>, hpsi_dma_lds_limit(),  \
                        "k_hpsi_fd_dma");                                                   \
        hipLaunchKernelGGL((k_hpsi_fd_dma<double, TR, WA>), dim3(ntasks), dim3(kDmaThreads), \
                           dshm, st, m, psi, psi_out, prev, fields, qn, tasks, TY, ntile_y, \
                           nzs, kin0, kin1, dth, mu, eparts);                               \
    }
    if (trc) {
        if (w_antisym) MOCCA_HPSI_DMA(true, true) else MOCCA_HPSI_DMA(true, false)
    } else {
        if (w_antisym) MOCCA_HPSI_DMA(false, true) else MOCCA_HPSI_DMA(false, false)
    }
#undef MOCCA_HPSI_DMA
}

void launch_hpsi_kinetic_dma(const MeshK<double> &m, const double *psi, const int *qn,
                             const Task *tasks, int ntasks, int TY, int ntile_y, int nzs,
                             int trc, double *eparts, hipStream_t st) {
    // the LDS of the state buffers only (no potentials)
    const size_t dshm = sizeof(double) * ((size_t)2 * 4 * TY * (m.nx + 8) +
                                          (size_t)2 * 4 * (TY + 8) * m.nx);
#define MOCCA_KIN_DMA(TR)                                                                    \
    {                                                                                        \
        static const bool lds_set = require_lds(                                             \
            (const void *)k_hpsi_fd_dma<double, TR, false, true>, hpsi_dma_lds_limit(),      \
            "k_hpsi_fd_dma");                                                                \
        if (!lds_set)                                                                        \
            require_lds((const void *)k_hpsi_fd_dma<double, TR, false, true>,                \
                        hpsi_dma_lds_limit(), "k_hpsi_fd_dma");                              \
        hipLaunchKernelGGL((k_hpsi_fd_dma<double, TR, false, true>), dim3(ntasks),           \
                           dim3(kDmaThreads), dshm, st, m, psi, nullptr, nullptr, nullptr, qn, \
                           tasks, TY, ntile_y, nzs, 0.0, 0.0, 0.0, 0.0, eparts);             \
    }
    if (trc) MOCCA_KIN_DMA(true) else MOCCA_KIN_DMA(false)
#undef MOCCA_KIN_DMA
}

template void launch_hpsi_fd<double>(const MeshK<double> &, const double *, double *,
                                     double *, double *, const double *, const int *,
                                     const Task *, int, int, int, int, double, double,
                                     double, double, int, double *, hipStream_t,
                                     const double *, bool);
template void launch_hpsi_fd<float>(const MeshK<float> &, const float *, float *, float *,
                                    float *, const float *, const int *, const Task *, int,
                                    int, int, int, float, float, float, float, int,
                                    double *, hipStream_t, const float *, bool);


// ---------------------------------------------------------------------------
// Densities (ComputeDensity, Densities.f90:464-534) with the same z-march:
// a task is (y-tile, z-slab of TZ planes, chunk of states of one isospin).
// For every state of the chunk the slab is walked with a 7-plane register
// queue (first derivatives only) and the occupation-weighted contributions
// are summed in registers; the chunk's partial fields go to
// part[chunk][f][N], f = rho, tau, nabla.J [, j_x, j_y, j_z, s_x, s_y, s_z].
namespace {

template <typename T>
__device__ __forceinline__ T imd(const T a[4], const T b[4]) {
    return a[0] * b[1] - a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
}

template <typename T>
__device__ __forceinline__ void sig(const T a[4], int dir, T o[4]) {
    if (dir == 0) { o[0] = a[2]; o[1] = a[3]; o[2] = a[0]; o[3] = a[1]; }
    else if (dir == 1) { o[0] = a[3]; o[1] = -a[2]; o[2] = -a[1]; o[3] = a[0]; }
    else { o[0] = a[0]; o[1] = a[1]; o[2] = -a[2]; o[3] = -a[3]; }
}

template <typename T>
__device__ __forceinline__ T fd1_7(const MeshK<T> &m, const T v[7]) {
    return (T(45) * (v[4] - v[2]) - T(9) * (v[5] - v[1]) + (v[6] - v[0])) * m.inv60dx;
}

}  // namespace

template <typename T, int NF, int TZ, int HMAX>
__global__ __launch_bounds__(kBlock) void k_density_fd(
    MeshK<T> m, const T *__restrict__ psi, const int *__restrict__ qn,
    const double *__restrict__ occ, const int *__restrict__ order,
    const Chunk *__restrict__ dchunks, const DTask *__restrict__ tasks, int TY,
    T *__restrict__ part) {
    constexpr int G = 3;  // ghost width; HMAX = halo loads per thread (24 nx / 256)
    extern __shared__ double smem_d[];
    T *sp = reinterpret_cast<T *>(smem_d);  // [4][TY+6][nx+6]
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const long N = m.N, plane = (long)nx * ny;
    const int W = nx + 2 * G;
    const DTask tk = tasks[blockIdx.x];
    const Chunk ch = dchunks[tk.chunk];
    const int tid = threadIdx.x, tx = tid % nx, tyl = tid / nx;
    const int y0 = tk.tile * TY, y = y0 + tyl;
    const bool active = (tyl < TY) && (y < ny);
    const int rows = TY + 6;
    const int z0 = tk.slab * TZ;
    const long col = (long)(active ? y : 0) * nx + tx;
    T acc[TZ][NF];
#pragma unroll
    for (int a = 0; a < TZ; ++a)
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[a][f] = 0;
    const int nhalo = 4 * 6 * nx;
    // halo element r of this thread (fixed for the task): source offset in
    // the state less the plane, LDS offset, and component | yy < 0 |
    // yy >= ny | left ghost band | right ghost band | x << 8
    int hg[HMAX], hl[HMAX], hf[HMAX];
#pragma unroll
    for (int r = 0; r < HMAX; ++r) {
        const int idx = tid + r * kBlock;
        const int c = idx / (6 * nx), rem = idx % (6 * nx);
        const int hr = rem / nx, xx = rem % nx;
        const int rr = hr < 3 ? hr : TY + hr;
        const int yy = y0 - 3 + rr;
        const int yi = yy < 0 ? -1 - yy : (yy < ny ? yy : ny - 1);
        const bool in = idx < nhalo;
        hg[r] = in ? c * (int)N + yi * nx + xx : 0;
        hl[r] = in ? (c * rows + rr) * W + xx + G : -1;
        hf[r] = (in ? c : 0) | (yy < 0 ? 4 : 0) | (yy >= ny ? 8 : 0) | (xx < G ? 16 : 0) |
                (xx >= nx - G ? 32 : 0) | (xx << 8);
    }
    for (int q = 0; q < ch.count; ++q) {
        const int s = order[ch.start + q];
        if (occ[s] == 0.0) continue;
        const T o = (T)occ[s];
        const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
        T sgx[4], sgz[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sgx[c] = T(comp_sign_x(c, S));
            sgz[c] = T(comp_sign_z(c, S, P));
        }
        const T *ps = psi + (long)s * 4 * N;
        T qz[4][7];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int zz = z0 - 3 + k;
                const int zi = zz < 0 ? -1 - zz : (zz < nz ? zz : nz - 1);
                const T val = ps[c * N + zi * plane + col];
                qz[c][k] = !active || zz >= nz ? T(0) : (zz < 0 ? sgz[c] * val : val);
            }
#pragma unroll
        for (int zl = 0; zl < TZ; ++zl) {
            const int z = z0 + zl;
            if (z >= nz) break;
            T hv[HMAX];
#pragma unroll
            for (int r = 0; r < HMAX; ++r) hv[r] = ps[(long)z * plane + hg[r]];
            const int zn = z + 4 < nz ? z + 4 : nz - 1;
            T nxt[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) nxt[c] = ps[c * N + zn * plane + col];
            if (tyl < TY) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    T *row = sp + (c * rows + tyl + 3) * W;
                    const T v = active ? qz[c][3] : T(0);
                    row[tx + G] = v;
                    if (tx < G) row[G - 1 - tx] = sgx[c] * v;
                    if (tx >= nx - G) row[tx + 2 * G] = T(0);
                }
            }
#pragma unroll
            for (int r = 0; r < HMAX; ++r) {
                if (hl[r] >= 0) {
                    const int f = hf[r], c = f & 3, xx = f >> 8;
                    const T sy = T(comp_sign_y(c, TS)), sx = T(comp_sign_x(c, S));
                    const T v = (f & 4) ? sy * hv[r] : ((f & 8) ? T(0) : hv[r]);
                    T *e = sp + hl[r];
                    e[0] = v;
                    if (f & 16) e[-1 - 2 * xx] = sx * v;
                    if (f & 32) e[G] = T(0);
                }
            }
            __syncthreads();
            if (active) {
                T v[4], d[3][4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T *row = sp + (c * rows + tyl + 3) * W + tx + G;
                    T w[7];
#pragma unroll
                    for (int k = -3; k <= 3; ++k) w[k + 3] = row[k];
                    d[0][c] = fd1_7(m, w);
#pragma unroll
                    for (int k = -3; k <= 3; ++k) w[k + 3] = row[k * W];
                    d[1][c] = fd1_7(m, w);
                    d[2][c] = fd1_7(m, qz[c]);
                    v[c] = qz[c][3];
                }
                T t[4];
                acc[zl][0] += o * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
                T ta = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    ta += d[a][0] * d[a][0] + d[a][1] * d[a][1] + d[a][2] * d[a][2] + d[a][3] * d[a][3];
                acc[zl][1] += o * ta;
                // nabla.J = sum eps_ijk Im(d_i psi^+ sigma_k d_j psi) (Spwf.f90:2206-2243)
                T nj = 0;
                sig(d[1], 2, t); nj += imd(d[0], t);
                sig(d[2], 1, t); nj -= imd(d[0], t);
                sig(d[1], 0, t); nj -= imd(d[2], t);
                sig(d[0], 1, t); nj += imd(d[2], t);
                sig(d[0], 2, t); nj -= imd(d[1], t);
                sig(d[2], 0, t); nj += imd(d[1], t);
                acc[zl][2] += o * nj;
                if (NF == 9) {
                    // j (Spwf.f90:2034) and s (Spwf.f90:2071)
#pragma unroll
                    for (int a = 0; a < 3; ++a) acc[zl][3 + a] += o * imd(v, d[a]);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        sig(v, a, t);
                        acc[zl][6 + a] += o * (v[0] * t[0] + v[1] * t[1] + v[2] * t[2] + v[3] * t[3]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int k = 0; k < 6; ++k) qz[c][k] = qz[c][k + 1];
                qz[c][6] = (active && z + 4 < nz) ? nxt[c] : T(0);
            }
        }
    }
    if (active) {
#pragma unroll
        for (int zl = 0; zl < TZ; ++zl) {
            const int z = z0 + zl;
            if (z >= nz) break;
#pragma unroll
            for (int f = 0; f < NF; ++f)
                part[(long)(tk.chunk * NF + f) * N + z * plane + col] = acc[zl][f];
        }
    }
}

// The same densities with one barrier per plane: while plane z is computed,
// the 3 + 3 halo rows of the next plane (this state's z+1, or the first
// plane of the chunk's next occupied state) move to the other LDS buffer by
// direct-to-LDS loads.  Wave w moves halo blocks w and w+4 (block k:
// component k/2, the rows above (k even) or below the tile).  The tile's own
// rows come from the register queue.  Requires nx * sizeof(T) to be a
// multiple of 16 and three rows to fit two 1 KiB wave loads.
template <typename T, int NF, int TZ>
__global__ __launch_bounds__(kBlock) void k_density_fd_dma(
    MeshK<T> m, const T *__restrict__ psi, const int *__restrict__ qn,
    const double *__restrict__ occ, const int *__restrict__ order,
    const Chunk *__restrict__ dchunks, const DTask *__restrict__ tasks, int TY,
    T *__restrict__ part) {
    constexpr int G = 3;
    constexpr int VPL = 16 / (int)sizeof(T);   // values per lane of one 16-byte load
    constexpr int EPI = 64 * VPL;              // values per wave load (1 KiB)
    static_assert(kBlock == 256, "halo blocks are spread over four waves");
    extern __shared__ double smem_d[];
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const long N = m.N, plane = (long)nx * ny;
    const int W = nx + 2 * G, YR = TY + 2 * G;
    T *const Yb = reinterpret_cast<T *>(smem_d);  // [2][4][TY+6][nx]
    T *const Xb = Yb + 2 * 4 * YR * nx;           // [2][4][TY][nx+6]
    const DTask tk = tasks[blockIdx.x];
    const Chunk ch = dchunks[tk.chunk];
    const int tid = threadIdx.x, lane = tid % 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int tx = tid % nx, tyl = tid / nx;
    const int y0 = tk.tile * TY, y = y0 + tyl;
    const bool inplane = tyl < TY;
    const bool active = inplane && (y < ny);
    const bool edge = y0 < G || y0 + TY + G > ny;
    const int zb = tk.slab * TZ, ze = min(zb + TZ, nz);
    const long col = (long)(active ? y : 0) * nx + tx;
    const int nhb = (G * nx + EPI - 1) / EPI;
    int hrow[2], hx[2];
    bool hval[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int e = k * EPI + lane * VPL;
        hval[k] = k < nhb && e < G * nx;
        hrow[k] = e / nx;
        hx[k] = e % nx;
    }
    auto issue = [&](const T *ps, int zz, int b) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int blk = wave + 4 * j;
            const int c = blk >> 1, rr0 = (blk & 1) ? TY + G : 0;
            T *const dst = Yb + ((b * 4 + c) * YR + rr0) * nx;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (hval[k]) {
                    const int yy = y0 - G + rr0 + hrow[k];
                    const int yi = yy < 0 ? -1 - yy : (yy < ny ? yy : ny - 1);
                    dma16(ps + (long)c * N + (long)zz * plane + (long)yi * nx + hx[k],
                          dst + k * EPI);
                }
            }
        }
    };
    auto next_occupied = [&](int q) {
        while (q < ch.count && occ[order[ch.start + q]] == 0.0) ++q;
        return q;
    };
    T acc[TZ][NF];
#pragma unroll
    for (int a = 0; a < TZ; ++a)
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[a][f] = 0;
    int q = next_occupied(0);
    int b = 0;
    if (q < ch.count) issue(psi + (long)order[ch.start + q] * 4 * N, zb, 0);
    while (q < ch.count) {
        const int s = order[ch.start + q];
        const int qnext = next_occupied(q + 1);
        const T *const psn =
            qnext < ch.count ? psi + (long)order[ch.start + qnext] * 4 * N : nullptr;
        const T o = (T)occ[s];
        const int P = qn[4 * s + 1], S = qn[4 * s + 2], TS = qn[4 * s + 3];
        T sgx[4], sgy[4], sgz[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sgx[c] = T(comp_sign_x(c, S));
            sgy[c] = T(comp_sign_y(c, TS));
            sgz[c] = T(comp_sign_z(c, S, P));
        }
        const T *ps = psi + (long)s * 4 * N;
        T qz[4][7];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int zz = zb - G + k;
                const int zi = zz < 0 ? -1 - zz : (zz < nz ? zz : nz - 1);
                const T val = ps[c * N + zi * plane + col];
                qz[c][k] = !active || zz >= nz ? T(0) : (zz < 0 ? sgz[c] * val : val);
            }
#pragma unroll
        for (int zl = 0; zl < TZ; ++zl) {
            const int z = zb + zl;
            if (z >= ze) break;
            T *const Xc = Xb + b * 4 * TY * W;
            T *const Yc = Yb + b * 4 * YR * nx;
            if (inplane) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T v = active ? qz[c][3] : T(0);
                    T *xr = Xc + (c * TY + tyl) * W;
                    xr[tx + G] = v;
                    if (tx < G) xr[G - 1 - tx] = sgx[c] * v;
                    if (tx >= nx - G) xr[tx + 2 * G] = T(0);
                    Yc[(c * YR + tyl + G) * nx + tx] = v;
                }
            }
            // this wave's direct-to-LDS loads of buffer b are complete before
            // the barrier makes the buffer visible to the other waves
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
            if (z + 1 < ze) issue(ps, z + 1, b ^ 1);
            else if (psn) issue(psn, zb, b ^ 1);
            const int zn = z + 4 < nz ? z + 4 : nz - 1;
            T nxt[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) nxt[c] = ps[c * N + zn * plane + col];
            if (active) {
                T v[4], d[3][4];
                const int rowi = tyl;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T *row = Xc + (c * TY + rowi) * W + tx + G;
                    T w[7];
#pragma unroll
                    for (int k = -3; k <= 3; ++k) w[k + 3] = row[k];
                    d[0][c] = fd1_7(m, w);
                    const T *colp = Yc + (c * YR + rowi + G) * nx + tx;
#pragma unroll
                    for (int k = -3; k <= 3; ++k) w[k + 3] = colp[k * nx];
                    if (edge) {
#pragma unroll
                        for (int k = -3; k <= 3; ++k) {
                            const int yy = y + k;
                            w[k + 3] *= yy < 0 ? sgy[c] : (yy < ny ? T(1) : T(0));
                        }
                    }
                    d[1][c] = fd1_7(m, w);
                    d[2][c] = fd1_7(m, qz[c]);
                    v[c] = qz[c][3];
                }
                T t[4];
                acc[zl][0] += o * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
                T ta = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    ta += d[a][0] * d[a][0] + d[a][1] * d[a][1] + d[a][2] * d[a][2] +
                          d[a][3] * d[a][3];
                acc[zl][1] += o * ta;
                // nabla.J = sum eps_ijk Im(d_i psi^+ sigma_k d_j psi) (Spwf.f90:2206-2243)
                T nj = 0;
                sig(d[1], 2, t); nj += imd(d[0], t);
                sig(d[2], 1, t); nj -= imd(d[0], t);
                sig(d[1], 0, t); nj -= imd(d[2], t);
                sig(d[0], 1, t); nj += imd(d[2], t);
                sig(d[0], 2, t); nj -= imd(d[1], t);
                sig(d[2], 0, t); nj += imd(d[1], t);
                acc[zl][2] += o * nj;
                if (NF == 9) {
                    // j (Spwf.f90:2034) and s (Spwf.f90:2071)
#pragma unroll
                    for (int a = 0; a < 3; ++a) acc[zl][3 + a] += o * imd(v, d[a]);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        sig(v, a, t);
                        acc[zl][6 + a] +=
                            o * (v[0] * t[0] + v[1] * t[1] + v[2] * t[2] + v[3] * t[3]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int k = 0; k < 6; ++k) qz[c][k] = qz[c][k + 1];
                qz[c][6] = (active && z + 4 < nz) ? nxt[c] : T(0);
            }
            b ^= 1;
        }
        q = qnext;
    }
    if (active) {
#pragma unroll
        for (int zl = 0; zl < TZ; ++zl) {
            const int z = zb + zl;
            if (z >= ze) break;
#pragma unroll
            for (int f = 0; f < NF; ++f)
                part[(long)(tk.chunk * NF + f) * N + z * plane + col] = acc[zl][f];
        }
    }
}

template <typename T>
static bool density_dma_ok(const MeshK<T> &m, int TY) {
    return (m.nx * sizeof(T)) % 16 == 0 && 3 * m.nx * sizeof(T) <= 2048 &&
           TY * m.nx <= kBlock && m.ny > 3 && m.nz > 3;
}

template <typename T>
void launch_density_fd(const MeshK<T> &m, const T *psi, const int *qn, const double *occ,
                       const int *order, const Chunk *dchunks, const DTask *tasks,
                       int ntasks, int TY, int trc, T *part, hipStream_t st) {
    // (the time-odd variant keeps the two-barrier kernel: with its nine
    // accumulators per plane the pipelined one holds a single wave per SIMD)
    if (trc && density_dma_ok(m, TY)) {
        const size_t dshm =
            sizeof(T) * ((size_t)2 * 4 * (TY + 6) * m.nx + (size_t)2 * 4 * TY * (m.nx + 6));
        hipLaunchKernelGGL((k_density_fd_dma<T, 3, kDensSlab>), dim3(ntasks), dim3(kBlock),
                           dshm, st, m, psi, qn, occ, order, dchunks, tasks, TY, part);
        return;
    }
    const size_t shm = sizeof(T) * 4 * (size_t)(TY + 6) * (m.nx + 6);
#define MOCCA_DENS_LAUNCH(NF, TZ, HM)                                                      \
    hipLaunchKernelGGL((k_density_fd<T, NF, TZ, HM>), dim3(ntasks), dim3(kBlock), shm, st, m, \
                       psi, qn, occ, order, dchunks, tasks, TY, part)
    constexpr int TZ9 = sizeof(T) == 4 ? kDensSlabTodd32 : kDensSlab;
    if (m.nx <= 32) {
        if (trc) MOCCA_DENS_LAUNCH(3, kDensSlab, 3); else MOCCA_DENS_LAUNCH(9, TZ9, 3);
    } else {
        if (trc) MOCCA_DENS_LAUNCH(3, kDensSlab, 6); else MOCCA_DENS_LAUNCH(9, TZ9, 6);
    }
#undef MOCCA_DENS_LAUNCH
}

template void launch_density_fd<double>(const MeshK<double> &, const double *, const int *,
                                        const double *, const int *, const Chunk *,
                                        const DTask *, int, int, int, double *, hipStream_t);
template void launch_density_fd<float>(const MeshK<float> &, const float *, const int *,
                                       const double *, const int *, const Chunk *,
                                       const DTask *, int, int, int, float *, hipStream_t);

}  // namespace mocca
