// Gram-Schmidt with no allocation (mocca_gram_schmidt_noalloc): the same
// Cholesky-QR orthonormalisation as mocca_gram_schmidt (Gramschmidt,
// SpwfStorage.f90:372-480), on FP64 states that the caller keeps in device
// memory, with its workspace in one static device array per GPU instead of
// buffers of a solve handle.  It is meant for callers that cannot allocate:
// code under HIP graph capture, tight serving loops.
//
// The workspace `scratch` holds the partial and reduced Gram matrices, the
// Cholesky factors and their inverses, the global-memory Cholesky's work
// matrix, the block tables, the energy order and a status word.  Each GPU
// holds its own copy for the life of the process; its device address is
// looked up once per device (hipGetSymbolAddress) and only that address is
// passed to kernels.  Every call on a GPU uses the same array, so calls that
// overlap on one GPU (different streams, different threads) race; each call
// clears what must start at zero, so a call never sees the previous call's
// values.
#include "mocca_internal.h"

#include <algorithm>
#include <mutex>

namespace mocca {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;   // per GPU
constexpr int kNaMaxStates = 512;
constexpr int kNaMaxBlocks = 8;                      // (isospin, parity, signature)
constexpr int kNaMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// Block tables of one call, passed by value: block k holds the states
// st[first[k] .. first[k] + blk_m[k] - 1] in ascending index order.
struct NaTables {
    int nblk, max_m;
    int blk_m[kNaMaxBlocks];
    int first[kNaMaxBlocks];
    unsigned short st[kNaMaxStates];
};

__global__ void k_na_tables(NaTables t, int *__restrict__ blk_states, int *__restrict__ blk_m,
                            int *__restrict__ status) {
    for (int e = threadIdx.x; e < t.nblk * t.max_m; e += blockDim.x) {
        const int k = e / t.max_m, q = e % t.max_m;
        blk_states[e] = q < t.blk_m[k] ? (int)t.st[t.first[k] + q] : 0;
    }
    for (int k = threadIdx.x; k < t.nblk; k += blockDim.x) blk_m[k] = t.blk_m[k];
    if (threadIdx.x == 0) *status = 0;
}

__global__ void k_na_zero(double *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

struct DeviceScratch {
    double *ptr = nullptr;   // device address of scratch on this GPU
    size_t chol_lds = 0;     // dynamic LDS of the in-LDS Cholesky kernels
};

std::mutex g_mu;
DeviceScratch g_dev[kNaMaxDevices];

// Looks up the current device's copy of scratch (once per device).
int device_scratch(DeviceScratch *&out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return detail::fail(nullptr, MOCCA_ERR_HIP, "hipGetDevice failed: %s",
                            hipGetErrorString(e));
    if (dev < 0 || dev >= kNaMaxDevices)
        return detail::fail(nullptr, MOCCA_ERR_UNSUPPORTED, "device %d (at most %d devices)", dev,
                            kNaMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch &d = g_dev[dev];
    if (!d.ptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return detail::fail(nullptr, MOCCA_ERR_HIP,
                                "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                hipGetErrorString(e));
        d.ptr = static_cast<double *>(p);
        d.chol_lds = cholesky_lds_limit();
    }
    out = &d;
    return 0;
}

size_t align_doubles(size_t n) { return (n + 31) / 32 * 32; }   // 256-byte pieces

}  // namespace
}  // namespace mocca

using namespace mocca;
using namespace mocca::detail;

extern "C" {

int mocca_gram_schmidt_noalloc(const double *psi, double *out, int nwt, long n, const int *qn,
                               const double *energy, double dv, int *status, void *stream) {
    if (!psi || !out || !qn || !energy)
        return fail(nullptr, MOCCA_ERR_ARG, "gram_schmidt_noalloc: null argument");
    if (nwt < 1 || nwt > kNaMaxStates)
        return fail(nullptr, MOCCA_ERR_ARG, "gram_schmidt_noalloc: %d states (1 .. %d)", nwt,
                    kNaMaxStates);
    if (n < 1) return fail(nullptr, MOCCA_ERR_ARG, "gram_schmidt_noalloc: %ld points", n);
    const long K4 = 4 * n;
    const char *a0 = (const char *)psi, *a1 = a0 + (size_t)nwt * K4 * sizeof(double);
    const char *b0 = (const char *)out, *b1 = b0 + (size_t)nwt * K4 * sizeof(double);
    if (a0 < b1 && b0 < a1)
        return fail(nullptr, MOCCA_ERR_ARG, "gram_schmidt_noalloc: out overlaps psi");
    // Blocks in the order of mocca_set_states_meta: (isospin, parity,
    // signature), each with its states in ascending index.
    NaTables t{};
    int mmax = 0, pos = 0;
    for (int iso = -1; iso <= 1; iso += 2)
        for (int P = -1; P <= 1; P += 2)
            for (int S = -1; S <= 1; S += 2) {
                const int first = pos;
                for (int i = 0; i < nwt; ++i)
                    if (qn[4 * i] == iso && qn[4 * i + 1] == P && qn[4 * i + 2] == S)
                        t.st[pos++] = (unsigned short)i;
                if (pos == first) continue;
                t.first[t.nblk] = first;
                t.blk_m[t.nblk] = pos - first;
                mmax = std::max(mmax, pos - first);
                ++t.nblk;
            }
    if (pos != nwt)
        return fail(nullptr, MOCCA_ERR_ARG,
                    "gram_schmidt_noalloc: quantum numbers must be -1 or +1 (isospin, parity, "
                    "signature)");
    if (mmax > kMaxBlock)
        return fail(nullptr, MOCCA_ERR_UNSUPPORTED,
                    "gram_schmidt_noalloc: a block holds %d states (max %d)", mmax, kMaxBlock);
    GsPlan gp = gs_plan(t.nblk, mmax, K4, nwt, sizeof(double));
    t.max_m = gp.max_m;
    // scratch layout (doubles, 256-byte pieces): G, T, work, partials, then
    // the int tables (block states, block sizes, energy order, status)
    const size_t nG = (size_t)t.nblk * gp.max_m * gp.max_m;
    const size_t ints = (size_t)2 * t.nblk * gp.max_m + kNaMaxBlocks + 1;
    const size_t fixed = 3 * align_doubles(nG) + align_doubles((ints + 1) / 2);
    const size_t cap = kScratchBytes / sizeof(double);
    // fewer K splits than a solve handle would take when the partial
    // matrices do not fit (the reduction order, not the result, changes)
    while (fixed + align_doubles(gp.gpart_doubles()) > cap &&
           (gp.nsplit > 1 || gp.nsplit_mfma > 1)) {
        gp.nsplit = std::max(1, gp.nsplit / 2);
        gp.nsplit_mfma = std::max(1, gp.nsplit_mfma / 2);
    }
    if (fixed + align_doubles(gp.gpart_doubles()) > cap)
        return fail(nullptr, MOCCA_ERR_UNSUPPORTED,
                    "gram_schmidt_noalloc: %d states in %d blocks (largest %d) need %zu bytes of "
                    "workspace, more than the %zu of the static scratch",
                    nwt, t.nblk, mmax,
                    (fixed + align_doubles(gp.gpart_doubles())) * sizeof(double),
                    kScratchBytes);
    DeviceScratch *ds = nullptr;
    if (int rc = device_scratch(ds)) return rc;
    double *G = ds->ptr, *Tm = G + align_doubles(nG), *work = Tm + align_doubles(nG);
    double *part = work + align_doubles(nG);
    int *blk_states = reinterpret_cast<int *>(part + align_doubles(gp.gpart_doubles()));
    int *blk_m = blk_states + (size_t)t.nblk * gp.max_m;
    int *perm = blk_m + kNaMaxBlocks;
    int *st_word = status ? status : perm + (size_t)t.nblk * gp.max_m;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const char *name = "gram_schmidt_noalloc";
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_na_tables, dim3(1), dim3(256), 0, st, t, blk_states, blk_m, st_word);
    hipLaunchKernelGGL(k_na_zero, dim3((unsigned)std::min<size_t>(1024, (3 * align_doubles(nG) + 255) / 256)),
                       dim3(256), 0, st, G, (long)(3 * align_doubles(nG)));
    if (gp.mfma_gram) {
        launch_gram_mfma<double>(psi, K4, blk_states, blk_m, t.nblk, gp.max_m, gp.mpad, part,
                                 gp.nsplit_mfma, dv, G, st);
    } else {
        launch_gram<double>(psi, K4, blk_states, blk_m, t.nblk, gp.max_m, gp.mmax, gp.gtile,
                            part, gp.nsplit, st);
        launch_gram_reduce(part, blk_m, t.nblk, gp.max_m, gp.mmax, gp.gtile, gp.nsplit, dv, G,
                           st);
    }
    launch_cholesky_ld(G, blk_states, blk_m, t.nblk, gp.max_m, gp.mmax, energy, Tm, perm, work,
                       st_word, ds->chol_lds, st);
    if (gp.mfma_apply)
        launch_gs_apply_mfma<double>(psi, out, K4, perm, blk_m, t.nblk, gp.max_m, gp.mpad, Tm,
                                     st);
    else
        launch_gs_apply<double>(psi, out, nullptr, K4, perm, blk_m, t.nblk, gp.max_m, Tm, 0, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(nullptr, MOCCA_ERR_HIP, "launch of %s failed: %s", name,
                    hipGetErrorString(e));
    if (const char *k = take_launch_failure())
        return fail(nullptr, MOCCA_ERR_HIP,
                    "launch of %s: the dynamic LDS limit of %s could not be raised", name, k);
    return 0;
}

}  // extern "C"
