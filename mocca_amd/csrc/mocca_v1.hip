// MOCCa v1 wavefunction files: the record layout WriteMOCCa_v1 writes and
// ReadMOCCa_v1 reads (InOut.f90:905-1416), as host code, plus the
// checkpoint of a solve handle into such a file and the restart of a
// handle from one.
//
// A file is a sequence of Fortran sequential unformatted records: int32
// byte count, the bytes, the same count again.  Records are streamed one
// at a time, so a file of any size passes through a bounded host buffer per
// record (the largest is one state's grid or the U, V pair).
#include "mocca_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <vector>

using namespace mocca;
using namespace mocca::detail;

namespace {

constexpr size_t kMaxRecord = 0x7fffffff;   // one int32 count per record

struct Piece {
    const void *p;   // NULL: zeros
    size_t n;        // bytes
};

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

class Writer {
public:
    explicit Writer(FILE *f) : f_(f) {}
    bool rec(std::initializer_list<Piece> parts) {
        size_t n = 0;
        for (const Piece &q : parts) n += q.n;
        if (n > kMaxRecord) return ok_ = false;
        const int32_t len = (int32_t)n;
        ok_ = ok_ && fwrite(&len, 4, 1, f_) == 1;
        for (const Piece &q : parts) {
            if (!ok_ || !q.n) continue;
            if (q.p) {
                ok_ = fwrite(q.p, 1, q.n, f_) == q.n;
            } else {
                static const char zero[4096] = {};
                for (size_t k = 0; k < q.n && ok_; k += sizeof zero) {
                    const size_t c = std::min(sizeof zero, q.n - k);
                    ok_ = fwrite(zero, 1, c, f_) == c;
                }
            }
        }
        ok_ = ok_ && fwrite(&len, 4, 1, f_) == 1;
        return ok_;
    }
    bool ok() const { return ok_; }

private:
    FILE *f_;
    bool ok_ = true;
};

class Reader {
public:
    explicit Reader(FILE *f) : f_(f) {}
    // Next record's length (false at the end of the file or on a bad count).
    bool begin(size_t &n) {
        int32_t len = 0;
        if (fread(&len, 4, 1, f_) != 1 || len < 0) return false;
        n_ = (size_t)len;
        n = n_;
        left_ = n_;
        return true;
    }
    // Read (dst != NULL) or skip the next bytes of the current record.
    bool take(void *dst, size_t n) {
        if (n > left_) return false;
        left_ -= n;
        if (dst) return fread(dst, 1, n, f_) == n;
        return fseek(f_, (long)n, SEEK_CUR) == 0;
    }
    // Skip the rest and check the trailing count.
    bool end() {
        if (left_ && fseek(f_, (long)left_, SEEK_CUR) != 0) return false;
        left_ = 0;
        int32_t len = 0;
        return fread(&len, 4, 1, f_) == 1 && (size_t)len == n_;
    }
    // A record of exactly the given pieces.
    bool rec(std::initializer_list<Piece> parts) {
        size_t n = 0, want = 0;
        for (const Piece &q : parts) want += q.n;
        if (!begin(n) || n != want) return false;
        for (const Piece &q : parts)
            if (!take(const_cast<void *>(q.p), q.n)) return false;
        return end();
    }
    bool at_end() {
        const int c = fgetc(f_);
        if (c == EOF) return true;
        ungetc(c, f_);
        return false;
    }

private:
    FILE *f_;
    size_t n_ = 0, left_ = 0;
};

int bad(const char *path, const char *what) {
    return fail(nullptr, MOCCA_ERR_ARG, "%s: not a MOCCa v1 file (%s)", path, what);
}

}  // namespace

extern "C" {

int mocca_v1_read(const char *path, mocca_v1_info *info, const mocca_v1_arrays *arr) {
    if (!path || !info) return fail(nullptr, MOCCA_ERR_ARG, "null argument");
    File fh(fopen(path, "rb"));
    if (!fh) return fail(nullptr, MOCCA_ERR_ARG, "%s: cannot open", path);
    Reader r(fh.get());
    mocca_v1_info &I = *info;
    std::memset(&I, 0, sizeof I);
    const mocca_v1_arrays none{};
    const mocca_v1_arrays &A = arr ? *arr : none;
    int version = 0;
    if (!r.rec({{&version, 4}}) || version != 1) return bad(path, "version record");
    if (!r.rec({{&I.energy, 8}, {I.denergy, 56}})) return bad(path, "energy record");
    int n3[3];
    if (!r.rec({{n3, 12}, {&I.dx, 8}})) return bad(path, "mesh record");
    I.nx = n3[0];
    I.ny = n3[1];
    I.nz = n3[2];
    int sym[5];
    if (!r.rec({{sym, 20}})) return bad(path, "symmetry record");
    I.trc = sym[0];
    I.pc = sym[1];
    I.sc = sym[2];
    I.tsc = sym[3];
    I.ic = sym[4];
    if (!r.rec({{&I.nwt, 4}, {&I.neutrons, 8}, {&I.protons, 8}})) return bad(path, "nwt record");
    if (I.nx <= 0 || I.ny <= 0 || I.nz <= 0 || I.nwt <= 0) return bad(path, "sizes");
    const size_t N = (size_t)I.nx * I.ny * I.nz, M = (size_t)I.nwt;
    // per state: the grid (ReadSpinor, Spinor.f90:759-776), then the scalars
    // and quantum numbers (ReadSpwf, Spwf.f90:2263-2287)
    for (size_t i = 0; i < M; ++i) {
        if (!r.rec({{A.grid ? A.grid + i * 4 * N : nullptr, 32 * N}}))
            return bad(path, "state grid record");
        double sc[7];
        int qn[5];
        if (!r.rec({{sc, 56}, {qn, 20}})) return bad(path, "state record");
        if (A.state) std::memcpy(A.state + 7 * i, sc, sizeof sc);
        if (A.qn) std::memcpy(A.qn + 5 * i, qn, sizeof qn);
    }
    // ReadDensity (Densities.f90:1242-1348)
    const mocca_density &d = A.dens;
    const size_t b2 = 16 * N, b6 = 48 * N, b18 = 144 * N;
    if (!r.rec({{d.rho, b2}, {d.derrho, b6}, {d.laprho, b2}})) return bad(path, "rho record");
    if (!r.rec({{d.tau, b2}})) return bad(path, "tau record");
    if (!r.rec({{d.nablaJ, b2}})) return bad(path, "nablaJ record");
    size_t n = 0;
    if (!r.begin(n) || !r.end()) return bad(path, "JMuNu record");   // not used here
    if (I.trc) {
        for (int k = 0; k < 5; ++k)
            if (!r.begin(n) || !r.end()) return bad(path, "time-odd record");
    } else {
        if (!r.rec({{d.vecj, b6}, {d.rotvecj, b6}})) return bad(path, "vecj record");
        if (!r.rec({{d.vecs, b6}, {d.rots, b6}, {d.ders, b18}})) return bad(path, "vecs record");
        for (int k = 0; k < 3; ++k)   // LapS .., VecT, VecF (no such terms here)
            if (!r.begin(n) || !r.end()) return bad(path, "time-odd record");
    }
    if (!r.rec({{I.fd_order, 12}})) return bad(path, "FD record");
    if (!r.rec({{I.force, 200}})) return bad(path, "force name record");
    if (!r.rec({{I.skyrme, 112}})) return bad(path, "force record");
    if (!r.rec({{nullptr, 200}, {&I.hbar, 8}, {I.hbm, 16}, {&I.com1body, 4}, {&I.com2body, 4}}))
        return bad(path, "constants record");
    if (!r.rec({{I.B, 224}})) return bad(path, "functional record");
    if (!r.rec({{&I.pairing_type, 4}, {I.pairing_strength, 16}, {I.pairing_alpha, 16},
                {I.pairing_cut, 16}}))
        return bad(path, "pairing record");
    if (!r.begin(n) || (n != 0 && n != 8 * M)) return bad(path, "cutoff record");
    I.has_cutoffs = n != 0;
    if (!r.take(I.has_cutoffs ? A.cutoffs : nullptr, n) || !r.end())
        return bad(path, "cutoff record");
    if (!r.rec({{I.fermi, 16}, {I.lnlambda, 16}})) return bad(path, "Fermi record");
    // HFB matrices (InOut.f90:1272-1290 / WriteMOCCa_v1 :1007-1011)
    const size_t S = I.trc ? 2 * M : M;
    I.hfb_size = (int)S;
    I.npar = I.pc ? 2 : 1;
    I.niso = I.ic ? 2 : 1;
    const size_t nb = (size_t)I.npar * I.niso;
    if (!r.rec({{I.blocksizes, 4 * nb}})) return bad(path, "blocksizes record");
    const size_t uv = S * 2 * S * 4 * 16, rk = S * S * 4 * 16;
    if (!r.rec({{A.u, uv}, {A.v, uv}})) return bad(path, "U, V record");
    if (!r.rec({{A.hfb_columns, 4 * S * nb}})) return bad(path, "HFBColumns record");
    if (!r.rec({{A.rho_hfb, rk}, {A.kappa_hfb, rk}})) return bad(path, "Rho, Kappa record");
    if (!r.rec({{A.can_transfo, S * S * nb * 16}})) return bad(path, "CanTransfo record");
    if (!r.rec({{I.omega, 24}, {I.crank_values, 24}})) return bad(path, "cranking record");
    // constrained moments to the end of the file (ReadMoment)
    while (!r.at_end()) {
        if (I.nmoments >= MOCCA_V1_MAX_MOMENTS) return bad(path, "too many moments");
        mocca_v1_moment &m = I.moments[I.nmoments];
        int hd[5];
        if (!r.rec({{hd, 20}})) return bad(path, "moment record");
        m.constraint_type = hd[0];
        m.isoswitch = hd[1];
        m.l = hd[2];
        m.m = hd[3];
        m.impart = hd[4];
        if (!r.rec({{m.intensity, 16}, {m.constraint, 16}, {m.true_constraint, 16},
                    {m.value, 16}, {m.old_value, 112}}))
            return bad(path, "moment values record");
        ++I.nmoments;
    }
    return 0;
}

int mocca_v1_write(const char *path, const mocca_v1_info *info, const mocca_v1_arrays *arr) {
    if (!path || !info || !arr) return fail(nullptr, MOCCA_ERR_ARG, "null argument");
    const mocca_v1_info &I = *info;
    const mocca_v1_arrays &A = *arr;
    if (I.nx <= 0 || I.ny <= 0 || I.nz <= 0 || I.nwt <= 0)
        return fail(nullptr, MOCCA_ERR_ARG, "v1 file: bad mesh or nwt");
    if (!A.grid || !A.state || !A.qn)
        return fail(nullptr, MOCCA_ERR_ARG, "v1 file: grid, state and qn are required");
    const size_t N = (size_t)I.nx * I.ny * I.nz, M = (size_t)I.nwt;
    const size_t S = I.trc ? 2 * M : M;
    const size_t nb = (size_t)(I.pc ? 2 : 1) * (I.ic ? 2 : 1);
    if ((size_t)I.hfb_size != S || (size_t)I.npar * I.niso != nb)
        return fail(nullptr, MOCCA_ERR_ARG, "v1 file: hfb_size / npar / niso do not follow "
                                            "from nwt and the symmetries");
    if (I.nmoments < 0 || I.nmoments > MOCCA_V1_MAX_MOMENTS)
        return fail(nullptr, MOCCA_ERR_ARG, "v1 file: nmoments out of range");
    if (32 * N > kMaxRecord || S * 2 * S * 4 * 16 * 2 > kMaxRecord || 144 * N + 96 * N > kMaxRecord)
        return fail(nullptr, MOCCA_ERR_UNSUPPORTED, "v1 file: a record would exceed 2 GiB");
    File fh(fopen(path, "wb"));
    if (!fh) return fail(nullptr, MOCCA_ERR_ARG, "%s: cannot create", path);
    Writer w(fh.get());
    const int version = 1;
    const int n3[3] = {I.nx, I.ny, I.nz};
    const int sym[5] = {I.trc, I.pc, I.sc, I.tsc, I.ic};
    w.rec({{&version, 4}});
    w.rec({{&I.energy, 8}, {I.denergy, 56}});
    w.rec({{n3, 12}, {&I.dx, 8}});
    w.rec({{sym, 20}});
    w.rec({{&I.nwt, 4}, {&I.neutrons, 8}, {&I.protons, 8}});
    for (size_t i = 0; i < M; ++i) {
        w.rec({{A.grid + i * 4 * N, 32 * N}});
        w.rec({{A.state + 7 * i, 56}, {A.qn + 5 * i, 20}});
    }
    const mocca_density &d = A.dens;
    const size_t b2 = 16 * N, b6 = 48 * N, b18 = 144 * N;
    w.rec({{d.rho, b2}, {d.derrho, b6}, {d.laprho, b2}});
    w.rec({{d.tau, b2}});
    w.rec({{d.nablaJ, b2}});
    w.rec({});   // JMuNu: not allocated without tensor terms
    if (I.trc) {
        for (int k = 0; k < 5; ++k) w.rec({});
    } else {
        w.rec({{d.vecj, b6}, {d.rotvecj, b6}});
        w.rec({{d.vecs, b6}, {d.rots, b6}, {d.ders, b18}});
        for (int k = 0; k < 3; ++k) w.rec({});
    }
    w.rec({{I.fd_order, 12}});
    w.rec({{I.force, 200}});
    w.rec({{I.skyrme, 112}});
    w.rec({{I.force, 200}, {&I.hbar, 8}, {I.hbm, 16}, {&I.com1body, 4}, {&I.com2body, 4}});
    w.rec({{I.B, 224}});
    w.rec({{&I.pairing_type, 4}, {I.pairing_strength, 16}, {I.pairing_alpha, 16},
           {I.pairing_cut, 16}});
    if (I.has_cutoffs)
        w.rec({{A.cutoffs, 8 * M}});
    else
        w.rec({});
    w.rec({{I.fermi, 16}, {I.lnlambda, 16}});
    const size_t uv = S * 2 * S * 4 * 16, rk = S * S * 4 * 16;
    w.rec({{I.blocksizes, 4 * nb}});
    w.rec({{A.u, uv}, {A.v, uv}});
    w.rec({{A.hfb_columns, 4 * S * nb}});
    w.rec({{A.rho_hfb, rk}, {A.kappa_hfb, rk}});
    w.rec({{A.can_transfo, S * S * nb * 16}});
    w.rec({{I.omega, 24}, {I.crank_values, 24}});
    for (int k = 0; k < I.nmoments; ++k) {
        const mocca_v1_moment &m = I.moments[k];
        const int hd[5] = {m.constraint_type, m.isoswitch, m.l, m.m, m.impart};
        w.rec({{hd, 20}});
        w.rec({{m.intensity, 16}, {m.constraint, 16}, {m.true_constraint, 16}, {m.value, 16},
               {m.old_value, 112}});
    }
    if (!w.ok() || fflush(fh.get()) != 0)
        return fail(nullptr, MOCCA_ERR_ARG, "%s: write failed", path);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Handle-level checkpoint and restart
namespace {

// (P, it) block of the reference's HFB matrices (PrepareHFBModule,
// HFB.f90:248-294): rows i = 1..S in order, state mod(i - 1, nwt) + 1, the
// time-reversed copy for i > nwt; Fortran flat index of element (a, b) of
// block (P, it) in an (S, S, 2, 2) array.
struct RefBlocks {
    size_t S;
    std::vector<int> rows[2][2];   // [P][it]: the rows (0-based i) of the block
    size_t at(int P, int it, size_t a, size_t b) const {
        return a + S * (b + S * (P + 2 * (size_t)it));
    }
};

RefBlocks ref_blocks(const std::vector<int> &qn, int M, bool trc) {
    RefBlocks rb;
    rb.S = trc ? 2 * (size_t)M : (size_t)M;
    for (size_t i = 0; i < rb.S; ++i) {
        const int s = (int)(i % M);
        const int P = (qn[4 * s + 1] + 3) / 2 - 1, it = (qn[4 * s] + 3) / 2 - 1;
        rb.rows[P][it].push_back((int)i);
    }
    return rb;
}

int dl(mocca_solve *h, void *dst, const void *src, size_t bytes) {
    HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int ul(mocca_solve *h, void *dst, const void *src, size_t bytes) {
    HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

extern "C" {

int mocca_v1_save(mocca_solve *h, const char *path, const mocca_v1_info *extra) {
    if (!h || !path) return fail(h, MOCCA_ERR_ARG, "null argument");
    if (!h->meta_set) return fail(h, MOCCA_ERR_STATE, "mocca_set_states_meta not called");
    if (!h->scf.set) return fail(h, MOCCA_ERR_STATE, "mocca_scf_setup not called");
    const int M = h->M;
    const size_t N = h->N;
    const bool trc = h->mesh.time_reversal != 0;
    const Scf &s = h->scf;
    const Hfb &p = h->hfb;
    HIPCHK(h, hipStreamSynchronize(h->st));
    auto info = std::make_unique<mocca_v1_info>();
    mocca_v1_info &I = *info;
    std::memset(&I, 0, sizeof I);
    // header (WriteMOCCa_v1 :957-968)
    I.energy = s.ehist[0];
    for (int k = 0; k < 7; ++k) I.denergy[k] = s.ehist[0] - s.ehist[k + 1];
    I.nx = h->mesh.nx;
    I.ny = h->mesh.ny;
    I.nz = h->mesh.nz;
    I.dx = h->mesh.dx;
    I.trc = trc;
    I.pc = I.sc = I.tsc = I.ic = 1;
    I.nwt = M;
    I.neutrons = s.sk.neutrons;
    I.protons = s.sk.protons;
    // states (WriteSpwf): grid, Occupation, Energy, Dispersion, Norm, J
    std::vector<double> grid((size_t)M * 4 * N), st((size_t)M * 7), am((size_t)M * 21);
    std::vector<int> qn((size_t)M * 5);
    if (int rc = mocca_download_states(h, 0, M, grid.data())) return rc;
    if (int rc = mocca_angular_momentum(h, am.data(), nullptr, nullptr)) return rc;
    std::vector<double> occ(M), E(M), D(M), nrm(M);
    int rc;
    if ((rc = dl(h, occ.data(), h->b.occ, M * sizeof(double))) ||
        (rc = dl(h, E.data(), h->b.E, M * sizeof(double))) ||
        (rc = dl(h, D.data(), h->b.D, M * sizeof(double))) ||
        (rc = dl(h, nrm.data(), h->b.nrm, M * sizeof(double))))
        return rc;
    for (int i = 0; i < M; ++i) {
        double *o = st.data() + 7 * (size_t)i;
        o[0] = occ[i];
        o[1] = E[i];
        o[2] = D[i];
        o[3] = nrm[i];
        for (int c = 0; c < 3; ++c) o[4 + c] = am[21 * (size_t)i + c];   // J (DiagAng)
        const int *q = h->qn.data() + 4 * (size_t)i;
        int *o2 = qn.data() + 5 * (size_t)i;
        o2[0] = q[0];   // Isospin
        o2[1] = q[3];   // TimeSimplex
        o2[2] = q[1];   // Parity
        o2[3] = q[2];   // Signature
        o2[4] = 1;      // TimeReversal (NewWaveFunction's)
    }
    // the current densities in the file's order
    std::vector<double> dens(18 * N), der(38 * N);
    if ((rc = dl(h, dens.data(), h->b.dens, dens.size() * sizeof(double))) ||
        (rc = dl(h, der.data(), h->b.derived, der.size() * sizeof(double))))
        return rc;
    std::vector<double> rho(2 * N), tau(2 * N), nj(2 * N), vj(6 * N), vs(6 * N);
    for (int it = 0; it < 2; ++it) {
        const double *src = dens.data() + 9 * N * it;
        std::copy(src, src + N, rho.begin() + N * it);
        std::copy(src + N, src + 2 * N, tau.begin() + N * it);
        std::copy(src + 2 * N, src + 3 * N, nj.begin() + N * it);
        std::copy(src + 3 * N, src + 6 * N, vj.begin() + 3 * N * it);
        std::copy(src + 6 * N, src + 9 * N, vs.begin() + 3 * N * it);
    }
    mocca_v1_arrays A{};
    A.grid = grid.data();
    A.state = st.data();
    A.qn = qn.data();
    A.dens.rho = rho.data();
    A.dens.derrho = der.data();
    A.dens.laprho = der.data() + 6 * N;
    A.dens.tau = tau.data();
    A.dens.nablaJ = nj.data();
    if (!trc) {
        A.dens.vecj = vj.data();
        A.dens.rotvecj = der.data() + 8 * N;
        A.dens.vecs = vs.data();
        A.dens.rots = der.data() + 14 * N;
        // after MixingScheme 0 the reference's density holds DerS = 0: its
        // scaling and sum of densities carry DerS only with LapS allocated
        // (Densities.f90:262-266,327-333), which no functional here needs
        A.dens.ders = s.mixing_scheme == 0 && s.iteration > 0 ? nullptr : der.data() + 20 * N;
    }
    // non-essential records
    I.fd_order[0] = h->mesh.lagrange ? -1 : 3;
    I.fd_order[1] = h->mesh.lagrange ? -1 : 4;
    I.fd_order[2] = extra ? extra->fd_order[2] : 2;
    if (extra) {
        std::memcpy(I.force, extra->force, sizeof I.force);
        std::memcpy(I.skyrme, extra->skyrme, sizeof I.skyrme);
        I.hbar = extra->hbar;
        I.com2body = extra->com2body;
        std::memcpy(I.omega, extra->omega, sizeof I.omega);
        std::memcpy(I.crank_values, extra->crank_values, sizeof I.crank_values);
    } else {
        std::memset(I.force, ' ', sizeof I.force);
        I.hbar = 6.58211928;
    }
    I.hbm[0] = s.sk.hbm[0];
    I.hbm[1] = s.sk.hbm[1];
    I.com1body = s.sk.com1body;
    const double Bs[20] = {s.sk.B1,  s.sk.B2,  s.sk.B3,   s.sk.B4,   s.sk.B5,   s.sk.B6,   s.sk.B7a,
                           s.sk.B7b, s.sk.B8a, s.sk.B8b,  s.sk.byt3a, s.sk.byt3b, s.sk.B9, s.sk.B9q,
                           s.sk.B10, s.sk.B11, s.sk.B12a, s.sk.B12b, s.sk.B13a, s.sk.B13b};
    std::copy(Bs, Bs + 20, I.B);   // B14 .. B21: zero (tensor terms are rejected)
    // pairing (:990-1002)
    I.pairing_type = p.on ? 2 : (p.bcs ? 1 : 0);
    if (p.in_set) {
        for (int it = 0; it < 2; ++it) {
            I.pairing_strength[it] = p.in.strength[it];
            I.pairing_alpha[it] = p.in.alpha[it];
            I.pairing_cut[it] = p.in.cut[it];
        }
    } else {
        I.pairing_cut[0] = I.pairing_cut[1] = 5.0;   // the namelist default (CutNeutron)
    }
    std::vector<double> cut;
    if (p.on || p.bcs) {
        // PCutoffs as SolvePairing leaves them: ComputePairingCutoffs with
        // the Fermi energies it found (Pairing.f90:849,
        // PairingInteraction.f90:182-199, SymmetricFermi)
        if ((rc = dl(h, I.fermi, p.fermi, 2 * sizeof(double)))) return rc;
        cut.resize(M);
        for (int i = 0; i < M; ++i) {
            const int it = (h->qn[4 * (size_t)i] + 3) / 2 - 1;
            const double up = (E[i] - I.fermi[it] - p.k.cut[it]) / p.k.mu[it];
            const double down = -(E[i] - I.fermi[it] + p.k.cut[it]) / p.k.mu[it];
            cut[i] = std::sqrt(std::sqrt(1.0 / (1.0 + std::exp(up)))) *
                     std::sqrt(std::sqrt(1.0 / (1.0 + std::exp(down))));
        }
        I.has_cutoffs = 1;
        A.cutoffs = cut.data();
    } else if (extra) {
        I.fermi[0] = extra->fermi[0];
        I.fermi[1] = extra->fermi[1];
    } else {
        I.fermi[0] = I.fermi[1] = -10.0;   // Pairing.f90:64
    }
    // HFB matrices (:1003-1011): blocks of the reference's row order
    const RefBlocks rb = ref_blocks(h->qn, M, trc);
    const size_t S = rb.S;
    I.hfb_size = (int)S;
    I.npar = I.niso = 2;
    for (int it = 0; it < 2; ++it)
        for (int P = 0; P < 2; ++P) I.blocksizes[P + 2 * it] = (int)rb.rows[P][it].size();
    std::vector<double> rk(2 * S * S * 4, 0.0), kap(2 * S * S * 4, 0.0);
    std::vector<int> cols(S * 4, 0);
    if (p.on) {
        // RhoHFB, KappaHFB of the last pairing step: the library keeps each
        // (isospin, parity) block of 2h rows (the h states, then their
        // time-reversed partners, the reference's order) row by row
        const size_t tot = (size_t)p.k.blk[p.k.nblk - 1].off +
                           4 * (size_t)p.k.blk[p.k.nblk - 1].h * p.k.blk[p.k.nblk - 1].h;
        std::vector<double> hr(tot), hk(tot);
        if ((rc = dl(h, hr.data(), p.rho, tot * sizeof(double))) ||
            (rc = dl(h, hk.data(), p.kappa, tot * sizeof(double))))
            return rc;
        for (int b = 0; b < p.k.nblk; ++b) {
            const HfbBlockK &bk = p.k.blk[b];
            const size_t n = 2 * (size_t)bk.h;
            for (size_t j = 0; j < n; ++j)
                for (size_t i = 0; i < n; ++i) {
                    const size_t at = rb.at(bk.P, bk.it, i, j);
                    rk[2 * at] = hr[bk.off + i * n + j];
                    kap[2 * at] = hk[bk.off + i * n + j];
                }
            // HFBColumns: the h quasiparticles of signature -1 with the
            // largest energies (columns 2n .. 2n - h + 1), then those of
            // signature +1 (h + 1 .. n)
            int *c = cols.data() + S * (bk.P + 2 * (size_t)bk.it);
            for (int a = 0; a < bk.h; ++a) c[a] = (int)(2 * n) - a;
            for (size_t a = bk.h; a < n; ++a) c[a] = (int)a + 1;
        }
        A.hfb_columns = cols.data();
    } else if (p.bcs) {
        // GuessHFBMatrices for PairingType 1 (HFB.f90:3719-3744): row i of
        // the first half of a block paired with row i + size/2,
        // kappa = +-sqrt(o (1 - o)), o = half the state's occupation
        for (int P = 0; P < 2; ++P)
            for (int it = 0; it < 2; ++it) {
                const std::vector<int> &r = rb.rows[P][it];
                const size_t half = r.size() / 2;
                for (size_t a = 0; a < half; ++a) {
                    const double o = 0.5 * occ[r[a] % M];
                    const double kv = std::sqrt(o * (1.0 - o));
                    kap[2 * rb.at(P, it, a, a + half)] = kv;
                    kap[2 * rb.at(P, it, a + half, a)] = -kv;
                }
            }
        A.hfb_columns = cols.data();
    } else {
        // GuessHFBMatrices for PairingType 0 (HFB.f90:3686-3715): kappa =
        // +-0.1 between rows of opposite signature (a time-reversed copy
        // has the opposite signature), + above the diagonal
        for (int P = 0; P < 2; ++P)
            for (int it = 0; it < 2; ++it) {
                const std::vector<int> &r = rb.rows[P][it];
                auto sig = [&](int i) {
                    const int sgn = h->qn[4 * (i % M) + 2];
                    return i >= M ? -sgn : sgn;
                };
                for (size_t a = 0; a < r.size(); ++a)
                    for (size_t b = a + 1; b < r.size(); ++b)
                        if (sig(r[a]) * sig(r[b]) < 0) {
                            kap[2 * rb.at(P, it, a, b)] = 0.1;
                            kap[2 * rb.at(P, it, b, a)] = -0.1;
                        }
            }
        A.hfb_columns = cols.data();
    }
    // A.u, A.v, A.can_transfo: NULL (zeros)
    A.rho_hfb = rk.data();
    A.kappa_hfb = kap.data();
    // constrained moments (:1017-1034 WriteMoment), in the order of the
    // moment list
    if (s.ncon) {
        std::vector<ConstraintK> ck(s.ncon);
        std::vector<double> hist(16 * (size_t)s.ncon);
        if ((rc = dl(h, ck.data(), s.con, ck.size() * sizeof(ConstraintK))) ||
            (rc = dl(h, hist.data(), s.chist, hist.size() * sizeof(double))))
            return rc;
        std::vector<int> order(s.ncon);
        for (int c = 0; c < s.ncon; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return ck[a].idx < ck[b].idx; });
        for (int c : order) {
            const ConstraintK &k = ck[c];
            if (k.type == 0) continue;
            if (I.nmoments >= MOCCA_V1_MAX_MOMENTS)
                return fail(h, MOCCA_ERR_UNSUPPORTED, "more than %d constraints",
                            MOCCA_V1_MAX_MOMENTS);
            mocca_v1_moment &m = I.moments[I.nmoments++];
            m.constraint_type = k.type;
            m.isoswitch = k.iso;
            m.l = s.lm[2 * k.idx];
            m.m = s.lm[2 * k.idx + 1];
            m.impart = 0;
            for (int q = 0; q < 2; ++q) {
                const int qq = k.iso == 1 ? 0 : q;
                m.intensity[q] = k.inten[qq];
                m.constraint[q] = k.con[qq];
                m.true_constraint[q] = k.truec[qq];
                m.value[q] = hist[16 * (size_t)c + q];
            }
            for (int q = 0; q < 14; ++q) m.old_value[q] = hist[16 * (size_t)c + 2 + q];
        }
    }
    const int wrc = mocca_v1_write(path, &I, &A);
    if (wrc) return fail(h, wrc, "%s", mocca_last_error(nullptr));
    return 0;
}

int mocca_v1_load(mocca_solve *h, const char *path, int recalc, mocca_v1_info *info_out) {
    if (!h || !path) return fail(h, MOCCA_ERR_ARG, "null argument");
    auto info = std::make_unique<mocca_v1_info>();
    mocca_v1_info &I = *info;
    if (int rc = mocca_v1_read(path, &I, nullptr)) return fail(h, rc, "%s", mocca_last_error(nullptr));
    const int M = h->M;
    const size_t N = h->N;
    const bool trc = h->mesh.time_reversal != 0;
    // checks of ReadMOCCa_v1 (:1111-1133) and what this library requires:
    // the same mesh and number of states (no transformation), no symmetry
    // broken on file that the solve conserves
    if (I.nx != h->mesh.nx || I.ny != h->mesh.ny || I.nz != h->mesh.nz)
        return fail(h, MOCCA_ERR_ARG, "%s: mesh %dx%dx%d, the solve's %dx%dx%d", path, I.nx,
                    I.ny, I.nz, h->mesh.nx, h->mesh.ny, h->mesh.nz);
    if (I.nwt != M) return fail(h, MOCCA_ERR_ARG, "%s: nwt %d, the solve's %d", path, I.nwt, M);
    if (!I.pc || !I.sc || !I.tsc || !I.ic)
        return fail(h, MOCCA_ERR_UNSUPPORTED, "%s: parity, signature, time simplex or isospin "
                                              "broken on file", path);
    if ((I.trc != 0) != trc)
        return fail(h, MOCCA_ERR_UNSUPPORTED, "%s: time reversal %s on file, %s in the solve",
                    path, I.trc ? "conserved" : "broken", trc ? "conserved" : "broken");
    const size_t S = trc ? 2 * (size_t)M : (size_t)M;
    std::vector<double> grid((size_t)M * 4 * N), st((size_t)M * 7);
    std::vector<int> qf((size_t)M * 5);
    std::vector<double> rho(2 * N), tau(2 * N), nj(2 * N), der(38 * N, 0.0), vj(6 * N),
        vs(6 * N);
    std::vector<double> rk, kp;
    mocca_v1_arrays A{};
    A.grid = grid.data();
    A.state = st.data();
    A.qn = qf.data();
    if (!recalc) {
        A.dens.rho = rho.data();
        A.dens.derrho = der.data();
        A.dens.laprho = der.data() + 6 * N;
        A.dens.tau = tau.data();
        A.dens.nablaJ = nj.data();
        if (!trc) {
            A.dens.vecj = vj.data();
            A.dens.rotvecj = der.data() + 8 * N;
            A.dens.vecs = vs.data();
            A.dens.rots = der.data() + 14 * N;
            A.dens.ders = der.data() + 20 * N;
        }
    }
    const Hfb &p = h->hfb;
    if (p.on) {
        rk.resize(2 * S * S * 4);
        kp.resize(2 * S * S * 4);
        A.rho_hfb = rk.data();
        A.kappa_hfb = kp.data();
    }
    if (int rc = mocca_v1_read(path, &I, &A)) return fail(h, rc, "%s", mocca_last_error(nullptr));
    // quantum numbers (iso, P, S, TS) and occupations; with HFB set up the
    // block tables depend on them, so they must not change
    std::vector<int> qn(4 * (size_t)M);
    std::vector<double> occ(M), E(M), D(M), nrm(M);
    for (int i = 0; i < M; ++i) {
        const int *q = qf.data() + 5 * (size_t)i;
        qn[4 * i] = q[0];
        qn[4 * i + 1] = q[2];
        qn[4 * i + 2] = q[3];
        qn[4 * i + 3] = q[1];
        occ[i] = st[7 * (size_t)i];
        E[i] = st[7 * (size_t)i + 1];
        D[i] = st[7 * (size_t)i + 2];
        nrm[i] = st[7 * (size_t)i + 3];
    }
    int rc;
    if (h->meta_set && qn == h->qn) {
        if ((rc = mocca_set_occupations(h, occ.data()))) return rc;
    } else if (p.on) {
        return fail(h, MOCCA_ERR_STATE, "%s: quantum numbers differ from the solve's, which "
                                        "its HFB blocks were set up with", path);
    } else if ((rc = mocca_set_states_meta(h, qn.data(), occ.data()))) {
        return rc;
    }
    if ((rc = mocca_upload_states(h, 0, M, grid.data())) ||
        (rc = mocca_set_energies(h, E.data())) ||
        (rc = ul(h, h->b.D, D.data(), M * sizeof(double))) ||
        (rc = ul(h, h->b.nrm, nrm.data(), M * sizeof(double))))
        return rc;
    if (p.bcs && (rc = ul(h, p.fermi, I.fermi, 2 * sizeof(double)))) return rc;
    if (p.on) {
        // Fermi energies and the HFB density matrices (Input :175-183 with
        // TransformHFBMatrices' copy for unchanged symmetries)
        if ((rc = ul(h, p.fermi, I.fermi, 2 * sizeof(double)))) return rc;
        const RefBlocks rb = ref_blocks(h->qn, M, trc);
        const size_t tot = (size_t)p.k.blk[p.k.nblk - 1].off +
                           4 * (size_t)p.k.blk[p.k.nblk - 1].h * p.k.blk[p.k.nblk - 1].h;
        std::vector<double> hr(tot, 0.0), hk(tot, 0.0);
        for (int b = 0; b < p.k.nblk; ++b) {
            const HfbBlockK &bk = p.k.blk[b];
            const size_t n = 2 * (size_t)bk.h;
            for (size_t j = 0; j < n; ++j)
                for (size_t i = 0; i < n; ++i) {
                    const size_t at = rb.at(bk.P, bk.it, i, j);
                    hr[bk.off + i * n + j] = rk[2 * at];
                    hk[bk.off + i * n + j] = kp[2 * at];
                }
        }
        if ((rc = ul(h, p.rho, hr.data(), tot * sizeof(double))) ||
            (rc = ul(h, p.kappa, hk.data(), tot * sizeof(double))))
            return rc;
    }
    if (!recalc) {
        std::vector<double> dens(18 * N);
        for (int it = 0; it < 2; ++it) {
            double *dst = dens.data() + 9 * N * it;
            std::copy(rho.begin() + N * it, rho.begin() + N * (it + 1), dst);
            std::copy(tau.begin() + N * it, tau.begin() + N * (it + 1), dst + N);
            std::copy(nj.begin() + N * it, nj.begin() + N * (it + 1), dst + 2 * N);
            if (trc) {
                std::fill(dst + 3 * N, dst + 9 * N, 0.0);
            } else {
                std::copy(vj.begin() + 3 * N * it, vj.begin() + 3 * N * (it + 1), dst + 3 * N);
                std::copy(vs.begin() + 3 * N * it, vs.begin() + 3 * N * (it + 1), dst + 6 * N);
            }
        }
        if ((rc = ul(h, h->b.dens, dens.data(), dens.size() * sizeof(double))) ||
            (rc = ul(h, h->b.derived, der.data(), der.size() * sizeof(double))))
            return rc;
        h->scf.dens_loaded = true;
    }
    if (info_out) *info_out = I;
    return 0;
}

}  // extern "C"
