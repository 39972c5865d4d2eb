"""mocca_gram_schmidt_noalloc (the allocation-free Gram-Schmidt) against
mocca_gram_schmidt through a solve handle, on two calls in a row, the second
with fewer states, fewer points and a different block structure: the static
workspace keeps the first call's contents, which the second call must not
see.  Device buffers come from torch (plumbing only); every computation is
the library's."""
import numpy as np
import pytest

import mocca_amd

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

# Tolerance of the step + Gram-Schmidt parity tests (relative, FP64).
TOL = 1e-10


def qn_blocks(sizes):
    """Quantum numbers for blocks of the given sizes, in the order
    (isospin, parity, signature) of the library, interleaved over the
    state index so that blocks are not contiguous."""
    keys = [(i, p, s) for i in (-1, 1) for p in (-1, 1) for s in (-1, 1)]
    qn = []
    left = list(sizes)
    while any(left):
        for k, n in enumerate(left):
            if n:
                qn.append([keys[k][0], keys[k][1], keys[k][2], 1])
                left[k] -= 1
    return np.array(qn, dtype=np.int32)


def handle_gram_schmidt(psi, qn, energy, nx, dv):
    M = psi.shape[0]
    s = mocca_amd.Solve(nx, nx, nx, 0.8, dv, M)
    try:
        s.set_states_meta(qn, np.zeros(M))
        s.upload_states(psi)
        s.set_energies(energy)
        s.gram_schmidt()
        return s.download_states()
    finally:
        s.close()


def noalloc_gram_schmidt(psi, qn, energy, dv):
    M, _, N = psi.shape
    dev = torch.device("cuda:0")
    p = torch.from_numpy(np.ascontiguousarray(psi)).to(dev)
    o = torch.full_like(p, float("nan"))
    e = torch.from_numpy(np.ascontiguousarray(energy)).to(dev)
    st = torch.full((1,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    mocca_amd.gram_schmidt_noalloc(p.data_ptr(), o.data_ptr(), M, N, qn, e.data_ptr(), dv,
                                   st.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return o.cpu().numpy(), int(st.cpu()[0])


@pytest.mark.parametrize("first, second", [
    # (mesh edge, block sizes): MFMA Gram, in-LDS Cholesky and MFMA apply
    # (100 states) first, then small register-Cholesky blocks
    ((16, [100, 0, 0, 0, 0, 60, 0, 0]), (10, [0, 9, 0, 8, 7, 0, 0, 0])),
    # many small blocks first, then one of 130 states (three-job Gram)
    ((12, [12, 10, 8, 6, 12, 10, 8, 6]), (14, [0, 0, 130, 0, 0, 0, 0, 0])),
    # a block of 200 states (vector Gram, global-memory Cholesky, vector
    # apply), then four blocks of 5
    ((10, [0, 0, 0, 200, 0, 0, 0, 0]), (8, [5, 0, 5, 0, 5, 0, 5, 0])),
])
def test_noalloc_gram_schmidt_matches_handle(first, second):
    rng = np.random.default_rng(11)
    for nx, sizes in (first, second):
        qn = qn_blocks(sizes)
        M, N = len(qn), nx ** 3
        dv = 0.512 * 8
        psi = rng.standard_normal((M, 4, N)) * 0.05
        energy = rng.uniform(-40.0, 10.0, M)
        print(f"noalloc case: mesh {nx}^3, {M} states, blocks {sizes}", flush=True)
        ref = handle_gram_schmidt(psi, qn, energy, nx, dv)
        out, status = noalloc_gram_schmidt(psi, qn, energy, dv)
        assert status == 0
        assert np.all(np.isfinite(out))
        err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
        print(f"  max relative difference vs mocca_gram_schmidt: {err:.2e}", flush=True)
        assert err < TOL
        # orthonormal per block: dv <psi_i|psi_j> = delta_ij
        for key in {tuple(q[:3]) for q in qn}:
            idx = [i for i in range(M) if tuple(qn[i][:3]) == key]
            Q = out[idx].reshape(len(idx), -1)
            G = dv * Q @ Q.T
            assert np.max(np.abs(G - np.eye(len(idx)))) < 1e-12


def test_noalloc_rejects_sizes_beyond_the_workspace():
    dev = torch.device("cuda:0")
    x = torch.zeros(8, dtype=torch.float64, device=dev)
    qn = np.tile(np.array([[-1, 1, 1, 1]], dtype=np.int32), (513, 1))
    with pytest.raises(mocca_amd.MoccaError, match="states"):
        mocca_amd.gram_schmidt_noalloc(x.data_ptr(), x.data_ptr() + 64, 513, 1, qn,
                                       x.data_ptr(), 1.0)
