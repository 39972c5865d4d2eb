"""MOCCa v1 wavefunction files (WriteMOCCa_v1 / ReadMOCCa_v1,
src/InOut.f90:905-1416) on the host.

The fixtures hold files the reference itself wrote (oracle/ref_driver.f90,
task wv1, after an HF run with time reversal, an HF run with time-odd
fields, a Lagrange mesh and two constrained moments, an HFB run and a BCS
run).

- The oracle's record-by-record restatement parses each file and rebuilds
  it byte for byte, and its contents equal what the same reference run
  dumped independently (final states, single-particle energies,
  occupations, densities, energies, Fermi energies); the HF KappaHFB guess
  of GuessHFBMatrices equals the file's.
- The library's host reader and writer (mocca_v1_read / mocca_v1_write,
  no device use) read every record as the oracle does and write the file
  back identical byte for byte; malformed files are refused.
"""
import os

import numpy as np
import pytest

import mocca_amd
import oracle_lib as ol
from oracle_lib import golden

V1_CASES = ["scf_trc_fd", "scfcon_todd_lag", "scfhfb_trc_fd", "scfbcs_trc_fd"]


@pytest.mark.parametrize("case", V1_CASES)
def test_oracle_parses_reference_files_and_rebuilds_them(case):
    g = golden(case)
    m = g["meta"]
    raw = g["v1_file"].tobytes()
    f = ol.v1_parse(raw)
    assert ol.v1_build(f) == raw
    M, n = m["nwt"], m["nx"]
    N = n ** 3
    assert f["version"] == 1 and list(f["n"]) == [n, n, n] and f["dx"][0] == m["dx"]
    assert list(f["sym"]) == [int(m["trc"]), 1, 1, 1, 1]
    assert int(f["nwt"][0]) == M and list(f["particles"]) == [m["neutrons"], m["protons"]]
    # the states as the run left them
    assert np.array_equal(f["grid"], g["final_psi"].reshape(M, 4, N))
    assert np.array_equal(f["state_energy"], g["final_eo"][:, 0])
    assert np.array_equal(f["state_occupation"], g["final_eo"][:, 1])
    qn = g["qn"]   # (isospin, parity, signature, time simplex)
    assert np.array_equal(f["state_qn"], np.stack([qn[:, 0], qn[:, 3], qn[:, 1], qn[:, 2],
                                                   np.ones(M, np.int32)], axis=1))
    # Gram-Schmidt norms before normalisation: near 1 for small steps
    assert np.all(np.abs(f["state_norm"] - 1.0) < 0.2)
    # energy and its history: the last trajectory row, then decreasing
    assert f["energy"] == g["traj"][-1, 0]
    assert np.allclose(f["energy"] - f["denergy"], g["traj"][-2:-9:-1, 0], rtol=0, atol=1e-12)
    # densities (the mixed ones the run ends with)
    d = f["dens_rec"]
    assert np.array_equal(d[0][:2 * N], g["final_rho"].reshape(-1))
    assert np.array_equal(d[0][2 * N:8 * N], g["final_derrho"].reshape(-1))
    assert np.array_equal(d[0][8 * N:], g["final_laprho"].reshape(-1))
    assert np.array_equal(d[1], g["final_tau"].reshape(-1))
    assert np.array_equal(d[2], g["final_nablaJ"].reshape(-1))
    assert d[3].size == 0
    if m["trc"]:
        assert all(x.size == 0 for x in d[4:])
    else:
        assert d[4].size == 12 * N and d[5].size == 30 * N and all(x.size == 0 for x in d[6:])
    assert np.array_equal(f["B"], g["coef"])
    assert list(f["fd_order"][:2]) == ([-1, -1] if m["lagrange"] else [3, 4])
    S, rows = ol.v1_blocks(qn.tolist(), m["trc"])
    assert f["hfb_size"] == S
    assert np.array_equal(f["blocksizes"],
                          np.array([[len(rows[(P, it)]) for P in (1, 2)] for it in (1, 2)]))
    if m["pairing_type"] == 2:
        assert int(f["pairing_type"][0]) == 2
        assert np.array_equal(f["fermi"], g["pairing"][-1, :2])
        assert f["cutoffs"].size == M and np.all((f["cutoffs"] >= 0) & (f["cutoffs"] <= 1))
        # the HFB density matrices are real here, rho symmetric, kappa
        # antisymmetric, and the traces of rho over the blocks of an
        # isospin count its particles
        assert np.abs(f["rho_hfb"].imag).max() == 0 and np.abs(f["kappa_hfb"].imag).max() == 0
        tr = np.zeros(2)
        for (P, it), r in rows.items():
            rb = f["rho_hfb"][(it - 1) * 2 + P - 1].real
            kb = f["kappa_hfb"][(it - 1) * 2 + P - 1].real
            assert np.abs(rb - rb.T).max() < 1e-14 and np.abs(kb + kb.T).max() < 1e-14
            tr[it - 1] += np.trace(rb)
        assert np.allclose(tr, [m["neutrons"], m["protons"]], rtol=0, atol=1e-6)
    elif m["pairing_type"] == 1:
        # GuessHFBMatrices(1) (HFB.f90:3719-3744): row i of the first half of
        # a block with row i + size/2, kappa = +-sqrt(o (1 - o)), o = half
        # the occupation; the rest zero
        assert int(f["pairing_type"][0]) == 1 and f["cutoffs"].size == M
        assert np.array_equal(f["fermi"], g["pairing"][-1, :2])
        K = np.zeros((4, S, S), np.complex128)
        for (P, it), r in rows.items():
            half = len(r) // 2
            for a in range(half):
                o = 0.5 * f["state_occupation"][r[a] % M]
                K[(it - 1) * 2 + P - 1, a + half, a] = np.sqrt(o * (1.0 - o))
                K[(it - 1) * 2 + P - 1, a, a + half] = -np.sqrt(o * (1.0 - o))
        assert np.array_equal(f["kappa_hfb"], K)
        assert not f["rho_hfb"].any() and not f["U"].any() and not f["hfb_columns"].any()
    else:
        assert int(f["pairing_type"][0]) == 0 and f["cutoffs"].size == 0
        assert list(f["fermi"]) == [-10.0, -10.0]
        K = ol.v1_hf_kappa_guess(qn.tolist(), m["trc"]).reshape(4, S, S)
        assert np.array_equal(f["kappa_hfb"], K)
        assert not f["rho_hfb"].any() and not f["U"].any() and not f["can_transfo"].any()
    if m.get("constraints"):
        assert [(x["l"], x["m"], x["isoswitch"], x["constraint_type"]) for x in f["moments"]] \
            == [(2, 0, 2, 2), (2, 2, 1, 1)]
        assert list(f["moments"][0]["true_constraint"]) == [9.0, 7.0]
    else:
        assert f["moments"] == []


@pytest.mark.parametrize("case", V1_CASES)
def test_library_reads_and_writes_reference_files_byte_for_byte(case, tmp_path):
    g = golden(case)
    raw = g["v1_file"].tobytes()
    src = tmp_path / "ref.v1"
    src.write_bytes(raw)
    info, a = mocca_amd.v1_read(str(src))
    f = ol.v1_parse(raw)
    assert info["nwt"] == int(f["nwt"][0]) and info["trc"] == int(f["sym"][0])
    assert info["energy"] == f["energy"] and info["denergy"] == list(f["denergy"])
    assert np.array_equal(a["grid"], f["grid"])
    assert np.array_equal(a["state"][:, 3], f["state_norm"])
    assert np.array_equal(a["state"][:, 4:], f["state_J"])
    assert np.array_equal(a["qn"], f["state_qn"])
    assert np.array_equal(a["rho"].reshape(-1), f["dens_rec"][0][:a["rho"].size])
    assert np.array_equal(a["kappa_hfb"].reshape(-1), f["kappa_hfb"].reshape(-1))
    assert np.array_equal(a["hfb_columns"].reshape(-1), f["hfb_columns"].reshape(-1))
    assert info["fermi"] == list(f["fermi"]) and info["B"] == list(f["B"])
    assert len(info["moments"]) == len(f["moments"])
    for x, y in zip(info["moments"], f["moments"]):
        assert x["old_value"] == list(y["old_value"]) and x["value"] == list(y["value"])
    out = tmp_path / "out.v1"
    mocca_amd.v1_write(str(out), info, a)
    assert out.read_bytes() == raw


def test_library_refuses_malformed_files(tmp_path):
    raw = bytearray(golden("scf_trc_fd")["v1_file"].tobytes())
    p = tmp_path / "bad.v1"
    p.write_bytes(bytes(raw[:len(raw) // 2]))              # truncated
    with pytest.raises(mocca_amd.MoccaError, match="not a MOCCa v1 file"):
        mocca_amd.v1_read(str(p))
    bad = bytearray(raw)
    bad[4:8] = np.array([2], "<i4").tobytes()              # version 2
    p.write_bytes(bytes(bad))
    with pytest.raises(mocca_amd.MoccaError, match="version"):
        mocca_amd.v1_read(str(p))
    bad = bytearray(raw)
    bad[-4:] = np.array([7], "<i4").tobytes()              # trailing count
    p.write_bytes(bytes(bad))
    with pytest.raises(mocca_amd.MoccaError, match="not a MOCCa v1 file"):
        mocca_amd.v1_read(str(p))
    with pytest.raises(mocca_amd.MoccaError, match="cannot open"):
        mocca_amd.v1_read(str(tmp_path / "missing.v1"))
