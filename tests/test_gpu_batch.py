"""mocca_batch_run: the Evolve iterations of several independent solves
(uneven meshes and state counts, alternating Q20 constraints) issued in turn
on their own streams.  Each solve's energy rows must equal, bit for bit, those
of the same solve iterated alone with mocca_iterate: the batch changes only
the order in which the host issues the solves' work."""
import numpy as np
import pytest

import mocca_amd
from mocca_amd import workloads

pytestmark = pytest.mark.gpu


def _start(w, precision):
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"], precision=precision)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    s.scf_setup(**w["functional"])
    s.set_constraints(w["constraints"])
    s.gram_schmidt()
    s.scf_start()
    return s


@pytest.mark.parametrize("precision", [mocca_amd.FP64, mocca_amd.MIXED])
def test_batch_matches_solves_run_alone(precision):
    ws = workloads.batch_scan(seed=7, nsolves=4, nmin=12, nmax=18, mmin=20, mmax=40)
    niter = 5
    alone = []
    for w in ws:
        s = _start(w, precision)
        alone.append(s.iterate(niter, w["dt_over_hbar"], w["momentum"]))
        s.close()
    solves = [_start(w, precision) for w in ws]
    rows = mocca_amd.batch_run(solves, niter, [w["dt_over_hbar"] for w in ws],
                               [w["momentum"] for w in ws])
    for s in solves:
        s.close()
    for r, a in zip(rows, alone):
        assert np.all(np.isfinite(r))
        assert np.array_equal(r, a)
    # the constraints pulled the solves apart: different quadrupole targets
    assert len({round(r[-1, 0], 6) for r in rows}) == len(rows)


def test_batch_reports_the_failing_solve():
    ws = workloads.batch_scan(seed=8, nsolves=2, nmin=12, nmax=12, mmin=20, mmax=20)
    solves = [_start(w, mocca_amd.FP64) for w in ws]
    with pytest.raises(mocca_amd.MoccaError):
        mocca_amd.batch_run(solves, -1, 0.001, 0.3)
    for s in solves:
        s.close()
