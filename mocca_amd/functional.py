"""Skyrme functionals for the self-consistent loop.

``edf_coefficients`` maps Skyrme parameters (t0..t3, x0..x3, the density
exponent and the spin-orbit strengths) to the coefficients B1..B21 of the
energy density functional, with the relations of CalcEDFCoef
(Force.f90:467-552) for forces without tensor terms.  ``scf_arguments``
gives the keyword arguments of :meth:`mocca_amd.Solve.scf_setup`, with the
physical constants of Force.f90:103-105 (hbm doubled and masses averaged as
ReadForceInfo does for these forces, Force.f90:242-244).
"""
import numpy as np

E2 = 1.43996446                  # e^2 (MeV fm)
HBAR = 6.58211928                # hbar (MeV 10^-22 s)
HBM_HALF = 20.73551910           # hbar^2 / 2m (MeV fm^2)
NUCLEON_MASS = (939.565379, 938.272046)

# Parameter sets (forces.param); t in MeV fm^{3(1+alpha)}, w in MeV fm^5.
FORCES = {
    "SLy4": dict(t0=-2488.913, t1=486.818, t2=-546.395, t3a=13777.0, t3b=0.0,
                 x0=0.834, x1=-0.344, x2=-1.0, x3a=1.354, x3b=0.0,
                 yt3a=0.166666666666666666667, yt3b=0.0, wso=123.0, wsoq=123.0,
                 com1body=2),
}


def edf_coefficients(p, time_reversal):
    """B[28] = B1, B2, B3, B4, B5, B6, B7a, B7b, B8a, B8b, byt3a, byt3b, B9,
    B9q, B10, B11, B12a, B12b, B13a, B13b, B14 .. B21 (tensor terms zero)."""
    t0, t1, t2 = p["t0"], p["t1"], p["t2"]
    x0, x1, x2 = p["x0"], p["x1"], p["x2"]
    t3a, t3b, x3a, x3b = p["t3a"], p.get("t3b", 0.0), p["x3a"], p.get("x3b", 0.0)
    B = np.zeros(28)
    B[0] = t0 * (1.0 + 1 / 2.0 * x0) / 2.0
    B[1] = -t0 * (1 / 2.0 + x0) / 2.0
    B[2] = (t1 * (1 + 1 / 2.0 * x1) + t2 * (1 + 1 / 2.0 * x2)) / 4.0
    B[3] = -(t1 * (1 / 2.0 + x1) - t2 * (1 / 2.0 + x2)) / 4.0
    B[4] = -(3.0 * t1 * (1 + x1 / 2.0) - t2 * (1 + x2 / 2.0)) / 16.0
    B[5] = (3.0 * t1 * (x1 + 1 / 2.0) + t2 * (x2 + 1 / 2.0)) / 16.0
    B[6] = t3a * (1 + x3a / 2.0) / 12.0
    B[7] = t3b * (1 + x3b / 2.0) / 12.0
    B[8] = -t3a * (x3a + 1 / 2.0) / 12.0
    B[9] = -t3b * (x3b + 1 / 2.0) / 12.0
    B[10] = p["yt3a"]
    B[11] = p.get("yt3b", 0.0)
    B[12] = -p["wso"] / 2.0
    B[13] = -p["wsoq"] / 2.0
    if not time_reversal:
        B[14] = t0 * x0 / 4.0
        B[15] = -t0 / 4.0
        B[16] = t3a * x3a / 24.0
        B[17] = t3b * x3b / 24.0
        B[18] = -t3a / 24.0
        B[19] = -t3b / 24.0
    return B


def scf_arguments(name, neutrons, protons, time_reversal, **overrides):
    """Keyword arguments of Solve.scf_setup for a named force (Coulomb by the
    direct solve unless overridden with coulomb=1, the reference's CG)."""
    p = FORCES[name]
    m = 0.5 * (NUCLEON_MASS[0] + NUCLEON_MASS[1])
    args = dict(B=edf_coefficients(p, time_reversal), hbm=(2 * HBM_HALF, 2 * HBM_HALF),
                mass=(m, m), e2=E2, neutrons=neutrons, protons=protons,
                com1body=p["com1body"], cexchange=2, coulomb=2, coulomb_maxit=500,
                max_moment=10, damping=0.75)
    args.update(overrides)
    return args
