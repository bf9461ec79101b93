"""FP64 CPU oracle for the dipole + quadrupole (8 unknowns per particle)
spectral-Ewald operator of Capacitance_Quadrupole/.

TEST INFRASTRUCTURE ONLY (see ewald_oracle.py's header): only tests/ may use
it, as the test reference.

A NumPy restatement of the reference's MATLAB, same data flow: spread of the
3 dipole and 5 quadrupole components with the Gaussian kernel of the dipole
method, complex 3D FFTs, the four k-space blocks of Scale.m, contraction,
and the real-space sum of RealSpace.m with its nine tables.  k vectors follow
MatrixVectorMultiply.m:42-46 (fftshift order, k = -ceil((N-1)/2) ..
floor((N-1)/2)); here the FFT output stays in natural order and every bin
carries the same k (numpy's fftfreq), which is the same elementwise map.

Pinned by (tests/test_quad_oracle.py): the shared dipole tables are
identical to the pinned ones of ewald_oracle (Quad/RealSpaceTable.m repeats
Het's expressions); the eight new tables by xi-independence of the total
operator (real + wave space, all four blocks), their bare multipole limits at
small xi, continuity at r = 2 and their r -> 0 limits against the self terms;
the k-space blocks by the same xi-independence and by the Hermitian identity
for real input; the dipole block against ewald_oracle's operator; the
quadrupole diagonal (dielectric term 15/(8 pi (eps-1)), RealSpace.m:29, with
the self terms) by the l = 2 Clausius-Mossotti response of a lone sphere to
an imposed field gradient, Q = 4 pi/3 (eps-1)/(eps+3/2) g, at four
permittivities.  The
reference's MATLAB cannot be run here (no MATLAB/Octave), so there are no
golden vectors for this operator.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres
from scipy.special import erfc as _erfc
from scipy.special import spherical_jn as _sjn

from . import ewald_oracle as eo

PI = np.pi
SP = np.sqrt(np.pi)

TABLE_NAMES = ("pot_charge", "pot_dip", "pot_quad", "field_dip_1", "field_dip_2", "field_dip_force_1",
               "field_dip_force_2", "field_quad_1", "field_quad_2", "field_quad_3", "grad_quad_1",
               "grad_quad_2", "grad_quad_3", "grad_quad_4")
OPERATOR_TABLES = ("field_dip_1", "field_dip_2", "field_quad_1", "field_quad_2", "field_quad_3",
                   "grad_quad_1", "grad_quad_2", "grad_quad_3", "grad_quad_4")


def _combine(r, x, ep, em, e0, fp, fm, f0, reg, reg_on=True):
    return (ep * np.exp(-(r + 2) ** 2 * x * x) + em * np.exp(-(r - 2) ** 2 * x * x) + e0 * np.exp(-r * r * x * x)
            + fp * _erfc((r + 2) * x) + fm * _erfc((r - 2) * x) + f0 * _erfc(r * x) + (r < 2) * reg * reg_on)


def _quad_tables(r, x, regularize=True):
    """The eight couplings Quad/RealSpaceTable.m adds to the dipole ones
    (:66-85, :173-356), written out as in the reference; regularize=False
    drops the (r < 2) overlap terms, as Compute_Field_Quadrupole/
    RealSpaceTable.m does (its regpoly lines are commented out)."""
    def comb(*args):
        return _combine(*args, reg_on=regularize)

    p32 = PI ** 1.5
    out = {}
    # potential / quadrupole (unused by the operator; the reference notes its regulariser is wrong)
    out["pot_quad"] = comb(
        r, x,
        3 / (1024 * p32 * x ** 5 * r ** 3) * (20 * x ** 4 * r ** 5 + 24 * x ** 4 * r ** 4 + 2 * x ** 2 * (10 - 4 * x ** 2) * r ** 3
                                            - 16 * x ** 2 * (2 - x ** 2) * r ** 2 - (15 - 12 * x ** 2 + 32 * x ** 4) * r
                                            + 2 * (9 - 4 * x ** 2 + 32 * x ** 4)),
        3 / (1024 * p32 * x ** 5 * r ** 3) * (20 * x ** 4 * r ** 5 - 24 * x ** 4 * r ** 4 + 2 * x ** 2 * (10 - 4 * x ** 2) * r ** 3
                                            + 16 * x ** 2 * (2 - x ** 2) * r ** 2 - (15 - 12 * x ** 2 + 32 * x ** 4) * r
                                            - 2 * (9 - 4 * x ** 2 + 32 * x ** 4)),
        15 / (512 * p32 * x ** 5 * r ** 2) * (4 * x ** 4 * r ** 4 + 4 * x ** 2 * (1 + 2 * x ** 2) * r ** 2 - 3 * (1 + 4 * x ** 2)),
        3 / (2048 * PI * x ** 6 * r ** 3) * (40 * x ** 6 * r ** 6 + 128 * x ** 6 * r ** 5 + 20 * x ** 4 * (3 + 4 * x ** 2) * r ** 4
                                           - 10 * x ** 2 * (3 + 8 * x ** 2) * r ** 2 + 15 + 60 * x ** 2 + 256 * x ** 6),
        3 / (2048 * PI * x ** 6 * r ** 3) * (40 * x ** 6 * r ** 6 - 128 * x ** 6 * r ** 5 + 20 * x ** 4 * (3 + 4 * x ** 2) * r ** 4
                                           - 10 * x ** 2 * (3 + 8 * x ** 2) * r ** 2 + 15 + 60 * x ** 2 + 256 * x ** 6),
        15 / (1024 * PI * x ** 6 * r ** 3) * (-8 * x ** 6 * r ** 6 - 4 * x ** 4 * (3 + 4 * x ** 2) * r ** 4
                                            - 2 * x ** 2 * (3 + 8 * x ** 2) * r ** 2 - 3 * (1 + 4 * x ** 2)),
        15 / (1024 * PI * x ** 6 * r ** 3) * (12 * x ** 4 * r ** 4 - 2 * x ** 2 * (3 + 8 * x ** 2) * r ** 2 + 3 + 12 * x ** 2))
    # field / quadrupole (= field gradient / dipole), rrr.S component
    out["field_quad_1"] = comb(
        r, x,
        15 / (16384 * p32 * x ** 7 * r ** 4) * (-24 * x ** 6 * r ** 7 + 48 * x ** 6 * r ** 6 - 4 * x ** 4 * (9 - 8 * x ** 2) * r ** 5
                                              + 8 * x ** 4 * (3 - 8 * x ** 2) * r ** 4
                                              + 2 * x ** 2 * (21 - 80 * x ** 2 + 64 * x ** 4) * r ** 3
                                              - 4 * x ** 2 * (3 - 8 * x ** 2) ** 2 * r ** 2
                                              - (45 - 120 * x ** 2 + 192 * x ** 4 - 512 * x ** 6) * r
                                              - 2 * (45 + 24 * x ** 2 - 64 * x ** 4 + 512 * x ** 6)),
        15 / (16384 * p32 * x ** 7 * r ** 4) * (-24 * x ** 6 * r ** 7 - 48 * x ** 6 * r ** 6 - 4 * x ** 4 * (9 - 8 * x ** 2) * r ** 5
                                              - 8 * x ** 4 * (3 - 8 * x ** 2) * r ** 4
                                              + 2 * x ** 2 * (21 - 80 * x ** 2 + 64 * x ** 4) * r ** 3
                                              + 4 * x ** 2 * (3 - 8 * x ** 2) ** 2 * r ** 2
                                              - (45 - 120 * x ** 2 + 192 * x ** 4 - 512 * x ** 6) * r
                                              + 2 * (45 + 24 * x ** 2 - 64 * x ** 4 + 512 * x ** 6)),
        15 / (8192 * p32 * x ** 7 * r ** 3) * (24 * x ** 6 * r ** 6 + 4 * x ** 4 * (9 - 32 * x ** 2) * r ** 4
                                             - 2 * x ** 2 * (21 - 128 * x ** 2) * r ** 2 + 45 - 480 * x ** 2),
        15 / (32768 * PI * x ** 8 * r ** 4) * (48 * x ** 8 * r ** 8 + 32 * x ** 6 * (3 - 8 * x ** 2) * r ** 6
                                             - 24 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 + 72 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                             - 45 + 480 * x ** 2 + 4096 * x ** 8),
        15 / (32768 * PI * x ** 8 * r ** 4) * (48 * x ** 8 * r ** 8 + 32 * x ** 6 * (3 - 8 * x ** 2) * r ** 6
                                             - 24 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 + 72 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                             - 45 + 480 * x ** 2 + 4096 * x ** 8),
        15 / (16384 * PI * x ** 8 * r ** 4) * (-48 * x ** 8 * r ** 8 - 32 * x ** 6 * (3 - 8 * x ** 2) * r ** 6
                                             + 24 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 - 72 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                             + 45 - 480 * x ** 2),
        -15 / (4 * PI * r ** 4) + 15 * r ** 2 / (64 * PI) * (1 - 3 * r ** 2 / 16))
    # Ir.S + Sr + rS component
    out["field_quad_2"] = comb(
        r, x,
        3 / (16384 * p32 * x ** 7 * r ** 4) * (-40 * x ** 6 * r ** 7 + 80 * x ** 6 * r ** 6 - 20 * x ** 4 * (11 - 24 * x ** 2) * r ** 5
                                             + 8 * x ** 4 * (45 + 8 * x ** 2) * r ** 4
                                             - 2 * x ** 2 * (45 - 80 * x ** 2 + 64 * x ** 4) * r ** 3
                                             - 4 * x ** 2 * (15 + 48 * x ** 2 - 64 * x ** 4) * r ** 2
                                             + (45 - 120 * x ** 2 + 192 * x ** 4 - 512 * x ** 6) * r
                                             + 2 * (45 + 24 * x ** 2 - 64 * x ** 4 + 512 * x ** 6)),
        3 / (16384 * p32 * x ** 7 * r ** 4) * (-40 * x ** 6 * r ** 7 - 80 * x ** 6 * r ** 6 - 20 * x ** 4 * (11 - 24 * x ** 2) * r ** 5
                                             - 8 * x ** 4 * (45 + 8 * x ** 2) * r ** 4
                                             - 2 * x ** 2 * (45 - 80 * x ** 2 + 64 * x ** 4) * r ** 3
                                             + 4 * x ** 2 * (15 + 48 * x ** 2 - 64 * x ** 4) * r ** 2
                                             + (45 - 120 * x ** 2 + 192 * x ** 4 - 512 * x ** 6) * r
                                             - 2 * (45 + 24 * x ** 2 - 64 * x ** 4 + 512 * x ** 6)),
        15 / (8192 * p32 * x ** 7 * r ** 3) * (8 * x ** 6 * r ** 6 + 4 * x ** 4 * (11 - 32 * x ** 2) * r ** 4
                                             + 2 * x ** 2 * (9 - 64 * x ** 2) * r ** 2 - 9 + 96 * x ** 2),
        3 / (32768 * PI * x ** 8 * r ** 4) * (80 * x ** 8 * r ** 8 + 160 * x ** 6 * (3 - 8 * x ** 2) * r ** 6 - 2048 * x ** 8 * r ** 5
                                            + 120 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 - 120 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                            + 45 - 480 * x ** 2 - 4096 * x ** 8),
        3 / (32768 * PI * x ** 8 * r ** 4) * (80 * x ** 8 * r ** 8 + 160 * x ** 6 * (3 - 8 * x ** 2) * r ** 6 + 2048 * x ** 8 * r ** 5
                                            + 120 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 - 120 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                            + 45 - 480 * x ** 2 - 4096 * x ** 8),
        15 / (16384 * PI * x ** 8 * r ** 4) * (-16 * x ** 8 * r ** 8 - 32 * x ** 6 * (3 - 8 * x ** 2) * r ** 6
                                             - 24 * x ** 4 * (3 - 16 * x ** 2) * r ** 4 + 24 * x ** 2 * (1 - 8 * x ** 2) * r ** 2
                                             - 9 + 96 * x ** 2),
        3 / (4 * PI * r ** 4) - 3 * r / (8 * PI) * (1 - 5 * r / 8 + 5 * r ** 3 / 128))
    # Ir.S component
    out["field_quad_3"] = comb(
        r, x,
        5 / (1024 * p32 * x ** 5 * r ** 2) * (4 * x ** 4 * r ** 5 - 8 * x ** 4 * r ** 4 + 16 * x ** 2 * (1 - 2 * x ** 2) * r ** 3
                                            - 24 * x ** 2 * r ** 2 + 3 * r + 6),
        5 / (1024 * p32 * x ** 5 * r ** 2) * (4 * x ** 4 * r ** 5 + 8 * x ** 4 * r ** 4 + 16 * x ** 2 * (1 - 2 * x ** 2) * r ** 3
                                            + 24 * x ** 2 * r ** 2 + 3 * r - 6),
        5 / (512 * p32 * x ** 5 * r) * (-4 * x ** 4 * r ** 4 - 16 * x ** 2 * (1 - 3 * x ** 2) * r ** 2 - 3 + 24 * x ** 2),
        5 / (2048 * PI * x ** 6 * r ** 2) * (-8 * x ** 6 * r ** 6 - 12 * x ** 4 * (3 - 8 * x ** 2) * r ** 4 + 128 * x ** 6 * r ** 3
                                           - 6 * x ** 2 * (3 - 16 * x ** 2) * r ** 2 + 3 - 24 * x ** 2),
        5 / (2048 * PI * x ** 6 * r ** 2) * (-8 * x ** 6 * r ** 6 - 12 * x ** 4 * (3 - 8 * x ** 2) * r ** 4 - 128 * x ** 6 * r ** 3
                                           - 6 * x ** 2 * (3 - 16 * x ** 2) * r ** 2 + 3 - 24 * x ** 2),
        5 / (1024 * PI * x ** 6 * r ** 2) * (8 * x ** 6 * r ** 6 + 12 * x ** 4 * (3 - 8 * x ** 2) * r ** 4
                                           + 6 * x ** 2 * (3 - 16 * x ** 2) * r ** 2 - 3 + 24 * x ** 2),
        5 * r / (8 * PI) * (1 - 3 * r / 4 + r ** 3 / 16))
    # field gradient / quadrupole: Irr:Q component
    x2, x4, x6, x8, x10 = x ** 2, x ** 4, x ** 6, x ** 8, x ** 10
    gq1_f = (16 * x8 * r ** 8 + (96 * x6 - 128 * x8) * r ** 6 + (72 * x4 - 192 * x6 + 256 * x8) * r ** 4
             - (24 * x2 - 96 * x4 + 256 * x6) * r ** 2 + 9 - 48 * x2 + 192 * x4)
    out["grad_quad_1"] = comb(
        r, x,
        75 / (16384 * p32 * x ** 7 * r ** 3) * (8 * x6 * r ** 7 - 16 * x6 * r ** 6 + (44 * x4 - 32 * x6) * r ** 5
                                              - (72 * x4 - 64 * x6) * r ** 4 + (18 * x2 + 32 * x4) * r ** 3
                                              + (12 * x2 - 64 * x4) * r ** 2 - (9 + 24 * x2) * r - (18 - 48 * x2)),
        75 / (16384 * p32 * x ** 7 * r ** 3) * (8 * x6 * r ** 7 + 16 * x6 * r ** 6 + (44 * x4 - 32 * x6) * r ** 5
                                              + (72 * x4 - 64 * x6) * r ** 4 + (18 * x2 + 32 * x4) * r ** 3
                                              - (12 * x2 - 64 * x4) * r ** 2 - (9 + 24 * x2) * r + (18 - 48 * x2)),
        -75 / (8192 * p32 * x ** 7 * r ** 2) * (8 * x6 * r ** 6 + (44 * x4 - 64 * x6) * r ** 4
                                              + (18 * x2 - 64 * x4 + 128 * x6) * r ** 2 - 9 + 48 * x2 - 192 * x4),
        -75 / (32768 * PI * x8 * r ** 3) * gq1_f,
        -75 / (32768 * PI * x8 * r ** 3) * gq1_f,
        75 / (16384 * PI * x8 * r ** 3) * gq1_f,
        75 * r * (4 - r ** 2) ** 2 / (1024 * PI))
    # Q component
    gq2_f = (288 * x10 * r ** 10 + (3600 * x8 - 4800 * x10) * r ** 8 + (10800 * x6 - 28800 * x8 + 38400 * x10) * r ** 6
             + (5400 * x4 - 21600 * x6 + 57600 * x8) * r ** 4 - (1350 * x2 - 7200 * x4 + 28800 * x6) * r ** 2
             + 405 - 2700 * x2 + 14400 * x4)
    out["grad_quad_2"] = comb(
        r, x,
        3 / (32768 * p32 * x ** 9 * r ** 5) * (-48 * x8 * r ** 9 + 96 * x8 * r ** 8 - (576 * x6 - 608 * x8) * r ** 7
                                             + (1056 * x6 - 1216 * x8) * r ** 6
                                             - (1536 * x4 - 2576 * x6 + 3968 * x8) * r ** 5
                                             + (2160 * x4 - 4320 * x6 - 256 * x8) * r ** 4
                                             - (360 * x2 + 360 * x4 + 640 * x6 - 512 * x8) * r ** 3
                                             - (360 * x2 - 1680 * x4 - 768 * x6 + 1024 * x8) * r ** 2
                                             + (135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8) * r
                                             + (270 - 1080 * x2 - 192 * x4 + 512 * x6 - 4096 * x8)),
        3 / (32768 * p32 * x ** 9 * r ** 5) * (-48 * x8 * r ** 9 - 96 * x8 * r ** 8 - (576 * x6 - 608 * x8) * r ** 7
                                             - (1056 * x6 - 1216 * x8) * r ** 6
                                             - (1536 * x4 - 2576 * x6 + 3968 * x8) * r ** 5
                                             - (2160 * x4 - 4320 * x6 - 256 * x8) * r ** 4
                                             - (360 * x2 + 360 * x4 + 640 * x6 - 512 * x8) * r ** 3
                                             + (360 * x2 - 1680 * x4 - 768 * x6 + 1024 * x8) * r ** 2
                                             + (135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8) * r
                                             - (270 - 1080 * x2 - 192 * x4 + 512 * x6 - 4096 * x8)),
        1 / (16384 * p32 * x ** 9 * r ** 4) * (144 * x8 * r ** 8 + (1728 * x6 - 2400 * x8) * r ** 6
                                             + (4608 * x4 - 13200 * x6 + 19200 * x8) * r ** 4
                                             + (1080 * x2 - 5400 * x4 + 19200 * x6) * r ** 2 - 405 + 2700 * x2 - 14400 * x4),
        1 / (65536 * PI * x10 * r ** 5) * (gq2_f + 49152 * x10 * r ** 5 + 49152 * x10),
        1 / (65536 * PI * x10 * r ** 5) * (gq2_f - 49152 * x10 * r ** 5 + 49152 * x10),
        -1 / (32768 * PI * x10 * r ** 5) * gq2_f,
        -3 / (2 * PI) * (1 / r ** 5 - 1 + 25 * r / 32 - 25 * r ** 3 / 256 + 3 * r ** 5 / 512))
    # Irr:Q + 2 Q.rr + 2 rr.Q component
    gq3_f = (-96 * x10 * r ** 10 + (-720 * x8 + 960 * x10) * r ** 8 - (720 * x6 - 1920 * x8 + 2560 * x10) * r ** 6
             + (360 * x4 - 1440 * x6 + 3840 * x8) * r ** 4 - (270 * x2 - 1440 * x4 + 5760 * x6) * r ** 2
             + 135 - 900 * x2 + 4800 * x4)
    out["grad_quad_3"] = comb(
        r, x,
        -15 / (65536 * p32 * x ** 9 * r ** 5) * (48 * x8 * r ** 9 - 96 * x8 * r ** 8 + (336 * x6 - 288 * x8) * r ** 7
                                               - 576 * (x6 - x8) * r ** 6 + (216 * x4 + 144 * x6 + 128 * x8) * r ** 5
                                               - (480 * x6 + 256 * x8) * r ** 4
                                               - (180 * x2 - 120 * x4 + 640 * x6 - 512 * x8) * r ** 3
                                               + (720 * x4 + 768 * x6 - 1024 * x8) * r ** 2
                                               + (135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8) * r
                                               - (-270 + 1080 * x2 + 192 * x4 - 512 * x6 + 4096 * x8)),
        -15 / (65536 * p32 * x ** 9 * r ** 5) * (48 * x8 * r ** 9 + 96 * x8 * r ** 8 + (336 * x6 - 288 * x8) * r ** 7
                                               + 576 * (x6 - x8) * r ** 6 + (216 * x4 + 144 * x6 + 128 * x8) * r ** 5
                                               + (480 * x6 + 256 * x8) * r ** 4
                                               - (180 * x2 - 120 * x4 + 640 * x6 - 512 * x8) * r ** 3
                                               - (720 * x4 + 768 * x6 - 1024 * x8) * r ** 2
                                               + (135 + 180 * x2 + 480 * x4 - 768 * x6 + 2048 * x8) * r
                                               + (-270 + 1080 * x2 + 192 * x4 - 512 * x6 + 4096 * x8)),
        15 / (32768 * p32 * x ** 9 * r ** 4) * (48 * x8 * r ** 8 + (336 * x6 - 480 * x8) * r ** 6
                                              + (216 * x4 - 720 * x6 + 1280 * x8) * r ** 4
                                              - (180 * x2 - 840 * x4 + 2560 * x6) * r ** 2 + 135 - 900 * x2 + 4800 * x4),
        -15 / (131072 * PI * x10 * r ** 5) * (gq3_f + 16384 * x10),
        -15 / (131072 * PI * x10 * r ** 5) * (gq3_f + 16384 * x10),
        15 / (65536 * PI * x10 * r ** 5) * gq3_f,
        -15 * (r ** 2 - 4) ** 3 * (3 * r ** 4 + 6 * r ** 2 + 8) / (2048 * PI * r ** 5))
    # rrrr:Q component
    gq4_f = (288 * x10 * r ** 10 + (720 * x8 - 960 * x10) * r ** 8 - (720 * x6 - 1920 * x8 + 2560 * x10) * r ** 6
             + (1080 * x4 - 4320 * x6 + 11520 * x8) * r ** 4 - (1350 * x2 - 7200 * x4 + 28800 * x6) * r ** 2
             + 945 - 6300 * x2 + 33600 * x4)
    out["grad_quad_4"] = comb(
        r, x,
        -15 / (65536 * p32 * x ** 9 * r ** 5) * (144 * x8 * r ** 9 - 288 * x8 * r ** 8 + (288 * x6 + 96 * x8) * r ** 7
                                               - (288 * x6 + 192 * x8) * r ** 6
                                               - (432 * x4 - 912 * x6 + 896 * x8) * r ** 5
                                               + (720 * x4 - 480 * x6 + 1792 * x8) * r ** 4
                                               + (720 * x2 - 2280 * x4 + 4480 * x6 - 3584 * x8) * r ** 3
                                               - (1080 * x2 + 2160 * x4 + 5376 * x6 - 7168 * x8) * r ** 2
                                               - (945 + 1260 * x2 + 3360 * x4 - 5376 * x6 + 14336 * x8) * r
                                               - (1890 - 7560 * x2 - 1344 * x4 + 3584 * x6 - 28672 * x8)),
        -15 / (65536 * p32 * x ** 9 * r ** 5) * (144 * x8 * r ** 9 + 288 * x8 * r ** 8 + (288 * x6 + 96 * x8) * r ** 7
                                               + (288 * x6 + 192 * x8) * r ** 6
                                               - (432 * x4 - 912 * x6 + 896 * x8) * r ** 5
                                               - (720 * x4 - 480 * x6 + 1792 * x8) * r ** 4
                                               + (720 * x2 - 2280 * x4 + 4480 * x6 - 3584 * x8) * r ** 3
                                               + (1080 * x2 + 2160 * x4 + 5376 * x6 - 7168 * x8) * r ** 2
                                               - (945 + 1260 * x2 + 3360 * x4 - 5376 * x6 + 14336 * x8) * r
                                               + (1890 - 7560 * x2 - 1344 * x4 + 3584 * x6 - 28672 * x8)),
        15 / (32768 * p32 * x ** 9 * r ** 4) * (144 * x8 * r ** 8 + (288 * x6 - 480 * x8) * r ** 6
                                              - (432 * x4 - 1200 * x6 + 1280 * x8) * r ** 4
                                              + (720 * x2 - 3000 * x4 + 6400 * x6) * r ** 2 - 945 + 6300 * x2 - 33600 * x4),
        15 / (131072 * PI * x10 * r ** 5) * (gq4_f + 114688 * x10),
        15 / (131072 * PI * x10 * r ** 5) * (gq4_f + 114688 * x10),
        -15 / (65536 * PI * x10 * r ** 5) * gq4_f,
        -15 * (3584 - 80 * r ** 6 - 30 * r ** 8 + 9 * r ** 10) / (2048 * PI * r ** 5))
    return out


def grad_quad_2_self(x):
    """Self term of the Q coupling (Quad/RealSpaceTable.m:372)."""
    return 1 / (2 * PI) * (3 * (3 - 10 * x ** 2 + 20 * x ** 4) / (8 * SP * x ** 5)
                           - 3 * (3 + 2 * x ** 2 + 4 * x ** 4) * np.exp(-4 * x ** 2) / (8 * SP * x ** 5)
                           + 3 * _erfc(2 * x))


def real_space_table(r, xi, regularize=True):
    """The 14 tables of Quad/RealSpaceTable.m in its output order, evaluated
    at every r given (r = 0 gives non-finite couplings, as in the
    reference); when r[0] == 0 the self terms are prepended (:358-378), so
    every table then has len(r) + 1 entries.  regularize=False: the tables
    of Compute_Field_Quadrupole/RealSpaceTable.m (no overlap terms, no self
    terms prepended)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    x = float(xi)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        rp = np.where(r > 0, r, 1.0)
        six = eo.real_space_table(rp, x, regularize=regularize)  # shared expressions (self first)
        base = dict(zip(eo.TABLE_NAMES, (t[1:] for t in six)))
        quad = _quad_tables(r, x, regularize)
    tabs = {}
    for name in TABLE_NAMES:
        v = base[name] if name in base else quad[name]
        v = np.where(r > 0, v, np.nan)
        tabs[name] = v
    if r.size and r[0] == 0.0 and regularize:
        e4 = np.exp(-4 * x * x)
        tail = _erfc(2 * x) / (4 * PI)
        self_fd = (-1 + 6 * x * x + (1 - 2 * x * x) * e4) / (16 * PI ** 1.5 * x ** 3) + tail
        selfs = {n: 0.0 for n in TABLE_NAMES}
        selfs["pot_charge"] = (1 - e4) / (8 * PI ** 1.5 * x) + tail
        selfs["field_dip_1"] = self_fd
        selfs["field_dip_2"] = self_fd
        selfs["grad_quad_2"] = grad_quad_2_self(x)
        for name in TABLE_NAMES:
            tabs[name] = np.concatenate(([selfs[name]], tabs[name]))
    return tuple(tabs[n] for n in TABLE_NAMES)


def operator_tables(r, xi):
    """(9 operator tables with self terms, rvals) on the MATLAB driver's grid
    for r > 0 values r: the same alignment as the dipole path (rvals[0] = 0
    pairs with the self entry)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    rv = np.concatenate(([0.0], r))
    t = dict(zip(TABLE_NAMES, real_space_table(rv, xi)))
    # entry 1 is the coupling at r = 0 (not finite); drop it so that tables align with rv
    tabs = np.stack([np.concatenate((t[n][:1], t[n][2:])) for n in OPERATOR_TABLES])
    return tabs, rv


# ---------------------------------------------------------------------------
# spread / scale / contract (Quad/Spread.m, Scale.m, Contract.m)
# ---------------------------------------------------------------------------
def spread(m8, Ngrid, linnode, w_spread):
    """(N,8) moments -> (Nx,Ny,Nz,8) grids; F-order linear node index."""
    G = int(np.prod(Ngrid))
    out = np.zeros((G, 8), dtype=np.complex128)
    for c in range(8):
        v = (w_spread * m8[:, c][:, None]).reshape(-1)
        out[:, c] = (np.bincount(linnode.reshape(-1), weights=np.real(v), minlength=G)
                     + 1j * np.bincount(linnode.reshape(-1), weights=np.imag(v), minlength=G))
    return out.reshape(tuple(Ngrid[::-1]) + (8,)).transpose(2, 1, 0, 3)


def kgrid(Ngrid, box):
    """k vectors of every FFT bin (natural order), (Nx,Ny,Nz,3)."""
    ks = [np.fft.fftfreq(int(n), d=1.0 / int(n)) * 2 * PI / L for n, L in zip(Ngrid, box)]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    return np.stack([KX, KY, KZ], axis=-1)


def scale(fH, k, xi, eta):
    """The four k-space blocks of Quad/Scale.m:18-72 on (Nx,Ny,Nz,8) spectra."""
    k2 = np.sum(k * k, axis=-1)
    kmag = np.sqrt(k2)
    zero = k2 == 0
    with np.errstate(divide="ignore", invalid="ignore"):
        khat = np.where(zero[..., None], 0.0, k / np.where(zero, 1.0, kmag)[..., None])
        u0, u1, u2 = khat[..., 0], khat[..., 1], khat[..., 2]
        qf = np.stack([u0 * u0 - 1 / 3, u0 * u1, u0 * u2, u1 * u1 - 1 / 3, u1 * u2], axis=-1)
        qd = np.stack([u0 * u0 - u2 * u2, 2 * u0 * u1, 2 * u0 * u2, u1 * u1 - u2 * u2, 2 * u1 * u2], axis=-1)
        etak2 = np.sum((1.0 - np.asarray(eta))[None, None, None, :] * k * k, axis=-1)
        j1 = _sjn(1, kmag)
        j2 = _sjn(2, kmag)
        e = np.exp(-etak2 / (4 * xi * xi)) / np.where(zero, 1.0, k2)
    cES = np.where(zero, 0.0, 9 * j1 * j1 * e)
    cEQ = np.where(zero, 0.0, -45 / 2 * 1j * j1 * j2 * e)
    cGS = np.where(zero, 0.0, 45 * 1j * j2 * j1 * e)
    cGQ = np.where(zero, 0.0, 225 / 2 * j2 * j2 * e)
    sS = np.sum(khat * fH[..., :3], axis=-1)
    sQ = np.sum(qd * fH[..., 3:], axis=-1)
    E = (cES * sS + cEQ * sQ)[..., None] * khat
    G = (cGS * sS + cGQ * sQ)[..., None] * qf
    return np.concatenate([E, G], axis=-1)


def contract(Ht, Ngrid, linnode, w_contract):
    G = int(np.prod(Ngrid))
    flat = Ht.transpose(2, 1, 0, 3).reshape(G, 8)
    return np.stack([np.sum(w_contract * flat[:, c][linnode], axis=1) for c in range(8)], axis=1)


# ---------------------------------------------------------------------------
# real space (Quad/RealSpace.m:27-100)
# ---------------------------------------------------------------------------
def _v5_sym(a, b):
    """[a1 b1, a1 b2, a1 b3, a2 b2, a2 b3] for (M,3) a, b."""
    return np.stack([a[:, 0] * b[:, 0], a[:, 0] * b[:, 1], a[:, 0] * b[:, 2], a[:, 1] * b[:, 1],
                     a[:, 1] * b[:, 2]], axis=1)


def pair_terms(S_j, Q_j, rh, coefs):
    """Field (M,3) and gradient (M,5) at p1 from the moments of p2 for unit
    separation vectors rh and interpolated coefficients A..I (9, M)."""
    A, B, C, D, E, F, G, H, I = coefs
    Sr = np.sum(S_j * rh, axis=1)
    r2v = np.stack([rh[:, 0] ** 2, rh[:, 0] * rh[:, 1], rh[:, 0] * rh[:, 2], rh[:, 1] ** 2, rh[:, 1] * rh[:, 2]],
                   axis=1)
    Ivec = np.array([1.0, 0.0, 0.0, 1.0, 0.0])
    rrrdotS = r2v * Sr[:, None]
    IrdotS = Ivec[None, :] * Sr[:, None]
    S_r = _v5_sym(S_j, rh)
    r_S = _v5_sym(rh, S_j)
    q1, q2, q3, q4, q5 = (Q_j[:, t] for t in range(5))
    Qdotr = np.stack([q1 * rh[:, 0] + q2 * rh[:, 1] + q3 * rh[:, 2],
                      q2 * rh[:, 0] + q4 * rh[:, 1] + q5 * rh[:, 2],
                      q3 * rh[:, 0] + q5 * rh[:, 1] - (q1 + q4) * rh[:, 2]], axis=1)
    qdd = ((rh[:, 0] ** 2 - rh[:, 2] ** 2) * q1 + 2 * rh[:, 0] * rh[:, 1] * q2 + 2 * rh[:, 0] * rh[:, 2] * q3
           + (rh[:, 1] ** 2 - rh[:, 2] ** 2) * q4 + 2 * rh[:, 1] * rh[:, 2] * q5)
    IrrQ = Ivec[None, :] * qdd[:, None]
    Qrr = _v5_sym(Qdotr, rh)
    rrQ = _v5_sym(rh, Qdotr)
    rrrrQ = r2v * qdd[:, None]
    Ef = (A[:, None] * (S_j - Sr[:, None] * rh) + (B * Sr)[:, None] * rh
          - 0.5 * ((C * qdd)[:, None] * rh + 2 * D[:, None] * Qdotr))
    Gf = (C[:, None] * rrrdotS + D[:, None] * (S_r + r_S) + (D + E)[:, None] * IrdotS
          + 0.5 * (F[:, None] * IrrQ + G[:, None] * Q_j + H[:, None] * (IrrQ + 2 * Qrr + 2 * rrQ)
                   + I[:, None] * rrrrQ))
    return Ef, Gf


def interpolate(d, tabs, rvals):
    """Linear interpolation between the first r_vals >= d and the entry
    before it (RealSpace.m:70-92); tabs (T, Nr)."""
    idx = np.searchsorted(rvals, d, side="left")
    if np.any(idx <= 0) or np.any(idx >= rvals.size):
        raise ValueError("pair separation outside the real-space table range")
    lo, hi = rvals[idx - 1], rvals[idx]
    wlo = (hi - d) / (hi - lo)
    whi = (d - lo) / (hi - lo)
    return tabs[:, idx - 1] * wlo + tabs[:, idx] * whi


class QuadGeometry:
    """Everything of MatrixVectorMultiply.m that depends only on positions."""

    def __init__(self, x, box, Ngrid, P, xi, rc, tabs, rvals):
        self.x = np.asarray(x, dtype=np.float64)
        self.box = np.asarray(box, dtype=np.float64)
        self.Ngrid = np.asarray(Ngrid, dtype=np.int64)
        self.h = self.box / self.Ngrid
        self.P = int(P)
        self.xi = float(xi)
        self.eta = self.P * (self.h * self.xi) ** 2 / PI
        self.rc = float(rc)
        self.tabs = np.asarray(tabs, dtype=np.float64)
        self.rvals = np.asarray(rvals, dtype=np.float64)
        offset, offsetxyz = eo.pre_calculations(self.P, self.h)
        self.linnode, self.w_spread, self.w_contract = eo.prepare_grid_kernel(
            self.x, self.Ngrid, self.h, self.xi, self.eta, self.P, offset, offsetxyz)
        cell, ncell = eo.cell_list(self.x, self.box, self.rc)
        p1, p2 = eo.neighbor_list(cell, ncell)
        r = self.x[p1] - self.x[p2]
        r = r - self.box * np.trunc(2 * r / self.box)
        d = np.sqrt(np.sum(r * r, axis=1))
        keep = d < self.rc
        self.p1, self.p2, d, r = p1[keep], p2[keep], d[keep], r[keep]
        self.rhat = r / d[:, None]
        self.coefs = interpolate(d, self.tabs, self.rvals)
        self.k = kgrid(self.Ngrid, self.box)


def field(geo: QuadGeometry, S, Q, lam, parts=False):
    """(E (N,3), G (N,5)) complex: MatrixVectorMultiply.m for moments S, Q."""
    S = np.asarray(S, dtype=np.complex128)
    Q = np.asarray(Q, dtype=np.complex128)
    lam = np.asarray(lam).reshape(-1)
    m8 = np.concatenate([S, Q], axis=1)
    H = spread(m8, geo.Ngrid, geo.linnode, geo.w_spread)
    fH = np.fft.fftn(H, axes=(0, 1, 2))
    Ht = np.fft.ifftn(scale(fH, geo.k, geo.xi, geo.eta), axes=(0, 1, 2))
    K = contract(Ht, geo.Ngrid, geo.linnode, geo.w_contract)
    # real space
    Er = (3.0 / (4 * PI * (lam - 1)))[:, None] * S + S * geo.tabs[0, 0]
    Gr = (15.0 / (8 * PI * (lam - 1)))[:, None] * Q + 0.5 * Q * geo.tabs[6, 0]
    Ef, Gf = pair_terms(S[geo.p2], Q[geo.p2], geo.rhat, geo.coefs)
    Er = Er + np.stack([np.bincount(geo.p1, weights=np.real(Ef[:, c]), minlength=S.shape[0])
                        + 1j * np.bincount(geo.p1, weights=np.imag(Ef[:, c]), minlength=S.shape[0])
                        for c in range(3)], axis=1)
    Gr = Gr + np.stack([np.bincount(geo.p1, weights=np.real(Gf[:, c]), minlength=S.shape[0])
                        + 1j * np.bincount(geo.p1, weights=np.imag(Gf[:, c]), minlength=S.shape[0])
                        for c in range(5)], axis=1)
    if parts:
        return (K[:, :3], K[:, 3:]), (Er, Gr)
    return K[:, :3] + Er, K[:, 3:] + Gr


def ewald_parameters(box, xi, errortol):
    """rc, kcut, Ngrid, h, P, eta (Quad/Capacitance.m:25-31)."""
    box = np.asarray(box, dtype=np.float64)
    rc = np.sqrt(-np.log(errortol)) / xi
    kcut = 2 * xi * np.sqrt(-np.log(errortol))
    Ngrid = np.ceil(1 + box * kcut / PI).astype(np.int64)
    h = box / Ngrid
    P = int(np.ceil(-2 * np.log(errortol) / PI))
    eta = P * (h * xi) ** 2 / PI
    return rc, kcut, Ngrid, h, P, eta


def solve(geo: QuadGeometry, lam, E0, S_guess, Q_guess, tol, restart=10, maxit=100):
    """MatrixVectorSolve.m:35-43: GMRES(10) on [E0; 0] with the 8N unknowns
    ordered [S (3N); Q (5N)]."""
    n = geo.x.shape[0]
    lam = np.asarray(lam).reshape(-1)

    def mv(v):
        S = v[:3 * n].reshape(n, 3)
        Q = v[3 * n:].reshape(n, 5)
        E, G = field(geo, S, Q, lam)
        return np.concatenate([E.reshape(-1), G.reshape(-1)])

    b = np.concatenate([np.tile(np.asarray(E0, dtype=np.complex128), n), np.zeros(5 * n, dtype=np.complex128)])
    x0 = np.concatenate([np.asarray(S_guess, dtype=np.complex128).reshape(-1),
                         np.asarray(Q_guess, dtype=np.complex128).reshape(-1)])
    A = LinearOperator((8 * n, 8 * n), matvec=mv, dtype=np.complex128)
    sol, _ = gmres(A, b, x0=x0, rtol=tol, atol=0.0, restart=restart, maxiter=maxit)
    return sol[:3 * n].reshape(n, 3), sol[3 * n:].reshape(n, 5)


def dense_operator(geo: QuadGeometry, lam):
    """The 8N x 8N matrix in the [S; Q] ordering of MatrixVectorSolve.m."""
    n = geo.x.shape[0]
    A = np.zeros((8 * n, 8 * n), dtype=np.complex128)
    for k in range(8 * n):
        v = np.zeros(8 * n)
        v[k] = 1.0
        E, G = field(geo, v[:3 * n].reshape(n, 3), v[3 * n:].reshape(n, 5), lam)
        A[:, k] = np.concatenate([E.reshape(-1), G.reshape(-1)])
    return A


# ---------------------------------------------------------------------------
# field at probe points (Compute_Field_Quadrupole/FieldMap.m, MatrixVectorMultiply.m,
# RealSpace.m)
# ---------------------------------------------------------------------------
def field_map(x_E, x_p, S, Q, box, xi):
    """E (N_E,3) complex at the probe points x_E from the dipoles S (N_p,3)
    and quadrupoles Q (N_p,5) at x_p: errortol 1e-3 and the parameter rule
    (FieldMap.m:17-36), the 14 tables on r = [0, 1e-4:1e-4:10] without the
    overlap terms (FieldMap.m:23-25; Compute_Field_Quadrupole/RealSpaceTable.m
    prepends no self terms), the particles' moments spread, scaled by the
    four k-space blocks and contracted at the probes
    (MatrixVectorMultiply.m:31-76), the real-space sum over probe-particle
    pairs with 0 < d < rc (RealSpace.m:27-100, brute force over all pairs in
    place of NeighborList_Types), and E = -(Ek + Er) (FieldMap.m:49-52)."""
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    S = np.asarray(S, dtype=np.complex128).reshape(-1, 3)
    Q = np.asarray(Q, dtype=np.complex128).reshape(-1, 5)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    errortol = 1e-3
    rvals = np.concatenate(([0.0], np.arange(1, 100001) * 1e-4))
    t = dict(zip(TABLE_NAMES, real_space_table(rvals, xi, regularize=False)))
    tabs = np.stack([t[n] for n in OPERATOR_TABLES])
    rc, _, Ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff larger than half the box size")
    offset, offsetxyz = eo.pre_calculations(P, h)
    lin_p, w_p, _ = eo.prepare_grid_kernel(x_p, Ngrid, h, xi, eta, P, offset, offsetxyz)
    H = spread(np.concatenate([S, Q], axis=1), Ngrid, lin_p, w_p)
    fH = np.fft.fftn(H, axes=(0, 1, 2))
    Ht = np.fft.ifftn(scale(fH, kgrid(Ngrid, box), xi, eta), axes=(0, 1, 2))
    lin_e, _, wc_e = eo.prepare_grid_kernel(x_E, Ngrid, h, xi, eta, P, offset, offsetxyz)
    Ek = contract(Ht, Ngrid, lin_e, wc_e)[:, :3]
    Er = np.zeros((x_E.shape[0], 3), dtype=np.complex128)
    for i in range(x_E.shape[0]):
        rr = x_E[i][None, :] - x_p
        rr = rr - box[None, :] * np.trunc(2.0 * rr / box[None, :])
        d = np.sqrt(np.sum(rr * rr, axis=1))
        keep = (d < rc) & (d > 0)
        if not np.any(keep):
            continue
        d, rr = d[keep], rr[keep]
        Ef, _ = pair_terms(S[keep], Q[keep], rr / d[:, None], interpolate(d, tabs, rvals))
        Er[i] = np.sum(Ef, axis=0)
    return -(Ek + Er)
