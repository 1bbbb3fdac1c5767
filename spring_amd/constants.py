"""Physical constants of the scattering-rate path, in the reference's unit system.

Frequencies are angular, in rad/ps; force constants in eV/Angstrom^n; masses in amu.
Each constant is recomputed from the same CODATA-style factors the reference uses
(src/Kappa/Src/constants.f90:42-65) so that results agree to the last bits.
"""
import math

PI = 4.0 * math.atan(1.0)
ONE_BY_SQRT_2PI = 1.0 / math.sqrt(2.0 * PI)          # constants.f90:30
# constants.f90:40 writes 1.0E-8 without a kind suffix, so it is the single-precision
# value widened to double.
EPS = 9.99999993922529e-09

# sqrt(eV / (A^2 amu)) -> rad/ps                                  constants.f90:45
W_THZ = math.sqrt(1.60217733 * 6.02214076 * 10.0) * 10.0
THZ_CONV = W_THZ / (2.0 * PI)
VG_MS = W_THZ * 100.0                                             # constants.f90:49
# hbar*omega/(k_B T) prefactor for omega in sqrt(eV/A^2/amu)     constants.f90:54-55
BE_CONV = math.sqrt(1.60217733 * 6.02214076 * 10.0) * 6.582119569 * 100.0 / 8.617333262
# |V3|^2 -> rad/ps scattering factor                               constants.f90:59-60
SCTR3_THZ2 = 1.054571817 * (1.60217733 ** 2) * (6.02214076 ** 3) * 1000000.0 * PI * 0.0625
# |V4|^2 -> rad/ps                                                 constants.f90:62-63
SCTR4_THZ2 = 0.125 * PI * ((1.054571817 * 1.602176565) ** 2) * (6.022140857 ** 4) * 1000000.0
# kappa prefactor, W/(m K)                                         constants.f90:65
KW_M_K = (1.054571817 ** 2) * 0.001 / 1.380649
