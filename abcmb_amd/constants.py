"""Physical constants used by the perturbation + line-of-sight path.

The values are the ones ABCMB uses (``abcmb/constants.py:8-39``); they must agree
digit for digit because every background closed form (rho, H, aH, n_H, tau_c) is
built from them, and the HIP kernels receive the same numbers through
``AbcmbCosmo``.
"""
import math

c = 29979245800.0              # speed of light, cm/s
c_Mpc_over_s = 9.71561e-15     # speed of light, Mpc/s
H0_over_h = 3.24078e-18        # 100 km/s/Mpc in 1/s
hbar = 6.582119569509e-16      # eV s
G = 1.18980205e-40             # G/c^2 in cm^3 eV^-1 s^-2
kB = 8.617343e-5               # eV/K

mp = 938271999.0
mn = 939565413.0
me = 510998.9461
TCMB_today = 2.34865418e-4     # eV
mH = mp + me + 13.5982860719383
mHe = 3.72839e9
mu_e = mp * me / mH

E21 = 10.198714553953742
rydberg = 13.598286071938324
lya_eng = rydberg * 3.0 / 4.0
lya_freq = lya_eng / (2.0 * math.pi * hbar)
thomson_xsec = 6.652458734e-25
stef_bolt = math.pi ** 2 / (60.0 * hbar ** 3 * c ** 2)
R2s1s = 8.2206

# critical-density unit: rho = omega * RHO_UNIT  (eV cm^-3)
RHO_UNIT = 3.0 * H0_over_h ** 2 / 8.0 / math.pi / G

# massive-neutrino momentum quadratures (abcmb/species.py:791-796)
Q_3P = (0.913201, 3.37517, 7.79184)
W_3P = (0.0687359, 3.31435, 2.29911)
Q_5P = (0.583165, 2.0, 4.0, 7.26582, 13.0)
W_5P = (0.0081201, 0.689407, 2.8063, 2.05156, 0.12681)
