"""Writes eh_nowiggle_512.dat: the Eisenstein & Hu (1998, ApJ 496, 605)
zero-baryon-oscillation ("no-wiggle") transfer function, eqs. 26, 28-31,
tabulated on 512 log-spaced k in [1e-4, 1e2] h/Mpc for Omega_m=0.31,
Omega_b=0.049, h=0.677, T_CMB=2.7255 K, in the 7-column CAMB transfer format
genetIC reads (every transfer column holds T_EH(k); genetIC uses columns 1
and 6 divided by column 1 of the first row).
"""
import math
import os

import numpy as np


def eh_nowiggle(k_hmpc, om=0.31, ob=0.049, h=0.677, tcmb=2.7255):
    theta = tcmb / 2.7
    omh2 = om * h * h
    obh2 = ob * h * h
    fb = ob / om
    s = 44.5 * math.log(9.83 / omh2) / math.sqrt(1 + 10 * obh2 ** 0.75)        # eq. 26, Mpc
    alpha = 1 - 0.328 * math.log(431 * omh2) * fb + 0.38 * math.log(22.3 * omh2) * fb * fb  # eq. 31
    k = np.asarray(k_hmpc) * h                                                   # Mpc^-1
    gamma_eff = omh2 * (alpha + (1 - alpha) / (1 + (0.43 * k * s) ** 4))         # eq. 30 (Mpc^-1 units of Gamma h)
    q = k * theta * theta / gamma_eff                                           # eq. 28
    L0 = np.log(2 * math.e + 1.8 * q)
    C0 = 14.2 + 731.0 / (1 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)                                               # eq. 29


def main():
    k = np.logspace(-4, 2, 512)
    t = eh_nowiggle(k)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eh_nowiggle_512.dat")
    with open(path, "w") as f:
        for ki, ti in zip(k, t):
            f.write(" ".join(["%.10e" % ki] + ["%.10e" % ti] * 6) + "\n")


if __name__ == "__main__":
    main()
