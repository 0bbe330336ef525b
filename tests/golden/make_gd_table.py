"""Write tests/golden/gd270_bh.npz: the B(H) table of gadolinium at 270 K
that Magnetostatics/NewtonRaphson.jl:40-53 loads (loadMaterial(data,
"Materials", "Gd_MFT", "Gd", 7.9, 270.0)), from the data files the reference
ships in Magnetostatics/Materials/Gd_MFT/ (HofM.dat in Oe, TofM.dat in K,
M.dat in emu/g).

Steps of src/magneticProperties.jl:44-76 and 107-130, restated here:
  - M [A/m] = M [emu/g] * density * 1e3, H [A/m] = H [Oe] * 1e-4 / mu0;
  - M(H, 270 K) from the bicubic interpolating spline of M over (H, T), which
    Dierckx.jl's Spline2D(H, T, M) builds with FITPACK's regrid (kx = ky = 3,
    s = 0): here scipy's RectBivariateSpline, the same FITPACK routine;
  - B = mu0 (H + M).
The fixture holds only the resulting data (62 values of H and of B) and
these provenance notes. Run in a checkout that has the reference at
/root/reference: python tests/golden/make_gd_table.py
"""
import os

import numpy as np
from scipy.interpolate import RectBivariateSpline

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/Magnetostatics/Materials/Gd_MFT"
MU0 = np.pi * 4e-7
DENSITY = 7.9  # g/cm3, NewtonRaphson.jl:45
T = 270.0      # K, NewtonRaphson.jl:29


def main():
    H_oe = np.loadtxt(os.path.join(SRC, "HofM.dat"))
    TofM = np.loadtxt(os.path.join(SRC, "TofM.dat"))
    M_emug = np.loadtxt(os.path.join(SRC, "M.dat"))
    if M_emug.shape != (H_oe.size, TofM.size):
        M_emug = np.loadtxt(os.path.join(SRC, "M.dat"), delimiter=",")
    M = M_emug * DENSITY * 1e3
    H = H_oe * 1e-4 / MU0
    spl = RectBivariateSpline(H, TofM, M, kx=3, ky=3, s=0)
    MT = spl(H, T, grid=False)
    B = MU0 * (H + MT)
    np.savez(os.path.join(HERE, "gd270_bh.npz"), H=H, B=B,
             note=np.array("Gd at 270 K, density 7.9 g/cm3, from Magnetostatics/Materials/Gd_MFT "
                           "(HofM.dat, TofM.dat, M.dat); B = mu0 (H + M(H, 270 K)), M from the "
                           "bicubic FITPACK interpolating spline over (H, T)"))
    print(f"H: {H.size} values in [{H[0]:.3g}, {H[-1]:.3g}] A/m; B in [{B.min():.3g}, {B.max():.3g}] T")


if __name__ == "__main__":
    main()
