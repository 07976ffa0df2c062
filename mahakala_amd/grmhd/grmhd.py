"""Base class of fluid models (reference: mahakala/grmhd/grmhd.py:23-55)."""
import numpy as np

KB = 1.3807e-16
CL = 2.99792458e10
ME = 9.1094e-28
MP = 1.6726e-24
EC = 4.8032e-10
HPL = 6.6261e-27
GNEWT = 6.6743e-8


class GRMHDFluidModel:

    def __init__(self):
        pass

    def get_units(self, M_BH, mass_scale):
        """Code -> cgs units for a black hole of mass M_BH (g) and a density
        scale mass_scale (g): L = G M / c^2, T = L / c, rho = mass_scale / L^3,
        Ne = rho / (m_p + m_e), B = c sqrt(4 pi rho)  (grmhd.py:40-55)."""
        L_unit = GNEWT * M_BH / CL ** 2
        T_unit = L_unit / CL
        dens_unit = mass_scale / L_unit ** 3
        Ne_unit = dens_unit / (MP + ME)
        B_unit = CL * np.sqrt(4. * np.pi * dens_unit)
        return dict(L_unit=L_unit, T_unit=T_unit, dens_unit=dens_unit, Ne_unit=Ne_unit, B_unit=B_unit)
