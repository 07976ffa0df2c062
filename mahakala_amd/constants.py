"""Physical constants in cgs, with the reference's values (constants.py:23-31).

These are the reference's own numbers (not CODATA), kept so that units and
emissivities match it to the last digit.
"""
# universal constants
EE = 4.8032e-10
KB = 1.3807e-16
CL = 2.99792458e10
ME = 9.1094e-28
HPL = 6.6261e-27
GNEWT = 6.6743e-8

# other quantities
Msun = 1.989e33
