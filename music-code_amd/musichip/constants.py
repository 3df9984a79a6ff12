"""Physical and numerical constants of the reference (defaults.F90:33-145,
ewald.F90:60, coulomb.F90:292-300).  pi is the reference's truncated value."""
PI = 3.1415927
TWOPI = 2.0 * PI
KJMOLE_KB = 8.314e-3
CALTOJ = 4.184
SCALEPE = 4.184
KCALMOLE_KB = KJMOLE_KB / CALTOJ
NAV = 6.0221367e23
RGAS = 8.314
HPLANCK = 6.626e-34
MAX_EXP = 30.0
LOW_DIST_CUTOFF = 1.0e-3
HIGH_DIST_CUTOFF = 13.0
VERY_LRG_NRG = 1.0e10
TOKCAL = 332.064
COUL_CONST = 331.5
EWALD_NKMAX = 50000
