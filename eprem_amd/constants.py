"""Physical constants and unit conversions of the reference
(src/baseTypes.h:50-115), in the same expression order so the float
values match the C macros bit for bit."""
import math

PI = 3.1415926535897932
TWO_PI = 6.283185307179586
VERYSMALL = 1.0e-33
MP = 1.6726e-24          # g
EV = 1.6022e-12          # erg
MEV = 1.6022e-6          # erg
GEV = 1.6022e-3
Q = 4.80320425e-10
C = 2.99792458e10        # cm/s
MZERO = MP * C * C / GEV
AU = 1.495978707e13      # cm
RSUN = 6.96e10
RSAU = RSUN / AU
TAU = AU / C             # code time unit [s]
DAY = TAU / (24.0 * 60.0 * 60.0)
MHD_DENSITY_NORM = 1.0
MHD_B_NORM = math.sqrt(MP * MHD_DENSITY_NORM) * C
MAS_TIME_NORM = 1445.87003080685
MAS_LENGTH_NORM = 6.96e10
MAS_RHO_NORM = 1.6726e-16
MAS_V_CONVERT = (MAS_LENGTH_NORM / MAS_TIME_NORM) / C
MAS_RHO_CONVERT = (MAS_RHO_NORM / MP) / MHD_DENSITY_NORM
MAS_B_CONVERT = math.sqrt(4.0 * PI * MAS_RHO_NORM) * MAS_LENGTH_NORM / MAS_TIME_NORM / MHD_B_NORM
