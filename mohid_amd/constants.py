"""Numeric constants of MOHID's tracer-transport step.

Values are the reference's (Software/MOHIDBase1/ModuleGlobalData.F90), taken
verbatim so that callers can pass the same enums and sentinels as the Fortran
code: GD:143 ``null_real``, GD:1810-1812 advection methods and TVD limiters,
GD:1831-1832 explicit/implicit weights, GD:180-199 status codes.
"""

NULL_REAL = -9.9e15          # GD:143
MIN_VALUE = 1.0e-16          # GD:1812

# GD:1810 advection methods
UPWIND_ORDER1 = 1
UPWIND_ORDER2 = 2
UPWIND_ORDER3 = 3
P2_TVD = 4
CENTRAL_DIF = 5
LEAP_FROG = 6

# GD:1811 TVD limiters
MINMOD = 1
VANLEER = 2
MUSCL = 3
SUPERBEE = 4
PDM = 5

# GD:1831-1832 (AdvectionDiffusion receives 1 - file value, WP:9305-9660)
EXPLICIT_SCHEME = 0.0
IMPLICIT_SCHEME = 1.0

# GD:180-199 status codes
SUCCESS_ = 0
OFF_ERR_ = 1
UNKNOWN_ = -99

# ModuleAdvectionDiffusion.F90:152-158 boundary conditions
MASS_CONSERVATION = 1
IMPOSED_VALUE = 2
NULL_GRADIENT = 4
SUB_MODEL = 5
ORLANSKI = 6
MASS_CONSERV_NULL_GRAD = 7
CYCLIC = 8
