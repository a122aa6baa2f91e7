"""Small seeded cases shared by the parity tests (sized for the oracle).

Every case runs on a stretched, slightly skewed grid (fixtures.stretched_corners)
with dx != dy, so DUX, DVY, DZX and DZY differ from cell to cell along both
axes: an index slip in a horizontal metric changes the results."""
from mohid_amd import constants as C
from mohid_amd import fixtures as fx

# name -> (base config, overrides)
SMALL = {
    "upwind_flat": ("c1", dict(ni=18, nj=14, nk=6)),
    "upwind_land_ragged": ("c3", dict(ni=20, nj=17, nk=7, kfloor=(1, 7), land_frac=0.3,
                                      adv_method=C.UPWIND_ORDER1, tracers=("lognormal",))),
    "tvd_sigma": ("c2", dict(ni=16, nj=19, nk=8, depth=(5.0, 60.0))),
    "tvd_ragged_opt": ("c3", dict(ni=22, nj=18, nk=9, kfloor=(2, 9), land_frac=0.3,
                                  tracers=("temperature", "salinity", "lognormal"))),
    "tvd_stretched_stiff": ("c4", dict(ni=14, nj=12, nk=10, depth=(5.0, 40.0),
                                       tracers=("lognormal", "temperature"))),
    "tvd_land_sigma": ("c5", dict(ni=24, nj=20, nk=7, depth=(2.0, 80.0),
                                  tracers=("temperature", "lognormal"))),
}


GRID = dict(metrics="stretched", dx=500.0, dy=350.0)


def make(name, **extra):
    base, ov = SMALL[name]
    ov = dict(GRID, **ov)
    ov.update(extra)
    return fx.make_case(fx.config(base, **ov))
