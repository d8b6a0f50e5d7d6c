"""CGS units of the simulation (the reference's hamunits.h:8-68 table).

Values are formed by the same expressions so that every parsed parameter
rounds to the same double as in the reference's Param.
"""


class CGS:
    erg = 1.0
    cm = 1.0
    sec = 1.0
    pi = 3.141592653589793238462643383279502884197
    halfpi = 1.570796326794896619231321691639751442099
    twopi = 6.283185307179586476925286766559005768394
    rad = pi / 180.0
    m = 1.0e2 * cm
    pc = 3.0856775806e18 * cm
    kpc = 3.0856775806e21 * cm
    eV = 1.60217733e-12 * erg
    GeV = 1.0e9 * eV
    GHz = 1.0e9 / sec
    muGauss = 1.0e-6
    ccm = cm * cm * cm
    c_light = 2.99792458e10
    h_planck = 6.626075540e-27
    qe = 4.80320425e-10
    mec2 = 0.51099907e-3 * GeV
    mec = mec2 / c_light
    kB = 1.380622e-16
