"""The ``./execute`` flag surface: command-line flags to the ``statevars``
vector, as the reference's ``fillargs`` (args.f90:29-176) fills it through the
setters of state.f90.

``fillargs(argv)`` returns the 49 statevars (1-based slot ``k`` at index
``k - 1``) that the reference holds after ``clearState`` (every slot -1,
state.f90:43-45) and ``fillargs`` on the same arguments; pass ``state`` to
start from another vector.  Where the reference stops the program (an
unknown equation set, operator, process, placement, initial condition,
coordinate system, silent/dumps/density/artificial-term choice, a dimension
outside 1..3 or a negative snapshot count) this raises ``ArgsError`` with the
reference's message.  Unknown flags are skipped, as the reference warns and
goes on.
"""
import re

import numpy as np

from . import consts as C


class ArgsError(ValueError):
    """A flag value the reference rejects (it prints the message and stops)."""


# set_equations, state.f90:91-112 (no 'manual' there: it is not a flag value)
_EQUATIONS = {
    "hydro": C.eeq_hydro,
    "magnetohydro": C.eeq_magnetohydro,
    "diffusion": C.eeq_diffusion,
    "mhd": C.eeq_magnetohydrodiffusion,
    "hydi": C.eeq_hydrodiffusion,
    "kd2": C.eeq_kd2,
    "hyrad": C.eeq_hyrad,
}
# sinitvar, state.f90:174-213
_ICS = {
    "sin3": C.ett_sin3,
    "mti": C.ett_mti,
    "mtilowres": C.ett_mtilowres,
    "shock12": C.ett_shock12,
    "pulse": C.ett_pulse,
    "ring": C.ett_ring,
    "soundwave": C.ett_soundwave,
    "hydroshock": C.ett_hydroshock,
    "alfvenwave": C.ett_alfvenwave,
    "otvortex": C.ett_OTvortex,
    "boilingtank": C.ett_boilingtank,
    "fld_gauss": C.ett_fld_gauss,
    "none": C.e_none,
}
# splacement, state.f90:497-512
_PLACEMENT = {"uniform": C.epl_uniform, "random": C.epl_random, "closepacked": C.epl_closepacked}

# flags that write their slot as soon as they are read (args.f90:74-75,
# 105-122): --resolution and the box bounds take the number as it is
_DIRECT = {
    "--resolution": C.ec_resolution,
    "--xmin": C.ec_xmin, "--xmax": C.ec_xmax,
    "--ymin": C.ec_ymin, "--ymax": C.ec_ymax,
    "--zmin": C.ec_zmin, "--zmax": C.ec_zmax,
}

_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?$")


def _real(text, flag):
    """A list-directed ``read(argval, *) x`` of one real: the first item up
    to a blank, comma or slash; Fortran's d exponent is accepted."""
    item = re.split(r"[\s,/]", text.strip(), maxsplit=1)[0]
    if not _NUM.match(item):
        raise ArgsError("bad number for %s: %r" % (flag, text))
    return float(item.replace("d", "e").replace("D", "E"))


def _yes_no(value, message, yes=1.0, no=0.0, empty=1.0):
    """The yes / no / '' setters of state.f90 (setAdvancedDensity :219-232,
    setArtificialTerms :238-250, setdiffisotropic :293-305, setsilentmode
    :360-373, setUseDumps)."""
    if value == "yes":
        return yes
    if value == "no":
        return no
    if value == "":
        return empty
    raise ArgsError(message + ": " + value)


def clear_state():
    """``clearState`` (state.f90:43-45): every slot -1."""
    return np.full(C.ec_total, -1.0)


def fillargs(argv, state=None):
    """statevars after the reference's ``fillargs`` on ``argv`` (the flags,
    without the program name).  See the module docstring."""
    s = clear_state() if state is None else np.array(state, dtype=np.float64, copy=True)
    if s.shape != (C.ec_total,):
        raise ArgsError("state must hold %d values" % C.ec_total)

    def put(slot, v):
        s[slot - 1] = float(v)

    # defaults, args.f90:40-58
    dim, eqs, pspc1, ddwtype, tfinish, npic, hfac = 0, "", 0.0, "", 1.0, 0, 1.2
    silentstr = adden = artts = coordsysstr = usedumps = process = difiso = ""
    au = 0.0
    # key / value pairs; a last key without its value is not read
    # (args.f90:60-125: the loop runs while the key's position < the count)
    argv = [str(a) for a in argv]
    i = 0
    while i < len(argv) - 1:
        key, val = argv[i].lstrip(), argv[i + 1]
        v = val.lstrip()
        i += 2
        if key == "--dim":
            dim = int(_real(val, key))
        elif key == "--equations":
            eqs = v
        elif key == "--spacing":
            pspc1 = _real(val, key)
        elif key == "--nsnapshots":
            npic = int(_real(val, key))
        elif key in _DIRECT:
            put(_DIRECT[key], _real(val, key))
        elif key == "--ddw":
            ddwtype = v
        elif key == "--tfinish":
            tfinish = _real(val, key)
        elif key == "--hfac":
            hfac = _real(val, key)
        elif key == "--silent":
            silentstr = v
        elif key == "--ics":
            name = v.rstrip()
            if name not in _ICS:
                raise ArgsError("There is no such initial variable | " + name)
            put(C.ec_ics, _ICS[name])
        elif key == "--adden":
            adden = v
        elif key == "--artts":
            artts = v
        elif key == "--coordsys":
            coordsysstr = v
        elif key == "--usedumps":
            usedumps = v
        elif key == "--process":
            process = v
        elif key == "--au":
            au = _real(val, key)
        elif key == "--disotropic":
            difiso = v
        elif key == "--placement":
            name = v.rstrip()
            if name not in _PLACEMENT:
                raise ArgsError("There is no such initial particles distribution | " + name)
            put(C.ec_placement, _PLACEMENT[name])
        # anything else: the reference warns "Argument not found" and goes on

    # the setters, in the reference's order (args.f90:128-172)
    rs = str.rstrip
    if npic != 0:
        if npic < 0:
            raise ArgsError("Number of snapshots cannot be negative | %d" % npic)
        put(C.ec_npics, npic)
    if tfinish != 0.0:
        put(C.ec_tfinish, tfinish)
    if rs(coordsysstr) != "":
        c = rs(coordsysstr)
        if c not in ("cartesian", "cylindric"):
            raise ArgsError("There is no such coordinate system: " + c)
        put(C.ec_coordsys, 1 if c == "cartesian" else 2)
    if rs(artts) != "":
        put(C.ec_artts, _yes_no(rs(artts), "There is no such default Artificail Term"))
    if rs(eqs) != "":
        if rs(eqs) not in _EQUATIONS:
            raise ArgsError("Equation is not set | " + rs(eqs))
        put(C.ec_eqs, _EQUATIONS[rs(eqs)])
    if rs(ddwtype) != "":
        if rs(ddwtype) not in C.DDW:
            raise ArgsError("Second derivative kernel is not set | " + rs(ddwtype))
        put(C.ec_ddw, C.DDW[rs(ddwtype)])
    put(C.ec_hfac, hfac)
    if rs(silentstr) != "":
        put(C.ec_silent, _yes_no(rs(silentstr), "There is no such silent mode", 1, 0, 0))
    if pspc1 != 0.0:
        put(C.ec_spacing, pspc1)
    if rs(adden) != "":
        put(C.ec_adden, _yes_no(rs(adden), "There is no such advanced density choise"))
    if rs(usedumps) != "":
        put(C.ec_usedumps, _yes_no(rs(usedumps), "There is no such mode for Use Dumps", 1, 0, 1))
    if dim != 0:
        if dim < 1 or dim > 3:
            raise ArgsError("Dimmention is wrong | %d" % dim)
        put(C.ec_dim, dim)
    if rs(process) != "":
        if rs(process) not in C.PROCESS:
            raise ArgsError("The process is not set | " + rs(process))
        put(C.ec_process, C.PROCESS[rs(process)])
    put(C.ec_au, au)
    put(C.ec_disotropic, _yes_no(rs(difiso), "There is no such diso mode"))
    put(C.ec_lastprint, 0)
    return s
