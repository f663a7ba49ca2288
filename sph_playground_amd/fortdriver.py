"""The Python runner's ``Context`` (playground/runner/fortdriver/core.py:46-470)
on this library instead of h5py and ``./execute``.

The reference's runners drive a study through a ``Context``: read the
enumerators of const.f90 (``ReadConstFile``, core.py:235-319), open a state
dump (``ReadDumpHDF``, core.py:321-326: datasets ``state`` and ``store``,
``store[i, es - 1]``), edit ``setup`` / ``state`` / ``store``
(``SetState``, ``AddParticles``, ``ModifyParticlesProperties``) and write it
back (``Apply``, core.py:465-470), then run ``./execute``, which restores from
the dump (main.f90:112-116).  Here the same methods, with the same names,
arguments and effects on the dump, read and write the dump through
libsph_dump.so (the reference's HDF5 layout, include/sph_dump.h), and
``Iterate`` / ``Leapfrog`` run the hot path on the GPU (libsph_hip.so) on the
context's store in place of ``./execute``.

The process helpers of the reference (make, logs, ``os.system`` copies) have
no counterpart: nothing here builds or launches a program.
"""
import copy
import os
import re

import numpy as np

from . import consts as C
from . import dump as D

# enumerator families of const.f90 that the reference's Context keeps
# (core.py:208-233), by prefix
_FAMILIES = ("ec", "es", "ept", "ett", "ebc", "eeq", "esd", "edi", "eqs", "epc", "eif", "ests")


class Setup:
    """Attribute bag of the statevars by name (core.py:33-44)."""
    stderrfile = "__stderr.info"
    stdoutfile = "__stdout.info"
    logfile = "__runslog.info"
    kernel = "m6"

    def toArgs(self):
        args = []
        for k in self.__dict__.keys():
            args.append("--" + k)
            args.append(str(self.__dict__[k]))
        return args


class Context:
    """A study's working copy of one state dump.

    ``constf90``: a const.f90 to read the enumerators from (as core.py:235
    does); by default they come from ``consts`` (the same values)."""

    def __init__(self, constf90=None):
        for fam in _FAMILIES:
            setattr(self, fam, {})
            setattr(self, fam + "Rev", {})
        self.setup = Setup()
        self.store = np.zeros((0, C.es_total))
        self.state = np.zeros(C.ec_total)
        self.directory = None
        self.ReadConstFile(constf90)
        for key in self.ec.keys():
            self.setup.__dict__[key] = 0.0

    # ---- enumerators
    def ReadConstFile(self, constf90=None):
        """Enumerator tables: ``name = value,&`` lines of const.f90 by family
        prefix (core.py:235-319), or the same values from ``consts``."""
        if constf90 is not None:
            pat = re.compile(r"^\s*(" + "|".join(_FAMILIES) + r")_(\w+)\s*=\s*([-+0-9.eEdD]+)\s*(,&)?\s*$")
            with open(constf90) as f:
                for line in f:
                    m = pat.match(line)
                    if not m:
                        continue
                    fam, name, val = m.group(1), m.group(2), m.group(3).replace("d", "e").replace("D", "e")
                    if fam == "es" and name == "total":
                        continue
                    v = float(val) if fam in ("eif", "ests") else int(float(val))
                    self._add(fam, name, v)
            return
        for key, v in vars(C).items():
            fam, _, name = key.partition("_")
            if fam in _FAMILIES and name and isinstance(v, (int, float)) and not (fam == "es" and name == "total"):
                self._add(fam, name, float(v) if fam in ("eif", "ests") else int(v))

    def _add(self, fam, name, v):
        getattr(self, fam)[name] = v
        getattr(self, fam + "Rev")[v] = name

    def idx(self, prop, i):
        """(i, es - 1): the index of particle i's ``prop`` in ``store``."""
        return (i, self.es[prop] - 1)

    # ---- the dump
    def ReadDumpHDF(self, filename):
        """Open a dump (``<dir>/dump_full.h5`` or its directory): ``store``,
        ``state`` and ``setup`` from its datasets (core.py:321-326)."""
        d = filename if os.path.isdir(filename) else os.path.dirname(os.path.abspath(filename))
        if not os.path.isdir(filename) and os.path.basename(filename) != "dump_full.h5":
            raise ValueError("dumps are read as <dir>/dump_full.h5 (dumper.f90:30-38)")
        self.state, self.store = D.restore(d)
        self.directory = d
        for i in range(1, len(self.state) + 1):
            if i in self.ecRev:
                self.setup.__dict__[self.ecRev[i]] = self.state[i - 1]

    def _flush(self):
        if self.directory is None:
            raise RuntimeError("no dump is open (ReadDumpHDF)")
        D.dump(self.directory, self.state, self.store, float(self.state[C.ec_time - 1]))

    def Apply(self):
        """``setup`` back into ``state`` and the dump written (core.py:465-470)."""
        for i in self.ecRev:
            e = self.ecRev[i]
            if e in self.setup.__dict__.keys():
                self.state[i - 1] = self.setup.__dict__[e]
        self._flush()

    # ---- state and setup
    def GetState(self, key):
        return self.state[self.ec[key] - 1]

    def SetState(self, key, val):
        self.state[self.ec[key] - 1] = val

    def PrintPropertyKeys(self):
        print("Properties: " + str(self.es.keys()))

    def PrintSetup(self):
        for i in self.ecRev:
            e = self.ecRev[i]
            if e in self.setup.__dict__.keys():
                print("\t |", "{0:^15}".format(e), "|", "{0:^40}".format(self.setup.__dict__[e]), "|")

    def PrintState(self):
        for i in self.ecRev:
            print("\t |", "{0:^15}".format(self.ecRev[i]), "|", "{0:^40}".format(self.state[i - 1]), "|")

    # ---- particles
    def ParticlesWithWrongPositions(self):
        """Real particles outside the box (core.py:344-360)."""
        n = int(self.setup.realpn)
        r = self.store[:n, 0:3]
        lo = np.array([self.setup.xmin, self.setup.ymin, self.setup.zmin])
        hi = np.array([self.setup.xmax, self.setup.ymax, self.setup.zmax])
        return [int(i) for i in np.nonzero(~np.all((r >= lo) & (r <= hi), axis=1))[0]]

    def AddParticles(self, dim: str, type: str, method: str):
        """Append, for every real particle within ``bordsize`` of an end of
        axis ``dim``, its periodic image across the box as a particle of
        ``type`` (core.py:383-416): x -> max + (x - min) below min + bordsize,
        x -> min + (x - max) above max - bordsize, kept when it moved by more
        than 1e-3 of the spacing; ``fixedpn`` is set for fixed types."""
        iprop = self.es["r" + dim] - 1
        itype = self.es["type"] - 1
        bs = self.setup.bordsize
        lo = self.setup.__dict__[dim + "min"]
        hi = self.setup.__dict__[dim + "max"]
        dr = (hi - lo) / self.setup.resolution
        nreal = int(self.setup.realpn)
        old = self.store[:nreal, iprop]
        new = old.copy()
        low, high = old < lo + bs, old > hi - bs
        new[low] = hi + old[low] - lo
        new[high] = lo + old[high] - hi
        moved = np.abs(old - new) > dr * 1e-3
        add = self.store[:nreal][moved].copy()
        add[:, itype] = self.ept[type]
        add[:, iprop] = new[moved]
        self.store = np.concatenate([self.store, add])
        if type in ("fixed", "fixedreal"):
            self.setup.fixedpn = add.shape[0]

    def CalcUniformMass(self, rho):
        """Particle mass of density rho over the box's dim-volume (core.py:418-428)."""
        lx = ly = lz = 1.0
        if self.setup.dim > 0:
            lx = abs(self.setup.xmax - self.setup.xmin)
        if self.setup.dim > 1:
            ly = abs(self.setup.ymax - self.setup.ymin)
        if self.setup.dim > 2:
            lz = abs(self.setup.zmax - self.setup.zmin)
        return lx * ly * lz * rho / self.setup.realpn

    def CopyContext(self):
        return copy.deepcopy(self)

    def ModifyParticlesProperties(self, condition=[lambda rx: True], conditionarg=["rx"], properties=["rx"],
                                  value=[lambda rx: 1.0], valuearg=["rx"]):
        """properties[j] of every real and fixed particle set to
        value[j](store[:, valuearg]) (core.py:436-463; like the reference,
        ``condition`` is accepted and not applied), then the dump written."""
        n = int(self.setup.realpn) + int(self.setup.fixedpn)
        new = self.store.copy()
        for j, prop in enumerate(properties):
            arg = valuearg[0] if len(valuearg) == 1 else valuearg[j]
            x1 = self.store[:n, self.es[arg] - 1]
            new[:n, self.es[prop] - 1] = np.vectorize(value[j], otypes=[np.float64])(x1)
        self.store = new
        self._flush()

    # ---- the hot path in place of ./execute
    def _engine(self, kernel, **kw):
        from .engine import Engine
        return Engine(self.state, kernel=kernel or self.setup.kernel, **kw)

    def Iterate(self, kernel=None, **kw):
        """One ``iterate(n, store, maxconsenrg)`` (iterator.f90:44) of the
        context's real and fixed particles on the GPU; ``store`` is updated in
        place.  Returns maxconsenrg (0)."""
        n = int(self.state[C.ec_realpn - 1]) + int(self.state[C.ec_fixedpn - 1])
        eng = self._engine(kernel, **kw)
        try:
            st = np.ascontiguousarray(self.store[:n])
            e = eng.iterate(st)
            self.store[:n] = st
            return e
        finally:
            eng.close()

    def Leapfrog(self, steps, dtcap, kernel=None, **kw):
        """``steps`` passes of the leapfrog loop after the initial iterate
        (main.f90:216-362) on the GPU; ``store`` and ec_time are updated.
        Returns the time steps taken."""
        n = int(self.state[C.ec_realpn - 1]) + int(self.state[C.ec_fixedpn - 1])
        eng = self._engine(kernel, **kw)
        try:
            eng.load(np.ascontiguousarray(self.store[:n]))
            dts = [eng.leapfrog(dtcap) for _ in range(steps)]
            self.store[:n] = eng.save()
        finally:
            eng.close()
        self.state[C.ec_time - 1] += sum(dts)
        self.setup.time = self.state[C.ec_time - 1]
        return dts
