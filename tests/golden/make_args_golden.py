"""Writes tests/golden/args_statevars.json: the statevars vector the
reference's own ./execute flag parser (args.f90 fillargs, state.f90 setters)
produces for each argument list below, or the message it stops with.

Runs oracle/_ref/m6/args_harness (make -f oracle/ref.mk KERNEL=m6 args), the
reference's args.f90/state.f90/const.f90 compiled from /root/reference with
this repository's driver oracle/args_harness.f90.  Only the outputs are
committed.

    python tests/golden/make_args_golden.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "m6", "args_harness")

FULL = ["--dim", "2", "--equations", "hydro", "--ddw", "fab", "--hfac", "1.0", "--tfinish", "0.5",
        "--adden", "yes", "--artts", "no", "--au", "0.1", "--disotropic", "no", "--process", "fullyperiodic",
        "--placement", "uniform", "--resolution", "64", "--spacing", "0.03125", "--xmin", "-1", "--xmax", "1",
        "--ymin", "-1", "--ymax", "1", "--zmin", "0", "--zmax", "0", "--nsnapshots", "10", "--silent", "yes",
        "--usedumps", "no", "--coordsys", "cartesian", "--ics", "none"]

CASES = [
    [],
    FULL,
    ["--dim", "3", "--equations", "mhd", "--ddw", "2nw-+", "--process", "borderless", "--placement", "random"],
    ["--dim", "1", "--equations", "diffusion", "--ddw", "2nw+-", "--process", "backcompatibility",
     "--ics", "soundwave", "--disotropic", "yes"],
    ["--equations", "magnetohydro", "--ddw", "n2w", "--placement", "closepacked", "--coordsys", "cylindric"],
    ["--equations", "hydi", "--ddw", "fw", "--adden", "no", "--artts", "yes", "--silent", "no",
     "--usedumps", "yes"],
    ["--equations", "kd2"], ["--equations", "hyrad"],
    ["--ics", "sin3"], ["--ics", "mti"], ["--ics", "mtilowres"], ["--ics", "shock12"], ["--ics", "pulse"],
    ["--ics", "ring"], ["--ics", "hydroshock"], ["--ics", "alfvenwave"], ["--ics", "otvortex"],
    ["--ics", "boilingtank"], ["--ics", "fld_gauss"],
    # number formats of a list-directed read
    ["--dim", "3.0", "--hfac", "1.2d0", "--tfinish", "2e-1", "--au", ".5", "--spacing", "7,8",
     "--resolution", "5.", "--xmin", "-1.5E+00", "--nsnapshots", "3.9"],
    # zeros leave their slots alone (tfinish, spacing, snapshots, dim)
    ["--tfinish", "0", "--spacing", "0", "--nsnapshots", "0", "--dim", "0"],
    # an unknown flag is skipped; a last key without a value is not read
    ["--foo", "bar", "--dim", "2"], ["--dim", "2", "--hfac"], ["--hfac"],
    # values the reference stops on
    ["--equations", "manual"], ["--equations", "euler"], ["--ddw", "xyz"], ["--dim", "4"],
    ["--process", "periodic"], ["--adden", "maybe"], ["--coordsys", "polar"], ["--ics", "foo"],
    ["--placement", "grid"], ["--nsnapshots", "-1"], ["--artts", "sometimes"], ["--silent", "loud"],
    ["--usedumps", "often"], ["--disotropic", "partly"],
]


def run(argv):
    p = subprocess.run([HARNESS] + argv, capture_output=True, text=True, timeout=60)
    out = p.stdout.splitlines()
    if "STATEVARS" in out:
        k = out.index("STATEVARS")
        return {"argv": argv, "statevars": [float(x) for x in out[k + 1:k + 50]], "error": None}
    msg = [ln for ln in out if "<!>" in ln or "There is no such" in ln]
    return {"argv": argv, "statevars": None, "error": (msg[0].strip() if msg else p.stdout.strip())[:200]}


def main():
    res = [run(c) for c in CASES]
    with open(os.path.join(HERE, "args_statevars.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(len(res), "cases,", sum(r["error"] is not None for r in res), "stops")


if __name__ == "__main__":
    main()
