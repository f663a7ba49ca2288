"""Interface enumerators of the reference engine, kept value-for-value.

These are the numbers that cross the drop-in boundary: the row indices of the
particle ``store`` array, the slots of the ``statevars`` vector and the codes
stored in them.  They follow ``playground/src/const.f90`` exactly
(const.f90:3-193); indices here are 1-based like the Fortran source, so row
``es_h`` of a NumPy ``store[i, :]`` view is ``store[i, es_h - 1]``.
"""

# physical constants, const.f90:3-8 (CGS)
pi = 3.141592653589793
lightspeed = 2.997924e10
Rg = 8.314e7
sigmaB = 5.67051e-5

e_none = -1

# statevars slots, const.f90:14-66
ec_dim = 1
ec_eqs = 2
ec_ddw = 3
ec_ics = 4
ec_adden = 5
ec_artts = 6
ec_coordsys = 7
ec_disotropic = 8
ec_dcondconst = 9
ec_muzero = 10
ec_tfinish = 11
ec_npics = 12
ec_hfac = 13
ec_silent = 14
ec_resolution = 15
ec_spacing = 16
ec_xmin = 17
ec_xmax = 18
ec_ymin = 19
ec_ymax = 20
ec_zmin = 21
ec_zmax = 22
ec_bordsize = 23
ec_padding = 24
ec_realpn = 25
ec_fixedpn = 26
ec_gamma = 27
ec_lastprint = 28
ec_usedumps = 29
ec_dtprint = 30
ec_process = 31
ec_au = 32
ec_placement = 33
ec_eqonhydro = 34
ec_eqondiff = 35
ec_eqonfluxlim = 36
ec_eqonradexch = 37
ec_eqonsts = 38
ec_udist = 39
ec_umass = 40
ec_utime = 41
ec_uvelocity = 42
ec_udensity = 43
ec_upressure = 44
ec_uacceller = 45
ec_time = 46
ec_molecularmass = 47
ec_ststype = 48
ec_stsfixeds = 49
ec_total = 49

# store rows, const.f90:69-103
es_rx = 1
es_ry = 2
es_rz = 3
es_vx = 4
es_vy = 5
es_vz = 6
es_ax = 7
es_ay = 8
es_az = 9
es_m = 10
es_den = 11
es_h = 12
es_dh = 13
es_p = 14
es_u = 15
es_du = 16
es_t = 17
es_kappa = 18
es_dtdx = 19
es_dtdy = 20
es_dtdz = 21
es_ddt = 22
es_bx = 23
es_by = 24
es_bz = 25
es_dbx = 26
es_dby = 27
es_dbz = 28
es_c = 29
es_om = 30
es_type = 31
es_vfld = 32
es_total = 32

# particle types, const.f90:106-112
ept_empty = 0
ept_real = 1
ept_fixed = 2
ept_fixedreal = 21
ept_fixedmirror = 22
ept_periodic = 3

# legacy initial setups, const.f90:115-127
ett_sin3 = 100
ett_mti = 101
ett_shock12 = 102
ett_pulse = 103
ett_ring = 104
ett_soundwave = 105
ett_hydroshock = 106
ett_alfvenwave = 107
ett_OTvortex = 108
ett_boilingtank = 109
ett_mtilowres = 110
ett_fld_gauss = 111

# border selectors, const.f90:130-140
ebc_all = 200
ebc_x = 201
ebc_y = 202
ebc_z = 203
ebc_x1 = 204
ebc_y1 = 205
ebc_z1 = 206
ebc_x2 = 207
ebc_y2 = 208
ebc_z2 = 209

# equation sets, const.f90:143-151
eeq_hydro = 300
eeq_diffusion = 301
eeq_magnetohydro = 302
eeq_magnetohydrodiffusion = 303
eeq_hydrodiffusion = 304
eeq_kd2 = 305
eeq_hyrad = 306
eeq_manual = 307

# second-derivative operators, const.f90:154-159
esd_fab = 400
esd_n2w = 401
esd_fw = 402
esd_2nw_ds = 403
esd_2nw_sd = 404

# processes, const.f90:161-164
epc_backcompatibility = 600
epc_fullyperiodic = 601
epc_borderless = 602

# placements, const.f90:166-169
epl_uniform = 700
epl_random = 701
epl_closepacked = 702

eif_yes = 1.0
eif_no = 0.0

# equation components, const.f90:175-182 (1-based flag slots)
eqs_hydro = 1
eqs_magneto = 2
eqs_diff = 3
eqs_fld = 4
eqs_radexch = 5
eqs_sts = 6
eqs_total = 6

# STS stage-count modes, const.f90:191-193
ests_auto = 0
ests_fixeds = 1

# string tables of the ./execute flags, state.f90:91-112,174-184,479-491
EQUATIONS = {
    "hydro": eeq_hydro,
    "magnetohydro": eeq_magnetohydro,
    "diffusion": eeq_diffusion,
    "mhd": eeq_magnetohydrodiffusion,
    "hydi": eeq_hydrodiffusion,
    "kd2": eeq_kd2,
    "hyrad": eeq_hyrad,
    "manual": eeq_manual,
}
DDW = {
    "n2w": esd_n2w,
    "fab": esd_fab,
    "2nw-+": esd_2nw_ds,
    "2nw+-": esd_2nw_sd,
    "fw": esd_fw,
}
PROCESS = {
    "borderless": epc_borderless,
    "fullyperiodic": epc_fullyperiodic,
    "backcompatibility": epc_backcompatibility,
}
YESNO = {"yes": eif_yes, "no": eif_no, "": eif_yes}

# compile-time kernel choice of the reference (Makefile:14 kernel=m4|m6|gauss)
KERNEL_M4 = 0
KERNEL_M6 = 1
KERNEL_GAUSS = 2
KERNELS = {"m4": KERNEL_M4, "m6": KERNEL_M6, "gauss": KERNEL_GAUSS}
# krad and maxneibnum per kernel, base_m4.f90:12-18, base_m6.f90:13-17,
# base_gauss.f90:9-13
KRAD = {KERNEL_M4: 2.0, KERNEL_M6: 3.0, KERNEL_GAUSS: 10.0}
MAXNEIB = {KERNEL_M4: (35, 250, 150), KERNEL_M6: (50, 10000, 150),
           KERNEL_GAUSS: (100, 5000, 50000)}


def eq_components(state):
    """Equation-set flags the loops read, as ``getEqComponent`` (state.f90:117-149).

    Returns a dict of booleans keyed hydro/magneto/diff/fld/radexch/sts.
    """
    eq = int(state[ec_eqs - 1])
    f = dict(hydro=False, magneto=False, diff=False, fld=False,
             radexch=False, sts=False)
    if eq == eeq_hydro:
        f["hydro"] = True
    elif eq == eeq_magnetohydro:
        f.update(hydro=True, magneto=True)
    elif eq == eeq_diffusion:
        f["diff"] = True
    elif eq == eeq_hyrad:
        f.update(fld=True, hydro=True, radexch=True)
    elif eq == eeq_magnetohydrodiffusion:
        f.update(hydro=True, magneto=True, diff=True)
    elif eq == eeq_hydrodiffusion:
        f.update(hydro=True, diff=True)
    elif eq == eeq_manual:
        f["diff"] = int(state[ec_eqondiff - 1]) == 1
        f["hydro"] = int(state[ec_eqonhydro - 1]) == 1
        f["fld"] = int(state[ec_eqonfluxlim - 1]) == 1
        f["radexch"] = int(state[ec_eqonradexch - 1]) == 1
        f["sts"] = int(state[ec_eqonsts - 1]) == 1
    return f
