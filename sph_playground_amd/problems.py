"""Seeded synthetic particle sets in the reference's ``store``/``statevars`` form.

The reference's own placements are unseeded (placeRandom.f90:65) and its
``closepacked`` lattice is not periodic in a cube (placeClose.f90:54,103), so
the workloads are generated here with a stated PRNG: splitmix64 for uniform
doubles and Box-Muller for Gaussians.  Every generator returns
``(state, store)``: ``state`` is the 49-slot FP64 ``statevars`` vector
(const.f90:14-66) and ``store`` a C-contiguous ``(n, 32)`` FP64 array whose
row ``i`` is the reference's column ``store(:, i)`` (const.f90:69-103).

Field defaults follow the reference CLI (args.f90:36-63: hfac 1.2, au 0,
tfinish 1) and set-up (place.f90:56-67: bordersize = (int(krad*hfac)+1)*dx).
"""
import math

import numpy as np

from . import consts as C

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """First ``n`` outputs of splitmix64 started at ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """Uniform doubles in (0, 1): the top 53 bits of splitmix64, centred."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def gaussian(seed, n):
    """Standard normal doubles by Box-Muller on pairs of ``uniform`` draws."""
    u = uniform(seed, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def make_state(*, dim, equations="hydro", ddw="fab", adden="yes", artts="yes",
               au=0.0, disotropic="yes", dcond=1.0, muzero=1.0, hfac=1.2,
               box=(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), spacing=None,
               bordersize=None, n_real=0, n_fixed=0, gamma=5.0 / 3.0,
               process="fullyperiodic", kernel=C.KERNEL_M6, tfinish=1.0,
               dtprint=1.0, eqon=None, molecularmass=1.0, ststype=C.ests_fixeds,
               stsfixeds=3, ics=C.e_none):
    """Build a ``statevars`` vector the way args.f90:29-176 and place.f90 fill it."""
    s = np.full(C.ec_total, -1.0)

    def put(slot, v):
        s[slot - 1] = float(v)

    put(C.ec_dim, dim)
    put(C.ec_eqs, C.EQUATIONS[equations])
    put(C.ec_ddw, C.DDW[ddw])
    put(C.ec_ics, ics)
    put(C.ec_adden, C.YESNO[adden])
    put(C.ec_artts, C.YESNO[artts])
    put(C.ec_coordsys, 1)
    put(C.ec_disotropic, C.YESNO[disotropic])
    put(C.ec_dcondconst, dcond)
    put(C.ec_muzero, muzero)
    put(C.ec_tfinish, tfinish)
    put(C.ec_npics, 1)
    put(C.ec_hfac, hfac)
    put(C.ec_silent, 1)
    xmin, xmax, ymin, ymax, zmin, zmax = box
    if dim < 2:
        ymin = ymax = 0.0
    if dim < 3:
        zmin = zmax = 0.0
    put(C.ec_xmin, xmin)
    put(C.ec_xmax, xmax)
    put(C.ec_ymin, ymin)
    put(C.ec_ymax, ymax)
    put(C.ec_zmin, zmin)
    put(C.ec_zmax, zmax)
    if spacing is not None:
        put(C.ec_spacing, spacing)
        put(C.ec_resolution, round((xmax - xmin) / spacing))
    if bordersize is None and spacing is not None:
        bordersize = (int(C.KRAD[kernel] * hfac) + 1) * spacing
    put(C.ec_bordsize, bordersize if bordersize is not None else 0.0)
    put(C.ec_padding, 0.0)
    put(C.ec_realpn, n_real)
    put(C.ec_fixedpn, n_fixed)
    put(C.ec_gamma, gamma)
    put(C.ec_lastprint, 0)
    put(C.ec_usedumps, 0)
    put(C.ec_dtprint, dtprint)
    put(C.ec_process, C.PROCESS[process])
    put(C.ec_au, au)
    put(C.ec_placement, C.epl_uniform)
    eqon = eqon or {}
    put(C.ec_eqonhydro, 1 if eqon.get("hydro") else 0)
    put(C.ec_eqondiff, 1 if eqon.get("diff") else 0)
    put(C.ec_eqonfluxlim, 1 if eqon.get("fld") else 0)
    put(C.ec_eqonradexch, 1 if eqon.get("radexch") else 0)
    put(C.ec_eqonsts, 1 if eqon.get("sts") else 0)
    for slot in (C.ec_udist, C.ec_umass, C.ec_utime, C.ec_uvelocity,
                 C.ec_udensity, C.ec_upressure, C.ec_uacceller):
        put(slot, 1.0)
    put(C.ec_time, 0.0)
    put(C.ec_molecularmass, molecularmass)
    put(C.ec_ststype, ststype)
    put(C.ec_stsfixeds, stsfixeds)
    return s


def empty_store(n):
    st = np.zeros((n, C.es_total), dtype=np.float64)
    st[:, C.es_type - 1] = C.ept_real
    st[:, C.es_om - 1] = 1.0
    return st


def _eos_fill(st, gamma):
    rho = st[:, C.es_den - 1]
    p = (gamma - 1.0) * rho * st[:, C.es_u - 1]
    st[:, C.es_p - 1] = p
    st[:, C.es_c - 1] = np.sqrt(gamma * p / rho)


def lattice_positions(dim, n_side, box, kind="square"):
    """Cell-centred lattice in ``box``; ``kind='staggered'`` (3D) shifts x by
    dx/2 on every other (y+z) row, a close-packed-like arrangement that tiles a
    periodic cube for even ``n_side``: x = (i + 1/4 + ((j+k) mod 2)/2) dx,
    y = (j + 1/2) dx, z = (k + 1/2) dx."""
    lo = np.array(box[0::2][:dim])
    hi = np.array(box[1::2][:dim])
    dx = (hi[0] - lo[0]) / n_side
    idx = np.indices((n_side,) * dim).reshape(dim, -1)[::-1].T.astype(np.float64)
    pos = np.zeros((idx.shape[0], 3))
    if kind == "staggered":
        assert dim == 3 and n_side % 2 == 0
        par = (idx[:, 1] + idx[:, 2]) % 2
        pos[:, 0] = lo[0] + (idx[:, 0] + 0.25 + 0.5 * par) * dx
        pos[:, 1] = lo[1] + (idx[:, 1] + 0.5) * dx
        pos[:, 2] = lo[2] + (idx[:, 2] + 0.5) * dx
    else:
        for a in range(dim):
            pos[:, a] = lo[a] + (idx[:, a] + 0.5) * dx
    return pos, dx


def hydro2d(n_side=64, seed=1, jitter=0.01, hfac=1.0, artts="yes", adden="yes",
            kernel=C.KERNEL_M6, au=0.0):
    """2D hydro on a jittered square lattice in periodic [-1,1]^2: rho=1, u=1,
    gamma=5/3, v=0, h=dx, m=4/N (the 64x64 one-leapfrog-step workload)."""
    box = (-1.0, 1.0, -1.0, 1.0, 0.0, 0.0)
    pos, dx = lattice_positions(2, n_side, box)
    n = pos.shape[0]
    g = gaussian(seed, 2 * n).reshape(n, 2)
    pos[:, :2] += jitter * dx * g
    st = empty_store(n)
    st[:, 0:3] = pos
    st[:, C.es_m - 1] = 4.0 / n
    st[:, C.es_den - 1] = 1.0
    st[:, C.es_h - 1] = dx * hfac
    st[:, C.es_u - 1] = 1.0
    gamma = 5.0 / 3.0
    _eos_fill(st, gamma)
    state = make_state(dim=2, equations="hydro", adden=adden, artts=artts,
                       hfac=hfac, box=box, spacing=dx, n_real=n, gamma=gamma,
                       kernel=kernel, au=au)
    return state, st


def hydro1d(n=64, seed=2, jitter=0.01, hfac=1.2, kernel=C.KERNEL_M6, velocity=0.05):
    """1D hydro on a jittered line in periodic [-1,1] with a sine velocity."""
    box = (-1.0, 1.0, 0.0, 0.0, 0.0, 0.0)
    pos, dx = lattice_positions(1, n, box)
    pos[:, 0] += jitter * dx * gaussian(seed, n)
    st = empty_store(n)
    st[:, 0:3] = pos
    st[:, C.es_vx - 1] = velocity * np.sin(math.pi * pos[:, 0])
    st[:, C.es_m - 1] = 2.0 / n
    st[:, C.es_den - 1] = 1.0
    st[:, C.es_h - 1] = dx * hfac
    st[:, C.es_u - 1] = 1.0
    gamma = 5.0 / 3.0
    _eos_fill(st, gamma)
    state = make_state(dim=1, equations="hydro", hfac=hfac, box=box, spacing=dx,
                       n_real=n, gamma=gamma, kernel=kernel, au=1.0)
    return state, st


def diffusion3d(n_side=128, hfac=1.0, ddw="fab", b=(1.0, 1.0, 0.0),
                disotropic="no", dcond=1.0, kernel=C.KERNEL_M6, kind="staggered",
                adden="yes", sts=True, h_scale=1.0):
    """3D anisotropic diffusion in the periodic unit cube: T = sin(2 pi x)
    sin(2 pi y) sin(2 pi z), B = b (field direction for the tensor), rho=1,
    h = hfac*dx, m = 1/N (the 2M-particle density+force workload at 128^3)."""
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    pos, dx = lattice_positions(3, n_side, box, kind=kind)
    n = pos.shape[0]
    st = empty_store(n)
    st[:, 0:3] = pos
    st[:, C.es_m - 1] = 1.0 / n
    st[:, C.es_den - 1] = 1.0
    st[:, C.es_h - 1] = dx * hfac * h_scale
    st[:, C.es_t - 1] = (np.sin(2 * math.pi * pos[:, 0]) * np.sin(2 * math.pi * pos[:, 1])
                         * np.sin(2 * math.pi * pos[:, 2]))
    st[:, C.es_kappa - 1] = 1.0
    st[:, C.es_bx - 1:C.es_bz] = np.asarray(b, dtype=np.float64)
    state = make_state(dim=3, equations="manual", ddw=ddw, adden=adden, artts="no",
                       disotropic=disotropic, dcond=dcond, hfac=hfac, box=box,
                       spacing=dx, n_real=n, kernel=kernel,
                       eqon=dict(diff=True, sts=sts))
    return state, st


def diffusion2d(n_side=32, hfac=1.0, ddw="fab", b=(1.0, 0.0, 0.0), disotropic="yes",
                dcond=1.0, kernel=C.KERNEL_M6, seed=None, jitter=0.0, adden="yes",
                h_scale=1.0):
    """2D diffusion in periodic [0,1]^2 with T = sin(2 pi x) sin(2 pi y)."""
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 0.0)
    pos, dx = lattice_positions(2, n_side, box)
    n = pos.shape[0]
    if seed is not None and jitter > 0:
        pos[:, :2] += jitter * dx * gaussian(seed, 2 * n).reshape(n, 2)
    st = empty_store(n)
    st[:, 0:3] = pos
    st[:, C.es_m - 1] = 1.0 / n
    st[:, C.es_den - 1] = 1.0
    st[:, C.es_h - 1] = dx * hfac * h_scale
    st[:, C.es_t - 1] = np.sin(2 * math.pi * pos[:, 0]) * np.sin(2 * math.pi * pos[:, 1])
    st[:, C.es_kappa - 1] = 1.0
    st[:, C.es_bx - 1:C.es_bz] = np.asarray(b, dtype=np.float64)
    state = make_state(dim=2, equations="diffusion", ddw=ddw, adden=adden, artts="no",
                       disotropic=disotropic, dcond=dcond, hfac=hfac, box=box,
                       spacing=dx, n_real=n, kernel=kernel)
    return state, st


def batch2d_specs(nprob=256, seed=11, n_min=1024, n_max=262144, ratios=(1.0, 0.1, 0.01, 0.001)):
    """A batched 2D anisotropic-diffusion study: problem k has
    N_k = round(exp(U[ln n_min, ln n_max])) particles on a periodic square
    lattice (n_side = round(sqrt(N_k))), field angle theta ~ U[0, pi) and
    kappa_perp / kappa_par drawn from ``ratios``.  Returns a list of dicts
    (n_side, theta, ratio)."""
    u = uniform(seed, 3 * nprob).reshape(nprob, 3)
    lo, hi = math.log(n_min), math.log(n_max)
    specs = []
    for k in range(nprob):
        nk = round(math.exp(lo + (hi - lo) * u[k, 0]))
        specs.append(dict(n_side=max(4, int(round(math.sqrt(nk)))), theta=math.pi * float(u[k, 1]),
                          ratio=float(ratios[min(int(u[k, 2] * len(ratios)), len(ratios) - 1)])))
    return specs


def batch2d_problem(spec):
    """One problem of ``batch2d_specs``: 2D diffusion in periodic [0,1]^2 with
    T = sin(2 pi x) sin(2 pi y), b = (cos theta, sin theta, 0), kappa_par = 1,
    the full tensor (kappa_perp = ratio).  Returns (state, store, K) with
    K = kappa_perp I + (kappa_par - kappa_perp) b b as (xx, yy, zz, xy, xz, yz)."""
    t, r = spec["theta"], spec["ratio"]
    b = (math.cos(t), math.sin(t), 0.0)
    state, st = diffusion2d(spec["n_side"], ddw="fab", b=b, disotropic="no", dcond=1.0)
    K = np.array([r + (1.0 - r) * b[0] * b[0], r + (1.0 - r) * b[1] * b[1], r,
                  (1.0 - r) * b[0] * b[1], 0.0, 0.0])
    return state, st, K


def random_slab3d(n, seed=3, box=(-1.0, 1.0, 0.0, 0.125, 0.0, 0.125), hfac=1.2,
                  equations="magnetohydro", b=(0.75, 1.0, 0.0), kernel=C.KERNEL_M6,
                  fld=False, adden="yes", artts="yes", au=1.0, process="fullyperiodic",
                  muzero=1.0, velocity_seed=None, lattice=None):
    """Uniform-random particles in a slab with a left/right shock-tube state
    (rho 1 | 0.125, P 1 | 0.1, E_rad 1 | 0.1, equal number density so m ~ rho),
    uniform B.  ``fld=True`` turns on diffusion+FLD (manual equation set, the
    only way the reference enables FLD in loop 2, state.f90:140-145).
    ``lattice=(nx, ny, nz)`` places the n = nx ny nz particles on a cell-centred
    lattice jittered by 0.05 of its spacing (Gaussian, ``seed``) instead."""
    lo = np.array(box[0::2])
    hi = np.array(box[1::2])
    if lattice is None:
        u = uniform(seed, 3 * n).reshape(n, 3)
    else:
        nl = np.array(lattice)
        assert int(np.prod(nl)) == n
        idx = np.indices(tuple(nl[::-1])).reshape(3, -1)[::-1].T.astype(np.float64)
        u = (idx + 0.5 + 0.05 * gaussian(seed, 3 * n).reshape(n, 3)) / nl
    pos = lo + (hi - lo) * u
    vol = float(np.prod(hi - lo))
    left = pos[:, 0] < 0.5 * (lo[0] + hi[0])
    rho = np.where(left, 1.0, 0.125)
    pres = np.where(left, 1.0, 0.1)
    erad = np.where(left, 1.0, 0.1)
    gamma = 5.0 / 3.0
    st = empty_store(n)
    st[:, 0:3] = pos
    st[:, C.es_m - 1] = rho * vol / n
    st[:, C.es_den - 1] = rho
    dxm = (vol / n) ** (1.0 / 3.0)
    st[:, C.es_h - 1] = hfac * dxm * (1.0 / rho) ** (1.0 / 3.0) * np.where(left, 1.0, 0.5)
    st[:, C.es_u - 1] = pres / ((gamma - 1.0) * rho)
    st[:, C.es_t - 1] = erad / rho
    st[:, C.es_kappa - 1] = 1.0
    st[:, C.es_bx - 1:C.es_bz] = np.asarray(b, dtype=np.float64)
    if velocity_seed is not None:
        st[:, C.es_vx - 1:C.es_vz] = 0.1 * gaussian(velocity_seed, 3 * n).reshape(n, 3)
    _eos_fill(st, gamma)
    eqon = None
    if fld:
        equations = "manual"
        eqon = dict(diff=True, fld=True, hydro=True)
    state = make_state(dim=3, equations=equations, adden=adden, artts=artts, au=au,
                       hfac=hfac, box=tuple(box), spacing=dxm, n_real=n, gamma=gamma,
                       kernel=kernel, process=process, eqon=eqon, muzero=muzero,
                       bordersize=(int(C.KRAD[kernel] * hfac) + 1) * dxm * 2.0)
    return state, st


def fixed_walls(state, st, side=C.ebc_all, return_source=False):
    """Append fixed wall particles to ``st`` as ``createFixedBorders(store,
    side)`` does (BC.f90:477-640, after findInsideBorderParticles,
    BC.f90:196-306): every real or fixed particle strictly inside the band
    (emin, emin + bordersize) of a walled axis is mirrored to
    emin + (emin - x) (and likewise at emax), x walls first, then y, then z,
    with a mirrored particle that lies in the y or z band mirrored again by
    the later axis (the corner walls).  Only positions and the type are set
    (the reference leaves every other field to the set-up); the new rows copy
    nothing else and are zero.  Returns (state, store) with ec_fixedpn
    updated, and with ``return_source`` also the index of the particle each
    new row mirrors (traced back to a particle of ``st``)."""
    state = np.array(state, dtype=np.float64, copy=True)
    dim = int(state[C.ec_dim - 1])
    bs = state[C.ec_bordsize - 1]
    emin = np.array([state[C.ec_xmin - 1], state[C.ec_ymin - 1], state[C.ec_zmin - 1]])
    emax = np.array([state[C.ec_xmax - 1], state[C.ec_ymax - 1], state[C.ec_zmax - 1]])
    bmin, bmax = emin + bs, emax - bs
    n0 = st.shape[0]
    types = st[:, C.es_type - 1].astype(int)
    live = np.isin(types, (C.ept_real, C.ept_fixed, C.ept_fixedreal))
    pos = st[:, 0:3].copy()
    # findInsideBorderParticles: per axis and end, the sources in index order
    lists = {}
    for a in range(3):
        lo = live & (pos[:, a] < bmin[a]) & (pos[:, a] > emin[a])
        hi = live & ~(pos[:, a] < bmin[a]) & (pos[:, a] > bmax[a]) & (pos[:, a] < emax[a])
        lists[(a, 1)] = list(np.nonzero(lo)[0]) if a < dim else []
        lists[(a, 2)] = list(np.nonzero(hi)[0]) if a < dim else []
    blocks, origin_blocks = [pos], [np.arange(n0)]
    origin = np.arange(n0)
    total = n0

    def walls(a, end):
        nonlocal total, origin
        src = np.array(lists[(a, end)], dtype=np.int64)
        if src.size == 0:
            return
        allpos = np.concatenate(blocks)
        q = allpos[src].copy()
        q[:, a] = emax[a] + (emax[a] - q[:, a]) if end == 2 else emin[a] + (emin[a] - q[:, a])
        idx = np.arange(total, total + src.size)
        blocks.append(q)
        org = np.concatenate(origin_blocks)[src]
        origin_blocks.append(org)
        total += src.size
        # a new particle in a later axis' band joins that axis' list (corner walls)
        for b in range(a + 1, dim):
            lo = q[:, b] < bmin[b]
            hi = ~lo & (q[:, b] > bmax[b])
            lists[(b, 1)].extend(idx[lo].tolist())
            lists[(b, 2)].extend(idx[hi].tolist())

    if side in (C.ebc_all, C.ebc_x):
        walls(0, 1)
        walls(0, 2)
    if dim > 1:
        if side in (C.ebc_all, C.ebc_y, C.ebc_y1):
            walls(1, 1)
        if side in (C.ebc_all, C.ebc_y, C.ebc_y2):
            walls(1, 2)
        if dim == 3 and side in (C.ebc_all, C.ebc_z):
            walls(2, 1)
            walls(2, 2)
    new = np.concatenate(blocks[1:]) if len(blocks) > 1 else np.zeros((0, 3))
    extra = np.zeros((new.shape[0], st.shape[1]))
    extra[:, 0:3] = new
    extra[:, C.es_type - 1] = C.ept_fixed
    out = np.concatenate([st, extra])
    state[C.ec_fixedpn - 1] = (out.shape[0] - int(state[C.ec_realpn - 1]))
    if return_source:
        return state, out, np.concatenate(origin_blocks[1:]) if len(blocks) > 1 else np.zeros(0, np.int64)
    return state, out


def hydro2d_mti_walls(n_side=16, seed=1, jitter=0.01, hfac=1.0, kernel=C.KERNEL_M6, au=1.0,
                      velocity_seed=None):
    """The 2D hydro lattice of ``hydro2d`` run as the reference's legacy mti
    set-up (IC/IC.f90:149-174, iterator.f90:134-143): periodic along x,
    fixed walls below and above y (createFixedBorders(store, ebc_y)), gravity
    a_y -= 1.  The wall particles carry the real particles' mass, density,
    h, u, p and c (the set-up initialises every real and fixed particle) and
    no velocity."""
    state, st = hydro2d(n_side, seed=seed, jitter=jitter, hfac=hfac, kernel=kernel, au=au)
    if velocity_seed is not None:
        n = st.shape[0]
        st[:, C.es_vx - 1:C.es_vy] = 0.05 * gaussian(velocity_seed, 2 * n).reshape(n, 2)
    state[C.ec_process - 1] = C.epc_backcompatibility
    state[C.ec_ics - 1] = C.ett_mti
    state, full, src = fixed_walls(state, st, C.ebc_y, return_source=True)
    nr = st.shape[0]
    for f in (C.es_m, C.es_den, C.es_h, C.es_u, C.es_p, C.es_c, C.es_om):
        full[nr:, f - 1] = st[src, f - 1]
    return state, full


def fld2d_gauss_walls(n_side=20, hfac=1.2, kernel=C.KERNEL_M6, ddw="fab"):
    """The reference's fld_gauss set-up in 2D (IC/setups/place_fld_gauss.f90,
    init_fld_gauss.f90): a uniform lattice in [-10, 10]^2 with fixed walls on
    every side (createFixedBorders(store, ebc_all)), no periodic images and no
    EOS (iterator.f90:88-91); rho = 1, P = 1, gamma = 2, opacity 0.1 and
    xi = T a Gaussian pulse (width 5/3 on this box, plus a floor of 1e-3),
    with diffusion + flux-limited radiation on (manual equation set).  Every
    particle, real or fixed, gets the set-up's fields."""
    box = (-10.0, 10.0, -10.0, 10.0, 0.0, 0.0)
    pos, dx = lattice_positions(2, n_side, box)
    n = pos.shape[0]
    st = empty_store(n)
    st[:, 0:3] = pos
    gamma = 2.0
    state = make_state(dim=2, equations="manual", ddw=ddw, adden="yes", artts="no", hfac=hfac, box=box,
                       spacing=dx, n_real=n, gamma=gamma, kernel=kernel, process="backcompatibility",
                       ics=C.ett_fld_gauss, eqon=dict(diff=True, fld=True))
    state, full = fixed_walls(state, st, C.ebc_all)
    p = full[:, 0:3]
    full[:, C.es_h - 1] = hfac * dx
    full[:, C.es_m - 1] = dx ** 2
    full[:, C.es_den - 1] = 1.0
    full[:, C.es_p - 1] = 1.0
    r2 = (p[:, 0] ** 2 + p[:, 1] ** 2) / 25.0
    full[:, C.es_t - 1] = np.exp(-0.5 * r2 / (1.0 / 3.0) ** 2) + 1e-3
    full[:, C.es_u - 1] = 1.0 / (gamma - 1.0)
    full[:, C.es_kappa - 1] = 0.1
    full[:, C.es_om - 1] = 1.0
    return state, full


def shock_slab_walls(n, seed=3, **kw):
    """``random_slab3d`` run as the reference's hydroshock set-up
    (IC/IC.f90:227-253, iterator.f90:102-112): periodic in y and z, fixed
    walls beyond both x ends (createFixedBorders(store, ebc_x)).  A wall
    particle carries the fields of the particle it mirrors (the left or right
    state of its side) and no velocity."""
    kw.setdefault("process", "backcompatibility")
    state, st = random_slab3d(n, seed=seed, **kw)
    state[C.ec_ics - 1] = C.ett_hydroshock
    state, full, src = fixed_walls(state, st, C.ebc_x, return_source=True)
    for f in range(4, C.es_total + 1):
        if f in (C.es_vx, C.es_vy, C.es_vz, C.es_type):
            continue
        full[n:, f - 1] = st[src, f - 1]
    return state, full


def shock_tube3d(n=1048576, half="mhd", seed=3, hfac=1.0, velocity_seed=None):
    """The 3D MHD + flux-limited-radiation shock tube: n uniform-random
    particles (``seed``) in [-1,1] x [0,0.125]^2, periodic in y and z, left
    rho = 1, P = 1, E_rad = 1, right rho = 0.125, P = 0.1, E_rad = 0.1,
    B = (0.75, 1, 0), opacity 1, Levermore-Pomraning limiter, run as the
    reference's hydroshock set-up (fixed walls beyond both x ends,
    IC/IC.f90:227-253, iterator.f90:102-112).  No reference equation set has
    MHD and FLD together (state.f90:117-149), so the workload runs as two
    halves: ``half='mhd'`` (magnetohydro: hydro + AV + MHD) and ``half='fld'``
    (manual: hydro + AV + diffusion with the FLD conductivity)."""
    if half == "mhd":
        return shock_slab_walls(n, seed=seed, hfac=hfac, equations="magnetohydro", velocity_seed=velocity_seed)
    if half == "fld":
        return shock_slab_walls(n, seed=seed, hfac=hfac, fld=True, velocity_seed=velocity_seed)
    raise ValueError("half must be 'mhd' or 'fld'")


def energy_budget(state, store):
    """Energies of the real particles (code units): kinetic sum m |v|^2 / 2,
    thermal sum m u, magnetic sum m |B|^2 / (2 rho mu0) when the equation set
    has MHD, radiation sum m xi (xi = E_rad / rho, the T row) when it has FLD;
    ``total`` is their sum.  The quantities whose rates loop 2's energy sum
    adds up (circuit2.f90:490-492)."""
    nr = int(state[C.ec_realpn - 1])
    s = store[:nr]
    m = s[:, C.es_m - 1]
    v = s[:, C.es_vx - 1:C.es_vz]
    out = {"kinetic": float(np.sum(0.5 * m * np.sum(v * v, axis=1))),
           "thermal": float(np.sum(m * s[:, C.es_u - 1]))}
    eq = int(state[C.ec_eqs - 1])
    if eq in (C.eeq_magnetohydro, C.eeq_magnetohydrodiffusion):
        b = s[:, C.es_bx - 1:C.es_bz]
        out["magnetic"] = float(np.sum(0.5 * m * np.sum(b * b, axis=1) / (s[:, C.es_den - 1]
                                                                          * state[C.ec_muzero - 1])))
    if eq == C.eeq_manual and int(state[C.ec_eqonfluxlim - 1]) == 1:
        out["radiation"] = float(np.sum(m * s[:, C.es_t - 1]))
    out["total"] = float(sum(out.values()))
    return out


def anisotropic_decay_exact(pos, t, kpar=1.0, kperp=1e-3):
    """Exact solution of dT/dt = div(K grad T), K = kpar b b + kperp (I - b b),
    b = (1, 1, 0)/sqrt(2), from T0 = sin(2 pi x) sin(2 pi y) sin(2 pi z) in
    the periodic unit cube.  T0 = (cos 2pi(x-y) - cos 2pi(x+y)) sin(2 pi z) / 2,
    and each term is an eigenmode: k = 2pi(1, -1, 1) has k.b = 0, rate
    3 kperp (2 pi)^2; k = 2pi(1, 1, 1) has (k.b)^2 = 2 (2 pi)^2, rate
    (2 kpar + kperp) (2 pi)^2."""
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    w2 = (2.0 * math.pi) ** 2
    return 0.5 * (np.cos(2 * math.pi * (x - y)) * math.exp(-3.0 * kperp * w2 * t)
                  - np.cos(2 * math.pi * (x + y)) * math.exp(-(2.0 * kpar + kperp) * w2 * t)) \
        * np.sin(2 * math.pi * z)
