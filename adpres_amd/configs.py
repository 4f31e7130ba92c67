"""Seeded synthetic static cores (the workloads the engine is measured on).

Each generator returns a :class:`GeneratedCore`: the :class:`Problem` handed
to the engine plus the deck data (materials, assembly layout) from which the
reference build reads the same core.  Node-wise perturbations that no deck
can express are carried by ``Problem`` and written for the reference as a
per-node cross-section override file.

Cores
-----
``core2d_small``  2D 2-group, 9x9 assemblies of 20 cm with 2x2 nodes
                  (18x18x2 nodes), 4 fuel types in a radius-4-assembly disk,
                  reflector ring, vacuum (zero incoming current) outside.
``core3d_2g``     3D 2-group quarter core, 17x17 assemblies of 21.5 cm with
                  2x2 nodes, 24 fuel planes of 15 cm + 1 reflector plane top
                  and bottom (34x34x26 nodes), 8 fuel types, per-node +-2%
                  perturbation of every cross section, reflective on the
                  quarter-core faces (west, south).
``core3d_8g``     3D 8-group quarter core, same layout with 4x4 radial nodes
                  and 24 axial layers of 15 cm split in 4 (68x68x96 nodes),
                  full downscatter plus upscatter into groups 6-8, nu-sigf
                  rescaled per fuel material to a sampled k-infinity.
``batch_cores``   the batching workload: 256 uneven cores (2D and 3D, 2-4
                  groups, 5k-60k nodes, each with its own nodal-update
                  interval) with instance-wide and per-node cross-section
                  perturbations (see batch_core).
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .problem import Problem, build_geometry, node_xs_from_materials, round_deck

NU = 2.43          # nu-sigf / sigf, i.e. fixed kappa/nu for the power


@dataclasses.dataclass
class GeneratedCore:
    problem: Problem
    deck: dict                   # keyword arguments of problem.write_deck
    perturbed: bool              # True if node XS differ from the material tables
    seed: int


def _mk_problem(name, ng, mats, geo, bc, chi, perturb_rng=None, perturb=0.0, **iter_kw):
    g = build_geometry(**geo)
    xsigtr, xsiga, xnuf, xsigf, xsigs = (round_deck(a) for a in mats)
    chi = round_deck(chi)
    sigtr, siga, nuf, sigf, sigs = node_xs_from_materials(g["mat"], xsigtr, xsiga, xnuf, xsigf, xsigs)
    if perturb_rng is not None and perturb > 0:
        def pert(a):
            return a * perturb_rng.uniform(1.0 - perturb, 1.0 + perturb, size=a.shape)
        sigtr, siga, nuf, sigf, sigs = pert(sigtr), pert(siga), pert(nuf), pert(sigf), pert(sigs)
    nmat = len(xsigtr)
    chi_arr = np.ascontiguousarray(np.asarray(chi, dtype=np.float64).T)   # (G, nmat)
    p = Problem(ng=ng, nmat=nmat, nxx=g["nxx"], nyy=g["nyy"], nzz=g["nzz"],
                xdel=g["xdel"], ydel=g["ydel"], zdel=g["zdel"],
                ystag_min=g["ystag_min"], ystag_max=g["ystag_max"],
                xstag_min=g["xstag_min"], xstag_max=g["xstag_max"], bc=tuple(bc),
                ix=g["ix"], iy=g["iy"], iz=g["iz"], mat=g["mat"],
                sigtr=np.ascontiguousarray(sigtr), siga=np.ascontiguousarray(siga),
                nuf=np.ascontiguousarray(nuf), sigf=np.ascontiguousarray(sigf),
                sigs=np.ascontiguousarray(sigs), chi=chi_arr,
                dc=np.ones((6, ng, len(g["ix"]))), name=name, **iter_kw)
    deck = dict(ng=ng, xsigtr=xsigtr, xsiga=xsiga, xnuf=xnuf, xsigf=xsigf, xchi=chi,
                xsigs=xsigs, bc=bc, **geo)
    return p, deck


def _two_group_material(rng, fuel: bool):
    D1 = rng.uniform(1.2, 1.6); D2 = rng.uniform(0.3, 0.45)
    sa1 = rng.uniform(0.008, 0.012)
    sa2 = rng.uniform(0.06, 0.12) if fuel else rng.uniform(0.02, 0.04)
    nf1 = rng.uniform(0.004, 0.008) if fuel else 0.0
    nf2 = rng.uniform(0.09, 0.16) if fuel else 0.0
    s12 = rng.uniform(0.015, 0.02)
    sigtr = [1.0 / (3.0 * D1), 1.0 / (3.0 * D2)]
    return (sigtr, [sa1, sa2], [nf1, nf2], [nf1 / NU, nf2 / NU], [[0.0, s12], [0.0, 0.0]])


def _stack(mats):
    return tuple(np.asarray([m[i] for m in mats], dtype=np.float64) for i in range(5))


def core2d_small(seed: int = 1, **iter_kw) -> GeneratedCore:
    rng = np.random.default_rng(seed)
    nfuel = 4
    mats = [_two_group_material(rng, True) for _ in range(nfuel)]
    mats.append(_two_group_material(rng, False))
    refl = nfuel + 1
    nx = 9
    pl = np.full((nx, nx), refl, dtype=np.int32)
    c = (nx + 1) / 2.0
    for j in range(1, nx + 1):
        for i in range(1, nx + 1):
            if (i - c) ** 2 + (j - c) ** 2 <= 16.0:
                pl[j - 1, i - 1] = rng.integers(1, nfuel + 1)
    geo = dict(xsize=[20.0] * nx, xdiv=[2] * nx, ysize=[20.0] * nx, ydiv=[2] * nx,
               zsize=[10.0, 10.0], zdiv=[1, 1], planars=[pl], zpln=[1, 1])
    chi = [[1.0, 0.0]] * (nfuel + 1)
    bc = (1, 1, 1, 1, 2, 2)
    p, deck = _mk_problem("core2d_small", 2, _stack(mats), geo, bc, chi, **iter_kw)
    return GeneratedCore(p, deck, False, seed)


def _quarter_layout(rng, nx, nfuel, radius):
    refl = nfuel + 1
    pl = np.full((nx, nx), refl, dtype=np.int32)
    for j in range(1, nx + 1):
        for i in range(1, nx + 1):
            if (i - 0.5) ** 2 + (j - 0.5) ** 2 <= radius ** 2:
                pl[j - 1, i - 1] = rng.integers(1, nfuel + 1)
    return pl


def core3d_2g(seed: int = 1, perturb: float = 0.02, **iter_kw) -> GeneratedCore:
    rng = np.random.default_rng(seed)
    nfuel = 8
    mats = [_two_group_material(rng, True) for _ in range(nfuel)]
    mats.append(_two_group_material(rng, False))
    refl = nfuel + 1
    nx = 17
    fuel_pl = _quarter_layout(rng, nx, nfuel, 14.5)
    refl_pl = np.full((nx, nx), refl, dtype=np.int32)
    nz = 26
    geo = dict(xsize=[21.5] * nx, xdiv=[2] * nx, ysize=[21.5] * nx, ydiv=[2] * nx,
               zsize=[15.0] * nz, zdiv=[1] * nz, planars=[refl_pl, fuel_pl],
               zpln=[1] + [2] * (nz - 2) + [1])
    chi = [[1.0, 0.0]] * (nfuel + 1)
    bc = (1, 2, 1, 2, 1, 1)
    kw = dict(nout=3000, nin=20, serc=1e-5, ferc=1e-5, nac=1000, nupd=22)
    kw.update(iter_kw)
    p, deck = _mk_problem("core3d_2g", 2, _stack(mats), geo, bc, chi,
                          perturb_rng=np.random.default_rng(seed + 1000), perturb=perturb, **kw)
    return GeneratedCore(p, deck, perturb > 0, seed)


def kinf(nuf, chi, sigr, sigs_from_to):
    """k-infinity of one material: nuf^T M^-1 chi with
    M(g,h) = delta_gh sigr_g - (1-delta_gh) sigs(h->g)."""
    G = len(nuf)
    M = -np.asarray(sigs_from_to, dtype=np.float64).T.copy()
    for g in range(G):
        M[g, g] = sigr[g]
    return float(np.asarray(nuf) @ np.linalg.solve(M, np.asarray(chi)))


def _eight_group_material(rng, fuel: bool, G: int = 8):
    D = np.sort(np.exp(rng.uniform(np.log(0.3), np.log(2.0), G)))[::-1]
    sa = np.sort(np.exp(rng.uniform(np.log(1e-3), np.log(0.15), G)))
    if not fuel:
        sa = sa * 0.3
    sigs = np.zeros((G, G))
    for g in range(G):
        for h in range(g + 1, G):
            sigs[g, h] = rng.uniform(0.002, 0.03) * 0.5 ** (h - g - 1)
        if g >= 6:
            for h in range(5, g):
                sigs[g, h] = rng.uniform(0.0005, 0.003)
    nf = np.sort(rng.uniform(0.0, 0.2, G)) if fuel else np.zeros(G)
    chi = np.zeros(G)
    chi[:3] = rng.dirichlet([6.0, 3.0, 1.0])
    sigr = sa + sigs.sum(axis=1) - np.diag(sigs)
    if fuel:
        target = rng.uniform(1.0, 1.2)
        nf = nf * (target / kinf(nf, chi, sigr, sigs))
    sigtr = 1.0 / (3.0 * D)
    return (list(sigtr), list(sa), list(nf), list(nf / NU), sigs.tolist()), list(chi)


def core3d_8g(seed: int = 2, **iter_kw) -> GeneratedCore:
    rng = np.random.default_rng(seed)
    nfuel = 8
    mats, chis = [], []
    for m in range(nfuel + 1):
        mt, ch = _eight_group_material(rng, m < nfuel)
        mats.append(mt); chis.append(ch)
    chis[-1] = chis[0]
    refl = nfuel + 1
    nx = 17
    fuel_pl = _quarter_layout(rng, nx, nfuel, 14.5)
    refl_pl = np.full((nx, nx), refl, dtype=np.int32)
    nz = 24
    geo = dict(xsize=[21.5] * nx, xdiv=[4] * nx, ysize=[21.5] * nx, ydiv=[4] * nx,
               zsize=[15.0] * nz, zdiv=[4] * nz, planars=[refl_pl, fuel_pl],
               zpln=[1] + [2] * (nz - 2) + [1])
    bc = (1, 2, 1, 2, 1, 1)
    kw = dict(nout=5000, nin=20, serc=1e-5, ferc=1e-5, nac=1000, nupd=93)
    kw.update(iter_kw)
    p, deck = _mk_problem("core3d_8g", 8, _stack(mats), geo, bc, chis, **kw)
    return GeneratedCore(p, deck, False, seed)


def _few_group_material(rng, fuel: bool, G: int):
    """The 2-group material of the small cores for G = 2; for G = 3, 4 its
    fast group split into G - 1 groups (D decreasing in [1.2, 2.0], small
    absorption, a down-scatter chain) above the same thermal group."""
    if G == 2:
        return _two_group_material(rng, fuel)
    D = list(np.sort(rng.uniform(1.2, 2.0, G - 1))[::-1]) + [rng.uniform(0.3, 0.45)]
    sa = list(rng.uniform(0.002, 0.012, G - 1)) + [rng.uniform(0.06, 0.12) if fuel else rng.uniform(0.02, 0.04)]
    nf = (list(rng.uniform(0.001, 0.008, G - 1)) + [rng.uniform(0.09, 0.16)]) if fuel else [0.0] * G
    sigs = np.zeros((G, G))
    for g in range(G - 1):
        sigs[g, g + 1] = rng.uniform(0.015, 0.05) if g < G - 2 else rng.uniform(0.015, 0.02)
    sigtr = [1.0 / (3.0 * d) for d in D]
    return (sigtr, sa, nf, [x / NU for x in nf], sigs.tolist())


def batch_core(seed: int, index: int, nin: int = 20) -> GeneratedCore:
    """Instance `index` of the seeded batch of uneven static cores: 2D (two
    10 cm planes, reflective top and bottom; 17-20 assemblies across) or 3D
    (9-17 assemblies across, 6-24 fuel planes of 15 cm between two reflector
    planes), 2-3 radial nodes per 21.5 cm assembly, 2-4 groups, 5k-60k nodes,
    quarter-core layout (reflective west/south), SANM with the reference's
    default nodal-update interval ceil((nxx+nyy+nzz)/2.5).  Cross sections:
    the 2-group ranges of the small cores (split fast groups for G > 2), an
    instance-wide +-5% factor on the thermal absorption and thermal
    nu-sigf, and per-node multipliers U[0.97, 1.03] on every cross section."""
    rng = np.random.default_rng([seed, index])
    G = int(rng.integers(2, 5))
    three_d = bool(rng.integers(0, 2))
    target = float(np.exp(rng.uniform(np.log(5000.0), np.log(60000.0))))
    if three_d:
        shapes = [(nx, d, nzf) for nx in range(9, 18) for d in range(2, 4) for nzf in range(6, 25)]
        nodes = np.array([(nx * d) ** 2 * (nzf + 2) for nx, d, nzf in shapes], float)
    else:
        shapes = [(nx, d, 0) for nx in range(17, 21) for d in range(2, 4)]
        nodes = np.array([(nx * d) ** 2 * 2 for nx, d, _ in shapes], float)
    ok = (nodes >= 5000) & (nodes <= 60000)
    near = np.nonzero(ok & (np.abs(nodes - target) <= 0.15 * target))[0]
    if len(near) == 0:
        near = [int(np.argmin(np.where(ok, np.abs(nodes - target), np.inf)))]
    nx, ndiv, nzf = shapes[int(rng.choice(near))]
    nfuel = int(rng.integers(4, 9))
    mats = [_few_group_material(rng, True, G) for _ in range(nfuel)]
    mats.append(_few_group_material(rng, False, G))
    fa, ff = rng.uniform(0.95, 1.05, 2)
    for m in mats[:nfuel]:
        m[1][G - 1] *= fa
        m[2][G - 1] *= ff
        m[3][G - 1] = m[2][G - 1] / NU
    refl = nfuel + 1
    fuel_pl = _quarter_layout(rng, nx, nfuel, nx - 2.5)
    if three_d:
        refl_pl = np.full((nx, nx), refl, dtype=np.int32)
        nz = nzf + 2
        geo = dict(xsize=[21.5] * nx, xdiv=[ndiv] * nx, ysize=[21.5] * nx, ydiv=[ndiv] * nx,
                   zsize=[15.0] * nz, zdiv=[1] * nz, planars=[refl_pl, fuel_pl], zpln=[1] + [2] * nzf + [1])
        bc = (1, 2, 1, 2, 1, 1)
    else:
        nz = 2
        geo = dict(xsize=[21.5] * nx, xdiv=[ndiv] * nx, ysize=[21.5] * nx, ydiv=[ndiv] * nx,
                   zsize=[10.0, 10.0], zdiv=[1, 1], planars=[fuel_pl], zpln=[1, 1])
        bc = (1, 2, 1, 2, 2, 2)
    chi = [[1.0] + [0.0] * (G - 1)] * (nfuel + 1)
    nupd = int(np.ceil((2 * nx * ndiv + nz) / 2.5))
    kw = dict(nout=3000, nin=nin, serc=1e-5, ferc=1e-5, nac=1000, nupd=nupd)
    p, deck = _mk_problem(f"batch_s{seed}_{index}", G, _stack(mats), geo, bc, chi,
                          perturb_rng=np.random.default_rng([seed, index, 1]), perturb=0.03, **kw)
    return GeneratedCore(p, deck, True, seed)


def batch_cores(seed: int = 3, n: int = 256):
    """The batching workload: `n` uneven cores of batch_core."""
    return [batch_core(seed, i) for i in range(n)]


def small_multigroup_core(G: int, seed: int) -> GeneratedCore:
    """A 10x10x12-node core of G-group materials (the 8-group core's material
    generator at G groups: three fuel materials in a quarter layout of
    radius 4.5 assemblies, a reflector around them and above and below),
    for the cooperative nodal kernels' instances of 5 to 7 groups."""
    rng = np.random.default_rng(seed)
    nfuel, nx, nz = 3, 5, 6
    mats, chis = [], []
    for m in range(nfuel + 1):
        mt, ch = _eight_group_material(rng, m < nfuel, G)
        mats.append(mt)
        chis.append(ch)
    chis[-1] = chis[0]
    fuel_pl = _quarter_layout(rng, nx, nfuel, 4.5)
    refl_pl = np.full((nx, nx), nfuel + 1, dtype=np.int32)
    geo = dict(xsize=[21.5] * nx, xdiv=[2] * nx, ysize=[21.5] * nx, ydiv=[2] * nx,
               zsize=[20.0] * nz, zdiv=[2] * nz, planars=[refl_pl, fuel_pl], zpln=[1] + [2] * (nz - 2) + [1])
    p, deck = _mk_problem(f"small_{G}g", G, _stack(mats), geo, (1, 2, 1, 2, 1, 1), chis)
    return GeneratedCore(p, deck, False, seed)


CORES = {"core2d_small": core2d_small, "core3d_2g": core3d_2g, "core3d_8g": core3d_8g}


def feedback_tables(gc: GeneratedCore, seed: int = 7, nb: int = 4):
    """Feedback tables in the form of the reference's XS_updt
    (src/mod_xsec.f90:11-46) for a generated core: its material cross
    sections as the base tables; boron, fuel-temperature, moderator-
    temperature and coolant-density coefficients that move each entry by up
    to 1e-5 per ppm, 1e-3 per sqrt(K), 1e-3 per K and 5e-2 per g/cc of its
    base value; nb control-rod banks over random full-height columns, each
    adding 2 % to sigtr and 10 % to siga.  Returns (tables, parameters)
    where the parameters (bcon, ftem, mtem, cden, bpos) are a typical state
    about the reference values."""
    p, d = gc.problem, gc.deck
    rng = np.random.default_rng(seed)
    base = [np.ascontiguousarray(np.asarray(d[k], dtype=np.float64).T) for k in ("xsigtr", "xsiga", "xnuf", "xsigf")]
    base.append(np.ascontiguousarray(np.transpose(np.asarray(d["xsigs"], dtype=np.float64), (2, 1, 0))))
    coreh = float(np.sum(p.zdel))
    t = dict(nmat=p.nmat, nb=nb, rbcon=1000.0, rftem=900.0, rmtem=560.0, rcden=0.75, coreh=coreh, pos0=0.0,
             ssize=coreh / 200.0, base=base)
    for kind, scale in (("boron", 1e-5), ("ftem", 1e-3), ("mtem", 1e-3), ("cden", 5e-2)):
        t[kind] = [b * rng.uniform(-scale, scale, b.shape) for b in base]
    t["rod"] = [0.02 * base[0], 0.1 * base[1]] + [np.zeros_like(b) for b in base[2:]]
    cols = np.zeros((p.nyy, p.nxx), np.int64)
    np.add.at(cols, (p.iy - 1, p.ix - 1), 1)
    full = np.argwhere(cols == p.nzz)
    fb = np.zeros((p.nyy, p.nxx), np.int32)
    pick = full[rng.permutation(len(full))[:max(nb, len(full) // 8)]]
    fb[pick[:, 0], pick[:, 1]] = 1 + np.arange(len(pick)) % nb
    t["fbmap"] = fb
    N = p.nnod
    par = (1200.0, 900.0 * rng.uniform(0.8, 1.4, N), 560.0 + rng.uniform(-5.0, 15.0, N),
           0.75 * rng.uniform(0.97, 1.02, N), rng.uniform(0.0, 200.0, nb))
    return t, par


def generate(name: str, seed: Optional[int] = None, **kw) -> GeneratedCore:
    fn = CORES[name]
    return fn(seed, **kw) if seed is not None else fn(**kw)


def branch_tables(gc: GeneratedCore, seed: int = 9, nb: int = 4):
    """Branch tables in the flat form of the reference's XStab_updt
    (src/mod_xsec.f90:48-84) for a generated core: three branch points of
    coolant density, boron, fuel and moderator temperature per material,
    each record the material's cross sections scaled by up to +-3 % per
    parameter step (unit discontinuity factors scaled the same way), and a
    rodded set (2 % more sigtr, 10 % more siga) for every material, with
    feedback_tables' rod banks.  Returns (tables, parameters) with the
    parameters inside the branch ranges."""
    p, d = gc.problem, gc.deck
    rng = np.random.default_rng(seed)
    G, M = p.ng, p.nmat
    pts = [np.array([0.66, 0.71, 0.76]), np.array([0.0, 1000.0, 2000.0]), np.array([560.0, 900.0, 1320.0]),
           np.array([550.0, 570.0, 590.0])]
    recs, rrecs = [], []
    for m in range(M):
        b = np.concatenate([np.asarray(d[k], dtype=np.float64)[m] for k in ("xsigtr", "xsiga", "xnuf", "xsigf")]
                           + [np.asarray(d["xsigs"], dtype=np.float64)[m].ravel(order="F")]    # (from, to), from fastest
                           + [np.ones(6 * G)])
        slope = rng.uniform(-0.03, 0.03, (4, b.size))
        for v in range(3):
            for u in range(3):
                for t in range(3):
                    for s_ in range(3):
                        f = 1.0 + slope[0] * (s_ - 1) + slope[1] * (t - 1) + slope[2] * (u - 1) + slope[3] * (v - 1)
                        r = b * f
                        recs.append(r)
                        rr = r.copy()
                        rr[:G] *= 1.02
                        rr[G:2 * G] *= 1.10
                        rrecs.append(rr)
    ft, _ = feedback_tables(gc, seed, nb)
    t = {"nmat": M, "nb": nb, "dims": np.tile(np.array([3, 3, 3, 3, 1], np.int32), (M, 1)),
         "points": np.concatenate([np.concatenate(pts)] * M), "xsec": np.concatenate(recs),
         "rxsec": np.concatenate(rrecs), "coreh": ft["coreh"], "pos0": 0.0, "ssize": ft["ssize"],
         "fbmap": ft["fbmap"]}
    N = p.nnod
    par = (1200.0, rng.uniform(600.0, 1300.0, N), rng.uniform(552.0, 588.0, N), rng.uniform(0.67, 0.75, N),
           rng.uniform(0.0, 200.0, nb))
    return t, par
