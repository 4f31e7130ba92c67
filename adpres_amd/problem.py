"""Static-core problem description on the host side of the C-ABI.

A :class:`Problem` carries exactly the data the reference's CMFD/nodal hot
path reads from its global state module (``src/mod_data.f90``): node
geometry, the staggered radial footprint, boundary codes, node-wise cross
sections and the iteration controls of the ``%ITER`` and ``%KERN`` cards.

Array convention.  The reference stores Fortran column-major arrays such as
``sigtr(nnod, ng)`` or ``sigs(nnod, ng, ng)``.  Here every such array is a
C-contiguous numpy array whose shape is the Fortran shape reversed
(``sigtr[g, n]``, ``sigs[to, from, n]``, ``dc[face, g, n]``), so its memory is
byte-identical to the Fortran array and can be handed across the C-ABI (or
written for the reference) without any transposition.  Indices stored in
arrays (``ix``, ``iy``, ``iz``, ``mat``, stagger bounds) are 1-based, as in
the reference.
"""

from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

KERN_IDS = {"FDM": 0, "PNM": 1, "SANM": 2}
KERN_NAMES = {v: k for k, v in KERN_IDS.items()}
BC_NAMES = ("east", "west", "north", "south", "bottom", "top")


@dataclasses.dataclass
class Problem:
    ng: int
    nmat: int
    nxx: int
    nyy: int
    nzz: int
    xdel: np.ndarray            # (nxx,) cm
    ydel: np.ndarray            # (nyy,)
    zdel: np.ndarray            # (nzz,)
    ystag_min: np.ndarray       # (nyy,) first i of row j (1-based)
    ystag_max: np.ndarray       # (nyy,)
    xstag_min: np.ndarray       # (nxx,) first j of column i
    xstag_max: np.ndarray       # (nxx,)
    bc: Tuple[int, int, int, int, int, int]   # east, west, north, south, bottom, top
    ix: np.ndarray              # (N,) 1-based
    iy: np.ndarray
    iz: np.ndarray
    mat: np.ndarray             # (N,) 1-based material of node
    sigtr: np.ndarray           # (G, N)
    siga: np.ndarray            # (G, N)
    nuf: np.ndarray             # (G, N)
    sigf: np.ndarray            # (G, N)
    sigs: np.ndarray            # (G_to, G_from, N)
    chi: np.ndarray             # (G, nmat)
    dc: np.ndarray              # (6, G, N)  assembly discontinuity factors
    exsrc: Optional[np.ndarray] = None      # (G, N)
    kern: str = "SANM"
    mode: str = "FORWARD"
    nout: int = 500
    nin: int = 2
    serc: float = 1e-5
    ferc: float = 1e-5
    nac: int = 5
    nupd: Optional[int] = None
    name: str = "problem"
    D: Optional[np.ndarray] = None          # derived (G, N)
    sigr: Optional[np.ndarray] = None       # derived (G, N)

    def __post_init__(self):
        if self.exsrc is None:
            self.exsrc = np.zeros((self.ng, self.nnod))
        if self.nupd is None:
            self.nupd = default_nupd(self.nxx, self.nyy, self.nzz)
        if self.D is None or self.sigr is None:
            self.D, self.sigr = d_sigr(self.sigtr, self.siga, self.sigs)

    @property
    def nnod(self) -> int:
        return int(self.ix.shape[0])

    @property
    def vdel(self) -> np.ndarray:
        """Node volumes, ``misc`` (src/mod_io.f90:1336-1340)."""
        return (self.xdel[self.ix - 1] * self.ydel[self.iy - 1]) * self.zdel[self.iz - 1]

    @property
    def kern_id(self) -> int:
        return KERN_IDS[self.kern]

    @property
    def unknowns(self) -> int:
        return self.nnod * self.ng

    def with_iter(self, **kw) -> "Problem":
        return dataclasses.replace(self, **kw)

    def xyz(self) -> np.ndarray:
        """Dense (nzz, nyy, nxx) map to 1-based node number, 0 outside."""
        m = np.zeros((self.nzz, self.nyy, self.nxx), dtype=np.int32)
        m[self.iz - 1, self.iy - 1, self.ix - 1] = np.arange(1, self.nnod + 1, dtype=np.int32)
        return m


def default_nupd(nxx: int, nyy: int, nzz: int) -> int:
    """Default nodal update interval (src/mod_io.f90:1363)."""
    return int(math.ceil((nxx + nyy + nzz) / 2.5))


def d_sigr(sigtr: np.ndarray, siga: np.ndarray, sigs: np.ndarray):
    """D = 1/(3 sigtr), sigr = siga + sum_{h!=g} sigs(g->h)  (src/mod_xsec.f90:199-226).

    The sum runs in the reference's order (h ascending, starting from 0) so
    that the result is bit-identical.
    """
    if np.any(sigtr < 1e-5):
        raise ValueError("Negative diffusion coefficient encountered")
    ng = sigtr.shape[0]
    D = 1.0 / (3.0 * sigtr)
    sigr = np.empty_like(siga)
    for g in range(ng):
        dum = np.zeros(siga.shape[1])
        for h in range(ng):
            if h != g:
                dum = dum + sigs[h, g]       # sigs(n, from=g, to=h)
        sigr[g] = siga[g] + dum
    return D, sigr


# --------------------------------------------------------------------------
# Geometry: assemblies -> nodes, staggered footprint, node numbering.
# --------------------------------------------------------------------------

def build_geometry(xsize, xdiv, ysize, ydiv, zsize, zdiv, planars, zpln):
    """Node sizes, footprint and numbering from assembly data.

    Mirrors ``inp_geom1`` (src/mod_io.f90:900-950: node sizes), ``inp_geom2``
    (src/mod_io.f90:1010-1110: material per node and stagger bounds taken
    from the first plane) and ``misc`` (src/mod_io.f90:1310-1323: nodes
    numbered with k slowest, then j, then i from smin to smax).

    ``planars[p]`` is an (ny, nx) integer array indexed [j-1, i-1] (j=1 is
    the south row); ``zpln[k]`` is the 1-based planar type of axial layer k.
    Returns a dict of geometry arrays.
    """
    xsize = np.asarray(xsize, float); ysize = np.asarray(ysize, float); zsize = np.asarray(zsize, float)
    xdiv = np.asarray(xdiv, int); ydiv = np.asarray(ydiv, int); zdiv = np.asarray(zdiv, int)
    xdel = np.repeat(xsize / xdiv.astype(np.float64), xdiv)
    ydel = np.repeat(ysize / ydiv.astype(np.float64), ydiv)
    zdel = np.repeat(zsize / zdiv.astype(np.float64), zdiv)
    nxx, nyy, nzz = len(xdel), len(ydel), len(zdel)
    # node material map mnum[k, j, i]
    xa = np.repeat(np.arange(len(xdiv)), xdiv)
    ya = np.repeat(np.arange(len(ydiv)), ydiv)
    za = np.repeat(np.arange(len(zdiv)), zdiv)
    mnum = np.zeros((nzz, nyy, nxx), dtype=np.int32)
    for k in range(nzz):
        pl = np.asarray(planars[zpln[za[k]] - 1], dtype=np.int32)
        mnum[k] = pl[np.ix_(ya, xa)]
    plane1 = mnum[0]
    ystag_min = np.full(nyy, nxx, dtype=np.int32)
    ystag_max = np.zeros(nyy, dtype=np.int32)
    for j in range(nyy):
        nz_i = np.nonzero(plane1[j] != 0)[0]
        if len(nz_i):
            ystag_min[j] = nz_i[0] + 1
            ystag_max[j] = nz_i[-1] + 1
    xstag_min = np.full(nxx, nyy, dtype=np.int32)
    xstag_max = np.zeros(nxx, dtype=np.int32)
    for i in range(nxx):
        nz_j = np.nonzero(plane1[:, i] != 0)[0]
        if len(nz_j):
            xstag_min[i] = nz_j[0] + 1
            xstag_max[i] = nz_j[-1] + 1
    ix, iy, iz, mat = [], [], [], []
    for k in range(nzz):
        for j in range(nyy):
            for i in range(ystag_min[j], ystag_max[j] + 1):
                ix.append(i); iy.append(j + 1); iz.append(k + 1)
                mat.append(mnum[k, j, i - 1])
    mat = np.asarray(mat, dtype=np.int32)
    if np.any(mat == 0):
        raise ValueError("Zero material found inside core. Check material assignment")
    return dict(nxx=nxx, nyy=nyy, nzz=nzz, xdel=xdel, ydel=ydel, zdel=zdel,
                ystag_min=ystag_min, ystag_max=ystag_max,
                xstag_min=xstag_min, xstag_max=xstag_max,
                ix=np.asarray(ix, np.int32), iy=np.asarray(iy, np.int32),
                iz=np.asarray(iz, np.int32), mat=mat)


def node_xs_from_materials(mat, xsigtr, xsiga, xnuf, xsigf, xsigs):
    """Per-node cross sections from material tables, ``base_updt``
    (src/mod_xsec.f90:172-195).  Material tables are (nmat, G) and
    (nmat, G_from, G_to) in natural numpy indexing.
    """
    m = np.asarray(mat) - 1
    sigtr = np.ascontiguousarray(np.asarray(xsigtr)[m].T)
    siga = np.ascontiguousarray(np.asarray(xsiga)[m].T)
    nuf = np.ascontiguousarray(np.asarray(xnuf)[m].T)
    sigf = np.ascontiguousarray(np.asarray(xsigf)[m].T)
    # xsigs[mat, from, to] -> sigs[to, from, n]
    sigs = np.ascontiguousarray(np.transpose(np.asarray(xsigs)[m], (2, 1, 0)))
    return sigtr, siga, nuf, sigf, sigs


# --------------------------------------------------------------------------
# Deck writer (for the reference build) and reader of the reference dump.
# --------------------------------------------------------------------------

def round_deck(v):
    """Round to the 10 significant digits a deck line can hold.

    The reference reads each card line into a 200-character buffer
    (src/mod_io.f90:470,1012), so an 8-group %XSEC row (13 values) only fits
    with ~10 digits per value.  Material data are rounded at generation time
    so that the deck and the in-memory problem hold the same doubles.
    """
    return np.vectorize(lambda x: float("%.10g" % x))(np.asarray(v, dtype=np.float64))


def _fmt(v: float) -> str:
    return "%.10g" % float(v)


def write_deck(path: str, *, ng, xsigtr, xsiga, xnuf, xsigf, xchi, xsigs,
               xsize, xdiv, ysize, ydiv, zsize, zdiv, planars, zpln, bc,
               kern="SANM", mode="FORWARD", iter_card=None, case="GRAFT"):
    """Write an ADPRES deck with %MODE %CASE %XSEC %GEOM [%KERN] [%ITER].

    Card syntax follows docs/card-desc.md and the readers ``inp_xsec``
    (src/mod_io.f90:683), ``inp_geom1/2`` (:809, :958), ``inp_iter``
    (:1522) and ``inp_kern`` (:1593).  Values are written with full
    round-trip precision.
    """
    nmat = len(xsigtr)
    L: List[str] = []
    L += ["%MODE", mode, "", "%CASE", case, "synthetic core", ""]
    L += ["%XSEC", f"{ng} {nmat}"]
    for m in range(nmat):
        for g in range(ng):
            row = [xsigtr[m][g], xsiga[m][g], xnuf[m][g], xsigf[m][g], xchi[m][g]]
            row += [xsigs[m][g][h] for h in range(ng)]
            L.append(" ".join(_fmt(v) for v in row))
    L += ["", "%GEOM", f"{len(xsize)} {len(ysize)} {len(zsize)}"]
    L.append(" ".join(_fmt(v) for v in xsize)); L.append(" ".join(str(int(v)) for v in xdiv))
    L.append(" ".join(_fmt(v) for v in ysize)); L.append(" ".join(str(int(v)) for v in ydiv))
    L.append(" ".join(_fmt(v) for v in zsize)); L.append(" ".join(str(int(v)) for v in zdiv))
    L.append(str(len(planars)))
    L.append(" ".join(str(int(v)) for v in zpln))
    for pl in planars:
        pl = np.asarray(pl)
        for j in range(pl.shape[0] - 1, -1, -1):      # rows written north to south
            L.append(" ".join(str(int(v)) for v in pl[j]))
    L.append(" ".join(str(int(b)) for b in bc))
    if kern != "SANM":
        L += ["", "%KERN", kern]
    if iter_card is not None:
        nout, nin, serc, ferc, nac, nupd = iter_card
        L += ["", "%ITER", f"{nout} {nin} {serc!r} {ferc!r} {nac} {nupd} 30 20"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_xs_override(path: str, p: Problem) -> None:
    """Per-node XS file read by oracle/ref_dump.f90 (Fortran order float64)."""
    with open(path, "wb") as f:
        for a in (p.sigtr, p.siga, p.nuf, p.sigf, p.sigs):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def write_state_file(path: str, nin: int, Ke: float, f0: np.ndarray, dn: np.ndarray,
                     b: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(np.int32(nin).tobytes())
        f.write(np.float64(Ke).tobytes())
        for a in (f0, dn, b):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


class _Reader:
    def __init__(self, buf: bytes):
        self.buf = buf
        self.off = 0

    def take(self, dtype, count):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.buf, dtype=dt, count=count, offset=self.off)
        self.off += dt.itemsize * count
        return a.copy()


def read_ref_dump(path: str, action: str = "solve"):
    """Read the raw stream written by oracle/ref_dump.f90.

    Returns ``(problem, results)``.  The problem is exactly what the
    reference held in its global state when the hot path started.
    """
    with open(path, "rb") as f:
        r = _Reader(f.read())
    magic, N, G, nxx, nyy, nzz, nmat, kid = r.take("<i4", 8)
    if magic != 20240601:
        raise ValueError("not a ref_dump file")
    bc = tuple(int(v) for v in r.take("<i4", 6))
    nout, nin, nac, nupd = (int(v) for v in r.take("<i4", 4))
    serc, ferc = (float(v) for v in r.take("<f8", 2))
    xdel = r.take("<f8", nxx); ydel = r.take("<f8", nyy); zdel = r.take("<f8", nzz)
    ys = r.take("<i4", 2 * nyy); xs = r.take("<i4", 2 * nxx)
    idx = r.take("<i4", 4 * N).reshape(4, N)
    sh2 = (G, N)
    sigtr = r.take("<f8", G * N).reshape(sh2)
    siga = r.take("<f8", G * N).reshape(sh2)
    nuf = r.take("<f8", G * N).reshape(sh2)
    sigf = r.take("<f8", G * N).reshape(sh2)
    sigs = r.take("<f8", G * G * N).reshape(G, G, N)
    chi = r.take("<f8", nmat * G).reshape(G, nmat)
    dc = r.take("<f8", 6 * G * N).reshape(6, G, N)
    D = r.take("<f8", G * N).reshape(sh2)
    sigr = r.take("<f8", G * N).reshape(sh2)
    vdel = r.take("<f8", N)
    p = Problem(ng=int(G), nmat=int(nmat), nxx=int(nxx), nyy=int(nyy), nzz=int(nzz),
                xdel=xdel, ydel=ydel, zdel=zdel,
                ystag_min=ys[:nyy], ystag_max=ys[nyy:], xstag_min=xs[:nxx], xstag_max=xs[nxx:],
                bc=bc, ix=idx[0], iy=idx[1], iz=idx[2], mat=idx[3],
                sigtr=sigtr, siga=siga, nuf=nuf, sigf=sigf, sigs=sigs, chi=chi, dc=dc,
                kern=KERN_NAMES[int(kid)], nout=nout, nin=nin, serc=serc, ferc=ferc,
                nac=nac, nupd=nupd, D=D, sigr=sigr)
    res: Dict[str, object] = {"vdel": vdel}
    if action == "solve":
        res["Ke"] = float(r.take("<f8", 1)[0])
        res["f0"] = r.take("<f8", G * N).reshape(G, N)
        res["fs0"] = r.take("<f8", N)
        res["pow"] = r.take("<f8", N)
        res["df"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["dn"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        if r.off < len(r.buf):      # a fixed-source deck: its %ESRC source follows (outer_fs)
            p.exsrc = r.take("<f8", G * N).reshape(sh2)
            p.mode = "FIXEDSRC"
    elif action == "nodal":
        res["df"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["dn"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["ndmax"] = float(r.take("<f8", 1)[0])
        res["loc"] = tuple(int(v) for v in r.take("<i4", 3))
    elif action == "bicg":
        res["x"] = r.take("<f8", G * N).reshape(G, N)
    elif action == "fwdadj":
        res["Ke_s"] = float(r.take("<f8", 1)[0])
        res["f0_s"] = r.take("<f8", G * N).reshape(G, N)
        res["fs0_s"] = r.take("<f8", N)
        res["Ke"] = float(r.take("<f8", 1)[0])
        res["f0"] = r.take("<f8", G * N).reshape(G, N)
        res["fs0"] = r.take("<f8", N)
        res["dn"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["ndmax"] = float(r.take("<f8", 1)[0])
    elif action == "xsupd":
        flags = [int(v) for v in r.take("<i4", 6)]
        # the feedback tables in the form oracle.xs_update and
        # Engine.set_xs_tables take
        fl = dict(zip(("boron", "ftem", "mtem", "cden", "rod"), flags[:5]))
        nb = flags[5]
        M = int(nmat)
        tab = {"nmat": M, "nb": nb}
        tab.update(zip(("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize"),
                       (float(v) for v in r.take("<f8", 7))))

        def tables():
            t = [r.take("<f8", M * G).reshape(G, M) for _ in range(4)]
            t.append(r.take("<f8", M * G * G).reshape(G, G, M))
            return t
        tab["base"] = tables()
        for key in ("boron", "ftem", "mtem", "cden", "rod"):
            tab[key] = tables() if fl[key] else None
        tab["fbmap"] = r.take("<i4", int(nxx) * int(nyy)).reshape(int(nyy), int(nxx)) if fl["rod"] else None
        res["tables"] = tab
        states = []
        for _ in range(2):
            st = {"bcon": float(r.take("<f8", 1)[0])}
            for key in ("ftem", "mtem", "cden"):
                if fl[key]:
                    st[key] = r.take("<f8", N)
            if fl["rod"]:
                st["bpos"] = r.take("<f8", nb)
            for key in ("sigtr", "siga", "nuf", "sigf"):
                st[key] = r.take("<f8", G * N).reshape(G, N)
            st["sigs"] = r.take("<f8", G * G * N).reshape(G, G, N)
            st["D"] = r.take("<f8", G * N).reshape(G, N)
            st["sigr"] = r.take("<f8", G * N).reshape(G, N)
            states.append(st)
        res["states"] = states
    elif action == "xtab":
        # branch tables in the flat form of branch_tables_from_dict
        bcrod, nb = (int(v) for v in r.take("<i4", 2))
        M = int(nmat)
        dims = r.take("<i4", 5 * M).reshape(M, 5)
        E = 4 * G + G * G + 6 * G
        pts, xs, rxs = [], [], []
        for mm in range(M):
            nd, nbr, nf, nm, trod = (int(v) for v in dims[mm])
            pts.append(r.take("<f8", nd + nbr + nf + nm))
            xs.append(r.take("<f8", nd * nbr * nf * nm * E))
            if trod == 1:
                rxs.append(r.take("<f8", nd * nbr * nf * nm * E))
        coreh, pos0, ssize = (float(v) for v in r.take("<f8", 3))
        fb = r.take("<i4", int(nxx) * int(nyy)).reshape(int(nyy), int(nxx)) if bcrod else None
        res["branch"] = {"nmat": M, "nb": nb if bcrod else 0, "dims": dims.astype(np.int32),
                         "points": np.concatenate(pts), "xsec": np.concatenate(xs),
                         "rxsec": np.concatenate(rxs) if rxs else None,
                         "coreh": coreh, "pos0": pos0, "ssize": ssize, "fbmap": fb}
        states = []
        for _ in range(2):
            st = {"bcon": float(r.take("<f8", 1)[0])}
            for key in ("ftem", "mtem", "cden"):
                st[key] = r.take("<f8", N)
            if bcrod:
                st["bpos"] = r.take("<f8", nb)
            for key in ("sigtr", "siga", "nuf", "sigf"):
                st[key] = r.take("<f8", G * N).reshape(G, N)
            st["sigs"] = r.take("<f8", G * G * N).reshape(G, G, N)
            st["D"] = r.take("<f8", G * N).reshape(G, N)
            st["sigr"] = r.take("<f8", G * N).reshape(G, N)
            st["dc"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
            states.append(st)
        res["states"] = states
    elif action == "trstep":
        res["Ke_s"] = float(r.take("<f8", 1)[0])
        res["f0_s"] = r.take("<f8", G * N).reshape(G, N)
        res["fs0_s"] = r.take("<f8", N)
        res["dn_s"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["ht"] = float(r.take("<f8", 1)[0])
        res["tbeta"] = r.take("<f8", int(nmat))
        res["sigr_tr"] = r.take("<f8", G * N).reshape(G, N)
        res["dfis"] = r.take("<f8", N)
        res["exsrc"] = r.take("<f8", G * N).reshape(G, N)
        res["Ke"] = float(r.take("<f8", 1)[0])
        res["f0"] = r.take("<f8", G * N).reshape(G, N)
        res["fs0"] = r.take("<f8", N)
        res["dn"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["df"] = r.take("<f8", 6 * G * N).reshape(6, G, N)
        res["maxi"] = int(r.take("<i4", 1)[0])
    return p, res


def parse_trace(text: str, fixed_source: bool = False):
    """Per-outer trace lines 'p k ser fer' printed by outer (src/mod_cmfd.f90:491-494),
    or 'p ser fer' printed by outer_fs (:583-584), returned as (p, k, ser, fer)
    with k = NaN for the latter."""
    rows = []
    width = 3 if fixed_source else 4
    for line in text.splitlines():
        t = line.split()
        if len(t) == width and t[0].isdigit():
            try:
                v = [float(x) for x in t[1:]]
            except ValueError:
                continue
            rows.append((int(t[0]), *(([float("nan")] + v) if fixed_source else v)))
    return rows


# --------------------------------------------------------------------------
# Problem <-> compact npz (fixtures)
# --------------------------------------------------------------------------

_ARR_FIELDS = ("xdel", "ydel", "zdel", "ystag_min", "ystag_max", "xstag_min", "xstag_max",
               "ix", "iy", "iz", "mat", "sigtr", "siga", "nuf", "sigf", "sigs", "chi", "dc",
               "exsrc", "D", "sigr")
_SCALAR_FIELDS = ("ng", "nmat", "nxx", "nyy", "nzz", "kern", "mode", "nout", "nin", "serc",
                  "ferc", "nac", "nupd", "name")


def problem_to_dict(p: Problem, prefix: str = "p_") -> Dict[str, np.ndarray]:
    d = {prefix + k: np.asarray(getattr(p, k)) for k in _ARR_FIELDS}
    for k in _SCALAR_FIELDS:
        d[prefix + k] = np.asarray(getattr(p, k))
    d[prefix + "bc"] = np.asarray(p.bc, dtype=np.int32)
    return d


def problem_from_dict(d, prefix: str = "p_") -> Problem:
    kw = {k: np.asarray(d[prefix + k]) for k in _ARR_FIELDS}
    for k in _SCALAR_FIELDS:
        v = np.asarray(d[prefix + k]).item()
        kw[k] = v
    kw["bc"] = tuple(int(v) for v in np.asarray(d[prefix + "bc"]))
    for k in ("ng", "nmat", "nxx", "nyy", "nzz", "nout", "nin", "nac", "nupd"):
        kw[k] = int(kw[k])
    return Problem(**kw)


XS_KINDS = ("base", "boron", "ftem", "mtem", "cden", "rod")
XS_TABLE_NAMES = ("sigtr", "siga", "nuf", "sigf", "sigs")
XS_REFS = ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize")


def xs_tables_from_dict(d, prefix: str = "t_"):
    """The feedback tables of xsec::XS_updt (src/mod_xsec.f90:11-46) from a
    fixture's arrays: {"nmat", "nb", the reference values of XS_REFS, per
    kind of XS_KINDS the five tables of XS_TABLE_NAMES (sigtr..sigf as
    (G, nmat), sigs as (G_to, G_from, nmat); None where the feedback is
    off) and "fbmap" (nyy, nxx) or None}."""
    files = set(getattr(d, "files", d.keys()))
    t = {"nmat": int(np.asarray(d[prefix + "nmat"])), "nb": int(np.asarray(d[prefix + "nb"]))}
    for k in XS_REFS:
        t[k] = float(np.asarray(d[prefix + k]))
    for kind in XS_KINDS:
        key = f"{prefix}{kind}_sigtr"
        t[kind] = [np.asarray(d[f"{prefix}{kind}_{x}"]) for x in XS_TABLE_NAMES] \
            if key in files and np.asarray(d[key]).ndim == 2 else None
    fb = np.asarray(d[prefix + "fbmap"]) if prefix + "fbmap" in files else None
    t["fbmap"] = fb if fb is not None and fb.ndim == 2 else None
    return t


def branch_tables_from_dict(d, prefix: str = "b_"):
    """Flat branch tables of xsec::XStab_updt (src/mod_xsec.f90:48-84) from
    a fixture's arrays: {"nmat", "nb", "dims" (nmat, 5: nd, nb, nf, nm,
    trod), "points", "xsec", "rxsec" or None, "coreh", "pos0", "ssize",
    "fbmap" (nyy, nxx) or None}; the layouts are those of adg_xtab_tables."""
    t = {"nmat": int(np.asarray(d[prefix + "nmat"])), "nb": int(np.asarray(d[prefix + "nb"]))}
    for k in ("dims", "points", "xsec"):
        t[k] = np.asarray(d[prefix + k])
    for k in ("coreh", "pos0", "ssize"):
        t[k] = float(np.asarray(d[prefix + k]))
    for k in ("rxsec", "fbmap"):
        a = np.asarray(d[prefix + k])
        t[k] = None if a.dtype == object or a.ndim == 0 else a
    return t


def deck_layout(path: str):
    """The assembly layout of a deck's %GEOM card (src/mod_io.f90's
    inp_geom1): nx, ny, nz and the node divisions xdiv (nx), ydiv (ny),
    zdiv (nz), reading "k*v" repeats and dropping "!" comments."""
    toks, on = [], False
    for line in open(path):
        body = line.split("!")[0].strip()
        if body.startswith("%"):
            if on:
                break
            on = body.upper().startswith("%GEOM")
            continue
        if on and body:
            for t in body.split():
                if "*" in t:
                    k, v = t.split("*")
                    toks += [v] * int(k)
                else:
                    toks.append(t)
    nx, ny, nz = (int(t) for t in toks[:3])
    pos = 3

    def take(n):
        nonlocal pos
        out = toks[pos:pos + n]
        pos += n
        return out
    take(nx)
    xdiv = [int(float(t)) for t in take(nx)]
    take(ny)
    ydiv = [int(float(t)) for t in take(ny)]
    take(nz)
    zdiv = [int(float(t)) for t in take(nz)]
    return nx, ny, nz, np.array(xdiv, np.int32), np.array(ydiv, np.int32), np.array(zdiv, np.int32)
