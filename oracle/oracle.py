"""ctypes binding of the FP64 CPU restatement in liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry and by the
benchmark's CPU-baseline leg, always as the checker.  The product path never
imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_OK, ORC_E_MAX_OUTER, ORC_E_NODAL_UNSTABLE, ORC_E_LU_PIVOT, ORC_E_POWER_NONPOS, ORC_E_ARG = range(6)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [("nnod", C.c_int), ("ng", C.c_int), ("nxx", C.c_int), ("nyy", C.c_int),
                ("nzz", C.c_int), ("nmat", C.c_int),
                ("ix", _ip), ("iy", _ip), ("iz", _ip), ("mat", _ip),
                ("ystag_min", _ip), ("ystag_max", _ip), ("xstag_min", _ip), ("xstag_max", _ip),
                ("xdel", _dp), ("ydel", _dp), ("zdel", _dp), ("bc", C.c_int * 6),
                ("D", _dp), ("sigr", _dp), ("nuf", _dp), ("sigf", _dp), ("sigs", _dp),
                ("chi", _dp), ("dc", _dp), ("exsrc", _dp), ("wtr", _dp)]


class OrcXsTables(C.Structure):
    _fields_ = [("nmat", C.c_int), ("nb", C.c_int)] + \
        [(k, C.c_double) for k in ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize")] + \
        [("tab", (_dp * 5) * 6), ("fbmap", _ip)]


class OrcXtabTables(C.Structure):
    _fields_ = [("nmat", C.c_int), ("nb", C.c_int), ("dims", _ip), ("points", _dp), ("xsec", _dp),
                ("rxsec", _dp), ("coreh", C.c_double), ("pos0", C.c_double), ("ssize", C.c_double),
                ("fbmap", _ip)]


class OrcIter(C.Structure):
    _fields_ = [("nout", C.c_int), ("nin", C.c_int), ("nac", C.c_int), ("nupd", C.c_int),
                ("kern", C.c_int), ("serc", C.c_double), ("ferc", C.c_double)]


class OrcState(C.Structure):
    _fields_ = [("Ke", C.c_double), ("f0", _dp), ("fs0", _dp), ("dn", _dp), ("df", _dp),
                ("ndmax", C.c_double), ("im", C.c_int), ("jm", C.c_int), ("km", C.c_int),
                ("first", C.c_int)]


class OrcTrace(C.Structure):
    _fields_ = [("cap", C.c_int), ("n", C.c_int), ("k", _dp), ("ser", _dp), ("fer", _dp),
                ("ndmax", _dp), ("flags", _ip), ("loc", _ip)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_outer.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState), C.POINTER(OrcIter),
                                C.c_int, C.POINTER(OrcTrace)]
        L.orc_nodal_update.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState), C.c_int, C.c_int]
        L.orc_bicg.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState), C.c_int, C.c_int, _dp, _dp]
        L.orc_leakage.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState), _dp]
        L.orc_matvec.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState), C.c_int, _dp, _dp]
        L.orc_powdis.argtypes = [C.POINTER(OrcProblem), _dp, _dp]
        L.orc_coup_coef.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcState)]
        L.orc_sanm_coef.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_sanm_coef.restype = None
        L.orc_flux_map.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_int, _ip, _ip, _dp, C.c_double, _dp, _ip]
        L.orc_power_maps.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp,
                                     _dp, _ip, _dp, _ip]
        L.orc_xtab_update.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcXtabTables), C.c_double,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_xs_update.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcXsTables), C.c_double,
                                    _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def sanm_coef(sigr, D, h):
    """(A, B, E, F, G, H) of get_ABEFGH for one node and group."""
    out = np.zeros(6)
    lib().orc_sanm_coef(float(sigr), float(D), float(h), _d(out))
    return out


class Oracle:
    """Holds contiguous copies of a Problem's arrays and the solver state."""

    def __init__(self, p):
        self.p = p
        self._keep = {}
        ints = {k: np.ascontiguousarray(getattr(p, k), dtype=np.int32)
                for k in ("ix", "iy", "iz", "mat", "ystag_min", "ystag_max", "xstag_min", "xstag_max")}
        dbls = {k: np.ascontiguousarray(getattr(p, k), dtype=np.float64)
                for k in ("xdel", "ydel", "zdel", "D", "sigr", "nuf", "sigf", "sigs", "chi", "dc", "exsrc")}
        self._keep.update(ints); self._keep.update(dbls)
        pb = OrcProblem()
        pb.nnod, pb.ng, pb.nxx, pb.nyy, pb.nzz, pb.nmat = p.nnod, p.ng, p.nxx, p.nyy, p.nzz, p.nmat
        for k, v in ints.items():
            setattr(pb, k, _i(v))
        for k, v in dbls.items():
            setattr(pb, k, _d(v))
        pb.bc = (C.c_int * 6)(*p.bc)
        self.set_transient(getattr(p, "wtr", None), pb)
        self.pb = pb
        N, G = p.nnod, p.ng
        self.f0 = np.ones((G, N)); self.fs0 = np.zeros(N)
        self.dn = np.zeros((6, G, N)); self.df = np.zeros((6, G, N))
        st = OrcState()
        st.Ke = 1.0; st.f0 = _d(self.f0); st.fs0 = _d(self.fs0); st.dn = _d(self.dn)
        st.df = _d(self.df); st.ndmax = 0.0; st.first = 1
        self.st = st

    @property
    def Ke(self):
        return self.st.Ke

    def set_transient(self, wtr, pb=None):
        """Per-node weight 1 - tbeta + dfis of the transient source and
        nodal update (None: no transient form)."""
        pb = pb if pb is not None else self.pb
        if wtr is None:
            pb.wtr = None
            return
        w = np.ascontiguousarray(wtr, dtype=np.float64)
        self._keep["wtr"] = w
        pb.wtr = _d(w)

    def outer(self, adjoint=None, nout=None, nin=None, serc=None, ferc=None, nac=None, nupd=None,
              kern=None, trace_cap=None, fixed_source=None, transient=False, nodal=True):
        p = self.p
        it = OrcIter(nout if nout is not None else p.nout, nin if nin is not None else p.nin,
                     nac if nac is not None else p.nac, nupd if nupd is not None else p.nupd,
                     {"FDM": 0, "PNM": 1, "SANM": 2}[kern or p.kern],
                     serc if serc is not None else p.serc, ferc if ferc is not None else p.ferc)
        if adjoint is None:
            adjoint = p.mode == "ADJOINT"
        if fixed_source is None:
            fixed_source = p.mode == "FIXEDSRC"
        mode = 3 if transient else (2 if fixed_source else (int(bool(adjoint)) if nodal or not adjoint else 4))
        cap = trace_cap or it.nout
        self.trace = {k: np.zeros(cap) for k in ("k", "ser", "fer", "ndmax")}
        self.trace["flags"] = np.zeros(cap, np.int32)
        self.trace["loc"] = np.zeros(3 * cap, np.int32)
        tr = OrcTrace(cap, 0, _d(self.trace["k"]), _d(self.trace["ser"]), _d(self.trace["fer"]),
                      _d(self.trace["ndmax"]), _i(self.trace["flags"]), _i(self.trace["loc"]))
        rc = lib().orc_outer(C.byref(self.pb), C.byref(self.st), C.byref(it), mode, C.byref(tr))
        self.niter = tr.n
        for k in ("k", "ser", "fer", "ndmax", "flags"):
            self.trace[k] = self.trace[k][:tr.n]
        self.trace["loc"] = self.trace["loc"][:3 * tr.n].reshape(-1, 3)
        return rc

    def leakage(self):
        """Lxyz of every node and group of the current state: (3, G, N)."""
        out = np.zeros((3, self.p.ng, self.p.nnod))
        rc = lib().orc_leakage(C.byref(self.pb), C.byref(self.st), _d(out))
        if rc != 0:
            raise RuntimeError(f"orc_leakage: {rc}")
        return out

    def nodal_update(self, Ke, f0, dn, kern=None, cmode=1):
        self.st.Ke = Ke
        self.f0[...] = f0
        self.dn[...] = dn
        rc = lib().orc_nodal_update(C.byref(self.pb), C.byref(self.st),
                                    {"FDM": 0, "PNM": 1, "SANM": 2}[kern or self.p.kern], cmode)
        return rc, self.dn.copy(), self.st.ndmax, (self.st.im, self.st.jm, self.st.km)

    def bicg(self, g, nin, b, x, dn=None):
        if dn is not None:
            self.dn[...] = dn
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        rc = lib().orc_bicg(C.byref(self.pb), C.byref(self.st), g, nin, _d(b), _d(x))
        return rc, x

    def matvec(self, g, x, dn=None):
        if dn is not None:
            self.dn[...] = dn
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        rc = lib().orc_matvec(C.byref(self.pb), C.byref(self.st), g, _d(x), _d(y))
        return rc, y

    def powdis(self, f0=None):
        f0 = np.ascontiguousarray(self.f0 if f0 is None else f0, dtype=np.float64)
        out = np.zeros(self.p.nnod)
        rc = lib().orc_powdis(C.byref(self.pb), _d(f0), _d(out))
        return rc, out


XS_KINDS = ("base", "boron", "ftem", "mtem", "cden", "rod")


def xs_update(p, tables, bcon, ftem=None, mtem=None, cden=None, bpos=None):
    """xsec::XS_updt (src/mod_xsec.f90:11-46) of Problem p's nodes with the
    feedback tables (a dict: "nmat", "nb", the reference values "rbcon",
    "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize", per present kind of
    XS_KINDS a list of the five tables sigtr, siga, nuf, sigf as (G, nmat)
    and sigs as (G_to, G_from, nmat), and "fbmap" (nyy, nxx) with the rod).
    Returns (status, dict of sigtr, siga, nuf, sigf, D, sigr as (G, N) and
    sigs as (G_to, G_from, N))."""
    N, G = p.nnod, p.ng
    keep = []

    def dp(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _d(a)
    t = OrcXsTables()
    t.nmat, t.nb = int(tables["nmat"]), int(tables.get("nb", 0))
    for k in ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize"):
        setattr(t, k, float(tables.get(k, 0.0)))
    for i, kind in enumerate(XS_KINDS):
        if tables.get(kind) is not None:
            for j in range(5):
                t.tab[i][j] = dp(tables[kind][j])
    if tables.get("fbmap") is not None:
        fb = np.ascontiguousarray(tables["fbmap"], dtype=np.int32)
        keep.append(fb)
        t.fbmap = _i(fb)
    pb = Oracle(p).pb
    out = {k: np.zeros((G, N)) for k in ("sigtr", "siga", "nuf", "sigf", "D", "sigr")}
    out["sigs"] = np.zeros((G, G, N))
    rc = lib().orc_xs_update(C.byref(pb), C.byref(t), float(bcon), dp(ftem), dp(mtem), dp(cden), dp(bpos),
                             *[_d(out[k]) for k in ("sigtr", "siga", "nuf", "sigf", "sigs", "D", "sigr")])
    return rc, out


def xtab_update(p, tables, bcon, ftem, mtem, cden, bpos=None):
    """xsec::XStab_updt (src/mod_xsec.f90:48-84) of Problem p's nodes with
    branch tables in flat form (a dict: "nmat", "nb", "dims" (nmat, 5),
    "points", "xsec", "rxsec" or None, "coreh", "pos0", "ssize", "fbmap"
    (nyy, nxx) or None).  Returns (status, dict of sigtr, siga, nuf, sigf,
    D, sigr as (G, N), sigs as (G_to, G_from, N) and dc as (6, G, N))."""
    N, G = p.nnod, p.ng
    keep = []

    def dp(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _d(a)

    def ip(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return _i(a)
    t = OrcXtabTables(int(tables["nmat"]), int(tables.get("nb", 0)), ip(tables["dims"]), dp(tables["points"]),
                      dp(tables["xsec"]), dp(tables.get("rxsec")), float(tables.get("coreh", 0.0)),
                      float(tables.get("pos0", 0.0)), float(tables.get("ssize", 0.0)), ip(tables.get("fbmap")))
    pb = Oracle(p).pb
    out = {k: np.zeros((G, N)) for k in ("sigtr", "siga", "nuf", "sigf", "D", "sigr")}
    out["sigs"] = np.zeros((G, G, N))
    out["dc"] = np.zeros((6, G, N))
    rc = lib().orc_xtab_update(C.byref(pb), C.byref(t), float(bcon), dp(ftem), dp(mtem), dp(cden), dp(bpos),
                               *[_d(out[k]) for k in ("sigtr", "siga", "nuf", "sigf", "sigs", "D", "sigr", "dc")])
    return rc, out


def power_maps(p, nx, ny, nz, xdiv, ydiv, zdiv, pw):
    """io::AsmPow / AxiPow numbers (src/mod_io.f90:3267-3498) of node powers
    pw: (status, fasm (ny, nx), (i, j) of its maximum, faxi (nz,), plane of
    its maximum), positions 1-based."""
    xd, yd, zd = (np.ascontiguousarray(a, dtype=np.int32) for a in (xdiv, ydiv, zdiv))
    pw = np.ascontiguousarray(pw, dtype=np.float64)
    fasm = np.zeros((ny, nx)); faxi = np.zeros(nz)
    pm = np.zeros(2, np.int32); am = np.zeros(1, np.int32)
    o = Oracle(p)
    rc = lib().orc_power_maps(C.byref(o.pb), nx, ny, nz, _i(xd), _i(yd), _i(zd), _d(pw), _d(fasm), _i(pm),
                              _d(faxi), _i(am))
    return rc, fasm, (int(pm[0]), int(pm[1])), faxi, int(am[0])


def flux_map(p, nx, ny, xdiv, ydiv, f0, norm=1.0):
    """io::AsmFlux numbers (src/mod_io.f90:3502-3644) of the flux f0 (G, N):
    (status, fasm (G, ny, nx), negative-flux flag)."""
    xd, yd = (np.ascontiguousarray(a, dtype=np.int32) for a in (xdiv, ydiv))
    f0 = np.ascontiguousarray(f0, dtype=np.float64)
    fasm = np.zeros((p.ng, ny, nx))
    negf = np.zeros(1, np.int32)
    rc = lib().orc_flux_map(C.byref(Oracle(p).pb), nx, ny, _i(xd), _i(yd), _d(f0), float(norm), _d(fasm),
                            _i(negf))
    return rc, fasm, int(negf[0])
