"""Python host mirror of the reference's static hot path over the C-ABI.

:class:`Engine` plays the role of the reference's ``cmfd`` module for one
core: ``outer()`` is ``cmfd::outer`` / ``cmfd::outer_ad``
(src/mod_cmfd.f90:415,602), ``powdis()`` is ``cmfd::PowDis`` (:1290), and
``nodal_update``, ``bicg``, ``matvec`` are the kernel-level routines.  The
solver state (flux, fission source, k, D-hat) lives on the device between
calls, as the reference keeps it in SAVEd module variables, so repeated
``outer()`` calls warm-start exactly like the reference's BCSEARCH loop.

Errors are raised as :class:`ADGError` carrying the reference's message
(where the reference would print and STOP).  There is no CPU fallback: if
``libadpres_gpu.so`` is missing or no GPU is present, construction fails.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .problem import Problem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ADG_LIBRARY") or os.path.join(HERE, "libadpres_gpu.so")

ADG_OK, ADG_E_MAX_OUTER, ADG_E_NODAL_UNSTABLE, ADG_E_LU_PIVOT, ADG_E_POWER_NONPOS = 0, 1, 2, 3, 4
ADG_E_ARG, ADG_E_HIP, ADG_E_NOT_SET, ADG_E_XS, ADG_E_DEVICE_SYNC = 5, 6, 7, 8, 9
ADG_FORWARD, ADG_ADJOINT, ADG_FIXED_SOURCE, ADG_TRANSIENT, ADG_ADJOINT_FROZEN = 0, 1, 2, 3, 4   # enum adg_mode
FP64, FP32_REFINE = 0, 1
PRECISIONS = {"fp64": FP64, "fp32_refine": FP32_REFINE}
KERNS = {"FDM": 0, "PNM": 1, "SANM": 2}

EXPORTS = ("adg_create", "adg_destroy", "adg_set_xs", "adg_set_state", "adg_get_state", "adg_outer",
           "adg_powdis", "adg_nodal_update", "adg_bicg", "adg_matvec", "adg_last_timing",
           "adg_set_profiling", "adg_read_profile", "adg_autotune",
           "adg_create_batch", "adg_core_count", "adg_set_xs_core", "adg_set_state_core", "adg_get_state_core",
           "adg_outer_batch", "adg_trace_core", "adg_powdis_core",
           "adg_adjoint", "adg_get_adjoint", "adg_set_adjoint", "adg_backward", "adg_nodal_vjp", "adg_backward_nodal",
           "adg_set_transient", "adg_set_transient_core", "adg_leakage",
           "adg_set_xs_tables", "adg_set_xs_tables_core", "adg_update_xs", "adg_update_xs_core",
           "adg_get_xs", "adg_get_xs_core", "adg_set_xtab_tables", "adg_set_xtab_tables_core",
           "adg_power_maps", "adg_power_maps_core", "adg_flux_map", "adg_flux_map_core",
           "adg_power_maps_noalloc", "adg_solver_info", "adg_persistent_plan", "adg_status_string",
           "adg_last_error")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class AdgGeom(C.Structure):
    _fields_ = [("nnod", C.c_int), ("ng", C.c_int), ("nmat", C.c_int),
                ("nxx", C.c_int), ("nyy", C.c_int), ("nzz", C.c_int),
                ("ix", _ip), ("iy", _ip), ("iz", _ip),
                ("xdel", _dp), ("ydel", _dp), ("zdel", _dp),
                ("bc", C.c_int * 6), ("precision", C.c_int)]


class AdgXsTables(C.Structure):
    _fields_ = [("nmat", C.c_int), ("nb", C.c_int)] + \
        [(k, C.c_double) for k in ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize")] + \
        [("tab", (_dp * 5) * 6), ("fbmap", _ip)]


class AdgXtabTables(C.Structure):
    _fields_ = [("nmat", C.c_int), ("nb", C.c_int), ("dims", _ip), ("points", _dp), ("xsec", _dp),
                ("rxsec", _dp), ("coreh", C.c_double), ("pos0", C.c_double), ("ssize", C.c_double),
                ("fbmap", _ip)]


class AdgIter(C.Structure):
    _fields_ = [("nout", C.c_int), ("nin", C.c_int), ("nac", C.c_int), ("nupd", C.c_int),
                ("kern", C.c_int), ("serc", C.c_double), ("ferc", C.c_double)]


class AdgTrace(C.Structure):
    _fields_ = [("cap", C.c_int), ("n", C.c_int), ("k", _dp), ("ser", _dp), ("fer", _dp),
                ("ndmax", _dp), ("flags", _ip), ("loc", _ip)]


class ADGError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the engine library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run python -m adpres_amd._build")
    L = C.CDLL(path)
    L.adg_create.argtypes = [C.POINTER(AdgGeom), C.POINTER(C.c_void_p)]
    L.adg_create.restype = C.c_int
    L.adg_destroy.argtypes = [C.c_void_p]
    L.adg_destroy.restype = None
    L.adg_set_xs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]
    L.adg_set_state.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, C.c_int]
    L.adg_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.adg_outer.argtypes = [C.c_void_p, C.c_int, C.POINTER(AdgIter), C.POINTER(AdgTrace)]
    L.adg_powdis.argtypes = [C.c_void_p, _dp]
    L.adg_nodal_update.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _ip]
    L.adg_bicg.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
    L.adg_matvec.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.adg_last_timing.argtypes = [C.c_void_p, _dp]
    L.adg_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.adg_read_profile.argtypes = [C.c_void_p, _dp]
    L.adg_autotune.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_solver_info.argtypes = [C.c_void_p, _ip]
    L.adg_persistent_plan.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip]
    L.adg_create_batch.argtypes = [C.POINTER(AdgGeom), C.c_int, C.POINTER(C.c_void_p)]
    L.adg_core_count.argtypes = [C.c_void_p]
    L.adg_set_xs_core.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]
    L.adg_set_state_core.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double, _dp, C.c_int]
    L.adg_get_state_core.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.adg_outer_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(AdgIter), C.c_int, _ip, _ip, _dp]
    L.adg_trace_core.argtypes = [C.c_void_p, C.c_int, C.POINTER(AdgTrace)]
    L.adg_powdis_core.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_set_transient.argtypes = [C.c_void_p, _dp]
    L.adg_set_transient_core.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_leakage.argtypes = [C.c_void_p, _dp]
    L.adg_set_xs_tables.argtypes = [C.c_void_p, C.POINTER(AdgXsTables), _ip, _dp, _dp]
    L.adg_set_xs_tables_core.argtypes = [C.c_void_p, C.c_int, C.POINTER(AdgXsTables), _ip, _dp, _dp]
    L.adg_power_maps.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip, _dp, _ip]
    L.adg_power_maps_core.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip,
                                      _dp, _ip]
    L.adg_power_maps_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _ip, _dp, _ip]
    L.adg_flux_map.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, C.c_double, _dp, _ip]
    L.adg_flux_map_core.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_double, _dp, _ip]
    L.adg_set_xtab_tables.argtypes = [C.c_void_p, C.POINTER(AdgXtabTables), _ip, _dp]
    L.adg_set_xtab_tables_core.argtypes = [C.c_void_p, C.c_int, C.POINTER(AdgXtabTables), _ip, _dp]
    L.adg_update_xs.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp]
    L.adg_update_xs_core.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    L.adg_get_xs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
    L.adg_get_xs_core.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.adg_adjoint.argtypes = [C.c_void_p, C.POINTER(AdgIter), _dp]
    L.adg_get_adjoint.argtypes = [C.c_void_p, _dp, _dp]
    L.adg_set_adjoint.argtypes = [C.c_void_p, _dp]
    L.adg_backward.argtypes = [C.c_void_p, C.POINTER(AdgIter), C.c_double, _dp, _dp]
    L.adg_nodal_vjp.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.adg_backward_nodal.argtypes = [C.c_void_p, C.POINTER(AdgIter), C.c_int, C.c_double, _dp, C.c_double, C.c_int,
                                     _dp, _ip, _dp]
    L.adg_status_string.argtypes = [C.c_int]
    L.adg_status_string.restype = C.c_char_p
    L.adg_last_error.argtypes = [C.c_void_p]
    L.adg_last_error.restype = C.c_char_p
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _geom(p: Problem, precision: str, keep: dict):
    ix = np.ascontiguousarray(p.ix, np.int32); iy = np.ascontiguousarray(p.iy, np.int32)
    iz = np.ascontiguousarray(p.iz, np.int32)
    xd = np.ascontiguousarray(p.xdel, np.float64); yd = np.ascontiguousarray(p.ydel, np.float64)
    zd = np.ascontiguousarray(p.zdel, np.float64)
    keep.setdefault("geom", []).append((ix, iy, iz, xd, yd, zd))
    return AdgGeom(p.nnod, p.ng, p.nmat, p.nxx, p.nyy, p.nzz, _i(ix), _i(iy), _i(iz), _d(xd), _d(yd), _d(zd),
                   (C.c_int * 6)(*p.bc), PRECISIONS[precision])


def _xs_args(p: Problem):
    arrs = {k: np.ascontiguousarray(getattr(p, k), np.float64)
            for k in ("D", "sigr", "sigs", "nuf", "sigf", "chi", "dc", "exsrc")}
    mat = np.ascontiguousarray(p.mat, np.int32)
    return (arrs, mat), (_d(arrs["D"]), _d(arrs["sigr"]), _d(arrs["sigs"]), _d(arrs["nuf"]), _d(arrs["sigf"]),
                         _d(arrs["chi"]), _i(mat), _d(arrs["dc"]), _d(arrs["exsrc"]))


XS_KINDS = ("base", "boron", "ftem", "mtem", "cden", "rod")
XS_REFS = ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize")


def _xs_tables(t: dict, keep: list) -> "AdgXsTables":
    """adg_xs_tables from the dict form of the feedback tables: "nmat",
    "nb", the reference values XS_REFS, per kind of XS_KINDS a list of five
    arrays (sigtr, siga, nuf, sigf as (G, nmat), sigs as (G_to, G_from,
    nmat)) or None, and "fbmap" (nyy, nxx) or None."""
    s = AdgXsTables()
    s.nmat, s.nb = int(t["nmat"]), int(t.get("nb", 0))
    for k in XS_REFS:
        setattr(s, k, float(t.get(k, 0.0)))
    for i, kind in enumerate(XS_KINDS):
        if t.get(kind) is None:
            continue
        for j in range(5):
            a = np.ascontiguousarray(t[kind][j], np.float64)
            keep.append(a)
            s.tab[i][j] = _d(a)
    if t.get("fbmap") is not None:
        fb = np.ascontiguousarray(t["fbmap"], np.int32)
        keep.append(fb)
        s.fbmap = _i(fb)
    return s


def _opt(a, keep: list):
    if a is None:
        return None
    a = np.ascontiguousarray(a, np.float64)
    keep.append(a)
    return _d(a)


def _iter(p: Problem, ov: dict):
    return AdgIter(ov.get("nout", p.nout), ov.get("nin", p.nin), ov.get("nac", p.nac), ov.get("nupd", p.nupd),
                   KERNS[ov.get("kern", p.kern)], ov.get("serc", p.serc), ov.get("ferc", p.ferc))


class Engine:
    """One core on one GPU behind the C-ABI."""

    def __init__(self, p: Problem, precision: str = "fp64"):
        self.lib = load_library()
        self.p = p
        self.precision = precision
        self._keep = {}
        g = _geom(p, precision, self._keep)
        h = C.c_void_p()
        rc = self.lib.adg_create(C.byref(g), C.byref(h))
        if rc != ADG_OK:
            raise ADGError(rc, f"adg_create failed: {self.lib.adg_status_string(rc).decode()}")
        self.h = h
        self.set_xs(p)
        self.niter = 0
        self.trace = None

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.adg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != ADG_OK:
            msg = self.lib.adg_last_error(self.h).decode()
            raise ADGError(rc, f"{what}: {self.lib.adg_status_string(rc).decode()} ({msg})")

    # -- data ------------------------------------------------------------
    def set_xs(self, p: Problem):
        _, args = _xs_args(p)
        self._check(self.lib.adg_set_xs(self.h, *args), "adg_set_xs")

    def set_xs_tables(self, tables: dict):
        """Install the feedback tables of xsec::XS_updt on the device (with
        the problem's material map, chi and discontinuity factors); the node
        cross sections then come from update_xs."""
        keep = []
        t = _xs_tables(tables, keep)
        mat = np.ascontiguousarray(self.p.mat, np.int32)
        chi = np.ascontiguousarray(self.p.chi, np.float64)
        dc = np.ascontiguousarray(self.p.dc, np.float64)
        self._check(self.lib.adg_set_xs_tables(self.h, C.byref(t), _i(mat), _d(chi), _d(dc)), "adg_set_xs_tables")

    def set_xtab_tables(self, tables: dict):
        """Install the branch tables of a %XTAB deck (xsec::XStab_updt) in the
        flat form of adpres_amd.problem.branch_tables_from_dict, with the
        problem's material map and chi; update_xs then interpolates them."""
        keep = []

        def ip(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.int32)
            keep.append(a)
            return _i(a)
        t = AdgXtabTables(int(tables["nmat"]), int(tables.get("nb", 0)), ip(tables["dims"]),
                          _opt(tables["points"], keep), _opt(tables["xsec"], keep), _opt(tables.get("rxsec"), keep),
                          float(tables.get("coreh", 0.0)), float(tables.get("pos0", 0.0)),
                          float(tables.get("ssize", 0.0)), ip(tables.get("fbmap")))
        mat = np.ascontiguousarray(self.p.mat, np.int32)
        chi = np.ascontiguousarray(self.p.chi, np.float64)
        self._check(self.lib.adg_set_xtab_tables(self.h, C.byref(t), _i(mat), _d(chi)), "adg_set_xtab_tables")

    def update_xs(self, bcon: float, ftem=None, mtem=None, cden=None, bpos=None):
        """xsec::XS_updt on the device for the given T-H and rod parameters."""
        keep = []
        self._check(self.lib.adg_update_xs(self.h, float(bcon), _opt(ftem, keep), _opt(mtem, keep),
                                           _opt(cden, keep), _opt(bpos, keep)), "adg_update_xs")

    def get_xs(self) -> dict:
        """The node cross sections the context holds: D, sigr, nuf, sigf
        (G, N) and sigs (G_to, G_from, N), and the discontinuity factors dc
        (6, G, N)."""
        N, G = self.p.nnod, self.p.ng
        out = {k: np.zeros((G, N)) for k in ("D", "sigr", "nuf", "sigf")}
        out["sigs"] = np.zeros((G, G, N))
        out["dc"] = np.zeros((6, G, N))
        self._check(self.lib.adg_get_xs(self.h, *[_d(out[k]) for k in ("D", "sigr", "nuf", "sigf", "sigs", "dc")]),
                    "adg_get_xs")
        return out

    def reset(self):
        """Start the next outer() from f0 = 1, k = 1, D-hat = 0 (the reference's first call)."""
        self._check(self.lib.adg_set_state(self.h, None, None, 1.0, None, 1), "adg_set_state")

    def set_state(self, f0, fs0, Ke, dn):
        f0 = np.ascontiguousarray(f0, np.float64); fs0 = np.ascontiguousarray(fs0, np.float64)
        dn = np.ascontiguousarray(dn, np.float64)
        self._check(self.lib.adg_set_state(self.h, _d(f0), _d(fs0), float(Ke), _d(dn), 0), "adg_set_state")

    def state(self):
        N, G = self.p.nnod, self.p.ng
        f0 = np.zeros((G, N)); fs0 = np.zeros(N); dn = np.zeros((6, G, N)); df = np.zeros((6, G, N))
        Ke = C.c_double()
        self._check(self.lib.adg_get_state(self.h, _d(f0), _d(fs0), C.byref(Ke), _d(dn), _d(df)), "adg_get_state")
        return dict(f0=f0, fs0=fs0, Ke=Ke.value, dn=dn, df=df)

    @property
    def Ke(self) -> float:
        Ke = C.c_double()
        self._check(self.lib.adg_get_state(self.h, None, None, C.byref(Ke), None, None), "adg_get_state")
        return Ke.value

    # -- hot path --------------------------------------------------------
    def outer(self, adjoint: Optional[bool] = None, raise_on_error: bool = True,
              fixed_source: Optional[bool] = None, transient: bool = False, nodal: bool = True, **ov):
        """cmfd::outer, outer_ad (adjoint), outer_fs (fixed_source: the
        forward outer with the external source and k held at its value) or
        outer_tr (transient: the external source and the fission weight of
        set_transient, no k); the default follows the problem's %MODE.
        nodal=False with adjoint: outer_ad called with popt = 0, the D-hats of
        the forward state kept (no nodal update)."""
        p = self.p
        it = _iter(p, ov)
        if adjoint is None:
            adjoint = p.mode == "ADJOINT"
        if fixed_source is None:
            fixed_source = p.mode == "FIXEDSRC"
        mode = ADG_TRANSIENT if transient else (
            ADG_FIXED_SOURCE if fixed_source else (
                (ADG_ADJOINT if nodal else ADG_ADJOINT_FROZEN) if adjoint else ADG_FORWARD))
        cap = it.nout
        t = {k: np.zeros(cap) for k in ("k", "ser", "fer", "ndmax")}
        t["flags"] = np.zeros(cap, np.int32)
        t["loc"] = np.zeros(3 * cap, np.int32)
        tr = AdgTrace(cap, 0, _d(t["k"]), _d(t["ser"]), _d(t["fer"]), _d(t["ndmax"]), _i(t["flags"]), _i(t["loc"]))
        rc = self.lib.adg_outer(self.h, mode, C.byref(it), C.byref(tr))
        n = tr.n
        self.niter = n
        self.trace = {k: (v[:n] if k != "loc" else v[:3 * n].reshape(-1, 3)) for k, v in t.items()}
        if raise_on_error:
            self._check(rc, "adg_outer")
        return rc

    def set_transient(self, wtr) -> None:
        """Per-node fission weight 1 - beta + dfis of the transient outer."""
        w = np.ascontiguousarray(wtr, np.float64)
        self._check(self.lib.adg_set_transient(self.h, _d(w)), "adg_set_transient")

    def leakage(self) -> np.ndarray:
        """nodal::Lxyz of every node and group: (3, G, N)."""
        out = np.zeros((3, self.p.ng, self.p.nnod))
        self._check(self.lib.adg_leakage(self.h, _d(out)), "adg_leakage")
        return out

    def powdis(self) -> np.ndarray:
        out = np.zeros(self.p.nnod)
        self._check(self.lib.adg_powdis(self.h, _d(out)), "adg_powdis")
        return out

    def power_maps(self, nx: int, ny: int, nz: int, xdiv, ydiv, zdiv, noalloc: bool = False):
        """The radial and axial power maps the reference prints (io::AsmPow,
        io::AxiPow) from the device's PowDis: (fasm (ny, nx), (i, j) of its
        maximum, faxi (nz,), plane of its maximum), positions 1-based.
        noalloc: through adg_power_maps_noalloc (no device allocation)."""
        xd, yd, zd = (np.ascontiguousarray(a, np.int32) for a in (xdiv, ydiv, zdiv))
        fasm = np.zeros((ny, nx)); faxi = np.zeros(nz)
        pm = np.zeros(2, np.int32); am = np.zeros(1, np.int32)
        fn = self.lib.adg_power_maps_noalloc if noalloc else self.lib.adg_power_maps
        self._check(fn(self.h, nx, ny, nz, _i(xd), _i(yd), _i(zd), _d(fasm), _i(pm), _d(faxi), _i(am)),
                    "adg_power_maps_noalloc" if noalloc else "adg_power_maps")
        return fasm, (int(pm[0]), int(pm[1])), faxi, int(am[0])

    def flux_map(self, nx: int, ny: int, xdiv, ydiv, norm: float = 1.0):
        """The radial flux map the reference prints (io::AsmFlux) from the
        device's flux: (fasm (G, ny, nx), negative-flux flag); norm <= 0
        leaves it unnormalised."""
        xd, yd = (np.ascontiguousarray(a, np.int32) for a in (xdiv, ydiv))
        fasm = np.zeros((self.p.ng, ny, nx))
        negf = np.zeros(1, np.int32)
        self._check(self.lib.adg_flux_map(self.h, nx, ny, _i(xd), _i(yd), float(norm), _d(fasm), _i(negf)),
                    "adg_flux_map")
        return fasm, int(negf[0])

    def nodal_update(self, kern: Optional[str] = None, cmode: int = 1):
        nd = C.c_double()
        loc = np.zeros(3, np.int32)
        rc = self.lib.adg_nodal_update(self.h, KERNS[kern or self.p.kern], cmode, C.byref(nd), _i(loc))
        self._check(rc, "adg_nodal_update")
        return nd.value, tuple(int(v) for v in loc)

    def bicg(self, g: int, nin: int, b, x):
        b = np.ascontiguousarray(b, np.float64)
        x = np.array(x, np.float64, copy=True)
        self._check(self.lib.adg_bicg(self.h, g, nin, _d(b), _d(x)), "adg_bicg")
        return x

    def matvec(self, g: int, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._check(self.lib.adg_matvec(self.h, g, _d(x), _d(y)), "adg_matvec")
        return y

    def autotune(self, nin: Optional[int] = None) -> dict:
        """Pick the fastest BiCGSTAB form and launch shape (see adg_autotune);
        shape 0 is the persistent kernel."""
        out = np.zeros(4)
        self._check(self.lib.adg_autotune(self.h, int(nin or self.p.nin), _d(out)), "adg_autotune")
        return dict(shape=int(out[0]), group_solve_ms=float(out[1]), grid=int(out[2]), candidates=int(out[3]))

    def solver_info(self) -> dict:
        """How the group solves run (adg_solver_info)."""
        i = np.zeros(8, np.int32)
        self._check(self.lib.adg_solver_info(self.h, _i(i)), "adg_solver_info")
        return dict(persistent=bool(i[0]), blocks=int(i[1]), threads=int(i[2]), brick=(int(i[3]), int(i[4]), int(i[5])),
                    nodes_per_thread=int(i[6]), halo=int(i[7]))

    # -- differentiation ----------------------------------------------------
    def adjoint(self, **ov) -> float:
        """Discrete adjoint of the converged CMFD operator (D-hat fixed); returns its eigenvalue."""
        it = _iter(self.p, ov)
        ka = C.c_double()
        self._check(self.lib.adg_adjoint(self.h, C.byref(it), C.byref(ka)), "adg_adjoint")
        return ka.value

    def adjoint_flux(self):
        N, G = self.p.nnod, self.p.ng
        psi = np.zeros((G, N))
        pfp = C.c_double()
        self._check(self.lib.adg_get_adjoint(self.h, _d(psi), C.byref(pfp)), "adg_get_adjoint")
        return psi, pfp.value

    def set_adjoint_flux(self, psi):
        psi = np.ascontiguousarray(psi, np.float64)
        self._check(self.lib.adg_set_adjoint(self.h, _d(psi)), "adg_set_adjoint")

    def backward(self, dJdk: float, dJdp=None, **ov) -> np.ndarray:
        """Per-node gradient (P, N) of dJdk * k + <dJdp, node powers>; rows D_g, sigma_a_g,
        nu-sigma_f_g, sigma_s(g->h) (g != h)."""
        N, G = self.p.nnod, self.p.ng
        it = _iter(self.p, ov)
        P = 3 * G + G * (G - 1)
        out = np.zeros((P, N))
        dp = None if dJdp is None else np.ascontiguousarray(dJdp, np.float64)
        self._check(self.lib.adg_backward(self.h, C.byref(it), float(dJdk), None if dp is None else _d(dp), _d(out)),
                    "adg_backward")
        return out

    def nodal_vjp(self, w, kern: Optional[str] = None):
        """Transpose of one nodal D-hat update at the current state for the
        cotangent w (6, G, N) of the new D-hat: (flux cotangent (G, N), k
        cotangent, old-D-hat cotangent (6, G, N), parameter cotangents (P, N))."""
        N, G = self.p.nnod, self.p.ng
        P = 3 * G + G * (G - 1)
        w = np.ascontiguousarray(w, np.float64)
        phib = np.zeros((G, N)); dnb = np.zeros((6, G, N)); grad = np.zeros((P, N))
        kb = C.c_double()
        self._check(self.lib.adg_nodal_vjp(self.h, KERNS[kern or self.p.kern], _d(w), _d(phib), C.byref(kb), _d(dnb),
                                           _d(grad)), "adg_nodal_vjp")
        return phib, kb.value, dnb, grad

    def backward_nodal(self, dJdk: float, dJdp=None, kern: Optional[str] = None, tol: float = 1e-9,
                       maxit: int = 5000, **ov):
        """Per-node gradient (P, N) of dJdk * k + <dJdp, node powers> for the
        full nodal model (D-hat moving with the cross sections through the
        nodal update), with the fixed-point steps and the last relative change."""
        N, G = self.p.nnod, self.p.ng
        it = _iter(self.p, ov)
        P = 3 * G + G * (G - 1)
        out = np.zeros((P, N))
        dp = None if dJdp is None else np.ascontiguousarray(dJdp, np.float64)
        nit = C.c_int()
        res = C.c_double()
        self._check(self.lib.adg_backward_nodal(self.h, C.byref(it), KERNS[kern or self.p.kern], float(dJdk),
                                                None if dp is None else _d(dp), float(tol), int(maxit), _d(out),
                                                C.byref(nit), C.byref(res)), "adg_backward_nodal")
        self.nodal_adjoint_steps, self.nodal_adjoint_resid = nit.value, res.value
        return out

    def last_timing_ms(self):
        t = np.zeros(4)
        self._check(self.lib.adg_last_timing(self.h, _d(t)), "adg_last_timing")
        return t

    def set_profiling(self, on: bool):
        self._check(self.lib.adg_set_profiling(self.h, 1 if on else 0), "adg_set_profiling")

    def read_profile(self):
        t = np.zeros(6)
        self._check(self.lib.adg_read_profile(self.h, _d(t)), "adg_read_profile")
        return dict(k1_ms=t[0], k1_n=int(t[1]), gsolve_ms=t[2], gsolve_n=int(t[3]), nodal_ms=t[4],
                    nodal_n=int(t[5]))


class Batch:
    """Many independent cores with one group count in one context: every
    kernel launch and every replay of the captured outer iteration covers all
    cores, and each core keeps its own k, schedule, convergence test and
    status (the batching form; the reference solves one core per process).
    Each core runs with its own %ITER (nac, nupd, serc, ferc); nout, nin and
    the nodal kernel must agree, and overrides apply to every core."""

    def __init__(self, problems, precision: str = "fp64"):
        self.lib = load_library()
        self.problems = list(problems)
        if not self.problems or len({q.ng for q in self.problems}) != 1:
            raise ValueError("a batch needs at least one core and one group count")
        self.precision = precision
        self._keep = {}
        geoms = (AdgGeom * len(self.problems))(*[_geom(q, precision, self._keep) for q in self.problems])
        h = C.c_void_p()
        rc = self.lib.adg_create_batch(geoms, len(self.problems), C.byref(h))
        if rc != ADG_OK:
            raise ADGError(rc, f"adg_create_batch failed: {self.lib.adg_status_string(rc).decode()}")
        self.h = h
        for b, q in enumerate(self.problems):
            _, args = _xs_args(q)
            self._check(self.lib.adg_set_xs_core(self.h, b, *args), "adg_set_xs_core")
        self.status = self.niter = self.keff = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.adg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != ADG_OK:
            msg = self.lib.adg_last_error(self.h).decode()
            raise ADGError(rc, f"{what}: {self.lib.adg_status_string(rc).decode()} ({msg})")

    def __len__(self):
        return len(self.problems)

    def reset(self):
        for b in range(len(self)):
            self._check(self.lib.adg_set_state_core(self.h, b, None, None, 1.0, None, 1), "adg_set_state_core")

    def outer(self, adjoint: bool = False, fixed_source: bool = False, **ov):
        """Solve every core; returns (status, outers, k) arrays.  Per-core
        failures are reported in `status`, not raised."""
        n = len(self)
        its = (AdgIter * n)(*[_iter(q, ov) for q in self.problems])
        st = np.zeros(n, np.int32); ni = np.zeros(n, np.int32); ke = np.zeros(n)
        mode = ADG_FIXED_SOURCE if fixed_source else (ADG_ADJOINT if adjoint else ADG_FORWARD)
        rc = self.lib.adg_outer_batch(self.h, mode, its, n, _i(st), _i(ni), _d(ke))
        if rc in (ADG_E_ARG, ADG_E_NOT_SET, ADG_E_HIP):
            self._check(rc, "adg_outer_batch")
        self.status, self.niter, self.keff = st, ni, ke
        return st, ni, ke

    def set_profiling(self, on: bool):
        self._check(self.lib.adg_set_profiling(self.h, 1 if on else 0), "adg_set_profiling")

    def read_profile(self):
        t = np.zeros(6)
        self._check(self.lib.adg_read_profile(self.h, _d(t)), "adg_read_profile")
        return dict(k1_ms=t[0], k1_n=int(t[1]), gsolve_ms=t[2], gsolve_n=int(t[3]), nodal_ms=t[4],
                    nodal_n=int(t[5]))

    def autotune(self, nin: Optional[int] = None) -> dict:
        """Pick the fastest cells-per-block of the flat kernels (see adg_autotune)."""
        out = np.zeros(4)
        self._check(self.lib.adg_autotune(self.h, int(nin or self.problems[0].nin), _d(out)), "adg_autotune")
        return dict(shape=int(out[0]), group_solve_ms=float(out[1]), grid=int(out[2]), candidates=int(out[3]))

    def powdis(self, b: int) -> np.ndarray:
        out = np.zeros(self.problems[b].nnod)
        self._check(self.lib.adg_powdis_core(self.h, b, _d(out)), "adg_powdis_core")
        return out

    def set_xs_tables(self, b: int, tables: dict):
        """Feedback tables of core b (see Engine.set_xs_tables)."""
        q, keep = self.problems[b], []
        t = _xs_tables(tables, keep)
        mat = np.ascontiguousarray(q.mat, np.int32)
        chi = np.ascontiguousarray(q.chi, np.float64)
        dc = np.ascontiguousarray(q.dc, np.float64)
        self._check(self.lib.adg_set_xs_tables_core(self.h, b, C.byref(t), _i(mat), _d(chi), _d(dc)),
                    "adg_set_xs_tables_core")

    def update_xs(self, b: int, bcon: float, ftem=None, mtem=None, cden=None, bpos=None):
        """xsec::XS_updt of core b on the device (see Engine.update_xs)."""
        keep = []
        self._check(self.lib.adg_update_xs_core(self.h, b, float(bcon), _opt(ftem, keep), _opt(mtem, keep),
                                                _opt(cden, keep), _opt(bpos, keep)), "adg_update_xs_core")

    def state(self, b: int):
        q = self.problems[b]
        N, G = q.nnod, q.ng
        f0 = np.zeros((G, N)); fs0 = np.zeros(N); dn = np.zeros((6, G, N)); df = np.zeros((6, G, N))
        Ke = C.c_double()
        self._check(self.lib.adg_get_state_core(self.h, b, _d(f0), _d(fs0), C.byref(Ke), _d(dn), _d(df)),
                    "adg_get_state_core")
        return dict(f0=f0, fs0=fs0, Ke=Ke.value, dn=dn, df=df)

    def trace(self, b: int, cap: int):
        t = {k: np.zeros(cap) for k in ("k", "ser", "fer", "ndmax")}
        t["flags"] = np.zeros(cap, np.int32)
        t["loc"] = np.zeros(3 * cap, np.int32)
        tr = AdgTrace(cap, 0, _d(t["k"]), _d(t["ser"]), _d(t["fer"]), _d(t["ndmax"]), _i(t["flags"]), _i(t["loc"]))
        self._check(self.lib.adg_trace_core(self.h, b, C.byref(tr)), "adg_trace_core")
        n = tr.n
        return {k: (v[:n] if k != "loc" else v[:3 * n].reshape(-1, 3)) for k, v in t.items()}


def persistent_plan(nxx: int, nyy: int, nzz: int, precision: str = "fp64", ncu: int = 256) -> dict:
    """The persistent BiCGSTAB kernel's brick plan for a node box, made and
    checked on the host (adg_persistent_plan; no GPU needed)."""
    L = load_library()
    out = np.zeros(8, np.int32)
    rc = L.adg_persistent_plan(int(nxx), int(nyy), int(nzz), PRECISIONS[precision], int(ncu), _i(out))
    if rc != ADG_OK:
        raise ADGError(rc, "adg_persistent_plan: bad argument")
    return dict(ok=bool(out[0]), blocks=int(out[1]), brick=(int(out[2]), int(out[3]), int(out[4])),
                nodes_per_thread=int(out[5]), halo=int(out[6]), checked=out[7] == 0)
