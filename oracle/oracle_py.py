"""ctypes binding of the CPU restatement (oracle/libfesom_oracle.so) -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Dict, Iterable, Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfesom_oracle.so"

HOR = {"UPW1": 0, "MUSCL": 1, "MFCT": 2}
VER = {"UPW1": 0, "QR4C": 1, "CDIFF": 2, "PPM": 3}
LIM = {"NONE": 0, "FCT": 1}

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class FoMesh(ctypes.Structure):
    _fields_ = [("nl", ctypes.c_int), ("nod2D", ctypes.c_int), ("elem2D", ctypes.c_int),
                ("edge2D", ctypes.c_int), ("max_adj", ctypes.c_int),
                ("elem2D_nodes", _ip), ("edges", _ip), ("edge_tri", _ip),
                ("nod_in_elem2D", _ip), ("nod_in_elem2D_num", _ip),
                ("nlevels", _ip), ("ulevels", _ip),
                ("nlevels_nod2D", _ip), ("ulevels_nod2D", _ip),
                ("nlevels_nod2D_min", _ip), ("ulevels_nod2D_max", _ip),
                ("edge_up_dn_tri", _ip), ("nboundary_lay", _ip),
                ("elem_area", _dp), ("area", _dp), ("areasvol", _dp),
                ("edge_dxdy", _dp), ("edge_cross_dxdy", _dp), ("elem_cos", _dp),
                ("gradient_sca", _dp), ("hnode", _dp), ("hnode_new", _dp), ("helem", _dp),
                ("Z_3d_n", _dp), ("zbar_3d_n", _dp)]


class FoScheme(ctypes.Structure):
    _fields_ = [("hor", ctypes.c_int), ("ver", ctypes.c_int), ("lim", ctypes.c_int),
                ("ph", ctypes.c_double), ("pv", ctypes.c_double), ("use_wsplit", ctypes.c_int)]


class FoWork(ctypes.Structure):
    _fields_ = [(name, _dp) for name in (
        "tr_xy", "edge_up_dn_grad", "fct_LO", "adv_flux_hor", "adv_flux_ver", "fct_ttf_max",
        "fct_ttf_min", "fct_plus", "fct_minus", "tvert_max", "tvert_min", "dttf_h", "dttf_v")]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.fo_tracer_step.argtypes = [ctypes.POINTER(FoMesh), ctypes.POINTER(FoScheme),
                                     ctypes.c_double, _dp, _dp, _dp, _dp, _dp,
                                     ctypes.POINTER(FoWork)]
        L.fo_tracer_step.restype = ctypes.c_int
        L.fo_do_oce_adv_tra.argtypes = [ctypes.POINTER(FoMesh), ctypes.POINTER(FoScheme),
                                        ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.POINTER(FoWork)]
        L.fo_do_oce_adv_tra.restype = ctypes.c_int
        L.fo_tracer_step_ab.argtypes = [ctypes.POINTER(FoMesh), ctypes.POINTER(FoScheme),
                                        ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.POINTER(FoWork)]
        L.fo_tracer_step_ab.restype = ctypes.c_int
        L.fo_adv_tra_vert_impl.argtypes = [ctypes.POINTER(FoMesh), ctypes.c_double, _dp, _dp]
        L.fo_adv_tra_vert_impl.restype = ctypes.c_int
        L.fo_work_alloc.argtypes = [ctypes.POINTER(FoMesh), ctypes.POINTER(FoWork)]
        L.fo_work_alloc.restype = ctypes.c_int
        L.fo_work_free.argtypes = [ctypes.POINTER(FoWork)]
        L.fo_tracer_gradients.argtypes = [ctypes.POINTER(FoMesh), _dp, _dp]
        L.fo_fill_up_dn_grad.argtypes = [ctypes.POINTER(FoMesh), _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def ab2_interpolate(values_old: np.ndarray, values: np.ndarray, epsilon: float) -> np.ndarray:
    """FESOM2's second-order Adams-Bashforth time level of the high-order fluxes,
    valuesAB = -(0.5+epsilon)*valuesold + (1.5+epsilon)*values (epsilon of lib/oce_modules.F90:75,
    valuesold of lib/MOD_TRACER.F90:13), evaluated in that operation order."""
    return -(0.5 + epsilon) * values_old + (1.5 + epsilon) * values


class Oracle:
    """CPU restatement bound to one mesh and velocity field."""

    def __init__(self, case):
        if getattr(case, "dt_node", None) is not None:
            raise ValueError("the restatement advances every node with one time step; "
                             "run the members of a union with their own dt separately")
        m = case.mesh
        self.case = case
        self._keep = []

        def ints(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return _p(a, ctypes.c_int)

        def dbls(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _p(a, ctypes.c_double)

        self.mesh = FoMesh(
            m.nl, m.nod2D, m.elem2D, m.edge2D, m.max_adj,
            ints(m.elem2D_nodes), ints(m.edges), ints(m.edge_tri), ints(m.nod_in_elem2D),
            ints(m.nod_in_elem2D_num), ints(m.nlevels), ints(m.ulevels), ints(m.nlevels_nod2D),
            ints(m.ulevels_nod2D), ints(m.nlevels_nod2D_min), ints(m.ulevels_nod2D_max),
            ints(m.edge_up_dn_tri), ints(m.nboundary_lay),
            dbls(m.elem_area), dbls(m.area), dbls(m.areasvol), dbls(m.edge_dxdy),
            dbls(m.edge_cross_dxdy), dbls(m.elem_cos), dbls(m.gradient_sca), dbls(m.hnode),
            dbls(m.hnode_new), dbls(m.helem), dbls(m.Z_3d_n), dbls(m.zbar_3d_n))
        self.uv = np.ascontiguousarray(case.uv, dtype=np.float64)
        self.w = np.ascontiguousarray(case.w, dtype=np.float64)
        self.wi = np.ascontiguousarray(case.wi, dtype=np.float64)
        self.we = np.ascontiguousarray(case.we, dtype=np.float64)
        self.use_wsplit = int(bool(getattr(case, "use_wsplit", False)))
        self.work = FoWork()
        if lib().fo_work_alloc(ctypes.byref(self.mesh), ctypes.byref(self.work)) != 0:
            raise MemoryError("oracle work allocation failed")

    def __del__(self):
        try:
            lib().fo_work_free(ctypes.byref(self.work))
        except Exception:
            pass

    def scheme(self, sc) -> FoScheme:
        return FoScheme(HOR[sc.hor], VER[sc.ver], LIM[sc.lim], float(sc.ph), float(sc.pv),
                        self.use_wsplit)

    def step(self, values: np.ndarray, sc, dt: float) -> None:
        """Advance one tracer (N, nl-1) float64 C-contiguous array in place by one step."""
        assert values.dtype == np.float64 and values.flags.c_contiguous
        s = self.scheme(sc)
        rc = lib().fo_tracer_step(ctypes.byref(self.mesh), ctypes.byref(s), dt, _p(self.uv, ctypes.c_double),
                                  _p(self.w, ctypes.c_double), _p(self.wi, ctypes.c_double),
                                  _p(self.we, ctypes.c_double), _p(values, ctypes.c_double),
                                  ctypes.byref(self.work))
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")

    def step_ab(self, values: np.ndarray, values_ab: np.ndarray, sc, dt: float) -> None:
        """One step of one tracer with a separate valuesAB (both (N, nl-1) float64)."""
        assert values.dtype == np.float64 and values.flags.c_contiguous
        tab = np.ascontiguousarray(values_ab, dtype=np.float64)
        s = self.scheme(sc)
        rc = lib().fo_tracer_step_ab(ctypes.byref(self.mesh), ctypes.byref(s), dt, _p(self.uv, ctypes.c_double),
                                     _p(self.w, ctypes.c_double), _p(self.wi, ctypes.c_double),
                                     _p(self.we, ctypes.c_double), _p(values, ctypes.c_double),
                                     _p(tab, ctypes.c_double), ctypes.byref(self.work))
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")

    def work_arrays(self) -> Dict[str, np.ndarray]:
        m = self.case.mesh
        L = m.nl - 1
        shapes = {"fct_LO": (m.nod2D, L), "adv_flux_hor": (m.edge2D, L),
                  "adv_flux_ver": (m.nod2D, m.nl), "dttf_h": (m.nod2D, L), "dttf_v": (m.nod2D, L),
                  "edge_up_dn_grad": (max(m.edge2D, m.elem2D), L, 4), "tr_xy": (m.elem2D, L, 2)}
        out = {}
        for k, shp in shapes.items():
            ptr = getattr(self.work, k)
            out[k] = np.ctypeslib.as_array(ptr, shape=(int(np.prod(shp)),)).reshape(shp).copy()
        out["edge_up_dn_grad"] = out["edge_up_dn_grad"][: m.edge2D]
        return out

    def run(self, nsteps: Optional[int] = None, save_steps: Iterable[int] = (),
            work_step1: bool = False, ab_mode: int = 0, ab_epsilon: float = 0.1):
        """Advance all tracers; returns ({step: (ntr, N, nl-1)}, work arrays of tracer 0 step 1).

        ab_mode 0: valuesAB := values every step; 1: valuesAB frozen at the initial values (the
        reference's drivers, src/fesom_analytic.F90:335 and lib/tracer_c_interface.F90:613);
        2: FESOM2's AB2 interpolation valuesAB = -(0.5+eps)*valuesold + (1.5+eps)*values, then
        valuesold := values (eps as lib/oce_modules.F90:75), valuesold starting at the values."""
        case = self.case
        ns = case.nsteps if nsteps is None else nsteps
        vals = [np.ascontiguousarray(case.tracers[i], dtype=np.float64).copy()
                for i in range(case.ntracers)]
        ab = [v.copy() for v in vals]
        old = [v.copy() for v in vals]
        save = set(save_steps)
        out = {}
        work = None
        for s in range(1, ns + 1):
            for i in range(case.ntracers):
                if ab_mode != 0:
                    if ab_mode == 2:
                        ab[i] = ab2_interpolate(old[i], vals[i], ab_epsilon)
                        old[i] = vals[i].copy()
                    self.step_ab(vals[i], ab[i], case.schemes[i], case.dt)
                elif work_step1 and s == 1 and i == 0:
                    # gradients are clobbered by the limiter: capture them before advection
                    self.step_with_capture(vals[i], case.schemes[i], case.dt)
                    work = self._captured
                else:
                    self.step(vals[i], case.schemes[i], case.dt)
            if s in save:
                out[s] = np.stack(vals)
        return out, work

    def step_with_capture(self, values, sc, dt):
        m = self.case.mesh
        L = m.nl - 1
        g = np.zeros((m.elem2D, L, 2))
        egrad = np.zeros((max(m.edge2D, m.elem2D), L, 4))
        lib().fo_tracer_gradients(ctypes.byref(self.mesh), _p(values, ctypes.c_double), _p(g, ctypes.c_double))
        lib().fo_fill_up_dn_grad(ctypes.byref(self.mesh), _p(g, ctypes.c_double), _p(egrad, ctypes.c_double))
        self.step(values, sc, dt)
        self._captured = self.work_arrays()
        self._captured["edge_up_dn_grad"] = egrad[: m.edge2D]


class DropIn:
    """do_oce_adv_tra on the CPU restatement with independent values / valuesAB."""

    def __init__(self, oracle: Oracle):
        self.o = oracle

    def run(self, values: np.ndarray, values_ab: np.ndarray, sc, dt: float):
        o = self.o
        m = o.case.mesh
        L = m.nl - 1
        ttf = np.ascontiguousarray(values, dtype=np.float64)
        tab = np.ascontiguousarray(values_ab, dtype=np.float64)
        wk = o.work
        lib().fo_tracer_gradients(ctypes.byref(o.mesh), _p(tab, ctypes.c_double), wk.tr_xy)
        lib().fo_fill_up_dn_grad(ctypes.byref(o.mesh), wk.tr_xy, wk.edge_up_dn_grad)
        for name in ("dttf_h", "dttf_v"):
            arr = np.ctypeslib.as_array(getattr(wk, name), shape=(m.nod2D * L,))
            arr[:] = 0.0
        s = o.scheme(sc)
        rc = lib().fo_do_oce_adv_tra(ctypes.byref(o.mesh), ctypes.byref(s), dt,
                                     _p(o.uv, ctypes.c_double), _p(o.w, ctypes.c_double),
                                     _p(o.wi, ctypes.c_double), _p(o.we, ctypes.c_double),
                                     _p(ttf, ctypes.c_double), _p(tab, ctypes.c_double),
                                     ctypes.byref(wk))
        if rc != 0:
            raise RuntimeError("oracle do_oce_adv_tra failed")
        return o.work_arrays()
