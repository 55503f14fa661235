"""Continuous-batching engine for independent Dyson solves (the C ABI's cntrx_engine_*).

Each request is one propagation of the reference's `dyson(...)` sequence
(libcntr/cntr/cntr_dyson_impl.hpp:1725-1734): dyson_mat, dyson_start and dyson_timestep
for n = SolveOrder+1 .. nt, with H(t) and a bath self-energy Sigma = V^2 g[H_bath]
(cntr::green_from_H, libcntr/cntr/cntr_equilibrium_impl.hpp:873-953).  The engine keeps
all admitted solves resident in one device pool and advances every live solve by one
timestep per step(); requests queue until their footprint fits.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CNTRX_MAT_FIXPOINT, CntrxHerm, CntrxRequest, check

NO_GRAPH = 1


def make_request(size1, nt, ntau, beta, h, H, Hbath, V, mu=0.0, sig=-1, SolveOrder=5,
                 mat_method=CNTRX_MAT_FIXPOINT):
    """request dict; H: complex (nt+2, size1, size1) = H(t), t = -1..nt; Hbath: (size1, size1)"""
    H = np.ascontiguousarray(H, np.complex128).reshape(nt + 2, size1, size1)
    Hbath = np.ascontiguousarray(Hbath, np.complex128).reshape(size1, size1)
    return {"size1": size1, "nt": nt, "ntau": ntau, "beta": beta, "h": h, "H": H, "Hbath": Hbath,
            "V": V, "mu": mu, "sig": sig, "SolveOrder": SolveOrder, "mat_method": mat_method}


def _c_request(r):
    c = CntrxRequest()
    c.size1, c.nt, c.ntau, c.sig = r["size1"], r["nt"], r["ntau"], r["sig"]
    c.SolveOrder, c.mat_method = r["SolveOrder"], r["mat_method"]
    c.beta, c.h, c.mu, c.V = r["beta"], r["h"], r["mu"], r["V"]
    c.H = r["H"].ctypes.data
    c.Hbath = r["Hbath"].ctypes.data
    return c


def footprint(r):
    """device bytes the engine reserves for request r"""
    return int(_lib.load().cntrx_request_footprint(ctypes.byref(_c_request(r))))


class Engine:
    def __init__(self, budget_bytes, max_live=0, graph=True):
        self._h = ctypes.c_void_p()
        check(_lib.load().cntrx_engine_create(int(budget_bytes), int(max_live),
                                              0 if graph else NO_GRAPH, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(_lib.load().cntrx_engine_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, r):
        rid = ctypes.c_long(0)
        check(_lib.load().cntrx_engine_submit(self._h, ctypes.byref(_c_request(r)), ctypes.byref(rid)))
        return rid.value

    def step(self):
        live, queued = ctypes.c_int(0), ctypes.c_int(0)
        check(_lib.load().cntrx_engine_step(self._h, ctypes.byref(live), ctypes.byref(queued)))
        return live.value, queued.value

    def poll(self, cap=4096):
        ids = np.zeros(cap, np.int64)
        lat = np.zeros(cap, np.float64)
        n = ctypes.c_long(0)
        check(_lib.load().cntrx_engine_poll(self._h, cap, ids.ctypes.data, lat.ctypes.data, ctypes.byref(n)))
        return list(zip(ids[:n.value].tolist(), lat[:n.value].tolist()))

    def drain(self):
        check(_lib.load().cntrx_engine_drain(self._h))

    def result_desc(self, rid):
        d = CntrxHerm()
        check(_lib.load().cntrx_engine_result(self._h, rid, ctypes.byref(d)))
        return d

    def result(self, rid):
        """completed G as host arrays {ret, les, tv, mat} (packed layout)"""
        from .lattice import herm_to_host
        return herm_to_host(self.result_desc(rid))

    def release(self, rid):
        check(_lib.load().cntrx_engine_release(self._h, rid))

    def tuning(self):
        """autotuner decisions made while this engine ran"""
        rows = np.zeros((64, 5))
        n = ctypes.c_int(0)
        check(_lib.load().cntrx_engine_tuning(self._h, 64, rows.ctypes.data, ctypes.byref(n)))
        return [{"size1": int(r[0]), "history_from": int(r[1]), "panel": int(r[2]),
                 "ms_default": float(r[3]), "ms_alternative": float(r[4])} for r in rows[:n.value]]

    def stats(self):
        out = (ctypes.c_double * 8)()
        check(_lib.load().cntrx_engine_stats(self._h, out))
        keys = ("steps", "timesteps", "admitted", "completed", "epochs", "pool_used", "pool_peak",
                "pool_cap")
        return dict(zip(keys, list(out)))
