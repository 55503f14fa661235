"""Continuous batching of independent Dyson solves: the C ABI's cntrx_server_* (C++, in
libcntrx.so) behind a small Python class.

Each request is the reference's dyson(...) sequence (libcntr/cntr/cntr_dyson_impl.hpp:
1725-1734) on its own G: dyson_mat, dyson_start, then dyson_timestep for n = SolveOrder+1..nt,
with H(t) and a bath self-energy Sigma = V^2 g[H_bath] (green_from_H,
libcntr/cntr/cntr_equilibrium_impl.hpp:873-953).  The server keeps every admitted solve
resident in one device pool; each step() retires solves whose last timestep has completed,
admits queued requests (first come first served) while their footprint fits, and advances
every live solve one timestep in one batched call over solves of any size1 and n.  No
reference counterpart: it replaces a caller's loop over solves.  A solve's result does not
depend on what else is in the batch, bit for bit.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CntrxHerm, CntrxRequest, check


def _c_request(r):
    c = CntrxRequest()
    c.size1, c.nt, c.ntau, c.sig = r["size1"], r["nt"], r["ntau"], r["sig"]
    c.SolveOrder, c.mat_method = r["SolveOrder"], r["mat_method"]
    c.beta, c.h, c.mu, c.V = r["beta"], r["h"], r["mu"], r["V"]
    c.H = r["H"].ctypes.data
    c.Hbath = r["Hbath"].ctypes.data
    return c


def footprint(r):
    """device bytes the server holds for request r (G, Sigma, H); 0 if r is invalid"""
    return int(_lib.load().cntrx_request_footprint(ctypes.byref(_c_request(r))))


class Result:
    """a completed request's G in the server's pool (valid until release)"""

    def __init__(self, d):
        self.desc = d
        self.nt, self.ntau, self.size1, self.sig = d.nt, d.ntau, d.size1, d.sig

    def to_numpy(self):
        from .lattice import herm_to_host
        return herm_to_host(self.desc)


class BatchServer:
    def __init__(self, budget_bytes):
        self._h = ctypes.c_void_p()
        check(_lib.load().cntrx_server_create(int(budget_bytes), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(_lib.load().cntrx_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, r):
        rid = ctypes.c_long(0)
        check(_lib.load().cntrx_server_submit(self._h, ctypes.byref(_c_request(r)), ctypes.byref(rid)))
        return rid.value

    def step(self):
        """one tick; returns (live, queued)"""
        live, queued = ctypes.c_int(0), ctypes.c_int(0)
        check(_lib.load().cntrx_server_step(self._h, ctypes.byref(live), ctypes.byref(queued)))
        return live.value, queued.value

    def drain(self):
        check(_lib.load().cntrx_server_drain(self._h))

    def poll(self, cap=4096):
        """[(id, seconds from submission to completion)] since the last poll"""
        ids = np.zeros(cap, np.int64)
        lat = np.zeros(cap, np.float64)
        n = ctypes.c_long(0)
        check(_lib.load().cntrx_server_poll(self._h, cap, ids.ctypes.data, lat.ctypes.data, ctypes.byref(n)))
        return list(zip(ids[:n.value].tolist(), lat[:n.value].tolist()))

    def result(self, rid):
        d = CntrxHerm()
        check(_lib.load().cntrx_server_result(self._h, rid, ctypes.byref(d)))
        return Result(d)

    def release(self, rid):
        check(_lib.load().cntrx_server_release(self._h, rid))

    def stats(self):
        out = (ctypes.c_double * 8)()
        check(_lib.load().cntrx_server_stats(self._h, out))
        keys = ("ticks", "timesteps", "admitted", "completed", "pool_used", "pool_peak", "pool_cap", "live")
        return dict(zip(keys, list(out)))

    @property
    def ticks(self):
        return int(self.stats()["ticks"])

    @property
    def timesteps(self):
        return int(self.stats()["timesteps"])


def autotune(enable=True):
    """turn the autotuner of the batched calls on or off (cntrx_autotune)"""
    check(_lib.load().cntrx_autotune(1 if enable else 0))


def autotune_decisions():
    """decisions made so far: size1, history bucket, what was tuned, ms per candidate, choice"""
    rows = np.zeros((64, 8))
    n = ctypes.c_int(0)
    check(_lib.load().cntrx_autotune_decisions(64, rows.ctypes.data, ctypes.byref(n)))
    what = {0: "recurrence panel size", 1: "GEMM split-K workgroup target"}
    out = []
    for r in rows[:min(n.value, 64)]:
        nc = int(r[3])
        out.append({"size1": int(r[0]), "history_bucket": int(r[1]), "tuned": what.get(int(r[2]), "?"),
                    "ms": [float(x) for x in r[5:5 + nc]], "choice": int(r[4])})
    return out
