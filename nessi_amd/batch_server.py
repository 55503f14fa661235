"""Continuous batching of independent Dyson solves over the per-solve entry points.

Each request is the reference's dyson(...) sequence (libcntr/cntr/cntr_dyson_impl.hpp:
1725-1734) on its own G: dyson_mat, dyson_start, then dyson_timestep for n = SolveOrder+1..nt,
with H(t) and a bath self-energy Sigma = V^2 g[H_bath] (green_from_H,
libcntr/cntr/cntr_equilibrium_impl.hpp:873-953).  The server keeps every admitted solve
resident on the device:

  admit  queued requests, first come first served, while their footprint (G, Sigma, H)
         fits the budget: build Sigma, dyson_mat, dyson_start;
  tick   every live solve advances one timestep: one cntrx_dyson_timestep_batched call per
         size1 class (the batched entry point takes one size1 per call; n, ntau, beta, h,
         mu, H and Sigma are per solve), and one cntrx_dyson_timestep call per size-1 solve;
  retire solves that reached nt leave the live set; their G stays until released.

No reference counterpart: it replaces a caller's loop over solves.  Each solve's kernels
read only its own G, Sigma and H, so its result does not depend on what else is in the
batch, bit for bit.  It uses no stream capture and none of cntrx_engine_*.
"""
import time

import torch

from . import cntr


def footprint(r):
    """device bytes of one request: G and Sigma in the packed layout, and H(t)"""
    s2 = r["size1"] * r["size1"]
    herm = cntr.herm_matrix.storage_elems(r["nt"], r["ntau"], r["size1"]) * s2 * 16
    return 2 * herm + (r["nt"] + 2) * s2 * 16


class BatchServer:
    def __init__(self, budget_bytes, SolveOrder=5):
        self.budget, self.used, self.k = int(budget_bytes), 0, SolveOrder
        self.next_id = 1
        self.queue = []          # ids in arrival order
        self.reqs = {}           # id -> state dict
        self.live = []           # ids admitted and not finished
        self.completed = []      # (id, seconds from submission to completion)
        self._classes = None     # per size1: (ids, Gs, Hs, Sigmas) while the live set holds
        self.ticks = self.timesteps = 0

    def submit(self, r):
        if footprint(r) > self.budget:
            raise ValueError("request footprint exceeds the budget")
        rid = self.next_id
        self.next_id += 1
        self.reqs[rid] = {"r": r, "t_submit": time.perf_counter(), "done": False}
        self.queue.append(rid)
        return rid

    def _admit(self, rid):
        st = self.reqs[rid]
        r = st["r"]
        S, nt, ntau, k = r["size1"], r["nt"], r["ntau"], self.k
        G = cntr.herm_matrix(nt, ntau, S, r["sig"])
        Sig = cntr.herm_matrix(nt, ntau, S, r["sig"])
        H = cntr.function(nt, S)
        H.data.copy_(torch.from_numpy(r["H"]))
        cntr.green_from_H(Sig, r["mu"], r["Hbath"], r["beta"], r["h"])
        Sig._storage.mul_(r["V"] * r["V"])
        cntr.dyson_mat(G, r["mu"], H, Sig, r["beta"], k, r["mat_method"])
        cntr.dyson_start(G, r["mu"], H, Sig, r["beta"], r["h"], k)
        st.update(G=G, Sig=Sig, H=H, n=k, bytes=footprint(r))
        self.used += st["bytes"]
        self.live.append(rid)
        self._classes = None

    def step(self):
        """admit what fits, then advance every live solve one timestep; returns (live, queued)"""
        while self.queue:
            rid = self.queue[0]
            if self.used + footprint(self.reqs[rid]["r"]) > self.budget:
                break
            self.queue.pop(0)
            self._admit(rid)
        if self.live:
            if self._classes is None:
                cl = {}
                for rid in self.live:
                    st = self.reqs[rid]
                    c = cl.setdefault(st["r"]["size1"], ([], [], [], []))
                    c[0].append(rid)
                    c[1].append(st["G"])
                    c[2].append(st["H"])
                    c[3].append(st["Sig"])
                self._classes = cl
            for size1, (ids, Gs, Hs, Sigs) in self._classes.items():
                rs = [self.reqs[i]["r"] for i in ids]
                ns = [self.reqs[i]["n"] + 1 for i in ids]
                if size1 == 1:
                    # a size-1 solve alone sums its lesser convolution as a one-column GEMM
                    # (k_lesconv_mm), which solves with different Sigma cannot share: one
                    # call each keeps every size-1 solve on that path, bit for bit
                    for n, G, H, Sg, r in zip(ns, Gs, Hs, Sigs, rs):
                        cntr.dyson_timestep(n, G, r["mu"], H, Sg, r["beta"], r["h"], self.k)
                    continue
                cntr.dyson_timestep_batched(ns, Gs, [r["mu"] for r in rs], Hs, Sigs,
                                            [r["beta"] for r in rs], [r["h"] for r in rs], self.k)
            self.ticks += 1
            finished = []
            for rid in self.live:
                st = self.reqs[rid]
                st["n"] += 1
                self.timesteps += 1
                if st["n"] == st["r"]["nt"]:
                    finished.append(rid)
            if finished:
                torch.cuda.current_stream().synchronize()
                now = time.perf_counter()
                for rid in finished:
                    st = self.reqs[rid]
                    st["done"] = True
                    self.live.remove(rid)
                    self.completed.append((rid, now - st["t_submit"]))
                    # Sigma and H are no longer needed; G stays until released
                    r = st["r"]
                    del st["Sig"], st["H"]
                    g_bytes = cntr.herm_matrix.storage_elems(r["nt"], r["ntau"], r["size1"]) * \
                        r["size1"] * r["size1"] * 16
                    self.used -= st["bytes"] - g_bytes
                    st["bytes"] = g_bytes
                self._classes = None
        return len(self.live), len(self.queue)

    def drain(self):
        while self.live or self.queue:
            self.step()

    def poll(self):
        out, self.completed = self.completed, []
        return out

    def result(self, rid):
        st = self.reqs[rid]
        if not st["done"]:
            raise ValueError("request %d not done" % rid)
        return st["G"]

    def release(self, rid):
        st = self.reqs.pop(rid)
        if not st["done"]:
            raise ValueError("request %d not done" % rid)
        self.used -= st["bytes"]
