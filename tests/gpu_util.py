"""Helpers shared by the GPU parity tests: host<->device copies and the parity metric."""
import numpy as np

COMPONENTS = ("ret", "les", "tv", "mat")


def to_gpu(cntr, X):
    """oracle HostHerm -> nessi_amd herm_matrix (same packed layout)."""
    G = cntr.herm_matrix(X.nt, X.ntau, X.size1, X.sig)
    G.from_numpy({c: getattr(X, c) for c in COMPONENTS})
    return G


def to_host(orc, G):
    X = orc.HostHerm(G.nt, G.ntau, G.size1, G.sig)
    a = G.to_numpy()
    for c in COMPONENTS:
        getattr(X, c)[...] = a[c]
    return X


def func_to_gpu(cntr, F):
    f = cntr.function(F.nt, F.size1)
    import torch
    f.data.copy_(torch.from_numpy(F.data))
    return f


def rel_err(a, b):
    """max|a-b| / max|b| (the per-component parity metric)."""
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    return num / den if den > 0 else num


def component_errors(orc, G, X, tsteps=None):
    """relative error per component of GPU G against oracle X; restricted to the
    entries of timesteps `tsteps` when given."""
    Gh = to_host(orc, G)
    out = {}
    if tsteps is None:
        for c in COMPONENTS:
            out[c] = rel_err(getattr(Gh, c), getattr(X, c))
        return out
    for name, get in (("ret", lambda H, t: H.ret_row(t)), ("les", lambda H, t: H.les_col(t)),
                      ("tv", lambda H, t: H.tv_row(t))):
        a = np.concatenate([get(Gh, t).ravel() for t in tsteps])
        b = np.concatenate([get(X, t).ravel() for t in tsteps])
        out[name] = rel_err(a, b)
    return out
