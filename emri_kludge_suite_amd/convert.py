"""Sampler-to-physical parameter conversion for the likelihood caller
(behaviour of pygpuAAK/convert.py:21-120).

Keys name the 14 AAK parameters, optionally with a ``ln_``, ``cos_`` or
``sin_`` prefix meaning the sampler works in log, cosine or sine of it; the
angles gamma, psi, alph (alpha), phi_S and phi_K are wrapped into [0, 2 pi).
Two behaviours of the reference are not kept: its recycle step matches only
the last underscore-separated part of each key (so phi_S and phi_K are never
wrapped), and a distance key is multiplied by the metres in a Gpc although
the generator takes D in Gpc.  Here the whole parameter name is matched and
D stays in Gpc.
"""
import numpy as np

PARAMS = ("iota", "s", "p", "e", "M", "mu", "gamma", "psi", "alph", "theta_S", "phi_S", "theta_K", "phi_K", "D")
WRAP_2PI = ("gamma", "psi", "alph", "alpha", "phi_S", "phi_K")
_ALIASES = {"alpha": "alph", "a": "s", "distance": "D"}


def _base_name(key):
    for pre in ("ln_", "cos_", "sin_"):
        if key.startswith(pre):
            return pre[:-1], key[len(pre):]
    return None, key


class Converter:
    def __init__(self, source_type, key_order, **kwargs):
        if source_type != "emri":
            raise ValueError("only the 'emri' source type is supported")
        self.key_order = list(key_order)
        self.conversions = []
        self.recycles = []
        self.index = {}
        for i, key in enumerate(self.key_order):
            pre, name = _base_name(key)
            name = _ALIASES.get(name, name)
            if name not in PARAMS:
                raise ValueError("unknown parameter key %r" % key)
            self.index[name] = i
            if pre is not None:
                self.conversions.append((pre, i))
            elif name in WRAP_2PI:
                self.recycles.append(i)

    def convert(self, x):
        """Sampler coordinates to physical values (in place, returns x)."""
        for pre, i in self.conversions:
            x[i] = {"ln": np.exp, "cos": np.arccos, "sin": np.arcsin}[pre](x[i])
        return x

    def recycle(self, x):
        """Wrap the periodic angles into [0, 2 pi) (in place, returns x)."""
        for i in self.recycles:
            x[i] = x[i] % (2 * np.pi)
        return x

    def physical_order(self, x):
        """Rows of x (in key order) rearranged into the generator's order."""
        missing = [p for p in PARAMS if p not in self.index]
        if missing:
            raise ValueError("key_order lacks %s" % missing)
        return np.stack([np.asarray(x[self.index[p]], dtype=np.float64) for p in PARAMS])
