"""Per-type and per-type-pair parameter dictionaries (hoomd/data/typeparam.py semantics)."""

import itertools


def _expand(key):
    """Keys may be a name or a list of names; pairs are tuples of those."""
    if isinstance(key, str):
        return [key]
    return list(key)


class TypeParameter:
    """params['A'] = {...}; params[['A', 'B']] = {...}"""

    def __init__(self, default=None):
        self._d = {}
        self._default = default

    def __setitem__(self, key, value):
        for k in _expand(key):
            self._d[k] = value

    def __getitem__(self, key):
        if key in self._d:
            return self._d[key]
        if self._default is not None:
            return self._default
        raise KeyError(key)

    def __contains__(self, key):
        return key in self._d or self._default is not None

    def items(self):
        return self._d.items()


class TypePairParameter:
    """params[('A', 'B')] = {...}; unordered pairs; lists expand to every combination."""

    def __init__(self, default=None):
        self._d = {}
        self._default = default

    @staticmethod
    def _pairs(key):
        if isinstance(key, str):
            raise KeyError("type pair keys are tuples of two type names")
        a, b = key
        return [tuple(sorted(p)) for p in itertools.product(_expand(a), _expand(b))]

    def __setitem__(self, key, value):
        for p in self._pairs(key):
            self._d[p] = value

    def __getitem__(self, key):
        (p,) = self._pairs(key)
        if p in self._d:
            return self._d[p]
        if self._default is not None:
            return self._default
        raise KeyError(key)

    def get(self, a, b, default=None):
        p = tuple(sorted((a, b)))
        if p in self._d:
            return self._d[p]
        return self._default if self._default is not None else default

    def signature(self):
        return tuple(sorted((k, tuple(sorted(v.items())) if isinstance(v, dict) else v)
                            for k, v in self._d.items()))
