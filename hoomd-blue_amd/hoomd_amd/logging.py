"""Logging (hoomd/logging.py): named quantities of operations, collected for writers.

A loggable class lists its quantities in ``_loggables`` (name -> (category, default)) and its
namespace in ``_log_namespace``, e.g. ('md', 'compute', 'ThermodynamicQuantities'), the tuples
the reference derives from its module path (logging.py:286-296). ``Logger.add`` registers an
object's quantities under namespace + (name,), ``Logger.log`` evaluates them into the
reference's nested-dictionary form {namespace...: {name: (value, category)}}.
"""

import enum
from functools import reduce


class LoggerCategories(enum.Flag):
    """Kinds of logged values (logging.py:34-122)."""

    NONE = 0
    scalar = enum.auto()
    sequence = enum.auto()
    string = enum.auto()
    strings = enum.auto()
    object = enum.auto()
    angle = enum.auto()
    bond = enum.auto()
    constraint = enum.auto()
    dihedral = enum.auto()
    improper = enum.auto()
    pair = enum.auto()
    particle = enum.auto()

    @classmethod
    def any(cls, categories=None):
        if categories is None:
            return cls.ALL
        return reduce(lambda a, b: a | b,
                      (c if isinstance(c, cls) else cls[c] for c in categories), cls.NONE)


LoggerCategories.ALL = reduce(lambda a, b: a | b, list(LoggerCategories), LoggerCategories.NONE)


class Logger:
    """hoomd.logging.Logger(categories=None, only_default=True)."""

    def __init__(self, categories=None, only_default=True):
        if isinstance(categories, LoggerCategories):
            self._categories = categories
        else:
            self._categories = LoggerCategories.any(categories)
        self._only_default = bool(only_default)
        self._entries = {}  # namespace tuple -> (callable, category name)
        self._owner = {}    # namespace tuple -> id of the object it was added from

    @property
    def categories(self):
        return self._categories

    @property
    def string_categories(self):
        return [c.name for c in LoggerCategories if c in self._categories and c.value]

    @property
    def only_default(self):
        return self._only_default

    def _unique(self, ns):
        """The reference appends _1, _2, ... to the class level of a namespace in use
        (logging.py:608-613)."""
        if not any(k[:len(ns)] == ns for k in self._entries):
            return ns
        k = 1
        while True:
            cand = ns[:-1] + (f"{ns[-1]}_{k}",)
            if not any(e[:len(cand)] == cand for e in self._entries):
                return cand
            k += 1

    def add(self, obj, quantities=None, user_name=None):
        """Register the object's quantities; returns the namespaces added."""
        table = getattr(type(obj), "_loggables", None)
        if table is None:
            raise ValueError(f"{type(obj).__name__} has no loggable quantities")
        if quantities is None:
            names = [n for n, (cat, default) in table.items()
                     if LoggerCategories[cat] in self._categories
                     and (default or not self._only_default)]
        else:
            names = [quantities] if isinstance(quantities, str) else list(quantities)
            bad = [n for n in names if n not in table]
            if bad:
                raise ValueError(f"object {obj} has no loggable quantities {bad}.")
            names = [n for n in names if LoggerCategories[table[n][0]] in self._categories]
        base = tuple(type(obj)._log_namespace)
        if user_name is not None:
            base = base[:-1] + (user_name,)
        base = self._unique(base)
        added = []
        for n in names:
            ns = base + (n,)
            self._entries[ns] = (lambda o=obj, a=n: getattr(o, a), table[n][0])
            self._owner[ns] = id(obj)
            added.append(ns)
        return added

    def __iadd__(self, obj):
        self.add(obj)
        return self

    def remove(self, obj=None, quantities=None, user_name=None):
        """Remove the quantities added from obj (all, or those named in quantities), or the
        given namespaces when obj is None."""
        if obj is None:
            for q in quantities or ():
                self._entries.pop(tuple(q), None)
                self._owner.pop(tuple(q), None)
            return
        for k in [k for k, o in self._owner.items() if o == id(obj)
                  and (quantities is None or k[-1] in quantities)
                  and (user_name is None or k[-2] == user_name)]:
            del self._entries[k]
            del self._owner[k]

    def __setitem__(self, namespace, value):
        """Custom quantity: value is (callable, category) or (object, attribute, category)."""
        ns = (namespace,) if isinstance(namespace, str) else tuple(namespace)
        if len(value) == 2:
            fn, cat = value
        else:
            o, a, cat = value
            fn = (lambda o=o, a=a: (getattr(o, a)() if callable(getattr(o, a))
                                    else getattr(o, a)))
        if LoggerCategories[cat] not in self._categories:
            raise ValueError(f"category {cat} is not accepted by this logger")
        self._entries[ns] = (fn, cat)

    def __getitem__(self, namespace):
        return self._entries[tuple(namespace) if not isinstance(namespace, str)
                             else (namespace,)]

    def __contains__(self, namespace):
        return (tuple(namespace) if not isinstance(namespace, str)
                else (namespace,)) in self._entries

    def __iter__(self):
        return iter(self._entries)

    def __len__(self):
        return len(self._entries)

    def log(self):
        """Nested dictionary {ns0: {ns1: ... {name: (value, category)}}}."""
        out = {}
        for ns, (fn, cat) in self._entries.items():
            d = out
            for part in ns[:-1]:
                d = d.setdefault(part, {})
            d[ns[-1]] = (fn(), cat)
        return out


def flatten(log_dict, prefix=()):
    """{namespace tuple: (value, category)} from the nested form."""
    out = {}
    for k, v in log_dict.items():
        if isinstance(v, dict):
            out.update(flatten(v, prefix + (k,)))
        else:
            out[prefix + (k,)] = v
    return out
