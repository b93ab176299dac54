"""Small helpers the host side shares: a key-checked dict and a name registry.

The registry pattern mirrors the reference's plugin lookup by ``name``
(``peregrine/misc/subclass.py:1-11``); the key-checked dict mirrors its config
sections, where a misspelled key is refused on set
(``peregrine/misc/frozendict.py``).
"""


class LockedDict(dict):
    """A dict whose key set can be closed: setting an unknown key then raises."""

    _locked = False

    def __setitem__(self, key, value):
        if self._locked and key not in self:
            known = ", ".join(sorted(map(str, self.keys())))
            raise KeyError(f"{key!r} is not a setting here (known: {known}); check spelling and case")
        super().__setitem__(key, value)

    def update(self, *args, **kwargs):
        for key, value in dict(*args, **kwargs).items():
            self[key] = value

    def setdefault(self, key, default=None):
        if key not in self:
            self[key] = default
        return self[key]

    def lock(self):
        self._locked = True
        return self


def all_subclasses(cls):
    found = []
    for sub in cls.__subclasses__():
        found.append(sub)
        found.extend(all_subclasses(sub))
    return found


def find_subclass(cls, **attrs):
    """The first subclass of ``cls`` whose class attributes equal ``attrs``."""
    for sub in all_subclasses(cls):
        if all(getattr(sub, k, None) == v for k, v in attrs.items()):
            return sub
    wanted = ", ".join(f"{k}={v!r}" for k, v in attrs.items())
    raise KeyError(f"no {cls.__name__} with {wanted}")


def merge_disjoint(*dicts):
    out = {}
    for d in dicts:
        clash = set(out) & set(d)
        if clash:
            raise ValueError(f"keys defined twice: {sorted(clash)}")
        out.update(d)
    return out
