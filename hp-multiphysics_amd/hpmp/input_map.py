"""input_map: ``key: value`` configuration with block-prefixed overrides.

Mirrors the reference's input_map (input_map/input_map.h:18-113): a file of
``key: value`` lines, ``#`` comments, and lookups that try the block-prefixed
key (``b0_mu``) before the global one (``mu``), as tri_hp/ins/allocate.cpp:20-21
and tri_hp/allocate.cpp:275-276 do.  Values that are lists (``alpha: .25 .1667``)
are read as whitespace-separated numbers.
"""
from __future__ import annotations


class InputMap(dict):
    @classmethod
    def from_text(cls, text: str) -> "InputMap":
        m = cls()
        for raw in text.splitlines():
            line = raw.split("#", 1)[0].strip()
            if not line or ":" not in line:
                continue
            k, v = line.split(":", 1)
            m[k.strip()] = v.strip()
        return m

    @classmethod
    def from_file(cls, path: str) -> "InputMap":
        with open(path) as f:
            return cls.from_text(f.read())

    def lookup(self, key: str, prefix: str = "b0"):
        if prefix and f"{prefix}_{key}" in self:
            return self[f"{prefix}_{key}"]
        return self.get(key)

    def getf(self, key: str, default: float, prefix: str = "b0") -> float:
        v = self.lookup(key, prefix)
        return float(v) if v is not None else float(default)

    def geti(self, key: str, default: int, prefix: str = "b0") -> int:
        v = self.lookup(key, prefix)
        return int(float(v)) if v is not None else int(default)

    def getlist(self, key: str, default, n: int | None = None, prefix: str = "b0") -> list:
        v = self.lookup(key, prefix)
        vals = [float(x) for x in v.split()] if v is not None else list(default)
        return vals[:n] if n is not None else vals
