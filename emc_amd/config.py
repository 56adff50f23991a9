"""emc.conf reading through the C ABI (``emc_config_read``, configs.cpp:40-90 semantics)."""
from __future__ import annotations

import ctypes as C

from ._lib import Config, check, lib


def _to_dict(c: Config) -> dict:
    out = {}
    for name, _ in Config._fields_:
        v = getattr(c, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


def defaults() -> dict:
    c = Config()
    check(lib().emc_config_defaults(C.byref(c)))
    return _to_dict(c)


def parse(text: str) -> dict:
    c = Config()
    check(lib().emc_config_parse(text.encode(), C.byref(c)))
    return _to_dict(c)


def read(path: str) -> dict:
    c = Config()
    check(lib().emc_config_read(path.encode(), C.byref(c)))
    return _to_dict(c)
