"""MI355X (gfx950) triples engine with the public names of the reference ``distr_cc``.

``RCCSDT`` and ``rccsdt_q`` are loaded lazily as in the reference
(``distr_cc/__init__.py:15-34``).
"""
__all__ = ["RCCSDT"]


def __getattr__(name):
    if name == "RCCSDT":
        from distr_cc.rccsdt import RCCSDT
        return RCCSDT
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
