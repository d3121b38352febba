"""MI355X-native (gfx950) visibility -> dirty-image path of the BLINK imager.

The product is the C ABI library ``libblink_mi355x.so`` (include/blink_mi355x.h)
and the C++ drop-in imager class over it (``host/pacer_imager_mi355x.h``).
This Python package holds the build recipe, the ctypes binding used by tests
and the benchmark, and the seeded synthetic-observation generator.
"""
__all__ = ["capi", "synth", "build"]


def __getattr__(name):
    if name in __all__:
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
