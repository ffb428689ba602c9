"""Load the upstream heatsim2 package for fixture generation (this container only).

The package is assembled from two places: the upstream ``.py`` modules where
they lie under /root/reference/heatsim2, and the Cython/C extension modules
that ``make -C oracle ref`` compiles from the upstream sources into
oracle/_ref/heatsim2/.  Nothing is copied.  Only the scripts in tests/golden/
use this, and only to write the ``.npz`` fixtures committed beside them.
"""
import contextlib
import importlib.util
import io
import os
import sys

REF_PKG = "/root/reference/heatsim2"
HERE = os.path.dirname(os.path.abspath(__file__))
REF_BUILD = os.path.normpath(os.path.join(HERE, "..", "..", "oracle", "_ref", "heatsim2"))


def load_reference():
    """Return the upstream ``heatsim2`` module (registered in sys.modules)."""
    if "heatsim2" in sys.modules:
        return sys.modules["heatsim2"]
    if not os.path.isdir(REF_PKG):
        raise RuntimeError("upstream sources not mounted at %s" % REF_PKG)
    if not os.path.isdir(REF_BUILD):
        raise RuntimeError("run `make -C oracle ref` first (%s missing)" % REF_BUILD)
    sys.dont_write_bytecode = True  # never write into the read-only mount
    spec = importlib.util.spec_from_file_location(
        "heatsim2", os.path.join(REF_PKG, "__init__.py"),
        submodule_search_locations=[REF_BUILD, REF_PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["heatsim2"] = mod
    spec.loader.exec_module(mod)
    return mod


@contextlib.contextmanager
def quiet():
    """Swallow the upstream setup()'s per-row progress prints."""
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        yield
