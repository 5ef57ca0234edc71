"""Python side of the MI355X BEAGLE plugin: a ctypes binding of the C ABI
(api), synthetic problem generators (synth), a driver that runs a full
likelihood/gradient evaluation through any library exporting the API
(driver), and a torch.autograd binding (autograd)."""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.dirname(PKG_DIR)
REPO_DIR = os.path.dirname(LIB_DIR)
LIB_PATH = os.path.join(LIB_DIR, "libhmsbeagle_amd.so")


def load(path=None):
    """Load the MI355X library.  Raises if it has not been built: there is
    no fallback implementation."""
    from . import api
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing; build it with __graft_entry__.build()")
    return api.Library(p, prefix="beagle")
