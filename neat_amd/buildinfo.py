"""Build identity: a hash of the sources (and of the build recipe) each native library is compiled from.

The hash is compiled into libneat_amd.so (neat_build_id()) and _neat_host.so (build_id()); the loaders recompute it
from the sources beside the libraries and refuse a library built from other sources, so a stale build fails the
import instead of running old code.
"""
import hashlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

_PARTS = {
    "abi": [os.path.join(ROOT, "include"), os.path.join(HERE, "csrc", "kernels")],
    "host": [os.path.join(ROOT, "include"), os.path.join(HERE, "csrc", "host")],
}
_SUFFIXES = (".h", ".hpp", ".hip", ".cpp")


def source_id(kind):
    """16 hex digits of SHA-256 over the relative path and bytes of every source of `kind` ("abi" or "host"),
    in sorted order, followed by the build recipe (neat_amd/build.py)."""
    h = hashlib.sha256()
    files = []
    for d in _PARTS[kind]:
        for name in os.listdir(d):
            if name.endswith(_SUFFIXES):
                files.append(os.path.join(d, name))
    files.append(os.path.join(HERE, "build.py"))
    for f in sorted(files):
        h.update(os.path.relpath(f, ROOT).encode())
        h.update(b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()[:16]
