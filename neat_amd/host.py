"""Loader of the libtorch module layer (_neat_host.so, built in-tree from neat_amd/csrc/host).

libneat_amd.so is loaded first (globally) so the extension resolves its C-ABI symbols from the in-tree library. A module
built from other sources than those beside it is refused (neat_amd/buildinfo.py)."""
import importlib
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_MOD = None


def load():
    global _MOD
    if _MOD is None:
        from . import _abi
        _abi.lib()
        if not os.path.exists(os.path.join(_HERE, "_neat_host.so")):
            raise ImportError("neat_amd/_neat_host.so is not built; run `python -m neat_amd.build`")
        mod = importlib.import_module("neat_amd._neat_host")
        from .buildinfo import source_id
        built, want = mod.build_id(), source_id("host")
        if built != want:
            raise ImportError(f"neat_amd/_neat_host.so was built from other sources (build id {built}, sources "
                              f"{want}); rebuild with `python -m neat_amd.build`")
        _MOD = mod
    return _MOD
