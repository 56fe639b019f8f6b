"""Loader of the libtorch module layer (_neat_host.so, built in-tree from neat_amd/csrc/host).

libneat_amd.so is loaded first (globally) so the extension resolves its C-ABI symbols from the in-tree library."""
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
        _MOD = importlib.import_module("neat_amd._neat_host")
    return _MOD
