"""On-disk cache of compiled code objects.

The reference caches compiled kernels keyed by a configuration hash
(cubie_cache.py:451-880). Here the key is the SHA-256 of the generated
translation unit plus the hiprtc options, and the value is the gfx950 code
object. The default directory is ``cubie_amd/_kernel_cache`` inside the
package (override with ``CUBIE_KERNEL_CACHE_DIR``), so code objects built on
a CPU-only machine ship with the tree and need no recompilation on the GPU
machine.
"""

from __future__ import annotations

import hashlib
import os
import tempfile
from typing import Dict, List, Tuple

from . import native

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_DIR = os.path.join(_PKG, "_kernel_cache")

_memory: Dict[str, bytes] = {}
_modules: Dict[Tuple[str, str], "native.Module"] = {}


def cache_dir() -> str:
    return os.environ.get("CUBIE_KERNEL_CACHE_DIR", DEFAULT_DIR)


def key_for(source: str, options: List[str]) -> str:
    h = hashlib.sha256()
    h.update(source.encode())
    h.update(b"\0")
    h.update("\n".join(options).encode())
    return h.hexdigest()


def get_code(source: str, options: List[str]) -> Tuple[str, bytes]:
    key = key_for(source, options)
    if key in _memory:
        return key, _memory[key]
    path = os.path.join(cache_dir(), key[:32] + ".co")
    if os.path.exists(path):
        with open(path, "rb") as f:
            code = f.read()
        _memory[key] = code
        return key, code
    code = native.compile_source(source, options)
    try:
        os.makedirs(cache_dir(), exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=cache_dir(), suffix=".tmp")
        with os.fdopen(fd, "wb") as f:
            f.write(code)
        os.replace(tmp, path)
    except OSError:
        pass
    _memory[key] = code
    return key, code


def get_module(source: str, options: List[str], kernel: str) -> "native.Module":
    key, code = get_code(source, options)
    mk = (key, kernel)
    if mk not in _modules:
        _modules[mk] = native.Module(code, kernel)
    return _modules[mk]
