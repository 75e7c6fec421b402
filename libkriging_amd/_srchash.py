"""Hash of the sources liblkgpu.so is built from.

build.py compiles it into the library (lkgpu_build_id()); _native.lib()
compares it with the sources next to the library and refuses a library built
from other sources."""
import hashlib
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")
INCLUDE = os.path.join(_HERE, "..", "include")
_EXT = (".hip", ".cpp", ".hpp", ".h")


def source_files():
    files = []
    for root in (CSRC, INCLUDE):
        if os.path.isdir(root):
            for fn in sorted(os.listdir(root)):
                if fn.endswith(_EXT):
                    files.append(os.path.join(root, fn))
    return files


def source_hash():
    """sha256 (first 24 hex digits) over the names and bytes of the sources,
    or None when they are not present next to the package."""
    files = source_files()
    if not files:
        return None
    h = hashlib.sha256()
    for path in files:
        h.update(os.path.basename(path).encode() + b"\0")
        with open(path, "rb") as f:
            h.update(f.read())
        h.update(b"\0")
    return h.hexdigest()[:24]
