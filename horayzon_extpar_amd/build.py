"""Compile libhorayzon.so for gfx950 with hipcc (no GPU needed)."""

from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_CSRC = os.path.join(_HERE, "csrc")


def library_path() -> str:
    return os.path.join(_HERE, "libhorayzon.so")


def build_library(force: bool = True, jobs: int = 8) -> str:
    """Run the csrc Makefile; ``force`` rebuilds every object."""
    cmd = ["make", f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, cwd=_CSRC, check=True)
    return library_path()


def build_relink(force: bool = True) -> str:
    """Compile cython/horizon_svf_relink.pyx (the reference's Cython binding
    pattern against include/horayzon.h) into an extension module next to
    libhorayzon.so, linked to it with an $ORIGIN run path.  Returns the
    module path."""
    import glob
    import sysconfig

    import numpy as np
    from Cython.Build import cythonize

    src = os.path.join(_HERE, "cython", "horizon_svf_relink.pyx")
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(_HERE, "horizon_svf_relink" + ext)
    if not force and os.path.exists(out) and os.path.getmtime(out) > os.path.getmtime(src):
        return out
    bdir = os.path.join(os.path.dirname(_HERE), "_build", "relink")
    os.makedirs(bdir, exist_ok=True)
    c_file = os.path.join(bdir, "horizon_svf_relink.c")
    cythonize([src], quiet=True, force=True, language_level=3)
    gen = os.path.splitext(src)[0] + ".c"
    os.replace(gen, c_file)
    inc = [sysconfig.get_paths()["include"], np.get_include(),
           os.path.join(os.path.dirname(_HERE), "include")]
    cmd = (["gcc", "-O2", "-fPIC", "-shared", "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION"]
           + [f"-I{d}" for d in inc]
           + [c_file, "-L" + _HERE, "-l:libhorayzon.so", "-Wl,-rpath,$ORIGIN", "-o", out])
    subprocess.run(cmd, check=True)
    for f in glob.glob(os.path.join(_HERE, "cython", "*.c")):
        os.remove(f)
    return out
