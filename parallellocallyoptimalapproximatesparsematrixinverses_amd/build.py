"""Build steps of the package: the HIP library (csrc/Makefile) and the Cython
mirror modules (methods_cython/*.pyx), both in-tree."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
PKG_NAME = PKG.name


def build_native(jobs: int = 8) -> Path:
    subprocess.run(["make", "-s", "-C", str(PKG / "csrc"), f"-j{jobs}"], check=True)
    return PKG / "libspai_hip.so"


def build_cython() -> None:
    import numpy
    from Cython.Build import cythonize
    from setuptools import Extension
    from setuptools.dist import Distribution

    tmp = PKG / "csrc" / "_build" / "cython"
    tmp.mkdir(parents=True, exist_ok=True)
    exts = []
    for name in ("methods", "backward_error"):
        exts.append(Extension(
            f"{PKG_NAME}.methods_cython.{name}",
            [str(PKG / "methods_cython" / f"{name}.pyx")],
            include_dirs=[str(INCLUDE), numpy.get_include()],
            library_dirs=[str(PKG)],
            libraries=["spai_hip"],
            runtime_library_dirs=["$ORIGIN/.."],
            language="c++",
            extra_compile_args=["-O2", "-std=c++17"],
        ))
    modules = cythonize(exts, build_dir=str(tmp), quiet=True, language_level=3,
                        compiler_directives={"embedsignature": True})
    dist = Distribution({"ext_modules": modules, "script_args": ["build_ext"]})
    cmd = dist.get_command_obj("build_ext")
    cmd.inplace = True
    cmd.build_temp = str(tmp / "obj")
    cwd = os.getcwd()
    os.chdir(ROOT)
    try:
        dist.run_command("build_ext")
    finally:
        os.chdir(cwd)


def build_all() -> None:
    build_native()
    build_cython()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
