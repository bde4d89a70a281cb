"""Build step of the product: the HIP libraries (csrc/Makefile ->
lib/libgks1d.so, libgks2d.so, libgks3d.so) and the reference-named C++ call
surface with its example drivers (host/Makefile)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(path, *targets, jobs=4):
    cmd = ["make", "-C", path, f"-j{jobs}", *targets]
    subprocess.run(cmd, check=True)


def build_hip(force=True):
    csrc = os.path.join(PKG_DIR, "csrc")
    if force:
        _make(csrc, "clean")
    _make(csrc)
    # reference-named C++ call surface + example drivers, over the C ABI
    _make(os.path.join(PKG_DIR, "host"), "-B")
