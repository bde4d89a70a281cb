"""Build steps: the HIP library (csrc/Makefile -> lib/libgks2d.so) and the
oracle (oracle/Makefile; the oracle is test infrastructure, built here so the
tests, smoke() and the bench's CPU baseline find it)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG_DIR)


def _make(path, *targets, jobs=4):
    cmd = ["make", "-C", path, f"-j{jobs}", *targets]
    subprocess.run(cmd, check=True)


def build_hip(force=True):
    csrc = os.path.join(PKG_DIR, "csrc")
    if force:
        _make(csrc, "clean")
    _make(csrc)
    # reference-named C++ call surface + example driver, over the C ABI
    _make(os.path.join(PKG_DIR, "host"), "-B")


def build_oracle(with_reference=True):
    odir = os.path.join(ROOT, "oracle")
    _make(odir, "-B", "all")
    ref = os.environ.get("GKS_REFERENCE_DIR", "/root/reference")
    if with_reference and os.path.isdir(os.path.join(ref, "src")):
        _make(odir, "-B", "ref", f"REF={ref}")


def build_all():
    build_hip(force=True)
    build_oracle()
