"""Build recipes for the CPU oracle (test infrastructure only).

* ``build_oracle()`` compiles this directory's C restatement into ``oracle/liborc.so``.
* ``build_ref()`` compiles the reference's own host predicates
  (``GDelaunay/Common/predicates.c``, Shewchuk's adaptive orient3d/insphere) straight from
  the reference tree into ``oracle/_ref/libref_predicates.so``.  It needs the reference
  checkout, which exists only in the development container; the GPU box uses the file
  built here.  No reference source is copied into this repository.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = os.environ.get("GSTAR4D_REFERENCE", "/root/reference")
REF_PREDICATES = os.path.join(REF_ROOT, "GDelaunay", "Common", "predicates.c")

CFLAGS = ["-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math",
          "-fopenmp"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_oracle():
    out = os.path.join(HERE, "liborc.so")
    srcs = [os.path.join(HERE, f) for f in ("orc_predicates.c", "orc_pipeline.c")]
    _run(["gcc"] + CFLAGS + srcs + ["-o", out, "-lm"])
    return out


def build_ref():
    """Returns the path of the built library, or None when the reference is absent."""
    if not os.path.exists(REF_PREDICATES):
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libref_predicates.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-DREAL_TYPE_DOUBLE",
          "-w", REF_PREDICATES, "-o", out, "-lm"])
    return out


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref())
