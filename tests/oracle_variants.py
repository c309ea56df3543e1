"""Variant builds of the CPU oracle with one planted mistake each, for checking
that the oracle's pins fail when the oracle is wrong.

A variant is (file, exact snippet, replacement): the oracle's C sources are
copied to a temporary folder, the snippet (which must occur exactly once) is
replaced, and the copy is compiled with the oracle's own flags. The product
library is never involved."""
import os
import shutil
import subprocess
import tempfile

from oracle import rco

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SOURCES = ["rc_oracle.c", "rc_oracle_post.c", "rc_oracle_sppm.c"]
CFLAGS = ["-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "retrocam_amd", "csrc"),
          "-Wno-unused-function"]

_cache = {}


def build_variant(fname: str, old: str, new: str):
    """Returns the loaded oracle library with `old` replaced by `new` in `fname`."""
    key = (fname, old, new)
    if key in _cache:
        return _cache[key]
    d = tempfile.mkdtemp(prefix="rco_variant_")
    for f in SOURCES + ["rc_oracle.h"]:
        shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    src = open(path).read()
    if src.count(old) != 1:
        raise AssertionError(f"planted mistake: snippet occurs {src.count(old)} times in {fname}: {old!r}")
    open(path, "w").write(src.replace(old, new))
    out = os.path.join(d, "librc_oracle_variant.so")
    r = subprocess.run(["gcc", *CFLAGS, "-o", out, *[os.path.join(d, f) for f in SOURCES], "-lm"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise AssertionError("variant build failed:\n" + r.stdout)
    lib = rco.load(out)
    _cache[key] = lib
    return lib


def pin_fails(pin, fname: str, old: str, new: str) -> bool:
    """True when `pin` (a function that raises AssertionError on a mismatch)
    fails under the variant oracle."""
    lib = build_variant(fname, old, new)
    with rco.using(lib):
        try:
            pin()
        except AssertionError:
            return True
    return False
