"""Planted-mistake builds of the oracle, for tests that must be able to fail.

`build(tmp_dir, name, [(old, new), ...])` copies oracle/lrt_oracle.c into
`tmp_dir`, replaces each `old` text (which must occur exactly once, so a
rewrite of the oracle that removes the target fails loudly instead of planting
nothing) and compiles it with the oracle's own flags. The result loads with
`oracle_lib.load`, so a check written against the oracle can be run against
the mistake and shown to reject it.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import oracle_lib as O

SRC = os.path.join(O.ROOT, "oracle", "lrt_oracle.c")
HDR = os.path.join(O.ROOT, "oracle", "lrt_oracle.h")
CFLAGS = ["-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-w"]


def build(tmp_dir, name: str, replacements):
    text = open(SRC).read()
    for old, new in replacements:
        found = text.count(old)
        assert found == 1, f"planted mistake {name!r}: {old!r} occurs {found} times in the oracle"
        text = text.replace(old, new)
    d = os.path.join(str(tmp_dir), name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(HDR, d)
    with open(os.path.join(d, "lrt_oracle.c"), "w") as fh:
        fh.write(text)
    so = os.path.join(d, f"liblrt_oracle_{name}.so")
    subprocess.run(["gcc", *CFLAGS, "-shared", "-o", so, os.path.join(d, "lrt_oracle.c"), "-lm"], check=True)
    return O.load(so)
