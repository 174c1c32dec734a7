"""The cook-side glue (include/mdxglue.c) compiles against cook's own
headers: the shipped cook/exe/ewald/simopt.h, with and without POLAR
(the Drude fields of siteinfo_t exist only with it), in the compile-time
setting of the reference build (PRESSURETENSOR=7).  The code shown in
INTEGRATION.md is the file's.  Needs the reference sources (the build
container); skipped elsewhere."""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
GLUE = os.path.join(REPO, "include", "mdxglue.c")


def _compile(extra):
    inc = ["-I" + os.path.join(REPO, "include")] + [
        "-I" + os.path.join(REF, d) for d in ("cook/exe/ewald", "gen", "sim", "sim/lj", "cook")]
    cmd = ["gcc", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-variable",
           "-Wno-unused-function", "-DMDX", "-DSCR", "-DPRESSURETENSOR=7"] + extra + inc + [GLUE]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "cook")) or not shutil.which("gcc"),
                    reason="needs the reference sources and gcc")
@pytest.mark.parametrize("extra", [[], ["-DPOLAR=0"]], ids=["nonpolar", "drude"])
def test_glue_compiles_against_cooks_headers(extra):
    p = _compile(extra)
    assert p.returncode == 0, p.stderr


def test_integration_shows_the_glue_file():
    """the main, RDF and step code of INTEGRATION.md appear verbatim in
    include/mdxglue.c"""
    doc = open(os.path.join(REPO, "INTEGRATION.md")).read()
    src = open(GLUE).read()
    blocks = re.findall(r"```c\n(.*?)```", doc, re.S)
    shown = [b for b in blocks if "mdxinit" in b or "mdxrdfsave" in b or "mdxsteps" in b]
    assert len(shown) == 3
    for b in shown:
        assert b.strip() in src, b[:80]
