"""The drop-in boundary at link level: the reference's own generated harness
(main.cpp:137-209, which calls fherma_init / fherma_run / fherma_free,
fherma.h:46-48) compiles unchanged, against the reference's fherma.h and
against include/fherma.h, and links against libfherma_polymul.so +
libpolymul.so with every fherma symbol resolved.  Skipped where the reference
tree is absent; nothing here runs the reference's program."""
import os
import subprocess

import pytest

from conftest import PKG_DIR, ROOT

REF = "/root/reference"
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "main.cpp")),
                                reason="the reference tree is not present")


@pytest.mark.parametrize("header_dir", [REF, os.path.join(ROOT, "include")], ids=["reference_header", "our_header"])
def test_reference_main_links_against_library(tmp_path, header_dir):
    exe = tmp_path / "solution_ref"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", header_dir, os.path.join(REF, "main.cpp"),
                    "-L", PKG_DIR, "-lfherma_polymul", "-lpolymul", "-Wl,-rpath," + PKG_DIR,
                    "-Wl,--no-undefined", "-o", str(exe)], check=True, timeout=180)
    undefined = subprocess.run(["nm", "-C", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    defined = subprocess.run(["nm", "-D", "-C", "--defined-only", os.path.join(PKG_DIR, "libfherma_polymul.so")],
                             capture_output=True, text=True, check=True).stdout
    for sig in ["fherma_init(fherma::Point const&)", "fherma_run(void*, fherma::Inputs const&)",
                "fherma_free(void*)"]:
        assert sig in undefined, f"main.cpp does not call {sig}"
        assert sig in defined, f"libfherma_polymul.so does not define {sig}"
