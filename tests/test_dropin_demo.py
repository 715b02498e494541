"""Drop-in check of the boundary: the reference's own caller, its command-line
program (demo/main.cpp with main.h, options.h and lodepng), compiles
unchanged against this build's am::morph headers and links against
libatomorph.so.  The program is built, never run.

The reference sources are not copied: the test links them by symlink into a
temporary tree whose ../atomorph.h (the path demo/main.h includes) is this
build's header.  Skipped where the reference tree is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
HOST = os.path.join(ROOT, "atomorph_amd", "host")
LIB = os.path.join(ROOT, "atomorph_amd", "lib")
DEMO = ["main.cpp", "main.h", "options.h", "lodepng.h", "lodepng.cpp"]


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "demo", "main.cpp")), reason="reference tree absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_reference_demo_compiles_and_links_unchanged(tmp_path):
    if not os.path.isfile(os.path.join(LIB, "libatomorph.so")):
        pytest.skip("libatomorph.so not built")
    demo = tmp_path / "demo"
    demo.mkdir()
    for f in DEMO:
        os.symlink(os.path.join(REF, "demo", f), demo / f)
    os.symlink(os.path.join(HOST, "atomorph.h"), tmp_path / "atomorph.h")

    def run(cmd):
        r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout}"

    # no forced includes, no defines: the program as its authors wrote it
    run(["g++", "-std=c++17", "-O0", "-I", HOST, "-c", "demo/main.cpp", "-o", "main.o"])
    run(["g++", "-std=c++17", "-O0", "-c", "demo/lodepng.cpp", "-o", "lodepng.o"])
    run(["g++", "main.o", "lodepng.o", "-L", LIB, "-latomorph", "-latomorph_hip", f"-Wl,-rpath,{LIB}", "-o",
         "atomorph_demo"])
    assert (tmp_path / "atomorph_demo").is_file()
