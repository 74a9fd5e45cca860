"""CPU checks of the C ABI: the library loads, exports every entry point that
include/ipg.h declares, agrees with the Python binding on struct layouts,
and parses reference-format input files (no device calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ipg.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ipg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import ipglasma_amd as I
    out = subprocess.run(["nm", "-D", "--defined-only", I.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (ipg_\w+)", out))
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_struct_layouts_match_binding():
    import ipglasma_amd as I
    I.lib()  # raises on a layout mismatch


def test_read_input_setup_semantics():
    """Setup::StringFind pairwise scan to EndOfFile, IFind truncation,
    ULLIFind rounding, derived dtau (main.cpp:531-534)."""
    import ipglasma_amd as I
    p = I.Setup.read(os.path.join(ROOT, "tests", "golden", "input_AuAu_C1"))
    assert p.size == 256          # the key after EndOfFile is ignored
    assert p.seed == 14           # ULLIFind: (ull)(13.6 + 0.5)
    assert p.Projectile == b"Au" and p.Target == b"Au"
    assert p.R_WS == 6.38 and p.setWSDeformParams == 1
    a = 30.0 / 256
    steps = int(10 * 0.4 / a) + 1
    assert p.itmax == 35 and p.dtau == 0.4 / (steps * a)


def test_read_input_missing_file_is_an_error():
    import ipglasma_amd as I
    with pytest.raises(I.IPGError):
        I.Setup.read("/nonexistent/input")


def test_version():
    import ipglasma_amd as I
    assert I.lib().ipg_version() >= 100


def test_cxx_call_surface_compiles_and_links(tmp_path):
    """A driver written against the reference's class names (main.cpp:167-360)
    compiles against include/ipglasma.hpp and links to the library, with the
    Wilson-line files (READFROMFILE, Init::readV), the per-event id and MPI
    size that name them, and the writeOutputs files of Evolution::run."""
    src = tmp_path / "drv.cpp"
    src.write_text('''
#include "ipglasma.hpp"
using namespace ipglasma;
int main(int argc, char** argv) {
    Parameters param; Setup setup;
    if (argc > 1) setup.read(argv[1], &param);
    param.setMPISize(1);
    Random random; random.init_genrand64(param.getSeed());
    int nn[2] = {param.getSize(), param.getSize()};
    Init init(nn); Group group(param.getNc()); Glauber glauber;
    Evolution evolution(nn);
    for (int iev = 0; iev < 2; iev++) {
        param.setEventId(iev);
        param.setSuccess(0);
        Lattice lat(&param, param.getNc(), param.getSize());
        const int READFROMFILE = param.getReadInitialWilsonLines();
        while (param.getSuccess() == 0) {
            init.init(&lat, &group, &param, &random, &glauber, READFROMFILE);
            if (param.getSuccess() == 0) continue;
            evolution.run(&lat, &group, &param);
        }
        if (Init::vname(&param, 1, 2) != "V-" + std::to_string(iev + 1 + 2 * param.getSeed())) return 2;
    }
    return 0;
}
''')
    import ipglasma_amd as I
    libdir = os.path.dirname(I.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "drv"),
                    "-L", libdir, "-lipglasma_amd", "-Wl,-rpath," + libdir], check=True)
