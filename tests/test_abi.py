"""CPU checks of the C ABI: the library loads, exports every entry point that
include/ipg.h declares, agrees with the Python binding on struct layouts,
and parses reference-format input files (no device calls)."""
import os
import re
import subprocess

import numpy as np
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


READINPUT_DRIVER = r"""
#include <cstdio>
#include "ipglasma.hpp"
using namespace ipglasma;
using std::string;
// readInput as main.cpp:448-578 drives Setup and Parameters
int readInput(Setup *setup, Parameters *param, const string& file_name) {
    param->setNucleusQsTableFileName(setup->StringFind(file_name, "NucleusQsTableFileName"));
    param->setNucleonPositionsFromFile(setup->IFind(file_name, "nucleonPositionsFromFile"));
    param->setTarget(setup->StringFind(file_name, "Target"));
    param->setProjectile(setup->StringFind(file_name, "Projectile"));
    param->setMode(setup->IFind(file_name, "mode"));
    param->setRunningCoupling(setup->IFind(file_name, "runningCoupling"));
    param->setL(setup->DFind(file_name, "L"));
    param->setLOutput(setup->DFind(file_name, "LOutput"));
    param->setBG(setup->DFind(file_name, "BG"));
    param->setBGq(setup->DFind(file_name, "BGq"));
    param->setBGqVar(setup->DFind(file_name, "BGqVar"));
    param->setDqmin(setup->DFind(file_name, "dqMin"));
    param->setMuZero(setup->DFind(file_name, "muZero"));
    param->setc(setup->DFind(file_name, "c"));
    param->setSize(setup->IFind(file_name, "size"));
    param->setSizeOutput(setup->IFind(file_name, "sizeOutput"));
    param->setEtaSizeOutput(setup->IFind(file_name, "etaSizeOutput"));
    param->setDetaOutput(setup->DFind(file_name, "detaOutput"));
    param->setUseFluctuatingx(setup->IFind(file_name, "useFluctuatingx"));
    param->setNc(setup->IFind(file_name, "Nc"));
    param->setInverseQsForMaxTime(setup->IFind(file_name, "inverseQsForMaxTime"));
    param->setSeed(setup->ULLIFind(file_name, "seed"));
    param->setUseSeedList(setup->IFind(file_name, "useSeedList"));
    param->setNy(setup->IFind(file_name, "Ny"));
    param->setRoots(setup->DFind(file_name, "roots"));
    param->setNu(setup->DFind(file_name, "tDistNu"));
    param->setUseFatTails(setup->IFind(file_name, "useFatTails"));
    param->setg(setup->DFind(file_name, "g"));
    param->setm(setup->DFind(file_name, "m"));
    param->setJacobianm(setup->DFind(file_name, "Jacobianm"));
    param->setSigmaNN(setup->DFind(file_name, "SigmaNN"));
    param->setRmax(setup->DFind(file_name, "rmax"));
    param->setUVdamp(setup->DFind(file_name, "UVdamp"));
    param->setSetWSDeformParams(setup->IFind(file_name, "setWSDeformParams"));
    if (param->getSetWSDeformParams()) {
        param->setR_WS(setup->DFind(file_name, "R_WS"));
        param->setA_WS(setup->DFind(file_name, "a_WS"));
        param->setBeta2(setup->DFind(file_name, "beta2"));
        param->setBeta3(setup->DFind(file_name, "beta3"));
        param->setBeta4(setup->DFind(file_name, "beta4"));
        param->setGamma(setup->DFind(file_name, "gamma"));
        param->setForceDmin(setup->DFind(file_name, "force_dmin_flag"));
        param->setDmin(setup->DFind(file_name, "d_min"));
        param->setWSdR_np(setup->DFind(file_name, "dR_np"));
        param->setWSda_np(setup->DFind(file_name, "da_np"));
    }
    param->setbmin(setup->DFind(file_name, "bmin"));
    param->setbmax(setup->DFind(file_name, "bmax"));
    param->setRotateReactionPlane(setup->IFind(file_name, "rotateReactionPlane"));
    param->setComputeGluonMultiplicity(setup->IFind(file_name, "computeGluonMultiplicity"));
    param->setQsmuRatio(setup->DFind(file_name, "QsmuRatio"));
    param->setUsePseudoRapidity(setup->DFind(file_name, "usePseudoRapidity"));
    param->setRapidity(setup->DFind(file_name, "Rapidity"));
    param->setUseNucleus(setup->IFind(file_name, "useNucleus"));
    param->setUseGaussian(setup->IFind(file_name, "useGaussian"));
    param->setlightNucleusOption(setup->IFind(file_name, "lightNucleusOption"));
    param->setg2mu(setup->DFind(file_name, "g2mu"));
    param->setMaxtime(setup->DFind(file_name, "maxtime"));
    double lattice_a = param->getL() / static_cast<double>(param->getSize());
    int iTimeSteps = static_cast<int>(10 * param->getMaxtime() / lattice_a) + 1;
    param->setdtau(param->getMaxtime() / (iTimeSteps * lattice_a));
    param->setRunWithQs(setup->IFind(file_name, "runWith0Min1Avg2MaxQs"));
    param->setRunWithkt(setup->IFind(file_name, "runWithkt"));
    param->setRunWithLocalQs(setup->IFind(file_name, "runWithLocalQs"));
    param->setRunWithThisFactorTimesQs(setup->DFind(file_name, "runWithThisFactorTimesQs"));
    param->setxFromThisFactorTimesQs(setup->DFind(file_name, "xFromThisFactorTimesQs"));
    param->setLinearb(setup->IFind(file_name, "samplebFromLinearDistribution"));
    param->setWriteOutputs(setup->IFind(file_name, "writeOutputs"));
    param->setWriteOutputsToHDF5(setup->IFind(file_name, "writeOutputsToHDF5"));
    param->setWriteEvolution(setup->IFind(file_name, "writeEvolution"));
    param->setWriteInitialWilsonLines(setup->IFind(file_name, "writeInitialWilsonLines"));
    param->setReadInitialWilsonLines(setup->IFind(file_name, "readInitialWilsonLines"));
    param->setAverageOverNuclei(setup->IFind(file_name, "averageOverThisManyNuclei"));
    param->setUseTimeForSeed(setup->IFind(file_name, "useTimeForSeed"));
    param->setUseFixedNpart(setup->IFind(file_name, "useFixedNpart"));
    param->setSmearQs(setup->IFind(file_name, "smearQs"));
    param->setSmearingWidth(setup->DFind(file_name, "smearingWidth"));
    param->setGaussianWounding(setup->IFind(file_name, "gaussianWounding"));
    param->setReadMultFromFile(setup->IFind(file_name, "readMultFromFile"));
    param->setProtonAnisotropy(setup->DFind(file_name, "protonAnisotropy"));
    param->setUseConstituentQuarkProton(setup->DFind(file_name, "useConstituentQuarkProton"));
    param->setNqBase(setup->DFind(file_name, "useConstituentQuarkProton"));
    param->setNqFluc(setup->DFind(file_name, "NqFluc"));
    param->setUseSmoothNucleus(setup->IFind(file_name, "useSmoothNucleus"));
    param->setShiftConstituentQuarkProtonOrigin(setup->DFind(file_name, "shiftConstituentQuarkProtonOrigin"));
    param->setMinimumQs2ST(setup->IFind(file_name, "minimumQs2ST"));
    param->setSubNucleonParamType(setup->IFind(file_name, "SubNucleonParamType"));
    param->setSubNucleonParamSet(setup->IFind(file_name, "SubNucleonParamSet"));
    if (param->getSubNucleonParamType() > 0) param->loadPosteriorParameterSets(param->getSubNucleonParamType());
    return 0;
}
// the set-up of main.cpp:69-188 up to the objects of one event
int main(int argc, char** argv) {
    Parameters *param = new Parameters();
    param->setMPIRank(0);
    param->setMPISize(1);
    Setup setup;
    readInput(&setup, param, argv[1]);
    Random *random = new Random();
    unsigned long long rnum = param->getSeed();
    param->setRandomSeed(rnum + 0 * 1000);
    random->init_genrand64(rnum);
    random->gslRandomInit(rnum);
    int iev = 0;
    if (param->getSubNucleonParamType() > 0)
        param->setParamsWithPosteriorParameterSet(param->getSubNucleonParamType(), random->genrand64_int63());
    param->setEventId(param->getMPIRank() + iev * param->getMPISize());
    param->setSuccess(0);
    int nn[2] = {param->getSize(), param->getSize()};
    Glauber glauber;
    glauber.initGlauber(param->getSigmaNN(), param->getTarget(), param->getProjectile(), param->getb(),
                        param->getSetWSDeformParams(), param->getR_WS(), param->getA_WS(), param->getBeta2(),
                        param->getBeta3(), param->getBeta4(), param->getGamma(), param->getForceDmin(),
                        param->getDmin(), param->getWSdR_np(), param->getWSda_np(), 100);
    std::printf("%d %d %s %s %.17g %.17g %llu %d %.17g %.17g %d %.17g %d %d %.17g %s %d %d %.17g %lld\n",
                nn[0], param->getNy(), param->getProjectile().c_str(), param->getTarget().c_str(),
                param->getL(), param->getdtau(), (unsigned long long)param->getRandomSeed(), param->getItmax(),
                param->getR_WS(), param->getQsmuRatio(), (int)param->getSetWSDeformParams(), param->getSigmaNN(),
                param->getLinearb(), param->getRunWithQs(), param->getBGq(),
                param->getNucleusQsTableFileName().c_str(), param->getSubNucleonParamSet(),
                (int)param->getRotateReactionPlane(), param->getRunWithThisFactorTimesQs(),
                random->genrand64_int63());
    // a missing key is an error (the reference exits)
    try {
        setup.IFind(argv[1], "noSuchKey");
        return 3;
    } catch (const Error&) {
    }
    return 0;
}
"""


def test_readinput_driver_against_cxx_surface(tmp_path):
    """readInput (main.cpp:448-578) and the set-up of main.cpp:69-188 written
    against the reference's Setup/Parameters/Random/Glauber methods compile
    against include/ipglasma.hpp and read the input file as the reference
    does: StringFind's pair scan to EndOfFile (a key after it is ignored),
    IFind truncation, ULLIFind rounding, dtau derived from L, size, maxtime;
    the Python reader of the C ABI agrees"""
    import ipglasma_amd as I
    src = tmp_path / "readinput.cpp"
    src.write_text(READINPUT_DRIVER)
    libdir = os.path.dirname(I.LIB_PATH)
    exe = tmp_path / "readinput"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lipglasma_amd", "-Wl,-rpath," + libdir], check=True)
    inp = os.path.join(ROOT, "tests", "golden", "input_AuAu_C1")
    out = subprocess.run([str(exe), inp], capture_output=True, text=True, check=True).stdout.split()
    p = I.Setup.read(inp)
    a = 30.0 / 256
    steps = int(10 * 0.4 / a) + 1
    assert int(out[0]) == 256 and int(out[1]) == 50
    assert out[2] == "Au" and out[3] == "Au"
    assert float(out[4]) == 30.0
    assert float(out[5]) == 0.4 / (steps * a) == p.dtau
    assert int(out[6]) == 14 == p.seed            # ULLIFind: (ull)(13.6 + 0.5)
    assert int(out[7]) == 35 == p.itmax
    assert float(out[8]) == 6.38 and float(out[9]) == 0.8 and int(out[10]) == 1
    assert float(out[11]) == 42.0 and int(out[12]) == 1 and int(out[13]) == 2
    assert float(out[14]) == 0.3 and out[15] == "qs2Adj_vs_Tp_vs_Y_200.in"
    assert int(out[16]) == -1 and int(out[17]) == 1 and float(out[18]) == 0.5
    # genrand64_int63 of MT19937-64 seeded 14 (Random.cpp:67-120)
    import _oracle as O
    assert int(out[19]) == int(O.mt_stream(14, 0, 1).view(np.uint64)[0]) >> 1
