"""The drop-in solver classes keep the reference's declarations, and a
caller written like feMorph's own (solveExplicitNS, src/main.cpp:788-805;
solveIncompressibleNS, :687-705) compiles against them and links to the
shipped libraries (no GPU needed)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "femorph_amd", "csrc", "host")
LIBDIR = os.path.join(ROOT, "femorph_amd")
REF_SRC = os.environ.get("FEMX_REFERENCE", "/root/reference") + "/src"

CALLER = r"""
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "femModel.h"
#include "femSolver.h"
#include "femIncompressibleSolver.h"
#include "femSUPGExplicitFluidSolver.h"
#include "femVMSExplicitFluidSolver.h"
#include "femx.h"

static_assert(std::is_base_of<femSolver, femVMSExplicitFluidSolver>::value, "VMS derives from femSolver");
static_assert(std::is_base_of<femSolver, femSUPGExplicitFluidSolver>::value, "SUPG derives from femSolver");
static_assert(std::is_base_of<femSolver, femIncompressibleSolver>::value, "INS derives from femSolver");
static_assert(sizeof(femVMSExplicitFluidSolver) == sizeof(femSolver), "no data members beyond the base");
static_assert(sizeof(femSUPGExplicitFluidSolver) == sizeof(femSolver), "no data members beyond the base");
static_assert(sizeof(femIncompressibleSolver) == sizeof(femSolver), "no data members beyond the base");

// written as src/main.cpp:788-805
int solveExplicitNS(const char* inputFileName) {
  femModel* model = new femModel();
  model->ReadFromFEMTextFile(inputFileName);
  femSUPGExplicitFluidSolver* exNS = new femSUPGExplicitFluidSolver();
  exNS->solve(model);
  return 0;
}

int solveVMS(const char* inputFileName) {
  femModel* model = new femModel();
  model->ReadFromFEMTextFile(inputFileName);
  femVMSExplicitFluidSolver* exNS = new femVMSExplicitFluidSolver();
  exNS->solve(model);
  return 0;
}

int main(int argc, char** argv) {
  // the base's solve(femOption*, femModel*) throws, as src/femSolver.cpp:194-196
  femSolver* base = new femVMSExplicitFluidSolver();
  try {
    base->solve(nullptr, nullptr);
    return 3;
  } catch (femException& ex) {
    printf("base: %s", ex.what());
  }
  if (femx_default_precision() != FEMX_FP64 && !getenv("FEMX_PRECISION")) return 4;
  if (argc > 2) {
    try {
      if (!strcmp(argv[1], "vms")) solveVMS(argv[2]);
      else solveExplicitNS(argv[2]);
    } catch (femException& ex) {
      printf("solve: %s", ex.what());
      return 5;
    }
  }
  printf("ok\n");
  return 0;
}
"""


def _built():
    for f in ("libfemx.so", "libfemx_model.so"):
        if not os.path.exists(os.path.join(LIBDIR, f)):
            pytest.skip(f"{f} not built")


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    _built()
    d = tmp_path_factory.mktemp("caller")
    src = d / "main.cpp"
    src.write_text(CALLER)
    exe = d / "caller"
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror", "-I", HOST, "-I", os.path.join(ROOT, "include"), str(src),
           "-o", str(exe), "-L", LIBDIR, "-l:libfemx_model.so", "-l:libfemx.so", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_caller_compiles_links_and_runs(caller):
    r = subprocess.run([str(caller)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Solve Not Implemented for femSolver." in r.stdout and r.stdout.strip().endswith("ok")


def test_caller_solve_reaches_the_library(caller, tmp_path):
    """Without a GPU the solve stops at the femx device check and surfaces as
    femException; with one it runs (covered by the GPU drop-in tests)."""
    from femorph_amd import femx
    if femx.device_count() > 0:
        pytest.skip("a GPU is visible")
    import gzip
    model = tmp_path / "model.txt"
    model.write_bytes(gzip.open(os.path.join(ROOT, "tests", "golden", "ldc_coarse_model.txt.gz")).read())
    r = subprocess.run([str(caller), "vms", str(model)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 5 and "no HIP device" in r.stdout, r.stdout + r.stderr


def _decl(text, cls):
    text = re.sub(r"//[^\n]*", "", re.sub(r"/\*.*?\*/", "", text, flags=re.S))
    m = re.search(r"class\s+" + cls + r"\b[^;{]*\{.*?\};", text, flags=re.S)
    assert m, cls
    return re.sub(r"\s+", "", m.group(0))


@pytest.mark.parametrize("cls,hdr", [("femVMSExplicitFluidSolver", "femVMSExplicitFluidSolver.h"),
                                     ("femSUPGExplicitFluidSolver", "femSUPGExplicitFluidSolver.h"),
                                     ("femIncompressibleSolver", "femIncompressibleSolver.h")])
def test_declarations_equal_the_reference(cls, hdr):
    """Token-for-token the same class (comments and white space aside), so a
    feMorph build that keeps its own headers and swaps in the drop-in
    sources sees one definition of each class."""
    ref = os.path.join(REF_SRC, hdr)
    if not os.path.exists(ref):
        pytest.skip("reference sources not present")
    assert _decl(open(os.path.join(HOST, hdr)).read(), cls) == _decl(open(ref).read(), cls)


def test_solver_base_declares_the_reference_interface():
    ref = os.path.join(REF_SRC, "femSolver.h")
    if not os.path.exists(ref):
        pytest.skip("reference sources not present")
    ours = _decl(open(os.path.join(HOST, "femSolver.h")).read(), "femSolver")
    theirs = re.sub(r"#ifdef.*?#endif", "", _decl(open(ref).read(), "femSolver"), flags=re.S)
    assert ours == theirs


def test_dropin_sources_include_no_standalone_only_header():
    """The drop-in sources sit apart from the standalone model's headers, so
    in a feMorph build their includes resolve to the reference's."""
    dropin = os.path.join(ROOT, "femorph_amd", "csrc", "dropin")
    assert not any(f.endswith(".h") and os.path.exists(os.path.join(HOST, f)) for f in os.listdir(dropin))
    for f in os.listdir(dropin):
        if f.endswith(".cpp"):
            inc = re.findall(r'#include\s+"([^"]+)"', open(os.path.join(dropin, f)).read())
            assert set(inc) <= {"femModel.h", "femException.h", "femx.h", "femx_dropin.h", f[:-4] + ".h"}, (f, inc)
