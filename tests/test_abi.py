"""CPU checks of the C ABI: the library loads, exports every function that
include/sap_c.h declares, and the option/stat structs match the header.
No GPU call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sap_c.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sap_[a-z_0-9]+)\s*\(", text)
    return sorted(set(n for n in names if not n.startswith("sap_spmv_fn")))


@pytest.fixture(scope="module")
def sap():
    import saplibrary_amd
    if not os.path.exists(saplibrary_amd.LIB_PATH):
        pytest.skip("libsap_amd.so not built")
    saplibrary_amd.lib()
    return saplibrary_amd


def test_header_declares_the_reference_surface():
    names = _declared_functions()
    for n in ("sap_create", "sap_destroy", "sap_setup_banded", "sap_solve", "sap_get_stats",
              "sap_monitor_code", "sap_monitor_message", "sap_precond_apply", "sap_band_spmv"):
        assert n in names


def test_library_exports_every_declared_symbol(sap):
    L = ctypes.CDLL(sap.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(sap.EXPORTS) == _declared_functions()


def test_default_options_match_reference(sap):
    # sap::Options::Options() (sap/solver.h:232-256)
    o = sap.Options()
    assert o.solverType == sap.BiCGStab2 and o.maxNumIterations == 100
    assert o.relTol == 1e-6 and o.absTol == 0
    assert o.performReorder == 1 and o.performDB == 1 and o.applyScaling == 1
    assert o.maxBandwidth == 2 ** 31 - 1 and o.gpuCount == 1 and o.dropOffFraction == 0
    assert o.factMethod == sap.LU_only and o.precondType == sap.Spike
    assert o.safeFactorization == 0 and o.variableBandwidth == 1 and o.trackReordering == 0
    assert o.useBCR == 0 and o.ilu_level == -1
    assert o.precValueType == sap.PREC_FLOAT and o.refine == 0


def test_struct_layouts_match_header(sap):
    text = open(HEADER).read()

    def fields(struct):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        return re.findall(r"\b(?:int|double|float|const char\*)\s+([A-Za-z_0-9]+)\s*;", body)

    assert [f for f, _ in sap.Options._fields_] == fields("sap_options")
    assert [f for f, _ in sap.Stats._fields_] == fields("sap_stats")


def test_create_rejects_unsupported_options(sap):
    # ILU / several GPUs / the CUSP Krylov solvers are not part of this build,
    # and BCR does not take SPD matrices (as in the reference): rejected at create
    for kw in ({"useBCR": 1, "isSPD": 1}, {"ilu_level": 0}, {"gpuCount": 2}, {"solverType": sap.GMRES_C}):
        with pytest.raises(sap.SapError):
            sap.Solver(4, sap.Options(**kw))


def test_cpp_facade_compiles(tmp_path):
    # include/sap/solver.hpp: the reference-style C++ interface over the ABI
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "facade.cpp"
    src.write_text(
        '#include "sap/solver.hpp"\n'
        "int main() {\n"
        "  typedef sap::DeviceArray<double> Vec;\n"
        "  sap::Options opts; opts.factMethod = sap::LU_UL; opts.relTol = 1e-10;\n"
        "  Vec band(30), b(10), x(10);\n"
        "  sap::BandedMatrix<Vec> A(10, 1, band);\n"
        "  sap::Solver<Vec, float> solver(2, opts);\n"
        "  sap::MVBanded<Vec> spmv(A);\n"
        "  solver.setup(A);\n"
        "  bool ok = solver.solve(spmv, b, x);\n"
        "  sap::DeviceArray<int> rp(11), ci(28);\n"
        "  Vec val(28);\n"
        "  sap::CsrMatrix<sap::DeviceArray<int>, Vec> C(10, 28, rp, ci, val);\n"
        "  opts.variableBandwidth = false; opts.trackReordering = true;\n"
        "  sap::Solver<Vec, double> s2(2, opts);\n"
        "  sap::SpmvCusp<sap::CsrMatrix<sap::DeviceArray<int>, Vec> > cs(C);\n"
        "  s2.setup(C);\n"
        "  s2.update(val);\n"
        "  return (ok && s2.solve(cs, b, x)) ? 0 : 1;\n"
        "}\n")
    r = subprocess.run([hipcc, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o",
                        str(tmp_path / "facade.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
