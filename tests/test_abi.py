"""The C-ABI library loads, exports every entry point of include/adm.h, and
reports argument errors without a GPU (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(adm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from adamantine_amd import thermal
    lib = thermal.load_library()
    names = declared("adm.h")
    assert "adm_vmult" in names and "adm_evolve" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    """ctypes mirrors of the descriptor structs have the C sizes."""
    from adamantine_amd.problem import (BeamDesc, BeamState, MaterialDesc,
                                        MeshDesc, ProblemDesc, Segment)
    src = r'''
#include <stdio.h>
#include "adm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(adm_material_desc),
         sizeof(adm_mesh_desc), sizeof(adm_segment), sizeof(adm_beam_desc),
         sizeof(adm_beam_state), sizeof(adm_problem_desc));
  return 0;
}
'''
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "s.c"), "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o",
                        os.path.join(d, "s"), os.path.join(d, "s.c")],
                       check=True)
        out = subprocess.run([os.path.join(d, "s")], capture_output=True,
                             text=True, check=True).stdout.split()
    sizes = [C.sizeof(t) for t in (MaterialDesc, MeshDesc, Segment, BeamDesc,
                                   BeamState, ProblemDesc)]
    assert [int(x) for x in out] == sizes


def test_invalid_descriptor_is_an_error_not_a_crash():
    """adm_create validates before touching the device (ASSERT_THROW
    semantics: an error code and a message)."""
    from adamantine_amd import thermal
    from adamantine_amd.problem import Problem
    lib = thermal.load_library()
    for bad in (Problem(dim=1, n_cells=(4, 1, 5)),
                Problem(fe_degree=3),
                Problem(n_cells=(0, 4, 4))):
        d, keep = bad.to_c()
        h = C.c_void_p()
        rc = lib.adm_create(C.byref(d), C.byref(h))
        assert rc != 0
        assert lib.adm_last_error()
    assert lib.adm_vmult(None, None, None, 0, None) != 0
    assert b"context" in lib.adm_last_error()


def test_oracle_is_not_linked_by_the_library():
    """The product library must not depend on the oracle."""
    import subprocess
    from adamantine_amd import thermal
    out = subprocess.run(["ldd", thermal.LIB_PATH], capture_output=True,
                         text=True).stdout
    assert "oracle" not in out


def test_cxx_adapter_compiles_against_the_header_alone(tmp_path):
    """include/adm_cxx.hpp needs nothing but adm.h and the C++ standard
    library: a host compiler checks it without HIP headers."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "adm_cxx.hpp"\n'
                   'int f(const adm_problem_desc &d, double *x, double *y) {\n'
                   '  adm_cxx::ThermalPhysics ph(d, ADM_RK_FOURTH_ORDER);\n'
                   '  ph.get_thermal_operator().vmult(y, x);\n'
                   '  return (int)ph.get_thermal_operator().m();\n}\n')
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror",
                        "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_wrong_length_vector_is_refused_before_any_launch():
    """A device vector whose length is not the one the kernels index is
    refused by the wrapper, before the library is called."""
    import torch
    from adamantine_amd.thermal import AdmError, _dev_ptr
    with pytest.raises(AdmError, match="5 values where 6 are needed"):
        _dev_ptr(torch.zeros(5, dtype=torch.float64), 6)
    with pytest.raises(AdmError, match="where 6 are needed"):
        _dev_ptr(None, 6)
    # without a length only the kind of tensor is checked (a CPU tensor is
    # not a device vector)
    with pytest.raises(AdmError, match="HIP tensors"):
        _dev_ptr(torch.zeros(6, dtype=torch.float64), 6)
