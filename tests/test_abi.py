"""The C-ABI libraries load without a GPU and export every function that
include/*.h declares; argument validation happens on the host (CPU only)."""
import ctypes as C
import re
import subprocess

import pytest

from conftest import ROOT


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xcal[mw]?_[a-z0-9_]+)\s*\(", text)))


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header,so", [("xcal.h", "libxcalhip.so"),
                                       ("xcal_mesh.h", "libxcal_mesh.so"),
                                       ("xcal_io.h", "libxcal_mesh.so")])
def test_library_exports_every_declared_symbol(x, header, so):
    path = ROOT / "xcalibre.jl_amd" / so
    C.CDLL(str(path))  # loads on a host without a GPU
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if n not in exported(path)]
    assert not missing, missing


def test_hip_library_abi_version_and_errors(x):
    L = x.hip.lib()
    assert L.xcal_abi_version() == 2
    # argument checks run before any device call
    assert L.xcal_ctx_create(0, None) == x._abi.OK - 1
    assert b"NULL" in L.xcal_last_error()
    assert L.xcal_mesh_info(None, None, None, None, None, None) < 0
    assert L.xcal_field_upload(None, 0, None) < 0


def test_reference_shaped_api_validation(x, meshes):
    api = x.api
    m = meshes("cavity16")
    with pytest.raises(AssertionError, match="Inconsistent number of boundaries"):
        api.assign(region=m, U=[api.Wall("top", [1, 0, 0])])
    with pytest.raises(KeyError, match="not a recognised boundary name"):
        api.assign(region=m, p=[api.Wall("inlet")])
    with pytest.raises(ValueError, match="3 components"):
        api.assign(region=m, U=[api.Wall(n, [1, 0]) for n in m.patch_names])
    s = api.SolverSetup(solver=api.Cg(), preconditioner=api.DILU(), convergence=1e-7, relax=0.3)
    assert abs(s.atol - 2.220446049250313e-16 ** 0.9) < 1e-30 and s.rtol == 0.1 and s.itmax == 1000
    U = x.api.Physics(domain=m, fluid=api.Fluid(nu=1.0)).momentum.U
    api.initialise(U, [1, 2, 3])
    assert (U[:, 5] == [1, 2, 3]).all()


def test_ctypes_structs_match_the_c_header(x, tmp_path):
    """Every ctypes mirror of an include/xcal.h struct has the C struct's size
    and each field's offset, as gcc lays them out (host only)."""
    A = x._abi
    pairs = {"MeshHost": "xcal_mesh_host", "BC": "xcal_bc", "SolverSetupC": "xcal_solver_setup",
             "SolveStats": "xcal_solve_stats", "SimpleSetupC": "xcal_simple_setup",
             "TurbulenceSetupC": "xcal_turbulence_setup", "ServerConfigC": "xcal_server_config",
             "RequestStatusC": "xcal_request_status"}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "xcal.h"', "int main(void) {"]
    expect = {}
    for py, c in pairs.items():
        S = getattr(A, py)
        lines.append(f'  printf("{c} sizeof %zu\\n", sizeof({c}));')
        expect[f"{c} sizeof"] = C.sizeof(S)
        for f in S._fields_:
            lines.append(f'  printf("{c} {f[0]} %zu\\n", offsetof({c}, {f[0]}));')
            expect[f"{c} {f[0]}"] = getattr(S, f[0]).offset
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {" ".join(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}
    assert got == expect
