"""The C ABI library loads and exports every symbol the public headers
declare (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

from vidyut3d_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("vidyut3d_amd.h", "vidyut3d_amd_host.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(vd_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_the_c_header(tmp_path):
    """ctypes mirrors have the C sizes and field offsets (checked with gcc)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        import pytest
        pytest.skip("gcc not available")
    structs = {"vd_array4": abi.vd_array4, "vd_tagarray4": abi.vd_tagarray4, "vd_box": abi.vd_box,
               "vd_geom": abi.vd_geom, "vd_params": abi.vd_params, "vd_reaction": abi.vd_reaction,
               "vd_mech_desc": abi.vd_mech_desc, "vd_solve_info": abi.vd_solve_info,
               "vd_transport_fit": abi.vd_transport_fit,
               "vd_mlmg_problem": abi.vd_mlmg_problem, "vd_regrid_params": abi.vd_regrid_params}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vidyut3d_amd_host.h"', "int main(void){"]
    for n, st in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (n, n))
        for f in st._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (n, f[0], n, f[0]))
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for n, st in structs.items():
        assert int(out[n]) == ctypes.sizeof(st), n
        for f in st._fields_:
            assert int(out["%s.%s" % (n, f[0])]) == getattr(st, f[0]).offset, (n, f[0])


def test_params_default_matches_reference_defaults():
    p = abi.vd_params()
    native.lib().vd_params_default(ctypes.byref(p))
    assert p.hyp_order == 1 and p.weno_scheme == 3
    assert p.linsolve_reltol == 1e-10 and p.linsolve_abstol == 1e-10
    assert p.linsolve_maxiter == 100 and p.min_electron_density == 1e12
    assert p.gas_temperature == 300.0 and p.gas_pressure == 101325.0
    assert list(p.pot_bc_lo) == [abi.ROBINBC] * 3


def test_version_string():
    assert b"gfx950" in native.lib().vd_version()


def test_null_context_is_rejected_without_touching_the_device():
    """Memory entry points given no context return VD_EINVAL (a handle that
    outlives its context must not dereference it)."""
    L = native.lib()
    p = ctypes.c_void_p()
    buf = (ctypes.c_char * 8)()
    assert L.vd_free(None, None) == abi.VD_EINVAL
    assert L.vd_malloc(None, ctypes.c_size_t(8), ctypes.byref(p)) == abi.VD_EINVAL
    assert L.vd_memcpy_h2d(None, None, buf, ctypes.c_size_t(8)) == abi.VD_EINVAL
    assert L.vd_memcpy_d2h(None, buf, None, ctypes.c_size_t(8)) == abi.VD_EINVAL
    assert L.vd_memset_zero(None, None, ctypes.c_size_t(8)) == abi.VD_EINVAL
    assert L.vd_ctx_sync(None) == abi.VD_EINVAL
    assert L.vd_mech_destroy(None) == abi.VD_EINVAL
    assert L.vd_sim_destroy(None) == abi.VD_EINVAL
