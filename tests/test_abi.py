"""The C ABI library loads and exports every entry point include/moed.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "moed.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(moed_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for required in ("moed_hv", "moed_apply_c", "moed_lanczos", "moed_solve", "moed_green", "moed_chi2_grad",
                     "moed_minimize", "moed_basis", "moed_calc_dev"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import moed
    if not os.path.exists(moed.LIB_PATH):
        pytest.fail("libmoed.so is not built")
    lib = ctypes.CDLL(moed.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(moed.EXPORTS) == declared_functions()


def test_host_only_entry_points():
    """moed_omega_grid and argument validation run without a device."""
    import numpy as np
    import moed
    from moed.model import omega_grid
    om = moed.omega_grid(64, 25.0)
    assert np.array_equal(om, omega_grid(64, 25.0))
    lib = moed.lib()
    assert lib.moed_create(None, None) == -1
    assert lib.moed_omega_grid(-1, 1.0, None) == -1


def test_create_rejects_bad_sizes_before_any_device_call():
    """moed_create validates the configuration before it touches the device: more than 4 orbitals,
    more than 15 sites, fewer sites than orbitals and negative counts are MOED_EARG"""
    import numpy as np
    import moed
    for norb, nbath, kw in [(5, 2, {}), (2, 14, {}), (3, -1, {}), (2, 2, {"nstep": -1}), (2, 2, {"nev": -1})]:
        ns = max(norb + nbath, 1)
        with pytest.raises(moed.MoedError) as ei:
            moed.EDSolver(norb, nbath, 1.0, 0.1, np.zeros(ns), np.full((norb, max(nbath, 0)), 0.5), **kw)
        assert ei.value.code == moed.MOED_EARG, (norb, nbath, kw)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_create_fails_loudly_without_a_device():
    """with no GPU the library raises MOED_EHIP at moed_create: there is no host fallback"""
    import numpy as np
    import moed
    with pytest.raises(moed.MoedError) as ei:
        moed.EDSolver(2, 2, 1.0, 0.1, np.zeros(4), np.full((2, 2), 0.5))
    assert ei.value.code == moed.MOED_EHIP


def test_fortran_interface_compiles():
    """the ISO_C_BINDING module compiles with the image's Fortran compiler"""
    import shutil
    import subprocess
    import tempfile
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "mo-dmft_amd", "fortran", "moed_iface.f90")
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([fc, "-c", src, "-module-dir", td, "-o", os.path.join(td, "m.o")], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr


def test_fortran_bindings_name_declared_symbols():
    """every bind(C, name=...) in the Fortran module is a function include/moed.h declares"""
    src = open(os.path.join(ROOT, "mo-dmft_amd", "fortran", "moed_iface.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(moed_[a-z0-9_]+)"\)', src))
    assert bound and bound <= set(declared_functions()), sorted(bound - set(declared_functions()))
