"""The C-ABI library loads and exports every entry point include/amrmg.h declares
(no device call: runs without a GPU)."""
import ctypes
import os
import subprocess

import amrmg


def _ensure_built():
    if not os.path.exists(amrmg.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.dirname(amrmg.LIB_PATH)], check=True)


def test_header_declares_entry_points():
    syms = amrmg.header_symbols()
    for s in ["amrmg_ctx_create", "amrmg_solver_create", "amrmg_solver_solve", "amrmg_level_relax",
              "amrmg_quad_cf", "amrmg_bottom_solve", "amrmg_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    _ensure_built()
    L = ctypes.CDLL(amrmg.LIB_PATH)
    missing = [s for s in amrmg.header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_last_error_callable_without_device():
    _ensure_built()
    L = amrmg.lib()
    assert isinstance(L.amrmg_last_error(), bytes)


def test_cxx_caller_compiles_links_and_starts(tmp_path):
    """The header compiles as C++ and a C++ caller links against the library
    (tests/cxx/abi_driver.cpp, built the way tests/cxx/Makefile builds it); run
    without arguments it prints its usage and stops before any device call."""
    _ensure_built()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "abi_driver"
    libdir = os.path.dirname(amrmg.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"), "-o",
                    str(exe), os.path.join(root, "tests", "cxx", "abi_driver.cpp"), "-L", libdir, "-lamrmg",
                    "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
