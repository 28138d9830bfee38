"""The C ABI: libbfamx.so loads, exports every function include/bfamx.h
declares, the Python front end binds each of them, and the host-only entry
points (no device needed) honour the status-code contract.  CPU only: no
device call is made."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bfam_amd as bx

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "bfamx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(bfamx_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("bfamx_device_create", "bfamx_subdomain_upload", "bfamx_lsrk_init",
                 "bfamx_rhs", "bfamx_stage", "bfamx_step", "bfamx_batch_capture",
                 "bfamx_download_fields", "bfamx_reduce_energy", "bfamx_reduce_l2_diff"):
        assert must in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(bx.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binds_every_declared_symbol():
    assert set(declared_functions()) <= set(bx.SIGNATURES), \
        sorted(set(declared_functions()) - set(bx.SIGNATURES))


def test_status_codes_and_last_error():
    L = bx.lib()
    r = np.zeros(4)
    w = np.zeros(4)
    assert L.bfamx_lgl(-1, r, w) == bx.EINVAL
    msg = L.bfamx_last_error()
    assert msg and b"bfamx_lgl" in msg
    assert L.bfamx_lgl(3, r, w) == bx.OK
    with pytest.raises(bx.BfamxError) as e:
        # one element whose faces name a neighbour that does not exist
        bx.build_vmaps(3, 2, np.full(6, 5, np.int32), np.zeros(6, np.int8))
    assert e.value.code == bx.EINVAL


def test_subdomain_bytes_budget_formula():
    """Resident bytes of a subdomain grow with the state type: FP32 state and
    rates need about half of FP64 (checked without a device)."""
    K = 4096
    b64 = bx.subdomain_bytes(3, 7, K, bx.ELASTIC3D, bx.FP64)
    b32 = bx.subdomain_bytes(3, 7, K, bx.ELASTIC3D, bx.FP32)
    state64 = 9 * K * 512 * 8 * 2  # two state buffers; dq is rebuilt, not stored
    assert state64 <= b64 < 1.1 * state64 and b32 < 0.6 * b64
    b32r = bx.subdomain_bytes(3, 7, K, bx.ELASTIC3D, bx.FP32_RHS)
    assert b32r >= 9 * K * 512 * (8 * 2 + 4)  # FP32_RHS stores an FP32 dq


def test_subdomain_bytes_count_a_stored_dq_for_tables_that_need_one():
    """A table with A[0] != 0 cannot rebuild dq from the two states, so FP32
    and FP64 hold a dq as well (checked without a device)."""
    K, N = 1000, 5
    ndof = 9 * K * (N + 1) ** 3
    t = bx.Lsrk()
    t.n_stages = 2
    t.A[0], t.A[1] = 0.25, -0.5
    t.B[0], t.B[1] = 1.0, 0.5
    t.C[0], t.C[1], t.C[2] = 0.0, 1.0, 1.0
    kc = bx.Lsrk()
    bx.lib().bfamx_lsrk_init(C.byref(kc), bx.LSRK_KC54)
    rnd = lambda x: (x + 255) & ~255
    for prec, word in ((bx.FP64, 8), (bx.FP32, 4)):
        base = bx.subdomain_bytes(3, N, K, bx.ELASTIC3D, prec)
        assert bx.subdomain_bytes(3, N, K, bx.ELASTIC3D, prec, lsrk=kc) == base
        # a state or dq buffer holds 16 bytes past its values
        assert bx.subdomain_bytes(3, N, K, bx.ELASTIC3D, prec, lsrk=t) == base + rnd(word * ndof + 16)
    b = bx.subdomain_bytes(3, N, K, bx.ELASTIC3D, bx.FP32_RHS)
    assert bx.subdomain_bytes(3, N, K, bx.ELASTIC3D, bx.FP32_RHS, lsrk=t) == b
