"""bfam_amd -- MI355X-native DG right-hand side and LSRK time stepping.

Python front end over the C ABI in include/bfamx.h (libbfamx.so, built
in-tree by ``python -m bfam_amd.build``).  It mirrors the pieces of bfam a
caller touches on the hot path: the bfam_subdomain_dgx data (bfam.h:1364-1420),
bfamo_lsrk_init (bfam.c:8996-9090) and the device layer (bfam.c:8389-9092).
There is no CPU fallback: every compute call goes to the HIP library, and a
missing library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbfamx.so")

ACOUSTIC2D, ELASTIC3D = 0, 1
BC_NONE, BC_FREE, BC_RIGID, BC_ABSORB = 0, 1, 2, 3
FP64, FP64_PLAIN, FP32_RHS, FP32, FP32_BF16X3 = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": FP64, "fp64_plain": FP64_PLAIN, "fp32_rhs": FP32_RHS, "fp32": FP32,
              "fp32_bf16x3": FP32_BF16X3}
LSRK_KC54, LSRK_FE, LSRK_HEUN, LSRK_W33, LSRK_NOOP = 0, 1, 2, 3, 4
OK, EINVAL, EBUDGET, EHIP, ESTATE, EUNSUPPORTED = 0, 1, 2, 3, 4, 5
FLAG_SAME, FLAG_COARSEN, FLAG_REFINE = 0, 1, 2

ELASTIC_FIELDS = ("v1", "v2", "v3", "s11", "s22", "s33", "s12", "s13", "s23")
ACOUSTIC_FIELDS = ("p", "u", "v")

_lib = None


class BfamxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bfamx error {code}: {msg}")
        self.code = code


class Lsrk(C.Structure):
    _fields_ = [("n_stages", C.c_int), ("A", C.c_double * 8), ("B", C.c_double * 8),
                ("C", C.c_double * 9)]


class Dgx(C.Structure):
    _fields_ = [("dim", C.c_int), ("N", C.c_int), ("Np", C.c_int), ("K", C.c_int32),
                ("r", C.c_void_p), ("w", C.c_void_p), ("wi", C.c_void_p), ("Dr", C.c_void_p),
                ("vmapM", C.c_void_p), ("vmapP", C.c_void_p)]


class PhysDesc(C.Structure):
    _fields_ = [("physics", C.c_int), ("precision", C.c_int),
                ("elem_h", C.c_void_p), ("elem_x0", C.c_void_p),
                ("rho", C.c_void_p), ("cp", C.c_void_p), ("cs", C.c_void_p)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int), ("N", C.c_int), ("nsteps", C.c_int), ("seed", C.c_uint64),
                ("arrival", C.c_double)]


class OctreeDesc(C.Structure):
    _fields_ = [("n_base", C.c_int), ("levels", C.c_int), ("center", C.c_double * 3),
                ("radius", C.c_double), ("order_by_level", C.c_int * 4)]


class MortarRec(C.Structure):
    _fields_ = [("s_sub", C.c_int32), ("s_elem", C.c_int32), ("b_sub", C.c_int32),
                ("b_elem", C.c_int32), ("s_face", C.c_int8), ("b_face", C.c_int8),
                ("quad", C.c_int8), ("order", C.c_int8)]


MORTAR_DTYPE = np.dtype([("s_sub", np.int32), ("s_elem", np.int32), ("b_sub", np.int32),
                         ("b_elem", np.int32), ("s_face", np.int8), ("b_face", np.int8),
                         ("quad", np.int8), ("order", np.int8)])


class TransferMaps(C.Structure):
    _fields_ = [("dst_to_adapt_flags", C.c_void_p), ("dst_to_dst_chld_id", C.c_void_p),
                ("dst_to_src_subd_id", C.c_void_p), ("dst_to_src_elem_id", C.c_void_p),
                ("num_dst", C.c_int32), ("coarse_dst_to_src_chld_id", C.c_void_p),
                ("coarse_dst_to_src_subd_id", C.c_void_p),
                ("coarse_dst_to_src_elem_id", C.c_void_p), ("num_coarse_dst", C.c_int32)]


# (name, numpy dtype, length key) of the transfer-map arrays
_MAP_ARRAYS = [("dst_to_adapt_flags", np.uint8, "num_dst"),
               ("dst_to_dst_chld_id", np.int8, "num_dst"),
               ("dst_to_src_subd_id", np.int32, "num_dst"),
               ("dst_to_src_elem_id", np.int32, "num_dst"),
               ("coarse_dst_to_src_chld_id", np.int8, "num_coarse_dst"),
               ("coarse_dst_to_src_subd_id", np.int32, "num_coarse_dst"),
               ("coarse_dst_to_src_elem_id", np.int32, "num_coarse_dst")]


class DgxGlue(C.Structure):
    _fields_ = [("sub_m", C.c_int32), ("sub_p", C.c_int32), ("partner", C.c_int32),
                ("K", C.c_int32), ("N", C.c_int32), ("EToEm", C.c_void_p), ("EToFm", C.c_void_p),
                ("EToHm", C.c_void_p), ("EToEp", C.c_void_p), ("EToHp", C.c_void_p),
                ("EToOp", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("arrival", C.c_double), ("admitted", C.c_double), ("finished", C.c_double),
                ("latency", C.c_double), ("dof", C.c_int64), ("energy", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("dof_stage_per_s", C.c_double), ("p50_s", C.c_double),
                ("p99_s", C.c_double), ("max_s", C.c_double), ("peak_bytes", C.c_size_t),
                ("max_active", C.c_int), ("nlaunch", C.c_int), ("nbatches", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_PP = C.POINTER(C.c_void_p)

# exported symbol -> (restype, argtypes); also the list the ABI tests check
SIGNATURES = {
    "bfamx_last_error": (C.c_char_p, []),
    "bfamx_set_abort_on_error": (None, [C.c_int]),
    "bfamx_lsrk_init": (C.c_int, [C.POINTER(Lsrk), C.c_int]),
    "bfamx_lgl": (C.c_int, [C.c_int, _dp, _dp]),
    "bfamx_diff_matrix": (C.c_int, [C.c_int, _dp, _dp]),
    "bfamx_lagrange_eval": (C.c_int, [C.c_int, _dp, C.c_double, _dp]),
    "bfamx_mortar_ops": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]),
    "bfamx_face_mask": (C.c_int, [C.c_int, C.c_int, _ip]),
    "bfamx_build_vmaps": (C.c_int, [C.c_int, C.c_int, C.c_int32, _ip, _i8p, _ip, _ip]),
    "bfamx_face_connectivity": (C.c_int, [C.c_int, C.c_int, C.c_int32, _ip, _ip, _ip, _u8p]),
    "bfamx_pcg32_seed": (None, [_vp, C.c_uint64, C.c_uint64]),
    "bfamx_pcg32_next": (C.c_uint32, [_vp]),
    "bfamx_pcg32_bounded": (C.c_uint32, [_vp, C.c_uint32]),
    "bfamx_pcg32_uniform": (C.c_double, [_vp]),
    "bfamx_brick": (C.c_int, [C.c_int, _ip, _ip, _dp, C.c_int, _ip, _i8p, _ip, _dp, _dp]),
    "bfamx_device_create": (C.c_int, [C.c_int, C.c_size_t, _PP]),
    "bfamx_device_destroy": (C.c_int, [_vp]),
    "bfamx_device_memory": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "bfamx_device_synchronize": (C.c_int, [_vp]),
    "bfamx_device_stream": (C.c_void_p, [_vp]),
    "bfamx_device_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bfamx_subdomain_upload": (C.c_int, [_vp, C.POINTER(Dgx), C.POINTER(PhysDesc),
                                         C.POINTER(Lsrk), _PP]),
    "bfamx_subdomain_free": (C.c_int, [_vp]),
    "bfamx_subdomain_info": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "bfamx_subdomain_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int32, C.c_int, C.c_int]),
    "bfamx_subdomain_bytes_lsrk": (C.c_size_t, [C.c_int, C.c_int, C.c_int32, C.c_int, C.c_int,
                                                C.POINTER(Lsrk)]),
    "bfamx_set_boundary": (C.c_int, [_vp, _i8p]),
    "bfamx_set_source": (C.c_int, [_vp, C.c_int32, _dp, _dp, C.c_double, C.c_double]),
    "bfamx_set_forcing_mms": (C.c_int, [_vp, C.c_double, C.c_double, _dp]),
    "bfamx_set_fields": (C.c_int, [_vp, _PP, C.c_double, C.c_double]),
    "bfamx_download_fields": (C.c_int, [_vp, _PP]),
    "bfamx_get_time": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "bfamx_rhs": (C.c_int, [_vp, C.c_double, _PP]),
    "bfamx_rhs_noalloc": (C.c_int, [_vp, C.c_double, _PP]),
    "bfamx_stage": (C.c_int, [_vp]),
    "bfamx_step": (C.c_int, [_vp, C.c_int]),
    "bfamx_reduce_energy": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "bfamx_reduce_cfl_dt": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_double)]),
    "bfamx_reduce_l2_diff": (C.c_int, [_vp, _vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bfamx_reduce_mms_error": (C.c_int, [_vp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bfamx_reduce_mms_energy_error": (C.c_int, [_vp, _dp, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]),
    "bfamx_batch_create": (C.c_int, [_vp, _PP, C.c_int, _PP]),
    "bfamx_batch_free": (C.c_int, [_vp]),
    "bfamx_batch_step": (C.c_int, [_vp, C.c_int]),
    "bfamx_batch_capture": (C.c_int, [_vp, C.c_int]),
    "bfamx_batch_launch_graph": (C.c_int, [_vp]),
    "bfamx_batch_autotune": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "bfamx_batch_set_strategy": (C.c_int, [_vp, C.c_int]),
    "bfamx_set_kernel": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "bfamx_get_kernel": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "bfamx_set_occupancy": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bfamx_get_occupancy": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "bfamx_timer_start": (C.c_int, [_vp]),
    "bfamx_timer_stop": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "bfamx_set_tuning": (C.c_int, [_vp, C.c_int]),
    "bfamx_set_persistent": (C.c_int, [_vp, C.c_int]),
    "bfamx_octree_build": (C.c_int, [C.POINTER(OctreeDesc), _PP]),
    "bfamx_octree_free": (C.c_int, [_vp]),
    "bfamx_octree_info": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int32), _ip]),
    "bfamx_octree_sub": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int32),
                                   _vp, _vp, _vp, _vp, _vp, _vp]),
    "bfamx_octree_mortars": (C.c_int, [_vp, _vp]),
    "bfamx_octree_elements": (C.c_int, [_vp, _vp, _vp]),
    "bfamx_transfer_maps_init": (C.c_int, [C.POINTER(TransferMaps), _vp, _vp]),
    "bfamx_transfer_maps_free": (C.c_int, [C.POINTER(TransferMaps)]),
    "bfamx_transfer_fields": (C.c_int, [C.POINTER(TransferMaps), _PP, C.c_int, _PP, C.c_int,
                                        _ip, _ip, C.c_double]),
    "bfamx_glue_upload": (C.c_int, [_vp, _PP, C.c_int, _vp, C.c_int32, _PP]),
    "bfamx_glue_free": (C.c_int, [_vp]),
    "bfamx_batch_set_glue": (C.c_int, [_vp, _vp]),
    "bfamx_mortars_from_glue": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(C.c_int32)]),
    "bfamx_batch_rhs": (C.c_int, [_vp, C.c_double, _vp]),
    "bfamx_vtk_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int32, _dp, _dp, _dp,
                                  C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_int, _vp]),
    "bfamx_vtk_write_index": (C.c_int, [C.c_char_p, C.c_int, _vp, C.c_int, _vp, C.c_int]),
    "bfamx_serve_generate": (C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(Request)]),
    "bfamx_request_bytes": (C.c_size_t, [C.POINTER(Request), C.c_int]),
    "bfamx_request_build": (C.c_int, [_vp, C.POINTER(Request), C.c_int, _PP,
                                      C.POINTER(C.c_double)]),
    "bfamx_request_inputs": (C.c_int, [C.POINTER(Request), _vp, _vp, _vp, C.POINTER(C.c_int32),
                                       _vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    "bfamx_serve_cb": (C.c_int, [_vp, C.POINTER(Request), C.c_int, C.c_int, C.c_size_t, C.c_int,
                                 C.c_void_p, C.c_void_p, C.POINTER(RequestResult),
                                 C.POINTER(ServeStats)]),
    "bfamx_serve": (C.c_int, [_vp, C.POINTER(Request), C.c_int, C.c_int, C.c_size_t, C.c_int,
                              C.POINTER(RequestResult), C.POINTER(ServeStats)]),
}


def lib():
    """The loaded libbfamx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m bfam_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        msg = lib().bfamx_last_error()
        raise BfamxError(rc, msg.decode() if msg else "")


def _ptrs(arrays):
    return (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


# ---------------------------------------------------------------- host setup

def lsrk_init(method=LSRK_KC54):
    t = Lsrk()
    _check(lib().bfamx_lsrk_init(C.byref(t), method))
    n = t.n_stages
    return np.array(t.A[:n]), np.array(t.B[:n]), np.array(t.C[:n + 1])


def lgl(N):
    r, w = np.zeros(N + 1), np.zeros(N + 1)
    _check(lib().bfamx_lgl(N, r, w))
    return r, w


def diff_matrix(N, r):
    D = np.zeros((N + 1) ** 2)
    _check(lib().bfamx_diff_matrix(N, np.ascontiguousarray(r, dtype=np.float64), D))
    return D.reshape(N + 1, N + 1).T.copy()  # D[i, j] = l_j'(r_i)


def lagrange_eval(N, r, x):
    out = np.zeros(N + 1)
    _check(lib().bfamx_lagrange_eval(N, np.ascontiguousarray(r, dtype=np.float64), float(x), out))
    return out


def mortar_ops(Na, Nb):
    na, nb = Na + 1, Nb + 1
    P, MP, WMP = np.zeros(5 * na * nb), np.zeros(5 * na * nb), np.zeros(5 * na * nb)
    ident = C.c_int(0)
    _check(lib().bfamx_mortar_ops(Na, Nb, P, MP, WMP, C.byref(ident)))

    def blocks(a):
        return [a[k * na * nb:(k + 1) * na * nb].reshape(na, nb).T.copy() for k in range(5)]

    return {"prj": blocks(P), "mass_prj": blocks(MP), "wi_mass_prj": blocks(WMP),
            "prj0_identity": bool(ident.value)}


def face_mask(dim, N):
    nfp = (N + 1) ** (dim - 1)
    g = np.zeros(2 * dim * nfp, dtype=np.int32)
    _check(lib().bfamx_face_mask(dim, N, g))
    return g.reshape(2 * dim, nfp)


def build_vmaps(dim, N, EToE, EToF):
    EToE = np.ascontiguousarray(EToE, dtype=np.int32).ravel()
    EToF = np.ascontiguousarray(EToF, dtype=np.int8).ravel()
    K = EToE.size // (2 * dim)
    nfp = (N + 1) ** (dim - 1)
    vM = np.zeros(K * 2 * dim * nfp, dtype=np.int32)
    vP = np.zeros_like(vM)
    _check(lib().bfamx_build_vmaps(dim, N, K, EToE, EToF, vM, vP))
    return vM, vP


def face_connectivity(dim, N, vmapM, vmapP):
    vmapM = np.ascontiguousarray(vmapM, dtype=np.int32)
    vmapP = np.ascontiguousarray(vmapP, dtype=np.int32)
    nfp = (N + 1) ** (dim - 1)
    K = vmapM.size // (2 * dim * nfp)
    ne = np.zeros(K * 2 * dim, dtype=np.int32)
    nc = np.zeros(K * 2 * dim, dtype=np.uint8)
    _check(lib().bfamx_face_connectivity(dim, N, K, vmapM, vmapP, ne, nc))
    return ne, nc


def brick(dim, n, periodic, L=None, morton=True):
    """Uniform brick: returns dict(EToE, EToF, lex, x0, h, K)."""
    n3 = np.ascontiguousarray(list(n) + [1] * (3 - len(n)), dtype=np.int32)
    per = np.ascontiguousarray(list(periodic) + [0] * (3 - len(periodic)), dtype=np.int32)
    L = np.ascontiguousarray(list(L if L is not None else [1.0] * dim) + [1.0] * (3 - dim),
                             dtype=np.float64)
    K = int(np.prod(n3[:dim]))
    EToE = np.zeros(K * 2 * dim, dtype=np.int32)
    EToF = np.zeros(K * 2 * dim, dtype=np.int8)
    lex = np.zeros(K, dtype=np.int32)
    x0 = np.zeros(K * dim)
    h = np.zeros(K * dim)
    _check(lib().bfamx_brick(dim, n3, per, L, 1 if morton else 0, EToE, EToF, lex, x0, h))
    return {"EToE": EToE, "EToF": EToF, "lex": lex, "x0": x0.reshape(K, dim),
            "h": h.reshape(K, dim), "K": K, "n": list(n[:dim]), "L": list(L[:dim])}


class PCG32:
    """bfam_pcg32 (O'Neill's PCG32) through the library."""

    def __init__(self, initstate, initseq):
        self._st = (C.c_uint64 * 2)()
        lib().bfamx_pcg32_seed(self._st, initstate, initseq)

    def next(self):
        return lib().bfamx_pcg32_next(self._st)

    def bounded(self, bound):
        return lib().bfamx_pcg32_bounded(self._st, bound)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * lib().bfamx_pcg32_uniform(self._st)


# -------------------------------------------------------------------- device

class Device:
    def __init__(self, hip_device=0, budget_bytes=0):
        h = C.c_void_p()
        _check(lib().bfamx_device_create(hip_device, budget_bytes, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().bfamx_device_destroy(self.h)
            self.h = None

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        _check(lib().bfamx_device_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def info(self):
        cu, xcd = C.c_int(), C.c_int()
        _check(lib().bfamx_device_info(self.h, C.byref(cu), C.byref(xcd)))
        return cu.value, xcd.value

    def sync(self):
        _check(lib().bfamx_device_synchronize(self.h))

    def timer_start(self):
        _check(lib().bfamx_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_float()
        _check(lib().bfamx_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def set_tuning(self, xcd_remap=2):
        _check(lib().bfamx_set_tuning(self.h, int(xcd_remap)))

    def set_kernel(self, precision, N, kind):
        """elastic stage kernel for (precision, N): 0 line, 1 node, 2 node on a
        resident grid, 3 and 4 as 1 and 2 with the node kernel's other
        elements-per-block (N = 3..5), -1 unknown (the per-precision default
        until autotuned)"""
        _check(lib().bfamx_set_kernel(self.h, int(precision), int(N), int(kind)))

    def get_kernel(self, precision, N):
        """the elastic stage kernel that runs for (precision, N): 0 line, 1 node,
        2 node on a resident grid, 3 and 4 the same with the other block size"""
        k = C.c_int()
        _check(lib().bfamx_get_kernel(self.h, int(precision), int(N), C.byref(k)))
        return k.value

    def set_occupancy(self, precision, N, kind, level):
        """resident-block cap of stage kernel `kind` (0 line, 1 node, 2 node with
        its other block size) for (precision, N): 0 none, 1 three quarters,
        2 half"""
        _check(lib().bfamx_set_occupancy(self.h, int(precision), int(N), int(kind), int(level)))

    def get_occupancy(self, precision, N, kind):
        lv = C.c_int()
        _check(lib().bfamx_get_occupancy(self.h, int(precision), int(N), int(kind),
                                         C.byref(lv)))
        return lv.value

    def set_persistent(self, on=1):
        _check(lib().bfamx_set_persistent(self.h, int(on)))


def subdomain_bytes(dim, N, K, physics, precision, lsrk=None):
    """device bytes of an upload; lsrk: an Lsrk table (None: KC54)"""
    if lsrk is None:
        return lib().bfamx_subdomain_bytes(dim, N, K, physics, precision)
    return lib().bfamx_subdomain_bytes_lsrk(dim, N, K, physics, precision, C.byref(lsrk))


class Subdomain:
    """A volume subdomain on the device (bfamx_subdomain_upload)."""

    def __init__(self, dev, physics, precision, N, r, w, Dr, vmapM, vmapP, h, rho, cp,
                 cs=None, x0=None, lsrk_method=LSRK_KC54):
        dim = 2 if physics == ACOUSTIC2D else 3
        self.dev, self.physics, self.precision, self.N, self.dim = dev, physics, precision, N, dim
        self.nf = 3 if physics == ACOUSTIC2D else 9
        self.Np = (N + 1) ** dim
        keep = {}

        def arr(name, a, dt=np.float64):
            keep[name] = np.ascontiguousarray(a, dtype=dt).ravel()
            return keep[name].ctypes.data

        wi = 1.0 / np.asarray(w, dtype=np.float64)
        D = np.asarray(Dr, dtype=np.float64)
        vM = np.ascontiguousarray(vmapM, dtype=np.int32)
        self.K = vM.size // (2 * dim * (N + 1) ** (dim - 1))
        g = Dgx(dim, N, self.Np, self.K, arr("r", r), arr("w", w), arr("wi", wi),
                arr("Dr", D.T), arr("vM", vM, np.int32), arr("vP", vmapP, np.int32))
        p = PhysDesc(physics, precision, arr("h", h), arr("x0", x0) if x0 is not None else None,
                     arr("rho", rho), arr("cp", cp), arr("cs", cs) if cs is not None else None)
        if isinstance(lsrk_method, Lsrk):  # a custom table
            t = lsrk_method
        else:
            t = Lsrk()
            _check(lib().bfamx_lsrk_init(C.byref(t), lsrk_method))
        hh = C.c_void_p()
        _check(lib().bfamx_subdomain_upload(dev.h, C.byref(g), C.byref(p), C.byref(t),
                                            C.byref(hh)))
        self.h = hh
        self.nstages = t.n_stages

    @classmethod
    def _adopt(cls, dev, h, physics, precision, N, K, nstages=5):
        """wrap a handle the library created (bfamx_request_build)"""
        self = cls.__new__(cls)
        self.dev, self.physics, self.precision, self.N = dev, physics, precision, N
        self.dim = 2 if physics == ACOUSTIC2D else 3
        self.nf = 3 if physics == ACOUSTIC2D else 9
        self.Np = (N + 1) ** self.dim
        self.K, self.h, self.nstages = K, h, nstages
        return self

    def close(self):
        if self.h:
            lib().bfamx_subdomain_free(self.h)
            self.h = None

    def info(self):
        nf, b = C.c_int(), C.c_size_t()
        _check(lib().bfamx_subdomain_info(self.h, C.byref(nf), C.byref(b)))
        return nf.value, b.value

    def set_boundary(self, kind):
        """kind: (K, 2*dim) int8: 0 none, 1 free surface, 2 rigid, 3 absorbing"""
        k = np.ascontiguousarray(kind, dtype=np.int8).ravel()
        _check(lib().bfamx_set_boundary(self.h, k))

    def set_source(self, elem, weights, amp, f0, t0):
        a = np.zeros(9)
        a[:len(amp)] = amp
        _check(lib().bfamx_set_source(self.h, int(elem),
                                      np.ascontiguousarray(weights, dtype=np.float64), a,
                                      float(f0), float(t0)))

    def set_forcing_mms(self, k, omega, F):
        _check(lib().bfamx_set_forcing_mms(self.h, float(k), float(omega),
                                           np.ascontiguousarray(F, dtype=np.float64)))

    def set_fields(self, q, t0=0.0, dt=0.0):
        """q: (nfields, K*Np) in bfam's field layout."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.shape == (self.nf, self.K * self.Np), q.shape
        rows = [q[f] for f in range(self.nf)]
        _check(lib().bfamx_set_fields(self.h, _ptrs(rows), float(t0), float(dt)))

    def download(self):
        q = np.zeros((self.nf, self.K * self.Np))
        rows = [q[f] for f in range(self.nf)]
        _check(lib().bfamx_download_fields(self.h, _ptrs(rows)))
        return q

    def rhs(self, t=0.0):
        out = np.zeros((self.nf, self.K * self.Np))
        rows = [out[f] for f in range(self.nf)]
        _check(lib().bfamx_rhs(self.h, float(t), _ptrs(rows)))
        return out

    def rhs_noalloc(self, t=0.0):
        """bfamx_rhs_noalloc: the RHS staged in the library's fixed device
        array instead of a buffer allocated on first use"""
        out = np.zeros((self.nf, self.K * self.Np))
        rows = [out[f] for f in range(self.nf)]
        _check(lib().bfamx_rhs_noalloc(self.h, float(t), _ptrs(rows)))
        return out

    def stage(self):
        _check(lib().bfamx_stage(self.h))

    def step(self, n=1):
        _check(lib().bfamx_step(self.h, int(n)))

    def time(self):
        t, s = C.c_double(), C.c_int64()
        _check(lib().bfamx_get_time(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def energy(self):
        e = C.c_double()
        _check(lib().bfamx_reduce_energy(self.h, C.byref(e)))
        return e.value

    def cfl_dt(self, cfl=0.5):
        dt = C.c_double()
        _check(lib().bfamx_reduce_cfl_dt(self.h, float(cfl), C.byref(dt)))
        return dt.value

    def l2_diff(self, other):
        """(sum Jw|self - other|^2, sum Jw|other|^2)"""
        d, n = C.c_double(), C.c_double()
        _check(lib().bfamx_reduce_l2_diff(self.h, other.h, C.byref(d), C.byref(n)))
        return d.value, n.value

    def mms_energy_error(self, A):
        """(energy norm of q - exact, energy norm of exact)"""
        d, n = C.c_double(), C.c_double()
        _check(lib().bfamx_reduce_mms_energy_error(
            self.h, np.ascontiguousarray(A, dtype=np.float64), C.byref(d), C.byref(n)))
        return d.value, n.value

    def mms_error(self, A):
        d, n = C.c_double(), C.c_double()
        _check(lib().bfamx_reduce_mms_error(self.h, np.ascontiguousarray(A, dtype=np.float64),
                                            C.byref(d), C.byref(n)))
        return d.value, n.value


class Batch:
    """Many subdomains advanced by one launch per stage (bfamx_batch_*)."""

    def __init__(self, dev, subs):
        self.dev, self.subs = dev, list(subs)
        arr = (C.c_void_p * len(subs))(*[s.h for s in subs])
        h = C.c_void_p()
        _check(lib().bfamx_batch_create(dev.h, arr, len(subs), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().bfamx_batch_free(self.h)
            self.h = None

    def step(self, n=1):
        _check(lib().bfamx_batch_step(self.h, int(n)))

    def capture(self, nsteps):
        _check(lib().bfamx_batch_capture(self.h, int(nsteps)))

    def launch_graph(self):
        _check(lib().bfamx_batch_launch_graph(self.h))

    def autotune(self, trials=3, force=False):
        """0: one mixed-order launch per stage, 1: one launch per order"""
        c = C.c_int()
        _check(lib().bfamx_batch_autotune(self.h, int(trials), int(force), C.byref(c)))
        return c.value

    def set_strategy(self, per_order):
        _check(lib().bfamx_batch_set_strategy(self.h, int(per_order)))

    def set_glue(self, glue):
        _check(lib().bfamx_batch_set_glue(self.h, glue.h if glue is not None else None))
        self.glue = glue

    def rhs(self, t=0.0):
        """L(q, t) of every member: list of (nfields, K*Np) arrays"""
        outs = [np.zeros((s.nf, s.K * s.Np)) for s in self.subs]
        rows = [_ptrs([o[f] for f in range(o.shape[0])]) for o in outs]
        arr = (C.c_void_p * len(rows))(*[C.cast(r, C.c_void_p) for r in rows])
        _check(lib().bfamx_batch_rhs(self.h, float(t), arr))
        return outs


class Glue:
    """Mortars between subdomains on the device (bfamx_glue_upload)."""

    def __init__(self, dev, subs, mortars):
        m = np.ascontiguousarray(mortars, dtype=MORTAR_DTYPE)
        self.subs = list(subs)
        arr = (C.c_void_p * len(self.subs))(*[s.h for s in self.subs])
        h = C.c_void_p()
        _check(lib().bfamx_glue_upload(dev.h, arr, len(self.subs),
                                       m.ctypes.data if len(m) else None, len(m), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _check(lib().bfamx_glue_free(self.h))
            self.h = None


def mortars_from_glue(glues):
    """Mortar records (MORTAR_DTYPE) from reference-format glue subdomains
    (bfamx_mortars_from_glue): glues is a list of dicts with sub_m, sub_p,
    optionally the glue order N, and the arrays EToEm, EToFm, EToHm, EToEp,
    EToHp, EToOp; each glue's partner (the pair reversed) is found in the
    list."""
    keep = []
    arr = (DgxGlue * max(len(glues), 1))()
    index = {(int(g["sub_m"]), int(g["sub_p"])): i for i, g in enumerate(glues)}
    for i, g in enumerate(glues):
        d = arr[i]
        d.sub_m, d.sub_p = int(g["sub_m"]), int(g["sub_p"])
        d.partner = index.get((d.sub_p, d.sub_m), -1)
        d.K = len(g["EToEm"])
        d.N = int(g.get("N", 0))
        for name, dt in (("EToEm", np.int32), ("EToFm", np.int8), ("EToHm", np.int8),
                         ("EToEp", np.int32), ("EToHp", np.int8), ("EToOp", np.int8)):
            a = np.ascontiguousarray(g[name], dtype=dt)
            keep.append(a)
            setattr(d, name, a.ctypes.data if a.size else None)
    n = C.c_int32()
    _check(lib().bfamx_mortars_from_glue(C.cast(arr, C.c_void_p), len(glues), None,
                                         C.byref(n)))
    m = np.zeros(n.value, MORTAR_DTYPE)
    if n.value:
        _check(lib().bfamx_mortars_from_glue(C.cast(arr, C.c_void_p), len(glues),
                                             m.ctypes.data, C.byref(n)))
    return m


# ------------------------------------------------------------------- output

def vtk_write(path, dim, N, x0, h, r, fields, time=0.0, binary=True, Np_write=0,
              mpirank=0, subdomain_id=0, root_id=0):
    """fields: dict name -> (K*(N+1)^dim,) array in bfam node order.
    binary: 0 ascii, 1 (True) base64 binary, 2 zlib-compressed binary;
    mpirank, subdomain_id, root_id: the Int32 cell arrays of those names."""
    ids = np.array([mpirank, subdomain_id, root_id], np.int32)
    x0 = np.ascontiguousarray(x0, dtype=np.float64).ravel()
    h = np.ascontiguousarray(h, dtype=np.float64).ravel()
    K = x0.size // dim
    names = list(fields)
    arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in names]
    cn = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    _check(lib().bfamx_vtk_write(str(path).encode(), dim, N, K, x0, h,
                                 np.ascontiguousarray(r, dtype=np.float64), len(names),
                                 C.cast(cn, C.c_void_p), _ptrs(arrs) if arrs else None,
                                 float(time), int(binary), int(Np_write), ids.ctypes.data))


def vtk_write_index(path, pieces, names, binary=True):
    """.pvtu index over .vtu files (paths as they appear in the index)
    written by vtk_write with fields `names` and the same `binary`."""
    pc = (C.c_char_p * max(len(pieces), 1))(*[str(p).encode() for p in pieces])
    cn = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    _check(lib().bfamx_vtk_write_index(str(path).encode(), len(pieces), C.cast(pc, C.c_void_p),
                                       len(names), C.cast(cn, C.c_void_p), int(binary)))


# ------------------------------------------------------------- adapted mesh

def _octree_handle(n_base, levels, center, radius, order_by_level):
    d = OctreeDesc(n_base, levels, (C.c_double * 3)(*center), radius,
                   (C.c_int * 4)(*(list(order_by_level) + [1] * (4 - len(order_by_level)))))
    h = C.c_void_p()
    _check(lib().bfamx_octree_build(C.byref(d), C.byref(h)))
    return h


def transfer_maps(dst, src):
    """Transfer maps (bfamx_transfer_maps_init) from the forest `src` to the
    forest `dst`, each given as the argument tuple of octree(): a dict of
    numpy arrays named as bfamx_transfer_maps_t's fields, plus the
    destination's per-element subdomain and local index ("dst_subd",
    "dst_elem", bfamx_octree_elements)."""
    hd = _octree_handle(*dst)
    hs = None
    try:
        hs = _octree_handle(*src)
        m = TransferMaps()
        _check(lib().bfamx_transfer_maps_init(C.byref(m), hd, hs))
        try:
            out = {"num_dst": m.num_dst, "num_coarse_dst": m.num_coarse_dst}
            for name, dt, nkey in _MAP_ARRAYS:
                n = out[nkey] * (8 if nkey == "num_coarse_dst" else 1)
                ptr = getattr(m, name)
                a = np.zeros(n, dt)
                if n and ptr:
                    C.memmove(a.ctypes.data, ptr, a.nbytes)
                out[name] = a
        finally:
            lib().bfamx_transfer_maps_free(C.byref(m))
        K = out["num_dst"]
        out["dst_subd"], out["dst_elem"] = np.zeros(K, np.int32), np.zeros(K, np.int32)
        _check(lib().bfamx_octree_elements(hd, out["dst_subd"].ctypes.data,
                                           out["dst_elem"].ctypes.data))
        return out
    finally:
        lib().bfamx_octree_free(hd)
        if hs is not None:
            lib().bfamx_octree_free(hs)


def transfer_fields(maps, src, dst, dt):
    """Move the state of the Subdomains `src` to the Subdomains `dst` through
    maps (a dict from transfer_maps()), on the device (bfamx_transfer_fields);
    the destinations start at the sources' time with step dt."""
    keep = {name: np.ascontiguousarray(maps[name], dtype=dt) for name, dt, _ in _MAP_ARRAYS}
    m = TransferMaps()
    for name, _, _ in _MAP_ARRAYS:
        setattr(m, name, keep[name].ctypes.data if keep[name].size else None)
    m.num_dst, m.num_coarse_dst = int(maps["num_dst"]), int(maps["num_coarse_dst"])
    sa = (C.c_void_p * len(src))(*[s.h for s in src])
    da = (C.c_void_p * len(dst))(*[s.h for s in dst])
    _check(lib().bfamx_transfer_fields(C.byref(m), sa, len(src), da, len(dst),
                                       np.ascontiguousarray(maps["dst_subd"], dtype=np.int32),
                                       np.ascontiguousarray(maps["dst_elem"], dtype=np.int32),
                                       float(dt)))


def octree(n_base, levels, center, radius, order_by_level):
    """Adapted periodic brick forest (bfamx_octree_build): returns a dict with
    per-subdomain data (N, K, EToE, EToF, glued, x0, h, gid), the mortar
    records (structured array of MORTAR_DTYPE) and leaf counts per level."""
    h = _octree_handle(n_base, levels, center, radius, order_by_level)
    try:
        K, ns, nm = C.c_int32(), C.c_int(), C.c_int32()
        kl = np.zeros(4, np.int32)
        _check(lib().bfamx_octree_info(h, C.byref(K), C.byref(ns), C.byref(nm), kl))
        subs = []
        for s in range(ns.value):
            N, Ks = C.c_int(), C.c_int32()
            _check(lib().bfamx_octree_sub(h, s, C.byref(N), C.byref(Ks), None, None, None,
                                          None, None, None))
            k = Ks.value
            a = {"N": N.value, "K": k, "EToE": np.zeros(6 * k, np.int32),
                 "EToF": np.zeros(6 * k, np.int8), "glued": np.zeros(6 * k, np.int8),
                 "x0": np.zeros(3 * k), "h": np.zeros(3 * k), "gid": np.zeros(k, np.int32)}
            _check(lib().bfamx_octree_sub(h, s, C.byref(N), C.byref(Ks),
                                          *[a[k].ctypes.data for k in
                                            ("EToE", "EToF", "glued", "x0", "h", "gid")]))
            a["x0"] = a["x0"].reshape(k, 3)
            a["h"] = a["h"].reshape(k, 3)
            subs.append(a)
        m = np.zeros(nm.value, MORTAR_DTYPE)
        if nm.value:
            _check(lib().bfamx_octree_mortars(h, m.ctypes.data))
        return {"K": K.value, "subs": subs, "mortars": m, "K_level": kl}
    finally:
        lib().bfamx_octree_free(h)


# ------------------------------------------------------------------ serving

def serve_generate(seed, nreq, rate_per_s, n=(8, 32), N=(3, 7), steps=(200, 2000)):
    """Seeded Poisson stream of elastic requests (bfamx_serve_generate)."""
    arr = (Request * max(nreq, 1))()
    _check(lib().bfamx_serve_generate(seed, nreq, rate_per_s, n[0], n[1], N[0], N[1],
                                      steps[0], steps[1], arr))
    return [arr[i] for i in range(nreq)]


def request_bytes(req, precision):
    return lib().bfamx_request_bytes(C.byref(req), precision)


def request_build(dev, req, precision):
    """The request's subdomain at rest at t = 0, and its dt."""
    h, dt = C.c_void_p(), C.c_double()
    _check(lib().bfamx_request_build(dev.h, C.byref(req), precision, C.byref(h), C.byref(dt)))
    return Subdomain._adopt(dev, h, ELASTIC3D, precision, req.N, req.n ** 3), dt.value


def request_inputs(req):
    """The request's material, source and dt as bfamx_request_build sets
    them up (bfamx_request_inputs), without a device."""
    K, Np = req.n ** 3, (req.N + 1) ** 3
    rho, cp, cs, w = np.zeros(K), np.zeros(K), np.zeros(K), np.zeros(Np)
    e = C.c_int32()
    f0, t0, dt = C.c_double(), C.c_double(), C.c_double()
    _check(lib().bfamx_request_inputs(C.byref(req), rho.ctypes.data, cp.ctypes.data,
                                      cs.ctypes.data, C.byref(e), w.ctypes.data, C.byref(f0),
                                      C.byref(t0), C.byref(dt)))
    return {"rho": rho, "cp": cp, "cs": cs, "src_elem": e.value, "src_weights": w,
            "amp": [0, 0, 0, 1, 1, 1, 0, 0, 0], "f0": f0.value, "t0": t0.value, "dt": dt.value}


SERVE_DONE = C.CFUNCTYPE(None, C.c_int, C.c_void_p, C.c_void_p)


def serve(dev, reqs, precision=FP32_RHS, budget_bytes=32 << 30, max_chunk=10, on_done=None):
    """Run a request stream in real time; returns (per-request dicts, stats).
    on_done(index, subdomain), if given, sees each finished request's
    subdomain at its final state (bfamx_serve_cb; valid during the call)."""
    n = len(reqs)
    arr = (Request * max(n, 1))(*reqs)
    res = (RequestResult * max(n, 1))()
    st = ServeStats()
    cb = None
    if on_done is not None:
        def _done(i, h, _user):
            q = reqs[i]
            on_done(i, Subdomain._adopt(dev, C.c_void_p(h), ELASTIC3D, precision, q.N,
                                        q.n ** 3))
        cb = SERVE_DONE(_done)
    _check(lib().bfamx_serve_cb(dev.h, arr, n, precision, budget_bytes, max_chunk,
                                C.cast(cb, C.c_void_p) if cb else None, None, res,
                                C.byref(st)))
    out = [{k: getattr(res[i], k) for k, _ in RequestResult._fields_} for i in range(n)]
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    return out, stats
