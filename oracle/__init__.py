"""ctypes/numpy front end of the CPU FP64 oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The shipped
library (``bfam_amd``) never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

ACOUSTIC2D = 0
ELASTIC3D = 1
LSRK_KC54, LSRK_FE, LSRK_HEUN, LSRK_W33 = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [
        ("physics", C.c_int), ("N", C.c_int), ("K", C.c_int32),
        ("r", C.c_void_p), ("w", C.c_void_p), ("Dr", C.c_void_p),
        ("vmapM", C.c_void_p), ("vmapP", C.c_void_p),
        ("h", C.c_void_p), ("x0", C.c_void_p),
        ("rho", C.c_void_p), ("lam", C.c_void_p), ("mu", C.c_void_p),
        ("src_elem", C.c_int32), ("src_wt", C.c_void_p),
        ("src_amp", C.c_double * 9), ("src_f0", C.c_double), ("src_t0", C.c_double),
        ("mms", C.c_int), ("mms_k", C.c_double), ("mms_omega", C.c_double),
        ("mms_F", C.c_double * 9),
        ("glued", C.c_void_p),
        ("bc", C.c_void_p),
    ]


class Mortar(C.Structure):
    _fields_ = [("s_sub", C.c_int32), ("s_elem", C.c_int32), ("b_sub", C.c_int32),
                ("b_elem", C.c_int32), ("s_face", C.c_int8), ("b_face", C.c_int8),
                ("quad", C.c_int8), ("order", C.c_int8)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        L.orc_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.orc_grad_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.orc_jacobi_gauss.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_jacobi_gauss_lobatto.argtypes = [C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_vandermonde.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_interpolation.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_mass.argtypes = [C.c_int, _dp, _dp]
        L.orc_dgx_ops.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_interpolator.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_lsrk_table.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_gmask_faces.argtypes = [C.c_int, C.c_int, _ip]
        L.orc_buildmaps.argtypes = [C.c_int, C.c_int, C.c_int32, _ip, _i8p, _ip, _ip]
        L.orc_glue_mapOp.argtypes = [C.c_int, C.c_int, _ip]
        L.orc_p8est_orientation.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_pcg32_seed.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.orc_pcg32_next.argtypes = [C.c_void_p]
        L.orc_pcg32_next.restype = C.c_uint32
        L.orc_pcg32_bounded.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_pcg32_bounded.restype = C.c_uint32
        L.orc_brick.argtypes = [C.c_int, _ip, _ip, _ip, _i8p]
        L.orc_rhs.argtypes = [C.POINTER(Problem), C.c_double, _dp, _dp]
        L.orc_lsrk_run.argtypes = [C.POINTER(Problem), C.c_int, C.c_double, C.c_double,
                                   C.c_int, _dp, _dp, _dp]
        L.orc_energy.argtypes = [C.POINTER(Problem), _dp]
        L.orc_energy.restype = C.c_double
        L.orc_mms_exact.argtypes = [C.POINTER(Problem), _dp, C.c_double, _dp]
        L.orc_mms_forcing_coef.argtypes = [_dp, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, _dp]
        L.orc_ricker.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_ricker.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_transfer.argtypes = [C.c_int, C.c_int32] + [C.c_void_p] * 15
        L.orc_transfer.restype = C.c_int
        L.orc_set_num_threads.restype = None
        _lib = L
    return _lib


def jacobi_p(alpha, beta, N, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = np.empty_like(x)
    lib().orc_jacobi_p(alpha, beta, N, x.size, x, P)
    return P


def grad_jacobi_p(alpha, beta, N, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = np.empty_like(x)
    lib().orc_grad_jacobi_p(alpha, beta, N, x.size, x, P)
    return P


def gauss_lobatto(N, alpha=0.0, beta=0.0):
    r = np.zeros(N + 1)
    w = np.zeros(N + 1)
    if lib().orc_jacobi_gauss_lobatto(alpha, beta, N, r, w):
        raise RuntimeError("Newton did not converge")
    return r, w


def gauss(N, alpha=0.0, beta=0.0):
    """Gauss nodes of P_{N+1}^{(alpha,beta)} (N+1 points)."""
    r = np.zeros(N + 1)
    w = np.zeros(N + 1)
    if lib().orc_jacobi_gauss(alpha, beta, N, r, w):
        raise RuntimeError("Newton did not converge")
    return r, w


def vandermonde(N, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    V = np.zeros((N + 1) * x.size)
    lib().orc_vandermonde(N, x.size, x, V)
    return V.reshape(N + 1, x.size).T  # V[i, j] = p_j(x_i)


def mass(N):
    r, _ = gauss_lobatto(N)
    V = np.ascontiguousarray(vandermonde(N, r).T).ravel()
    M = np.zeros((N + 1) ** 2)
    lib().orc_mass(N, V, M)
    return M.reshape(N + 1, N + 1).T


def dgx_ops(N):
    """(r, w, wi, Dr) with Dr as an (N+1)x(N+1) matrix, Dr[i, j] = l_j'(r_i)."""
    n = N + 1
    r, w, wi, D = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n * n)
    if lib().orc_dgx_ops(N, r, w, wi, D):
        raise RuntimeError("operator construction failed")
    return r, w, wi, D.reshape(n, n).T.copy()


def interpolator(Na, Nb):
    """create_interpolators a->b: dict of prj/mass_prj/wi_mass_prj, 5 matrices
    of shape (Nb+1, Na+1) each; prj[0] is None when Na == Nb."""
    na, nb = Na + 1, Nb + 1
    P, MP, WMP = np.zeros(5 * na * nb), np.zeros(5 * na * nb), np.zeros(5 * na * nb)
    null = C.c_int(0)
    if lib().orc_interpolator(Na, Nb, P, MP, WMP, C.byref(null)):
        raise RuntimeError("interpolator construction failed")

    def blocks(a):
        return [a[k * na * nb:(k + 1) * na * nb].reshape(na, nb).T.copy() for k in range(5)]

    prj = blocks(P)
    if null.value:
        prj[0] = None
    return {"prj": prj, "mass_prj": blocks(MP), "wi_mass_prj": blocks(WMP)}


def lsrk_table(method=LSRK_KC54):
    A, B, Cc = np.zeros(6), np.zeros(6), np.zeros(6)
    ns = lib().orc_lsrk_table(method, A, B, Cc)
    return A[:ns].copy(), B[:ns].copy(), Cc[:ns + 1].copy()


def gmask_faces(dim, N):
    nfp = (N + 1) ** (dim - 1)
    g = np.zeros(2 * dim * nfp, dtype=np.int32)
    lib().orc_gmask_faces(dim, N, g)
    return g.reshape(2 * dim, nfp)


def buildmaps(dim, N, EToE, EToF):
    EToE = np.ascontiguousarray(EToE, dtype=np.int32).ravel()
    EToF = np.ascontiguousarray(EToF, dtype=np.int8).ravel()
    K = EToE.size // (2 * dim)
    nfp = (N + 1) ** (dim - 1)
    vM = np.zeros(K * 2 * dim * nfp, dtype=np.int32)
    vP = np.zeros_like(vM)
    lib().orc_buildmaps(dim, N, K, EToE, EToF, vM, vP)
    return vM, vP


def glue_mapOp(dim_glue, N):
    no = 2 if dim_glue == 1 else 8
    m = np.zeros(no * (N + 1) ** dim_glue, dtype=np.int32)
    lib().orc_glue_mapOp(dim_glue, N, m)
    return m.reshape(no, -1)


def p8est_orientation(f, nf, o):
    return lib().orc_p8est_orientation(f, nf, o)


class PCG32:
    def __init__(self, initstate, initseq):
        self._st = (C.c_uint64 * 2)()
        lib().orc_pcg32_seed(self._st, initstate, initseq)

    def next(self):
        return lib().orc_pcg32_next(self._st)

    def bounded(self, bound):
        return lib().orc_pcg32_bounded(self._st, bound)


def brick(dim, n, periodic):
    n = np.ascontiguousarray(list(n) + [1] * (3 - len(n)), dtype=np.int32)
    per = np.ascontiguousarray(list(periodic) + [0] * (3 - len(periodic)), dtype=np.int32)
    K = int(np.prod(n[:dim]))
    EToE = np.zeros(K * 2 * dim, dtype=np.int32)
    EToF = np.zeros(K * 2 * dim, dtype=np.int8)
    lib().orc_brick(dim, n, per, EToE, EToF)
    return EToE, EToF


def ricker(f0, t0, t):
    return lib().orc_ricker(f0, t0, t)


def mms_forcing_coef(A, k, omega, rho, lam, mu):
    F = np.zeros(9)
    lib().orc_mms_forcing_coef(np.ascontiguousarray(A, dtype=np.float64), k, omega,
                               rho, lam, mu, F)
    return F


def transfer(maps, src_qs, src_N, dst_K, dst_N, nf=9):
    """Fields moved to the destination subdomains through transfer maps
    (a dict named as bfamx_transfer_maps_t's fields, plus dst_subd and
    dst_elem): src_qs[s] is (nf, K_s * Np_s) in bfam layout; returns one
    (nf, K_d * Np_d) array per destination subdomain."""
    srcs = [np.ascontiguousarray(q, dtype=np.float64) for q in src_qs]
    dsts = [np.zeros((nf, K * (N + 1) ** 3)) for K, N in zip(dst_K, dst_N)]
    src_K = np.array([q.shape[1] // (N + 1) ** 3 for q, N in zip(srcs, src_N)], np.int32)
    sp = (C.c_void_p * len(srcs))(*[q.ctypes.data for q in srcs])
    dp = (C.c_void_p * len(dsts))(*[q.ctypes.data for q in dsts])
    a = {k: np.ascontiguousarray(maps[k], dtype=dt) for k, dt in (
        ("dst_to_adapt_flags", np.uint8), ("dst_to_dst_chld_id", np.int8),
        ("dst_to_src_subd_id", np.int32), ("dst_to_src_elem_id", np.int32),
        ("coarse_dst_to_src_chld_id", np.int8), ("coarse_dst_to_src_subd_id", np.int32),
        ("coarse_dst_to_src_elem_id", np.int32), ("dst_subd", np.int32),
        ("dst_elem", np.int32))}
    a["src_N"] = np.array(src_N, np.int32)
    a["src_K"] = src_K
    a["dst_N"] = np.array(dst_N, np.int32)
    a["dst_K"] = np.array(dst_K, np.int32)
    ptr = lambda x: x.ctypes.data  # the arrays stay referenced by `a`
    rc = lib().orc_transfer(nf, int(maps["num_dst"]), ptr(a["dst_to_adapt_flags"]),
                            ptr(a["dst_to_dst_chld_id"]), ptr(a["dst_to_src_subd_id"]),
                            ptr(a["dst_to_src_elem_id"]), ptr(a["coarse_dst_to_src_chld_id"]),
                            ptr(a["coarse_dst_to_src_subd_id"]),
                            ptr(a["coarse_dst_to_src_elem_id"]),
                            ptr(a["src_N"]), ptr(a["src_K"]), sp,
                            ptr(a["dst_subd"]), ptr(a["dst_elem"]),
                            ptr(a["dst_N"]), ptr(a["dst_K"]), dp)
    if rc:
        raise RuntimeError("oracle transfer failed")
    return dsts


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    """OpenMP threads of the oracle's parallel loops"""
    lib().orc_set_num_threads(int(n))


class OracleProblem:
    """Holds numpy arrays alive for the C orc_problem_t.

    Fields are laid out as in bfam (one K*Np array per field, stacked):
    q[f, e*Np + n].
    """

    def __init__(self, physics, N, vmapM, vmapP, h, rho, lam, mu=None, x0=None,
                 src=None, mms=None, ops=None, glued=None, bc=None):
        dim = 2 if physics == ACOUSTIC2D else 3
        self.dim, self.N, self.physics = dim, N, physics
        self.Np = (N + 1) ** dim
        self.nf = 3 if physics == ACOUSTIC2D else 9
        r, w, wi, D = ops if ops is not None else dgx_ops(N)
        self._r = np.ascontiguousarray(r, dtype=np.float64)
        self._w = np.ascontiguousarray(w, dtype=np.float64)
        self._D = np.ascontiguousarray(np.asarray(D).T.ravel(), dtype=np.float64)  # col-major
        self._vM = np.ascontiguousarray(vmapM, dtype=np.int32)
        self._vP = np.ascontiguousarray(vmapP, dtype=np.int32)
        self.K = self._vM.size // (2 * dim * (N + 1) ** (dim - 1))
        self._h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        self._x0 = np.ascontiguousarray(x0 if x0 is not None else np.zeros(self.K * dim),
                                        dtype=np.float64).ravel()
        self._rho = np.ascontiguousarray(rho, dtype=np.float64)
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._mu = np.ascontiguousarray(mu if mu is not None else np.zeros(self.K),
                                        dtype=np.float64)
        p = Problem()
        p.physics, p.N, p.K = physics, N, self.K
        p.r, p.w, p.Dr = self._r.ctypes.data, self._w.ctypes.data, self._D.ctypes.data
        p.vmapM, p.vmapP = self._vM.ctypes.data, self._vP.ctypes.data
        p.h, p.x0 = self._h.ctypes.data, self._x0.ctypes.data
        p.rho, p.lam, p.mu = self._rho.ctypes.data, self._lam.ctypes.data, self._mu.ctypes.data
        p.src_elem = -1
        if src is not None:
            self._swt = np.ascontiguousarray(src["weights"], dtype=np.float64)
            p.src_elem = int(src["elem"])
            p.src_wt = self._swt.ctypes.data
            for f, a in enumerate(src["amp"]):
                p.src_amp[f] = float(a)
            p.src_f0, p.src_t0 = float(src["f0"]), float(src["t0"])
        if mms is not None:
            p.mms = 1
            p.mms_k, p.mms_omega = float(mms["k"]), float(mms["omega"])
            for f, a in enumerate(mms["F"]):
                p.mms_F[f] = float(a)
        if glued is not None:
            self._glued = np.ascontiguousarray(glued, dtype=np.int8).ravel()
            p.glued = self._glued.ctypes.data
        if bc is not None:
            self._bc = np.ascontiguousarray(bc, dtype=np.int8).ravel()
            p.bc = self._bc.ctypes.data
        self.p = p

    def zeros(self):
        return np.zeros((self.nf, self.K * self.Np))

    def rhs(self, q, t=0.0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros_like(q)
        lib().orc_rhs(C.byref(self.p), t, q.ravel(), out.ravel())
        return out

    def run(self, q, dt, nsteps, t0=0.0, method=LSRK_KC54):
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        dq = np.zeros_like(q)
        work = np.zeros_like(q)
        lib().orc_lsrk_run(C.byref(self.p), method, t0, dt, nsteps, q.ravel(), dq.ravel(),
                           work.ravel())
        return q

    def energy(self, q):
        return lib().orc_energy(C.byref(self.p), np.ascontiguousarray(q, dtype=np.float64).ravel())

    def mms_exact(self, A, t):
        q = self.zeros()
        lib().orc_mms_exact(C.byref(self.p), np.ascontiguousarray(A, dtype=np.float64), t,
                            q.ravel())
        return q


class OracleMulti:
    """Elastic subdomains (OracleProblem, each built with its glued-face
    flags) coupled by mortars (records of bfamx_mortar_t layout)."""

    def __init__(self, problems, mortars):
        self.problems = list(problems)
        m = np.ascontiguousarray(mortars)
        self._m = (Mortar * max(len(m), 1))()
        for i, r in enumerate(m):
            order = int(r["order"]) if "order" in (r.dtype.names or ()) else 0
            self._m[i] = Mortar(*[int(r[k]) for k in ("s_sub", "s_elem", "b_sub", "b_elem",
                                                     "s_face", "b_face", "quad")], order)
        self.nm = len(m)
        self._subs = (C.POINTER(Problem) * len(self.problems))(
            *[C.pointer(P.p) for P in self.problems])

    @staticmethod
    def _rows(arrs):
        return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])

    def rhs(self, qs, t=0.0):
        qs = [np.ascontiguousarray(q, dtype=np.float64) for q in qs]
        out = [np.zeros_like(q) for q in qs]
        lib().orc_rhs_multi(self._subs, len(qs), self._m, self.nm, C.c_double(t),
                            self._rows(qs), self._rows(out))
        return out

    def run(self, qs, dt, nsteps, t0=0.0, method=LSRK_KC54):
        qs = [np.ascontiguousarray(q, dtype=np.float64).copy() for q in qs]
        dq = [np.zeros_like(q) for q in qs]
        work = [np.zeros_like(q) for q in qs]
        lib().orc_lsrk_run_multi(self._subs, len(qs), self._m, self.nm, method,
                                 C.c_double(t0), C.c_double(dt), nsteps, self._rows(qs),
                                 self._rows(dq), self._rows(work))
        return qs

    def energy(self, qs):
        return sum(P.energy(q) for P, q in zip(self.problems, qs))
