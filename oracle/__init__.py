"""CPU oracle for the Born forward / RTM adjoint path -- TEST INFRASTRUCTURE.

ctypes bindings over oracle/liboracle.so (a C restatement of the reference's
base/cpu_prop.cc and of the preparation code around it) and over
oracle/_ref/libsinc_ref.so (the reference's own base/sinc_bob.cc, compiled in
place).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import this package, and only as the checker: nothing in the shipped
library calls it.

The functions whose parity is not pinned by a closed form, a brute force or
the compiled reference are listed in the header of oracle/born_oracle.c; the
random band of rand_subspace is "parity unpinned" (the reference draws rand()
inside a parallel loop, vel_fd_3d.cc:50-122).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcProp(C.Structure):
    """Mirror of orc_prop (oracle/born_oracle.h), i.e. of cpuProp's state
    (base/cpu_prop.h:36-53)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("coeffs", C.c_float * 13),
        ("nbound", C.c_int),
        ("bound", C.c_float * 256),
        ("npts_s", C.c_int), ("nt_src", C.c_int),
        ("locs_s", C.c_void_p), ("src", C.c_void_p),
        ("nsinc_s", C.c_int), ("jts", C.c_int), ("tab_s", C.c_void_p),
        ("nrec", C.c_int), ("nt_rec", C.c_int),
        ("locs_r", C.c_void_p), ("rec", C.c_void_p),
        ("nsinc_d", C.c_int), ("jtd", C.c_int), ("tab_d", C.c_void_p),
        ("vel1", C.c_void_p), ("vel2", C.c_void_p),
        ("jt", C.c_int), ("dir", C.c_int), ("nthreads", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = _bind(C.CDLL(path))
    return _LIB


_VARIANTS = {}


def variant_lib(name):
    """The same C source built for speed on the host that loads it
    (oracle/Makefile target `native`: "native" = -O3 -march=native, "ofast" =
    -Ofast -march=native, the reference's CMakeLists.txt:10 flags).  Timing
    only: FP contraction changes the rounding, so parity checks use lib()."""
    if name not in _VARIANTS:
        path = os.path.join(_HERE, "_native", "liboracle_%s.so" % name)
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (make -C oracle native)" % path)
        _VARIANTS[name] = _bind(C.CDLL(path))
    return _VARIANTS[name]


def _bind(L):
    """Set the ctypes signatures of every exported oracle function."""
    P = C.POINTER(OrcProp)
    sig = {
        "orc_create_space": (None, [P, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]),
        "orc_set_sponge": (None, [P, C.c_int, C.c_float]),
        "orc_prop_step": (None, [P, f32p, f32p, f32p]),
        "orc_damp": (None, [P, f32p, f32p]),
        "orc_inject_source": (None, [P, C.c_int, C.c_int, f32p]),
        "orc_inject_receivers": (None, [P, C.c_int, C.c_int, f32p]),
        "orc_data_extract": (None, [P, C.c_int, C.c_int, f32p]),
        "orc_image_add": (None, [P, f32p, f32p, f32p]),
        "orc_image_condition": (None, [P, f32p, f32p, f32p]),
        "orc_rtm_forward": (None, [P, C.c_int, f32p, C.c_int]),
        "orc_source_prop": (None, [P, f32p, f32p, C.c_int, C.c_int]),
        "orc_rtm_adjoint": (None, [P, f32p, f32p, f32p, C.c_int]),
        "orc_born_forward_d": (None, [P, C.c_int, C.c_int, f32p, C.c_int, f64p]),
        "orc_exact_adjoint_d": (None, [P, C.c_int, C.c_int, f64p, C.c_int, f64p]),
        "orc_exact_adjoint_f": (None, [P, C.c_int, C.c_int, f32p, C.c_int, f32p]),
        "orc_nonlinear_d": (None, [P, C.c_int, f64p, C.c_int, f64p]),
        "orc_sinc_table": (None, [C.c_int, C.c_int, f32p]),
        "orc_calc_stability": (None, [C.c_float, C.c_float, f32p, C.c_int, C.c_int,
                                      C.POINTER(C.c_float), C.POINTER(C.c_int)]),
        "orc_padded_size": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "orc_wavelet_deriv": (None, [C.c_int, C.c_float, f32p, f32p]),
        "orc_source_func": (None, [C.c_int, C.c_int] + [C.c_float] * 9 +
                            [f32p, C.c_int, C.c_int, i32p, f32p]),
        "orc_rand_subspace": (C.c_float, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                          C.c_int, C.c_int, C.c_int] + [C.c_float] * 6 +
                              [f32p, C.c_int, C.c_int, C.c_int] + [C.c_float] * 6 + [f32p]),
        "orc_receiver_locs": (None, [C.c_int, C.c_int] + [C.c_float] * 5 +
                              [C.c_int, C.c_int] + [C.c_float] * 6 + [i32p]),
        "orc_map_data_forward": (None, [C.c_int] * 4 + [f32p, f32p]),
        "orc_map_data_adjoint": (None, [C.c_int] * 4 + [f32p, f32p]),
        "orc_deriv_forward": (None, [C.c_int] * 3 + [C.c_float, f32p, f32p]),
        "orc_deriv_adjoint": (None, [C.c_int] * 3 + [C.c_float, f32p, f32p]),
        "orc_laplac_forward": (None, [C.c_int] * 3 + [f32p, f32p]),
        "orc_laplac_adjoint": (None, [C.c_int] * 3 + [f32p, f32p]),
        "orc_extract_sub": (None, [C.c_int] * 3 + [C.c_float] * 6 + [f32p] +
                            [C.c_int] * 3 + [C.c_float] * 3 + [f32p]),
        "orc_add_image": (None, [C.c_int] * 3 + [C.c_float] * 6 + [f32p] +
                          [C.c_int] * 3 + [C.c_float] * 3 + [f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def ref_lib():
    """The reference's base/sinc_bob.cc compiled into oracle/_ref (or None)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libsinc_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_sinc_table.restype = C.c_int
        L.ref_sinc_table.argtypes = [C.c_int, C.c_int, f32p]
        _REF = L
    return _REF


def ref_sinc_table(jt, nsinc=8, variant=""):
    path = os.path.join(_HERE, "_ref", "libsinc_ref%s.so" % variant)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_sinc_table.restype = C.c_int
    L.ref_sinc_table.argtypes = [C.c_int, C.c_int, f32p]
    out = np.zeros((jt, nsinc), np.float32)
    assert L.ref_sinc_table(jt, nsinc, out) == 0
    return out


def sinc_table(jt, nsinc=8):
    out = np.zeros((jt, nsinc), np.float32)
    lib().orc_sinc_table(jt, nsinc, out)
    return out


def default_threads():
    """OMP_NUM_THREADS when set, else the CPUs this process may run on"""
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env else len(os.sched_getaffinity(0))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Prop:
    """Oracle counterpart of the reference's cpuProp: same call order
    (createSpace, transferSincTableS/D, transferSourceFunc, transferVelFunc1/2,
    transferReceiverFunc, then rtmForward or sourceProp + rtmAdjoint).
    Arrays handed over are kept alive here, as the reference aliases them."""

    def __init__(self, nthreads=0, variant=None):
        self.s = OrcProp()
        self.s.nthreads = nthreads or default_threads()
        self._keep = {}
        self.L = lib() if variant is None else variant_lib(variant)

    @property
    def ptr(self):
        return C.byref(self.s)

    def _hold(self, key, arr, dtype):
        a = np.ascontiguousarray(arr, dtype=dtype)
        self._keep[key] = a
        return a.ctypes.data

    def create_space(self, d1, d2, d3, nx, ny, nz, bc_a=50.0, bc_b=0.0005, bc_y=0.0005):
        self.L.orc_create_space(self.ptr, d1, d2, d3, bc_a, bc_b, bc_y, nx, ny, nz)

    def set_sponge(self, width, bcB=0.0005):
        self.L.orc_set_sponge(self.ptr, width, bcB)

    @property
    def n(self):
        return self.s.nx * self.s.ny * self.s.nz

    @property
    def shape(self):
        return (self.s.nz, self.s.ny, self.s.nx)

    def transfer_sinc_table_s(self, table):
        t = np.asarray(table, np.float32)
        self.s.tab_s = self._hold("tab_s", t, np.float32)
        self.s.jts, self.s.nsinc_s = t.shape

    def transfer_sinc_table_d(self, table):
        t = np.asarray(table, np.float32)
        self.s.tab_d = self._hold("tab_d", t, np.float32)
        self.s.jtd, self.s.nsinc_d = t.shape

    def transfer_source_func(self, locs, vals):
        v = np.asarray(vals, np.float32)
        self.s.npts_s, self.s.nt_src = v.shape
        self.s.locs_s = self._hold("locs_s", locs, np.int32)
        self.s.src = self._hold("src", v, np.float32)

    def transfer_vel_func1(self, v):
        self.s.vel1 = self._hold("vel1", np.ravel(v), np.float32)

    def transfer_vel_func2(self, v):
        self.s.vel2 = self._hold("vel2", np.ravel(v), np.float32)

    def transfer_receiver_func(self, locs, rec):
        """rec is [nrec][nt_rec] float32; like the reference, the forward
        writes into this very array."""
        assert rec.dtype == np.float32 and rec.flags["C_CONTIGUOUS"]
        self.s.nrec, self.s.nt_rec = rec.shape
        self.s.locs_r = self._hold("locs_r", locs, np.int32)
        self._keep["rec"] = rec
        self.s.rec = rec.ctypes.data

    def rtm_forward(self, jt, img, nt):
        self.L.orc_rtm_forward(self.ptr, jt, f32(np.ravel(img)), nt)

    def source_prop(self, p0, p1, jts, nt):
        self.L.orc_source_prop(self.ptr, p0, p1, jts, nt)

    def rtm_adjoint(self, p0, p1, img, nt):
        self.L.orc_rtm_adjoint(self.ptr, p0, p1, img, nt)

    def prop_step(self, p0, p1, vel):
        self.L.orc_prop_step(self.ptr, p0, p1, vel)

    def damp(self, p0, p1):
        self.L.orc_damp(self.ptr, p0, p1)

    def born_forward_d(self, jt, img, nt, scatter=0):
        out = np.zeros(self.s.nrec * self.s.nt_rec, np.float64)
        self.L.orc_born_forward_d(self.ptr, jt, scatter, f32(np.ravel(img)), nt, out)
        return out.reshape(self.s.nrec, self.s.nt_rec)

    def exact_adjoint_d(self, jt, rec, nt, scatter=0):
        out = np.zeros(self.n, np.float64)
        self.L.orc_exact_adjoint_d(self.ptr, jt, scatter,
                                  np.ascontiguousarray(rec, np.float64).ravel(), nt, out)
        return out.reshape(self.shape)

    def exact_adjoint_f(self, jt, rec, nt, scatter=0):
        out = np.zeros(self.n, np.float32)
        self.L.orc_exact_adjoint_f(self.ptr, jt, scatter, f32(rec).ravel(), nt, out)
        return out.reshape(self.shape)

    def nonlinear_d(self, jt, vel2dt2, nt):
        out = np.zeros(self.s.nrec * self.s.nt_rec, np.float64)
        self.L.orc_nonlinear_d(self.ptr, jt, np.ascontiguousarray(np.ravel(vel2dt2), np.float64),
                              nt, out)
        return out.reshape(self.s.nrec, self.s.nt_rec)
