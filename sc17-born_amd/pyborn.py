"""ctypes binding of libborn_hip.so (include/born_hip.h).

`Propagator` mirrors the reference's baseProp interface (base/base_prop.h):
the same method names and argument meaning, host numpy arrays in and out,
errors raised as BornError where the reference would call exit(-1).  The
`dev_*` methods take device tensors (torch, ROCm) and queue work on the
library's stream.

There is no fallback: importing works anywhere, but constructing a
Propagator without the built library or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libborn_hip.so")

BORN_OK = 0
_ERRNAMES = {-1: "BORN_ERR_ARG", -2: "BORN_ERR_STATE", -3: "BORN_ERR_HIP",
             -4: "BORN_ERR_NOMEM", -5: "BORN_ERR_UNSUPPORTED"}

# every symbol declared in include/born_hip.h, with its ctypes signature
_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
SIGNATURES = {
    "born_create": (_i, [C.POINTER(_vp), _i]),
    "born_destroy": (_i, [_vp]),
    "born_last_error": (C.c_char_p, [_vp]),
    "born_create_space": (_i, [_vp, _f, _f, _f, _f, _f, _f, _i, _i, _i]),
    "born_set_ntblock": (_i, [_vp, _i]),
    "born_transfer_sinc_table_s": (_i, [_vp, _i, _i, _vp]),
    "born_transfer_sinc_table_d": (_i, [_vp, _i, _i, _vp]),
    "born_transfer_source_func": (_i, [_vp, _i, _i, _vp, _vp]),
    "born_transfer_vel_func1": (_i, [_vp, _i, _i, _i, _vp]),
    "born_transfer_vel_func2": (_i, [_vp, _i, _i, _i, _vp]),
    "born_transfer_receiver_func": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "born_rtm_forward": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i, _i, _i, _i]),
    "born_source_prop": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _i]),
    "born_rtm_adjoint": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _i]),
    "born_set_sponge": (_i, [_vp, _i, _f]),
    "born_adjoint_exact": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i]),
    "born_forward_scatter": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i]),
    "born_dev_forward": (_i, [_vp, _i, _i, _i, _vp, _vp, _i]),
    "born_dev_adjoint_exact": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "born_dev_source_prop": (_i, [_vp, _i, _i, _vp, _vp]),
    "born_dev_rtm_adjoint": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "born_set_checkpoint_budget": (_i, [_vp, C.c_size_t]),
    "born_last_checkpoint_plan": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "born_set_zchunk": (_i, [_vp, _i]),
    "born_get_stream": (_vp, [_vp]),
    "born_synchronize": (_i, [_vp]),
}

_LIB = None


class BornError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libborn_hip.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise BornError("%s is not built (run make -C sc17-born_amd or "
                            "__graft_entry__.build())" % path)
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _dptr(t):
    """device pointer of a torch tensor (must be contiguous)"""
    if not t.is_contiguous():
        raise BornError("device tensor must be contiguous")
    return C.c_void_p(t.data_ptr())


class Propagator:
    """HIP counterpart of the reference's gpuProp/cpuProp (base/base_prop.h)."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.born_create(C.byref(h), device)
        if rc != BORN_OK:
            raise BornError("born_create(%d) failed: %s" % (device, _ERRNAMES.get(rc, rc)))
        self.h = h
        self._keep = {}
        self.shape = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.born_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != BORN_OK:
            msg = self.lib.born_last_error(self.h)
            raise BornError("%s: %s: %s" % (what, _ERRNAMES.get(rc, rc),
                                            msg.decode() if msg else ""))

    # ---- baseProp interface -------------------------------------------
    def create_space(self, d1, d2, d3, nx, ny, nz, bc_a=50.0, bc_b=0.0005, bc_y=0.0005):
        self._chk(self.lib.born_create_space(self.h, d1, d2, d3, bc_a, bc_b, bc_y,
                                             nx, ny, nz), "createSpace")
        self.shape = (nz, ny, nx)

    def set_ntblock(self, nb):
        self._chk(self.lib.born_set_ntblock(self.h, nb), "setNtblock")

    def set_sponge(self, width, bcB=0.0005):
        self._chk(self.lib.born_set_sponge(self.h, width, bcB), "set_sponge")

    def transfer_sinc_table_s(self, table):
        t = np.ascontiguousarray(table, np.float32)
        self._chk(self.lib.born_transfer_sinc_table_s(self.h, t.shape[1], t.shape[0], _ptr(t)),
                  "transferSincTableS")

    def transfer_sinc_table_d(self, table):
        t = np.ascontiguousarray(table, np.float32)
        self._chk(self.lib.born_transfer_sinc_table_d(self.h, t.shape[1], t.shape[0], _ptr(t)),
                  "transferSincTableD")

    def transfer_source_func(self, locs, vals):
        l = np.ascontiguousarray(locs, np.int32)
        v = np.ascontiguousarray(vals, np.float32)
        self._chk(self.lib.born_transfer_source_func(self.h, v.shape[0], v.shape[1], _ptr(l),
                                                     _ptr(v)), "transferSourceFunc")

    def transfer_vel_func1(self, v):
        a = np.ascontiguousarray(v, np.float32)
        nz, ny, nx = self.shape
        self._chk(self.lib.born_transfer_vel_func1(self.h, nx, ny, nz, _ptr(a)), "transferVelFunc1")

    def transfer_vel_func2(self, v):
        a = np.ascontiguousarray(v, np.float32)
        nz, ny, nx = self.shape
        self._chk(self.lib.born_transfer_vel_func2(self.h, nx, ny, nz, _ptr(a)), "transferVelFunc2")

    def transfer_receiver_func(self, rec_nx, rec_ny, locs, traces):
        """traces: [rec_nx*rec_ny][nt_big] float32, kept and written in place
        by rtm_forward, read by rtm_adjoint (base/cpu_prop.cc:281-288)."""
        if traces.dtype != np.float32 or not traces.flags["C_CONTIGUOUS"]:
            raise BornError("traces must be a C-contiguous float32 array")
        l = np.ascontiguousarray(locs, np.int32)
        self._keep["traces"] = traces
        self._chk(self.lib.born_transfer_receiver_func(self.h, rec_nx, rec_ny, traces.shape[1],
                                                       _ptr(l), _ptr(traces)),
                  "transferReceiverFunc")

    def rtm_forward(self, jt, img, nt):
        nz, ny, nx = self.shape
        a = np.ascontiguousarray(img, np.float32)
        tr = self._keep["traces"]
        self._chk(self.lib.born_rtm_forward(self.h, nx, ny, nz, jt, _ptr(a), _ptr(tr), 9, nt,
                                            tr.shape[1], 1, tr.shape[0]), "rtmForward")

    def source_prop(self, p0, p1, jts, nt):
        nz, ny, nx = self.shape
        self._chk(self.lib.born_source_prop(self.h, nx, ny, nz, 0, 1, _ptr(p0), _ptr(p1), jts,
                                            9, nt), "sourceProp")

    def rtm_adjoint(self, p0, p1, img, nt, jtd=1):
        nz, ny, nx = self.shape
        self._chk(self.lib.born_rtm_adjoint(self.h, nx, ny, nz, jtd, _ptr(p0), _ptr(p1),
                                            _ptr(img), 9, nt), "rtmAdjoint")

    # ---- extensions -----------------------------------------------------
    def forward_scatter(self, jt, img, rec, nt, scatter):
        nz, ny, nx = self.shape
        a = np.ascontiguousarray(img, np.float32)
        self._chk(self.lib.born_forward_scatter(self.h, nx, ny, nz, jt, _ptr(a), _ptr(rec), nt,
                                                scatter), "forward_scatter")

    def adjoint_exact(self, jt, rec, img, nt, scatter=0):
        nz, ny, nx = self.shape
        r = np.ascontiguousarray(rec, np.float32)
        self._chk(self.lib.born_adjoint_exact(self.h, nx, ny, nz, jt, _ptr(r), _ptr(img), nt,
                                              scatter), "adjoint_exact")

    def set_checkpoint_budget(self, nbytes):
        self._chk(self.lib.born_set_checkpoint_budget(self.h, int(nbytes)), "checkpoint budget")

    def last_checkpoint_plan(self):
        s, r = C.c_int(), C.c_int()
        self._chk(self.lib.born_last_checkpoint_plan(self.h, C.byref(s), C.byref(r)), "plan")
        return s.value, r.value

    def set_zchunk(self, zc):
        self._chk(self.lib.born_set_zchunk(self.h, zc), "zchunk")

    @property
    def stream(self):
        return self.lib.born_get_stream(self.h)

    def synchronize(self):
        self._chk(self.lib.born_synchronize(self.h), "synchronize")

    # ---- device-resident (torch tensors on the context's device) ---------
    def dev_forward(self, jt, nt, model, traces, scatter=0, keep_background=False):
        self._chk(self.lib.born_dev_forward(self.h, jt, nt, scatter, _dptr(model),
                                            _dptr(traces), int(keep_background)), "dev_forward")

    def dev_adjoint_exact(self, jt, nt, traces, img, scatter=0):
        self._chk(self.lib.born_dev_adjoint_exact(self.h, jt, nt, scatter, _dptr(traces),
                                                  _dptr(img)), "dev_adjoint_exact")

    def dev_source_prop(self, jts, nt, p0, p1):
        self._chk(self.lib.born_dev_source_prop(self.h, jts, nt, _dptr(p0), _dptr(p1)),
                  "dev_source_prop")

    def dev_rtm_adjoint(self, nt, p0, p1, img, traces):
        self._chk(self.lib.born_dev_rtm_adjoint(self.h, nt, _dptr(p0), _dptr(p1), _dptr(img),
                                                _dptr(traces)), "dev_rtm_adjoint")
