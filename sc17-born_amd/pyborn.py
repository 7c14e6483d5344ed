"""ctypes binding of libborn_hip.so (include/born_hip.h).

`Propagator` mirrors the reference's baseProp interface (base/base_prop.h):
the same method names and argument meaning, host numpy arrays in and out,
errors raised as BornError where the reference would call exit(-1).  The
`dev_*` methods take device tensors (torch, ROCm) and queue work on the
library's stream.

There is no fallback: importing works anywhere, but constructing a
Propagator without the built library or without a GPU raises.

A process that also uses torch on the GPU must import torch before the
library is loaded: torch ships its own libamdhip64, and loading it second
would bring a second HIP runtime into the process, which then sees no
device.  Loaded after torch, the library binds to torch's runtime.
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
    "born_dev_source_prop_noalloc": (_i, [_vp, _i, _i, _vp, _vp]),
    "born_set_checkpoint_budget": (_i, [_vp, C.c_size_t]),
    "born_last_checkpoint_plan": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "born_last_background_sweep": (_i, [_vp]),
    "born_set_zchunk": (_i, [_vp, _i]),
    "born_set_storage": (_i, [_vp, _i]),
    "born_set_step_pairs": (_i, [_vp, _i]),
    "born_set_placement": (_i, [_vp, _i, C.c_longlong]),
    "born_last_placement": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "born_last_step_pairs": (_i, [_vp]),
    "born_set_autotune": (_i, [_vp, _i]),
    "born_get_tuning": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                             C.POINTER(C.c_float)]),
    "born_get_tile": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                           C.POINTER(_i)]),
    "born_set_tuning": (_i, [_vp, _i, _i, _i, _i, _i, _i]),
    "born_get_stream": (_vp, [_vp]),
    "born_batch_create": (_i, [_vp, _i, _vp, _f, _f, _f, C.POINTER(_vp)]),
    "born_batch_destroy": (_i, [_vp]),
    "born_batch_forward": (_i, [_vp, _i, _i, _i]),
    "born_batch_traces": (_i, [_vp, _i, _vp]),
    "born_batch_adjoint_exact": (_i, [_vp, _i, _i, _i]),
    "born_batch_set_traces": (_i, [_vp, _i, _vp]),
    "born_batch_image": (_i, [_vp, _i, _vp]),
    "born_batch_last_checkpoint_plan": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "born_batch_get_tile": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                 C.POINTER(C.c_float)]),
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

    def last_background_sweep(self):
        """background steps the last exact adjoint swept before its
        segments (0 when the forward left the history or the checkpoints)"""
        n = self.lib.born_last_background_sweep(self.h)
        self._chk(min(n, 0), "background sweep")
        return n

    def set_zchunk(self, zc):
        self._chk(self.lib.born_set_zchunk(self.h, zc), "zchunk")

    def set_autotune(self, on):
        self._chk(self.lib.born_set_autotune(self.h, int(on)), "autotune")

    KERNELS = ("fwd", "fwd_dV", "source", "rtm_adjoint", "background", "background_dV",
               "adjoint_state", "fwd_f16", "fwd_dV_f16", "background_f16",
               "background_dV_f16", "adjoint_state_f16", "fwd_pair", "fwd_dV_pair",
               "adjoint_state_pair", "fwd_f24", "fwd_dV_f24", "background_f24",
               "background_dV_f24", "adjoint_state_f24")
    STORAGE = {"f32": 0, "f16": 1, "f24": 2}

    def set_storage(self, kind):
        """Wavefield storage of the Born forward / exact adjoint: "f32",
        "f16" (FP16 levels with per-level power-of-two scales) or "f24"
        (FP32 values rounded to 24 bits: 16 significant bits, 3 bytes)."""
        self._chk(self.lib.born_set_storage(self.h, self.STORAGE[kind]), "storage")

    PLACEMENT = {"off": 0, "timed": 1, "cycle": 2}

    def set_placement(self, mode="timed", window_bytes=0):
        """Where the two levels of r (forward) and of the adjoint state sit
        in HBM for FP32 histories longer than one window: "timed" (default)
        the place timed fastest per window of history levels, "cycle" every
        place in turn, "off" the work buffers and a history without gaps;
        window_bytes: history bytes per window (0: 32 GiB)."""
        self._chk(self.lib.born_set_placement(self.h, self.PLACEMENT[mode], int(window_bytes)),
                  "placement")

    def last_placement(self):
        """(forward, adjoint) moves of the level pair in the last calls"""
        f, a = C.c_int(), C.c_int()
        self._chk(self.lib.born_last_placement(self.h, C.byref(f), C.byref(a)), "placement")
        return f.value, a.value

    def set_step_pairs(self, mode):
        """Two time steps per launch in the FP32 Born forward: 0 / False
        (default) single steps, 1 pairs unless timed slower, 2 / True pairs."""
        mode = {True: 2, False: 0}.get(mode, mode) if isinstance(mode, bool) else int(mode)
        self._chk(self.lib.born_set_step_pairs(self.h, mode), "step_pairs")

    @property
    def last_step_pairs(self):
        """True when the last Born forward ran in two-step launches."""
        return bool(self.lib.born_last_step_pairs(self.h) & 1)

    @property
    def last_adjoint_pairs(self):
        """True when the last exact adjoint ran in two-step launches."""
        return bool(self.lib.born_last_step_pairs(self.h) & 2)

    def tuning(self):
        out = {}
        for k, name in enumerate(self.KERNELS):
            ry, wy, zc, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
            self._chk(self.lib.born_get_tuning(self.h, k, C.byref(ry), C.byref(wy), C.byref(zc),
                                               C.byref(ms)), "tuning")
            tx, tyy, vx, la = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            self._chk(self.lib.born_get_tile(self.h, k, C.byref(tx), C.byref(tyy), C.byref(vx),
                                             C.byref(la)), "tile")
            out[name] = {"tile": [tx.value, tyy.value], "rows_per_thread": ry.value,
                         "waves": wy.value, "cells_per_lane": vx.value,
                         "lookahead": la.value, "zchunk": zc.value,
                         "ms": round(ms.value, 4)}
        return out

    def set_tuning(self, kernel, rows_per_thread, waves, cells_per_lane, lookahead, zchunk):
        k = self.KERNELS.index(kernel)
        self._chk(self.lib.born_set_tuning(self.h, k, rows_per_thread, waves, cells_per_lane,
                                           lookahead, zchunk), "set_tuning")

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

    def dev_source_prop_noalloc(self, jts, nt, p0, p1):
        """dev_source_prop without any allocation (scratch in a fixed device
        array; see born_dev_source_prop_noalloc)"""
        self._chk(self.lib.born_dev_source_prop_noalloc(self.h, jts, nt, _dptr(p0), _dptr(p1)),
                  "dev_source_prop_noalloc")

    def dev_rtm_adjoint(self, nt, p0, p1, img, traces):
        self._chk(self.lib.born_dev_rtm_adjoint(self.h, nt, _dptr(p0), _dptr(p1), _dptr(img),
                                                _dptr(traces)), "dev_rtm_adjoint")


# ---------------------------------------------------------------------------
# Born operator (include/born_host.h): rtm_zero_op on the HIP propagator
HOST_LIB_PATH = os.path.join(_HERE, "libborn_host.so")
_HOST = None


class OpDesc(C.Structure):
    """born_op_desc of include/born_host.h"""

    _fields_ = [("vn", C.c_int * 3), ("vo", C.c_float * 3), ("vd", C.c_float * 3),
                ("vel", C.c_void_p), ("nw", C.c_int), ("wdt", C.c_float),
                ("wavelet", C.c_void_p), ("nshots", C.c_int), ("src_x", C.c_void_p),
                ("src_y", C.c_void_p), ("src_depth", C.c_float), ("nt", C.c_int),
                ("t0", C.c_float), ("dt", C.c_float), ("rec_nx", C.c_int), ("rec_ny", C.c_int),
                ("rec_ox", C.c_float), ("rec_oy", C.c_float), ("rec_dx", C.c_float),
                ("rec_dy", C.c_float), ("rec_depth", C.c_float), ("nbound", C.c_int),
                ("mode", C.c_int), ("scatter", C.c_int), ("device", C.c_int),
                ("storage", C.c_int)]


OP_SIGNATURES = {
    "born_op_create": (_i, [C.POINTER(_vp), C.POINTER(OpDesc)]),
    "born_op_destroy": (_i, [_vp]),
    "born_op_last_error": (C.c_char_p, [_vp]),
    "born_op_sizes": (_i, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                           C.POINTER(_i), C.POINTER(_f), C.POINTER(_i)]),
    "born_op_forward": (_i, [_vp, _i, _vp, _vp]),
    "born_op_adjoint": (_i, [_vp, _i, _vp, _vp]),
    "born_op_dot_test": (_i, [_vp, C.c_double, C.POINTER(C.c_double), C.POINTER(_i)]),
    "born_op_lsrtm": (_i, [_vp, _vp, _vp, _i, C.POINTER(C.c_double)]),
}


def load_host_library(path=HOST_LIB_PATH):
    global _HOST
    if _HOST is None:
        if not os.path.exists(path):
            raise BornError("%s is not built" % path)
        lib = C.CDLL(path)
        for name, (res, args) in OP_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _HOST = lib
    return _HOST


class BornOp:
    """rtm_zero_op (base/rtm_zero_op_3d.h) on the MI355X: forward(model) ->
    data [shot][rec_y][rec_x][t], adjoint(data) -> model on the velocity grid.
    mode 0: the reference's RTM adjoint; mode 1: the exact transpose."""

    def __init__(self, vel, vo, vd, wavelet, wdt, src_x, src_y, src_depth, nt, t0, dt,
                 rec_n, rec_o, rec_d, rec_depth, nbound=50, mode=0, scatter=0, device=0,
                 storage="f32"):
        self.lib = load_host_library()
        vel = np.ascontiguousarray(vel, np.float32)
        w = np.ascontiguousarray(wavelet, np.float32)
        sx = np.ascontiguousarray(src_x, np.float32)
        sy = np.ascontiguousarray(src_y, np.float32)
        self._keep = (vel, w, sx, sy)
        d = OpDesc()
        d.vn[:] = [vel.shape[2], vel.shape[1], vel.shape[0]]
        d.vo[:] = list(vo)
        d.vd[:] = list(vd)
        d.vel = vel.ctypes.data
        d.nw, d.wdt, d.wavelet = len(w), wdt, w.ctypes.data
        d.nshots, d.src_x, d.src_y, d.src_depth = len(sx), sx.ctypes.data, sy.ctypes.data, src_depth
        d.nt, d.t0, d.dt = nt, t0, dt
        d.rec_nx, d.rec_ny = rec_n
        d.rec_ox, d.rec_oy = rec_o
        d.rec_dx, d.rec_dy = rec_d
        d.rec_depth, d.nbound, d.mode, d.scatter, d.device = rec_depth, nbound, mode, scatter, device
        d.storage = Propagator.STORAGE[storage]
        h = C.c_void_p()
        rc = self.lib.born_op_create(C.byref(h), C.byref(d))
        self.h = h
        if rc != 0:
            msg = self.lib.born_op_last_error(h) if h else b"create failed"
            self.close()
            raise BornError("born_op_create: %s" % msg.decode())
        mn, dn, ntp, dtp = C.c_longlong(), C.c_longlong(), C.c_int(), C.c_float()
        dom = (C.c_int * 3)()
        self.lib.born_op_sizes(self.h, C.byref(mn), C.byref(dn), C.byref(ntp), C.byref(dtp), dom)
        self.model_shape = vel.shape
        self.data_shape = (len(sx), rec_n[1], rec_n[0], nt)
        self.nt, self.dt, self.domain = ntp.value, dtp.value, tuple(dom)

    def close(self):
        if getattr(self, "h", None):
            self.lib.born_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise BornError("%s: %s" % (what, self.lib.born_op_last_error(self.h).decode()))

    def forward(self, model, data=None, add=False):
        m = np.ascontiguousarray(model, np.float32)
        out = data if data is not None else np.zeros(self.data_shape, np.float32)
        self._chk(self.lib.born_op_forward(self.h, int(add), _ptr(m), _ptr(out)), "forward")
        return out

    def adjoint(self, data, model=None, add=False):
        d = np.ascontiguousarray(data, np.float32)
        out = model if model is not None else np.zeros(self.model_shape, np.float32)
        self._chk(self.lib.born_op_adjoint(self.h, int(add), _ptr(out), _ptr(d)), "adjoint")
        return out

    def dot_test(self, tol=1e-4):
        dots = (C.c_double * 4)()
        ok = C.c_int()
        self._chk(self.lib.born_op_dot_test(self.h, tol, dots, C.byref(ok)), "dot_test")
        return bool(ok.value), list(dots)

    def lsrtm(self, data, niter):
        d = np.ascontiguousarray(data, np.float32)
        m = np.zeros(self.model_shape, np.float32)
        res = (C.c_double * max(niter, 1))()
        self._chk(self.lib.born_op_lsrtm(self.h, _ptr(d), _ptr(m), niter, res), "lsrtm")
        return m, list(res)[:niter]


class _Shot(C.Structure):
    """born_shot (include/born_hip.h)"""
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("vel", _vp), ("model", _vp),
                ("npts", _i), ("nt_big", _i), ("src_locs", _vp), ("src_vals", _vp),
                ("rec_nx", _i), ("rec_ny", _i), ("rec_locs", _vp)]


class Batch:
    """Shots with their own windowed domains run together (born_batch_*).

    shots: dicts with nx, ny, nz, vel ([nz][ny][nx] V = v^2 dt^2), model,
    src_locs, src_vals ([npts][nt_big]), rec_nx, rec_ny, rec_locs.  The
    propagator supplies the sinc tables, sponge, storage and stream."""

    def __init__(self, prop, shots, d):
        self.prop = prop
        self.lib = prop.lib
        self._keep = []
        arr = (_Shot * len(shots))()
        self.nrec, self.ntb, self.cells = [], [], []
        for k, sh in enumerate(shots):
            vel = np.ascontiguousarray(sh["vel"], np.float32)
            model = np.ascontiguousarray(sh["model"], np.float32)
            sl = np.ascontiguousarray(sh["src_locs"], np.int32)
            sv = np.ascontiguousarray(sh["src_vals"], np.float32)
            rl = np.ascontiguousarray(sh["rec_locs"], np.int32)
            self._keep += [vel, model, sl, sv, rl]
            arr[k] = _Shot(sh["nx"], sh["ny"], sh["nz"], _ptr(vel), _ptr(model), len(sl),
                           sv.shape[-1], _ptr(sl), _ptr(sv), sh["rec_nx"], sh["rec_ny"],
                           _ptr(rl))
            self.nrec.append(sh["rec_nx"] * sh["rec_ny"])
            self.cells.append(sh["nx"] * sh["ny"] * sh["nz"])
            self.ntb.append(sv.shape[-1])
        h = _vp()
        d1, d2, d3 = d
        prop._chk(self.lib.born_batch_create(prop.h, len(shots), arr, d1, d2, d3, C.byref(h)),
                  "batch_create")
        self.h = h
        self._keep = []  # copied to the device

    def close(self):
        if getattr(self, "h", None):
            self.lib.born_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, jt, nt, scatter=0):
        self.prop._chk(self.lib.born_batch_forward(self.h, jt, nt, scatter), "batch_forward")

    def traces(self, shot):
        out = np.zeros((self.nrec[shot], self.ntb[shot]), np.float32)
        self.prop._chk(self.lib.born_batch_traces(self.h, shot, _ptr(out)), "batch_traces")
        return out

    def adjoint_exact(self, jt, nt, scatter=0):
        """images of every shot from the current traces (born_batch_adjoint_exact)"""
        self.prop._chk(self.lib.born_batch_adjoint_exact(self.h, jt, nt, scatter),
                       "batch_adjoint_exact")

    def set_traces(self, shot, rec):
        r = np.ascontiguousarray(rec, np.float32)
        assert r.shape == (self.nrec[shot], self.ntb[shot])
        self.prop._chk(self.lib.born_batch_set_traces(self.h, shot, _ptr(r)), "batch_set_traces")

    def image(self, shot):
        out = np.zeros(self.cells[shot], np.float32)
        self.prop._chk(self.lib.born_batch_image(self.h, shot, _ptr(out)), "batch_image")
        return out

    def last_checkpoint_plan(self):
        s, r = C.c_int(), C.c_int()
        self.prop._chk(self.lib.born_batch_last_checkpoint_plan(self.h, C.byref(s), C.byref(r)),
                       "batch_plan")
        return s.value, r.value

    def tile(self, scatter=0):
        tx, ty, zc, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        self.prop._chk(self.lib.born_batch_get_tile(self.h, scatter, C.byref(tx), C.byref(ty),
                                                    C.byref(zc), C.byref(ms)), "batch_tile")
        return {"tile": [tx.value, ty.value], "zchunk": zc.value, "ms": round(ms.value, 4)}
