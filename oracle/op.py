"""Oracle composition of the reference's Born operator -- TEST INFRASTRUCTURE.

rtm_zero_op::forward / adjoint (base/rtm_zero_op_3d.cc:101-299) composed from
the oracle's restatements: calc_stability (fd_prop_3d.cc:6-33), the padded
domain (source_func_3d.cc:4-74, origin computed once), the wavelet stamp
(source_func_3d.cc:76-127), rand_subspace (vel_fd_3d.cc:29-134), the receiver
map (data_rtm_3d.cc:59-90, map_data_3d.cc:4-35), extract_sub / laplac /
rtmForward / map_data / deriv for the forward and their counterparts for the
adjoint.  Single pass over the shots, each shot from a fresh propagator.

exact=True replaces the reference's RTM adjoint (sourceProp + rtmAdjoint)
by the FP64 transpose of the forward propagation (orc_exact_adjoint_d), the
operator pair the shipped library runs in its exact mode.
"""
import ctypes as C

import numpy as np

from . import Prop, lib, sinc_table


class OracleBornOp:
    def __init__(self, vn, vo, vd, vel, wavelet, wdt, src_x, src_y, src_depth, nt, t0, dt,
                 rec_n, rec_o, rec_d, rec_depth, nbound, nthreads=0, exact=False):
        L = lib()
        self.exact = exact
        self.vn, self.vo, self.vd = list(vn), list(vo), list(vd)
        self.vel = np.ascontiguousarray(vel, np.float32).ravel()
        self.src_x, self.src_y = list(src_x), list(src_y)
        self.src_depth = src_depth
        self.ntd, self.t0, self.dtd = nt, t0, dt
        self.rec_n, self.rec_o, self.rec_d = rec_n, rec_o, rec_d
        self.rec_depth = rec_depth
        self.nbound = nbound
        self.nthreads = nthreads
        dx, dy, dz = vd
        # vel_fd_3d.h:43-51 (min sampling as the reference computes it)
        if dz < dx and dz < dy:
            dmin = dz
        elif dy < dx:
            dmin = dz
        else:
            dmin = dx
        vmax = np.float32(self.vel.astype(np.float64).max())
        dtp, ntp = C.c_float(), C.c_int()
        L.orc_calc_stability(vmax, np.float32(dmin), np.array([dt, wdt], np.float32), 2, nt,
                             C.byref(dtp), C.byref(ntp))
        self.dt = np.float32(dtp.value)
        self.jts = int(np.float32(wdt) / self.dt)
        self.jtd = int(np.float32(dt) / self.dt)
        self.nt = (nt - 1) * self.jtd + 1
        self.nt_big = 8 + nt
        pad = nbound + 4
        self.dn = [L.orc_padded_size(n, nbound, 4) for n in vn]
        self.do = [np.float32(o) - np.float32(d) * pad for o, d in zip(vo, vd)]
        w = np.ascontiguousarray(wavelet, np.float32)
        self.wdev = np.zeros_like(w)
        L.orc_wavelet_deriv(len(w), np.float32(wdt), w, self.wdev)
        self.tab_s = sinc_table(self.jts)
        self.tab_d = sinc_table(self.jtd)

    @property
    def model_shape(self):
        return (self.vn[2], self.vn[1], self.vn[0])

    @property
    def data_shape(self):
        return (len(self.src_x), self.rec_n[1], self.rec_n[0], self.ntd)

    def _shot(self, ishot):
        L = lib()
        n1, n2, n3 = self.dn
        o1, o2, o3 = self.do
        d1, d2, d3 = self.vd
        locs = np.zeros(9, np.int32)
        vals = np.zeros(9 * self.nt_big, np.float32)
        L.orc_source_func(n1, n2, o1, o2, o3, d1, d2, d3, self.src_x[ishot], self.src_y[ishot],
                          self.src_depth, self.wdev, len(self.wdev), self.nt_big, locs, vals)
        seed = int(locs[3]) - (int(locs[3]) // 10000) * 10000
        v = []
        for nr in (self.nbound, 0):
            out = np.zeros(n1 * n2 * n3, np.float32)
            L.orc_rand_subspace(seed, nr, nr, nr, self.dt, *self.vn, *self.vo, *self.vd, self.vel,
                                n1, n2, n3, o1, o2, o3, d1, d2, d3, out)
            v.append(out)
        rl = np.zeros(self.rec_n[0] * self.rec_n[1], np.int32)
        L.orc_receiver_locs(self.rec_n[0], self.rec_n[1], self.rec_o[0], self.rec_o[1],
                            self.rec_d[0], self.rec_d[1], self.rec_depth, n1, n2, o1, o2, o3,
                            d1, d2, d3, rl)
        P = Prop(self.nthreads)
        P.create_space(d1, d2, d3, n1, n2, n3)
        P.transfer_sinc_table_s(self.tab_s)
        P.transfer_sinc_table_d(self.tab_d)
        P.transfer_source_func(locs, vals.reshape(9, self.nt_big))
        P.transfer_vel_func1(v[0])
        P.transfer_vel_func2(v[1])
        return P, rl

    def forward(self, model):
        L = lib()
        n1, n2, n3 = self.dn
        rnx, rny = self.rec_n
        model = np.ascontiguousarray(model, np.float32).ravel()
        data = np.zeros(self.data_shape, np.float32)
        for s in range(len(self.src_x)):
            P, rl = self._shot(s)
            img = np.zeros(n1 * n2 * n3, np.float32)
            L.orc_extract_sub(*self.vn, *self.vo, *self.vd, model, n1, n2, n3, *self.do, img)
            lap = np.zeros_like(img)
            L.orc_laplac_forward(n1, n2, n3, img, lap)
            rec = np.zeros((rnx * rny, self.nt_big), np.float32)
            P.transfer_receiver_func(rl, rec)
            P.rtm_forward(self.jtd, lap, self.nt)
            sd = np.zeros(rnx * rny * self.ntd, np.float32)
            L.orc_map_data_forward(self.ntd, rnx, rny, self.nt_big, rec.ravel(), sd)
            shot = np.zeros_like(sd)
            L.orc_deriv_forward(self.ntd, rnx, rny, np.float32(self.dtd), sd, shot)
            data[s] += shot.reshape(rny, rnx, self.ntd)
        return data

    def adjoint(self, data):
        L = lib()
        n1, n2, n3 = self.dn
        rnx, rny = self.rec_n
        data = np.ascontiguousarray(data, np.float32)
        image = np.zeros(self.vn[0] * self.vn[1] * self.vn[2], np.float32)
        for s in range(len(self.src_x)):
            P, rl = self._shot(s)
            sd = np.zeros(rnx * rny * self.ntd, np.float32)
            L.orc_deriv_adjoint(self.ntd, rnx, rny, np.float32(self.dtd), sd,
                                np.ascontiguousarray(data[s]).ravel())
            rec = np.zeros(rnx * rny * self.nt_big, np.float32)
            L.orc_map_data_adjoint(self.ntd, rnx, rny, self.nt_big, rec, sd)
            rec = rec.reshape(rnx * rny, self.nt_big)
            P.transfer_receiver_func(rl, rec)
            if self.exact:
                img = P.exact_adjoint_d(self.jtd, rec.astype(np.float64), self.nt)
                img = np.ascontiguousarray(img, np.float32).ravel()
            else:
                p0 = np.zeros(n1 * n2 * n3, np.float32)
                p1 = np.zeros_like(p0)
                P.source_prop(p0, p1, self.jts, self.nt)
                img = np.zeros_like(p0)
                P.rtm_adjoint(p1, p0, img, self.nt)
            lap = np.zeros_like(img)
            L.orc_laplac_adjoint(n1, n2, n3, lap, img)
            L.orc_add_image(*self.vn, *self.vo, *self.vd, image, n1, n2, n3, *self.do, lap)
        return image.reshape(self.model_shape)
