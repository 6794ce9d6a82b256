"""ctypes bindings for the CPU oracle (liboracle.so) and the reference KLU build (_ref/libklu_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s ``cpu_baseline`` leg.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compiles liboracle.so (and _ref/libklu_ref.so where the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libklu_ref.so"))


def ref():
    global _REF
    if _REF is None:
        _REF = ctypes.CDLL(os.path.join(HERE, "_ref", "libklu_ref.so"))
        _REF.kref_create.restype = ctypes.c_void_p
        _REF.kref_create.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int]
        _REF.kref_destroy.argtypes = [ctypes.c_void_p]
        _REF.kref_counts.argtypes = [ctypes.c_void_p, _i32p]
        _REF.kref_symbolic.argtypes = [ctypes.c_void_p, _i32p, _i32p, _i32p]
        _REF.kref_extract.argtypes = [ctypes.c_void_p] + [_i32p, _i32p, _f64p] * 3 + [_i32p, _i32p, _i32p]
        _REF.kref_refactor_solve.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p, _f64p]
        _REF.kref_solve.argtypes = [ctypes.c_void_p, _f64p]
    return _REF


def c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ybus(net):
    """Merged Y-bus as CSR (rowptr, col, complex values)."""
    L = lib()
    cap = 4 * (net.line_from.size + net.tr_from.size) + net.n_bus
    rowptr = np.zeros(net.n_bus + 1, np.int32)
    col = np.zeros(cap, np.int32)
    yre = np.zeros(cap)
    yim = np.zeros(cap)
    ys = c64(np.stack([net.tr_ys.real, net.tr_ys.imag], 1).ravel()) if net.tr_from.size else np.zeros(2)
    ysh = c64(np.stack([net.tr_ysh.real, net.tr_ysh.imag], 1).ravel()) if net.tr_from.size else np.zeros(2)
    tap = c64(np.stack([net.tr_tap.real, net.tr_tap.imag], 1).ravel()) if net.tr_from.size else np.zeros(2)
    lf, lt, lr, lx, lb, lg = (c32(net.line_from), c32(net.line_to), c64(net.line_r), c64(net.line_x),
                              c64(net.line_b), c64(net.line_g))
    tf, tt = c32(net.tr_from), c32(net.tr_to)
    nnz = L.orc_ybus(net.n_bus, lf.size, _p(lf, _i32p), _p(lt, _i32p), _p(lr, _f64p), _p(lx, _f64p), _p(lb, _f64p),
                     _p(lg, _f64p), tf.size, _p(tf, _i32p), _p(tt, _i32p), _p(ys, _f64p), _p(ysh, _f64p),
                     _p(tap, _f64p), cap, _p(rowptr, _i32p), _p(col, _i32p), _p(yre, _f64p), _p(yim, _f64p))
    assert nnz >= 0
    return rowptr, col[:nnz].copy(), yre[:nnz] + 1j * yim[:nnz]


class Oracle:
    """One topology on the CPU oracle."""

    def __init__(self, net, zip_coef=None):
        """zip_coef: optional [n_bus, 6] per-bus (P, Q) of constant-power, -current and -impedance
        ZIP terms (defaults to ``net.zip_coef``)."""
        self.net = net
        self.n = net.n_bus
        self.rowptr, self.col, self.y = ybus(net)
        self.yre = c64(self.y.real)
        self.yim = c64(self.y.imag)
        self.type = c32(net.bus_type)
        if zip_coef is None:
            zip_coef = getattr(net, "zip_coef", None)
        self.zip = None if zip_coef is None else c64(zip_coef).reshape(self.n, 6)

    def _ya(self):
        return (_p(self.rowptr, _i32p), _p(self.col, _i32p), _p(self.yre, _f64p), _p(self.yim, _f64p))

    def injections(self, vm, va):
        p = np.zeros(self.n)
        q = np.zeros(self.n)
        lib().orc_power_injections(self.n, *self._ya(), _p(c64(vm), _f64p), _p(c64(va), _f64p), _p(p, _f64p),
                                   _p(q, _f64p))
        return p, q

    def jacobian(self, vm, va, p, q):
        cap = 4 * self.col.size
        rows = np.zeros(cap, np.int32)
        cols = np.zeros(cap, np.int32)
        vals = np.zeros(cap)
        vm, va, p, q = c64(vm), c64(va), c64(p), c64(q)
        nnz = lib().orc_jacobian(self.n, *self._ya(), _p(self.type, _i32p), _p(vm, _f64p), _p(va, _f64p),
                                 _p(p, _f64p), _p(q, _f64p), cap, _p(rows, _i32p), _p(cols, _i32p), _p(vals, _f64p))
        assert nnz >= 0
        return rows[:nnz].copy(), cols[:nnz].copy(), vals[:nnz].copy()

    def linear_initial_guess(self, p_spec, q_spec, vm, va):
        vm = c64(vm).copy()
        va = c64(va).copy()
        rc = lib().orc_linear_initial_guess_zip(self.n, *self._ya(), _p(self.type, _i32p), _p(self.zip, _f64p),
                                                _p(c64(p_spec), _f64p), _p(c64(q_spec), _f64p), _p(vm, _f64p),
                                                _p(va, _f64p))
        return vm, va, rc

    def batch_solve(self, p_spec, q_spec, vm0, va0, tol=1e-8, max_iter=20, init=0, threads=1, order=None):
        """order: optional (q, p) elimination order of the first factorization (column order and
        preferred rows of the unknowns, e.g. the device's BTF + AMD orders); None = bus-major."""
        p_spec = c64(np.atleast_2d(p_spec))
        S = p_spec.shape[0]
        q_spec = c64(np.atleast_2d(q_spec))
        vm0 = c64(np.broadcast_to(vm0, (S, self.n)))
        va0 = c64(np.broadcast_to(va0, (S, self.n)))
        vm = np.zeros((S, self.n))
        va = np.zeros((S, self.n))
        status = np.zeros(S, np.int32)
        iters = np.zeros(S, np.int32)
        hist = np.zeros((S, max_iter))
        colq = rowp = None
        if order is not None:
            colq, rowp = c32(order[0]), c32(order[1])
        lib().orc_batch_solve_ordered(self.n, *self._ya(), _p(self.type, _i32p), _p(self.zip, _f64p), S,
                                      _p(p_spec, _f64p), _p(q_spec, _f64p), _p(vm0, _f64p), _p(va0, _f64p), init,
                                      ctypes.c_double(tol), max_iter, threads, _p(colq, _i32p), _p(rowp, _i32p),
                                      _p(vm, _f64p), _p(va, _f64p), _p(status, _i32p), _p(iters, _i32p),
                                      _p(hist, _f64p))
        return vm, va, status, iters, hist

    def components(self):
        """Connected components of the Y-bus graph (src/network.rs:121-145)."""
        comp = np.zeros(self.n, np.int32)
        nc = lib().orc_components(self.n, _p(self.rowptr, _i32p), _p(self.col, _i32p), _p(comp, _i32p))
        return comp, nc

    def persistent_solve(self, p_spec, q_spec, vm0, va0, tol=1e-8, max_iter=20, init=1):
        """One scenario the way PowerFlowModel.solve runs it (src/python.rs:256-274): the linear
        initial guess on the buses as given, then every component without a Slack bus pinned to V = 0
        with no injection (classify / mark_unreferenced_islands, src/network.rs:176-223), then the
        Newton loop.  Returns (vm, va, status, iters, history)."""
        vm = c64(vm0).copy()
        va = c64(va0).copy()
        p = c64(p_spec).copy()
        q = c64(q_spec).copy()
        if init:
            vm, va, _ = self.linear_initial_guess(p, q, vm, va)
        comp, nc = self.components()
        types = self.type.copy()
        for c in range(nc):
            members = np.nonzero(comp == c)[0]
            if not (types[members] == 0).any():
                types[members] = 0
                vm[members] = va[members] = p[members] = q[members] = 0.0
        marked = Oracle.__new__(Oracle)
        marked.__dict__.update(self.__dict__)
        marked.type = c32(types)
        out = marked.batch_solve(p, q, vm, va, tol, max_iter, init=0)
        return tuple(a[0] for a in out)


def sparse_solve(ap, ai, ax, b):
    ap, ai, ax, b = c32(ap), c32(ai), c64(ax), c64(b)
    x = np.zeros_like(b)
    rc = lib().orc_sparse_solve(ap.size - 1, _p(ap, _i32p), _p(ai, _i32p), _p(ax, _f64p), _p(b, _f64p), _p(x, _f64p))
    return x, rc


class KluRef:
    """The reference's vendored C KLU on one CSC matrix (analyze + factor)."""

    def __init__(self, ap, ai, ax, scale=0):
        self.ap, self.ai, self.ax = c32(ap), c32(ai), c64(ax)
        self.n = self.ap.size - 1
        R = ref()
        self.h = R.kref_create(self.n, _p(self.ap, _i32p), _p(self.ai, _i32p), _p(self.ax, _f64p), scale)
        if not self.h:
            raise RuntimeError("klu_analyze/klu_factor failed")

    def __del__(self):
        if getattr(self, "h", None):
            ref().kref_destroy(self.h)
            self.h = None

    def counts(self):
        c = np.zeros(4, np.int32)
        ref().kref_counts(self.h, _p(c, _i32p))
        return dict(nblocks=int(c[0]), lnz=int(c[1]), unz=int(c[2]), nzoff=int(c[3]))

    def symbolic(self):
        P = np.zeros(self.n, np.int32)
        Q = np.zeros(self.n, np.int32)
        R = np.zeros(self.n + 1, np.int32)
        ref().kref_symbolic(self.h, _p(P, _i32p), _p(Q, _i32p), _p(R, _i32p))
        nb = self.counts()["nblocks"]
        return P, Q, R[:nb + 1].copy()

    def extract(self):
        c = self.counts()
        n = self.n
        Lp, Li, Lx = np.zeros(n + 1, np.int32), np.zeros(c["lnz"], np.int32), np.zeros(c["lnz"])
        Up, Ui, Ux = np.zeros(n + 1, np.int32), np.zeros(c["unz"], np.int32), np.zeros(c["unz"])
        nzo = max(c["nzoff"], 1)
        Fp, Fi, Fx = np.zeros(n + 1, np.int32), np.zeros(nzo, np.int32), np.zeros(nzo)
        P, Q, R = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n + 1, np.int32)
        ok = ref().kref_extract(self.h, _p(Lp, _i32p), _p(Li, _i32p), _p(Lx, _f64p), _p(Up, _i32p), _p(Ui, _i32p),
                                _p(Ux, _f64p), _p(Fp, _i32p), _p(Fi, _i32p), _p(Fx, _f64p), _p(P, _i32p),
                                _p(Q, _i32p), _p(R, _i32p))
        assert ok
        return dict(Lp=Lp, Li=Li, Lx=Lx, Up=Up, Ui=Ui, Ux=Ux, Fp=Fp, Fi=Fi[:c["nzoff"]], Fx=Fx[:c["nzoff"]],
                    P=P, Q=Q, R=R[:c["nblocks"] + 1])

    def solve(self, b):
        x = c64(b).copy()
        ok = ref().kref_solve(self.h, _p(x, _f64p))
        assert ok
        return x

    def refactor_solve(self, ax, b):
        ax = c64(ax)
        x = c64(b).copy()
        ok = ref().kref_refactor_solve(self.h, _p(self.ap, _i32p), _p(self.ai, _i32p), _p(ax, _f64p), _p(x, _f64p))
        return x, bool(ok)


def triplets_to_csc(n, rows, cols, vals):
    """CSC with rows sorted per column and duplicates summed (build_csc_structure semantics,
    src/klu_native/mod.rs:88-114)."""
    order = np.lexsort((rows, cols))
    r, c, v = rows[order], cols[order], vals[order]
    key = c.astype(np.int64) * n + r
    uniq, first = np.unique(key, return_index=True)
    sums = np.add.reduceat(v, first) if v.size else v
    ai = (uniq % n).astype(np.int32)
    cc = (uniq // n).astype(np.int64)
    ap = np.zeros(n + 1, np.int32)
    np.add.at(ap, cc + 1, 1)
    return np.cumsum(ap).astype(np.int32), ai, sums
