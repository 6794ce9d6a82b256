"""The allocation-free system solve (gx_sys_factor_and_solve_values_noalloc) against the allocating
one (gx_sys_factor_and_solve_values) through the C ABI: bit-identical solutions and statuses on two
calls in a row, the second with a smaller system, a singular system, and the size limit."""
import numpy as np
import pytest

from gridoxide_amd import _lib, config_grid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _jacobian(oracle_mod, cfg, seed):
    net = config_grid(cfg)
    o = oracle_mod.Oracle(net)
    rng = np.random.default_rng(seed)
    vm = net.vm * (1.0 + 0.01 * rng.standard_normal(net.n_bus))
    va = net.va + 0.01 * rng.standard_normal(net.n_bus)
    p, q = o.injections(vm, va)
    r, c, v = o.jacobian(vm, va, p, q)
    n = int(max(r.max(), c.max())) + 1
    return n, r.astype(np.int32), c.astype(np.int32), v


class _System:
    def __init__(self, n, r, c, v):
        self.L = _lib.lib()
        self.n = n
        self.h = self.L.gx_sys_create(n, r.size, _lib.ptr(r, _lib._i32p), _lib.ptr(c, _lib._i32p), _lib.ptr(v))
        assert self.h, self.L.gx_last_error()

    def host(self, vals, rhs):
        x = np.zeros(self.n)
        rc = self.L.gx_sys_factor_and_solve_values(self.h, _lib.ptr(np.ascontiguousarray(vals, np.float64)),
                                                   _lib.ptr(np.ascontiguousarray(rhs, np.float64)), _lib.ptr(x))
        return rc, x

    def noalloc(self, vals, rhs):
        import torch
        dv = torch.as_tensor(np.ascontiguousarray(vals, np.float64)).cuda()
        db = torch.as_tensor(np.ascontiguousarray(rhs, np.float64)).cuda()
        dx = torch.full((self.n,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        rc = self.L.gx_sys_factor_and_solve_values_noalloc(self.h, dv.data_ptr(), db.data_ptr(), dx.data_ptr(), None)
        return rc, dx.cpu().numpy()

    def close(self):
        self.L.gx_sys_destroy(self.h)


def test_noalloc_equals_allocating_solve(oracle_mod):
    import scipy.sparse as sp
    big = _System(*_jacobian(oracle_mod, 2, 1))      # 2,000-bus grid: n = 3,698
    small = _System(*_jacobian(oracle_mod, 1, 2))    # 200-bus grid
    try:
        rng = np.random.default_rng(7)
        for name, sysm, seed in (("2000-bus Jacobian", big, 11), ("200-bus Jacobian (smaller, second)", small, 12)):
            _say(f"case: {name}, new values and right-hand side")
            n, r, c, v = _jacobian(oracle_mod, 2 if sysm is big else 1, seed)
            b = rng.standard_normal(sysm.n)
            rc0, x0 = sysm.host(v, b)
            rc1, x1 = sysm.noalloc(v, b)
            assert rc0 == rc1 == 0, (rc0, rc1, _lib.lib().gx_last_error())
            assert np.array_equal(x0, x1)
            A = sp.csc_matrix((v, (r, c)), shape=(sysm.n, sysm.n))
            assert np.abs(A @ x1 - b).max() <= 1e-9 * max(1.0, np.abs(b).max())
        _say("case: 2000-bus Jacobian again after the smaller system")
        n, r, c, v = _jacobian(oracle_mod, 2, 13)
        b = rng.standard_normal(big.n)
        assert np.array_equal(big.host(v, b)[1], big.noalloc(v, b)[1])
        _say("case: zero values (a zero pivot): both report Singular")
        z = np.zeros(r.size)
        assert big.host(z, b)[0] == big.noalloc(z, b)[0] == _lib.GX_SINGULAR
    finally:
        big.close()
        small.close()


def test_noalloc_rejects_systems_beyond_its_array():
    _say("case: tridiagonal system of 800,000 rows (2.4 M L+U entries) -> GX_ERR_CAPACITY")
    n = 800_000
    i = np.arange(n, dtype=np.int32)
    r = np.r_[i, i[1:], i[:-1]].astype(np.int32)
    c = np.r_[i, i[:-1], i[1:]].astype(np.int32)
    v = np.r_[np.full(n, 4.0), np.full(n - 1, -1.0), np.full(n - 1, -1.0)]
    s = _System(n, r, c, v)
    try:
        rc, _ = s.noalloc(v, np.ones(n))
        assert rc == _lib.GX_ERR_CAPACITY
        assert b"32 B per L+U entry" in _lib.lib().gx_last_error()
        rc, x = s.host(v, np.ones(n))
        assert rc == 0 and np.all(np.isfinite(x))
    finally:
        s.close()
