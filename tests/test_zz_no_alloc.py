"""The no-allocation PCG entry point (g2ohip_linear_solve_pcg_device): its
kernels take their scratch from a per-GPU static array, so two calls in a row
- the second on a smaller system, after the first left its vectors there -
must give the allocating entry point's result (g2ohip_linear_solve_pcg on the
same matrix from host arrays): the same iterations, residual memory and x.

The test's own device buffers come from the HIP runtime the library itself
links (hipMalloc through ctypes), not from a second runtime in the process."""
import ctypes as ct
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """hipMalloc'd copies of host arrays (freed on close), through the HIP
    runtime libg2o_hip.so is linked against."""

    def __init__(self):
        import g2o_amd
        g2o_amd.lib()  # loads that runtime first
        path = "/opt/rocm/lib/libamdhip64.so.7"
        self.hip = ct.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.hip.hipFree.argtypes = [ct.c_void_p]
        self.hip.hipSetDevice.argtypes = [ct.c_int]
        assert self.hip.hipSetDevice(0) == 0
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ct.c_void_p()
        assert self.hip.hipMalloc(ct.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ct.c_void_p(p), out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _solve_device(ctx, nb, bd, F, b, x, tol):
    import g2o_amd
    return g2o_amd.linear_solve_pcg_device(ctx, nb, bd, F[0], F[1], F[2], F[3], b, x, tol=tol, max_iter=-1,
                                           residual=-1.0)


def _spd_system(nb, bd, seed):
    rng = np.random.default_rng(seed)
    rows, cols, blocks = [], [], []
    for i in range(nb):
        for j in range(i, nb):
            if i == j or (abs(i - j) <= 3 and rng.random() < 0.7) or rng.random() < 8.0 / nb:
                rows.append(i)
                cols.append(j)
                blocks.append(rng.normal(size=(bd, bd)))
    rows, cols = np.array(rows), np.array(cols)
    blocks = np.array(blocks)
    # diagonal blocks: SPD and dominant enough to make the whole matrix SPD
    deg = np.bincount(np.concatenate([rows, cols]), minlength=nb)
    for k in np.flatnonzero(rows == cols):
        a = rng.normal(size=(bd, bd))
        blocks[k] = a @ a.T + (2.0 * bd * deg[rows[k]] + 1.0) * np.eye(bd)
    b = rng.normal(size=nb * bd)
    return rows, cols, blocks, b


@pytest.mark.parametrize("bd", [6, 3])
def test_no_alloc_pcg_matches_allocating_entry(bd):
    import g2o_amd
    ctx = g2o_amd.Context(0)
    dev = DeviceArrays()
    for call, nb in enumerate((1400, 300)):  # the second call is smaller
        rows, cols, blocks, b = _spd_system(nb, bd, 100 + call)
        print(f"bd {bd} call {call}: nb {nb} upper blocks {len(rows)}", flush=True)
        x_ref, it_ref, res_ref = g2o_amd.linear_solve_pcg(ctx, nb, bd, rows, cols, blocks, b, tol=1e-10,
                                                          max_iter=-1, residual=-1.0)
        F = [dev.put(a) for a in g2o_amd.full_block_rows(nb, bd, rows, cols, blocks)]
        x0 = np.full(nb * bd, np.nan)
        xd = dev.put(x0)
        it, res = _solve_device(ctx, nb, bd, F, dev.put(b), xd, 1e-10)
        xh = dev.get(xd, x0)
        print(f"  iterations {it} / {it_ref}, max |dx| {np.abs(xh - x_ref).max():.3e}", flush=True)
        assert it == it_ref and it > 0
        assert res == res_ref
        np.testing.assert_array_equal(xh, x_ref)
        # and it solved the system
        A = np.zeros((nb * bd, nb * bd))
        for r_, c_, B in zip(rows, cols, blocks):
            A[r_ * bd:(r_ + 1) * bd, c_ * bd:(c_ + 1) * bd] = B
            A[c_ * bd:(c_ + 1) * bd, r_ * bd:(r_ + 1) * bd] = B.T
        assert np.linalg.norm(A @ xh - b) <= 1e-4 * np.linalg.norm(b)
    dev.close()


def test_no_alloc_pcg_rejects_systems_above_its_limit():
    """The size check runs on the host before any launch."""
    import g2o_amd
    ctx = g2o_amd.Context(0)
    dev = DeviceArrays()
    z = dev.put(np.zeros(8))
    with pytest.raises(g2o_amd.G2oHipError) as ei:
        _solve_device(ctx, 20000, 6, [z, z, z, z], z, z, 1e-6)
    assert ei.value.status == g2o_amd._lib.ERR_UNSUPPORTED
    dev.close()
