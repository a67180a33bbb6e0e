"""The allocation-free two-site update (tebd_two_site_op_noalloc) against the context's
tebd_two_site_op on the same bond.

The entry point takes the bond's tensors in device memory and keeps its workspace in
a fixed device array; the context path allocates its workspace at tebd_create.  Both
run the same kernels on the same task, so lambda, the kept rank, the truncation error
and both new Gammas must agree to rounding (FP64: 1e-13).  Two calls in a row, the
second with a smaller bond, check that nothing of the first call's workspace contents
leaks into the second.
"""
import ctypes

import numpy as np
import pytest

from tebd_amd import Context, _lib

TOL = 1e-13


def _lib_noalloc():
    lib = _lib.lib()
    P, I = ctypes.c_void_p, ctypes.c_int
    f = lib.tebd_two_site_op_noalloc
    f.argtypes = [I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P]
    f.restype = I
    return f


class DeviceBuffers:
    """Device memory for the test's inputs and outputs, from the HIP runtime that
    libtebdhip.so itself uses (a second runtime in the process, e.g. torch's bundled
    one, would not see the device)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def random_state(L, d, chi, seed):
    """Padded Vidal state [site, b, i, a] with bond dims 1, chi, ..., chi, 1."""
    rng = np.random.default_rng(seed)
    dims = [1] + [chi] * (L - 1) + [1]
    gam = np.zeros((L, chi, d, chi), np.complex128)
    lam = np.zeros((L + 1, chi))
    for s in range(L):
        cl, cr = dims[s], dims[s + 1]
        gam[s, :cr, :, :cl] = (rng.normal(size=(cr, d, cl)) + 1j * rng.normal(size=(cr, d, cl))) / np.sqrt(cl * d)
    for b in range(L + 1):
        v = np.sort(rng.uniform(0.05, 1.0, dims[b]))[::-1]
        lam[b, :dims[b]] = v / np.linalg.norm(v)
    return gam, lam


def test_noalloc_rejects_sizes_beyond_its_workspace():
    f = _lib_noalloc()
    k, tr = ctypes.c_int(), ctypes.c_double()
    # d chi_l = 4 * 128 > 256: refused before any device call
    rc = f(4, 128, 64, 64, 64, *([None] * 9), None, ctypes.byref(k), ctypes.byref(tr), None)
    assert rc == _lib.TEBD_ERR_SHAPE


@pytest.mark.gpu
def test_noalloc_matches_context_two_site_op():
    f = _lib_noalloc()
    for call, (d, chi, seed) in enumerate([(4, 64, 11), (3, 20, 12)]):
        print(f"[gpu] noalloc call {call}: d={d} chi={chi}", flush=True)
        L, link = 4, 2  # sites 2, 3 (1-based): chi_l = chi_c = chi_r = chi
        gam, lam = random_state(L, d, chi, seed)
        rng = np.random.default_rng(seed + 100)
        op = rng.normal(size=(d * d, d * d)) + 1j * rng.normal(size=(d * d, d * d))
        with Context(L, d, chi, mode="fp64") as ctx:
            ctx.upload(gam, lam)
            tr_ctx = ctx.two_site_op(link, op)
            k_ctx = int(ctx.bond_dims()[link])
            g_ctx, l_ctx = ctx.download(chi)

        dev = DeviceBuffers()
        try:
            gl = dev.put(gam[link - 1, :chi, :, :chi])  # packed Gamma_l(a, i, b), a fastest
            gr = dev.put(gam[link, :chi, :, :chi])
            ll, lc, lr = (dev.put(lam[b, :chi]) for b in (link - 1, link, link + 1))
            o = dev.put(op.T)  # column-major d^2 x d^2
            zeros_g = np.zeros(chi * d * chi, np.complex128)
            gl_out, gr_out = dev.put(zeros_g), dev.put(zeros_g)
            lc_out = dev.put(np.zeros(chi))
            k, tr = ctypes.c_int(), ctypes.c_double()
            rc = f(d, chi, chi, chi, chi, gl, gr, ll, lc, lr, o, gl_out, gr_out, lc_out, ctypes.byref(k),
                   ctypes.byref(tr), None)
            assert rc == 0
            lam_new = dev.get(lc_out, np.zeros(chi))
            gl_all = dev.get(gl_out, zeros_g)
            gr_all = dev.get(gr_out, zeros_g)
        finally:
            dev.close()
        assert k.value == k_ctx
        kk = k.value
        assert abs(tr.value - tr_ctx) <= TOL
        assert np.max(np.abs(lam_new - l_ctx[link])) <= TOL
        gl_new = gl_all[:chi * d * kk].reshape(kk, d, chi)   # [b, i, a]
        gr_new = gr_all[:kk * d * chi].reshape(chi, d, kk)
        err_l = np.max(np.abs(gl_new - g_ctx[link - 1, :kk, :, :chi]))
        err_r = np.max(np.abs(gr_new - g_ctx[link, :chi, :, :kk]))
        print(f"[gpu] noalloc call {call}: k={kk} lambda err {np.max(np.abs(lam_new - l_ctx[link])):.1e} "
              f"Gamma errs {err_l:.1e} {err_r:.1e}", flush=True)
        assert err_l <= TOL * max(1.0, np.max(np.abs(gl_new))) and err_r <= TOL * max(1.0, np.max(np.abs(gr_new)))
