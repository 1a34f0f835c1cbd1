"""Multidimensional plans (kfr_dft_create_{2d,3d,md}_plan_*, kfr_dft_real_create_{2d,3d,md}_plan_*)
against FP64 numpy and the reference engine's dft_plan_md / dft_plan_md_real (gpu)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import kfr_amd
import oracle
from tests.common import bound_fp64

pytestmark = pytest.mark.gpu

# power-of-two shapes whose leading axes have 2^7..2^14 points run those axes in place along their
# stride: (1024, 128), (128, 256, 64), (256, 1, 512), (2048, 1024), (16384, 128) (f32; f64 stops at 2^13)
SHAPES = [(6, 8, 10), (32, 64), (3, 1, 5, 4), (101, 16), (1024, 128), (64, 64, 64), (7, 9, 11), (1, 30),
          (128, 256, 64), (256, 1, 512), (2048, 1024), (16384, 128)]


def _bound(shape, dtype):
    return bound_fp64(int(np.prod(shape)), dtype)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_md_complex(shape, dtype):
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    p = kfr_amd.dft_plan_md(shape, dtype)
    assert p.size == x.size
    ref_engine = oracle.reference()
    for inverse in (False, True):
        y = np.empty_like(x)
        p.execute(y, x, inverse=inverse)
        xx = x.astype(np.complex128)
        ref = np.fft.ifftn(xx) * x.size if inverse else np.fft.fftn(xx)
        err = oracle.rel_rms(y, ref)
        print(f"{shape} {np.dtype(dtype).name} inv={inverse} rel_rms={err:.3e}", flush=True)
        assert err <= _bound(shape, dtype), (shape, inverse, err)
        if ref_engine is not None:
            k, temp = ref_engine.dft_md(x, inverse)
            assert oracle.rel_rms(y, k) <= 2 * _bound(shape, dtype)
            assert p.temp_size == temp, (p.temp_size, temp)
    # in place
    z = x.copy()
    p.execute(z, z)
    assert oracle.rel_rms(z, np.fft.fftn(x.astype(np.complex128))) <= _bound(shape, dtype)


def test_md_complex_device_and_2d_3d_creators():
    torch = pytest.importorskip("torch")
    L = kfr_amd.lib()
    rng = np.random.default_rng(1)
    for shape in [(48, 20), (12, 10, 8)]:
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        h = L.kfr_dft_create_2d_plan_f32(*shape) if len(shape) == 2 else L.kfr_dft_create_3d_plan_f32(*shape)
        assert h, kfr_amd.last_error()
        assert L.kfr_dft_get_size_f32(h) == x.size
        xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
        yd = torch.empty_like(xd)
        L.kfr_dft_execute_f32(h, yd.data_ptr(), xd.data_ptr(), None)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(np.complex64).reshape(shape)
        assert oracle.rel_rms(y, np.fft.fftn(x.astype(np.complex128))) <= 1e-6
        L.kfr_dft_delete_plan_f32(h)


REAL_SHAPES = [(6, 8, 10), (32, 64), (16, 2), (101, 20), (3, 1, 4, 6), (128, 1024), (9, 30)]


@pytest.mark.parametrize("shape", REAL_SHAPES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("roe", [False, True])
def test_md_real(shape, dtype, roe):
    rng = np.random.default_rng(7 + sum(shape))
    x = rng.standard_normal(shape).astype(dtype)
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    p = kfr_amd.dft_plan_md_real(shape, roe, dtype)
    assert p.size == x.size
    X = np.empty(p.complex_size(), dtype=cdt)
    p.execute(X, x)
    ref = np.fft.rfftn(x.astype(np.float64))
    err = oracle.rel_rms(X, ref)
    print(f"{shape} {np.dtype(dtype).name} roe={roe} forward rel_rms={err:.3e}", flush=True)
    assert err <= _bound(shape, dtype), (shape, err)
    # inverse (unscaled): the real output buffer is sized as the reference requires with
    # real_out_is_enough (real_out_size = 2 * complex values)
    out = np.full(2 * X.size if roe else x.size, np.nan, dtype=dtype)
    p.execute(out, X)
    xr = out[: x.size].reshape(shape) / x.size
    assert oracle.rel_rms(xr, x) <= 4 * _bound(shape, dtype)
    ref_engine = oracle.reference()
    if ref_engine is not None:
        K, temp = ref_engine.rdft_md(x, roe)
        assert oracle.rel_rms(X, K) <= 2 * _bound(shape, dtype)
        assert p.temp_size == temp, (p.temp_size, temp)
        k, _ = ref_engine.irdft_md(X, shape, roe)
        assert oracle.rel_rms(out[: x.size].reshape(shape), k) <= 4 * _bound(shape, dtype)


def test_md_real_rejects_odd_last_axis_and_bad_shapes():
    L = kfr_amd.lib()
    assert not L.kfr_dft_real_create_2d_plan_f32(4, 9, False)
    assert "even" in kfr_amd.last_error()
    sh = (C.c_uint * 2)(4, 0)
    assert not L.kfr_dft_create_md_plan_f32(2, C.cast(sh, C.c_void_p))
    assert "positive" in kfr_amd.last_error()
    assert not L.kfr_dft_create_md_plan_f64(0, None)
    assert kfr_amd.last_error()


def test_md_dump_lists_every_axis(capfd):
    p = kfr_amd.dft_plan_md((12, 16), np.complex64)
    p.dump()
    out = capfd.readouterr().out
    assert "axis 0 (12)" in out and "axis 1 (16)" in out


def test_md_strided_device_batch_and_schedule():
    """device arrays through the batched entry point, two arrays in one call, both directions"""
    torch = pytest.importorskip("torch")
    shape = (512, 2048)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)).astype(np.complex64)
    p = kfr_amd.dft_plan_md(shape, np.complex64)
    xd = torch.from_numpy(x.view(np.float32).copy()).cuda()
    yd = torch.empty_like(xd)
    for inverse in (False, True):
        p.execute_batch(yd, xd, 2, inverse)
        torch.cuda.synchronize()
        y = yd.cpu().numpy().view(np.complex64).reshape((2,) + shape)
        xx = x.astype(np.complex128)
        ref = np.fft.ifftn(xx, axes=(1, 2)) * (512 * 2048) if inverse else np.fft.fftn(xx, axes=(1, 2))
        assert oracle.rel_rms(y, ref) <= _bound(shape, np.complex64)
