"""ctypes binding of libmvr_hip.so (the C ABI declared in include/mvr_hip.h).

The library is built in-tree by ``multiview-reconstruction_amd/build.py``.  Loading
fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libmvr_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "mvr_hip.h")

PSF_TYPES = {"OPTIMIZATION_II": 0, "OPTIMIZATION_I": 1, "EFFICIENT_BAYESIAN": 2, "INDEPENDENT": 3}
SCHEMES = {"SEQ": 0, "MUL": 1}
PASSES = ["x_fwd", "z_fwd", "y_conv", "z_inv", "x_ratio", "x_update", "x_store", "update_mul", "reduce"]

_lib = None
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)


class MvrError(RuntimeError):
    pass


def header_functions():
    """Names of all functions declared in include/mvr_hip.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mvr_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MvrError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP library has no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.mvr_last_error.restype = ctypes.c_char_p
        L.mvr_get_num_devices.restype = ctypes.c_int
        L.mvr_get_device_name.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.mvr_get_mem_device.argtypes = [ctypes.c_int]
        L.mvr_get_mem_device.restype = ctypes.c_longlong
        L.mvr_get_free_mem_device.argtypes = [ctypes.c_int]
        L.mvr_get_free_mem_device.restype = ctypes.c_longlong
        L.mvr_get_arch_major.argtypes = [ctypes.c_int]
        L.mvr_get_arch_minor.argtypes = [ctypes.c_int]
        L.mvr_convolution3d_fft_inplace.argtypes = [vp, _i32p, vp, _i32p, ctypes.c_int]
        L.mvr_fft_convolve.argtypes = [ctypes.c_int, vp, _i32p, ctypes.c_int, ctypes.c_float, vp, _i32p, vp]
        L.mvr_fft_lengths.argtypes = [vp, ctypes.c_int]
        L.mvr_fft_convolve_noalloc.argtypes = [ctypes.c_int, vp, _i32p, ctypes.c_int, ctypes.c_float, vp, _i32p, vp,
                                               vp]
        L.mvr_compute_quotient.argtypes = [ctypes.c_int, vp, vp, ctypes.c_longlong]
        L.mvr_compute_final_values.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_longlong, ctypes.c_double,
                                               ctypes.c_float, ctypes.c_float, _f64p]
        L.mvr_compute_final_values_mul.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(vp),
                                                   ctypes.POINTER(vp), ctypes.c_longlong, ctypes.c_double,
                                                   ctypes.c_float, ctypes.c_float, _f64p]
        L.mvr_decon_run_view.argtypes = [vp, ctypes.c_int]
        L.mvr_init_psfs.argtypes = [ctypes.c_int, vp, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, vp, vp]
        L.mvr_block_thread_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _i32p, ctypes.c_float,
                                              ctypes.c_float, ctypes.c_int]
        L.mvr_block_thread_set_view_kernels.argtypes = [vp, ctypes.c_int, vp, vp, _i32p]
        L.mvr_block_run_iteration_seq.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_float, _f64p]
        L.mvr_block_run_iteration_mul.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                                  _f32p, _f64p]
        L.mvr_block_thread_set_view_data.argtypes = [vp, ctypes.c_int, vp, vp, _i32p]
        L.mvr_block_run_iteration_seq_at.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_longlong),
                                                     ctypes.c_float, _f64p]
        L.mvr_block_run_iteration_mul_at.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_longlong), _f32p,
                                                     _f64p]
        L.mvr_block_thread_destroy.argtypes = [vp]
        L.mvr_block_thread_destroy.restype = None
        L.mvr_decon_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.mvr_decon_set_view.argtypes = [vp, ctypes.c_int, vp, vp, vp, _i32p]
        L.mvr_decon_set_view_kernels.argtypes = [vp, ctypes.c_int, vp, vp, _i32p]
        L.mvr_decon_set_max.argtypes = [vp, _f32p]
        L.mvr_decon_set_block_size.argtypes = [vp, _i32p]
        L.mvr_decon_get_blocks.argtypes = [vp, _i32p, ctypes.c_int]
        L.mvr_decon_init.argtypes = [vp, ctypes.c_int, vp]
        L.mvr_decon_run.argtypes = [vp, ctypes.c_int]
        L.mvr_decon_get_psi.argtypes = [vp, vp]
        L.mvr_decon_get_max.argtypes = [vp, _f32p, _f64p]
        L.mvr_decon_get_kernels.argtypes = [vp, ctypes.c_int, vp, vp]
        L.mvr_decon_get_stats.argtypes = [vp, _f64p, ctypes.c_int]
        L.mvr_decon_reset_psi.argtypes = [vp]
        L.mvr_decon_get_geometry.argtypes = [vp, _i32p]
        L.mvr_decon_profile.argtypes = [vp, ctypes.c_uint]
        L.mvr_decon_profile_sampled.argtypes = [vp, ctypes.c_uint, ctypes.c_int]
        L.mvr_decon_pass_times.argtypes = [vp, _f64p, ctypes.POINTER(ctypes.c_longlong)]
        L.mvr_decon_destroy.argtypes = [vp]
        L.mvr_decon_destroy.restype = None
        L.mvr_decon_set_reference.argtypes = [vp, vp]
        L.mvr_decon_rel_l2.argtypes = [vp, _f64p]
        L.mvr_decon_run_until.argtypes = [vp, ctypes.c_int, ctypes.c_double, _i32p, _f64p]
        L.mvr_decon_init_learned.argtypes = [vp, vp]
        L.mvr_decon_autotune.argtypes = [vp, ctypes.c_int, _i32p, _f64p]
        L.mvr_decon_use_graph.argtypes = [vp, ctypes.c_int]
        L.mvr_decon_get_tile_widths.argtypes = [vp, _i32p]
        L.mvr_decon_set_tile_widths.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.mvr_transform_view.argtypes = [ctypes.c_int, vp, _i32p, _f64p, ctypes.POINTER(ctypes.c_longlong), _i32p,
                                         ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, vp]
        L.mvr_transform_blending.argtypes = [ctypes.c_int, _i32p, _f32p, _f32p, _f64p,
                                             ctypes.POINTER(ctypes.c_longlong), _i32p, vp]
        L.mvr_fuse_group.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp), _i32p, _f64p, _f64p, _f32p,
                                     _f32p, ctypes.POINTER(ctypes.c_longlong), _i32p, vp, vp]
        L.mvr_normalize_weights.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, _i32p, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.mvr_transform_psf_size.argtypes = [_i32p, _f64p, _i32p]
        L.mvr_transform_psf.argtypes = [ctypes.c_int, vp, _i32p, _f64p, vp]
        L.mvr_unet_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong]
        L.mvr_unet_param_count.argtypes = [vp]
        L.mvr_unet_param_count.restype = ctypes.c_longlong
        L.mvr_unet_get_params.argtypes = [vp, vp]
        L.mvr_unet_set_params.argtypes = [vp, vp]
        L.mvr_unet_get_grads.argtypes = [vp, vp]
        L.mvr_unet_add_volume.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp, _i32p]
        L.mvr_unet_clear_volumes.argtypes = [vp]
        L.mvr_unet_train_step.argtypes = [vp, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                          ctypes.c_int, _f64p]
        L.mvr_unet_infer.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), _i32p, ctypes.c_float, vp]
        L.mvr_unet_set_inference_window.argtypes = [vp, ctypes.c_longlong]
        L.mvr_unet_inference_halo.argtypes = []
        L.mvr_unet_load.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_char_p]
        L.mvr_unet_read_weights.argtypes = [ctypes.c_char_p, vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int)]
        L.mvr_unet_read_weights.restype = ctypes.c_longlong
        L.mvr_unet_destroy.argtypes = [vp]
        L.mvr_unet_destroy.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise MvrError(lib().mvr_last_error().decode())
    return rc


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(shape_zyx):
    return (ctypes.c_int * 3)(shape_zyx[2], shape_zyx[1], shape_zyx[0])


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def num_devices() -> int:
    return lib().mvr_get_num_devices()


def device_name(dev=0) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(lib().mvr_get_device_name(dev, buf, 256))
    return buf.value.decode()


def read_weights(path):
    """(params float32 vector, views) of a weights file, parsed by the library on the host (no GPU)."""
    v = ctypes.c_int()
    n = lib().mvr_unet_read_weights(os.fsencode(path), None, 0, ctypes.byref(v))
    if n < 0:
        raise MvrError(lib().mvr_last_error().decode())
    out = np.empty(n, np.float32)
    _check(lib().mvr_unet_read_weights(os.fsencode(path), _ptr(out), n, ctypes.byref(v)))
    return out, v.value


def fft_lengths():
    """FFT lengths with compiled kernels (ascending)."""
    n = lib().mvr_fft_lengths(None, 0)
    out = (ctypes.c_int * n)()
    lib().mvr_fft_lengths(out, n)
    return list(out)


def fft_convolve(img, kernel, ext="mirror", value=0.0, dev=0):
    img = _f32(img)
    kernel = _f32(kernel)
    out = np.empty_like(img)
    e = 0 if ext == "mirror" else 1
    _check(lib().mvr_fft_convolve(dev, _ptr(img), _dims(img.shape), e, value, _ptr(kernel), _dims(kernel.shape),
                                  _ptr(out)))
    return out


def fft_convolve_noalloc(img_ptr, shape_zyx, kernel_ptr, kshape_zyx, out_ptr, ext="mirror", value=0.0, stream=None,
                         dev=0):
    """mvr_fft_convolve_noalloc on device pointers (e.g. torch tensors' data_ptr()); enqueued on `stream`
    (None = the default stream), no device memory allocated.  The caller synchronises before reading out."""
    e = 0 if ext == "mirror" else 1
    _check(lib().mvr_fft_convolve_noalloc(dev, ctypes.c_void_p(img_ptr), _dims(shape_zyx), e, value,
                                          ctypes.c_void_p(kernel_ptr), _dims(kshape_zyx), ctypes.c_void_p(out_ptr),
                                          ctypes.c_void_p(stream) if stream else None))


def convolution3d_fft_inplace(img, kernel, dev=0):
    """Same call shape as CUDAFourierConvolution.convolution3DfftCUDAInPlace (dims as z,y,x)."""
    img = _f32(img).copy()
    kernel = _f32(kernel)
    imd = (ctypes.c_int * 3)(*img.shape)
    kd = (ctypes.c_int * 3)(*kernel.shape)
    _check(lib().mvr_convolution3d_fft_inplace(_ptr(img), imd, _ptr(kernel), kd, dev))
    return img


def init_psfs(kernels, psf_type, nT=8, dev=0):
    """DeconViewPSF.init for all views: returns (kernel1 normalised, kernel2), each [V, kz, ky, kx]."""
    k = _f32(kernels)
    V = k.shape[0]
    k1 = np.empty_like(k)
    k2 = np.empty_like(k)
    _check(lib().mvr_init_psfs(dev, _ptr(k), V, _dims(k.shape[1:]), PSF_TYPES[psf_type], nT, _ptr(k1), _ptr(k2)))
    return k1, k2


def compute_quotient(img, blurred, dev=0):
    """DeconvolutionMethods.computeQuotient over a flat array (mvr_compute_quotient)."""
    img = _f32(img).ravel()
    out = _f32(blurred).ravel().copy()
    _check(lib().mvr_compute_quotient(dev, _ptr(img), _ptr(out), out.size))
    return out


def compute_final_values(psi, integral, weight, lam, min_i, max_i, dev=0):
    """DeconvolutionMethods.computeFinalValues: (new psi, (sumChange, maxChange))."""
    out = _f32(psi).ravel().copy()
    st = (ctypes.c_double * 2)()
    _check(lib().mvr_compute_final_values(dev, _ptr(out), _ptr(_f32(integral).ravel()), _ptr(_f32(weight).ravel()),
                                          out.size, float(lam), float(min_i), float(max_i), st))
    return out, (st[0], st[1])


def compute_final_values_mul(psi, integrals, weights, lam, min_i, max_i, dev=0):
    """DeconvolutionMethods.computeFinalValuesMul: integrals / weights are [V, n]."""
    out = _f32(psi).ravel().copy()
    ii = [_f32(a).ravel() for a in integrals]
    ww = [_f32(a).ravel() for a in weights]
    st = (ctypes.c_double * 2)()
    _check(lib().mvr_compute_final_values_mul(dev, _ptr(out), len(ii), _ptr_array(ii), _ptr_array(ww), out.size,
                                              float(lam), float(min_i), float(max_i), st))
    return out, (st[0], st[1])


class BlockThread:
    """ComputeBlockSeqThread / ComputeBlockMulThread running on one GPU."""

    def __init__(self, block_size_xyz, min_value=1e-4, lam=0.0, thread_id=0, dev=0):
        h = ctypes.c_void_p()
        _check(lib().mvr_block_thread_create(ctypes.byref(h), dev, (ctypes.c_int * 3)(*block_size_xyz), min_value,
                                             lam, thread_id))
        self.h = h
        self.block_size = tuple(block_size_xyz)

    def set_view_kernels(self, view, k1, k2):
        k1, k2 = _f32(k1), _f32(k2)
        _check(lib().mvr_block_thread_set_view_kernels(self.h, view, _ptr(k1), _ptr(k2), _dims(k1.shape)))

    def run_iteration(self, view, psi_block, img_block, weight_block, max_intensity):
        psi = _f32(psi_block).copy()
        st = (ctypes.c_double * 2)()
        _check(lib().mvr_block_run_iteration_seq(self.h, view, _ptr(psi), _ptr(_f32(img_block)),
                                                 _ptr(_f32(weight_block)), float(max_intensity), st))
        return psi, (st[0], st[1])

    def run_iteration_mul(self, psi_block, img_blocks, weight_blocks, max_intensities):
        psi = _f32(psi_block).copy()
        imgs = [_f32(a) for a in img_blocks]
        ws = [_f32(a) for a in weight_blocks]
        n = len(imgs)
        ip = (ctypes.c_void_p * n)(*[a.ctypes.data for a in imgs])
        wp = (ctypes.c_void_p * n)(*[a.ctypes.data for a in ws])
        mx = np.ascontiguousarray(max_intensities, np.float32)
        st = (ctypes.c_double * 2)()
        _check(lib().mvr_block_run_iteration_mul(self.h, n, _ptr(psi), ip, wp, mx.ctypes.data_as(_f32p), st))
        return psi, (st[0], st[1])

    def set_view_data(self, view, img, weight):
        """The view's whole image and weights, resident on the device for the thread's lifetime."""
        img, weight = _f32(img), _f32(weight)
        assert img.shape == weight.shape
        _check(lib().mvr_block_thread_set_view_data(self.h, view, _ptr(img), _ptr(weight), _dims(img.shape)))

    def run_iteration_at(self, view, psi_block, offset_xyz, max_intensity):
        """run_iteration with the image and weight blocks cut on the device from the resident view data;
        offset_xyz = volume coordinate of the block's first voxel (may be negative)."""
        psi = _f32(psi_block).copy()
        st = (ctypes.c_double * 2)()
        off = (ctypes.c_longlong * 3)(*[int(o) for o in offset_xyz])
        _check(lib().mvr_block_run_iteration_seq_at(self.h, view, _ptr(psi), off, float(max_intensity), st))
        return psi, (st[0], st[1])

    def run_iteration_mul_at(self, psi_block, offset_xyz, max_intensities):
        psi = _f32(psi_block).copy()
        mx = np.ascontiguousarray(max_intensities, np.float32)
        st = (ctypes.c_double * 2)()
        off = (ctypes.c_longlong * 3)(*[int(o) for o in offset_xyz])
        _check(lib().mvr_block_run_iteration_mul_at(self.h, mx.size, _ptr(psi), off, mx.ctypes.data_as(_f32p), st))
        return psi, (st[0], st[1])

    def close(self):
        if self.h:
            lib().mvr_block_thread_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Decon:
    """Resident MultiViewDeconvolution{Seq,Mul} (psi, views and spectra in HBM)."""

    def __init__(self, shape_zyx, n_views, psf_type="EFFICIENT_BAYESIAN", scheme="SEQ", lam=0.0, min_value=1e-4,
                 nT=8, dev=0):
        h = ctypes.c_void_p()
        _check(lib().mvr_decon_create(ctypes.byref(h), dev, _dims(shape_zyx), n_views, PSF_TYPES[psf_type],
                                      SCHEMES[scheme], lam, min_value, nT))
        self.h = h
        self.shape = tuple(shape_zyx)
        self.V = n_views
        self.scheme = scheme

    def set_view(self, view, img, weight, kernel):
        img, weight, kernel = _f32(img), _f32(weight), _f32(kernel)
        assert img.shape == self.shape and weight.shape == self.shape
        _check(lib().mvr_decon_set_view(self.h, view, _ptr(img), _ptr(weight), _ptr(kernel), _dims(kernel.shape)))

    def set_view_device(self, view, img_ptr, weight_ptr, kernel, ):
        kernel = _f32(kernel)
        _check(lib().mvr_decon_set_view(self.h, view, ctypes.c_void_p(img_ptr), ctypes.c_void_p(weight_ptr),
                                        _ptr(kernel), _dims(kernel.shape)))

    def set_view_kernels(self, view, kernel1, kernel2):
        """Final kernels of a view whose DeconViewPSF.init ran on the caller's side (used as they are by init)."""
        k1, k2 = _f32(kernel1), _f32(kernel2)
        _check(lib().mvr_decon_set_view_kernels(self.h, view, _ptr(k1), _ptr(k2), _dims(k1.shape)))

    def set_max(self, max_per_view):
        """PsiInit.getMax of an initialiser that ran on the caller's side."""
        mx = np.ascontiguousarray(max_per_view, np.float32)
        assert mx.size == self.V
        _check(lib().mvr_decon_set_max(self.h, mx.ctypes.data_as(_f32p)))

    def set_block_size(self, block_size_xyz):
        """The reference's block structure (DeconView block size B per axis, x y z); (0, 0, 0): one frame."""
        _check(lib().mvr_decon_set_block_size(self.h, (ctypes.c_int * 3)(*[int(b) for b in block_size_xyz])))

    def blocks(self):
        """(blocks in the volume, [blocks with content per view])."""
        per = (ctypes.c_int * self.V)()
        n = _check(lib().mvr_decon_get_blocks(self.h, per, self.V))
        return n, list(per)

    PSI_INITS = {"AVG_PRECISE": 0, "FROM_ARRAY": 1, "AVG_APPROX": 2, "BLURRED_FUSED": 3}

    def init(self, psi0=None, mode="AVG_PRECISE"):
        """PsiInit: AVG_PRECISE (default), AVG_APPROX, BLURRED_FUSED, or FROM_ARRAY when psi0 is given."""
        if psi0 is None:
            _check(lib().mvr_decon_init(self.h, self.PSI_INITS[mode], None))
        else:
            p = _f32(psi0)
            _check(lib().mvr_decon_init(self.h, 1, _ptr(p)))

    def run(self, iterations):
        _check(lib().mvr_decon_run(self.h, iterations))

    def run_view(self, view):
        """One Seq view update of psi (MultiViewDeconvolutionSeq's per-view pass)."""
        _check(lib().mvr_decon_run_view(self.h, view))

    def psi(self):
        out = np.empty(self.shape, np.float32)
        _check(lib().mvr_decon_get_psi(self.h, _ptr(out)))
        return out

    def max_avg(self):
        mx = np.zeros(self.V, np.float32)
        avg = ctypes.c_double()
        _check(lib().mvr_decon_get_max(self.h, mx.ctypes.data_as(_f32p), ctypes.byref(avg)))
        return mx, avg.value

    def kernels(self, view, kshape_zyx):
        k1 = np.zeros(kshape_zyx, np.float32)
        k2 = np.zeros(kshape_zyx, np.float32)
        _check(lib().mvr_decon_get_kernels(self.h, view, _ptr(k1), _ptr(k2)))
        return k1, k2

    def stats(self):
        n = _check(lib().mvr_decon_get_stats(self.h, None, 0))
        buf = np.zeros(max(n, 1), np.float64)
        lib().mvr_decon_get_stats(self.h, buf.ctypes.data_as(_f64p), n)
        per = self.V if self.scheme == "SEQ" else 1
        if n % (2 * per):        # single view updates (run_view) that do not fill whole iterations
            return buf[:n].reshape(-1, 2)
        return buf[:n].reshape(-1, per, 2) if n else np.zeros((0, per, 2))

    def reset_psi(self):
        _check(lib().mvr_decon_reset_psi(self.h))

    def geometry(self):
        g = (ctypes.c_int * 4)()
        _check(lib().mvr_decon_get_geometry(self.h, g))
        tw = (ctypes.c_int * 2)()
        _check(lib().mvr_decon_get_tile_widths(self.h, tw))
        return dict(P=(g[0], g[1], g[2]), row_pitch=g[3], tile_widths_yz=(tw[0], tw[1]))

    def profile(self, passes=None, every=1):
        """Time the given passes (names of PASSES; None = all, [] = off), the first launch of every group of
        `every` launches of each; clears the totals."""
        if passes is None:
            mask = (1 << len(PASSES)) - 1
        else:
            mask = 0
            for p in passes:
                mask |= 1 << PASSES.index(p)
        if every == 1:
            _check(lib().mvr_decon_profile(self.h, mask))
        else:
            _check(lib().mvr_decon_profile_sampled(self.h, mask, int(every)))

    def pass_times(self):
        ms = (ctypes.c_double * 9)()
        n = (ctypes.c_longlong * 9)()
        _check(lib().mvr_decon_pass_times(self.h, ms, n))
        return {PASSES[i]: (ms[i], n[i]) for i in range(9)}

    def autotune(self, candidates=3):
        """Times candidate FFT lengths per axis and keeps the fastest; returns ((Px, Py, Pz), ms per update)."""
        ch = (ctypes.c_int * 3)()
        ms = ctypes.c_double()
        _check(lib().mvr_decon_autotune(self.h, candidates, ch, ctypes.byref(ms)))
        return (ch[0], ch[1], ch[2]), ms.value

    def set_tile_widths(self, wy, wz):
        """Tile widths (columns) of the y and z passes: 0 (default for the length), 8 or 16."""
        _check(lib().mvr_decon_set_tile_widths(self.h, int(wy), int(wz)))

    def use_graph(self, enable=True):
        _check(lib().mvr_decon_use_graph(self.h, 1 if enable else 0))

    def init_learned(self, net: "UNet"):
        """PsiInit from the network (PsiInitFromRAI pattern: max[] and avg from the AvgPrecise pass)."""
        _check(lib().mvr_decon_init_learned(self.h, net.h))

    def set_reference(self, ref):
        r = _f32(ref)
        assert r.shape == self.shape
        _check(lib().mvr_decon_set_reference(self.h, _ptr(r)))

    def rel_l2(self):
        out = ctypes.c_double()
        _check(lib().mvr_decon_rel_l2(self.h, ctypes.byref(out)))
        return out.value

    def run_until(self, tol, max_iterations):
        """Iterate until rel-L2 to the reference <= tol; returns (iterations, rel-L2)."""
        it = ctypes.c_int()
        rel = ctypes.c_double()
        _check(lib().mvr_decon_run_until(self.h, max_iterations, tol, ctypes.byref(it), ctypes.byref(rel)))
        return it.value, rel.value

    def close(self):
        if self.h:
            lib().mvr_decon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr_array(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class UNet:
    """Learned initial estimate: 3-level 3D U-Net (16/32/64 channels) trained and run by the HIP kernels.

    Parameters are one flat FP32 vector: per conv layer the weights [27][cout][cin] then the bias,
    last the 1x1 head (16 weights, bias).  Volumes are numpy arrays in z, y, x order.
    """

    LAYERS_CIN = (16, 16, 16, 32, 32, 64, 96, 32, 48, 16)
    LAYERS_COUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16)

    @staticmethod
    def inference_halo():
        """Planes of context on either side of an inference slab."""
        return int(lib().mvr_unet_inference_halo())

    @classmethod
    def load(cls, path, dev=0):
        """A trained network from its weights file (read by the library: mvr_unet_load)."""
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        _check(lib().mvr_unet_load(ctypes.byref(h), dev, os.fsencode(path)))
        self.h = h
        self.V = read_weights(path)[1]
        return self

    def __init__(self, n_views, seed=0, dev=0):
        h = ctypes.c_void_p()
        _check(lib().mvr_unet_create(ctypes.byref(h), dev, n_views, seed))
        self.h = h
        self.V = n_views
        self._keep = []

    def param_count(self):
        return _check(lib().mvr_unet_param_count(self.h))

    def params(self):
        out = np.empty(self.param_count(), np.float32)
        _check(lib().mvr_unet_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.param_count()
        _check(lib().mvr_unet_set_params(self.h, _ptr(p)))

    def grads(self):
        out = np.empty(self.param_count(), np.float32)
        _check(lib().mvr_unet_get_grads(self.h, _ptr(out)))
        return out

    def add_volume(self, imgs, weights, target):
        imgs = [_f32(a) for a in imgs]
        weights = [_f32(a) for a in weights]
        target = _f32(target)
        assert len(imgs) == self.V and len(weights) == self.V
        return _check(lib().mvr_unet_add_volume(self.h, _ptr_array(imgs), _ptr_array(weights), _ptr(target),
                                                _dims(target.shape)))

    def clear_volumes(self):
        _check(lib().mvr_unet_clear_volumes(self.h))

    def train_step(self, vol_idx, origins_zyx, patch, lr, update=True):
        """One step on patches; origins are (z, y, x) per patch.  Returns the loss (normalised MSE)."""
        vi = np.ascontiguousarray(vol_idx, np.int32)
        o = np.ascontiguousarray(np.asarray(origins_zyx, np.int32)[:, ::-1])  # -> x, y, z
        loss = ctypes.c_double()
        _check(lib().mvr_unet_train_step(self.h, vi.ctypes.data_as(_i32p), o.ctypes.data_as(_i32p), len(vi), patch,
                                         lr, 1 if update else 0, ctypes.byref(loss)))
        return loss.value

    def infer(self, imgs, weights, min_value=1e-4):
        imgs = [_f32(a) for a in imgs]
        weights = [_f32(a) for a in weights]
        out = np.empty(imgs[0].shape, np.float32)
        _check(lib().mvr_unet_infer(self.h, _ptr_array(imgs), _ptr_array(weights), _dims(out.shape), min_value,
                                    _ptr(out)))
        return out

    def set_inference_window(self, voxels):
        """Voxel budget of one inference window; larger volumes run in z slabs with inference_halo()
        planes of context on either side (<= 0 restores the default, 2^28)."""
        _check(lib().mvr_unet_set_inference_window(self.h, int(voxels)))

    def close(self):
        if self.h:
            lib().mvr_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- input preparation
def _affine(a):
    a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(3, 4))
    return a.ctypes.data_as(_f64p), a


def transform_view(img, affine, bb_min_xyz, bb_shape_zyx, min_value=1.0, has_min=True, outside=0.0,
                   interpolation=1, dev=0):
    """TransformView.transformView: the view resampled into the bounding box (z, y, x array).
    affine: 3 x 4 input pixel (x, y, z) -> world map; bb_min_xyz: world minimum of the box."""
    img = _f32(img)
    out = np.empty(tuple(bb_shape_zyx), np.float32)
    ap, keep = _affine(affine)
    bb = (ctypes.c_longlong * 3)(*[int(v) for v in bb_min_xyz])
    _check(lib().mvr_transform_view(dev, _ptr(img), _dims(img.shape), ap, bb, _dims(out.shape), 1 if has_min else 0,
                                    min_value, outside, interpolation, _ptr(out)))
    return out


def transform_blending(img_shape_zyx, border_xyz, blending_xyz, affine, bb_min_xyz, bb_shape_zyx, dev=0):
    """TransformWeight.transformBlending: blending weights of a view in the bounding box (z, y, x array)."""
    out = np.empty(tuple(bb_shape_zyx), np.float32)
    ap, keep = _affine(affine)
    bb = (ctypes.c_longlong * 3)(*[int(v) for v in bb_min_xyz])
    bo = (ctypes.c_float * 3)(*border_xyz)
    bl = (ctypes.c_float * 3)(*blending_xyz)
    _check(lib().mvr_transform_blending(dev, _dims(img_shape_zyx), bo, bl, ap, bb, _dims(out.shape), _ptr(out)))
    return out


def fuse_group(images, affines, bb_min_xyz, bb_shape_zyx, intensity_adjust=None, fusion_blending=None,
               decon_blending=None, dev=0):
    """ProcessInputImages.fuseGroups for one group: (fused image, summed decon weights), z, y, x arrays.

    images / affines: the group's views and their registrations; intensity_adjust: per view (m00, m01) or None;
    fusion_blending / decon_blending: per view (border_xyz, range_xyz) or None (constant weight 1)."""
    imgs = [_f32(a) for a in images]
    n = len(imgs)
    dims = np.ascontiguousarray([[a.shape[2], a.shape[1], a.shape[0]] for a in imgs], np.int32)
    aff = np.ascontiguousarray([np.asarray(a, np.float64).reshape(12) for a in affines])

    def arr(x, dtype, k):
        if x is None:
            return None, None
        a = np.ascontiguousarray(np.asarray(x, dtype).reshape(n, k))
        return a.ctypes.data_as(_f64p if dtype == np.float64 else _f32p), a

    adj_p, adj = arr(intensity_adjust, np.float64, 2)
    fb_p, fb = arr(fusion_blending, np.float32, 6)
    db_p, db = arr(decon_blending, np.float32, 6)
    out = np.empty(tuple(bb_shape_zyx), np.float32)
    wout = np.empty_like(out)
    bb = (ctypes.c_longlong * 3)(*[int(v) for v in bb_min_xyz])
    _check(lib().mvr_fuse_group(dev, n, _ptr_array(imgs), dims.ctypes.data_as(_i32p), aff.ctypes.data_as(_f64p),
                                adj_p, fb_p, db_p, bb, _dims(out.shape), _ptr(out), _ptr(wout)))
    return out, wout


def normalize_weights(weights, osem_speedup=1.0, smooth=False, max_diff_range=0.1, scaling_range=0.05, dev=0):
    """NormalizingRandomAccess for all views; returns the normalised weights [V, z, y, x]."""
    w = np.array(weights, np.float32, copy=True)
    ptrs = (ctypes.c_void_p * w.shape[0])(*[w[v].ctypes.data for v in range(w.shape[0])])
    _check(lib().mvr_normalize_weights(dev, ptrs, w.shape[0], _dims(w.shape[1:]), osem_speedup, 1 if smooth else 0,
                                       max_diff_range, scaling_range))
    return w


def transform_psf(psf, affine, dev=0):
    """PSFExtraction.getTransformedNormalizedPSF: the normalised PSF resampled by the view's affine."""
    psf = _f32(psf)
    ap, keep = _affine(affine)
    o = (ctypes.c_int * 3)()
    _check(lib().mvr_transform_psf_size(_dims(psf.shape), ap, o))
    out = np.empty((o[2], o[1], o[0]), np.float32)
    _check(lib().mvr_transform_psf(dev, _ptr(psf), _dims(psf.shape), ap, _ptr(out)))
    return out
