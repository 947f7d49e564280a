"""ctypes binding of libmemi.so (the HIP kernels) for torch device tensors.

Every wrapper takes torch tensors that must already live on the GPU, enqueues on torch's
current stream and raises ``RuntimeError`` with the library's message on a non-zero status.
There is no CPU fallback: if the library is missing or a tensor is on the CPU the call fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmemi.so")

PREC_FP32 = 0
PREC_SPLIT_BF16 = 1
FILTER_BUTTER = 1
FILTER_GAUSS = 2
FILTER_BSPLINE = 4

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class PdDesc(ctypes.Structure):
    _fields_ = [("row0", c_i64), ("d_off", c_i64), ("mask_off", c_i64), ("list_off", c_i64),
                ("n", c_i32), ("k", c_i32)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("n_pd", c_i64), ("n_items", c_i64), ("n_utiles", c_i64), ("n_ftiles", c_i64),
                ("n_rows", c_i64), ("off_pds", c_i64), ("off_items", c_i64),
                ("off_utiles", c_i64), ("off_ftiles", c_i64), ("off_rows", c_i64),
                ("plan_bytes", c_i64), ("d_elems", c_i64), ("mask_words", c_i64),
                ("list_elems", c_i64), ("ws_floats", c_i64), ("flag_cap", c_i64),
                ("chunks", c_i32), ("max_n", c_i32), ("gram_tile", c_i32),
                ("reserved", c_i32)]


class CtfParams(ctypes.Structure):
    _fields_ = [("pix_size", c_dbl), ("cs_mm", c_dbl), ("ekv", c_dbl), ("ampc", c_dbl),
                ("gauss_env", c_dbl)]


# name -> argument types (all return int status)
_SIGS = {
    "memi_load_normalize": [c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_vp],
    "memi_spectrum_scale": [c_vp, c_i64, c_int, c_int, c_dbl, c_int, c_vp],
    "memi_mask_multiply": [c_vp, c_i64, c_int, c_int, c_vp, c_vp],
    "memi_spline_rotate": [c_vp, c_vp, c_i64, c_int, c_vp, c_vp],
    "memi_align_images": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp],
    "memi_prep_images": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dbl,
                         c_int, ctypes.POINTER(CtfParams), c_int, c_i64, c_i64, c_vp, c_vp, c_vp,
                         c_vp, c_vp, c_vp, c_i64, c_vp],
    "memi_shift_images": [c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_i64, c_vp],
    "memi_project_mask": [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp],
    "memi_prep_images_noalloc": [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int,
                                 c_dbl, c_int, ctypes.POINTER(CtfParams), c_int, c_i64, c_i64,
                                 c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_ctf_pack": [c_vp, c_vp, c_i64, c_int, ctypes.POINTER(CtfParams), c_int, c_i64, c_i64,
                      c_vp, c_vp, c_vp, c_vp],
    "memi_phase_flip": [c_vp, c_vp, c_i64, c_int, c_vp, c_vp],
    "memi_wiener_average": [c_vp, c_vp, c_i64, c_int, c_dbl, c_vp, c_vp],
    "memi_image_moments": [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp],
    "memi_plan": [c_vp, c_int, c_int, c_int, c_vp, c_i64, ctypes.POINTER(PlanInfo)],
    "memi_ctf_distance": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_i64, c_int, c_dbl, c_vp,
                          c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp],
    "memi_knn_select": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_ferguson_sweep": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    "memi_ferguson_moments": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_int, c_dbl,
                              c_dbl, c_vp, c_vp, c_i64, c_vp],
    "memi_markov_normalize": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_vp, c_int,
                              c_vp, c_vp],
    "memi_tanh_fit": [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_int, c_vp, c_vp],
    "memi_gather_blocks": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_nearest_bins": [c_vp, c_i64, c_vp, c_int, c_vp, c_vp],
    "memi_ferguson_values": [c_vp, c_i64, c_vp, c_int, c_vp, c_vp],
    "memi_markov_coo": [c_vp, c_vp, c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp, c_vp, c_vp],
    "memi_lanczos_steps": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_int, c_vp, c_int, c_vp, c_int,
                           c_vp, c_vp, c_vp, c_vp, c_int, c_vp],
    "memi_ritz_vectors": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_int, c_vp, c_vp, c_int, c_vp,
                          c_vp],
    "memi_lanczos_ritz": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_int, c_int, c_vp, c_dbl, c_vp,
                          c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_nlsa_condist": [c_vp, c_int, c_vp, c_int, c_int, c_vp, c_vp],
    "memi_nlsa_supervectors": [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_int, c_dbl, c_vp, c_vp],
    "memi_tall_gram": [c_vp, c_i64, c_int, c_vp, c_vp],
    "memi_tall_apply": [c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp],
    "memi_nlsa_imgt": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp],
    "memi_center_gram": [c_vp, c_i64, c_int, c_vp, c_vp],
    "memi_nlsa_recon_dist": [c_vp, c_vp, c_int, c_int, c_i64, c_vp, c_vp, c_vp, c_vp],
    "memi_fit1d_tau": [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_kernel_rowsums": [c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp],
    "memi_kernel_grad_weights": [c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "memi_lanczos_init": [c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp],
    "memi_kernel_ell": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "memi_lanczos_steps_sparse": [c_vp, ctypes.POINTER(PlanInfo), c_vp, c_vp, c_vp, c_vp, c_vp,
                                  c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_int, c_vp],
    "memi_tanh_select": [c_vp, c_int, c_int, c_int, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp],
    "memi_flow_anisodiff": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_int, c_int, c_dbl, c_dbl,
                            c_dbl, c_dbl, c_dbl, c_int, c_vp],
    "memi_flow_block_average": [c_vp, c_i64, c_int, c_i64, c_int, c_vp, c_vp],
    "memi_gaussian_filter": [c_vp, c_int, c_vp, c_int, c_vp, c_i64, c_int, c_int, c_dbl, c_vp],
    "memi_hs_derivatives": [c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp],
    "memi_hs_iterate": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dbl, c_int, c_vp, c_vp, c_int,
                        c_vp, c_vp, c_vp],
    "memi_flow_sum": [c_vp, c_vp, c_int, c_i64, c_int, c_i64, c_int, c_vp, c_vp, c_vp, c_vp,
                      c_vp, c_vp],
}
EXPORTS = ["memi_version", "memi_last_error"] + list(_SIGS)

_lib = None


def load():
    """Load libmemi.so once (no GPU needed); raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m manifoldem_python_amd.build` (hipcc, gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.memi_version.restype = c_int
        lib.memi_last_error.restype = ctypes.c_char_p
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = c_int
        _lib = lib
    return _lib


def _check(status, name):
    if status != 0:
        msg = load().memi_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({status}): {msg}")


def _p(t):
    """Device pointer of a CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError("libmemi needs GPU tensors; got "
                           f"{type(t).__name__} on {getattr(t, 'device', 'host')}")
    if not t.is_contiguous():
        raise RuntimeError("libmemi needs contiguous tensors")
    return c_vp(t.data_ptr())


def _stream():
    return c_vp(torch.cuda.current_stream().cuda_stream)


def call(name, *args):
    _check(getattr(load(), name)(*args), name)


def _need(cond, msg):
    """Host-side size check before any launch: a mismatch raises instead of reaching a kernel."""
    if not cond:
        raise RuntimeError(msg)


def _numel(t):
    return 0 if t is None else t.numel()


def _bytes(t):
    return 0 if t is None else t.numel() * t.element_size()


# ------------------------------------------------------------------------- Stage 1 ---------
def load_normalize(stack, src, conj, N, out, src_max=None):
    """src_max: host-known largest index in src (checked against the stack); when None the
    indices are read back from the device to check them (one synchronisation)."""
    if src.numel():
        hi = int(src.max()) if src_max is None else int(src_max)
        lo = int(src.min()) if src_max is None else 0
        _need(lo >= 0 and hi < stack.shape[0], "load_normalize: image index outside the stack")
    _need(tuple(stack.shape[1:]) == (N, N) and tuple(out.shape) == (src.shape[0], N, N) and
          stack.dtype == torch.float32 and src.dtype == torch.int64 and
          (conj is None or (conj.dtype == torch.int32 and conj.numel() == src.numel())),
          "load_normalize: stack/src/conj/out shapes or dtypes do not match")
    call("memi_load_normalize", _p(stack), stack.shape[0], _p(src), _p(conj), src.shape[0], N,
         _p(out), _stream())


def spectrum_scale(spec, N, kind, Qc=1.0, order=1):
    _need(spec.dim() == 3 and tuple(spec.shape[1:]) == (N, N // 2 + 1) and spec.is_complex()
          and spec.dtype == torch.complex64, "spectrum_scale: spec must be complex64 (n, N, N/2+1)")
    call("memi_spectrum_scale", _p(spec), spec.shape[0], N, kind, float(Qc), int(order), _stream())


def mask_multiply(img, N, disk=True, msk2=None):
    _need(img.dim() == 3 and tuple(img.shape[1:]) == (N, N) and img.dtype == torch.float32,
          "mask_multiply: img must be float32 (n, N, N)")
    _need(msk2 is None or (msk2.numel() == N * N and msk2.dtype == torch.float32),
          "mask_multiply: msk2 must be float32 with N*N elements")
    call("memi_mask_multiply", _p(img), img.shape[0], N, int(bool(disk)), _p(msk2), _stream())


def spline_rotate(coef, angle_deg, N, out):
    n = coef.shape[0]
    _need(coef.dim() == 3 and tuple(coef.shape[1:]) == (N, N) and coef.dtype == torch.float32,
          "spline_rotate: coef must be float32 (n, N, N)")
    _need(out.shape == coef.shape and out.dtype == torch.float32, "spline_rotate: out shape")
    _need(angle_deg.numel() == n and angle_deg.dtype == torch.float64,
          "spline_rotate: angle_deg must be float64 (n,)")
    call("memi_spline_rotate", _p(coef), _p(angle_deg), coef.shape[0], N, _p(out), _stream())


ALIGN_MIN_N, ALIGN_MAX_N = 8, 196


def align_images(img, rot1_deg, rot2_deg, N, out, disk=True):
    """Disk mask (optional) then the two spline rotations, one fused kernel; out may be img."""
    n = img.shape[0]
    _need(img.dim() == 3 and tuple(img.shape[1:]) == (N, N) and img.dtype == torch.float32,
          "align_images: img must be float32 (n, N, N)")
    _need(out.shape == img.shape and out.dtype == torch.float32, "align_images: out shape")
    _need(rot1_deg.numel() == n and rot2_deg.numel() == n and rot1_deg.dtype == torch.float64
          and rot2_deg.dtype == torch.float64, "align_images: angles must be float64 (n,)")
    _need(ALIGN_MIN_N <= N <= ALIGN_MAX_N,
          f"align_images: N={N} outside [{ALIGN_MIN_N}, {ALIGN_MAX_N}]")
    call("memi_align_images", _p(img), _p(rot1_deg), _p(rot2_deg), n, N, int(bool(disk)),
         _p(out), _stream())


PREP_MAX_N = 160


def prep_supported(N):
    """True when memi_prep_images handles side N (8..160, prime factors 2, 3, 5, 7)."""
    if not 8 <= N <= PREP_MAX_N:
        return False
    n = N
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def prep_images(stack, src, conj, rot1, rot2, df, N, filter_kind, Qc, order, ctf, precision,
                Khp, row0, op_hi, op_lo, fy, ctf_half, ws, img_out=None, src_max=None):
    """memi_prep_images: a1-a6 for len(src) images in one kernel (see include/memi.h)."""
    n = int(src.shape[0])
    Kh = N * (N // 2 + 1)
    _need(prep_supported(N), f"prep_images: N={N} not supported by the fused kernel")
    _need(stack.dtype == torch.float32 and tuple(stack.shape[1:]) == (N, N),
          "prep_images: stack must be float32 (n_stack, N, N)")
    _need(src.dtype == torch.int64 and (conj is None or (conj.dtype == torch.int32 and
                                                        conj.numel() == n)),
          "prep_images: src int64 / conj int32")
    for t, name in ((rot1, "rot1"), (rot2, "rot2"), (df, "df")):
        _need(t.dtype == torch.float64 and t.numel() == n, f"prep_images: {name} float64 (n,)")
    if n:
        hi = int(src.max()) if src_max is None else int(src_max)
        lo = int(src.min()) if src_max is None else 0
        _need(lo >= 0 and hi < stack.shape[0], "prep_images: image index outside the stack")
    esz = 4 if precision == PREC_FP32 else 2
    _need(_bytes(op_hi) >= (row0 + n) * 4 * Khp * esz and
          (precision == PREC_FP32 or _bytes(op_lo) >= (row0 + n) * 4 * Khp * esz),
          "prep_images: operand smaller than row0 + n rows")
    _need(fy.dtype == torch.complex64 and fy.numel() >= n * Kh and ctf_half.numel() >= n * Kh,
          "prep_images: fy / ctf sizes")
    _need(ws.dtype == torch.float64 and ws.numel() >= 6 * Kh, "prep_images: ws too small")
    _need(img_out is None or (img_out.dtype == torch.float32 and img_out.numel() >= n * N * N),
          "prep_images: img_out size")
    call("memi_prep_images", _p(stack), stack.shape[0], _p(src), None if conj is None else _p(conj),
         _p(rot1), _p(rot2), _p(df), n, N, int(filter_kind), float(Qc), int(order),
         ctypes.byref(ctf), int(precision), int(Khp), int(row0), _p(op_hi),
         _p(op_lo) if op_lo is not None else None, _p(fy), _p(ctf_half),
         None if img_out is None else _p(img_out), _p(ws), ws.numel(), _stream())


def prep_images_noalloc(stack, src, conj, rot1, rot2, df, N, filter_kind, Qc, order, ctf,
                        precision, Khp, row0, op_hi, op_lo, fy, ctf_half, img_out=None):
    """memi_prep_images_noalloc: prep_images with the tables in the library's static device
    scratch instead of a workspace (calls that overlap on one GPU race; see include/memi.h)."""
    n = int(src.shape[0])
    Kh = N * (N // 2 + 1)
    _need(prep_supported(N), f"prep_images_noalloc: N={N} not supported by the fused kernel")
    _need(stack.dtype == torch.float32 and tuple(stack.shape[1:]) == (N, N),
          "prep_images_noalloc: stack must be float32 (n_stack, N, N)")
    _need(src.dtype == torch.int64 and (conj is None or (conj.dtype == torch.int32 and
                                                        conj.numel() == n)),
          "prep_images_noalloc: src int64 / conj int32")
    for t, name in ((rot1, "rot1"), (rot2, "rot2"), (df, "df")):
        _need(t.dtype == torch.float64 and t.numel() == n,
              f"prep_images_noalloc: {name} float64 (n,)")
    if n:
        _need(int(src.min()) >= 0 and int(src.max()) < stack.shape[0],
              "prep_images_noalloc: image index outside the stack")
    esz = 4 if precision == PREC_FP32 else 2
    _need(_bytes(op_hi) >= (row0 + n) * 4 * Khp * esz and
          (precision == PREC_FP32 or _bytes(op_lo) >= (row0 + n) * 4 * Khp * esz),
          "prep_images_noalloc: operand smaller than row0 + n rows")
    _need(fy.dtype == torch.complex64 and fy.numel() >= n * Kh and ctf_half.numel() >= n * Kh,
          "prep_images_noalloc: fy / ctf sizes")
    _need(img_out is None or (img_out.dtype == torch.float32 and img_out.numel() >= n * N * N),
          "prep_images_noalloc: img_out size")
    call("memi_prep_images_noalloc", _p(stack), stack.shape[0], _p(src),
         None if conj is None else _p(conj), _p(rot1), _p(rot2), _p(df), n, N, int(filter_kind),
         float(Qc), int(order), ctypes.byref(ctf), int(precision), int(Khp), int(row0),
         _p(op_hi), _p(op_lo) if op_lo is not None else None, _p(fy), _p(ctf_half),
         None if img_out is None else _p(img_out), _stream())


_SPLINE_LU = {}


def spline_lu(N, boundary="mirror"):
    """LU factors {m, 1/pivot} (float64 [2N]) of the cubic B-spline coefficient system
    (c[k-1] + 4 c[k] + c[k+1]) = 6 x[k] with the whole-sample mirror extension (c[-1] = c[1],
    c[N] = c[N-2]; boundary rows [4 2] / [2 4]) or the half-sample reflection (c[-1] = c[0],
    c[N] = c[N-1]; rows [5 1] / [1 5])."""
    key = (N, boundary)
    lu = _SPLINE_LU.get(key)
    if lu is None:
        if boundary == "mirror":
            d0, up0, low_last, dlast = 4.0, 2.0, 2.0, 4.0
        elif boundary == "reflect":
            d0, up0, low_last, dlast = 5.0, 1.0, 1.0, 5.0
        else:
            raise ValueError(boundary)
        m = np.zeros(N)
        piv = np.zeros(N)
        piv[0] = d0
        for k in range(1, N):
            a = low_last if k == N - 1 else 1.0
            cprev = up0 if k == 1 else 1.0
            m[k] = a / piv[k - 1]
            piv[k] = (dlast if k == N - 1 else 4.0) - m[k] * cprev
        lu = _SPLINE_LU[key] = np.concatenate([m, 1.0 / piv])
    return lu


MASK_NPAD = 12


def project_mask(vol_swapped, T):
    """memi_project_mask: (proj float64 (L, L), mask float32 (L, L)) of an axis-swapped volume
    (float32 (L, L, L) device tensor) under the affine map T (float64 (12,) host array)."""
    L = int(vol_swapped.shape[0])
    _need(vol_swapped.dtype == torch.float32 and tuple(vol_swapped.shape) == (L, L, L) and
          vol_swapped.is_contiguous(), "project_mask: volume must be contiguous float32 (L,L,L)")
    dev = vol_swapped.device
    Lp = L + 2 * MASK_NPAD
    lu = torch.as_tensor(spline_lu(Lp, "reflect"), device=dev)
    Td = torch.as_tensor(np.asarray(T, dtype=np.float64).reshape(12), device=dev)
    proj = torch.empty((L, L), dtype=torch.float64, device=dev)
    mask = torch.empty((L, L), dtype=torch.float32, device=dev)
    ws = torch.empty(Lp ** 3, dtype=torch.float64, device=dev)
    call("memi_project_mask", _p(vol_swapped), L, _p(Td), _p(lu), _p(proj), _p(mask), _p(ws),
         ws.numel(), _stream())
    return proj, mask


def shift_images(img, shift0, shift1, out, transpose_out=True, chunk=8192):
    """memi_shift_images in chunks of images (bounded float64 scratch): out[i] =
    ndimage.shift(img[i], (shift0[i], shift1[i]), order=3, mode='wrap'), transposed when
    transpose_out.  img, out: float32 (n, N, N) device tensors; shift0/1: float64 (n,)."""
    n, N = int(img.shape[0]), int(img.shape[1])
    _need(img.dtype == torch.float32 and tuple(img.shape[1:]) == (N, N) and img.is_contiguous(),
          "shift_images: img must be contiguous float32 (n, N, N)")
    _need(out.dtype == torch.float32 and out.shape == img.shape and out.is_contiguous(),
          "shift_images: out must match img")
    _need(out.data_ptr() != img.data_ptr(), "shift_images: in and out alias")
    for t in (shift0, shift1):
        _need(t.dtype == torch.float64 and t.numel() == n, "shift_images: shifts float64 (n,)")
    lu = torch.as_tensor(spline_lu(N), device=img.device)
    step = max(1, min(n, chunk))
    ws = torch.empty(step * N * N, dtype=torch.float64, device=img.device)
    for i0 in range(0, n, step):
        i1 = min(n, i0 + step)
        call("memi_shift_images", _p(img[i0:i1]), i1 - i0, N, _p(shift0[i0:i1]),
             _p(shift1[i0:i1]), _p(lu), int(bool(transpose_out)), _p(out[i0:i1]), _p(ws),
             ws.numel(), _stream())


def ctf_params(pix_size, cs, ekv, ampc, gauss_env):
    return CtfParams(float(pix_size), float(cs), float(ekv), float(ampc), float(gauss_env))


def ctf_pack(fy, df, N, ctf, precision, Khp, row0, op_hi, op_lo, ctf_half):
    n = fy.shape[0]
    Kh = N * (N // 2 + 1)
    _need(fy.dtype == torch.complex64 and fy.numel() == n * Kh, "ctf_pack: fy (n, N, N/2+1) c64")
    _need(df.dtype == torch.float64 and df.numel() == n, "ctf_pack: df must be float64 (n,)")
    _need(ctf_half.dtype == torch.float32 and ctf_half.numel() >= n * Kh,
          "ctf_pack: ctf_half needs n*Kh float32")
    need = (int(row0) + n) * 4 * int(Khp) * (4 if precision == PREC_FP32 else 2)
    _need(_bytes(op_hi) >= need and (precision == PREC_FP32 or _bytes(op_lo) >= need),
          "ctf_pack: operand too small for rows row0..row0+n")
    call("memi_ctf_pack", _p(fy), _p(df), fy.shape[0], N, ctypes.byref(ctf), precision, Khp, row0,
         _p(op_hi), _p(op_lo), _p(ctf_half), _stream())


def phase_flip(fy, ctf_half, N, out):
    n = fy.shape[0]
    Kh = N * (N // 2 + 1)
    _need(fy.dtype == torch.complex64 and fy.numel() == n * Kh and out.shape == fy.shape
          and out.dtype == fy.dtype and ctf_half.numel() >= n * Kh, "phase_flip: shapes")
    call("memi_phase_flip", _p(fy), _p(ctf_half), fy.shape[0], N, _p(out), _stream())


def wiener_average(spec, ctf_half, N, snr, out):
    n = spec.shape[0]
    Kh = N * (N // 2 + 1)
    _need(spec.dtype == torch.complex64 and spec.numel() == n * Kh and ctf_half.numel() >= n * Kh
          and out.dtype == torch.complex64 and out.numel() == Kh, "wiener_average: shapes")
    call("memi_wiener_average", _p(spec), _p(ctf_half), spec.shape[0], N, float(snr), _p(out),
         _stream())


def image_moments(img, s, s2):
    n = img.shape[0]
    npix = img.numel() // max(n, 1)
    _need(img.dtype == torch.float32 and s.numel() == npix and s2.numel() == npix
          and s.dtype == torch.float64 and s2.dtype == torch.float64, "image_moments: shapes")
    call("memi_image_moments", _p(img), n, img.numel() // max(n, 1), _p(s), _p(s2), _stream())


# ----------------------------------------------------------------------------- plans --------
class Plan:
    """Host-built batch plan (memi_plan) uploaded once to the device."""

    def __init__(self, rows0, ns, ks, chunks=1, device="cuda", gram_tile=128):
        npd = len(ns)
        descs = (PdDesc * max(npd, 1))()
        for p in range(npd):
            descs[p].row0 = int(rows0[p])
            descs[p].n = int(ns[p])
            descs[p].k = int(ks[p])
        info = PlanInfo()
        call("memi_plan", ctypes.cast(descs, c_vp), npd, chunks, int(gram_tile), None, 0,
             ctypes.byref(info))
        host = np.zeros(max(info.plan_bytes, 16), dtype=np.uint8)
        call("memi_plan", ctypes.cast(descs, c_vp), npd, chunks, int(gram_tile),
             c_vp(host.ctypes.data), host.nbytes, ctypes.byref(info))
        self.info = info
        self.host = host
        self.dev = torch.from_numpy(host).to(device)
        pds = np.frombuffer(host[info.off_pds:info.off_pds + 40 * npd].tobytes(),
                            dtype=np.dtype([("row0", "<i8"), ("d_off", "<i8"),
                                            ("mask_off", "<i8"), ("list_off", "<i8"),
                                            ("n", "<i4"), ("k", "<i4")]))
        self.pds = pds
        self.ns = np.asarray(ns, dtype=np.int64)
        self.ks = np.asarray(ks, dtype=np.int64)
        self.rows0 = np.asarray(rows0, dtype=np.int64)
        self.chunks = chunks
        self.gram_tile = int(gram_tile)

    @property
    def npd(self):
        return int(self.info.n_pd)

    def block(self, flat, p):
        """View of PD p's n x n block of a per-PD matrix buffer."""
        n = int(self.pds["n"][p])
        off = int(self.pds["d_off"][p])
        return flat[off:off + n * n].view(n, n)


def ctf_distance(plan, op_hi, op_lo, Khp, precision, kappa, fy, ctf_half, N, D, ws, flags,
                 flag_count, events=None, variant=0):
    I = plan.info
    Kh = N * (N // 2 + 1)
    esz = 4 if precision == PREC_FP32 else 2
    _need(_bytes(op_hi) >= I.n_rows * 4 * Khp * esz and
          (precision == PREC_FP32 or _bytes(op_lo) >= I.n_rows * 4 * Khp * esz),
          "ctf_distance: operand smaller than the plan's rows")
    _need(fy.numel() >= I.n_rows * Kh and ctf_half.numel() >= I.n_rows * Kh,
          "ctf_distance: fy/ctf_half smaller than the plan's rows")
    _need(D.dtype == torch.float32 and D.numel() >= I.d_elems, "ctf_distance: D too small")
    _need(ws.dtype == torch.float32 and ws.numel() >= I.ws_floats, "ctf_distance: ws too small")
    _need(flags.dtype == torch.int32 and flags.numel() >= 3 * I.flag_cap and
          flag_count.numel() >= 1, "ctf_distance: flag buffers too small")
    ev0, ev1 = (None, None) if events is None else (c_vp(events[0]), c_vp(events[1]))
    call("memi_ctf_distance", _p(plan.dev), ctypes.byref(plan.info), _p(op_hi), _p(op_lo), Khp,
         precision, float(kappa), _p(fy), _p(ctf_half), N, _p(D), _p(ws), _p(flags),
         _p(flag_count), int(variant), ev0, ev1, _stream())


def cu_masked_stream(keep, device=None):
    """A HIP stream (torch ExternalStream) whose kernels run only on the CUs whose mask bit i
    satisfies keep(i % 32) (hipExtStreamCreateWithCUMask).  Choosing the bits by i % 32 gives
    every XCD its share whether the mask enumerates CUs XCD by XCD or interleaved.  The stream
    is blocking with respect to the null stream; one per (device, bit pattern) lives until the
    process ends."""
    dev = torch.cuda.current_device() if device is None else device
    pattern = tuple(bool(keep(b)) for b in range(32))
    if not any(pattern):
        raise ValueError("cu_masked_stream: the mask keeps no CU")
    key = (dev, pattern)
    if key in _MASKED_STREAMS:
        return _MASKED_STREAMS[key]
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    nw = (n_cu + 31) // 32
    mask = (ctypes.c_uint32 * nw)()
    for i in range(n_cu):
        if pattern[i % 32]:
            mask[i // 32] |= 1 << (i % 32)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipExtStreamCreateWithCUMask.argtypes = [ctypes.POINTER(c_vp), ctypes.c_uint32,
                                                 ctypes.POINTER(ctypes.c_uint32)]
    st = c_vp()
    with torch.cuda.device(dev):
        err = hip.hipExtStreamCreateWithCUMask(ctypes.byref(st), nw, mask)
    if err or not st.value:
        raise RuntimeError(f"hipExtStreamCreateWithCUMask failed ({err})")
    _MASKED_STREAMS[key] = torch.cuda.ExternalStream(st.value, device=dev)
    return _MASKED_STREAMS[key]


_MASKED_STREAMS = {}


class HipEvents:
    """Raw hipEvent_t pairs (libamdhip64) for timing one kernel inside a library call on the
    stream it runs on."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(c_vp)]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), c_vp, c_vp]
        self.hip.hipEventSynchronize.argtypes = [c_vp]
        self.hip.hipEventDestroy.argtypes = [c_vp]
        self.pairs = []

    def new_pair(self):
        a, b = c_vp(), c_vp()
        if self.hip.hipEventCreate(ctypes.byref(a)) or self.hip.hipEventCreate(ctypes.byref(b)):
            raise RuntimeError("hipEventCreate failed")
        self.pairs.append((a.value, b.value))
        return self.pairs[-1]

    def elapsed_ms(self, pair):
        ms = ctypes.c_float()
        self.hip.hipEventSynchronize(c_vp(pair[1]))
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), c_vp(pair[0]), c_vp(pair[1])):
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)

    def close(self):
        for a, b in self.pairs:
            self.hip.hipEventDestroy(c_vp(a))
            self.hip.hipEventDestroy(c_vp(b))
        self.pairs = []


# ------------------------------------------------------------------------- Stage 2 ---------
def knn_select(plan, D, mask, idx_out=None, val_out=None):
    I = plan.info
    _need(D.dtype == torch.float32 and D.numel() >= I.d_elems, "knn_select: D too small")
    _need(mask.numel() >= I.mask_words and mask.element_size() == 4, "knn_select: mask too small")
    _need(idx_out is None or idx_out.numel() >= I.list_elems, "knn_select: idx_out too small")
    _need(val_out is None or val_out.numel() >= I.list_elems, "knn_select: val_out too small")
    call("memi_knn_select", _p(plan.dev), ctypes.byref(plan.info), _p(D), _p(mask), _p(idx_out),
         _p(val_out), _stream())


def gather_blocks(plan, src, src_off, src_n, sel, dst):
    """dst block p = src block (src_off[p], src_n[p]) at rows/cols sel[row0_p : row0_p + n_p]
    (host int arrays, checked here and uploaded)."""
    I = plan.info
    src_off = np.ascontiguousarray(src_off, dtype=np.int64)
    src_n = np.ascontiguousarray(src_n, dtype=np.int32)
    sel = np.ascontiguousarray(sel, dtype=np.int32)
    _need(src.dtype == torch.float32 and dst.dtype == torch.float32 and
          dst.numel() >= I.d_elems, "gather_blocks: float32 src / dst of d_elems")
    _need(len(src_off) == plan.npd and len(src_n) == plan.npd and len(sel) == I.n_rows,
          "gather_blocks: one source block per PD and one index per row")
    _need(np.all(src_off >= 0) and np.all(src_off + src_n.astype(np.int64) ** 2 <= src.numel()),
          "gather_blocks: source block outside src")
    for p in range(plan.npd):
        r0, n = int(plan.rows0[p]), int(plan.ns[p])
        s = sel[r0:r0 + n]
        _need(n == 0 or (s.min() >= 0 and s.max() < src_n[p]),
              "gather_blocks: index outside source block %d" % p)
    dev = dst.device
    t_off = torch.from_numpy(src_off).to(dev)
    t_n = torch.from_numpy(src_n).to(dev)
    t_sel = torch.from_numpy(sel).to(dev)
    call("memi_gather_blocks", _p(plan.dev), ctypes.byref(plan.info), _p(src), _p(t_off),
         _p(t_n), _p(t_sel), _p(dst), _stream())


def nearest_bins(pts, centres, idx):
    """memi_nearest_bins: idx[i] = argmin_j ||pts[i] - centres[j]|| (float64 device tensors
    (n, 3) and (m, 3); idx int32 (n,))."""
    _need(pts.dtype == torch.float64 and centres.dtype == torch.float64 and pts.dim() == 2 and
          centres.dim() == 2 and pts.shape[1] == 3 and centres.shape[1] == 3 and
          pts.is_contiguous() and centres.is_contiguous(), "nearest_bins: float64 (n,3), (m,3)")
    _need(idx.dtype == torch.int32 and idx.numel() >= pts.shape[0], "nearest_bins: idx int32 (n,)")
    _need(centres.shape[0] >= 1, "nearest_bins: no centres")
    call("memi_nearest_bins", _p(pts), pts.shape[0], _p(centres), centres.shape[0], _p(idx),
         _stream())


def ferguson_sweep(plan, D, mask, logEps, out):
    I = plan.info
    _need(D.numel() >= I.d_elems and mask.numel() >= I.mask_words, "ferguson_sweep: D/mask")
    _need(logEps.dtype == torch.float64 and out.dtype == torch.float64 and
          out.numel() >= I.n_pd * (3 * logEps.numel() + 2), "ferguson_sweep: out too small")
    call("memi_ferguson_sweep", _p(plan.dev), ctypes.byref(plan.info), _p(D), _p(mask),
         _p(logEps), logEps.shape[0], _p(out), _stream())


FM_WS_PER_PD = 6 + 8 + 544 * 18


def ferguson_moments(plan, D, mask, logEps, le0, dle, out, ws):
    """memi_ferguson_moments: the sweep for a log-uniform grid logEps[e] = le0 + e*dle."""
    I = plan.info
    _need(D.numel() >= I.d_elems and mask.numel() >= I.mask_words, "ferguson_moments: D/mask")
    _need(logEps.dtype == torch.float64 and out.dtype == torch.float64 and
          out.numel() >= I.n_pd * (3 * logEps.numel() + 2), "ferguson_moments: out too small")
    _need(ws.dtype == torch.float64 and ws.numel() >= I.n_pd * FM_WS_PER_PD,
          "ferguson_moments: ws too small")
    call("memi_ferguson_moments", _p(plan.dev), ctypes.byref(plan.info), _p(D), _p(mask),
         _p(logEps), logEps.shape[0], float(le0), float(dle), _p(out), _p(ws), ws.numel(),
         _stream())


def markov_normalize(plan, D, mask, sigma, L, deg):
    I = plan.info
    _need(D.numel() >= I.d_elems and mask.numel() >= I.mask_words, "markov_normalize: D/mask")
    _need(sigma.dtype == torch.float64 and sigma.numel() >= I.n_pd, "markov_normalize: sigma")
    _need(L.dtype in (torch.float32, torch.float64) and L.numel() >= I.d_elems,
          "markov_normalize: L must be float32 or float64 with d_elems elements")
    _need(deg.dtype == torch.float64 and deg.numel() >= 2 * I.n_rows, "markov_normalize: deg")
    call("memi_markov_normalize", _p(plan.dev), ctypes.byref(plan.info), _p(D), _p(mask),
         _p(sigma), _p(L), int(L.dtype == torch.float64), _p(deg), _stream())


def tanh_fit(x, y, p0, out, active=None, max_iter=400):
    npd, m = y.shape
    if x.numel() != m or p0.shape != (npd, 4) or out.numel() < 16 * npd:
        raise RuntimeError("tanh_fit: shape mismatch")
    call("memi_tanh_fit", _p(x), _p(y), m, npd, _p(p0), _p(active), int(max_iter), _p(out),
         _stream())


def lanczos_init(V, v0, idx, ab, j, active):
    """memi_lanczos_init: V[idx] = v0, ab = 0, j = 0, active = 1 (idx checked against V by the
    caller when it is built)."""
    _need(V.dtype == torch.float64 and v0.dtype == torch.float64 and idx.dtype == torch.int64
          and idx.numel() == v0.numel(), "lanczos_init: V/v0 float64, idx int64 of v0's size")
    _need(ab.dtype == torch.float64 and j.dtype == torch.int32 and j.numel() >= 2 and
          active.dtype == torch.int32, "lanczos_init: ab float64, j / active int32")
    call("memi_lanczos_init", _p(V), _p(v0), _p(idx), v0.numel(), _p(ab), ab.numel(), _p(j),
         _p(active), active.numel(), _stream())


def tanh_select(fit, npd, R, m, tune, max_res, ok, sigma, res=None):
    """memi_tanh_select on memi_tanh_fit's out of R starts per PD."""
    _need(fit.dtype == torch.float64 and fit.numel() >= npd * R * 16, "tanh_select: fit")
    _need(ok.dtype == torch.int32 and ok.numel() >= npd, "tanh_select: ok int32 (npd,)")
    _need(sigma.dtype == torch.float64 and sigma.numel() >= npd, "tanh_select: sigma")
    _need(res is None or (res.dtype == torch.float64 and res.numel() >= npd * R),
          "tanh_select: res")
    call("memi_tanh_select", _p(fit), int(npd), int(R), int(m), float(tune), float(max_res),
         _p(ok), _p(sigma), _p(res), _stream())


def ferguson_values(vals, logEps, out):
    call("memi_ferguson_values", _p(vals), vals.numel(), _p(logEps), logEps.numel(), _p(out),
         _stream())


def markov_coo(row, col, val, nS, sigma, alpha, out, deg):
    if not (row.numel() == col.numel() == val.numel() == out.numel()):
        raise RuntimeError("markov_coo: row/col/val/out lengths differ")
    if row.numel() and (int(row.min()) < 0 or int(row.max()) >= nS or int(col.min()) < 0
                        or int(col.max()) >= nS):
        raise RuntimeError("markov_coo: index outside [0, nS)")
    call("memi_markov_coo", _p(row), _p(col), _p(val), val.numel(), int(nS), float(sigma),
         float(alpha), _p(out), _p(deg), _stream())


def kernel_ell(plan, L, col, val, cnt, sel=None):
    """memi_kernel_ell: the sparse rows of the float64 kernel L (of the PDs with sel != 0
    when sel, int32 [npd], is given)."""
    I = plan.info
    _need(L.dtype == torch.float64 and L.numel() >= I.d_elems, "kernel_ell: L float64")
    _need(col.dtype == torch.int32 and col.numel() >= I.d_elems and val.dtype == torch.float64
          and val.numel() >= I.d_elems and cnt.dtype == torch.int32 and cnt.numel() >= I.n_rows,
          "kernel_ell: col int32 / val float64 [d_elems], cnt int32 [n_rows]")
    _need(sel is None or (sel.dtype == torch.int32 and sel.numel() >= I.n_pd),
          "kernel_ell: sel int32 [npd]")
    call("memi_kernel_ell", _p(plan.dev), ctypes.byref(plan.info), _p(L),
         None if sel is None else _p(sel), _p(col), _p(val), _p(cnt), _stream())


def lanczos_steps(plan, L, V, m, j_dev, nsteps, active, W, ab, H=None):
    """nsteps Lanczos steps from the step index held in j_dev (int32 [2] on the device: index,
    arrival counter; both zero before the first step).  H: float64 [npd*(m+2)] scratch selects
    the grid-parallel orthogonalisation (W then holds 2*n_rows); None selects one workgroup per
    PD."""
    I = plan.info
    _need(isinstance(L, tuple) or (L.dtype in (torch.float32, torch.float64) and
                                   L.numel() >= I.d_elems),
          "lanczos_steps: L must be float32 or float64 with d_elems elements")
    _need(V.numel() >= m * I.n_rows and ab.numel() >= I.n_pd * m * 2 and
          W.numel() >= (1 if H is None else 2) * I.n_rows and
          active.numel() >= I.n_pd and active.dtype == torch.int32, "lanczos_steps: buffers")
    _need(j_dev.dtype == torch.int32 and j_dev.numel() >= 2, "lanczos_steps: j_dev int32 [2]")
    _need(H is None or (H.dtype == torch.float64 and H.numel() >= I.n_pd * (m + 2)),
          "lanczos_steps: H must be float64 with npd*(m+2) elements")
    if isinstance(L, tuple):
        # (col, val, cnt) of kernel_ell for every PD, or (col, val, cnt, sel, L dense): the
        # PDs with sel != 0 sparse, the others dense
        col, val, cnt = L[:3]
        sel, Ld = (L[3], L[4]) if len(L) == 5 else (None, None)
        _need(sel is None or (sel.dtype == torch.int32 and sel.numel() >= I.n_pd and
                              Ld.dtype == torch.float64 and Ld.numel() >= I.d_elems),
              "lanczos_steps: sel int32 [npd] with a float64 dense L")
        call("memi_lanczos_steps_sparse", _p(plan.dev), ctypes.byref(plan.info), _p(col),
             _p(val), _p(cnt), None if sel is None else _p(sel),
             None if Ld is None else _p(Ld), _p(V), int(m), _p(j_dev), int(nsteps),
             _p(active), _p(W), None if H is None else _p(H), _p(ab), int(H is not None),
             _stream())
        return
    call("memi_lanczos_steps", _p(plan.dev), ctypes.byref(plan.info), _p(L),
         int(L.dtype == torch.float64), _p(V), int(m),
         _p(j_dev), int(nsteps), _p(active), _p(W), None if H is None else _p(H), _p(ab),
         int(H is not None), _stream())


RITZ_KMAX = 16


def lanczos_ritz(plan, ab, m, steps_done, ks, tol, active, theta, S, steps, res, piv,
                 summary=None):
    """GPU convergence check of every active PD (memi_lanczos_ritz); buffers as in memi.h.
    summary: optional float64 [3 * npd] device buffer for {active, res, max |theta|}."""
    I = plan.info
    K = RITZ_KMAX
    _need(1 <= steps_done <= m and ab.dtype == torch.float64 and ab.numel() >= I.n_pd * m * 2,
          "lanczos_ritz: ab / steps_done")
    _need(ks.dtype == torch.int32 and ks.numel() >= I.n_pd and active.dtype == torch.int32 and
          active.numel() >= I.n_pd and steps.dtype == torch.int32 and steps.numel() >= I.n_pd,
          "lanczos_ritz: ks / active / steps must be int32 [npd]")
    _need(theta.dtype == torch.float64 and theta.numel() >= I.n_pd * K and
          S.dtype == torch.float64 and S.numel() >= I.n_pd * m * K and
          res.dtype == torch.float64 and res.numel() >= I.n_pd and
          piv.dtype == torch.float64 and piv.numel() >= I.n_pd * K * 2 * m,
          "lanczos_ritz: theta / S / res / piv sizes")
    _need(summary is None or (summary.dtype == torch.float64 and summary.numel() >= 3 * I.n_pd),
          "lanczos_ritz: summary must be float64 [3 * npd]")
    call("memi_lanczos_ritz", _p(plan.dev), ctypes.byref(plan.info), _p(ab), int(m),
         int(steps_done), _p(ks), float(tol), _p(active), _p(theta), _p(S), _p(steps), _p(res),
         _p(piv), _p(summary), _stream())


def ritz_vectors(plan, V, m, S, meta_host, meta, kmax, Y):
    """Y[n_rows][kmax] = per-PD V_p S_p.  meta_host: int64 numpy [npd][3] {J, k, s_off} (checked
    here against the plan and S); meta: the same on the device."""
    I = plan.info
    mh = np.asarray(meta_host, dtype=np.int64).reshape(-1, 3)
    ns = np.asarray(plan.ns)
    _need(mh.shape[0] == I.n_pd and meta.numel() == 3 * I.n_pd and meta.dtype == torch.int64,
          "ritz_vectors: meta must be int64 [npd][3]")
    _need(1 <= kmax <= RITZ_KMAX and np.all(mh[:, 1] >= 0) and np.all(mh[:, 1] <= kmax) and
          np.all(mh[:, 0] >= 0) and np.all(mh[:, 0] <= np.minimum(ns, m)) and
          np.all(mh[:, 2] >= 0) and np.all(mh[:, 2] + mh[:, 0] * mh[:, 1] <= S.numel()),
          "ritz_vectors: J/k/s_off out of range for the plan, m, kmax or S")
    _need(V.dtype == torch.float64 and V.numel() >= m * I.n_rows and S.dtype == torch.float64
          and Y.dtype == torch.float64 and Y.numel() >= I.n_rows * kmax, "ritz_vectors: buffers")
    call("memi_ritz_vectors", _p(plan.dev), ctypes.byref(plan.info), _p(V), int(m), _p(S),
         _p(meta), int(kmax), _p(Y), _stream())


# ------------------------------------------------------------------------- Stage 3 ---------
def nlsa_condist(D, ord_, n, con_order, ConD):
    if D.dtype not in (torch.float32, torch.float64) or D.numel() != n * n:
        raise RuntimeError("nlsa_condist: D must be float32/float64 with n*n elements")
    if ord_.numel() != n or ConD.numel() != (n - con_order) ** 2:
        raise RuntimeError("nlsa_condist: ord / ConD sizes do not match n and ConOrder")
    call("memi_nlsa_condist", _p(D), int(D.dtype == torch.float64), _p(ord_), n, con_order,
         _p(ConD), _stream())


def nlsa_supervectors(spec, ctf_half, n, N, con_order, mu_psi, snr, out):
    Kh = N * (N // 2 + 1)
    ell = mu_psi.shape[1]
    _need(spec.dtype == torch.complex64 and spec.numel() == n * Kh, "nlsa_supervectors: spec")
    _need(ctf_half.dtype == torch.float32 and ctf_half.numel() >= n * Kh, "nlsa_supervectors: ctf")
    _need(mu_psi.dtype == torch.float64 and mu_psi.shape[0] == n - con_order,
          "nlsa_supervectors: mu_psi must be float64 (n - ConOrder, ell)")
    _need(out.dtype == torch.complex64 and out.numel() >= con_order * ell * Kh,
          "nlsa_supervectors: out too small")
    call("memi_nlsa_supervectors", _p(spec), _p(ctf_half), n, N, con_order, _p(mu_psi),
         mu_psi.shape[1], float(snr), _p(out), _stream())


def tall_gram(A, G):
    _need(A.dim() == 2 and A.dtype == torch.float64 and G.numel() >= A.shape[1] ** 2 and
          G.dtype == torch.float64, "tall_gram: A (M, c) float64, G (c, c)")
    call("memi_tall_gram", _p(A), A.shape[0], A.shape[1], _p(G), _stream())


def tall_apply(A, B_host, Y):
    B = np.ascontiguousarray(B_host, dtype=np.float64)
    _need(A.dim() == 2 and A.dtype == torch.float64 and B.shape[0] == A.shape[1] and
          Y.numel() >= A.shape[0] * B.shape[1] and Y.dtype == torch.float64, "tall_apply: shapes")
    call("memi_tall_apply", _p(A), A.shape[0], A.shape[1], c_vp(B.ctypes.data), B.shape[1],
         _p(Y), _stream())


def nlsa_recon_dist(Gq, H, npix, inv_sd, Drecon2, ws):
    q, ncols = H.shape
    _need(Gq.dtype == torch.float64 and Gq.numel() == q * q and H.dtype == torch.float64 and
          H.is_contiguous() and Gq.is_contiguous(), "nlsa_recon_dist: Gq (q, q), H (q, ncols)")
    _need(inv_sd.dtype == torch.float64 and inv_sd.numel() >= ncols and
          Drecon2.dtype == torch.float32 and Drecon2.numel() >= ncols * ncols and
          ws.dtype == torch.float64 and ws.numel() >= (q + 1) * ncols,
          "nlsa_recon_dist: inv_sd / Drecon2 / ws sizes")
    call("memi_nlsa_recon_dist", _p(Gq), _p(H), q, ncols, int(npix), _p(inv_sd), _p(Drecon2),
         _p(ws), _stream())


FIT1D_BLOCK = 256


def fit1d_ws_len(n):
    return 13 * (-(-int(n) // FIT1D_BLOCK))


def fit1d_tau(x, a, b, tau, flag, sums, ws):
    """One tau update of the 1-D fit for points x (n, 3) float64 on the device at curve (a, b)."""
    n = x.shape[0]
    _need(x.dtype == torch.float64 and x.is_contiguous() and x.dim() == 2 and x.shape[1] == 3,
          "fit1d_tau: x must be contiguous float64 (n, 3)")
    _need(tau.dtype == torch.float64 and tau.numel() >= n and flag.dtype == torch.int32 and
          flag.numel() >= n and sums.dtype == torch.float64 and sums.numel() >= 13 and
          ws.dtype == torch.float64 and ws.numel() >= fit1d_ws_len(n),
          "fit1d_tau: tau / flag / sums / ws sizes")
    ab = np.ascontiguousarray(np.concatenate([np.ravel(a), np.ravel(b)]), dtype=np.float64)
    _need(ab.size == 6, "fit1d_tau: a and b must have 3 entries each")
    call("memi_fit1d_tau", _p(x), n, c_vp(ab.ctypes.data), _p(tau), _p(flag), _p(sums), _p(ws),
         _stream())


def nlsa_imgt(Ubar, H, inv_sd, IMGT):
    _need(Ubar.dim() == 2 and H.dim() == 2 and H.shape[0] == Ubar.shape[1] and
          inv_sd.numel() >= H.shape[1] and IMGT.numel() >= Ubar.shape[0] * H.shape[1] and
          all(t.dtype == torch.float64 for t in (Ubar, H, inv_sd, IMGT)), "nlsa_imgt: shapes")
    call("memi_nlsa_imgt", _p(Ubar), _p(H), _p(inv_sd), Ubar.shape[0], Ubar.shape[1], H.shape[1],
         _p(IMGT), _stream())


def center_gram(U2, G):
    _need(U2.dim() == 2 and U2.dtype == torch.float64 and G.numel() >= U2.shape[1] ** 2 and
          G.dtype == torch.float64, "center_gram: shapes")
    call("memi_center_gram", _p(U2), U2.shape[0], U2.shape[1], _p(G), _stream())


# ------------------------------------------------------------------ differentiation ---------
def kernel_rowsums(D, eps, S, T):
    _need(D.dim() == 2 and D.shape[0] == D.shape[1] and D.dtype == torch.float32,
          "kernel_rowsums: D must be a float32 (n, n) matrix")
    n = D.shape[0]
    _need(eps.dtype == torch.float64 and S.dtype == torch.float64 and T.dtype == torch.float64
          and S.numel() >= eps.numel() * n and T.numel() >= eps.numel() * n,
          "kernel_rowsums: S/T must be float64 (n_eps, n)")
    call("memi_kernel_rowsums", _p(D), D.shape[0], _p(eps), eps.shape[0], _p(S), _p(T), _stream())


def kernel_grad_weights(D, eps, S, g, Wt):
    _need(D.dim() == 2 and D.shape[0] == D.shape[1] and D.dtype == torch.float32,
          "kernel_grad_weights: D must be a float32 (n, n) matrix")
    n = D.shape[0]
    _need(S.numel() >= eps.numel() * n and g.numel() >= eps.numel() and Wt.numel() >= n * n and
          Wt.dtype == torch.float32 and g.dtype == torch.float64, "kernel_grad_weights: shapes")
    call("memi_kernel_grad_weights", _p(D), D.shape[0], _p(eps), eps.shape[0], _p(S), _p(g),
         _p(Wt), _stream())


# --- optical flow of NLSA movies (CC stage) ---------------------------------------------------

def flow_anisodiff(inp, out, tmp, niter, kappa, gamma, step, option):
    """memi_flow_anisodiff on float32 (nmov, F, H, W) movies; the result lands in out."""
    _need(inp.dtype == torch.float32 and inp.dim() == 4, "flow_anisodiff: float32 (nmov,F,H,W)")
    _need(out.dtype == torch.float32 and out.shape == inp.shape, "flow_anisodiff: out shape")
    _need(tmp is None or (tmp.dtype == torch.float32 and tmp.numel() >= inp.numel()),
          "flow_anisodiff: tmp too small")
    nmov, F, H, W = inp.shape
    call("memi_flow_anisodiff", _p(inp), _p(out), _p(tmp), nmov, F, H, W, int(niter),
         float(kappa), float(gamma), float(step[0]), float(step[1]), float(step[2]),
         int(option), _stream())


def flow_block_average(inp, bs, out):
    """memi_flow_block_average: inp float32 (nmov, F, H, W) -> out float64 (nmov, A, H, W)."""
    _need(inp.dtype == torch.float32 and inp.dim() == 4, "flow_block_average: float32 input")
    nmov, F, H, W = inp.shape
    A = -(-F // bs)
    _need(out.dtype == torch.float64 and tuple(out.shape) == (nmov, A, H, W),
          "flow_block_average: out must be float64 (nmov, ceil(F/bs), H, W)")
    call("memi_flow_block_average", _p(inp), nmov, F, H * W, int(bs), _p(out), _stream())


def gaussian_filter(inp, out, tmp, sigma):
    """memi_gaussian_filter on (count, H, W) images: f64->f64, f64->f32 or f32->f32."""
    _need(inp.dim() == 3 and out.shape == inp.shape, "gaussian_filter: (count, H, W) in and out")
    _need(inp.dtype in (torch.float32, torch.float64) and
          out.dtype in (torch.float32, torch.float64), "gaussian_filter: float32/float64")
    _need(tmp.dtype == out.dtype and tmp.numel() >= inp.numel(), "gaussian_filter: tmp")
    count, H, W = inp.shape
    call("memi_gaussian_filter", _p(inp), int(inp.dtype == torch.float64), _p(out),
         int(out.dtype == torch.float64), _p(tmp), count, H, W, float(sigma), _stream())


def hs_derivatives(frames, pairs, fx, fy, ft):
    """memi_hs_derivatives: frames float32 (nF, H, W), pairs int32 (nP, 2) -> (nP, H, W) x 3."""
    _need(frames.dtype == torch.float32 and frames.dim() == 3, "hs_derivatives: frames")
    _need(pairs.dtype == torch.int32 and pairs.dim() == 2 and pairs.shape[1] == 2,
          "hs_derivatives: pairs int32 (nP, 2)")
    nF, H, W = frames.shape
    nP = pairs.shape[0]
    for t in (fx, fy, ft):
        _need(t.dtype == torch.float32 and tuple(t.shape) == (nP, H, W),
              "hs_derivatives: fx/fy/ft float32 (nP, H, W)")
    if nP:
        pr = pairs.cpu()
        _need(int(pr.min()) >= 0 and int(pr.max()) < nF, "hs_derivatives: frame index range")
    call("memi_hs_derivatives", _p(frames), _p(pairs), nP, H, W, _p(fx), _p(fy), _p(ft),
         _stream())


def hs_iterate(fx, fy, ft, alpha, niter, U, V, tmpU, tmpV):
    """memi_hs_iterate in place on U, V (nP, H, W), float64 or float32."""
    _need(fx.dim() == 3 and fx.dtype == torch.float32, "hs_iterate: fx float32 (nP, H, W)")
    for t in (fy, ft):
        _need(t.shape == fx.shape and t.dtype == torch.float32, "hs_iterate: fy/ft")
    for t in (U, V, tmpU, tmpV):
        _need(t.shape == fx.shape and t.dtype == U.dtype, "hs_iterate: U/V/tmp shapes")
    _need(U.dtype in (torch.float32, torch.float64), "hs_iterate: U float32 or float64")
    nP, H, W = fx.shape
    call("memi_hs_iterate", _p(fx), _p(fy), _p(ft), nP, H, W, float(alpha), int(niter), _p(U),
         _p(V), int(U.dtype == torch.float64), _p(tmpU), _p(tmpV), _stream())


def flow_sum(U, V, from_zero, vx, vy, orient=None, mag=None, orient_rev=None):
    """memi_flow_sum over the frame-pair axis of U, V (nmov, A, H, W) -> float64 (nmov, H, W)."""
    _need(U.dim() == 4 and V.shape == U.shape and V.dtype == U.dtype and
          U.dtype in (torch.float32, torch.float64), "flow_sum: U, V (nmov, A, H, W)")
    nmov, A, H, W = U.shape
    for t in (vx, vy, orient, mag, orient_rev):
        _need(t is None or (t.dtype == torch.float64 and t.numel() == nmov * H * W),
              "flow_sum: outputs float64 (nmov, H, W)")
    call("memi_flow_sum", _p(U), _p(V), int(U.dtype == torch.float64), nmov, A, H * W,
         int(bool(from_zero)), _p(vx), _p(vy), _p(orient), _p(mag), _p(orient_rev), _stream())
