"""ctypes loader for the CPU oracle (oracle/lib/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it checks results or is timed as the CPU
baseline.  The product (autom2_amd) never imports this module.

Array conventions follow TiltRec-mpi after RotateX
(src/TiltRec/src/mrcmx/mrcstack.cpp:566-579): projections [P][rows][X],
volume [rows][T][X].  Origins are (X/2, oy, T/2); TiltRec-mpi uses oy = T/2
(main-mpi.cpp:743-745), the CUDA and Z binaries use oy = Y/2.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liborc.so")
_lib = None

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "lib/liborc.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


def variant(name):
    """Loads lib/liborc_<name>.so, the same oracle built with other compiler
    flags (the Makefile's O0 and native targets), for CPU timing only."""
    return _load(os.path.join(_HERE, "lib", f"liborc_{name}.so"))


def _load(path):
    L = ctypes.CDLL(path)
    L.orc_coeffs_mpi.argtypes = [_F, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _D]
    L.orc_coeffs_cuda.argtypes = [_F, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _F]
    _DB = ctypes.c_double
    L.orc_bpt_f32.argtypes = [_D, _I, _I, _I, _I, _F, _F, _F, _DB]
    for n in ("orc_sirt_f32", "orc_sart_f32"):
        getattr(L, n).argtypes = [_D, _I, _I, _I, _I, _F, _F, _I, ctypes.c_float, _F, _DB, _DB]
    L.orc_reproject_f32.argtypes = [_D, _I, _I, _I, _F, _F, _F, _F]
    L.orc_reproject_f64.argtypes = [_D, _I, _I, _I, _F, _D, _D, _D]
    L.orc_backproject_sums_f64.argtypes = [_D, _I, _I, _I, _F, _D, _D, _D]
    L.orc_bpt_f64.argtypes = [_D, _I, _I, _I, _I, _F, _D, _D, _DB]
    for n in ("orc_sirt_f64", "orc_sart_f64"):
        getattr(L, n).argtypes = [_D, _I, _I, _I, _I, _F, _D, _I, ctypes.c_double, _D, _DB, _DB]
    L.orc_rweight_f64.argtypes = [_D, _I, _I]
    L.orc_optimal_dft_size.argtypes = [_I]
    L.orc_optimal_dft_size.restype = _I
    L.orc_spatial_taps.argtypes = [_I, _I, _F]
    L.orc_fbp_filter_f64.argtypes = [_D, _I, _I, _I]
    L.orc_admm_f32.argtypes = [_D, _I, _I, _I, _I, _F, _F, _I, _I, ctypes.c_float, ctypes.c_float, _F]
    L.orc_admm_f64.argtypes = [_D, _I, _I, _I, _I, _F, _D, _I, _I, ctypes.c_double, ctypes.c_double, _D]
    L.orc_bytes_from_stats.argtypes = [_F, ctypes.c_long, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    L.orc_float_to_byte.argtypes = [_F, ctypes.c_long, ctypes.c_void_p, _F]
    L.orc_xf_table.argtypes = [_I, _F]
    L.orc_xf_inverse.argtypes = [_F, _I, _I, _D]
    L.orc_image_mean.argtypes = [_F, ctypes.c_long]
    L.orc_image_mean.restype = ctypes.c_double
    L.orc_xf_coord.argtypes = [_D, _I, _I, ctypes.POINTER(ctypes.c_int)]
    L.orc_warp_affine.argtypes = [_F, _I, _I, _D, _I, ctypes.c_float, _F]
    L.orc_apply_xf.argtypes = [_F, _I, _I, _I, _F, _I, _F, _F]
    L.orc_gp_nchips.argtypes = [_I]
    L.orc_gp_nchips.restype = _I
    L.orc_gp_npairs.argtypes = [_I]
    L.orc_gp_npairs.restype = ctypes.c_longlong
    L.orc_gp_view.argtypes = [_F, _I, _I, _I, _I, _I, _F]
    L.orc_gp_normalize.argtypes = [_F, ctypes.c_longlong, _F]
    L.orc_gp_profiles.argtypes = [_F, _I, _I, _F]
    L.orc_gp_ncc.argtypes = [_F, _I, _I, _I]
    L.orc_gp_ncc.restype = ctypes.c_double
    _PI = ctypes.POINTER(ctypes.c_int)
    L.orc_gp_search.argtypes = [_F, _I, _PI, _PI, _PI, _PI]
    L.orc_gp_search.restype = ctypes.c_float
    L.orc_gp_view_result.argtypes = [_PI, _PI, _F, _I, _F]
    L.orc_geopar.argtypes = [_F, _I, _I, _I, _I, _F, _F, _F]
    return L


def _fp(a):
    return a.ctypes.data_as(_F)


def _dp(a):
    return a.ctypes.data_as(_D)


def _org(org):
    return (ctypes.c_float * 3)(*[float(v) for v in org])


def coeffs_mpi(angles, offset=0.0, pitch=0.0, zshift=0.0):
    """[P][8] double coefficients (a0..a3, b0..b3), main-mpi.cpp:85-183."""
    ang = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((len(ang), 8), dtype=np.float64)
    lib().orc_coeffs_mpi(_fp(ang), len(ang), offset, pitch, zshift, _dp(out))
    return out


def coeffs_cuda(angles, offset=0.0, pitch=0.0, zshift=0.0):
    """[P][8] float coefficients, main-cuda.cpp:111-168."""
    ang = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((len(ang), 8), dtype=np.float32)
    lib().orc_coeffs_cuda(_fp(ang), len(ang), offset, pitch, zshift, _fp(out))
    return out


def origin(X, T, oy):
    return (np.float32(X * 0.5), np.float32(oy), np.float32(T * 0.5))


# TiltRec-cuda's weight thresholds (update, residual) per method, as the
# double literals its kernels compare float sums with: BPT/WBP/FBP c > 1e-5
# (basic.cu:409), SIRT wtvol > 10e-6 (:184) with an exact residual test
# (:152), SART c > 10e-7 (:434) and reproj_wt > 1e-7 (:170).  TiltRec-mpi's
# tests are exact: (0, 0).
CUDA_THRESHOLDS = {"BPT": (1e-5, 0.0), "WBP": (1e-5, 0.0), "FBP": (1e-5, 0.0),
                   "SIRT": (10e-6, 0.0), "SART": (10e-7, 1e-7)}


def rules_origin(X, Y, T, rules):
    """TiltRec-mpi puts the y origin at T/2 (main-mpi.cpp:743-745), the CUDA
    and Z binaries at Y/2 (main-cuda.cpp:259-261)."""
    return origin(X, T, T / 2 if rules == 0 else Y / 2)


def _dims(proj, T):
    P, rows, X = proj.shape
    return P, X, rows, T


def reproject_f32(coef1, vol, org):
    """One tilt's Reproject (main-mpi.cpp:286-344): returns (val, wt) [rows][X]."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    rows, T, X = vol.shape
    rv = np.zeros((rows, X), np.float32)
    rw = np.zeros((rows, X), np.float32)
    c = np.ascontiguousarray(coef1, dtype=np.float64)
    lib().orc_reproject_f32(_dp(c), X, rows, T, _org(org), _fp(vol), _fp(rv), _fp(rw))
    return rv, rw


def reproject_f64(coef1, vol, org):
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    rows, T, X = vol.shape
    rv = np.zeros((rows, X), np.float64)
    rw = np.zeros((rows, X), np.float64)
    c = np.ascontiguousarray(coef1, dtype=np.float64)
    lib().orc_reproject_f64(_dp(c), X, rows, T, _org(org), _dp(vol), _dp(rv), _dp(rw))
    return rv, rw


def backproject_sums_f64(coef1, img, T, org):
    """Per-voxel BilinearValue sums (s, c) of one tilt, [rows][T][X]."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    rows, X = img.shape
    s = np.zeros((rows, T, X), np.float64)
    c = np.zeros((rows, T, X), np.float64)
    cc = np.ascontiguousarray(coef1, dtype=np.float64)
    lib().orc_backproject_sums_f64(_dp(cc), X, rows, T, _org(org), _dp(img), _dp(s), _dp(c))
    return s, c


def bpt_f32(coef, proj, T, org, vol=None, thr=(0.0, 0.0)):
    """thr = (update, residual) thresholds; BPT reads the update one."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float32) if vol is None else np.array(vol, np.float32)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    lib().orc_bpt_f32(_dp(c), P, X, rows, T, _org(org), _fp(proj), _fp(out), float(thr[0]))
    return out


def bpt_f64(coef, proj, T, org, vol=None, thr=(0.0, 0.0)):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float64) if vol is None else np.array(vol, np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    lib().orc_bpt_f64(_dp(c), P, X, rows, T, _org(org), _dp(proj), _dp(out), float(thr[0]))
    return out


def _iter_f32(name, coef, proj, T, org, iters, gamma, vol, thr, L=None):
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float32) if vol is None else np.array(vol, np.float32)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    getattr(L or lib(), name)(_dp(c), P, X, rows, T, _org(org), _fp(proj), iters, gamma, _fp(out), float(thr[0]),
                         float(thr[1]))
    return out


def _iter_f64(name, coef, proj, T, org, iters, gamma, vol, thr):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float64) if vol is None else np.array(vol, np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    getattr(lib(), name)(_dp(c), P, X, rows, T, _org(org), _dp(proj), iters, gamma, _dp(out), float(thr[0]),
                         float(thr[1]))
    return out


def sirt_f32(coef, proj, T, org, iters, gamma=1.0, vol=None, thr=(0.0, 0.0), L=None):
    """thr = (update, residual) thresholds: (0, 0) TiltRec-mpi, CUDA_THRESHOLDS
    TiltRec-cuda.  L: a library from variant() instead of the default build."""
    return _iter_f32("orc_sirt_f32", coef, proj, T, org, iters, gamma, vol, thr, L)


def sart_f32(coef, proj, T, org, iters, gamma=1.0, vol=None, thr=(0.0, 0.0)):
    return _iter_f32("orc_sart_f32", coef, proj, T, org, iters, gamma, vol, thr)


def sirt_f64(coef, proj, T, org, iters, gamma=1.0, vol=None, thr=(0.0, 0.0)):
    return _iter_f64("orc_sirt_f64", coef, proj, T, org, iters, gamma, vol, thr)


def sart_f64(coef, proj, T, org, iters, gamma=1.0, vol=None, thr=(0.0, 0.0)):
    return _iter_f64("orc_sart_f64", coef, proj, T, org, iters, gamma, vol, thr)


def rweight_f64(rows):
    """ApplyFilterInplace mode 2 on [..., X] rows, FP64 (filter_prj.cpp:652-736)."""
    a = np.array(rows, dtype=np.float64, copy=True)
    shp = a.shape
    a = np.ascontiguousarray(a.reshape(-1, shp[-1]))
    lib().orc_rweight_f64(_dp(a), shp[-1], a.shape[0])
    return a.reshape(shp)


def rel_l2(a, b):
    """||a - b|| / ||b|| in float64."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def admm_f32(coef, proj, T, org, outer, cgiter, mu, soft, vol=None):
    """TiltRecZ-mpi's ADMM (main-mpiz.cpp:603-1013), FP32 form."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float32) if vol is None else np.array(vol, np.float32)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    lib().orc_admm_f32(_dp(c), P, X, rows, T, _org(org), _fp(proj), int(outer), int(cgiter), float(mu), float(soft),
                       _fp(out))
    return out


def admm_f64(coef, proj, T, org, outer, cgiter, mu, soft, vol=None):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    P, X, rows, _ = _dims(proj, T)
    out = np.zeros((rows, T, X), np.float64) if vol is None else np.array(vol, np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    lib().orc_admm_f64(_dp(c), P, X, rows, T, _org(org), _dp(proj), int(outer), int(cgiter), float(mu), float(soft),
                       _dp(out))
    return out


def optimal_dft_size(n):
    return lib().orc_optimal_dft_size(int(n))


def spatial_taps(mode, L=251):
    """Shepp-Logan (mode 0) or Ram-Lak (mode 1) taps, filter_prj.cpp:137-178."""
    h = np.zeros(L, np.float32)
    lib().orc_spatial_taps(int(mode), int(L), _fp(h))
    return h


def fbp_filter_f64(rows, mode=0):
    """ApplyFilterInplace mode 0/1 on [..., X] rows, FP64 (filter_prj.cpp:266-446)."""
    a = np.array(rows, dtype=np.float64, copy=True)
    shp = a.shape
    a = np.ascontiguousarray(a.reshape(-1, shp[-1]))
    lib().orc_fbp_filter_f64(_dp(a), shp[-1], a.shape[0], int(mode))
    return a.reshape(shp)


def fbp_f32(coef, proj, T, org, mode=0, vol=None, thr=(0.0, 0.0)):
    """TiltRec's FBP: the spatial filter (FP64, rounded to float as the
    reference stores it) followed by the FP32 back-projection."""
    filt = fbp_filter_f64(proj.astype(np.float64), mode).astype(np.float32)
    return bpt_f32(coef, filt, T, org, vol, thr)

def float_to_byte(values):
    """--f2b conversion (recon_z.cu:253-262): (bytes uint8, (mean, sd))."""
    v = np.ascontiguousarray(values, np.float32).ravel()
    out = np.empty(v.size, np.uint8)
    st = np.zeros(2, np.float32)
    lib().orc_float_to_byte(_fp(v), v.size, out.ctypes.data, _fp(st))
    return out.reshape(np.shape(values)), (float(st[0]), float(st[1]))


def bytes_from_stats(values, mean, sd):
    """The byte mapping alone, for given float mean and deviation."""
    v = np.ascontiguousarray(values, np.float32).ravel()
    out = np.empty(v.size, np.uint8)
    lib().orc_bytes_from_stats(_fp(v), v.size, float(mean), float(sd), out.ctypes.data)
    return out.reshape(np.shape(values))

# ------------------------------------------------------------ .xf application
# mrcstack (src/mrctools/src/mrcstack/main.cpp:40-80) with OpenCV's
# warpAffine restated in xform_oracle.c.

def xf_table(lanczos):
    """1-D kernel table at fractions i/32: [32][4] cubic or [32][8] Lanczos-4."""
    K = 8 if lanczos else 4
    t = np.zeros((32, K), np.float32)
    lib().orc_xf_table(int(bool(lanczos)), _fp(t))
    return t


def xf_inverse(xf_row, W, H):
    """warpAffine's inverse of mrcstack's matrix for one .xf line."""
    x = np.ascontiguousarray(xf_row, np.float32)
    M = np.zeros(6, np.float64)
    lib().orc_xf_inverse(_fp(x), W, H, _dp(M))
    return M


def image_mean(img):
    a = np.ascontiguousarray(img, np.float32)
    return lib().orc_image_mean(_fp(a), a.size)


def xf_coord(M, x, y):
    """(ix, iy, table index fy*32 + fx) of output pixel (x, y)."""
    out = (ctypes.c_int * 3)()
    Md = np.ascontiguousarray(M, np.float64)
    lib().orc_xf_coord(_dp(Md), int(x), int(y), out)
    return out[0], out[1], out[2]


def warp_affine(src, M, lanczos, cval):
    s = np.ascontiguousarray(src, np.float32)
    H, W = s.shape
    Md = np.ascontiguousarray(M, np.float64)
    d = np.empty_like(s)
    lib().orc_warp_affine(_fp(s), W, H, _dp(Md), int(bool(lanczos)), float(cval), _fp(d))
    return d


def apply_xf(stack, xf, mode=1):
    """stack [P][H][W] float32, xf [P][6]; returns (out, border values)."""
    s = np.ascontiguousarray(stack, np.float32)
    P, H, W = s.shape
    x = np.ascontiguousarray(xf, np.float32).reshape(P, 6)
    out = np.empty_like(s)
    cv = np.empty(P, np.float32)
    lib().orc_apply_xf(_fp(s), W, H, P, _fp(x), int(mode), _fp(out), _fp(cv))
    return out, cv


# ---------------------------------------------------------------- GeoParGen

def gp_nchips(w):
    return lib().orc_gp_nchips(int(w))


def gp_npairs(n):
    return lib().orc_gp_npairs(int(n))


def gp_view(vol, axis, v):
    """View v of a [nz][ny][nx] volume: axis 0 ("x") -> [ny][nz], axis 1 -> [ny][nx]."""
    a = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = a.shape
    out = np.empty((ny, nz) if axis == 0 else (ny, nx), np.float32)
    lib().orc_gp_view(_fp(a), nx, ny, nz, int(axis), int(v), _fp(out))
    return out


def gp_normalize(img):
    a = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(a)
    lib().orc_gp_normalize(_fp(a), a.size, _fp(out))
    return out


def gp_profiles(norm):
    """[rows][cols] normalised view -> [nchips][rows] chip profiles."""
    a = np.ascontiguousarray(norm, np.float32)
    rows, cols = a.shape
    out = np.empty((gp_nchips(cols), rows), np.float32)
    lib().orc_gp_profiles(_fp(a), rows, cols, _fp(out))
    return out


def gp_ncc(prof, l, r):
    a = np.ascontiguousarray(prof, np.float32)
    return lib().orc_gp_ncc(_fp(a), a.size, int(l), int(r))


def gp_search(prof):
    """-> (z, thickness, ncc float, l, r) as getThicknessAndZ(prof, 30, 60)."""
    a = np.ascontiguousarray(prof, np.float32)
    z, t, lo, hi = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    nc = lib().orc_gp_search(_fp(a), a.size, ctypes.byref(z), ctypes.byref(t), ctypes.byref(lo),
                             ctypes.byref(hi))
    return z.value, t.value, np.float32(nc), lo.value, hi.value


def gp_view_result(z, thick, nc):
    zz = np.ascontiguousarray(z, np.int32)
    tt = np.ascontiguousarray(thick, np.int32)
    cc = np.ascontiguousarray(nc, np.float32)
    out = np.empty(3, np.float32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib().orc_gp_view_result(zz.ctypes.data_as(ip), tt.ctypes.data_as(ip), _fp(cc), len(cc), _fp(out))
    return out


def geopar(vol, axis=0):
    """GeoParGen on a [nz][ny][nx] float volume.  Returns (views [nviews][3]
    = angle, thickness, z; chips [nviews][nchips][4] = z, thickness, ncc, l;
    averages [3])."""
    a = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = a.shape
    nv, cols = (nx, nz) if axis == 0 else (nz, nx)
    nc = gp_nchips(cols)
    views = np.empty((nv, 3), np.float32)
    chips = np.empty((nv, nc, 4), np.float32)
    avg = np.empty(3, np.float32)
    lib().orc_geopar(_fp(a), nx, ny, nz, int(axis), _fp(views), _fp(chips), _fp(avg))
    return views, chips, avg
