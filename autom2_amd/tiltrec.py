"""Python binding of libtiltrec_hip.so (include/tiltrec.h) over ctypes.

Mirrors TiltRec's own interface for the reconstruction path: the
``--method NAME,iterations,gamma`` and ``--geometry offset,pitch,zshift,
thickness`` option strings (src/TiltRec/src/opts/opts.cpp:232-335), the
method names BPT/FBP/WBP/SIRT/SART/ADMM, and the two volume orientations
(TiltRec-* writes [Y][T][X], TiltRecZ-* writes [T][Y][X]).

Device memory and streams come from torch tensors (``tensor.data_ptr()`` and
``torch.cuda.current_stream().cuda_stream``); the arithmetic is all in the
HIP library.  If the library is missing the import of this module fails: there
is no CPU fallback.
"""
import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtiltrec_hip.so")

OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_NOMEM, ERR_DEVICE, ERR_IO, ERR_BUSY, ERR_NOT_READY = range(8)
BPT, FBP, WBP, SIRT, SART, ADMM = range(6)
STORE_F32, STORE_F16, STORE_BF16 = range(3)
LAYOUT_Y, LAYOUT_Z = range(2)
RULES_MPI, RULES_CUDA = range(2)

METHOD_NAMES = {"BPT": BPT, "FBP": FBP, "WBP": WBP, "SIRT": SIRT, "SART": SART, "ADMM": ADMM}
STORAGE_NAMES = {"f32": STORE_F32, "fp32": STORE_F32, "f16": STORE_F16, "fp16": STORE_F16,
                 "bf16": STORE_BF16}


class TrecError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class Geom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nproj", ctypes.c_int),
                ("thickness", ctypes.c_int), ("offset", ctypes.c_float), ("pitch", ctypes.c_float),
                ("zshift", ctypes.c_float), ("rules", ctypes.c_int)]


class Method(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("iterations", ctypes.c_int), ("gamma", ctypes.c_float),
                ("storage", ctypes.c_int), ("layout", ctypes.c_int), ("cg_iterations", ctypes.c_int),
                ("soft", ctypes.c_float)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int), ("zchunks", ctypes.c_int), ("band_cells", ctypes.c_int),
                ("back_tilts", ctypes.c_int), ("back_voxels_z", ctypes.c_int), ("slabs", ctypes.c_int),
                ("slab_rows", ctypes.c_int),
                ("workspace", ctypes.c_size_t)]


class Job(ctypes.Structure):
    _fields_ = [("geom", Geom), ("method", Method), ("angles_deg", ctypes.POINTER(ctypes.c_float)),
                ("stack", ctypes.c_void_p), ("initial", ctypes.c_void_p), ("volume", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("geom", Geom), ("method", Method), ("angles_deg", ctypes.POINTER(ctypes.c_float)),
                ("stack", ctypes.c_void_p), ("volume_out", ctypes.c_void_p), ("stack_half", ctypes.c_int),
                ("volume_half", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("footprint", ctypes.c_size_t), ("staged_s", ctypes.c_double),
                ("computed_s", ctypes.c_double)]


# (name, restype, argtypes) of every function include/tiltrec.h declares
_P = ctypes.c_void_p
_I = ctypes.c_int
_PF = ctypes.POINTER(ctypes.c_float)
SIGNATURES = [
    ("trec_context_create", _I, [_I, ctypes.POINTER(_P)]),
    ("trec_context_destroy", None, [_P]),
    ("trec_context_trim", _I, [_P]),
    ("trec_status_string", ctypes.c_char_p, [_I]),
    ("trec_last_error", ctypes.c_char_p, [_P]),
    ("trec_version", ctypes.c_char_p, []),
    ("trec_coefficients", _I, [ctypes.POINTER(Geom), _PF, ctypes.POINTER(ctypes.c_double)]),
    ("trec_workspace_size", _I, [ctypes.POINTER(Geom), ctypes.POINTER(Method), ctypes.POINTER(ctypes.c_size_t)]),
    ("trec_reconstruct", _I, [_P, ctypes.POINTER(Geom), _PF, _P, _P, ctypes.POINTER(Method), _P, _P]),
    ("trec_filter_rweight", _I, [_P, _P, _I, ctypes.c_longlong, _P]),
    ("trec_filter_rweight_static", _I, [_P, _P, _I, ctypes.c_longlong, _P]),
    ("trec_filter_fbp", _I, [_P, _P, _I, ctypes.c_longlong, _I, _P]),
    ("trec_float_to_byte", _I, [_P, _P, ctypes.c_size_t, _P, _PF, _P]),
    ("trec_forward_project", _I, [_P, ctypes.POINTER(Geom), _PF, _P, _P, _P, _P]),
    ("trec_backproject", _I, [_P, ctypes.POINTER(Geom), _PF, _P, _P, _P]),
    ("trec_sirt_iterate", _I, [_P, ctypes.POINTER(Geom), _PF, _P, ctypes.c_float, _P, _P]),
    ("trec_reconstruct_batch", _I, [_P, _I, ctypes.POINTER(Job), _P]),
    ("trec_set_workspace_limit", _I, [_P, ctypes.c_size_t]),
    ("trec_plan", _I, [_P, ctypes.POINTER(Geom), _PF, ctypes.POINTER(Method), ctypes.POINTER(PlanInfo)]),
    ("trec_autotune", _I, [_P, ctypes.POINTER(Geom), ctypes.POINTER(Method)]),
    ("trec_profile_enable", _I, [_P, _I]),
    ("trec_profile_read", _I, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]),
    ("trec_server_create", _I, [_I, ctypes.c_size_t, _I, ctypes.POINTER(_P)]),
    ("trec_server_submit", _I, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    ("trec_server_poll", _I, [_P, ctypes.c_uint64, ctypes.POINTER(_I)]),
    ("trec_server_wait", _I, [_P, ctypes.c_uint64, ctypes.POINTER(Result)]),
    ("trec_request_footprint", _I, [ctypes.POINTER(Geom), ctypes.POINTER(Method), ctypes.POINTER(ctypes.c_size_t)]),
    ("trec_server_destroy", None, [_P]),
    ("trec_apply_xf", _I, [_P, _I, _I, _I, _PF, _I, _P, _P, _P]),
    ("trec_convert", _I, [_P, _P, _I, _P, _I, ctypes.c_size_t, _P]),
    ("trec_geopar", _I, [_P, _P, _I, _I, _I, _I, _I, _P, _P, _PF, ctypes.POINTER(_I), _P]),
]

XF_CUBIC, XF_LANCZOS4 = 1, 2  # mrcstack -m: 2 selects Lanczos-4, anything else cubic
GP_VIEW_X, GP_VIEW_Y = 0, 1    # GeoParGen's "x" (default) and "y" views
# trec_dtype: element types of MRC data
DT_I8, DT_U8, DT_U8_DIV255, DT_I16, DT_U16, DT_F16, DT_F32 = range(7)
GP_EXHAUSTIVE = 1              # trec_geopar flag: exact NCC of every pair

_lib = None


def lib():
    """Loads libtiltrec_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "or `make -C autom2_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ------------------------------------------------------------------ options

def parse_method(text):
    """--method value -> (kind, iterations, gamma), as opts.cpp:287-335 reads it.

    "SIRT,20,1.0" and "SART,5,0.5" carry an iteration count and a relaxation;
    "BPT", "FBP" and "WBP" carry nothing; "ADMM,n,cg,gamma,soft" carries the
    outer iterations, the CG steps, the penalty and the soft threshold (the
    last two are returned by parse_admm).  Unlike the reference, a bare
    "SIRT" is an error instead of reading uninitialised values
    (opts.cpp:112-124).
    """
    parts = text.split(",")
    name = parts[0]
    if name not in METHOD_NAMES:
        raise ValueError(f"unknown method {name!r}")
    kind = METHOD_NAMES[name]
    if kind in (SIRT, SART):
        if len(parts) != 3:
            raise ValueError(f"{name} needs NAME,iterations,gamma")
        return kind, int(parts[1]), float(parts[2])
    if kind == ADMM:
        return kind, parse_admm(text)[0], parse_admm(text)[2]
    if len(parts) != 1:
        raise ValueError(f"{name} takes no parameters")
    return kind, 0, 1.0


def parse_admm(text):
    """"ADMM,iterations,cgiter,gamma,soft" -> (iterations, cgiter, gamma, soft)."""
    parts = text.split(",")
    if parts[0] != "ADMM" or len(parts) != 5:
        raise ValueError("ADMM needs ADMM,iterations,cgiter,gamma,soft")
    return int(parts[1]), int(parts[2]), float(parts[3]), float(parts[4])


def parse_geometry(text):
    """--geometry offset,pitch,zshift,thickness (opts.cpp:232-260)."""
    parts = text.split(",")
    if len(parts) != 4:
        raise ValueError("geometry needs offset,pitch,zshift,thickness")
    return float(parts[0]), float(parts[1]), float(parts[2]), int(parts[3])


def make_geom(nx, ny, nproj, thickness, offset=0.0, pitch=0.0, zshift=0.0, rules=RULES_MPI):
    return Geom(nx, ny, nproj, thickness, offset, pitch, zshift, rules)


def make_method(kind, iterations=0, gamma=1.0, storage=STORE_F32, layout=LAYOUT_Y, cg_iterations=0, soft=0.0):
    if isinstance(kind, str):
        text = kind
        kind, iterations, gamma = parse_method(text)
        if kind == ADMM:
            iterations, cg_iterations, gamma, soft = parse_admm(text)
    if isinstance(storage, str):
        storage = STORAGE_NAMES[storage]
    return Method(kind, iterations, gamma, storage, layout, cg_iterations, soft)


def _angles(angles):
    a = (ctypes.c_float * len(angles))(*[float(v) for v in angles])
    return ctypes.cast(a, _PF), a


def _check(status, ctx=None):
    if status != OK:
        msg = lib().trec_status_string(status).decode()
        if ctx:
            detail = lib().trec_last_error(ctx).decode()
            if detail:
                msg = f"{msg}: {detail}"
        raise TrecError(status, msg)


def coefficients(geom, angles):
    import numpy as np
    out = np.zeros((len(angles), 8), np.float64)
    pa, _keep = _angles(angles)
    _check(lib().trec_coefficients(ctypes.byref(geom), pa,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def workspace_size(geom, method):
    n = ctypes.c_size_t()
    _check(lib().trec_workspace_size(ctypes.byref(geom), ctypes.byref(method), ctypes.byref(n)))
    return n.value


def request_footprint(geom, method):
    n = ctypes.c_size_t()
    _check(lib().trec_request_footprint(ctypes.byref(geom), ctypes.byref(method), ctypes.byref(n)))
    return n.value


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


@dataclass
class Series:
    """One tilt series for a batch: device tensors and its options."""
    geom: Geom
    method: Method
    angles: list
    stack: object
    volume: object
    initial: object = None


class Context:
    """A device context: cached workspace, cached HIP graph, autotune table."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().trec_context_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().trec_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def reconstruct(self, geom, angles, stack, method, volume, initial=None, stream=None):
        pa, _keep = _angles(angles)
        _check(lib().trec_reconstruct(self._h, ctypes.byref(geom), pa, _ptr(stack), _ptr(initial),
                                      ctypes.byref(method), _ptr(volume), _stream(stream)), self._h)

    def reconstruct_batch(self, series, stream=None):
        jobs = (Job * len(series))()
        keep = []
        for j, s in zip(jobs, series):
            pa, k = _angles(s.angles)
            keep.append(k)
            j.geom = s.geom
            j.method = s.method
            j.angles_deg = pa
            j.stack = s.stack.data_ptr()
            j.initial = s.initial.data_ptr() if s.initial is not None else None
            j.volume = s.volume.data_ptr()
        _check(lib().trec_reconstruct_batch(self._h, len(series), jobs, _stream(stream)), self._h)

    def filter_rweight(self, rows, stream=None):
        nx = rows.shape[-1]
        _check(lib().trec_filter_rweight(self._h, _ptr(rows), nx, rows.numel() // nx, _stream(stream)), self._h)

    def filter_rweight_static(self, rows, stream=None):
        """The R-weighting without any device allocation (rows up to 7448 pixels)."""
        nx = rows.shape[-1]
        _check(lib().trec_filter_rweight_static(self._h, _ptr(rows), nx, rows.numel() // nx, _stream(stream)),
               self._h)

    def filter_fbp(self, rows, mode=0, stream=None):
        """FBP's spatial filter in place on [..., nx] rows (0: Shepp-Logan, 1: Ram-Lak)."""
        nx = rows.shape[-1]
        _check(lib().trec_filter_fbp(self._h, _ptr(rows), nx, rows.numel() // nx, int(mode), _stream(stream)),
               self._h)

    def float_to_byte(self, values, out, stream=None):
        """--f2b conversion of a float device tensor into a uint8 device
        tensor of the same size; returns (mean, sd)."""
        stats = (ctypes.c_float * 2)()
        _check(lib().trec_float_to_byte(self._h, _ptr(values), values.numel(), _ptr(out),
                                        ctypes.cast(stats, _PF), _stream(stream)), self._h)
        return stats[0], stats[1]

    def forward_project(self, geom, angles, volume, rval, rwt=None, stream=None):
        pa, _keep = _angles(angles)
        _check(lib().trec_forward_project(self._h, ctypes.byref(geom), pa, _ptr(volume), _ptr(rval), _ptr(rwt),
                                          _stream(stream)), self._h)

    def backproject(self, geom, angles, stack, volume, stream=None):
        pa, _keep = _angles(angles)
        _check(lib().trec_backproject(self._h, ctypes.byref(geom), pa, _ptr(stack), _ptr(volume),
                                      _stream(stream)), self._h)

    def sirt_iterate(self, geom, angles, stack, gamma, volume, stream=None):
        pa, _keep = _angles(angles)
        _check(lib().trec_sirt_iterate(self._h, ctypes.byref(geom), pa, _ptr(stack), gamma, _ptr(volume),
                                       _stream(stream)), self._h)

    def set_workspace_limit(self, nbytes):
        """Workspace budget of a call (0: the free device memory); larger
        pitch-0 series run in y slabs."""
        _check(lib().trec_set_workspace_limit(self._h, int(nbytes)), self._h)

    def plan(self, geom, angles, method):
        """How reconstruct() will run the series: a dict of trec_plan_info."""
        info = PlanInfo()
        pa, _keep = _angles(angles)
        _check(lib().trec_plan(self._h, ctypes.byref(geom), pa, ctypes.byref(method), ctypes.byref(info)), self._h)
        return {f: getattr(info, f) for f, _ in PlanInfo._fields_}

    def autotune(self, geom, method):
        _check(lib().trec_autotune(self._h, ctypes.byref(geom), ctypes.byref(method)), self._h)

    def profile(self, on=True):
        _check(lib().trec_profile_enable(self._h, 1 if on else 0), self._h)

    def profile_read(self, kernel):
        """(total milliseconds, launch groups) of one kernel since profile()."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(lib().trec_profile_read(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)), self._h)
        return ms.value, n.value

    def trim(self):
        _check(lib().trec_context_trim(self._h), self._h)

    def apply_xf(self, stack, xf, out, mode=XF_CUBIC, stream=None):
        """mrcstack's .xf application (trec_apply_xf): stack and out are
        [nsec][ny][nx] float32 device tensors, xf a host sequence of nsec
        rows a11 a12 a21 a22 dx dy."""
        import numpy as np
        nsec, ny, nx = stack.shape
        x = np.ascontiguousarray(np.asarray(xf, np.float32).reshape(nsec, 6))
        _check(lib().trec_apply_xf(self._h, nx, ny, nsec, x.ctypes.data_as(_PF), int(mode), _ptr(stack),
                                   _ptr(out), _stream(stream)), self._h)

    def convert(self, src, src_type, dst, dst_type, stream=None):
        """trec_convert: the elements of device tensor src (interpreted as
        src_type) into dst (DT_F32 or DT_F16), same element count."""
        _check(lib().trec_convert(self._h, _ptr(src), int(src_type), _ptr(dst), int(dst_type), dst.numel(),
                                  _stream(stream)), self._h)

    def geopar(self, volume, axis=GP_VIEW_X, views=None, chips=None, exhaustive=False, stream=None):
        """GeoParGen (trec_geopar) on a [nz][ny][nx] float32 device volume.
        views ([nviews][3]) and chips ([nviews][nchips][4]) are optional
        float32 device tensors.  Waits for the stream; returns (averages
        angle, thickness, z as floats, chips searched exhaustively)."""
        nz, ny, nx = volume.shape
        res = (ctypes.c_float * 3)()
        nex = _I()
        _check(lib().trec_geopar(self._h, _ptr(volume), nx, ny, nz, int(axis), GP_EXHAUSTIVE if exhaustive else 0,
                                 _ptr(views), _ptr(chips), ctypes.cast(res, _PF), ctypes.byref(nex),
                                 _stream(stream)), self._h)
        return (res[0], res[1], res[2]), nex.value


class Server:
    """Continuous-batching request loop (trec_server_*).  Stacks and output
    volumes are host numpy arrays; both are held until wait()."""

    def __init__(self, device=0, budget=0, max_active=8):
        self._h = ctypes.c_void_p()
        self._keep = {}
        _check(lib().trec_server_create(device, budget, max_active, ctypes.byref(self._h)))

    def submit(self, geom, method, angles, stack, out):
        import numpy as np
        assert stack.dtype in (np.float32, np.float16) and stack.flags["C_CONTIGUOUS"]
        assert out.dtype in (np.float32, np.float16) and out.flags["C_CONTIGUOUS"]
        pa, keep = _angles(angles)
        req = Request(geom, method, pa, stack.ctypes.data, out.ctypes.data, int(stack.dtype == np.float16),
                      int(out.dtype == np.float16))
        rid = ctypes.c_uint64()
        _check(lib().trec_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)))
        self._keep[rid.value] = (stack, out, keep)
        return rid.value

    def poll(self, rid):
        d = ctypes.c_int()
        _check(lib().trec_server_poll(self._h, rid, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, rid):
        r = Result()
        st = lib().trec_server_wait(self._h, rid, ctypes.byref(r))
        self._keep.pop(rid, None)
        _check(st)
        return {"status": r.status, "submit_s": r.submit_s, "admit_s": r.admit_s, "done_s": r.done_s,
                "footprint": r.footprint, "staged_s": r.staged_s, "computed_s": r.computed_s}

    def close(self):
        if self._h:
            lib().trec_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
