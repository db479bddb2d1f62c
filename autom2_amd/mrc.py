"""MRC files with TiltRec's header semantics, in numpy.

The 1024-byte header layout is the MRC format's (TiltRec's MRCheader,
src/TiltRec/src/mrcmx/mrcheader.h:24-101).  `write` follows
InitializeHeader / UpdateHeader (src/TiltRec/src/mrcmx/mrcstack.cpp:57-142):
mode 2, mx = my = mz = 1, cell lengths 1, angles 90, map 1,2,3, creatid
1000, no extended header, amin/amax/amean of section nz/2 with amax
starting at FLT_MIN, nlabl 0; with mode=12 it writes halves, and with
mrc2014=True an MRC2014 header (MAP and machine stamps, NVERSION 20140,
sampling mx/my/mz = nx/ny/nz, whole-map min/max/mean/RMS).  `read` accepts
modes 0 (signed bytes, as ReadBlock reads them), 1, 2, 6, 12 and 101
(4-bit), in either byte order (the machine stamp, or whichever order gives a
plausible header).
"""
import numpy as np

HEADER_DTYPE = np.dtype([
    ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"), ("mode", "<i4"),
    ("nxstart", "<i4"), ("nystart", "<i4"), ("nzstart", "<i4"),
    ("mx", "<i4"), ("my", "<i4"), ("mz", "<i4"),
    ("xlen", "<f4"), ("ylen", "<f4"), ("zlen", "<f4"),
    ("alpha", "<f4"), ("beta", "<f4"), ("gamma", "<f4"),
    ("mapc", "<i4"), ("mapr", "<i4"), ("maps", "<i4"),
    ("amin", "<f4"), ("amax", "<f4"), ("amean", "<f4"),
    ("ispg", "<i2"), ("nsymbt", "<i2"), ("next", "<i4"), ("creatid", "<i2"),
    ("blank", "S30"), ("nint", "<i2"), ("nreal", "<i2"), ("sub", "<i2"), ("zfac", "<i2"),
    ("min2", "<f4"), ("max2", "<f4"), ("min3", "<f4"), ("max3", "<f4"), ("min4", "<f4"), ("max4", "<f4"),
    ("idtype", "<i2"), ("lens", "<i2"), ("nd1", "<i2"), ("nd2", "<i2"), ("vd1", "<i2"), ("vd2", "<i2"),
    ("tiltangles", "<f4", (6,)), ("xorg", "<f4"), ("yorg", "<f4"), ("zorg", "<f4"),
    ("cmap", "S4"), ("stamp", "S4"), ("rms", "<f4"), ("nlabl", "<i4"), ("labels", "S80", (10,)),
])
assert HEADER_DTYPE.itemsize == 1024

MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16, 101: np.uint8}
FLT_MIN = np.float32(1.17549435e-38)


def new_header(nx, ny, nz):
    h = np.zeros((), HEADER_DTYPE)
    h["nx"], h["ny"], h["nz"], h["mode"] = nx, ny, nz, 2
    h["mx"] = h["my"] = h["mz"] = 1
    h["xlen"] = h["ylen"] = h["zlen"] = 1.0
    h["alpha"] = h["beta"] = h["gamma"] = 90.0
    h["mapc"], h["mapr"], h["maps"] = 1, 2, 3
    h["amin"], h["amax"], h["amean"] = 0.0, 255.0, 128.0
    h["creatid"] = 1000
    return h


def _plausible(h):
    lim = 1 << 28
    return (all(0 < int(h[k]) < lim for k in ("nx", "ny", "nz")) and int(h["mode"]) in MODES
            and int(h["next"]) >= 0)


def _header_from(raw):
    """(header in native order, big-endian?) from the first 1024 bytes."""
    h = np.frombuffer(raw, HEADER_DTYPE)[0].copy()
    s0 = raw[212]
    be = s0 == 0x11
    if s0 not in (0x11, 0x44) and not _plausible(h):
        hb = np.frombuffer(raw, HEADER_DTYPE.newbyteorder(">"))[0]
        be = _plausible(hb)
    if be:
        h = np.frombuffer(raw, HEADER_DTYPE.newbyteorder(">"))[0].astype(HEADER_DTYPE)
    return h, be


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(1024)
    if len(raw) != 1024:
        raise ValueError(f"short MRC header in {path}")
    return _header_from(raw)[0]


def read(path):
    """(header, float32 array [nz][ny][nx])"""
    with open(path, "rb") as f:
        raw = f.read(1024)
    if len(raw) != 1024:
        raise ValueError(f"short MRC header in {path}")
    h, be = _header_from(raw)
    mode = int(h["mode"])
    if mode not in MODES:
        raise ValueError(f"File type unknown! (MRC mode {mode})")
    nx, ny, nz = int(h["nx"]), int(h["ny"]), int(h["nz"])
    off = 1024 + int(h["next"])
    if mode == 101:
        rb = (nx + 1) // 2
        packed = np.fromfile(path, dtype=np.uint8, count=rb * ny * nz, offset=off)
        if packed.size != rb * ny * nz:
            raise ValueError(f"truncated MRC data in {path}")
        packed = packed.reshape(nz * ny, rb)
        vals = np.stack([packed & 15, packed >> 4], axis=2).reshape(nz * ny, 2 * rb)[:, :nx]
        return h, vals.astype(np.float32).reshape(nz, ny, nx)
    dt = np.dtype(MODES[mode]).newbyteorder(">" if be else "<")
    n = nx * ny * nz
    data = np.fromfile(path, dtype=dt, count=n, offset=off)
    if data.size != n:
        raise ValueError(f"truncated MRC data in {path}")
    return h, data.astype(np.float32).reshape(nz, ny, nx)


def header_stats(data):
    """UpdateHeader's statistics of section nz/2 (mrcstack.cpp:104-142)."""
    sec = np.asarray(data[int(data.shape[0] * 0.5)], np.float32).ravel()
    amax = max(FLT_MIN, sec.max()) if sec.size else FLT_MIN
    amin = min(np.float32(3.4028235e38), sec.min()) if sec.size else np.float32(3.4028235e38)
    return np.float32(amin), np.float32(amax), np.float32(np.longdouble(sec.astype(np.longdouble).sum()) / sec.size)


def set_2014(h, pixel=1.0):
    """MRC2014 fields: MAP stamp, little-endian machine stamp, NVERSION
    20140 (word 28, inside `blank`), sampling = size, cell = pixel * size."""
    h["cmap"] = b"MAP "
    h["stamp"] = b"\x44\x44\x00\x00"
    blank = bytearray(bytes(np.asarray(h["blank"]).item()).ljust(30, b"\0"))
    blank[10:14] = np.int32(20140).tobytes()
    h["blank"] = bytes(blank)
    h["mx"], h["my"], h["mz"] = h["nx"], h["ny"], h["nz"]
    p = pixel if pixel > 0 else 1.0
    h["xlen"], h["ylen"], h["zlen"] = p * h["nx"], p * h["ny"], p * h["nz"]
    h["alpha"] = h["beta"] = h["gamma"] = 90.0
    return h


def nversion(h):
    return int(np.frombuffer(bytes(np.asarray(h["blank"]).item()).ljust(30, b"\0")[10:14], "<i4")[0])


def write(path, data, header=None, mode=2, mrc2014=False):
    """Writes data [nz][ny][nx] as mode 2 (float32) or 12 (float16) with
    TiltRec's output header, or an MRC2014 header with mrc2014=True."""
    if mode not in (2, 12):
        raise ValueError("mode must be 2 or 12")
    data = np.ascontiguousarray(data, np.float32 if mode == 2 else np.float16)
    nz, ny, nx = data.shape
    h = new_header(nx, ny, nz) if header is None else header.copy()
    h["mode"] = mode
    if mrc2014:
        set_2014(h)
        v = data.astype(np.float64)
        h["amin"], h["amax"], h["amean"] = v.min(), v.max(), v.mean()
        h["rms"] = v.std()
    else:
        h["amin"], h["amax"], h["amean"] = header_stats(data.astype(np.float32))
        h["nlabl"] = 0
    h["next"] = 0
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(data.tobytes())
