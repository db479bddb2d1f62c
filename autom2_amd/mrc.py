"""MRC files with TiltRec's header semantics, in numpy.

The 1024-byte header layout is the MRC format's (TiltRec's MRCheader,
src/TiltRec/src/mrcmx/mrcheader.h:24-101).  `write` follows
InitializeHeader / UpdateHeader (src/TiltRec/src/mrcmx/mrcstack.cpp:57-142):
mode 2, mx = my = mz = 1, cell lengths 1, angles 90, map 1,2,3, creatid
1000, no extended header, amin/amax/amean of section nz/2 with amax
starting at FLT_MIN, nlabl 0.  `read` accepts modes 0 (signed bytes, as
ReadBlock reads them), 1, 2, 6 and 12.
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

MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
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


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(1024)
    if len(raw) != 1024:
        raise ValueError(f"short MRC header in {path}")
    return np.frombuffer(raw, HEADER_DTYPE)[0].copy()


def read(path):
    """(header, float32 array [nz][ny][nx])"""
    h = read_header(path)
    mode = int(h["mode"])
    if mode not in MODES:
        raise ValueError(f"File type unknown! (MRC mode {mode})")
    n = int(h["nx"]) * int(h["ny"]) * int(h["nz"])
    data = np.fromfile(path, dtype=MODES[mode], count=n, offset=1024 + int(h["next"]))
    if data.size != n:
        raise ValueError(f"truncated MRC data in {path}")
    return h, data.astype(np.float32).reshape(int(h["nz"]), int(h["ny"]), int(h["nx"]))


def header_stats(data):
    """UpdateHeader's statistics of section nz/2 (mrcstack.cpp:104-142)."""
    sec = np.asarray(data[int(data.shape[0] * 0.5)], np.float32).ravel()
    amax = max(FLT_MIN, sec.max()) if sec.size else FLT_MIN
    amin = min(np.float32(3.4028235e38), sec.min()) if sec.size else np.float32(3.4028235e38)
    return np.float32(amin), np.float32(amax), np.float32(np.longdouble(sec.astype(np.longdouble).sum()) / sec.size)


def write(path, data, header=None):
    """Writes float32 data [nz][ny][nx] with TiltRec's output header."""
    data = np.ascontiguousarray(data, np.float32)
    nz, ny, nx = data.shape
    h = new_header(nx, ny, nz) if header is None else header.copy()
    h["amin"], h["amax"], h["amean"] = header_stats(data)
    h["nlabl"] = 0
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(data.tobytes())
