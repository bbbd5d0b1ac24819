"""MetaImage (.mha) reading and writing, as the reference's mha_io does
(src/base/mha_io.cxx:26-257): uncompressed data stored after the header
(ElementDataFile = LOCAL); MET_FLOAT, MET_SHORT and MET_UCHAR volumes and
float vector fields (ElementNumberOfChannels = 3) on read, MET_UINT too on
write; TransformMatrix holds the direction cosines by columns; little- or
big-endian data (BinaryDataByteOrderMSB).  Images are numpy arrays shaped
(nz, ny, nx) or (nz, ny, nx, 3)."""
import numpy as np

_TYPES = {"MET_FLOAT": np.float32, "MET_SHORT": np.int16, "MET_UCHAR": np.uint8, "MET_UINT": np.uint32}


class MhaImage:
    """An image read from or written to an .mha file: data, origin (Offset),
    spacing (ElementSpacing), direction cosines row-major."""

    def __init__(self, img, spacing, origin=(0.0, 0.0, 0.0),
                 direction_cosines=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)):
        self.img = img
        self.spacing = tuple(float(v) for v in spacing)
        self.origin = tuple(float(v) for v in origin)
        self.direction_cosines = tuple(float(v) for v in direction_cosines)

    @property
    def dim(self):
        return (self.img.shape[2], self.img.shape[1], self.img.shape[0])


def _g(v):
    return "%g" % v


def write_mha(path, img, spacing, origin=(0.0, 0.0, 0.0),
              direction_cosines=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0), precise=False):
    """write_mha (src/base/mha_io.cxx:26-101): the reference's header, field
    for field, then the voxels little-endian.  The reference prints the
    geometry with %g (six significant digits); precise=True prints it
    exactly (shortest round-trip form) for fixtures that must keep it."""
    img = np.asarray(img)
    vf = img.ndim == 4
    if vf and (img.shape[3] != 3 or img.dtype != np.float32):
        raise ValueError("a vector field must be float32 with 3 components")
    etype = {np.dtype(np.float32): "MET_FLOAT", np.dtype(np.int16): "MET_SHORT",
             np.dtype(np.uint8): "MET_UCHAR", np.dtype(np.uint32): "MET_UINT"}.get(img.dtype)
    if etype is None:
        raise ValueError(f"unhandled type {img.dtype} (the reference writes uchar, short, uint32, float)")
    dc = direction_cosines
    _f = (lambda v: repr(float(v))) if precise else _g
    nz, ny, nx = img.shape[:3]
    header = ("ObjectType = Image\n"
              "NDims = 3\n"
              "BinaryData = True\n"
              "BinaryDataByteOrderMSB = False\n"
              f"TransformMatrix = {' '.join(_f(dc[i]) for i in (0, 3, 6, 1, 4, 7, 2, 5, 8))}\n"
              f"Offset = {' '.join(_f(v) for v in origin)}\n"
              "CenterOfRotation = 0 0 0\n"
              f"ElementSpacing = {' '.join(_f(v) for v in spacing)}\n"
              f"DimSize = {nx} {ny} {nz}\n"
              "AnatomicalOrientation = RAI\n"
              + ("ElementNumberOfChannels = 3\n" if vf else "")
              + f"ElementType = {etype}\n"
              "ElementDataFile = LOCAL\n")
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(np.ascontiguousarray(img, img.dtype.newbyteorder("<")).tobytes())


def read_mha(path):
    """read_mha (src/base/mha_io.cxx:103-245): header lines up to
    "ElementDataFile = LOCAL", then the data.  Raises ValueError where the
    reference gives up."""
    with open(path, "rb") as f:
        data = f.read()
    dim, origin, spacing = None, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    dc = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    etype, channels, big = None, 1, False
    pos = 0
    while True:
        end = data.find(b"\n", pos)
        if end < 0:
            raise ValueError(f"{path}: no ElementDataFile = LOCAL line")
        line = data[pos:end].decode("latin-1").rstrip()
        pos = end + 1
        if line == "ElementDataFile = LOCAL":
            break
        key, sep, val = line.partition(" = ")
        if not sep:
            continue
        vals = val.split()
        if key == "DimSize" and len(vals) >= 3:
            dim = tuple(int(v) for v in vals[:3])
        elif key == "Offset" and len(vals) >= 3:
            origin = tuple(float(np.float32(v)) for v in vals[:3])
        elif key == "ElementSpacing" and len(vals) >= 3:
            spacing = tuple(float(np.float32(v)) for v in vals[:3])
        elif key == "TransformMatrix" and len(vals) >= 9:
            m = [float(np.float32(v)) for v in vals[:9]]
            dc = (m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8])
        elif key == "ElementNumberOfChannels":
            channels = int(vals[0])
        elif key == "ElementType":
            etype = val.strip()
        elif key == "BinaryDataByteOrderMSB":
            big = val.strip() == "True"
    if dim is None:
        raise ValueError(f"{path}: no DimSize")
    if etype not in ("MET_FLOAT", "MET_SHORT", "MET_UCHAR"):
        raise ValueError(f"{path}: cannot interpret ElementType {etype}")
    if channels != 1 and etype != "MET_FLOAT":
        raise ValueError(f"{path}: vector fields must be MET_FLOAT")
    dt = np.dtype(_TYPES[etype]).newbyteorder(">" if big else "<")
    n = dim[0] * dim[1] * dim[2] * (3 if channels != 1 else 1)
    if len(data) - pos < n * dt.itemsize:
        raise ValueError(f"{path}: {len(data) - pos} data bytes for {n} values")
    a = np.frombuffer(data, dt, n, pos).astype(dt.newbyteorder("="))
    shape = (dim[2], dim[1], dim[0]) + ((3,) if channels != 1 else ())
    return MhaImage(a.reshape(shape), spacing, origin, dc)
