"""On-disk formats of the path tracer, byte-compatible with the C++ host API
(host/src/aurora_io.cpp):

- Portable FloatMap (PFM) dumps of FP32 radiance and AOV images: header
  "PF\\n<w> <h>\\n-1.0\\n" (colour, little-endian), scanlines from the bottom
  row up; arrays are (H, W, 3) float32 with row 0 at the top.
- the accumulation checkpoint of a progressive render ("AURACC01"): camera
  and render settings, the samples accumulated so far and the per-pixel sums,
  with a CRC-32 (the zlib / IEEE 802.3 polynomial) of the sums.  Resuming
  adds the next samples in sample order, so an interrupted render continues
  to the same bits.

Errors raise RuntimeError naming the file, the contract of the reference's
PPM loader (src/texture.cpp:9-50).
"""
import dataclasses
import struct
import zlib

import numpy as np

from .core import Camera, RenderSettings

MAGIC = b"AURACC01"
HEADER = 128


def write_pfm(path, rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    if a.ndim != 3 or a.shape[2] != 3 or a.shape[0] == 0 or a.shape[1] == 0:
        raise ValueError("PFM image must be (H, W, 3)")
    H, W = a.shape[:2]
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (W, H))
        f.write(np.ascontiguousarray(a[::-1]).astype("<f4").tobytes())


def read_pfm(path):
    try:
        data = open(path, "rb").read()
    except OSError:
        raise RuntimeError(f"Failed to open image file: {path}") from None
    tokens, pos = [], 0
    while len(tokens) < 4 and pos < len(data):
        while pos < len(data) and data[pos:pos + 1].isspace():
            pos += 1
        start = pos
        while pos < len(data) and not data[pos:pos + 1].isspace():
            pos += 1
        if pos > start:
            tokens.append(data[start:pos])
        if len(tokens) == 1 and tokens[0] != b"PF":
            break
    if not tokens or tokens[0] != b"PF":
        raise RuntimeError(f"Unsupported PFM format in file: {path}")
    try:
        W, H, scale = int(tokens[1]), int(tokens[2]), float(tokens[3])
    except (IndexError, ValueError):
        raise RuntimeError(f"Invalid PFM header in file: {path}") from None
    if W <= 0 or H <= 0 or scale == 0.0 or not np.isfinite(scale):
        raise RuntimeError(f"Invalid PFM header in file: {path}")
    head = pos + 1  # one whitespace byte ends the header
    n = W * H * 3
    if len(data) - head < 4 * n:
        raise RuntimeError(f"Unexpected end of file while reading pixel data in file: {path}")
    body = np.frombuffer(data, "<f4" if scale < 0 else ">f4", count=n, offset=head).astype(np.float32)
    return body.reshape(H, W, 3)[::-1].copy()


@dataclasses.dataclass
class Accumulation:
    """A progressive render's state: its camera and settings, samples
    [settings.sample_offset, settings.sample_offset + spp) accumulated, and
    the (H, W, 3) float32 sums (albedo/normal None without AOVs)."""
    camera: Camera
    settings: RenderSettings
    spp: int
    rgb_sum: np.ndarray
    albedo_sum: np.ndarray = None
    normal_sum: np.ndarray = None


def save_checkpoint(path, acc: Accumulation):
    c, r = acc.camera, acc.settings
    aov = acc.albedo_sum is not None
    sums = [acc.rgb_sum] + ([acc.albedo_sum, acc.normal_sum] if aov else [])
    for s in sums:
        if s is None or s.shape != (c.image_height, c.image_width, 3):
            raise ValueError("accumulation sums do not match the camera")
    payload = b"".join(np.ascontiguousarray(s, "<f4").tobytes() for s in sums)
    h = MAGIC + struct.pack("<IIII", HEADER, c.image_width, c.image_height, 1 if aov else 0)
    h += struct.pack("<12f", c.vfov, *c.lookfrom, *c.lookat, *c.vup, c.defocus_angle, c.focus_dist)
    h += struct.pack("<IIII", acc.spp, r.sample_offset, r.max_depth, r.seed)
    h += struct.pack("<4f", *r.background, r.tmin)
    h += b"\0" * (HEADER - 8 - len(h))
    h += struct.pack("<II", zlib.crc32(payload), 0)
    with open(path, "wb") as f:
        f.write(h)
        f.write(payload)


def load_checkpoint(path):
    try:
        data = open(path, "rb").read()
    except OSError:
        raise RuntimeError(f"Failed to open checkpoint file: {path}") from None
    if len(data) < HEADER or data[:8] != MAGIC or struct.unpack_from("<I", data, 8)[0] != HEADER:
        raise RuntimeError(f"Not an accumulation checkpoint: {path}")
    W, H, flags = struct.unpack_from("<III", data, 12)
    f = struct.unpack_from("<12f", data, 24)
    spp, off, depth, seed = struct.unpack_from("<IIII", data, 72)
    bg = struct.unpack_from("<3f", data, 88)
    tmin = struct.unpack_from("<f", data, 100)[0]
    want = struct.unpack_from("<I", data, HEADER - 8)[0]
    n = W * H * 3
    if n == 0:
        raise RuntimeError(f"Invalid checkpoint header in file: {path}")
    k = 3 if flags & 1 else 1
    if len(data) < HEADER + 4 * n * k:
        raise RuntimeError(f"Unexpected end of file while reading sums in file: {path}")
    if len(data) > HEADER + 4 * n * k:
        raise RuntimeError(f"Trailing data in checkpoint file: {path}")
    payload = data[HEADER:]
    if zlib.crc32(payload) != want:
        raise RuntimeError(f"Checksum mismatch in file: {path}")
    sums = [np.frombuffer(payload, "<f4", count=n, offset=4 * n * i).astype(np.float32).reshape(H, W, 3)
            for i in range(k)]
    cam = Camera(W, H, f[0], tuple(f[1:4]), tuple(f[4:7]), tuple(f[7:10]), f[10], f[11])
    rs = RenderSettings(spp=0, max_depth=depth, seed=seed, sample_offset=off, background=tuple(bg), aov=k == 3,
                        tmin=tmin)
    return Accumulation(cam, rs, spp, sums[0], sums[1] if k == 3 else None, sums[2] if k == 3 else None)


def accumulate(scene, acc: Accumulation, more):
    """Add `more` samples to acc (in place) with a host-buffer render:
    samples [sample_offset + spp, sample_offset + spp + more)."""
    if more <= 0:
        return acc
    rs = dataclasses.replace(acc.settings, spp=more, sample_offset=acc.settings.sample_offset + acc.spp,
                             aov=acc.albedo_sum is not None, spp_per_pass=0)
    rgb, alb, nrm, _ = scene.render(acc.camera, rs, acc.rgb_sum, acc.albedo_sum, acc.normal_sum)
    acc.rgb_sum, acc.albedo_sum, acc.normal_sum = rgb, (alb if rs.aov else None), (nrm if rs.aov else None)
    acc.spp += more
    return acc
