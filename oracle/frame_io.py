"""TEST INFRASTRUCTURE ONLY.  Python restatement of the reference's recorded-frame
file format, used by tests/ as the checker of the native frame ingest.

Foundation/NuiFileIOUtilities.cpp:13-52 writeFrameImage: int64 timestamp, uint32
width, uint32 height, then uint32 stored size and the payload -- the zlib stream
of the raw buffer (compress2 at Z_BEST_SPEED = level 1) when compression is on
and succeeds, else the raw buffer.  :77-110 readFrameImageBuffer: the payload is
raw when its stored size equals the raw size the reader expects, else it is
inflated.  :112-160 writeCamera / readCamera: four float32 fx, fy, cx, cy.
Frame/NuiFrameSaveManager.cpp:24-80: a recording is <dir>/<id>.depth / .color /
.cmap / .vertex / .camIntri.  Pixel sizes: UINT16, BGRQUAD, ColorSpacePoint
(2 floats), CameraSpacePoint (3 floats) (src/stdafx.h:122-140)."""
import os
import struct
import zlib

import numpy as np

PARTS = [(".depth", np.uint16, 1), (".color", np.uint8, 4), (".cmap", np.float32, 2), (".vertex", np.float32, 3)]
HEADER = struct.Struct("<qII")


def frame_image_bytes(timestamp, width, height, buffer, compressed):
    raw = np.ascontiguousarray(buffer).tobytes()
    head = HEADER.pack(timestamp, width, height)
    if compressed:
        z = zlib.compress(raw, 1)
        return head + struct.pack("<I", len(z)) + z
    return head + struct.pack("<I", len(raw)) + raw


def write_frame_image(path, timestamp, width, height, buffer, compressed):
    with open(path, "wb") as f:
        f.write(frame_image_bytes(timestamp, width, height, buffer, compressed))


def read_frame_image(path, dtype, channels):
    """(timestamp, width, height, array[h, w, channels]) or None when the file is missing."""
    if not os.path.exists(path):
        return None
    data = open(path, "rb").read()
    ts, w, h = HEADER.unpack_from(data, 0)
    (stored,) = struct.unpack_from("<I", data, HEADER.size)
    raw_bytes = w * h * np.dtype(dtype).itemsize * channels
    payload = data[HEADER.size + 4:HEADER.size + 4 + stored]
    raw = payload if stored == raw_bytes else zlib.decompress(payload)
    assert len(raw) == raw_bytes
    return ts, w, h, np.frombuffer(raw, dtype).reshape(h, w, channels)


def write_camera(path, fx, fy, cx, cy):
    with open(path, "wb") as f:
        f.write(struct.pack("<4f", fx, fy, cx, cy))


def read_camera(path):
    if not os.path.exists(path):
        return None
    return struct.unpack("<4f", open(path, "rb").read()[:16])


def write_recording_frame(dirname, frame_id, parts, compressed, camera=None):
    """parts: {".depth": (timestamp, array[h, w(, c)]), ...}."""
    for ext, dtype, ch in PARTS:
        if ext in parts:
            ts, arr = parts[ext]
            a = np.ascontiguousarray(arr, dtype)
            h, w = a.shape[0], a.shape[1]
            write_frame_image(os.path.join(dirname, f"{frame_id}{ext}"), ts, w, h, a, compressed)
    if camera is not None:
        write_camera(os.path.join(dirname, f"{frame_id}.camIntri"), *camera)


def read_recording(dirname):
    """Frames 0, 1, ... until the first id with no file (DeviceManager/FileLoader/
    NuiFrameLoader.cpp:26-91); each frame a dict of the parts that exist."""
    out = []
    while True:
        fid = len(out)
        fr = {}
        for ext, dtype, ch in PARTS:
            r = read_frame_image(os.path.join(dirname, f"{fid}{ext}"), dtype, ch)
            if r is not None and r[1] * r[2] > 0:
                fr[ext] = r
        cam = read_camera(os.path.join(dirname, f"{fid}.camIntri"))
        if cam is not None:
            fr[".camIntri"] = cam
        if not fr:
            return out
        out.append(fr)
