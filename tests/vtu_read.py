"""A plain-Python reader of VTU UnstructuredGrid pieces for the VTK tests: ASCII, base64 and raw
appended arrays, UInt32 or UInt64 headers, optional zlib blocks (VTK's XML format). Returns the
points, connectivity, offsets, types and the point and cell data as numpy arrays, plus the
VTKFile attributes."""
from __future__ import annotations

import base64
import re
import struct
import zlib
import xml.etree.ElementTree as ET

import numpy as np

_NP = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4", "UInt32": "<u4", "Int64": "<i8",
       "UInt8": "u1"}


def _decode_binary(raw: bytes, header: str, compressed: bool) -> bytes:
    w = 8 if header == "UInt64" else 4
    fmt = "<Q" if w == 8 else "<I"
    if not compressed:
        (n,) = struct.unpack(fmt, raw[:w])
        return raw[w:w + n]
    nb, bs, last = struct.unpack("<" + fmt[1] * 3, raw[:3 * w])
    sizes = struct.unpack("<" + fmt[1] * nb, raw[3 * w:3 * w + nb * w])
    pos, out = 3 * w + nb * w, b""
    for s in sizes:
        out += zlib.decompress(raw[pos:pos + s])
        pos += s
    return out


def _b64_binary(text: str, header: str, compressed: bool) -> bytes:
    w = 8 if header == "UInt64" else 4
    text = "".join(text.split())
    if not compressed:
        return _decode_binary(base64.b64decode(text), header, False)
    # The header words and the data are two base64 strings.
    nb_head = base64.b64decode(text[:((3 * w) + 2) // 3 * 4])[:w]
    nb = struct.unpack("<Q" if w == 8 else "<I", nb_head)[0]
    hlen = (3 + nb) * w
    htext = (hlen + 2) // 3 * 4
    head = base64.b64decode(text[:htext])
    data = base64.b64decode(text[htext:])
    return _decode_binary(head[:hlen] + data, header, True)


def read(path) -> dict:
    raw = open(path, "rb").read()
    appended = b""
    m = re.search(rb'<AppendedData encoding="raw">\s*_', raw)
    if m:
        # The raw block runs from the underscore to the last "\n  </AppendedData>".
        appended = raw[m.end():raw.rindex(b"\n  </AppendedData>")]
        xml_text = raw[:m.start()] + raw[raw.rindex(b"</VTKFile>"):]
    else:
        xml_text = raw
    root = ET.fromstring(xml_text)
    header = root.get("header_type", "UInt32")
    compressed = root.get("compressor") is not None
    out = {"attrs": dict(root.attrib), "point_data": {}, "cell_data": {}}

    def array(el):
        ty, fmt = el.get("type"), el.get("format")
        if fmt == "ascii":
            vals = np.array(el.text.split() if el.text else [], dtype=_NP[ty])
        elif fmt == "binary":
            vals = np.frombuffer(_b64_binary(el.text, header, compressed), dtype=_NP[ty])
        else:
            off = int(el.get("offset"))
            vals = np.frombuffer(_decode_binary(appended[off:], header, compressed), dtype=_NP[ty])
        return vals, int(el.get("NumberOfComponents", "1"))

    piece = root.find("UnstructuredGrid/Piece")
    out["num_points"] = int(piece.get("NumberOfPoints"))
    out["num_cells"] = int(piece.get("NumberOfCells"))
    pts, _ = array(piece.find("Points/DataArray"))
    out["points"] = pts.reshape(-1, 3)
    out["points_type"] = piece.find("Points/DataArray").get("type")
    for el in piece.findall("Cells/DataArray"):
        out[el.get("Name")] = array(el)[0]
    for sec, key in (("PointData", "point_data"), ("CellData", "cell_data")):
        for el in piece.findall(sec + "/DataArray"):
            v, c = array(el)
            out[key][el.get("Name")] = (v, c, el.get("type"))
    return out
