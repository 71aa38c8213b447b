"""Reader for the inline-binary XML .vtu files the solver writes (header_type
UInt64, little-endian), used by the tests to check what was written. Decodes
with Python's own base64 and xml modules, independently of the writer."""
import base64
import xml.etree.ElementTree as ET

import numpy as np

_TYPES = {"Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}


def _array(el):
    dt = _TYPES[el.get("type")]
    comps = int(el.get("NumberOfComponents", "1"))
    text = (el.text or "").strip()
    if el.get("format") == "ascii":
        vals = np.array(text.split(), dtype=dt) if text else np.zeros(0, dt)
    else:
        raw = base64.b64decode(text)
        nbytes = int(np.frombuffer(raw[:8], "<u8")[0])
        assert len(raw) == 8 + nbytes, (len(raw), nbytes)
        vals = np.frombuffer(raw[8:], dtype=np.dtype(dt).newbyteorder("<")).astype(dt)
    return vals.reshape(-1, comps) if comps > 1 else vals


def read(path):
    """{"time", "step", "points" (N,3), "cells": {connectivity, offsets,
    types}, "point_data": {name: array}, "n_points", "n_cells", "vectors"}."""
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt64"
    assert root.get("byte_order") == "LittleEndian"
    grid = root.find("UnstructuredGrid")
    fd = {a.get("Name"): _array(a) for a in grid.find("FieldData")}
    piece = grid.find("Piece")
    out = {
        "time": float(fd["TIME"][0]),
        "step": int(fd["STEP"][0]),
        "n_points": int(piece.get("NumberOfPoints")),
        "n_cells": int(piece.get("NumberOfCells")),
    }
    pts = _array(piece.find("Points").find("DataArray"))
    out["points"] = pts.reshape(-1, 3)
    out["cells"] = {a.get("Name"): _array(a) for a in piece.find("Cells")}
    pdn = piece.find("PointData")
    out["vectors"] = pdn.get("Vectors")
    out["point_data"] = {a.get("Name"): _array(a) for a in pdn}
    return out
