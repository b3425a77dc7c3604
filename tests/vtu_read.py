"""Minimal reader of the VTU pieces written by plens_output_write (inline base64 arrays with a
separately encoded UInt64 byte-count header) for the tests."""
import base64
import xml.etree.ElementTree as ET

import numpy as np

_DT = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}


def _array(el):
    text = "".join(el.text.split())
    nbytes = int(np.frombuffer(base64.b64decode(text[:12]), dtype=np.uint64)[0])
    data = base64.b64decode(text[12:])
    assert len(data) == nbytes, (len(data), nbytes)
    a = np.frombuffer(data, dtype=_DT[el.get("type")])
    nc = int(el.get("NumberOfComponents", "1"))
    return a.reshape(-1, nc) if nc > 1 else a


def read_vtu(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt64"
    piece = root.find("UnstructuredGrid/Piece")
    out = {"n_points": int(piece.get("NumberOfPoints")), "n_cells": int(piece.get("NumberOfCells")),
           "version": root.get("version")}
    out["points"] = _array(piece.find("Points/DataArray"))
    for el in piece.find("Cells"):
        out[el.get("Name")] = _array(el)
    out["point_data"] = {el.get("Name"): _array(el) for el in piece.find("PointData")}
    out["point_data_order"] = [el.get("Name") for el in piece.find("PointData")]
    return out
