"""Writes tests/golden/diverted_flood.npz and diverted_flood.json from the reference's
examples/diverted-flood inputs: the DEM (a single-strip Deflate Float32 GeoTIFF, decoded
here with zlib; nodata -9999 -> NaN, as crates/itr-cli/src/io.rs:33-104 reads it), its
georeferencing (ModelPixelScale / ModelTiepoint), the scenario's hydrology and objective
settings (scenario.toml), the protected-area and guard polygons (GeoJSON), and the
values the reference publishes for the no-change run: J = 0.0279 (README.md:90,
spec.md:1395-1396) and asset peak depths 0.033 and 0.062 m (README.md:15).

Run once in a container that has the reference checkout:
    python tests/golden/make_diverted_flood.py /root/reference/examples/diverted-flood
"""
from __future__ import annotations

import json
import os
import struct
import sys
import zlib

import numpy as np

try:
    import tomllib
except ImportError:   # Python < 3.11
    import tomli as tomllib

HERE = os.path.dirname(os.path.abspath(__file__))
TYPE_SIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 11: 4, 12: 8, 16: 8}
TYPE_FMT = {1: "B", 3: "H", 4: "I", 11: "f", 12: "d", 16: "Q"}


def read_tiff(path):
    """Baseline TIFF, one image: returns (data f32 [h][w] with nodata as NaN, tags)."""
    buf = open(path, "rb").read()
    e = "<" if buf[:2] == b"II" else ">"
    magic, off = struct.unpack(e + "HI", buf[2:8])
    assert magic == 42, "not a classic TIFF"
    n = struct.unpack(e + "H", buf[off:off + 2])[0]
    tags = {}
    for i in range(n):
        t, typ, cnt, val = struct.unpack(e + "HHI4s", buf[off + 2 + 12 * i: off + 14 + 12 * i])
        size = TYPE_SIZE[typ] * cnt
        data = val if size <= 4 else buf[struct.unpack(e + "I", val)[0]:][:size]
        if typ == 2:
            tags[t] = data[:cnt].rstrip(b"\0").decode()
        elif typ == 5:
            v = struct.unpack(e + "%dI" % (2 * cnt), data)
            tags[t] = tuple(v[2 * k] / v[2 * k + 1] for k in range(cnt))
        else:
            tags[t] = struct.unpack(e + "%d%s" % (cnt, TYPE_FMT[typ]), data[:size])
    w, h = tags[256][0], tags[257][0]
    assert tags[258][0] == 32 and tags[339][0] == 3, "expected Float32 samples"
    assert tags.get(317, (1,))[0] == 1, "predictor not supported"
    comp = tags.get(259, (1,))[0]
    raw = b""
    for o, c in zip(tags[273], tags[279]):
        s = buf[o:o + c]
        raw += zlib.decompress(s) if comp in (8, 32946) else s
    assert comp in (1, 8, 32946)
    a = np.frombuffer(raw, dtype=e + "f4", count=w * h).reshape(h, w).astype(np.float32)
    nd = tags.get(42113)
    if nd is not None:
        a = np.where(a == np.float32(float(nd)), np.float32(np.nan), a)
    return a, tags


def polygons(path):
    fc = json.load(open(path))
    out = []
    for f in fc["features"]:
        g = f["geometry"]
        rings = [g["coordinates"]] if g["type"] == "Polygon" else g["coordinates"]
        for poly in rings:
            out.append({"exterior": poly[0], "holes": poly[1:], "properties": f.get("properties", {})})
    return out


def main(src):
    dem, tags = read_tiff(os.path.join(src, "dem.tif"))
    s, t = tags[33550], tags[33922]
    geo = {"origin_x": t[3] - t[0] * s[0], "origin_y": t[4] + t[1] * s[1], "dx": s[0], "dy": s[1]}
    keys = tags.get(34735, ())
    assert not any(keys[4 + 4 * k] == 1025 and keys[4 + 4 * k + 3] == 2 for k in range(keys[3] if keys else 0)), \
        "pixel-is-point rasters are shifted by half a cell"
    sc = tomllib.load(open(os.path.join(src, "scenario.toml"), "rb"))
    hyd, obj = sc["hydrology"], sc["objectives"]
    fixture = {
        "source": "examples/diverted-flood (dem.tif, scenario.toml, protected.geojson, guard.geojson)",
        "geo": geo,
        "duration_s": hyd["duration_s"], "sync_interval_s": hyd["sync_interval_s"],
        "rain_mm_h": hyd["rainfall_hyetograph"], "manning_n": hyd["manning_n"],
        "infil_capacity_mm_h": hyd["infiltration"]["capacity_mm_h"],
        "boundaries": sc["boundaries"]["segments"],
        "depth_threshold_m": obj["depth_threshold_m"], "guard_tolerance_m": obj["guard_tolerance_m"],
        "tau_m": 0.02,   # ObjectiveDef::tau_m default (objectives.smoothing_m unset; problem.rs:159)
        "protected": polygons(os.path.join(src, sc["objectives"]["protected_areas"])),
        "guard": polygons(os.path.join(src, sc["objectives"]["downstream_guard_areas"])),
        "published": {"j_no_change": 0.0279, "asset_peaks_m": [0.033, 0.062]},
    }
    np.savez_compressed(os.path.join(HERE, "diverted_flood.npz"), dem=dem)
    with open(os.path.join(HERE, "diverted_flood.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print("dem", dem.shape, "nan", int(np.isnan(dem).sum()), "range", np.nanmin(dem), np.nanmax(dem), geo)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples/diverted-flood")
