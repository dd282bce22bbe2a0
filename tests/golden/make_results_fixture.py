"""Writes tests/golden/results_formats.json from the result files the
reference ships (results/wallmodel/test1/boat_results.csv and two records of
results/web3d/cache with their base profiles). Only data files are read
(CSV, JSON, numpy .npz without pickles); no reference code is imported.

The fixture keeps what pins the formats: the CSV text, every field of the
web-viewer records that depends only on the case geometry, the count and a
hash of the pooled cells marked solid (-128) in the int8 velocity field, the
sizes of the encoded arrays, and the names, shapes and dtypes of the base
profiles.

    python tests/golden/make_results_fixture.py [/root/reference]
"""
import base64
import hashlib
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "results_formats.json")
TAG = "e3_B50_Hb4_z2_Bf300_n0.03_h3_n2000000_L150"
GEOM = ("nx", "ny", "nz", "dxs", "dz", "y0", "L", "W", "h", "V", "Q", "S", "zb", "hulls", "props", "jet")


def record(key):
    c = json.load(open(os.path.join(REF, "results/web3d/cache", f"{key}_{TAG}.json"), encoding="utf-8"))
    q = np.frombuffer(base64.b64decode(c["u"]), np.int8).reshape(c["nz"], c["ny"], c["nx"], 3)
    solid = q[..., 0] == -128
    assert (q[solid] == -128).all()
    tq = np.frombuffer(base64.b64decode(c["tau"]), np.uint16)
    out = {k: c[k] for k in GEOM}
    out.update(solid_cells=int(solid.sum()),
               solid_sha256=hashlib.sha256(np.packbits(solid.ravel()).tobytes()).hexdigest(),
               u_b64_len=len(c["u"]), tau_len=int(tq.size), prof_len=len(c["prof"]["x"]),
               keys=sorted(c.keys()), stats_keys=sorted(c.get("stats", {}).keys()))
    return out


def main():
    csv_path = os.path.join(REF, "results/wallmodel/test1/boat_results.csv")
    csv_text = open(csv_path, encoding="utf-8-sig", newline="").read()
    npz = np.load(os.path.join(REF, "results/web3d/cache", f"base_profiles_{TAG}.npz"))
    fx = {"csv": {"source": "results/wallmodel/test1/boat_results.csv", "text": csv_text},
          "tag": TAG,
          "records": {k: record(k) for k in ("base", "t1c1")},
          "base_profiles": {k: [list(npz[k].shape), str(npz[k].dtype)] for k in npz.files}}
    with open(OUT, "w", encoding="utf-8") as f:
        json.dump(fx, f, ensure_ascii=False, indent=1)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
