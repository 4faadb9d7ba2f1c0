"""Write compact golden fixtures from the reference's native GRF fixtures.

Reads web/stellarator/fixtures/native/order{4,6}-direct.{bin,json} under the
reference checkout (layout: magic 'STGRF001', 4 x i64 (ntri, nsrc, nrender,
nfaces), f64 grfError, then sx, snx, sw, ub, ubn, u, render arrays;
web/stellarator/test/native-direct-parity.py:41-71) and stores, per order:
every 10th node's sx/snx/sw (geometry pin), the full post-RRQ GRF residual u,
grfError and the counts.  Data only; no reference code is copied.

Usage: python make_fixtures.py [/root/reference]
"""
import json
import struct
import sys
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent


def load(order):
    d = REF / "web/stellarator/fixtures/native"
    blob = (d / f"order{order}-direct.bin").read_bytes()
    meta = json.loads((d / f"order{order}-direct.json").read_text())
    magic, ntri, nsrc, nrender, nfaces, grf = struct.unpack_from("<8s4qd", blob)
    assert magic == b"STGRF001"
    off = 48
    f = {}
    for name, cnt in [("sx", 3 * nsrc), ("snx", 3 * nsrc), ("sw", nsrc), ("ub", nsrc), ("ubn", nsrc), ("u", nsrc)]:
        f[name] = np.frombuffer(blob, dtype="<f8", count=cnt, offset=off).copy()
        off += 8 * cnt
    return ntri, nsrc, grf, f, meta


for order in (4, 6):
    ntri, nsrc, grf, f, meta = load(order)
    sel = np.arange(0, nsrc, 10)
    np.savez_compressed(OUT / f"stellarator_order{order}.npz",
                        ntri=ntri, nsrc=nsrc, grf=grf, mp=meta["mp"], np_=meta["np"], order=order,
                        sel=sel, sx=f["sx"].reshape(-1, 3)[sel], snx=f["snx"].reshape(-1, 3)[sel], sw=f["sw"][sel],
                        ub=f["ub"][sel], ubn=f["ubn"][sel], u=f["u"])
    print(order, ntri, nsrc, grf)
