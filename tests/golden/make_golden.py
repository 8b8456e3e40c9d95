"""Writes the golden vectors under tests/golden/ from the reference's own
header code (oracle/_ref/libsmx_ref.so, built from /root/reference by
oracle/Makefile in the development container):

  rpc_vectors.npz    GetObjectToImageRPC_single (CoordConversion.hpp:26-87)
                     on fitted and perturbed RPCs, incl. the +-360 deg wrap
                     and the 0 / off+1.2*scale clamps
  pyr_vectors.npz    OriginalToPyramid_single (CoordConversion.hpp:96-105)
  gauss_vectors.npz  CreateImagePyramid<uint16>(9, 1.5) (Template.hpp:82-132)
  quant_vectors.npz  DoubleToSignedChar_result / _voxel (Template.hpp:964-992)
  layout.json        sizeof / offsetof of UGRID, NCCresult, VOXEL

Run: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from setsm_amd.synth import SceneSpec, make_scene  # noqa: E402


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libsmx_ref.so is not built")
    rng = np.random.default_rng(1234)
    sc = make_scene(SceneSpec(image_px=128, relief=50.0, seed=9))
    rpcs, adj, pts, out = [], [], [], []
    for k in range(6):
        rpc = sc.rpcs[k % 2].copy()
        if k >= 2:
            rpc[10:] *= 1.0 + 1e-3 * rng.standard_normal(80)
            rpc[30] = rpc[70] = 1.0
        a = rng.normal(0, 2.0, 2) if k >= 3 else np.zeros(2)
        lon0, lat0 = rpc[2], rpc[3]
        for j in range(64):
            lon = lon0 + rng.normal(0, 2 * rpc[7])
            lat = lat0 + rng.normal(0, 2 * rpc[8])
            h = rpc[4] + rng.normal(0, 2 * rpc[9])
            if j % 16 == 0:
                lon += 360.0 if lon < 0 else -360.0   # wrap branch
            if j % 16 == 1:
                lat += 30.0 * rpc[8]                  # far outside: clamp branch
            x, y = C.c_double(), C.c_double()
            R.ref_rpc_ground_to_image(rpc.ctypes.data, a.ctypes.data, lon, lat, h, C.byref(x), C.byref(y))
            rpcs.append(rpc)
            adj.append(a)
            pts.append((lon, lat, h))
            out.append((x.value, y.value))
    np.savez(os.path.join(HERE, "rpc_vectors.npz"), rpc=np.array(rpcs), adjust=np.array(adj), pts=np.array(pts),
             out=np.array(out))

    pin, pout = [], []
    for _ in range(64):
        x, y = rng.uniform(0, 20000, 2)
        sx, sy = rng.uniform(0, 500, 2)
        lv = int(rng.integers(0, 5))
        px, py = C.c_double(), C.c_double()
        R.ref_to_pyramid(x, y, sx, sy, lv, C.byref(px), C.byref(py))
        pin.append((x, y, sx, sy, lv))
        pout.append((px.value, py.value))
    np.savez(os.path.join(HERE, "pyr_vectors.npz"), inp=np.array(pin), out=np.array(pout))

    imgs, outs = {}, {}
    for k, (h, w) in enumerate([(37, 53), (64, 64), (20, 11)]):
        img = rng.integers(0, 2048, (h, w)).astype(np.uint16)
        if k == 1:
            img = sc.images[0][:64, :64].copy()
        out = np.zeros((h // 2, w // 2), np.uint16)
        R.ref_gauss_pyramid(img.ctypes.data, w, h, out.ctypes.data)
        imgs[f"img{k}"] = img
        outs[f"out{k}"] = out
    np.savez(os.path.join(HERE, "gauss_vectors.npz"), **imgs, **outs)

    vals = np.concatenate([np.array([0.0, 0.9999, -0.9999, 0.0005, -0.0005, 0.0015, -0.0015, 1.0, -1.0, -100.0,
                                     0.12345, -0.12355, 0.9995, -0.9995, 2.5e-4]), rng.uniform(-1, 1, 200)])
    qr = np.array([R.ref_q_result(float(v)) for v in vals], np.int16)
    qv = np.array([R.ref_q_voxel(float(v)) for v in vals], np.int16)
    np.savez(os.path.join(HERE, "quant_vectors.npz"), vals=vals, q_result=qr, q_voxel=qv)

    names = ["sizeof_UGRID", "UGRID.Matched_flag", "UGRID.ortho_ncc", "sizeof_NCCresult", "NCCresult.NumOfHeight",
             "NCCresult.check_height_change", "NCCresult.max_WNCC", "NCCresult.GNCC", "sizeof_VOXEL", "VOXEL.INCC",
             "NCCresult.minHeight", "NCCresult.result4"]
    lay = {n: int(R.ref_layout(i)) for i, n in enumerate(names)}
    with open(os.path.join(HERE, "layout.json"), "w") as f:
        json.dump(lay, f, indent=1)
    print("golden vectors written:", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
