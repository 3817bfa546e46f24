"""Writes boule_slice.npz: a slice of the reference's simulated ball data,
Tomography_3D/Simul_Data/proj_boule.data (720 FP32 views of 32 x 32 after a
512-byte header, made by that folder's sim_cone from boule.obj, a centred
ball) and boule.data (the 16^3 FP32 ball after a 512-byte header), read as
the bytes the files hold.

Kept: both headers, the views VIEWS of the projection file in stored order
([view][nv][nu], u fastest) and the whole volume ([nz][ny][nx]).  The file
sizes are kept too, so a reader test can rebuild a stand-in of the right
length when the reference is absent.

    python tests/golden/make_boule_slice.py   # needs /root/reference
"""
import os

import numpy as np

SRC = "/root/reference/Tomography_3D/Simul_Data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boule_slice.npz")
VIEWS = np.array([0, 1, 180, 360, 719])
NU = NV = 32
NSTORED = 720
N = 16
HEADER = 512


def main():
    with open(os.path.join(SRC, "proj_boule.data"), "rb") as f:
        pbytes = f.read()
    with open(os.path.join(SRC, "boule.data"), "rb") as f:
        vbytes = f.read()
    assert len(pbytes) == HEADER + NSTORED * NV * NU * 4
    assert len(vbytes) == HEADER + N ** 3 * 4
    proj = np.frombuffer(pbytes, dtype=np.float32, offset=HEADER).reshape(NSTORED, NV, NU)
    vol = np.frombuffer(vbytes, dtype=np.float32, offset=HEADER).reshape(N, N, N)
    np.savez_compressed(
        OUT, views=VIEWS, proj=proj[VIEWS].copy(), proj_header=np.frombuffer(pbytes[:HEADER], np.uint8),
        vol=vol.copy(), vol_header=np.frombuffer(vbytes[:HEADER], np.uint8),
        proj_size=len(pbytes), vol_size=len(vbytes))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
