"""Writes footprint_count_cone256.json: the number of footprint weights of
the cone 256^3 benchmark (256^3 voxels of 1 mm, 360 views over 360 degrees,
384 x 512 flat detector of 1 mm pixels, source 1000 mm and detector 1500 mm
from the source, cubic footprint), counted view by view with the oracle's
FP64 geometry pass (sum over the voxels of npix_u * npix_v, the window sizes
of trtt_3D_projectors.c:1293-1314).

Each weight costs the reference one multiply-add in the forward projection
and one in the back projection, so a forward + back pair does
4 * total floating-point operations in its footprint loops.

    python tests/golden/make_footprint_count.py   # about a minute on 8 cores
"""
import json
import math
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
OUT = os.path.join(HERE, "footprint_count_cone256.json")

NX = NY = NZ = 256
NU, NV = 384, 512
NVIEWS = 360
RSC, RSD = 1000.0, 1500.0


def bench_geometry():
    from yoricktrtt_amd.geometry import cone_geometry

    theta = np.arange(NVIEWS) * (2 * math.pi / NVIEWS)
    return cone_geometry(NX, NY, NZ, 1.0, NU, NV, 1.0, 1.0, theta, RSC, RSD)


def count_views(oracle, g, views, workers=8):
    """Footprint weights of each view in ``views`` (ctypes drops the GIL, so
    the views run on ``workers`` threads)."""
    def one(v):
        return oracle.count_pairs(g.nx, g.ny, g.nz, g.s_scl, g.nu, g.nv, g.u_scl, g.v_scl,
                                  g.xos[v], g.xsd[v], True)

    with ThreadPoolExecutor(workers) as ex:
        return [int(c) for c in ex.map(one, views)]


def main():
    from oracle.oracle import Oracle, build

    build()
    g = bench_geometry()
    per_view = count_views(Oracle(), g, range(NVIEWS))
    total = sum(per_view)
    doc = {
        "workload": "cone 256^3 x 1 mm, 360 views / 360 deg, 384 x 512 detector of 1 mm, "
                    "Rsc 1000 mm, Rsd 1500 mm, cubic footprint",
        "per_view": per_view,
        "total": total,
        "mean_per_voxel_view": total / (NX * NY * NZ * NVIEWS),
        "flops_per_pair": 4 * total,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote", OUT, "total", total, "mean", doc["mean_per_voxel_view"])


if __name__ == "__main__":
    main()
