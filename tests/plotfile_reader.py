"""Reader of the AMReX plotfiles the C++ host writes (tests only): the
top-level Header, Level_<l>/Cell_H and the FAB records of Cell_D_*."""
import os
import re

import numpy as np


def read(path):
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    assert lines[0] == "HyperCLaw-V1.1"
    nc = int(lines[1])
    names = lines[2:2 + nc]
    at = 2 + nc
    dim = int(lines[at]); time = float(lines[at + 1]); finest = int(lines[at + 2])
    prob_lo = [float(v) for v in lines[at + 3].split()]
    prob_hi = [float(v) for v in lines[at + 4].split()]
    steps = [int(v) for v in lines[at + 7].split()]
    at += 8 + (finest + 1)   # dx lines
    at += 2                   # coordinate system, boundary width
    levels = []
    for l in range(finest + 1):
        lev, nb, _ = lines[at].split()
        at += 2 + int(nb) * dim
        cell = lines[at].strip()
        at += 1
        with open(os.path.join(path, cell + "_H")) as f:
            h = f.read().split("\n")
        boxes = []
        for k in range(int(nb)):
            v = [int(x) for x in re.findall(r"-?\d+", h[5 + k])]
            boxes.append((tuple(v[:dim]), tuple(v[dim:2 * dim])))
        fabs = [ln.split() for ln in h if ln.startswith("FabOnDisk:")]
        data = []
        for (lo, hi), (_, fname, off) in zip(boxes, fabs):
            with open(os.path.join(path, "Level_%d" % l, fname), "rb") as f:
                f.seek(int(off))
                hdr = f.readline().decode()
                assert hdr.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
                n = int(np.prod([hi[d] - lo[d] + 1 for d in range(dim)]))
                data.append(np.frombuffer(f.read(8 * n * nc), dtype="<f8").reshape(nc, n))
        levels.append({"boxes": boxes, "data": data})
    return {"names": names, "dim": dim, "time": time, "finest": finest, "prob_lo": prob_lo,
            "prob_hi": prob_hi, "step": steps[0], "levels": levels}
