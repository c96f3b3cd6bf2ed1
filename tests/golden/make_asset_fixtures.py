"""Summarise the reference's target mass grids (plb/envs/assets/*.npy, data
files the reference ships) into a small JSON fixture: per grid its total mass
and the coordinates of its negative-mass nodes.  Run in a container that has
the reference mounted; the JSON is committed."""
import glob
import json
import os

import numpy as np

SRC = "/root/reference/plb/envs/assets"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "asset_grids.json")


def main():
    out = {}
    for f in sorted(glob.glob(os.path.join(SRC, "*.npy"))):
        a = np.load(f)  # allow_pickle defaults to False
        neg = np.argwhere(a < 0)
        out[os.path.basename(f)] = dict(shape=list(a.shape), sum=float(a.sum()), max=float(a.max()),
                                        nonzero=int((a != 0).sum()), negative_nodes=neg.tolist())
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", OUT, len(out))


if __name__ == "__main__":
    main()
