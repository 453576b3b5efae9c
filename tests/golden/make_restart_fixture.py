"""Write ``site_restart_4cells.npz``: four cells of the reference's saved state
``verify_out_stage1/site_transient_1996.npz`` (171 site cells, float32 / int32, 258 leaves),
kept in the reference's restart layout (``treedef_repr``, ``num_leaves``, ``leaf_<i>``,
``meta_<key>``; python-port/ibis/io/restart.py:19-55).  Per-cell leaves (first axis = 171)
are sliced to cells 0, 57, 114 and 170; global leaves and the metadata are copied as they
are.  Only ``numpy.load`` with ``allow_pickle=False`` touches the file.

Run from the repository root:  python tests/golden/make_restart_fixture.py
"""
import os

import numpy as np

SRC = "/root/reference/verify_out_stage1/site_transient_1996.npz"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "site_restart_4cells.npz")
CELLS = np.array([0, 57, 114, 170])


def main():
    z = np.load(SRC, allow_pickle=False)
    num = int(z["num_leaves"])
    ncell = z["leaf_0"].shape[0]
    out = {"treedef_repr": str(z["treedef_repr"]), "num_leaves": num}
    for i in range(num):
        if f"leaf_{i}_is_none" in z.files:
            out[f"leaf_{i}_is_none"] = True
            continue
        a = z[f"leaf_{i}"]
        out[f"leaf_{i}"] = a[CELLS] if a.ndim and a.shape[0] == ncell else a
    for k in z.files:
        if k.startswith("meta_"):
            out[k] = z[k]
    np.savez_compressed(DST, **out)
    print(f"{DST}: {len(CELLS)} of {ncell} cells, {num} leaves")


if __name__ == "__main__":
    main()
