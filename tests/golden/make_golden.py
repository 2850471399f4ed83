"""Records golden vectors from the reference's own CPU HashDAG
(svt/hash_dag.hpp, compiled from the reference sources into oracle/_ref by
oracle/Makefile). Run in the container that holds the reference tree:

    make -C oracle && python tests/golden/make_golden.py

Each case is an edit script of half-open boxes (x0 y0 z0 x1 y1 z1 fill) and
the canonical export of the reference's DAG after it (counts + FNV digest,
plus the full export for small cases). The 512^3 case applies the config-0
heightfield as column boxes and its 64 spheres as x-row boxes, in brush order.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def run_ref(depth, ops):
    d = O.RefDag(depth)
    for op in ops:
        d.box(op[0:3], op[3:6], bool(op[6]))
    return d.export()


def case_from_ops(name, depth, ops, keep_ops=True, keep_export=False):
    e = run_ref(depth, ops)
    c = {"name": name, "depth": depth, "nodes": int(e[4]), "leaves": int(e[5]), "digest": "%016x" % O.digest(e)}
    if keep_ops:
        c["ops"] = [list(map(int, op)) for op in ops]
    if keep_export:
        c["export"] = e.tolist()
    return c


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle)"
    cases = []
    # Known-answer scripts (test/svt/test_svt_hash_dag.cpp-style).
    cases.append(case_from_ops("single_voxel", 3, [(5, 6, 7, 6, 7, 8, 1)], keep_export=True))
    cases.append(case_from_ops("eight_partial_siblings", 3,
                               [(x * 4, y * 4, z * 4, x * 4 + 1, y * 4 + 1, z * 4 + 1, 1)
                                for z in range(2) for y in range(2) for x in range(2)], keep_export=True))
    cases.append(case_from_ops("fill_then_carve", 3, [(0, 0, 0, 32, 32, 32, 1), (3, 4, 5, 20, 9, 30, 0)],
                               keep_export=True))
    # Seeded random box scripts.
    for depth, nops, seed in ((4, 120, 11), (6, 300, 12)):
        w = 4 << depth
        r = O.gen_u32(seed, 99, nops * 7, 0, 0xFFFFFFFF)
        ops = []
        for i in range(nops):
            lo = [int(r[7 * i + a]) % w for a in range(3)]
            hi = [min(w, lo[a] + 1 + int(r[7 * i + 3 + a]) % (w // 2)) for a in range(3)]
            ops.append((*lo, *hi, int(r[7 * i + 6]) & 1))
        cases.append(case_from_ops(f"random_boxes_d{depth}", depth, ops))
    with open(os.path.join(os.path.dirname(__file__), "ref_hash_dag.json"), "w") as f:
        json.dump(cases, f)

    # Config 0 (512^3, inner depth 7), occupancy: terrain columns + 64 spheres.
    depth, world = 7, 512
    d = O.RefDag(depth)
    d.boxes(O.terrain_column_boxes(O.terrain_for_world(world, 1), world))
    e_terrain = d.export()
    brushes = O.gen_brushes(1, O.STREAM_BRUSH, 64, world, 0)
    d.boxes(O.sphere_row_boxes(brushes, world))
    e = d.export()
    c0 = {"name": "config0_occupancy", "depth": depth, "seed": 1, "brushes": 64, "mix": 0,
          "terrain": {"nodes": int(e_terrain[4]), "leaves": int(e_terrain[5]),
                      "digest": "%016x" % O.digest(e_terrain)},
          "edited": {"nodes": int(e[4]), "leaves": int(e[5]), "digest": "%016x" % O.digest(e)}}
    with open(os.path.join(os.path.dirname(__file__), "ref_config0.json"), "w") as f:
        json.dump(c0, f, indent=1)
    print("wrote", len(cases), "cases + config0", c0)


if __name__ == "__main__":
    main()
