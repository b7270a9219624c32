"""Writes tests/golden/rcb_vectors.json: the known-answer vectors for RCB.

No reference code is run: every expected array below was traced by hand
through coupe's algorithm (coupe/src/algorithms/recursive_bisection.rs
:456-705) and each trace is written next to its vector.  V1-V3 come from
the reference's own tests and examples; V4-V9 are built to hit each exit of
the split loop.  The AABB vectors are the reference's geometry tests.

Run:  python tests/golden/make_vectors.py
"""
import json
import os

VECTORS = [
    dict(
        id="V1",
        source="coupe/src/algorithms/recursive_bisection.rs:1077-1115 (test_rcb_basic)",
        points=[[-1.3, 6], [2, -4], [1, 1], [-3, -2.5], [-1.3, -0.3], [2, 1], [-3, 1], [1.3, -2]],
        weights=[1.0] * 8, weight_type="double", constant=False,
        iter_count=2, tolerance=0.05,
        expected=[1, 2, 3, 0, 0, 3, 1, 2],
        trace="x: bbox [-3,2], t=-0.5 -> left {0,3,4,6}, imbalance 0. "
              "y in both children: bbox [-4,6], t=1 -> {3,4}|{0,6} and {1,7}|{2,5}. "
              "leaf heap ids 3..6, offset 3.",
    ),
    dict(
        id="V2",
        source="coupe/src/algorithms/recursive_bisection.rs:739-768 (Rcb doctest)",
        points=[[1, 1], [-1, 1], [1, -1], [-1, -1]],
        weights=[1] * 4, weight_type="int", constant=False,
        iter_count=2, tolerance=0.0,
        expected=[3, 1, 2, 0],
        trace="Default tolerance 0.0 is met by the exact balance at t=0 on x then on y.",
    ),
    dict(
        id="V3a",
        source="coupe-ffi/examples/rcb.c:18-39 (iter_count 1)",
        points=[[0, 0], [0, 1], [1, 0], [1, 1]],
        weights=[1], weight_type="int", constant=True,
        iter_count=1, tolerance=0.05,
        expected=[0, 0, 1, 1],
        trace="x: t=0.5 splits {0,1}|{2,3}.",
    ),
    dict(
        id="V3b",
        source="coupe-ffi/examples/rcb.c:41-53 (iter_count 2)",
        points=[[0, 0], [0, 1], [1, 0], [1, 1]],
        weights=[1], weight_type="int", constant=True,
        iter_count=2, tolerance=0.05,
        expected=[0, 1, 2, 3],
        trace="then y: t=0.5 in both halves.",
    ),
    dict(
        id="V4",
        source="hand-derived: several candidates, exit on max <= t + d (:552-563)",
        points=[[x, 0] for x in range(8)],
        weights=[1.0] * 7 + [10.0], weight_type="double", constant=False,
        iter_count=1, tolerance=0.05,
        expected=[0, 0, 0, 0, 0, 0, 0, 1],
        trace="S=17. t=3.5: wl=4, imb .53, go right. t=5.25: wl=6, go right. "
              "t=6.125: nearest 7, d=0.875, max 7 <= 7 -> stop, left 0..6.",
    ),
    dict(
        id="V5",
        source="hand-derived: None-branch twice, all points left (:522-545)",
        points=[[0, 0], [0, 1], [10, 0], [10, 100]],
        weights=[1.0] * 4, weight_type="double", constant=False,
        iter_count=2, tolerance=0.05,
        expected=[0, 0, 2, 3],
        trace="x: t=5 balanced. Left child on y [0,100]: t=50 nothing right -> max=50; "
              "t=25 nothing right again -> all left, split_pos=50, heap id 4 empty (gap).",
    ),
    dict(
        id="V6",
        source="hand-derived: count_left == prev exit leaves the leftmost leaf empty",
        points=[[0, 90], [0, 95], [10, 0], [10, 100]],
        weights=[1.0] * 4, weight_type="double", constant=False,
        iter_count=2, tolerance=0.05,
        expected=[0, 0, 1, 2],
        trace="Left child on y [0,100]: t=50 count 0, imb 1, go right; t=75 count 0 == prev -> "
              "split with empty left: heap id 3 empty, offset 4.",
    ),
    dict(
        id="V7",
        source="hand-derived: ties at t go right, then exit on max <= t + d",
        points=[[0, 0], [1, 0], [1, 0], [2, 0]],
        weights=[1.0] * 4, weight_type="double", constant=False,
        iter_count=1, tolerance=0.05,
        expected=[0, 0, 0, 1],
        trace="t=1: only point 0 is < t, imb .5, go right; t=1.5: left {0,1,2}, nearest 2, "
              "d=0.5, max 2 <= 2 -> stop.",
    ),
    dict(
        id="V8",
        source="hand-derived: the imbalance stop at equality, imb == tolerance (:547-554)",
        points=[[0, 0], [1, 0], [2, 0], [6, 0], [7, 0], [10, 0]],
        weights=[1.0, 1.0, 1.0, 2.0, 2.0, 1.0], weight_type="double", constant=False,
        iter_count=1, tolerance=0.25,
        expected=[0, 0, 0, 1, 1, 1],
        trace="S=8. t=5: left {0,1,2}, wl=3, imb=|3-4|/4=0.25 <= 0.25 -> stop (nearest 6, "
              "6 <= t+d=6 < max 10, so only the imbalance test stops here). With imb < tol the loop "
              "would go right (3 < 5) to t=7.5 and stop there on max <= t + d with left {0,..,4}.",
    ),
    dict(
        id="V9",
        source="hand-derived: equal halves move left (wl < S - wl is false), negative tolerance (:564-571)",
        points=[[0, 0], [1, 0], [2, 0], [6, 0], [7, 0], [10, 0]],
        weights=[1.0, 1.0, 2.0, 1.0, 1.0, 2.0], weight_type="double", constant=False,
        iter_count=1, tolerance=-1.0,
        expected=[0, 0, 0, 1, 1, 1],
        trace="S=8. t=5: left {0,1,2}, wl=4 = S-wl, imb 0 is not <= -1, 6 < max 10: no stop; 4 < 4 is "
              "false -> max=5. t=2.5: left {0,1,2}, count 3 == prev -> stop.  Moving right on equality "
              "would reach t=7.5 and stop with left {0,..,4}.",
    ),
]

AABB = [
    dict(id="aabb2d", source="coupe/src/geometry.rs:341-355",
         points=[[1, 2], [0, 0], [3, 1], [5, 4], [4, 5]], p_min=[0, 0], p_max=[5, 5]),
    dict(id="aabb3d", source="coupe/src/geometry.rs:489-503",
         points=[[1, 2, 0], [0, 0, 5], [3, 1, 1], [5, 4, -2], [4, 5, 3]],
         p_min=[0, 0, -2], p_max=[5, 5, 5]),
]


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rcb_vectors.json")
    with open(out, "w") as f:
        json.dump({"rcb": VECTORS, "aabb": AABB}, f, indent=1)
        f.write("\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
