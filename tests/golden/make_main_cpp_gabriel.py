"""Write main_cpp_gabriel.json: the 11-point 2-D Gabriel fixture of the reference's main.cpp.

Input: the point table at main.cpp:34-46 (NaiveGabriel, r_max = 2.5 at main.cpp:17, default
gabriel_coefficient 0.8 from naive_gabriel.hpp:383). The reference prints the per-cell
neighbour counts (main.cpp:70-73) but stores no expected values, so the expected Gabriel
graph is computed here by brute force in float64 on the float32-rounded points, with the
predicate of naive_gabriel.hpp:68-80: pair (i, j) with |p_i - p_j| < r_max is kept unless some
k != i, j lies strictly inside the ball of radius c*|p_i - p_j|/2 around the midpoint. The
script also records the smallest margin of any decision so a test can see the float32 result
cannot differ.
"""
import json
from pathlib import Path

import numpy as np

POINTS = [
    (-3.1384, -1.4465), (0.0387, -1.8340), (2.2085, -2.5314), (-2.0018, -0.2325),
    (0.7878, -0.4391), (1.3044, -0.7232), (-3.5000, 1.3432), (-0.5554, 0.7491),
    (1.2528, 1.1365), (3.5000, 0.3875), (0.0129, 2.5314),
]
R_MAX = 2.5
C = 0.8


def main():
    p = np.asarray(POINTS, np.float32).astype(np.float64)
    n = len(p)
    pairs = []
    margin = np.inf
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            d2 = float(((p[i] - p[j]) ** 2).sum())
            margin = min(margin, abs(d2 - R_MAX * R_MAX))
            if d2 >= R_MAX * R_MAX:
                continue
            mid = p[i] - 0.5 * (p[i] - p[j])
            rho2 = d2 * 0.25 * C * C
            blocked = False
            for k in range(n):
                if k in (i, j):
                    continue
                o2 = float(((p[k] - mid) ** 2).sum())
                margin = min(margin, abs(o2 - rho2))
                if o2 < rho2:
                    blocked = True
            if not blocked:
                pairs.append([i, j])
    counts = [0] * n
    for i, _ in pairs:
        counts[i] += 1
    out = {
        "source": "main.cpp:34-46 point table, NaiveGabriel r_max=2.5 c=0.8, float64 brute force",
        "points": [list(map(float, q)) for q in np.asarray(POINTS, np.float32)],
        "r_max": R_MAX,
        "gabriel_coefficient": C,
        "ordered_pairs": pairs,
        "neighbour_counts": counts,
        "total_links": len(pairs) // 2,
        "min_decision_margin": margin,
    }
    Path(__file__).with_name("main_cpp_gabriel.json").write_text(json.dumps(out, indent=1))
    print(counts, len(pairs) // 2, margin)


if __name__ == "__main__":
    main()
