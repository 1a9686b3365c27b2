"""Derive the small committed fixtures from the reference's data files.

Run in the container that holds the reference checkout (read-only):
    python tests/golden/make_fixtures.py /root/reference/input_files

Everything here is computed with plain numpy directly from the data files
(independently of oracle/ and of the product), and only small derived
numbers are written -- never the data files themselves:

  grid_invariants.json   per shipped grid level (src/mesh.cpp:4016-4101 format):
                         counts, pentagon ids, a SHA-256 of the parsed arrays
                         (float64 radians = atof(deg) * pi/180), the sum of
                         d_e l_e / 4 pi R^2 (node-node arc times Voronoi-edge
                         arc, the face_area of src/mesh.cpp:1093), the sum of
                         spherical Voronoi-cell areas / 4 pi R^2, and the
                         circumcentre / centroid residuals of the tessellation.
  love_tables.json       h_s = 20 km column of the shipped Enceladus beta table
                         (rows l = 1..18), upsilon_2(20 km), nu_2(20 km).
"""
import hashlib
import json
import os
import sys

import numpy as np

RAD = 3.1415926535897932384626433832795028841971693993751058 / 180.0


def parse_grid(path):
    rows = {}
    with open(path) as fh:
        fh.readline()
        for line in fh:
            if "{" not in line:
                continue
            head, rest = line.split("{", 1)
            parts = head.split()
            fr_s, rest = rest.split("}", 1)
            fr = [int(x) for x in fr_s.split(",")]
            cen = rest.split("{", 1)[1].split("}", 1)[0]
            nums = [float(t) for t in cen.replace("(", " ").replace(")", " ").replace(",", " ").split()]
            rows[int(parts[0])] = (float(parts[1]), float(parts[2]), fr, nums)
    n = len(rows)
    lat = np.array([rows[i][0] for i in range(n)]) * RAD
    lon = np.array([rows[i][1] for i in range(n)]) * RAD
    fr = np.array([rows[i][2] for i in range(n)], np.int32)
    c = np.array([rows[i][3] for i in range(n)]).reshape(n, 6, 2) * RAD
    return lat, lon, fr, c[:, :, 0].copy(), c[:, :, 1].copy()


def xyz(lat, lon):
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], -1)


def arc(a, b):
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.sum(a * b, -1))


def tri_solid_angle(a, b, c):
    # Van Oosterom & Strackee (1983)
    num = np.abs(np.sum(a * np.cross(b, c), -1))
    den = 1 + np.sum(a * b, -1) + np.sum(b * c, -1) + np.sum(c * a, -1)
    return 2 * np.arctan2(num, den)


def grid_invariants(path):
    lat, lon, fr, clat, clon = parse_grid(path)
    n = len(lat)
    h = hashlib.sha256()
    for a in (lat, lon, fr.astype(np.int64), clat, clon):
        h.update(np.ascontiguousarray(a).tobytes())
    P = xyz(lat, lon)
    C = xyz(clat, clon)
    deg = np.where(fr[:, 5] < 0, 5, 6)
    sum_de_le = 0.0
    area = 0.0
    cc_res = 0.0
    cen_res = 0.0
    faces = 0
    for i in range(n):
        k = deg[i]
        acc = np.zeros(3)
        a_i = 0.0
        for j in range(k):
            f = fr[i, j]
            v0, v1 = C[i, (j - 1) % k], C[i, j]
            if i < f:
                sum_de_le += arc(P[i], P[f]) * arc(v0, v1)
                faces += 1
            t = tri_solid_angle(P[i], C[i, j], C[i, (j + 1) % k])
            a_i += t
            acc += t * (P[i] + C[i, j] + C[i, (j + 1) % k])
            # vertex j is equidistant from node, friend j and friend j+1
            dd = [arc(C[i, j], P[q]) for q in (i, fr[i, j], fr[i, (j + 1) % k])]
            cc_res = max(cc_res, max(dd) - min(dd))
        area += a_i
        cen = acc / np.linalg.norm(acc)
        spacing = np.mean([arc(P[i], P[fr[i, j]]) for j in range(k)])
        cen_res = max(cen_res, arc(cen, P[i]) / spacing)
    return {
        "n_cells": int(n),
        "n_faces": int(faces),
        "pentagons": [int(i) for i in np.where(fr[:, 5] < 0)[0]],
        "sha256_parsed": h.hexdigest(),
        "sum_de_le_over_4pi": sum_de_le / (4 * np.pi),
        "sum_cell_area_over_4pi": area / (4 * np.pi),
        "max_circumcentre_residual_rad": cc_res,
        "max_centroid_offset_in_spacings": cen_res,
        "min_node_dist_rad": float(min(arc(P[i], P[fr[i, j]]) for i in range(n) for j in range(deg[i]))),
    }


def main(inp):
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for lvl in (3, 4, 5, 6):
        p = os.path.join(inp, "grid_l%d.txt" % lvl)
        out["grid_l%d" % lvl] = grid_invariants(p)
        print(lvl, out["grid_l%d" % lvl])
    with open(os.path.join(here, "grid_invariants.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    d = os.path.join(inp, "LOVE_SHELL_COEFFS", "ENCELADUS")
    beta = np.loadtxt(os.path.join(d, "beta_hs_1km_to_50km_lmax30.txt"))
    ups = np.loadtxt(os.path.join(d, "upsilon_hs_1km_to_50km_lmax2.txt"))
    nu2 = np.loadtxt(os.path.join(d, "nu_hs1km_to_hs100km_l2.txt"))
    love = {"beta_hs20km_l1_to_l18": [float(x) for x in beta[:18, 19]],
            "upsilon2_hs20km": float(ups[19]), "nu2_hs20km": float(nu2[19]),
            "note": "row r of the beta table is degree l = r + 1; column c is h_s = c + 1 km"}
    with open(os.path.join(here, "love_tables.json"), "w") as fh:
        json.dump(love, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/input_files")
