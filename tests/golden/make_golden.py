#!/usr/bin/env python3
"""Writes femlib_ref_vectors.json: inputs and outputs of the reference's own
Femlib element routines (compiled from the reference sources by
oracle/Makefile into oracle/_ref/femlib_ref), for seeded jittered elements.
Run in the container that holds the reference tree; the JSON is committed."""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "femlib_ref")


def run(cmd, nums):
    inp = cmd + "\n" + " ".join("%.17e" % v for v in nums) + "\nquit\n"
    out = subprocess.run([REF], input=inp, capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), [float(x) for x in out[1:]]


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/femlib_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(2024)
    out = {"hexa31": [], "andes4": [], "fqs35": []}
    for k in range(4):
        h = [0.01, 0.02, 0.015][k % 3]
        X = np.array([0, 1, 1, 0, 0, 1, 1, 0], float) * h + rng.uniform(-0.1, 0.1, 8) * h
        Y = np.array([0, 0, 1, 1, 0, 0, 1, 1], float) * h + rng.uniform(-0.1, 0.1, 8) * h
        Z = np.array([0, 0, 0, 0, 1, 1, 1, 1], float) * h + rng.uniform(-0.1, 0.1, 8) * h
        E, nu = 2.1e11, [0.3, 0.29, 0.25, 0.33][k]
        ierr, v = run("hexa31", [E, nu, *X, *Y, *Z])
        out["hexa31"].append({"E": E, "nu": nu, "X": X.tolist(), "Y": Y.tolist(), "Z": Z.tolist(),
                              "ierr": ierr, "EK": v})
    for k in range(4):
        h = [0.05, 0.1, 0.02, 0.07][k]
        X = np.array([0, 1, 1, 0], float) * h + rng.uniform(-0.15, 0.15, 4) * h
        Y = np.array([0, 0, 1, 1], float) * h + rng.uniform(-0.15, 0.15, 4) * h
        Z = rng.uniform(-0.05, 0.05, 4) * h if k % 2 else np.zeros(4)
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0] if k >= 2 else np.eye(3)
        P = (R @ np.stack([X, Y, Z])).T + rng.normal(size=3)
        E, nu, t = 2.1e11, 0.3, [0.005, 0.01, 0.002, 0.004][k]
        ierr, v = run("andes4", [E, nu, t, *P[:, 0], *P[:, 1], *P[:, 2]])
        out["andes4"].append({"E": E, "nu": nu, "t": t, "xyz": P.tolist(), "ierr": ierr,
                              "trel": v[:12], "Kloc": v[12:12 + 576], "P": v[12 + 576:]})
        rho = 7850.0
        th = (t * (1 + 0.2 * rng.uniform(-1, 1, 4))).tolist()
        ierr, v = run("fqs35", [rho, *P[:, 0], *P[:, 1], *P[:, 2], *th])
        out["fqs35"].append({"rho": rho, "xyz": P.tolist(), "th": th, "EM": v})
    with open(os.path.join(HERE, "femlib_ref_vectors.json"), "w") as f:
        json.dump(out, f)
    print("wrote", {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
