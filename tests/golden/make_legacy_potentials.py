"""Writes legacy_potentials.json: digests of the oracle's legacy Java potentials
at the sketch's full size (160 x 160 x 242, scl 10, 0.001 %, cap 6000).

The oracle (oracle/oracle_legacy.c) runs the sketch's sweeps one node after
the other, which takes minutes per basis on one core, too long for a test. So
this script runs it once, the three bases in parallel processes, and stores
per basis the sweep count, the last change at the check point, the SHA-256 of
the float32 potentials (little-endian, node-major i + nx (j + ny k)) and of the
electrode flags, and the on-axis potential every 10 planes (for diagnosis).
The GPU regeneration must reproduce every digest.

Usage: python tests/golden/make_legacy_potentials.py
"""
import hashlib
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "multiq-workbench_amd"), os.path.join(ROOT, "oracle")]

SCL, PRECISION, MAX_ITER = 10.0, 0.001, 6000
NAMES = ("RF", "DC", "AC")


def one(basis: int) -> dict:
    import oracle

    nx, ny, nz, check = oracle.legacy_dims(SCL)
    phi0, el = oracle.legacy_rasterize(basis, SCL)
    phi, nit, dif, trace = oracle.legacy_sweeps(phi0, el, check, PRECISION, MAX_ITER)
    a = nx // 2
    return {
        "basis": NAMES[basis],
        "dims": [nx, ny, nz],
        "check": list(check),
        "sweeps": int(nit),
        "dif_pct": float(dif),
        "check_last": [float(v) for v in trace[-3:]],
        "sha256_phi": hashlib.sha256(phi.astype("<f4").tobytes()).hexdigest(),
        "sha256_electrode": hashlib.sha256(el.tobytes()).hexdigest(),
        "axis_every_10": [float(phi[k, a, a]) for k in range(0, nz, 10)],
    }


def main() -> None:
    with ProcessPoolExecutor(3) as ex:
        res = list(ex.map(one, range(3)))
    out = {"scl": SCL, "precision_pct": PRECISION, "max_iter": MAX_ITER, "bases": res}
    with open(os.path.join(HERE, "legacy_potentials.json"), "w") as f:
        json.dump(out, f, indent=1)
    for r in res:
        print(r["basis"], r["sweeps"], r["dif_pct"], r["sha256_phi"][:16])


if __name__ == "__main__":
    main()
