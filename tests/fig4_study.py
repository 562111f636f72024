"""The legacy Fig. 4 scan (channel.rs:739-760, compare_theirs 683-708) under the
variants that bear on the published columns (tests/golden/mass-scan.csv). A
study, not a test: run it on a GPU as `python tests/fig4_study.py [out.json]`;
it prints one line per run and writes every count to out.json.

* regenerated Java potentials (include/mq_legacy.h, Fields::from_theirs),
  dc [0, 5, 0, 0]: 20 sets of 1000 ions per m/z on the FP32 path (Philox
  seeds 1000 + m/z + 7919 k); the reference's seeds on both paths with the
  COMPAT streams; 12 sets per path at the flank m/z 489-505 (seeds
  1000 + m/z + 104729 k);
* converged default-geometry fields: dc [0, 5, -5, 0] (legacy-fig4), the
  downstream electrode at 0 V, trilinear fields (legacy_fields off), the
  COMPAT streams; the corrected S14 geometry.

FWHM follows channel.rs:512-520.
"""
import csv
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", "multiq-workbench_amd")]
import mqhip  # noqa: E402

ROWS = list(csv.DictReader(open(os.path.join(HERE, "golden", "mass-scan.csv"))))
MZ = [float(r["mz"]) for r in ROWS]
FLANK = [489.0, 491.0, 493.0, 495.0, 497.0, 499.0, 501.0, 503.0, 505.0]


def fwhm(xs, ys):
    """channel.rs:512-520: peak position, height and width at half height (linear between points)."""
    im = int(np.argmax(ys))
    ym = ys[im]
    half = ym / 2.0
    lo, hi = xs[0], xs[-1]
    for i in range(im, 0, -1):
        if ys[i - 1] < half:
            lo = xs[i - 1] + (half - ys[i - 1]) / (ys[i] - ys[i - 1]) * (xs[i] - xs[i - 1])
            break
    for i in range(im, len(xs) - 1):
        if ys[i + 1] < half:
            hi = xs[i] + (ys[i] - half) / (ys[i] - ys[i + 1]) * (xs[i + 1] - xs[i])
            break
    return xs[im], float(ym), hi - lo


def jobs(n, off=0, mzs=MZ):
    return [mqhip.job(mz * 5, 5.0, 0.0, 100.0, 500e3, 0.5, 100e3, n, 1000 + int(mz) + off) for mz in mzs]


def main():
    path_out = sys.argv[1] if len(sys.argv) > 1 else None
    out = {"mz": MZ, "flank_mz": FLANK}
    t0 = time.time()

    def log(key, counts, xs=MZ):
        out[key] = counts
        m = np.asarray(counts, float)
        m = m.mean(0) if m.ndim == 2 else m
        print(f"[{time.time() - t0:.0f}s] {key}: {np.round(m, 1).tolist()} peak/FWHM {fwhm(xs, list(m))}",
              flush=True)
        if path_out:
            json.dump(out, open(path_out, "w"), indent=1)

    ph = mqhip.Physics.legacy(1e-3 * 133.322)
    d = tempfile.mkdtemp()
    mqhip.legacy_export(d)
    e = mqhip.Engine(mqhip.fields_from_theirs(d))
    try:
        log("exported_fp32_20x1000", [[o.transmitted for o in e.run_batch(jobs(1000, 7919 * (k + 1)), [0, 5, 0, 0], ph)]
                                      for k in range(20)])
        for path in (mqhip.PATH_BATCHED, mqhip.PATH_PLAIN):
            log(f"exported_compat_refseeds_path{path}",
                [o.transmitted for o in e.run_batch(jobs(1000), [0, 5, 0, 0], ph.copy(rng=mqhip.RNG_COMPAT), path)])
            log(f"exported_flank_path{path}_12x1000",
                [[o.transmitted for o in e.run_batch(jobs(1000, 104729 * (k + 1), FLANK), [0, 5, 0, 0], ph, path)]
                 for k in range(12)], FLANK)
    finally:
        e.close()
    e = mqhip.Engine(mqhip.channel_grid(mqhip.Geom.default(), tol=1e-7))
    try:
        for key, dc, p in (("converged_m5", [0, 5, -5, 0], ph), ("converged_down0", [0, 5, 0, 0], ph),
                           ("converged_trilinear_m5", [0, 5, -5, 0], ph.copy(legacy_fields=0))):
            log(key + "_4x1000", [[o.transmitted for o in e.run_batch(jobs(1000, 7919 * (k + 1)), dc, p)]
                                  for k in range(4)])
        log("converged_compat_refseeds_m5",
            [o.transmitted for o in e.run_batch(jobs(1000), [0, 5, -5, 0], ph.copy(rng=mqhip.RNG_COMPAT))])
    finally:
        e.close()
    e = mqhip.Engine(mqhip.channel_grid(mqhip.Geom.s14(), tol=1e-7))
    try:
        log("s14_m5_4x1000", [[o.transmitted for o in e.run_batch(jobs(1000, 7919 * (k + 1)), [0, 5, -5, 0], ph)]
                              for k in range(4)])
    finally:
        e.close()


if __name__ == "__main__":
    main()
