"""The gather probe (mq_gather_probe): the achievable-bandwidth measurement
beside the batched worker's roofline."""
import sys

import pytest

import mqhip
import synth

pytestmark = pytest.mark.gpu


def test_gather_probe_needs_a_table_then_measures():
    e = mqhip.Engine(synth.c1_grid())
    try:
        with pytest.raises(mqhip.MqError, match="cell table"):
            e.gather_probe()
        jobs = synth.c1_jobs(mqhip, n=64)
        ph = mqhip.Physics.correct(0.133322).copy(dt=1e-7, max_steps=2000, injection=mqhip.INJ_GAUSSIAN,
                                                  inj_sigma_mm=0.1, legacy_force=1)
        e.run_batch(jobs, [0, 5, 0, 0], ph, mqhip.PATH_BATCHED)
        whole = e.gather_probe(0, 32)
        small = e.gather_probe(4096, 32)
        print(f"[gpu] gather probe: table {whole / 1e9:.0f} GB/s, 1.2 MB window {small / 1e9:.0f} GB/s", flush=True)
        assert 1e11 < whole < 1e14 and 1e11 < small < 1e14
        with pytest.raises(mqhip.MqError, match="iters"):
            e.gather_probe(0, 0)
    finally:
        e.close()
