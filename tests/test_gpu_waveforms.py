"""GPU: the waveform code of every transient kernel against krets' source
semantics (Pulse::value_at / Sinusoidal::value_at, voltage_source.rs:50-93,
and the lagged source time of transient.rs, quirk 1).

A resistor divider V(out) = 3/4 V(in) driven by one PULSE source (rise,
hold, fall, low and three period wraps) and one damped, delayed,
phase-shifted SIN source: with krets' backward Euler the divider node at
step n equals 3/4 * value_at(t_{n-1}); with the clean integrators,
3/4 * value_at(t_n). value_at is the oracle's, itself pinned to krets'
known answers (tests/test_oracle.py). A capacitor across the source gives
the kernels a dynamic element without touching either node voltage.
Checked through every transient mapping: level-scheduled (1..256 threads
per system, plans staged in LDS), LDS-resident (3000 + threads), and the
lane kernel specialised at run time (2001).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WAVES = ("PULSE(0.5 2.5 1u 1u 2u 3u 10u)", "SIN(0.5 2 100k 5u 2e4 90)")
NET = tuple(f"V1 in 0 0.25 {w}\nC1 in 0 1n\nR1 in out 1k\nR2 out 0 3k\n" for w in WAVES)
# the parsed parameters (v1 v2 td tr tf pw per; vo va f td theta phase): SI suffixes multiply,
# so td = 5u is 5 * 1e-6, one ulp below 5e-6, which decides the step at the sine's delay, and
# per = 10u is 10 * 1e-6, one ulp below 1e-5
PULSE = (0.5, 2.5, 1e-6, 1e-6, 2e-6, 3 * 1e-6, 10 * 1e-6)
SIN = (0.5, 2.0, 100e3, 5 * 1e-6, 2e4, 90.0)
H, NSTEPS = 0.1e-6, 400  # 40 us: four pulse periods, four sine periods after the delay


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch():
    import krets_amd as kg
    cs = [kg.Circuit(t) for t in NET]
    assert cs[0].names == cs[1].names
    assert tuple(cs[0].params[0, 3:10]) == PULSE and tuple(cs[1].params[0, 3:9]) == SIN
    ctx = kg.Context()
    ctx.load(cs[0].elements, np.stack([c.params for c in cs]), cs[0].n)
    return kg, ctx, cs[0].names.index("V(out)")


def expected(oracle, lag):
    out = np.zeros((2, NSTEPS + 1))
    for i, (kind, p) in enumerate(((1, PULSE), (2, SIN))):
        out[i, 0] = 0.75 * 0.25  # the operating point uses the dc value
        for n in range(1, NSTEPS + 1):
            out[i, n] = 0.75 * oracle.wave_value(kind, p, (n - 1 if lag else n) * H)
    return out


def test_expected_values_cover_every_branch(oracle):
    """The sampled times reach the pulse's rise, hold, fall and low segments
    in every period, and the sine's delay and damped part."""
    t = np.arange(NSTEPS + 1) * H
    tt = np.fmod(np.maximum(t - PULSE[2], 0), PULSE[6])
    after = t >= PULSE[2]
    for lo, hi in ((0, 1e-6), (1e-6, 4e-6), (4e-6, 6e-6), (6e-6, 10e-6)):
        assert ((tt > lo + 1e-9) & (tt < hi - 1e-9) & after & (t > 30e-6)).any()
    e = expected(oracle, lag=False)
    assert np.all(e[1, t < SIN[3]][1:] == 0.75 * SIN[0]) and np.ptp(e[1, t > 30e-6]) > 0.5


MAPS = [1, 4, 16, 64, 256, 1064, 1256, 3064, 3128, 3256, 3512, 2001]


@pytest.mark.parametrize("method", [0, 1, 2])  # BE_COMPAT, BE, TRAP
@pytest.mark.parametrize("tps", MAPS)
def test_divider_samples_the_waveforms(batch, oracle, tps, method):
    kg, ctx, k = batch
    say(f"waveform divider tps={tps} method={method}")
    cfg = kg.SolverConfig(threads_per_system=tps)
    xs, st, _ = ctx.transient(H, NSTEPS, method=method, config=cfg)
    assert (st == 0).all(), st
    e = expected(oracle, lag=method == kg.BE_COMPAT)
    for i in range(2):
        d = np.abs(xs[i][:, k] - e[i])
        assert np.max(d) <= 1e-12 * np.max(np.abs(e[i])), (i, int(np.argmax(d)), np.max(d))


@pytest.mark.parametrize("method", [0, 2])
def test_divider_samples_the_waveforms_mixed(batch, oracle, method):
    kg, ctx, k = batch
    say(f"waveform divider mixed method={method}")
    xs, st, _ = ctx.transient(H, NSTEPS, method=method, precision=kg.MIXED,
                              config=kg.SolverConfig(threads_per_system=64))
    assert (st == 0).all(), st
    e = expected(oracle, lag=method == kg.BE_COMPAT)
    for i in range(2):
        assert np.max(np.abs(xs[i][:, k] - e[i])) <= 1e-9 * np.max(np.abs(e[i]))
