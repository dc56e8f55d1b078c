"""Failure detection on the device: the diagnostic kernel (mf_check_lh and the summing
kernels' _check twins) counts events whose per-event terms are not finite, finds the first,
and counts events outside the observables' ranges; fits report them (mfh_fit_out) and fits of
other segments in the same batch are unaffected.  Expected counts are exact: they follow
from where the test puts the bad events.
"""
import sys

import numpy as np
import pytest

import morefit_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fp32", [False, True])
def test_check_counts_bad_and_outside_events(ctx, fp32):
    sizes = [50000, 3000, 70001]
    say("check", "fp32", fp32, sizes)
    mdl = ctx.model(M.MASS, fp32=fp32)
    b = mdl.batch(sizes)
    data = [O.generate(M.MASS, M.MASS_TRUTH, n, seed=20 + i, nstreams=4) for i, n in enumerate(sizes)]
    data[0][0, [40000, 17, 41000]] = [np.nan, 5.0, 5.6]  # one bad event, two exactly at the range ends
    data[2][0, [70000, 3]] = [np.inf, 6.5]                # one bad and out of range, one only out of range
    for i, d in enumerate(data):
        b.upload(i, d)
    r = b.check(M.MASS_TRUTH)
    np.testing.assert_array_equal(r, [[1, 40000, 1], [0, -1, 0], [1, 70000, 2]])
    # zero width: the sums are not finite, and the check finds events that make them so
    t0 = np.array(M.MASS_TRUTH)
    t0[1] = 0.0
    assert not np.all(np.isfinite(b.eval(M.EVAL_LH, np.tile(t0, (3, 1)))[1]))
    r0 = b.check(t0, segs=[1])
    assert r0[0, 0] >= 1 and 0 <= r0[0, 1] < 3000 and r0[0, 2] == 0
    b.close()
    mdl.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_batched_fit_reports_its_bad_segment_only(ctx, fp32):
    sizes = [20000, 20000, 20000]
    say("batched fit failure", "fp32", fp32)
    mdl = ctx.model(M.MASS, fp32=fp32)
    b = mdl.batch(sizes)
    data = [O.generate(M.MASS, M.MASS_TRUTH, n, seed=40 + i, nstreams=4) for i, n in enumerate(sizes)]
    data[1][0, 123] = np.nan
    for i, d in enumerate(data):
        b.upload(i, d)
    starts = np.tile(M.MASS_START, (3, 1))
    vals, errs, outs, stats = b.fit(starts)
    assert outs[1].status != 0 and outs[1].nonfinite_evals >= 1
    assert outs[1].nonfinite_events == 1 and outs[1].first_nonfinite == 123 and outs[1].out_of_range == 1
    # the same study with clean data in segment 1: the other fits are bitwise the same (a
    # segment's sums do not depend on the other segments of a launch)
    clean = mdl.batch(sizes)
    data[1][0, 123] = M.MASS_TRUTH[0]
    for i, d in enumerate(data):
        clean.upload(i, d)
    vc, ec, oc, _ = clean.fit(starts)
    assert oc[1].status == 0 and oc[1].nonfinite_evals == 0 and oc[1].out_of_range == 0
    for i in (0, 2):
        assert outs[i].status == 0 and outs[i].nonfinite_evals == 0
        np.testing.assert_array_equal(vals[i], vc[i])
        np.testing.assert_array_equal(errs[i], ec[i])
    clean.close()
    b.close()
    mdl.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_plain_device_fit_reports_bad_event(ctx, fp32):
    d = O.generate(M.MASS, M.MASS_TRUTH, 30000, seed=7, nstreams=4)
    d[0, 29999] = np.nan
    say("plain fit failure", "fp32", fp32)
    v, e, out, _ = M.fit_options(M.MASS, d, M.MASS_START, ctx=ctx, fp32=fp32)
    assert out.status != 0
    assert out.nonfinite_events == 1 and out.first_nonfinite == 29999 and out.out_of_range == 1
