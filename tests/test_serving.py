"""Serving's host logic on the CPU: the learned admission estimate."""
import pytest

from sparkalog_amd import batch, serving


def test_closure_estimator_learns_per_size():
    est = serving.ClosureEstimator(margin=2.0)
    req = serving.make_requests(1, rate_per_s=1.0, seed=3, min_log2_edges=12, max_log2_edges=12)[0]
    prior = est(req)
    assert prior == batch.estimate_closure(req.nodes, req.src.size)
    est.observe(req, int(prior / 30))
    assert est(req) == pytest.approx(prior * 2.0 / 30, rel=1e-3)
    est.observe(req, int(prior / 300))  # the largest factor of a bucket stays
    assert est(req) == pytest.approx(prior * 2.0 / 30, rel=1e-3)
    # a request 16x larger is in another bucket: the plain estimate again
    big = serving.make_requests(1, rate_per_s=1.0, seed=4, min_log2_edges=16, max_log2_edges=16)[0]
    assert est(big) == batch.estimate_closure(big.nodes, big.src.size)
    # never above the plain estimate
    est.observe(req, int(prior))
    assert est(req) == prior
