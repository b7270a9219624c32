"""The oracle's pins are sensitive to the mistakes a transcription of coupe's
split loop can plausibly make.

Each mutant of the pure-Python transcription (tests/pyref.py) plants one
mistake in a comparison or an update of par_rcb_split / rcb_recurse
(coupe/src/algorithms/recursive_bisection.rs:456-642).  The checks that pin
the C oracle -- the golden vectors and the agreement with the transcription
on seeded integer-weight cases -- must reject every mutant: had the oracle
made that mistake, one of them would have failed.  The unmutated
transcription passes the same checks (test_oracle.py).
"""
import pytest

from tests import cases, pyref


def _run(pts, w, iters, tol, kind, mutant):
    """The transcription's ids, or None when its split loop does not end."""
    try:
        return pyref.rcb(pts.tolist(), w.tolist(), iters, tol, kind, mutant=mutant)
    except RuntimeError:
        return None


def _golden_failures(mutant):
    vecs, _ = cases.golden()
    bad = 0
    for v, pts, w in vecs:
        kind = cases.kind_name(w)
        if v["constant"]:   # coupe_data_constant: one value for every point
            w = w[:1].repeat(pts.shape[0])
        bad += _run(pts, w, v["iter_count"], v["tolerance"], kind, mutant) != v["expected"]
    return bad


def _seeded_failures(oracle, mutant, seeds=range(240)):
    bad = 0
    for seed in seeds:
        pts, w, iters, tol = cases.random_case(seed, n_max=90)
        if not cases.integral(w):
            continue
        err, ids, _ = oracle.rcb(pts, w, iters, tol)
        assert err == 0
        bad += _run(pts, w, iters, tol, cases.kind_name(w), mutant) != ids.tolist()
    return bad


def test_unmutated_transcription_passes_golden_vectors():
    assert _golden_failures(None) == 0


@pytest.mark.parametrize("mutant", pyref.MUTANTS)
def test_mutant_is_caught(oracle_mod, mutant):
    golden = _golden_failures(mutant)
    seeded = _seeded_failures(oracle_mod, mutant)
    print(f"{mutant}: {golden} golden vectors and {seeded} seeded cases fail")
    assert golden + seeded >= 1, f"no check rejects the {mutant} mutant"
