"""The oracle's hash join and group-by, run as the 13 SSB queries over the
reference's scale-0.01 data, reproduce the answers the reference printed
(resources/result/ssb-sf0.01/ssb/original/*.r).  This pins orc_join_pairs and
orc_groupby (integer sums) against the reference's own output."""
import pytest

import ssb

CASE = {}


def _data():
    if not CASE:
        CASE["cols"], CASE["answers"] = ssb.load()
    return CASE["cols"], CASE["answers"]


@pytest.mark.parametrize("q", ssb.QUERIES)
def test_oracle_matches_reference_answer(q):
    cols, answers = _data()
    plan = ssb.PLANS[q]
    assert ssb.oracle_run(cols, plan) == ssb.expected_rows(plan, answers[q])


def test_answers_are_not_trivial():
    cols, answers = _data()
    assert sum(len(a["rows"]) for a in answers.values()) >= 40
    assert answers["1.1"]["rows"] == [[238810731]]
