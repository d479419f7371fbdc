"""The oracle's SSB restatement against the reference's own logged results.

Golden values are the `The result is:` / `Sum of results` / `Number of results (groups)` lines of
usecases/ssb/results/query_sf10_export.csv (:7-8 and following), query_sf50_export.csv (:6-7 and
following) and query_sf100_export.csv (:6-7 and following). They pin the generator's LCG call
order (SSB_def.hpp:179-399), the Q1 plan (01inmem.hpp:51-141, 01.cuh:26-54), the Q2 star
join + group-by (02inmem.hpp:58-206), and the Q3.x / Q4.x plans (03.hpp:611-677 with pipeline_q3,
03.cuh:118-170; 04.hpp:316-381 with pipeline_q4, 04.cuh:99-153): their predicates on the customer,
supplier, part and date attributes, their group keys (group counts) and SUM(lo_revenue) /
SUM(lo_revenue - lo_supplycost).
"""
import numpy as np
import pytest

GOLDEN = {
    10: dict(rows=60004738, q1=[4212228344835, 900952032730, 246824540411],
             q2_sum=[1737351738436, 347570089279, 43762644335], q2_groups=[280, 56, 7],
             q34_sum=[7511687354847, 301286847125, 12099576386, 140730476, 3440778955099, 981566989074,
                      19815949200],
             q34_groups=[150, 600, 24, 4, 35, 100, 798]),
    50: dict(rows=299998345, q1=[21039215126196, 4531516252170, 1234508031440],
             q2_sum=[8725371303394, 1739417333755, 209802496106], q2_groups=[280, 56, 7],
             q34_sum=[37224656601928, 1508847617616, 55560399329, 803033663, 17082562623215, 4868116900963,
                      96823394981],
             q34_groups=[150, 600, 24, 4, 35, 100, 800]),
    100: dict(rows=600016632, q1=[42056852386840, 9080947035194, 2468788349493],
              q2_sum=[17565422530734, 3483217621763, 447332943956], q2_groups=[280, 56, 7],
              q34_sum=[75063343008233, 2952340552244, 117332675301, 1502523597, 34141742450403, 9729298653664,
                       190602879891],
              q34_groups=[150, 600, 24, 4, 35, 100, 800]),
}


@pytest.mark.parametrize("sf", [10, 50, pytest.param(100, marks=pytest.mark.slow)])
def test_ssb_streaming_goldens(orc, sf):
    g = orc.ssb_golden(sf)
    assert g == GOLDEN[sf]


@pytest.fixture(scope="module")
def ssb10(orc):
    return orc.ssb_generate(10)


Q1_VARIANTS = [  # (date predicates, discount range, quantity range); 01.hpp:451-490
    (lambda d: [(d["year"], 1993, 1993)], (1, 3), (0, 24)),
    (lambda d: [(d["yearmonthnum"], 199401, 199401)], (4, 6), (26, 35)),
    (lambda d: [(d["weeknuminyear"], 6, 6), (d["year"], 1994, 1994)], (5, 7), (26, 35)),
]


@pytest.mark.parametrize("q", [0, 1, 2])
def test_q1_plan_reproduces_sf10_golden(orc, ssb10, q):
    lo, d, p, s, _ = ssb10
    preds, drng, qrng = Q1_VARIANTS[q]
    total, _ = orc.q1_plan(d["key"], preds(d), lo["orderdate"], lo["discount"], lo["quantity"],
                           lo["extendedprice"], drng, qrng, threads=4)
    assert total == GOLDEN[10]["q1"][q]


def test_star_groupby_reproduces_sf10_q21(orc, ssb10):
    """Q2.1: p_category = MFGR#12, s_region = AMERICA, group by (d_year, p_brand1)."""
    lo, d, p, s, _ = ssb10
    n_parts, n_supp = p["category"].size, s["region"].size
    dims = [
        (np.arange(n_parts, dtype=np.int32), [(p["category"], 12, 12)], p["brand1"].astype(np.int32)),
        (np.arange(n_supp, dtype=np.int32), [(s["region"], 1, 1)], None),
        (d["key"], [], d["year"].astype(np.int32)),
    ]
    res = orc.star_groupby(dims, [lo["partkey"].view(np.int32), lo["suppkey"].view(np.int32),
                                  lo["orderdate"]], lo["revenue"].view(np.int32), threads=4)
    keys, sums, counts = res
    assert keys.shape == (280, 2)
    assert int(sums.sum()) == GOLDEN[10]["q2_sum"][0]
    # canonical order: sorted by (brand1, year) as the group tuple is in dimension order
    assert np.all(np.diff(keys[:, 0] * 10000 + keys[:, 1]) > 0)
