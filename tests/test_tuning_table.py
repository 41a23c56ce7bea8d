"""The committed batched-QR tuning table (measured on an MI355X by
``python -m hybrid3_amd.batched``) is well formed and turns into plans
without a GPU."""
from hybrid3_amd import batched as B


def test_table_entries_are_candidates():
    t = B.load_table()
    assert t["chosen"] in t["measured"]["plans"]
    for b in t["buckets"]:
        assert tuple(b["m_range"]) in B.BUCKETS
        assert B._parse(b["best"]) in B.CANDIDATES
    # the chosen plan is the fastest one measured
    plans = t["measured"]["plans"]
    assert plans[t["chosen"]] == min(plans.values())


def test_plan_from_table_covers_every_size():
    t = B.load_table()
    sizes = B.uneven_sizes(500, seed=3)
    plan = B.plan_from_table(sizes, t)
    assert len(plan) == len(sizes) and all(p in B.CANDIDATES for p in plan)
    # a per-bucket table maps each matrix by its row count
    per = dict(t, chosen="per-bucket")
    plan = B.plan_from_table(sizes, per)
    best = {tuple(b["m_range"]): B._parse(b["best"]) for b in t["buckets"]}
    for (m, _), p in zip(sizes, plan):
        assert p == best[B.BUCKETS[B.bucket_of(m)]]


def test_names_round_trip():
    for c in B.CANDIDATES:
        assert B._parse(B._name(c)) == c
        assert B._key(c) == c
    assert B._key((64, 16)) == (64, 16, 0)
