"""Engine shootout TSV records (bf-new_amd/shootout.py).

The header lines equal the reference's (tests/golden/shootout_headers.txt,
read from src/engine_shootout.cpp by tests/golden/make_shootout_headers.py);
the row widths are the reference reader's (26 / 17 / 87 fields).  Round trips
are exact, the number format is 17 significant digits, the readers refuse
the rows the reference refuses, and quartile bins follow the rank rule.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "bf-new_amd"))

import shootout as so  # noqa: E402

FP = so.Fingerprint((2 ** 64 - 1, 17), (3, 4), 64)
WL = so.Workload(so.WORKLOAD_SYNTHETIC, 9, "c2 plain")


def manifest(n=8):
    feats = [so.QueryFeatures(i, 100 + i, 1000 - i, 1 + (i % 2), 1, 3 + i // 3) for i in range(n)]
    return so.Manifest(FP, WL, 5, 6, 1, 3, n, list(zip(feats, so.quartile_bins(feats))))


def catalog():
    return so.Catalog(FP, WL, [so.Tuning(2, 2, so.PER_ROUND, 1, 256, maximum_rounds=1 << 20),
                               so.Tuning(0, 0, so.CHUNKED, 8, 128, maximum_rounds=64),
                               so.Tuning(1, 1, so.PERSISTENT, 1, 512, maximum_rounds=7)])


def samples(m):
    out = []
    for cfg, engine, control in ((0, 0, so.CHUNKED), (1, 1, so.PERSISTENT)):
        for q in (3, 1):
            for rep in range(2):
                s = so.Sample(FP, WL, so.RUN_TIMING, len(out) * 7 + 1, rep, q, cfg, engine, control,
                              correctness_passed=True)
                s.status.update(converged=1, rounds_completed=12, valid_lane_mask=1, converged_lane_mask=1,
                                reached_target_mask=1, stop_reason=1)
                s.work["edges_examined"] = 2 ** 40 + q
                s.timing["wall"] = so.evidence(0.1 + rep)
                s.timing["timeline"] = so.evidence(1e-9 * (q + 1))
                s.timing["prep"] = so.evidence(state=so.EVIDENCE_NOT_APPLICABLE)
                s.profiler["l2_hit"] = so.evidence(87.5)
                out.append(s)
    return out


def test_headers_equal_the_reference():
    golden = open(os.path.join(ROOT, "tests", "golden", "shootout_headers.txt")).read().split("\n")
    assert [line.replace("\\t", "\t") for line in golden[:3]] == [
        so.MANIFEST_HEADER, so.CATALOG_HEADER, so.SAMPLES_HEADER]


def test_round_trips_are_exact():
    m, c = manifest(), catalog()
    text = so.serialize_manifest(m)
    assert so.deserialize_manifest(text) == m
    assert text.count("\n") == 1 + len(m.entries) and all(len(r.split("\t")) == 26 for r in text.splitlines())
    ctext = so.serialize_catalog(c)
    back = so.deserialize_catalog(ctext)
    assert sorted(back.tunings, key=lambda t: t.configuration_id) == sorted(c.tunings, key=lambda t: t.configuration_id)
    assert [int(r.split("\t")[9]) for r in ctext.splitlines()[1:]] == [0, 1, 2]  # ordered by id
    ss = samples(m)
    stext = so.serialize_samples(m, c.tunings, ss, so.RUN_TIMING)
    rows = stext.splitlines()[1:]
    assert all(len(r.split("\t")) == 87 for r in rows)
    back_s = so.deserialize_samples(m, c.tunings, stext, so.RUN_TIMING)
    key = lambda s: (s.execution_ordinal, s.query_id, s.configuration_id, s.repetition)  # noqa: E731
    assert so.samples_equal(back_s, sorted(ss, key=key))
    # 17 significant digits, as the reference writes doubles.
    assert "\t0.10000000000000001\t" in stext and "\t2.0000000000000001e-09\t" in stext
    assert "\t18446744073709551615\t" in text


def test_readers_refuse_what_the_reference_refuses():
    m, c = manifest(), catalog()
    text = so.serialize_manifest(m)
    lines = text.splitlines()
    bad = [
        text.replace("schema", "schemas", 1),                                    # header
        lines[0] + "\n",                                                         # no rows
        text.replace("\n", "\r\n"),                                              # CRLF
        "\n".join(lines[:1] + [lines[2], lines[1]] + lines[3:]) + "\n",          # ordinal order
        "\n".join(lines[:2] + [lines[2] + "\t0"] + lines[3:]) + "\n",           # width
        "\n".join(lines[:2] + [lines[2].replace("\t5\t6\t", "\t5\t7\t", 1)] + lines[3:]) + "\n",  # mixed metadata
        text.replace("\tc2 plain\t", "\t\t"),                                    # empty case name
        "\n".join(lines[:-1] + [lines[-1].rsplit("\t", 1)[0] + "\t4"]) + "\n",   # bin > 3
    ]
    for t in bad:
        with pytest.raises(so.ShootoutFormatError):
            so.deserialize_manifest(t)
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_manifest(so.Manifest(FP, so.Workload(so.WORKLOAD_LOGICNETS_JSCL, 1, "x"), 1, 1, 1, 1, 8,
                                          manifest().entries))  # logicnets needs 1000 queries
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_catalog(so.Catalog(FP, WL, [so.Tuning(0, 0, so.CHUNKED, 3, 256)]))  # chunk 3
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_catalog(so.Catalog(FP, WL, [so.Tuning(0, 0, so.PERSISTENT, 1, 64)]))  # block 64
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_catalog(so.Catalog(FP, WL, [so.Tuning(0, 0, so.PERSISTENT, 1, 128)] * 2))  # duplicate id
    ss = samples(m)
    ss[0].timing["prep"] = (so.EVIDENCE_UNAVAILABLE, 1.0)
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_samples(m, c.tunings, ss, so.RUN_TIMING)
    ss = samples(m)
    ss[0].configuration_id = 9
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_samples(m, c.tunings, ss, so.RUN_TIMING)
    ss = samples(m)
    with pytest.raises(so.ShootoutFormatError):
        so.serialize_samples(m, c.tunings, ss, so.RUN_TRACE)
    stext = so.serialize_samples(m, c.tunings, samples(m), so.RUN_TIMING)
    rows = stext.splitlines()
    with pytest.raises(so.ShootoutFormatError):  # rows out of canonical order
        so.deserialize_samples(m, c.tunings, "\n".join([rows[0], rows[2], rows[1]] + rows[3:]) + "\n", so.RUN_TIMING)


def test_quartile_bins_follow_the_rank_rule():
    # Values 1..8: thresholds at ranks 2, 4, 6 are 2, 4, 6.
    feats = [so.QueryFeatures(i, v, v, v, v, v) for i, v in enumerate(range(1, 9))]
    assert [b[0] for b in so.quartile_bins(feats)] == [0, 0, 1, 1, 2, 2, 3, 3]
    # One value: every threshold equals it, so its bin is 0.
    assert so.quartile_bins([so.QueryFeatures(0, 5, 5, 5, 5, 5)]) == [(0, 0, 0, 0, 0)]
    # Ties share a bin: all three thresholds are 7, so 9 lies above all three.
    feats = [so.QueryFeatures(i, v, 1, 1, 1, 1) for i, v in enumerate([7, 7, 7, 9])]
    assert [b[0] for b in so.quartile_bins(feats)] == [0, 0, 0, 3]


def test_committed_campaign_replays():
    """The per-query shootout campaign the autotuner ran on one MI355X (the
    64x64x256 graph's 64 queries, three engines, three timed repetitions)
    reads back from its committed files, and its decision -- the
    configuration of smallest median wall time -- is recomputed from them."""
    import statistics
    d = os.path.join(ROOT, "bf-new_amd", "tuning", "shootout_c2_plain_64x64x256_q64")
    m = so.deserialize_manifest(open(os.path.join(d, "manifest.v2.tsv")).read())
    c = so.deserialize_catalog(open(os.path.join(d, "catalog.v2.tsv")).read())
    ss = so.deserialize_samples(m, c.tunings, open(os.path.join(d, "samples.v2.tsv")).read(), so.RUN_TIMING)
    assert len(m.entries) == 64 and len(c.tunings) == 3 and len(ss) == 64 * 3 * 3
    assert all(s.status["converged"] == 1 and s.timing["wall"][0] == so.EVIDENCE_MEASURED for s in ss)
    med = {t.configuration_id: statistics.median(s.timing["wall"][1] for s in ss
                                                 if s.configuration_id == t.configuration_id) for t in c.tunings}
    best = min(med, key=med.get)
    assert c.tunings[best].engine == 1  # dense chaotic push
    # Writing the records again reproduces the files byte for byte.
    assert so.serialize_manifest(m) == open(os.path.join(d, "manifest.v2.tsv")).read()
    assert so.serialize_catalog(c) == open(os.path.join(d, "catalog.v2.tsv")).read()
    assert so.serialize_samples(m, c.tunings, ss, so.RUN_TIMING) == open(os.path.join(d, "samples.v2.tsv")).read()
