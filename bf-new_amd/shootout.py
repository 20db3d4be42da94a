"""Engine shootout records: manifest, configuration catalog and timing samples
as the reference's TSV files (reference: include/bfnew/engine_shootout.hpp:15-230,
src/engine_shootout.cpp:32-300, 693-783, 2129-2680).

A manifest names one workload (graph and query fingerprints, seeds,
repetition counts) and its selected queries with their features and
quartile bins; a catalog lists the run configurations by id; a samples file
holds one row per (execution, query, configuration, repetition) with the
run's status, work counters and evidence values (state + number).  The
autotuner writes all three, so a tuning decision can be audited and
replayed from files.

Formatting follows the reference writer: classic-locale integers, evidence
numbers with 17 significant digits, LF line endings, one header line; the
readers check the header, the field count of every row (manifest 26,
catalog 17, samples 87), canonical ordinals and that rows share their
metadata.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

SCHEMA_VERSION = 2
BLOCK_SIZES = (128, 256, 512)
CHUNK_SIZES = (2, 4, 8, 16, 32)
MINIMUM_LOGICNETS_QUERIES = 1000
WORKLOAD_LOGICNETS_JSCL, WORKLOAD_SYNTHETIC = 0, 1
RUN_WARMUP, RUN_CORRECTNESS, RUN_TIMING, RUN_COUNTERS, RUN_TRACE, RUN_PMC = range(6)
EVIDENCE_UNAVAILABLE, EVIDENCE_NOT_APPLICABLE, EVIDENCE_MEASURED = range(3)
PERSISTENT, CHUNKED, PER_ROUND = 0, 1, 2  # ControlMode values
OCCUPANCY_DERIVED = 0  # GridPolicy value

_PREFIX = ("schema", "workload_kind", "case_id", "case_name", "graph0", "graph1", "query0", "query1",
           "corpus_queries")
_SEEDS = ("selection_seed", "order_seed", "warmup_repetitions", "timing_repetitions")
MANIFEST_COLUMNS = _PREFIX + _SEEDS + (
    "requested", "selection_ordinal", "query_id", "selected_vertices", "selected_edges", "fanout",
    "source_count", "expected_rounds", "vertex_bin", "edge_bin", "fanout_bin", "source_bin", "round_bin")
CATALOG_COLUMNS = _PREFIX + ("configuration_id", "engine", "control", "rounds_per_chunk", "block_size",
                             "grid_policy", "blocks_per_wgp", "maximum_rounds")
STATUS_FIELDS = ("final_distance_slot", "converged", "rounds_completed", "reached_target_mask",
                 "bounding_box_miss_mask", "valid_lane_mask", "active_lane_mask", "converged_lane_mask",
                 "stop_reason", "error_bits")
WORK_FIELDS = ("edges_examined", "successful_decreases", "active_vertices", "active_lane_rounds",
               "maximum_queue_size", "host_checks", "host_synchronizations", "controller_copies",
               "kernel_dispatches", "expansion_count", "atomic_attempts", "successful_atomic_updates",
               "queue_claims", "duplicate_suppressions", "mask_operations", "overflow_events",
               "high_contention_destinations", "changed_flag_updates", "full_edge_rounds",
               "empty_frontier_rounds", "small_frontier_rounds")
TIMING_FIELDS = ("prep", "timeline", "transfer", "wall")
PROFILER_FIELDS = ("gpu_active", "l2_hit", "l2_read", "l2_write", "occupancy", "memory_busy", "waves",
                   "vector", "scalar", "memory_inst")


def _evidence_columns(names):
    return tuple(c for n in names for c in (n + "_state", n + "_value"))


SAMPLE_COLUMNS = (_PREFIX + _SEEDS + ("run_kind", "execution_ordinal", "repetition", "query_id",
                                      "configuration_id", "instrumentation", "correctness_passed",
                                      "distances_downloaded", "engine", "control")
                  + STATUS_FIELDS + WORK_FIELDS + _evidence_columns(TIMING_FIELDS)
                  + ("profiler_pass_id", "counter_set_id", "profiler_compatible")
                  + _evidence_columns(PROFILER_FIELDS)
                  + ("cooperative_grid_blocks", "cooperative_active_blocks_per_wgp"))
MANIFEST_HEADER = "\t".join(MANIFEST_COLUMNS)
CATALOG_HEADER = "\t".join(CATALOG_COLUMNS)
SAMPLES_HEADER = "\t".join(SAMPLE_COLUMNS)
assert (len(MANIFEST_COLUMNS), len(CATALOG_COLUMNS), len(SAMPLE_COLUMNS)) == (26, 17, 87)


class ShootoutFormatError(ValueError):
    pass


def _require(ok, message):
    if not ok:
        raise ShootoutFormatError(message)


@dataclass(frozen=True)
class Fingerprint:
    graph_words: tuple[int, int]
    query_words: tuple[int, int]
    corpus_query_count: int
    schema_version: int = SCHEMA_VERSION


@dataclass(frozen=True)
class Workload:
    kind: int
    case_id: int
    case_name: str


@dataclass(frozen=True)
class QueryFeatures:
    query_id: int
    selected_vertices: int
    selected_edges: int
    fanout: int
    source_count: int
    expected_rounds: int


@dataclass
class Manifest:
    fingerprint: Fingerprint
    workload: Workload
    selection_seed: int
    order_seed: int
    warmup_repetitions: int
    timing_repetitions: int
    requested_query_count: int
    entries: list = field(default_factory=list)  # [(QueryFeatures, (5 bins))]


@dataclass(frozen=True)
class Tuning:
    configuration_id: int
    engine: int
    control_mode: int
    rounds_per_chunk: int
    block_size: int
    grid_policy: int = OCCUPANCY_DERIVED
    blocks_per_wgp: int = 0
    maximum_rounds: int = 1


@dataclass
class Catalog:
    fingerprint: Fingerprint
    workload: Workload
    tunings: list = field(default_factory=list)


def evidence(value=None, state=None):
    """(state, number): measured when a number is given, else unavailable."""
    if state is None:
        state = EVIDENCE_MEASURED if value is not None else EVIDENCE_UNAVAILABLE
    return (state, 0.0 if value is None else float(value))


@dataclass
class Sample:
    fingerprint: Fingerprint
    workload: Workload
    run_kind: int
    execution_ordinal: int
    repetition: int
    query_id: int
    configuration_id: int
    engine: int
    control: int
    instrumentation: int = 0
    correctness_passed: bool = False
    distances_downloaded: bool = False
    status: dict = field(default_factory=lambda: {k: 0 for k in STATUS_FIELDS})
    work: dict = field(default_factory=lambda: {k: 0 for k in WORK_FIELDS})
    timing: dict = field(default_factory=lambda: {k: evidence() for k in TIMING_FIELDS})
    profiler_pass_id: int = 0
    counter_set_id: int = 0
    profiler_compatible: bool = False
    profiler: dict = field(default_factory=lambda: {k: evidence() for k in PROFILER_FIELDS})
    cooperative_grid_blocks: int = 0
    cooperative_active_blocks_per_wgp: int = 0


# ---------------------------------------------------------------------------
# validation (engine_shootout.cpp:81-300)
def _check_fingerprint(f):
    _require(f.schema_version == SCHEMA_VERSION, "shootout fingerprint schema version is unsupported")
    _require(f.corpus_query_count != 0, "shootout fingerprint requires a nonempty corpus")


def _check_workload(w):
    _require(w.kind in (WORKLOAD_LOGICNETS_JSCL, WORKLOAD_SYNTHETIC), "invalid shootout workload kind")
    _require(w.case_id != 0, "shootout workload requires a case ID")
    _require(w.case_name != "", "shootout workload requires a case name")
    _require(not any(c in w.case_name for c in "\t\r\n"), "shootout workload case name is not TSV-safe")


def _check_features(f):
    _require(f.selected_vertices != 0, "shootout query features require selected vertices")
    _require(f.fanout != 0, "shootout query features require fanout")
    _require(f.source_count != 0, "shootout query features require at least one source")
    _require(f.expected_rounds != 0, "shootout query features require positive pilot rounds")


def _check_manifest(m):
    _check_fingerprint(m.fingerprint)
    _check_workload(m.workload)
    _require(m.warmup_repetitions != 0 and m.timing_repetitions != 0,
             "shootout manifest requires positive warmup/timing repetitions")
    _require(m.requested_query_count != 0 and len(m.entries) == m.requested_query_count,
             "shootout manifest query count is inconsistent")
    _require(m.fingerprint.corpus_query_count >= len(m.entries),
             "shootout manifest exceeds its fingerprinted corpus")
    if m.workload.kind == WORKLOAD_LOGICNETS_JSCL:
        _require(m.requested_query_count >= MINIMUM_LOGICNETS_QUERIES,
                 "logicnets shootout manifest contains fewer than 1000 queries")
    ids = set()
    for features, bins in m.entries:
        _check_features(features)
        _require(features.query_id not in ids, "shootout manifest contains a duplicate query ID")
        ids.add(features.query_id)
        _require(len(bins) == 5 and all(0 <= b <= 3 for b in bins), "shootout manifest contains an invalid quantile bin")


def valid_tuning(t):
    if t.block_size not in BLOCK_SIZES or t.maximum_rounds == 0 or t.engine not in (0, 1, 2):
        return False
    if t.control_mode == PERSISTENT:
        return t.rounds_per_chunk == 1
    if t.control_mode == PER_ROUND:
        return t.rounds_per_chunk == 1 and t.grid_policy == OCCUPANCY_DERIVED and t.blocks_per_wgp == 0
    if t.control_mode == CHUNKED:
        return t.rounds_per_chunk in CHUNK_SIZES and t.grid_policy == OCCUPANCY_DERIVED and t.blocks_per_wgp == 0
    return False


def _tuning_index(tunings):
    _require(len(tunings) > 0, "shootout requires at least one tuning")
    index = {}
    for t in tunings:
        _require(valid_tuning(t), "shootout tuning is invalid")
        _require(t.configuration_id not in index, "shootout tuning IDs must be unique")
        index[t.configuration_id] = t
    return index


def _check_evidence(ev):
    state, value = ev
    _require(state in (EVIDENCE_UNAVAILABLE, EVIDENCE_NOT_APPLICABLE, EVIDENCE_MEASURED),
             "invalid shootout evidence state")
    if state == EVIDENCE_MEASURED:
        _require(math.isfinite(value) and value >= 0.0, "measured shootout evidence must be finite and nonnegative")
    else:
        _require(value == 0.0, "unavailable or nonapplicable evidence may not encode a number")


# ---------------------------------------------------------------------------
# quartile bins (engine_shootout.cpp:189-233)
FEATURE_DIMENSIONS = ("selected_vertices", "selected_edges", "fanout", "source_count", "expected_rounds")


def quartile_bins(features):
    """Per query, for each feature dimension, how many of the three quartile
    thresholds (the values at ranks ceil(n k / 4), k = 1..3, of the sorted
    values) lie strictly below its value."""
    _require(len(features) > 0, "shootout quantiles require candidates")
    thresholds = []
    for dim in FEATURE_DIMENSIONS:
        values = sorted(getattr(f, dim) for f in features)
        n = len(values)
        row = []
        for k in (1, 2, 3):
            rank = (n * k + 3) // 4
            row.append(values[min(n - 1, max(rank, 1) - 1)])
        thresholds.append(row)
    out = []
    for f in features:
        out.append(tuple(sum(1 for t in thresholds[d] if t < getattr(f, dim))
                         for d, dim in enumerate(FEATURE_DIMENSIONS)))
    return out


# ---------------------------------------------------------------------------
# TSV text
def _prefix(fp, wl):
    return [fp.schema_version, wl.kind, wl.case_id, wl.case_name, fp.graph_words[0], fp.graph_words[1],
            fp.query_words[0], fp.query_words[1], fp.corpus_query_count]


def _number(v):
    return "%.17g" % v


def _row(values):
    return "\t".join(str(v) if not isinstance(v, float) else _number(v) for v in values)


def _lines(text, header):
    _require(text != "", "shootout TSV is empty")
    first = text.find("\n")
    _require(first >= 0, "shootout TSV has no data rows")
    _require(text[:first] == header, "shootout TSV header mismatch")
    body = text[first + 1:]
    lines = body.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    for line in lines:
        _require(line != "", "shootout TSV contains an empty data row")
        _require(not line.endswith("\r"), "shootout TSV must use LF line endings")
    _require(len(lines) > 0, "shootout TSV has no data rows")
    return [line.split("\t") for line in lines]


def _unsigned(text, message):
    _require(text.isdigit() and text.isascii(), message)
    return int(text)


def _parse_prefix(f):
    fp = Fingerprint((_unsigned(f[4], "invalid graph fingerprint"), _unsigned(f[5], "invalid graph fingerprint")),
                     (_unsigned(f[6], "invalid query fingerprint"), _unsigned(f[7], "invalid query fingerprint")),
                     _unsigned(f[8], "invalid corpus count"), _unsigned(f[0], "invalid shootout schema"))
    _check_fingerprint(fp)
    wl = Workload(_unsigned(f[1], "invalid workload kind"), _unsigned(f[2], "invalid workload case ID"), f[3])
    _check_workload(wl)
    return fp, wl


def serialize_manifest(m):
    _check_manifest(m)
    rows = [MANIFEST_HEADER]
    for i, (f, bins) in enumerate(m.entries):
        rows.append(_row(_prefix(m.fingerprint, m.workload) + [
            m.selection_seed, m.order_seed, m.warmup_repetitions, m.timing_repetitions,
            m.requested_query_count, i, f.query_id, f.selected_vertices, f.selected_edges, f.fanout,
            f.source_count, f.expected_rounds, *bins]))
    return "\n".join(rows) + "\n"


def deserialize_manifest(text):
    m = None
    for i, f in enumerate(_lines(text, MANIFEST_HEADER)):
        _require(len(f) == 26, "manifest TSV row has wrong field count")
        fp, wl = _parse_prefix(f)
        meta = [_unsigned(f[k], "invalid manifest metadata") for k in range(9, 14)]
        _require(_unsigned(f[14], "invalid manifest ordinal") == i, "manifest TSV ordinal is not canonical")
        if m is None:
            m = Manifest(fp, wl, *meta)
        else:
            _require(fp == m.fingerprint and wl == m.workload and meta == [
                m.selection_seed, m.order_seed, m.warmup_repetitions, m.timing_repetitions,
                m.requested_query_count], "manifest TSV mixes metadata or workloads")
        features = QueryFeatures(*[_unsigned(f[k], "invalid manifest query field") for k in range(15, 21)])
        bins = tuple(_unsigned(f[k], "invalid manifest bin") for k in range(21, 26))
        m.entries.append((features, bins))
    _check_manifest(m)
    return m


def serialize_catalog(c):
    _check_fingerprint(c.fingerprint)
    _check_workload(c.workload)
    _tuning_index(c.tunings)
    rows = [CATALOG_HEADER]
    for t in sorted(c.tunings, key=lambda t: t.configuration_id):
        rows.append(_row(_prefix(c.fingerprint, c.workload) + [
            t.configuration_id, t.engine, t.control_mode, t.rounds_per_chunk, t.block_size, t.grid_policy,
            t.blocks_per_wgp, t.maximum_rounds]))
    return "\n".join(rows) + "\n"


def deserialize_catalog(text):
    c = None
    for f in _lines(text, CATALOG_HEADER):
        _require(len(f) == 17, "catalog TSV row has wrong field count")
        fp, wl = _parse_prefix(f)
        if c is None:
            c = Catalog(fp, wl)
        else:
            _require(fp == c.fingerprint and wl == c.workload, "catalog TSV mixes workloads")
        c.tunings.append(Tuning(*[_unsigned(f[k], "invalid catalog field") for k in range(9, 17)]))
    _tuning_index(c.tunings)
    return c


def _sample_key(s):
    return (s.execution_ordinal, s.query_id, s.configuration_id, s.repetition, s.profiler_pass_id,
            s.counter_set_id)


def _check_samples(manifest, tunings, samples, run_kind):
    _check_manifest(manifest)
    index = _tuning_index(tunings)
    ids = {f.query_id for f, _ in manifest.entries}
    for s in samples:
        _require(s.fingerprint == manifest.fingerprint and s.workload == manifest.workload,
                 "shootout sample belongs to another workload")
        _require(s.run_kind == run_kind, "shootout sample has another run kind")
        _require(s.query_id in ids, "shootout sample names a query outside the manifest")
        t = index.get(s.configuration_id)
        _require(t is not None, "shootout sample names an unknown configuration")
        _require(s.engine == t.engine and s.control == t.control_mode,
                 "shootout sample engine or control disagrees with its configuration")
        for ev in list(s.timing.values()) + list(s.profiler.values()):
            _check_evidence(ev)


def serialize_samples(manifest, tunings, samples, run_kind):
    _check_samples(manifest, tunings, samples, run_kind)
    rows = [SAMPLES_HEADER]
    for s in sorted(samples, key=_sample_key):
        values = _prefix(s.fingerprint, s.workload) + [
            manifest.selection_seed, manifest.order_seed, manifest.warmup_repetitions,
            manifest.timing_repetitions, s.run_kind, s.execution_ordinal, s.repetition, s.query_id,
            s.configuration_id, s.instrumentation, int(bool(s.correctness_passed)),
            int(bool(s.distances_downloaded)), s.engine, s.control]
        values += [s.status[k] for k in STATUS_FIELDS] + [s.work[k] for k in WORK_FIELDS]
        for k in TIMING_FIELDS:
            values += [s.timing[k][0], float(s.timing[k][1])]
        values += [s.profiler_pass_id, s.counter_set_id, int(bool(s.profiler_compatible))]
        for k in PROFILER_FIELDS:
            values += [s.profiler[k][0], float(s.profiler[k][1])]
        values += [s.cooperative_grid_blocks, s.cooperative_active_blocks_per_wgp]
        rows.append(_row(values))
    return "\n".join(rows) + "\n"


def _number_field(text, message):
    try:
        v = float(text)
    except ValueError as e:
        raise ShootoutFormatError(message) from e
    return v


def deserialize_samples(manifest, tunings, text, run_kind):
    samples = []
    for f in _lines(text, SAMPLES_HEADER):
        _require(len(f) == 87, "samples TSV row has wrong field count")
        fp, wl = _parse_prefix(f)
        seeds = [_unsigned(f[k], "invalid sample metadata") for k in range(9, 13)]
        _require(seeds == [manifest.selection_seed, manifest.order_seed, manifest.warmup_repetitions,
                           manifest.timing_repetitions], "samples TSV disagrees with its manifest")
        u = [None] * 87
        for k in range(13, 87):
            u[k] = f[k]
        at = 13
        head = [_unsigned(u[at + i], "invalid sample field") for i in range(10)]
        at += 10
        s = Sample(fp, wl, run_kind=head[0], execution_ordinal=head[1], repetition=head[2], query_id=head[3],
                   configuration_id=head[4], instrumentation=head[5], correctness_passed=bool(head[6]),
                   distances_downloaded=bool(head[7]), engine=head[8], control=head[9])
        _require(head[6] <= 1 and head[7] <= 1, "invalid sample boolean")
        s.status = {k: _unsigned(u[at + i], "invalid device status") for i, k in enumerate(STATUS_FIELDS)}
        at += len(STATUS_FIELDS)
        s.work = {k: _unsigned(u[at + i], "invalid work record") for i, k in enumerate(WORK_FIELDS)}
        at += len(WORK_FIELDS)
        for k in TIMING_FIELDS:
            s.timing[k] = (_unsigned(u[at], "invalid evidence state"), _number_field(u[at + 1], "invalid evidence number"))
            at += 2
        s.profiler_pass_id = _unsigned(u[at], "invalid profiler pass ID")
        s.counter_set_id = _unsigned(u[at + 1], "invalid profiler counter-set ID")
        s.profiler_compatible = bool(_unsigned(u[at + 2], "invalid profiler compatibility"))
        at += 3
        for k in PROFILER_FIELDS:
            s.profiler[k] = (_unsigned(u[at], "invalid evidence state"), _number_field(u[at + 1], "invalid evidence number"))
            at += 2
        s.cooperative_grid_blocks = _unsigned(u[at], "invalid cooperative grid")
        s.cooperative_active_blocks_per_wgp = _unsigned(u[at + 1], "invalid cooperative residency")
        samples.append(s)
    _check_samples(manifest, tunings, samples, run_kind)
    _require([_sample_key(s) for s in samples] == sorted(_sample_key(s) for s in samples),
             "samples TSV rows are not in canonical order")
    return samples


def samples_equal(a, b):
    return [dataclasses.asdict(x) for x in a] == [dataclasses.asdict(x) for x in b]
