"""Shared test helpers: fixture materialisation and oracle wrappers."""
import os
import struct
import sys

import numpy as np

import oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_fixtures as fixtures  # noqa: E402  (fixture data transcribed from the reference)

SYNTHETIC, SYN_EDGE = 1, 1


def fixture_arrays(case):
    v = np.zeros(len(case["vertices"]), oracle.VERTEX_DTYPE)
    for i, (x, y, loc, cls) in enumerate(case["vertices"]):
        v[i] = (x, y, loc, cls)
    e = np.zeros(len(case["edges"]), oracle.EDGE_DTYPE)
    for i, (s, d, w, rec) in enumerate(case["edges"]):
        e[i] = (s, d, w, SYNTHETIC, SYN_EDGE, 0, rec)
    return v, e


def core_case():
    return dict(vertices=fixtures.CORE_VERTICES, edges=fixtures.CORE_EDGES, config=[0, 0, 64, 64])


def spatial_case():
    return dict(vertices=fixtures.SPATIAL_VERTICES, edges=fixtures.SPATIAL_EDGES,
                config=[0, 0, 10, 10])


def plain_vertices(count):
    v = np.zeros(count, oracle.VERTEX_DTYPE)
    for i in range(count):
        v[i] = (i, 0, 1, 1)
    return v


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def f32_from_bits(values):
    return np.asarray(values, np.uint32).view(np.float32)


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def tile_mask(tile_count, tiles):
    m = np.zeros(tile_count, np.uint8)
    m[np.asarray(tiles, np.int64)] = 1
    return m


def oracle_query(view, sources, tiles):
    """Oracle Dijkstra distances + relaxation count for one query over a
    host-graph view (the product's CSR)."""
    sel = tile_mask(view["tile_count"], tiles)
    return oracle.dijkstra(view["vertex_count"], view["row_offsets"], view["destinations"],
                           view["weights"], view["owner_tiles"], sel, sources)


def certify_batch(layout, ws, sources, tiles, threads=16):
    """Certifies every query of a completed workspace with the oracle's region
    certificate: labels come off the device one query at a time (the
    workspace is single-stream), certificates run on `threads` host threads.
    Returns [(query, message)] of the failures."""
    from concurrent.futures import ThreadPoolExecutor
    failures = []
    with ThreadPoolExecutor(max_workers=threads) as pool:
        futures = []
        for q in range(len(tiles)):
            labels = ws.region_distances(q)
            futures.append((q, pool.submit(oracle.certify_region, layout, tiles[q], labels,
                                           sources[q])))
        for q, f in futures:
            msg, _ = f.result()
            if msg:
                failures.append((q, msg))
    return failures


def stratified(sizes, per_stratum=2):
    """Indices of the smallest, median and largest entries of `sizes`."""
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i], i))
    n = len(order)
    mid = n // 2
    picked = order[:per_stratum] + order[mid:mid + per_stratum] + order[n - per_stratum:]
    return sorted(set(picked))


def csr_view(layout):
    """The CSR part of a product layout in the form the oracle's Dijkstra takes."""
    return dict(vertex_count=layout["vertex_count"], tile_count=layout["tile_count"],
                row_offsets=layout["csr_row_offsets"].astype(np.uint64),
                destinations=layout["csr_destinations"], weights=layout["csr_weights"],
                owner_tiles=layout["owner_tiles"])


def oracle_region(view, layout, source, tiles):
    """Oracle Dijkstra labels of one query, in the region (tile by tile) order."""
    d, _ = oracle_query(view, [source], tiles)
    return d[oracle.region_vertices(layout["tile_vertex_offsets"], tiles)]
