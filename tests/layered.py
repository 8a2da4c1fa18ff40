"""Layered digraphs whose FW / BW layer sizes are set by hand, for known-answer checks of
the reference's direction switch (multistep/scc_fwbw.cpp:164-195, :347-378)."""
import numpy as np


def layered(msc, sizes, isolated=0):
    """Pivot -> layer 1 -> ... -> layer k -> pivot.  Every vertex of layer i has an out-edge
    into layer i+1 and every vertex of layer i+1 an in-edge from layer i, so a top-down level
    from the pivot discovers exactly one layer.  Deeper layers take smaller ids: a one-thread
    bottom-up sweep in ascending id order then claims one layer per forward level (a
    vertex's parents have larger ids and were visited before the sweep).  `isolated`
    vertices without edges come last; they count in n but are trimmed.  Returns
    (graph, pivot)."""
    layers, start = [None] * len(sizes), 0
    for i in range(len(sizes) - 1, -1, -1):
        layers[i] = np.arange(start, start + sizes[i], dtype=np.int64)
        start += sizes[i]
    pivot = start
    n = start + 1 + isolated
    src, dst = [np.full(sizes[0], pivot)], [layers[0]]
    for a, b in zip(layers, layers[1:]):
        src += [a, a[np.arange(b.size) % a.size]]
        dst += [b[np.arange(a.size) % b.size], b]
    src.append(layers[-1])
    dst.append(np.full(sizes[-1], pivot))
    g = msc.create_csr(n, np.concatenate(src), np.concatenate(dst), dedupe=True)
    return g, pivot
