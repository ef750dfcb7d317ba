"""tetragono's configuration files, one configuration per chain.

``write_configurations`` / ``read_configurations``
(tetragono/tetragono/utility.py:390-418) store, as raw little-endian int64,
a header ``[count, ndim, *shape]`` and then ``count`` arrays of that shape,
one per MPI rank: tetragono runs one chain per rank and writes what
``Configuration.export_configuration`` returns
(sampling_lattice/lattice.py:208-229), an ``[L1, L2, orbits]`` array of
physical edge indices (-1 no such orbit, -2 unset).  Here the chains of a
batch take the place of the ranks; for the spinless tV lattices the edge
index of the physical edge ``[(0, 1), (1, 1)]`` is the occupation.
"""
import numpy as np


def write_configurations(configs, path):
    """``configs``: (chains, L1, L2) occupations or (chains, L1, L2, orbits)
    edge indices."""
    c = np.asarray(configs, dtype=np.int64)
    if c.ndim == 3:
        c = c[..., None]
    if c.ndim != 4:
        raise ValueError("configurations must be (chains, L1, L2) or (chains, L1, L2, orbits)")
    head = np.array([c.shape[0], 3, *c.shape[1:]], dtype=np.int64)
    with open(path, "wb") as f:
        f.write(head.astype("<i8").tobytes())
        f.write(np.ascontiguousarray(c).astype("<i8").tobytes())


def read_configurations(path):
    """All configurations of a file: (count, L1, L2, orbits) int64."""
    raw = np.fromfile(path, dtype="<i8")
    if raw.size < 2:
        raise ValueError("not a configuration file")
    count, ndim = int(raw[0]), int(raw[1])
    if count < 1 or ndim < 1 or raw.size < 2 + ndim:
        raise ValueError("not a configuration file")
    shape = tuple(int(x) for x in raw[2:2 + ndim])
    body = raw[2 + ndim:]
    need = count * int(np.prod(shape))
    # the reference writer opens without truncation (MPI.File.Open with
    # WRONLY | CREATE), so a file rewritten with fewer configurations keeps
    # trailing bytes; like the reader (utility.py:402-418) read only what
    # the header announces
    if body.size < need:
        raise ValueError("configuration file is shorter than its header announces")
    return body[:need].reshape((count,) + shape).astype(np.int64)


def occupations(configs):
    """(chains, L1, L2) occupations of single-orbit edge indices (the tV
    lattices); raises on unset or missing orbits."""
    c = np.asarray(configs)
    if c.ndim == 4:
        if c.shape[-1] != 1:
            raise ValueError("one orbit per site expected")
        c = c[..., 0]
    if np.any(c < 0):
        raise ValueError("configuration has unset sites")
    return c.astype(np.int8)
