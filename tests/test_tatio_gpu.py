"""Drop-in Tensor dump / load / pickle (PyTAT/PyTAT.hpp:757-771) on device
tensors, against bytes TAT wrote (tests/golden/tatio_*.bin)."""
import os
import pickle

import numpy as np
import pytest

from test_tatio import GOLDEN, HERE, expected

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind_name", sorted(GOLDEN))
def test_load_dump_pickle(kind_name):
    from tnsp_amd import TAT, tatio

    kind, names, edges, storage = expected(kind_name)
    data = open(os.path.join(HERE, "golden", GOLDEN[kind_name]["file"]), "rb").read()
    t = getattr(TAT, kind_name).D.Tensor([], []).load(data)
    assert t.names == names
    assert list(t.edges) == edges
    assert np.array_equal(np.asarray(t.storage), storage)
    mine = t.dump()
    pad = set(tatio.padding_positions(kind, data))
    assert len(mine) == len(data)
    assert all(mine[i] == data[i] for i in range(len(data)) if i not in pad)
    u = pickle.loads(pickle.dumps(t))
    assert u.names == names and np.array_equal(np.asarray(u.storage), storage)
    # float32 tensors write float32 storage
    s = getattr(TAT, kind_name).S.Tensor(names, edges)
    s.storage = storage.astype(np.float32)
    back = getattr(TAT, kind_name).S.Tensor([], []).load(s.dump())
    assert np.array_equal(np.asarray(back.storage), storage.astype(np.float32))
