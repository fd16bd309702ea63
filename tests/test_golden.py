"""The oracle against the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_goldens.py). The fixtures pin the oracle's outputs on small cases: an edit to
the oracle that moves any NN choice, tetra vertex or image value fails here instead of silently
moving the bar the GPU tests compare against. The inputs are rebuilt from the library's seeded
generators and must match the fixtures' digests first."""
import os
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_goldens as G  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_matches_golden(name):
    import fomo_amd as F
    from oracle import oracle as O

    want = load(name)
    with tempfile.TemporaryDirectory() as tmp:
        got = G.make(name, F, O, tmp)
    assert str(got["input_digest"]) == str(want["input_digest"]), "the seeded generator changed its output"
    if "table_digest" in want:
        assert str(got["table_digest"]) == str(want["table_digest"]), "the table generator changed its output"
    assert set(got) == set(want)
    for k in want:
        if k.startswith(("idx", "vert", "views", "res", "rows", "nl")):
            assert np.array_equal(got[k], want[k]), k
        elif k.startswith(("image", "meta", "width")):
            np.testing.assert_allclose(got[k], want[k], rtol=1e-13, atol=0, err_msg=k)


def test_goldens_are_not_trivial():
    """Every fixture holds real content: non-zero images, NN choices that pick points, and (for
    the band cases) the band it claims."""
    for name, c in G.CASES.items():
        z = load(name)
        nv = len(c["views"])
        for v in range(nv):
            assert np.count_nonzero(z[f"image{v}"]) > 0.2 * z[f"image{v}"].size, (name, v)
            sel = z[f"vert{v}"] if c["kind"] == "tet" else z[f"idx{v}"]
            assert (sel >= 0).mean() > 0.2, (name, v)
            assert len(np.unique(sel[sel >= 0])) > 50, (name, v)
