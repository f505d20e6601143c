"""HIP library against the committed golden frames (tests/golden/frame_*.npz):
pixels, primary G-buffer and every pass's per-slot hit records must match
bit for bit.  The oracle is not run here; the vectors are data."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(rsh):
    e = rsh.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("name", sorted(make_golden.FRAMES))
def test_frame_matches_golden(engine, name):
    scene, cam, cfg = make_golden.frame_case(name)
    with np.load(os.path.join(HERE, "golden", "frame_%s.npz" % name), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    engine.load_scene(scene)
    st = engine.render(cam, cfg, flags=engine.KEEP_PASSES)
    assert st.pass_count == len(g["pass_kind"])
    for i in range(st.pass_count):
        assert st.pass_kind[i] == g["pass_kind"][i]
        assert st.pass_rays[i] == g["pass_rays"][i]
        if cfg.algorithm == 0:
            assert st.pass_chunks[i] == g["pass_chunks"][i]
        np.testing.assert_array_equal(engine.pass_hits(i), g["pass%d_hits" % i], err_msg="pass %d" % i)
    gb = engine.gbuffer()
    for k in ("pos", "nrm", "dif", "spec"):
        np.testing.assert_array_equal(gb[k].view(np.uint32), g["gbuffer_" + k].view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(engine.pixels(), g["pixels"])
