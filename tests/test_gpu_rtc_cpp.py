"""The C++ mirror of barney's rtc interface (include/bnx_rtc.hpp) on the GPU.

tests/cpp/rtc_trace (built by build()) drives the mirror the way barney's
native layer drives a backend -- buffers, GeomType::createGeom,
setVertices/setIndices, triangle and user-geometry groups, an instance
group (spheres, triangles and cylinders), buildAccel, a trace, then
setTransforms + refitAccel and a second trace, then a re-commit of moved
vertices into the same rtc::Geoms and a third trace over new groups -- and
checks that a mismatched createInstanceGroup throws.  Its hits must equal,
bit for bit, those of the same scene built through the Python binding of the
C-ABI and the oracle's, before and after the move and after the re-commit."""
import os
import subprocess

import numpy as np
import pytest

from barney_amd import abi, scenes
from tests import helpers

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "rtc_trace")


def _same(g, o):
    bad = helpers.hits_equal(g, o)
    assert len(bad) == 0, (len(bad), g[bad[:3]], o[bad[:3]])


def test_rtc_mirror_traces_like_the_c_abi(dev, orc, tmp_path):
    assert os.access(EXE, os.X_OK), f"{EXE} not built (run __graft_entry__.build())"
    verts, idx, col, cs, rs = helpers.soup_data()
    move = np.array([0.5, -0.25, 0.0], np.float32)
    st = scenes.sticks(300, extent=1.8)
    rays = np.ascontiguousarray(scenes.random_rays(4, 50000, radius=5.0, spread=2.0), np.float32)
    inp, out = tmp_path / "scene.bin", tmp_path / "hits.bin"
    with open(inp, "wb") as f:
        for a in (np.int32(len(verts)), verts, np.int32(len(idx)), idx.astype(np.int32), col,
                  np.int32(len(cs)), cs, rs, move, np.int32(len(st.vertices)), st.vertices,
                  np.int32(len(st.indices)), st.indices.astype(np.int32), st.radii, np.int32(len(rays)), rays):
            f.write(np.ascontiguousarray(a).tobytes())
    print(f"rtc_trace: {len(idx)} triangles, {len(cs)} spheres, {len(st.indices)} cylinders, {len(rays)} rays",
          flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "rtc_trace ok" in r.stdout, (r.returncode, r.stdout, r.stderr)
    hits = np.fromfile(out, dtype=abi.HIT_DTYPE)
    assert len(hits) == 3 * len(rays)
    n = len(rays)
    before, after, recommit = hits[:n], hits[n:2 * n], hits[2 * n:]
    spec = helpers.soup_spec(sticks=st)
    _same(before, spec.gpu(dev).trace(rays))
    _same(before, spec.oracle(orc).trace(rays)[0])
    moved = helpers.soup_spec(move=tuple(float(x) for x in move), sticks=st)
    _same(after, moved.gpu(dev).trace(rays))
    _same(after, moved.oracle(orc).trace(rays)[0])
    assert not np.array_equal(before["prim_id"], after["prim_id"])
    assert (before["obj_id"] == 3).sum() > 1000
    # re-committed vertices (moved in the vertex buffer, not by an instance)
    spec3 = helpers.soup_spec(sticks=st, vertex_move=tuple(float(x) for x in move))
    _same(recommit, spec3.gpu(dev).trace(rays))
    _same(recommit, spec3.oracle(orc).trace(rays)[0])
    assert not np.array_equal(before["prim_id"], recommit["prim_id"])
