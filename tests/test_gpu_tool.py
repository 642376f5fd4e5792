"""The reference's own command-line tool, DistanceOcclusionTool.cpp, built
unchanged two ways by oracle/Makefile (target `tool`):

  oracle/_ref/distocc_ref     with the reference's DistanceOcclusionLib.cpp
  oracle/_ref/distocc_dolhip  against include/DistanceOcclusionLib.h, linked
                              to libdolhip (every DOL:: call on the GPU)

Each command runs with both in fresh directories; the files they write (PNG
images) and what they print (the -c checks' error figures, -v logs) must be
byte-identical.  The binaries are built only where the reference sources
are; elsewhere the test skips.
"""
import os
import subprocess

import pytest

from oracle_bind import ROOT

pytestmark = pytest.mark.gpu

REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "distocc_ref")
HIP_TOOL = os.path.join(ROOT, "oracle", "_ref", "distocc_dolhip")

# an octahedron, for the mesh voxeliser of -F (MeshSupport.cpp, inside the tool)
OCTAHEDRON = """v 1 0 0
v -1 0 0
v 0 1 0
v 0 -1 0
v 0 0 1
v 0 0 -1
f 1 3 5
f 3 2 5
f 2 4 5
f 4 1 5
f 3 1 6
f 2 3 6
f 4 2 6
f 1 4 6
"""

COMMANDS = [
    # distance: Danielsson, FastSweep, JumpFlood, BruteForce, Chamfer, Felzenszwalb
    *[["-d", "-m", str(m), "-w", "96", "70", "-p", "40", "7", "-c"] for m in range(6)],
    ["-d", "-m", "2", "-i", "1", "-w", "64", "-b", "4"],
    # single-pass signed (border init), each method, each image type (the
    # reference's border check swaps x and y, Tool:474-475, so -c only passes
    # on square images)
    *[["-s", "-m", str(m), "-w", "64", "64", "-b", "3", "-p", "25", "3", "-c"] for m in range(4)],
    *[["-s", "-m", str(m), "-w", "80", "61", "-b", "3", "-p", "25", "3"] for m in range(4)],
    ["-s", "-m", "2", "-i", "1", "-w", "80", "61", "-b", "3"],
    ["-s", "-m", "2", "-i", "2", "-w", "80", "61", "-b", "3"],
    ["-s", "-m", "2", "-l", "20", "-w", "50", "-p", "9", "1"],
    # two-field signed
    *[["-S", "-m", str(m), "-w", "72", "50", "-b", "4", "-c"] for m in range(3)],
    # volumes (the 3D check prints to the -v stream and crashes without it,
    # Tool:991; -v prints the mask with h slices, Tool:1961, so it is used
    # on volumes with d = h only)
    *[["-d", "-m", str(m), "-W", "20", "18", "18", "-p", "15", "5", "-c", "-v"] for m in range(4)],
    *[["-d", "-m", str(m), "-W", "20", "18", "12", "-p", "15", "5"] for m in range(4)],
    ["-s", "-m", "2", "-W", "24", "17", "9", "-b", "3", "-c"],
    ["-s", "-m", "1", "-W", "24", "17", "9", "-b", "3"],
    ["-S", "-m", "2", "-W", "16", "16", "16", "-b", "2"],
    # occlusion, swept shadow
    *[["-x", "-m", str(m), "-w", "90", "64", "-b", "3", "-p", "30", "2"] for m in range(3)],
    # (3D -x -m 1 is left out: the tool sweeps into new[] memory whose first
    # cell of each plane the reference never writes, Tool:1567-1570, so its
    # images hold whatever the heap held)
    *[["-x", "-m", str(m), "-W", "18", "14", "10", "-b", "2"] for m in (0, 2)],
    ["-h", "0.8", "0.3", "1", "-w", "70", "-b", "3", "-p", "12", "4"],
    ["-h", "1", "-0.6", "-1", "-w", "50", "40", "-p", "30", "9"],
    # the tool's default sizes (256 x 256, 32^3)
    ["-s", "-m", "2", "-c"],
    ["-d", "-m", "2", "-W", "-c", "-v"],
    # a mesh source, with -v logging
    ["-d", "-m", "2", "-W", "20", "20", "20", "-F", "mesh.obj", "-o", "mesh"],
    ["-s", "-m", "2", "-W", "24", "24", "24", "-F", "mesh.obj", "-o", "mesh", "-v"],
]


def _run(tool, args, where):
    os.makedirs(where)
    with open(os.path.join(where, "mesh.obj"), "w") as f:
        f.write(OCTAHEDRON)
    r = subprocess.run([tool] + args, cwd=where, capture_output=True, timeout=90)
    files = {}
    for name in sorted(os.listdir(where)):
        with open(os.path.join(where, name), "rb") as f:
            files[name] = f.read()
    return r.returncode, r.stdout, files


@pytest.mark.skipif(not (os.path.exists(REF_TOOL) and os.path.exists(HIP_TOOL)),
                    reason="oracle/_ref tool binaries not built (reference sources absent)")
@pytest.mark.parametrize("args", COMMANDS, ids=lambda a: "_".join(a).replace("-", ""))
def test_tool_outputs_identical(args, tmp_path):
    rc0, out0, files0 = _run(REF_TOOL, args, str(tmp_path / "ref"))
    rc1, out1, files1 = _run(HIP_TOOL, args, str(tmp_path / "hip"))
    assert rc1 == rc0
    assert sorted(files1) == sorted(files0)
    assert len(files0) > 1, "the tool wrote nothing"
    for name in files0:
        assert files1[name] == files0[name], name
    assert out1 == out0
