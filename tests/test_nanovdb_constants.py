"""NanoVDB header facts the reference states in its own sources (tests/golden/nanovdb_constants.json,
from editor/FileHeaderInfo.cpp and raster/shaders/grid_build_blindmetadata.slang) against this
library's layout constants and the grids the oracle writes (the GPU grids equal those byte for byte
in the GPU tests): grid type codes, the grid magic, the 288-byte blind-metadata stride and the
data types of the two blind-data entries. The field layout of one blind-metadata entry restated in
the readers (data offset i64, value count u64, value size u32, semantic u32, data class u32, data
type u32, name char[256]) must fill exactly the pinned stride."""
import json
import os
import re

import numpy as np

import oracle
from helpers import preprocessed, small_uniform
from node2_reader import GRID_TYPE_NODE2, MAGIC, GRID_TYPE_ONINDEX, Node2Grid, OnIndexGrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C = json.load(open(os.path.join(ROOT, "tests", "golden", "nanovdb_constants.json")))
STRIDE = C["blind_metadata_stride"]


def _layout_defines():
    text = open(os.path.join(ROOT, "nanovdb-editor_amd", "csrc", "vdb_layout.h")).read()
    return {k: int(v, 0) for k, v in re.findall(r"#define (VDB_\w+) (0x[0-9a-fA-F]+|\d+)u?l?l?", text)}


def test_layout_constants_match_reference():
    d = _layout_defines()
    assert STRIDE == [288] and d["VDB_BLINDMETADATA_SIZE"] == 288
    assert d["VDB_MAGIC_GRID"] == C["magic_grid"] == MAGIC
    for name in ("FLOAT", "VEC3F", "RGBA8", "ONINDEX", "NODE2"):
        assert d["VDB_GRID_TYPE_" + name] == C["grid_types"][name], name
    assert GRID_TYPE_NODE2 == C["grid_types"]["NODE2"] and GRID_TYPE_ONINDEX == C["grid_types"]["ONINDEX"]
    # one GridBlindMetaData entry as the readers restate it fills the stride exactly
    assert 8 + 8 + 4 + 4 + 4 + 4 + 256 == STRIDE[0]


def test_written_grids_carry_the_reference_codes():
    c = small_uniform(300, 0.01, 99)
    m, op, q, s, col = preprocessed(c)
    grid, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    assert int(grid[:8].view("<u8")[0]) == C["magic_grid"]
    assert int(grid[636:640].view("<u4")[0]) == C["grid_types"]["NODE2"]
    g = Node2Grid(grid)
    assert g.blind_count == len(C["node2_blind_data_types"]) == 2
    # entries at the pinned stride, with the builder's data types in order (bboxes, values)
    for i, tname in enumerate(C["node2_blind_data_types"]):
        meta = g.blind_offset + STRIDE[0] * i
        assert int(grid[meta + 28:meta + 32].view("<u4")[0]) == C["grid_types"][tname]
        assert int(grid[meta + 20:meta + 28].view("<u4").sum()) == 0     # semantic 0, data class 0
    _, vcount, vsize, _ = g.blind(1)
    assert vsize == 4 and grid.size == g.blind(1)[0] + 4 * vcount
    oni = oracle.node2_to_onindex(grid)
    assert int(oni[636:640].view("<u4")[0]) == C["grid_types"]["ONINDEX"]
    o = OnIndexGrid(oni)
    off, count, size = o.blind()
    assert count == vcount and size == 4
    assert np.array_equal(oni[off:off + 4 * count], grid[g.blind(1)[0]:g.blind(1)[0] + 4 * vcount])


def test_version_word_encoding_recorded():
    """The editor's header display decodes the version as 16/8/8-bit fields, while the builder writes
    it through pnanovdb_make_version (GridBuild.cpp:128-130), whose definition lives in the fetched
    PNanoVDB.h; this library writes 32.7.0 as (32 << 21) | (7 << 10), so the version word stays
    unpinned. The test records the display formula the fixture holds."""
    dec = C["editor_version_decode"]
    assert (dec["major_shift"], dec["minor_shift"]) == (16, 8)
    c = small_uniform(50, 0.01, 98)
    m, op, q, s, col = preprocessed(c)
    grid, _ = oracle.raster_node2(m, q, s, col, op, c.voxel_size)
    assert int(grid[16:20].view("<u4")[0]) == (32 << 21) | (7 << 10)
