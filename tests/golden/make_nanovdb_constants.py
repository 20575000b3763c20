"""Writes nanovdb_constants.json: the NanoVDB header facts the reference states in its own sources,
read from /root/reference (run in the build container; the tests read only the JSON):
  - grid type codes, from the type-name switch of editor/FileHeaderInfo.cpp (getGridTypeName);
  - the grid magic ("NanoVDB1 (Grid)") from editor/FileHeaderInfo.cpp (getMagicTypeName);
  - the blind-metadata stride, from the `i * 288` address arithmetic of editor/FileHeaderInfo.cpp;
  - the data types the node2 builder gives its two blind-data entries, from
    raster/shaders/grid_build_blindmetadata.slang (set_data_type calls, in order);
  - the version decode of the editor's header display (getVersionString), for the record."""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def main():
    fhi = open(os.path.join(REF, "editor", "FileHeaderInfo.cpp")).read()
    types = {}
    body = fhi[fhi.index("getGridTypeName"):fhi.index("getGridTypeName") + 4000]
    body = body[:body.index("default:")]
    for code, name in re.findall(r"case\s+(\d+):\s*return\s+\"(\w+)\";", body):
        types[name] = int(code)
    magic = re.search(r"if \(magic == (0x[0-9a-fA-F]+)ULL\)\s*return \"NanoVDB1 \(Grid\)\"", fhi)
    strides = sorted(set(int(x) for x in re.findall(r"blindMetadataOffset \+ i \* (\d+)", fhi)))
    ver = re.search(r"major = \(version >> (\d+)\) & (0x[0-9A-Fa-f]+);\s*uint32_t minor = \(version >> (\d+)\) & "
                    r"(0x[0-9A-Fa-f]+);\s*uint32_t patch = version & (0x[0-9A-Fa-f]+);", fhi)
    slang = open(os.path.join(REF, "raster", "shaders", "grid_build_blindmetadata.slang")).read()
    blind_types = re.findall(r"set_data_type\(buf, (\w+), PNANOVDB_GRID_TYPE_(\w+)\)", slang)
    out = dict(grid_types=types, magic_grid=int(magic.group(1), 16), blind_metadata_stride=strides,
               node2_blind_data_types=[t for _, t in blind_types],
               editor_version_decode=dict(major_shift=int(ver.group(1)), major_mask=int(ver.group(2), 16),
                                          minor_shift=int(ver.group(3)), minor_mask=int(ver.group(4), 16),
                                          patch_mask=int(ver.group(5), 16)),
               sources=["editor/FileHeaderInfo.cpp", "raster/shaders/grid_build_blindmetadata.slang"])
    with open(os.path.join(HERE, "nanovdb_constants.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
