"""Writes abi_fields.json: the ordered field names of the reference's function tables and the
structs they take, read from the reference headers (nanovdb_editor/putil/*.h under
/root/reference). Run in the build container, where the reference tree is present; the tests read
only the JSON."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from abi_fields import REFERENCE_TABLES, struct_fields  # noqa: E402

PUTIL = "/root/reference/nanovdb_editor/putil"


def main():
    out = {}
    for header, names in REFERENCE_TABLES.items():
        path = os.path.join(PUTIL, header)
        found = struct_fields(open(path).read(), names)
        missing = set(names) - set(found)
        if missing:
            raise SystemExit(f"{header}: {sorted(missing)} not found")
        for name in names:
            out[name] = dict(found[name], source=f"nanovdb_editor/putil/{header}")
    with open(os.path.join(HERE, "abi_fields.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
