"""Writes qlookup_8bit.json: the 8-bit rows of the AV1 quantizer lookup
tables (Dc_Qlookup[0], Ac_Qlookup[0]) as the reference carries them in
src/spec_tables.h:608 and :692 (spec 7.12.2). Run in a container that holds
the reference sources; the JSON is the committed fixture."""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/spec_tables.h"


def table(text, name):
    m = re.search(name + r"\s*\[\s*3\s*\]\s*\[\s*256\s*\]\s*=\s*\{", text)
    body = text[m.end():]
    depth, end = 1, 0
    for i, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            end = i
            break
    vals = [int(v) for v in re.findall(r"-?\d+", re.sub(r"/\*.*?\*/", "", body[:end], flags=re.S))]
    assert len(vals) == 3 * 256, (name, len(vals))
    return vals[:256]


text = open(SRC).read()
out = {"source": "src/spec_tables.h (Dc_Qlookup[0], Ac_Qlookup[0])", "dc": table(text, "Dc_Qlookup"),
       "ac": table(text, "Ac_Qlookup")}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qlookup_8bit.json")
with open(path, "w") as f:
    json.dump(out, f)
print(path, out["dc"][:4], out["ac"][-2:])
