"""Writes nd_formula_table.inc: for every formula the library binds (nd_builtin_formula), the
file stem a scene names it by and its @param names and defaults in declaration order, which
is the order of FractalInstance.params (formula.zig:110-157 parses the same lines).

Saved scenes address parameters by name (scene_file.zig:586-601), so the loader needs the
table. Its source is the committed @param table of the formula files,
tests/golden/formula_params.json; the ids come from nd_hip.h's nd_builtin_formula.
Run it after either changes; the .inc is committed."""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    ids = {}
    for line in open(os.path.join(ROOT, "include", "nd_hip.h")):
        m = re.match(r"\s*ND_FORMULA_\w+\s*=\s*(\d+),?\s*/\*\s*(\S+)\.wgsl\s*\*/", line)
        if m:
            ids[int(m.group(1))] = os.path.basename(m.group(2))
    table = json.load(open(os.path.join(ROOT, "tests", "golden", "formula_params.json")))
    out = ["// Generated by gen_formula_table.py from the @param table of the formula files: do not edit.",
           "// id, file stem, parameter count, names, defaults (declaration order = params slot order)."]
    for i in sorted(ids):
        rows = table[ids[i]]
        assert len(rows) <= 16, ids[i]
        names = ", ".join('"%s"' % r["name"] for r in rows) or '""'
        vals = ", ".join(repr(float(r["default"])) + "f" for r in rows) or "0.0f"
        out.append('{%d, "%s", %d, {%s}, {%s}},' % (i, ids[i], len(rows), names, vals))
    with open(os.path.join(HERE, "nd_formula_table.inc"), "w") as fh:
        fh.write("\n".join(out) + "\n")
    print("wrote %d formulas" % len(ids))
    return 0


if __name__ == "__main__":
    sys.exit(main())
