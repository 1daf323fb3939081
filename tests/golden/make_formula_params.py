"""Writes tests/golden/formula_params.json: the @param table (name, min, max, default, int)
of every formula file, in declaration order = FractalInstance.params slot order
(formula.zig:110-157 parses the same lines). Reads the reference tree, so it runs only
where the reference is present; the JSON it writes is the committed fixture."""
import json
import os
import re
import sys

REF = os.environ.get("ND_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "formula_params.json")
LINE = re.compile(r"^//\s*@param\s+(.*?)\s+min=(\S+)\s+max=(\S+)\s+default=(\S+)(\s+int)?\s*$")


def table(path):
    rows = []
    for line in open(path):
        m = LINE.match(line.strip())
        if m:
            rows.append({"name": m.group(1), "min": float(m.group(2)), "max": float(m.group(3)),
                         "default": float(m.group(4)), "int": bool(m.group(5))})
    return rows


def main():
    files = {"mandelbrot": os.path.join(REF, "src/gpu/shaders/mandelbrot.wgsl")}
    fdir = os.path.join(REF, "formulas")
    for f in sorted(os.listdir(fdir)):
        if f.endswith(".wgsl"):
            files[f[:-5]] = os.path.join(fdir, f)
    out = {k: table(v) for k, v in files.items()}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", OUT, len(out), "formulas")
    return 0


if __name__ == "__main__":
    sys.exit(main())
