"""Renders one G-buffer case through the CPU emulation build and prints one JSON line: the
parity statistics against the FP64 oracle (as the GPU tests compute them) and a digest of the
outputs. Run in a process of its own (it imports a copy of the package bound to the emulated
library): python tests/simt/run_case.py <pkg_dir> <case> <persistent|plain> [cus]"""
import hashlib
import json
import os
import sys
import time

pkg_dir, case, kernel = sys.argv[1], sys.argv[2], sys.argv[3]
if len(sys.argv) > 4:
    os.environ["EMU_CUS"] = sys.argv[4]
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, pkg_dir)
sys.path.insert(1, ROOT)
import numpy as np  # noqa: E402
import numericdream_amd as nd  # noqa: E402
from numericdream_amd import abi  # noqa: E402
assert os.path.dirname(nd.__file__).startswith(pkg_dir), nd.__file__
import oracle  # noqa: E402
from tests import formula_table as ft  # noqa: E402
from tests.parity import compare, pointwise  # noqa: E402

if case == "c2":
    u, cm = nd.resize(nd.config_uniforms(2, 0), 48, 27), [0, 1, 2, 3]
elif case in ft.HYBRIDS:
    u, cm = ft.hybrid_test_scene(case)
else:
    u, cm = ft.test_scene(case)
r = nd.Renderer()
r.bind_formulas(cm)
W, H = int(u.resolution[0]), int(u.resolution[1])
g = nd.GBuffer(W, H)
print("[simt] %s %s %dx%d" % (case, kernel, W, H), file=sys.stderr, flush=True)
t0 = time.time()
kern = {"persistent": abi.KERNEL_PERSISTENT, "plain": abi.KERNEL_PLAIN}[kernel]
r.render_gbuffer(u, g, normal_mode=abi.NORMAL_ANALYTIC, shadows=True, shadow_max_steps=0, kernel=kern)
secs = time.time() - t0
got = g.numpy()
ref = oracle.render_gbuffer(u, normal_mode=0, shadows=True, code_map=cm)
st = compare(got, ref)
if st["hits_ref"] and st["hits_gpu"]:
    st.update(pointwise(got, u, 0, code_map=cm))
h = hashlib.sha256()
for k in ("hit", "depth", "normal", "trap", "shadow"):
    h.update(np.ascontiguousarray(got[k]).tobytes())
st = {k: float(v) for k, v in st.items()}
st.update(seconds=secs, digest=h.hexdigest())
print(json.dumps(st), flush=True)
