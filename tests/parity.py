"""Comparison metrics between a HIP G-buffer and the oracle's, with the tolerances of the
workloads: hit mask agreement >= 99.9 % away from silhouettes, depth within 1e-4 relative,
normals within 0.5 degrees, all stated as the fraction of pixels that must meet them."""
import numpy as np


def silhouette_mask(hit):
    """Pixels whose 3x3 neighbourhood mixes hits and misses (in the reference mask)."""
    h = hit.astype(bool)
    pad = np.pad(h, 1, mode="edge")
    H, W = h.shape
    anyh = np.zeros_like(h)
    allh = np.ones_like(h)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            win = pad[dy:dy + H, dx:dx + W]
            anyh |= win
            allh &= win
    return anyh & ~allh


def compare(gpu, ref):
    """Returns a dict of agreement statistics (fractions in [0, 1], errors in their units)."""
    gh = gpu["hit"].astype(bool)
    rh = ref["hit"].astype(bool)
    sil = silhouette_mask(rh) | silhouette_mask(gh)
    inner = ~sil
    both = gh & rh
    st = {"pixels": int(gh.size), "hits_ref": int(rh.sum()), "hits_gpu": int(gh.sum()),
          "hit_agree_all": float((gh == rh).mean()),
          "hit_agree_inner": float((gh == rh)[inner].mean()) if inner.any() else 1.0}
    if both.any():
        gd = gpu["depth"][both].astype(np.float64)
        rd = ref["depth"][both].astype(np.float64)
        rel = np.abs(gd - rd) / np.maximum(np.abs(rd), 1e-12)
        st["depth_rel_median"] = float(np.median(rel))
        st["depth_rel_max"] = float(rel.max())
        st["depth_rel_le_1e-4"] = float((rel <= 1e-4).mean())
        if "normal" in gpu and "normal" in ref:
            gn = gpu["normal"][both].astype(np.float64)
            rn = ref["normal"][both].astype(np.float64)
            cosang = np.clip(np.sum(gn * rn, axis=1) / np.maximum(
                np.linalg.norm(gn, axis=1) * np.linalg.norm(rn, axis=1), 1e-30), -1, 1)
            ang = np.degrees(np.arccos(cosang))
            st["normal_deg_median"] = float(np.median(ang))
            st["normal_deg_le_0.5"] = float((ang <= 0.5).mean())
        if "trap" in gpu and "trap" in ref:
            st["trap_equal"] = float(np.isclose(gpu["trap"][both], ref["trap"][both], rtol=1e-3, atol=1e-5).mean())
        if "shadow" in gpu and "shadow" in ref:
            ds = np.abs(gpu["shadow"][both].astype(np.float64) - ref["shadow"][both])
            st["shadow_abs_le_0.02"] = float((ds <= 0.02).mean())
            st["shadow_abs_median"] = float(np.median(ds))
    return st


def _angle_deg(a, b):
    c = np.clip(np.sum(a * b, axis=1) / np.maximum(np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1), 1e-30),
                -1, 1)
    return np.degrees(np.arccos(c))


def pointwise(gpu, u, normal_mode, code_map=None, floor=False):
    """Normal and trap of the oracle (FP64) evaluated at the GPU's own hit points.

    At the workloads' epsilon (1e-3 of a pixel) the DE gradient of a fractal changes
    direction over distances comparable to FP32 rounding of the hit point, so end-to-end
    normals mix march rounding with normal arithmetic. Evaluating the oracle at the exact
    FP32 hit point (the position output) isolates the normal and trap computation."""
    import oracle
    h = gpu["hit"].astype(bool)
    pos = gpu["position"][h].astype(np.float64)
    t = gpu["depth"][h].astype(np.float64)
    eps = float(u.epsilon_coefficient) * t + float(u.epsilon_floor)
    e = eps if normal_mode == 1 else 0.5 * eps
    ref_n = oracle.normal(u, pos, normal_mode, e, code_map=code_map)
    gn = gpu["normal"][h].astype(np.float64)
    cosang = np.clip(np.sum(gn * ref_n, axis=1) / np.maximum(
        np.linalg.norm(gn, axis=1) * np.linalg.norm(ref_n, axis=1), 1e-30), -1, 1)
    ang = np.degrees(np.arccos(cosang))
    _, ref_trap = oracle.scene_de(u, pos, trap=True, code_map=code_map)
    st = {"pw_hits": int(h.sum())}
    if h.any():
        st["pw_normal_deg_median"] = float(np.median(ang))
        st["pw_normal_deg_le_0.5"] = float((ang <= 0.5).mean())
        # discrete traps (iteration index) must match exactly; continuous ones to 1e-3
        st["pw_trap_equal"] = float(np.isclose(gpu["trap"][h], ref_trap, rtol=1e-3, atol=1e-5).mean())
    if floor and h.any():
        # What FP32 itself reaches at these points. The trap: the FP32 CPU restatement of
        # the same iteration against FP64. The stencil normals: their FP32 restatement. The
        # analytic normal has no FP32 restatement; its floor is how far the FP64 normal moves
        # when each coordinate of the hit point moves by one FP32 ulp (seeded signs).
        _, t32 = oracle.scene_de(u, pos, prec=1, trap=True, code_map=code_map)
        st["floor_trap_equal"] = float(np.isclose(t32, ref_trap, rtol=1e-3, atol=1e-5).mean())
        if normal_mode in (1, 2):
            n32 = oracle.normal(u, pos, normal_mode, e, code_map=code_map, prec=1)
        else:
            rng = np.random.default_rng(20261203)
            ulp = np.spacing(np.abs(pos).astype(np.float32)).astype(np.float64)
            n32 = oracle.normal(u, pos + ulp * rng.choice([-1.0, 1.0], size=pos.shape), 0, code_map=code_map)
        st["floor_normal_deg_le_0.5"] = float((_angle_deg(n32, ref_n) <= 0.5).mean())
    return st


def meets(value, bar, floor, margin=0.03):
    """value reaches the fixed bar, or comes within margin of the FP32 floor where the
    floor itself lies below the bar (an ill-conditioned quantity)."""
    return value >= bar or value >= floor - margin
