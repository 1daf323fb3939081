"""Depth of field in the oracle's path trace (cs_wf_generate's lens samples and
fs_wf_resolve's average, template.wgsl:5150-5182, 5373-5394), pinned by what it reduces to:

- dof_aperture (:4726-4730) is 0 inside half the focus range, the full aperture past it by
  the falloff max(4 aperture, 0.05), half of it midway (smoothstep);
- the lens disk sample (sample_unit_disk, :4709-4713) lies in the unit disk and covers it
  evenly over the seeds the renderer makes (E[r^2] = 1/2, centred);
- an aperture at DOF_MIN_APERTURE, the PART_DEPTH_OF_FIELD bit, or a focus range that holds
  every hit leave the pinhole render unchanged, bit for bit;
- pixels whose centre ray hits inside the focus range keep their pinhole colour exactly,
  while out-of-focus pixels are averages over the lens and so smoother.
"""
import numpy as np

import numericdream_amd as nd
import oracle

PART_DOF = 1 << 13


def H(op, args, n=4):
    return oracle.helper(op, args, n)


def test_dof_aperture_closed_forms():
    # the block holds aperture and focus range as float32
    ap, rng = float(np.float32(0.05)), float(np.float32(0.4))
    half, falloff = rng / 2, max(4 * ap, 0.05)
    assert H(60, [3.0, 3.1, ap, rng])[0] == 0.0                          # within range / 2
    assert H(60, [3.0, 3.0 + half + falloff, ap, rng])[0] == ap           # past it by the falloff
    assert abs(H(60, [3.0 + half + falloff / 2, 3.0, ap, rng])[0] - ap / 2) < 1e-12  # smoothstep midpoint
    tiny = float(np.float32(0.001))
    assert abs(H(60, [0.0, 0.025, tiny, 0.0])[0] - tiny / 2) < 1e-12     # falloff floor 0.05
    x = 0.3
    assert abs(H(60, [0.0, 0.05 * x, tiny, 0.0])[0] - tiny * x * x * (3 - 2 * x)) < 1e-12


def test_lens_disk_sample():
    pts = []
    for py in range(48):
        for px in range(48):
            for ms in (0.0, 5.0):
                bx, by = np.float32(px + 0.5) + np.float32(ms) * np.float32(91.73), np.float32(py + 0.5) + np.float32(ms) * np.float32(57.31)
                for j in range(4):
                    sx, sy = bx + np.float32(j) * np.float32(91.37), by + np.float32(j) * np.float32(51.91)
                    pts.append(H(61, [float(sx), float(sy)])[:3])
    pts = np.array(pts)
    assert np.all((pts[:, 0] >= 0) & (pts[:, 0] < 1))
    r2 = pts[:, 1] ** 2 + pts[:, 2] ** 2
    assert r2.max() <= 1.0 + 1e-12
    assert abs(r2.mean() - 0.5) < 0.02
    assert np.linalg.norm(pts[:, 1:3].mean(0)) < 0.03


def c3(w=48, h=27):
    u = nd.resize(nd.config_uniforms(3, 0), w, h)
    u.mc_enabled = 0.0  # direct lighting: deterministic, so exact comparisons hold
    u.mc_inner_samples = 8.0
    return u


def test_no_change_without_effective_aperture():
    u = c3()
    ref = oracle.render_mc(u, 2)["rgb"]
    for set_up in ("min_aperture", "part_bit", "wide_range"):
        v = c3()
        v.dof_enabled = 1.0
        v.aperture = 0.2
        v.focus_distance = 4.0
        v.focus_range = 0.5
        if set_up == "min_aperture":
            v.aperture = 0.0001
        elif set_up == "part_bit":
            v.debug_parts = float(int(v.debug_parts) | PART_DOF)
        else:
            v.focus_range = 1e6
        got = oracle.render_mc(v, 2)["rgb"]
        assert np.array_equal(got, ref), set_up


def test_in_focus_pixels_keep_their_colour_and_the_rest_blurs():
    u = c3(64, 36)
    pin = oracle.render_mc(u, 1)
    hit = pin["depth"] > 0
    t = np.median(pin["depth"][hit])
    v = c3(64, 36)
    v.dof_enabled = 1.0
    v.aperture = 0.15
    v.focus_distance = float(t)
    v.focus_range = 0.1
    dof = oracle.render_mc(v, 1)["rgb"]
    # the centre ray's hit against the focus: focus_t = focus_distance / (d . forward)
    d = np.array([oracle.primary_dir(u, x, y) for y in range(36) for x in range(64)]).reshape(36, 64, 3)
    fwd = np.array(u.camera_forward[:])
    ft = float(np.float32(t)) / np.maximum(d @ fwd, 1e-4)
    probe = np.where(hit, pin["depth"], float(u.max_dist))
    inside = np.abs(probe - ft) <= 0.05 - 1e-4  # within half the focus range: aperture 0
    assert inside.sum() > 20
    assert np.array_equal(dof[inside], pin["rgb"][inside])
    # out of focus: lens averages smooth the image (Laplacian energy of the luminance drops)
    lum = lambda img: img @ np.array([0.2126, 0.7152, 0.0722])
    lap = lambda y: (4 * y[1:-1, 1:-1] - y[:-2, 1:-1] - y[2:, 1:-1] - y[1:-1, :-2] - y[1:-1, 2:])
    far = ~inside[1:-1, 1:-1]
    e_pin = np.mean(lap(lum(pin["rgb"]))[far] ** 2)
    e_dof = np.mean(lap(lum(dof))[far] ** 2)
    assert e_dof < 0.7 * e_pin, (e_dof, e_pin)
