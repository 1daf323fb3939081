"""The scene of the radiance-adjoint tests: the differentiable workload's Mandelbulb + Mandelbox
union (one-pixel epsilon, cameras on radius 3, the default material), lit by a shadowing point
light, a shadowing global light and a ray light (a beam from above), with Monte Carlo off and
the reflection continuation cut, as the direct-lighting image is defined."""
import numericdream_amd as nd

PART_REFLECTIONS = 1 << 6


def radiance_scene(view=0, size=(96, 96)):
    u = nd.config_uniforms(5, view)
    nd.resize(u, *size)
    u.mc_enabled = 0.0
    u.debug_parts = float(PART_REFLECTIONS)
    lights = [
        # kind, colour, brightness, position or direction, softness, ray direction, spread
        (0, (1.0, 0.9, 0.8), 6.0, (2.0, 2.5, -1.5), 12.0, (0.0, -1.0, 0.0), 0.0),
        (1, (0.6, 0.7, 1.0), 0.8, (-0.4, 0.8, 0.45), 16.0, (0.0, -1.0, 0.0), 0.0),
        (2, (0.9, 1.0, 0.7), 3.0, (0.3, 2.6, 0.2), 8.0, (-0.1, -1.0, 0.0), 0.35),
    ]
    u.light_count = float(len(lights))
    for i, (kind, col, b, pos, soft, rd, spread) in enumerate(lights):
        L = u.lights[i]
        L.light_type = float(kind)
        L.color[:] = col
        L.brightness = b
        L.position_or_direction[:] = pos
        L.shadow_softness = soft
        L.cast_shadows = 1.0
        L.hard_shadows = 0.0
        L.ray_direction[:] = rd
        L.spread = spread
        L.waist = 0.05
    return u


SMOOTH_CODES = [4, 5]  # code 0 -> sphere_primitive, code 1 -> cube_primitive


def smooth_scene(size=(96, 96)):
    """The same lights over exact distance fields: a sphere of radius 0.8 on a flat slab (the
    cube primitive scaled to 4 x 0.1 x 4, 1.2 below), bound through SMOOTH_CODES. Its colour is
    smooth in the camera position at the scale of FP32 differences, unlike a fractal's, whose
    escape-time DE changes its iteration count across the surface."""
    u = radiance_scene(size=size)
    sph, slab = u.instances[0], u.instances[1]
    for inst, code, p0, off, sc in ((sph, 0, 0.8, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
                                    (slab, 1, 1.0, (0.0, -1.2, 0.0), (4.0, 0.1, 4.0))):
        inst.code = float(code)
        for k in range(16):
            inst.params[k] = 0.0
        inst.params[0] = p0
        inst.offset[:] = off
        inst.scale[:] = sc
        inst.combine_mode = 0.0
        inst.hidden = 0.0
        inst.mixin_count = 0.0
    return u
