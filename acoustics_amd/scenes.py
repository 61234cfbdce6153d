"""Seeded synthetic scenes for the five workloads of the FDTD engine.

Every scene follows the reference's static solver policy: a cubic (or, for
batches and serving, box-shaped) MAC grid centred at the origin
(``BoundingBox(h, n, center)``, src/geometry/BoundingBox.h:90-100), sound
speed 343 m/s, density 1.2 kg/m^3, dt = 0.5 h / (c sqrt 3), PML strength 1e6,
preset 0.  All randomness comes from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import math

import numpy as np

from .capi import (FDTD_SCHEME_STAGGERED, FDTD_BOUNDARY_FULLY_COUPLED, FDTD_BOUNDARY_RASTERIZE, FDTD_OBJ_ELLIPSOID,
                   FDTD_OBJ_SPHERE, FDTD_VIB_CHIRP, FDTD_VIB_HARMONIC, SceneBuffers)

C_SOUND = 343.0
RHO = 1.2
PML_STRENGTH = 1.0e6


def cfl_dt(h, c=C_SOUND, cfl=0.5):
    return cfl * h / (c * math.sqrt(3.0))


def base_settings(div, h, pml_width, boundary=FDTD_BOUNDARY_RASTERIZE, c=C_SOUND, rho=RHO, dt=None):
    div = tuple(int(x) for x in div)
    dt = cfl_dt(h, c) if dt is None else dt
    # BoundingBox(cellSize, cellDivisions, center) puts min = center - h*n/2
    min_bound = tuple(0.0 - h * float(n) / 2.0 for n in div)
    return dict(sound_speed=c, density=rho, cell_size=h, time_step=dt, stop_time=0.0,
                pml_width=float(pml_width), pml_strength=PML_STRENGTH, air_viscosity_alpha=0.0,
                start_time=0.0, boundary_condition_preset=0, boundary_handling=boundary, use_mesh=1,
                divisions=div, min_bound=min_bound)


def interior_bounds(settings, margin_cells):
    """Lower/upper corner of the region >= margin_cells inside the PML."""
    h = settings["cell_size"]
    w = settings["pml_width"]
    lo = np.array(settings["min_bound"]) + (w + margin_cells) * h
    hi = np.array(settings["min_bound"]) + (np.array(settings["divisions"]) - w - margin_cells) * h
    return lo, hi


def random_rotation(rng):
    """Uniform random rotation from a uniform unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def fibonacci_shell(center, radius, n):
    pts = []
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        z = 1.0 - 2.0 * (i + 0.5) / n
        r = math.sqrt(max(0.0, 1.0 - z * z))
        th = ga * i
        pts.append([center[0] + radius * r * math.cos(th), center[1] + radius * r * math.sin(th),
                    center[2] + radius * z])
    return np.array(pts)


def gaussian_pressure_source(position, width_time, c=C_SOUND, offset=None):
    ws = c * width_time
    return dict(position=tuple(position), width_time=width_time, width_space=ws,
                offset_time=4.0 * width_time if offset is None else offset,
                normalize_constant=1.0 / (math.sqrt(2.0 * math.pi) * ws) ** 3)


def _object_clear(p, objects, h, cells):
    for o in objects:
        r = max(o["semi_axes"])
        if np.linalg.norm(np.asarray(p) - np.asarray(o["center"])) < r + cells * h:
            return False
    return True


def _place_ellipsoids(rng, settings, count, axis_lo, axis_hi, margin_cells=2, gap_cells=2, tries=10000):
    h = settings["cell_size"]
    lo, hi = interior_bounds(settings, margin_cells)
    objs = []
    for _ in range(tries):
        if len(objs) == count:
            break
        axes = rng.uniform(axis_lo, axis_hi, size=3)
        r = axes.max()
        if np.any(hi - lo <= 2 * r):
            continue
        c = rng.uniform(lo + r, hi - r)
        if all(np.linalg.norm(c - np.asarray(o["center"])) > r + max(o["semi_axes"]) + gap_cells * h for o in objs):
            objs.append(dict(type=FDTD_OBJ_ELLIPSOID, center=tuple(c), semi_axes=tuple(axes),
                             rotation=random_rotation(rng)))
    if len(objs) != count:
        raise RuntimeError("could not place non-overlapping ellipsoids")
    return objs


def _random_listeners(rng, settings, objects, count, margin_cells=2, clear_cells=2):
    h = settings["cell_size"]
    lo, hi = interior_bounds(settings, margin_cells)
    pts = []
    while len(pts) < count:
        p = rng.uniform(lo, hi)
        if _object_clear(p, objects, h, clear_cells):
            pts.append(p)
    return np.array(pts)


def config1(seed=1, n=64, steps=500):
    """Free field: n^3 grid, h = 5 mm, 10-cell PML, one Gaussian pressure pulse
    (width 2e-4 s) at a seeded interior point, 8 seeded listeners."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), 0.005, 10 if n >= 40 else max(2, n // 6))
    lo, hi = interior_bounds(s, 2)
    src = gaussian_pressure_source(rng.uniform(lo, hi), 2e-4)
    listeners = rng.uniform(lo, hi, size=(8, 3))
    return SceneBuffers(s, pressure_sources=[src], listeners=listeners), steps


def staggered_free_field(seed=7, div=(40, 40, 40), h=0.005, pml=6, steps=300, n_listeners=8):
    """Free field in the reference's staggered (MAC) scheme: one Gaussian
    pressure pulse (width 2e-4 s, time-integrated by that scheme) at a seeded
    interior point, seeded listeners."""
    rng = np.random.default_rng(seed)
    s = base_settings(div, h, pml)
    s["scheme"] = FDTD_SCHEME_STAGGERED
    lo, hi = interior_bounds(s, 2)
    src = gaussian_pressure_source(rng.uniform(lo, hi), 2e-4)
    listeners = rng.uniform(lo, hi, size=(n_listeners, 3))
    return SceneBuffers(s, pressure_sources=[src], listeners=listeners), steps


def engquist_free_field(seed=8, div=(40, 44, 48), h=0.005, steps=300, n_listeners=8, use_mesh=0):
    """No PML: the Engquist absorbing boundary on the layer-1 cells
    (use_mesh = 0, MAC_Grid.cpp:521-919, 2457-2473).  One Gaussian pressure
    pulse (width 2e-4 s) at a seeded point, seeded listeners, plus one
    listener whose MLS cells include the face cells of layer 1 and one next
    to the (1, 1, 1) corner."""
    rng = np.random.default_rng(seed)
    s = base_settings(div, h, 0)
    s["use_mesh"] = use_mesh
    lo, hi = interior_bounds(s, 4)
    src = gaussian_pressure_source(rng.uniform(lo, hi), 2e-4)
    mb = np.array(s["min_bound"])
    mid = mb + 0.5 * h * np.array(div)
    near_face = mid.copy()
    near_face[0] = mb[0] + 1.7 * h
    near_corner = mb + 1.6 * h
    listeners = np.vstack([rng.uniform(lo, hi, size=(n_listeners - 2, 3)), near_face, near_corner])
    return SceneBuffers(s, pressure_sources=[src], listeners=listeners), steps


def engquist_sphere(seed=9, n=44, h=0.005, radius=0.025, shell=0.06, steps=300, n_listeners=10,
                    boundary=FDTD_BOUNDARY_RASTERIZE):
    """A rigid sphere with an impulse (as in config2) in a grid with a PML
    width of zero: object cells are classified (use_mesh = 1) and the layer-1
    cells take the Engquist boundary."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, 0, boundary)
    c = rng.uniform(-0.2, 0.2, size=3) * (n * h / 2 - radius - 4 * h)
    obj = dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(radius, radius, radius), rotation=np.eye(3))
    src = dict(object=0, amplitude=1e3, width=1e-4, onset=4e-4)
    listeners = fibonacci_shell(c, shell, n_listeners)
    return SceneBuffers(s, objects=[obj], impulse_sources=[src], listeners=listeners), steps


def moving_object(seed=10, n=48, h=0.005, pml=8, steps=400, kind="sphere", speed=20.0, spin=0.0, n_listeners=8,
                  pml_free=False):
    """One rigid object moving through the grid for `steps` steps (constant
    velocity of `speed` m/s in a seeded direction, and for an ellipsoid a
    constant angular velocity of `spin` rad/s about a seeded axis), with a
    Gaussian impulse (1e3 m/s^2, 1e-4 s, onset 3e-4 s); stop_time bounds the
    planned path.  Listeners on a shell around the path's midpoint, clear of
    every cell the object can reach.  pml_free: a PML width of zero (the
    Engquist boundary) instead of `pml` cells."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, 0 if pml_free else pml)
    dt = s["time_step"]
    s["stop_time"] = s["start_time"] + steps * dt
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    v = speed * u
    T = steps * dt
    mid = rng.uniform(-0.2, 0.2, size=3) * h
    c0 = mid - 0.5 * v * T
    if kind == "sphere":
        r = 0.02
        obj = dict(type=FDTD_OBJ_SPHERE, center=tuple(c0), semi_axes=(r, r, r), rotation=np.eye(3), velocity=tuple(v))
    else:
        axes = (0.024, 0.012, 0.017)
        r = max(axes)
        w = rng.normal(size=3)
        w = spin * w / np.linalg.norm(w)
        obj = dict(type=FDTD_OBJ_ELLIPSOID, center=tuple(c0), semi_axes=axes, rotation=random_rotation(rng),
                   velocity=tuple(v), angular_velocity=tuple(w))
    src = dict(object=0, amplitude=1e3, width=1e-4, onset=3e-4)
    # outside the path's box (the object's reach along each axis plus the
    # two-cell margin and the listener's own interpolation cells)
    half = 0.5 * np.abs(v) * T + r + 4.5 * h
    pts = fibonacci_shell(mid, float(np.max(half)) + 1.5 * h, 16 * n_listeners)
    keep = [p for p in pts if np.any(np.abs(p - mid) > half)]
    listeners = np.array(keep[::max(1, len(keep) // n_listeners)][:n_listeners])
    return SceneBuffers(s, objects=[obj], impulse_sources=[src], listeners=listeners), steps


def two_moving_objects(seed=14, n=56, h=0.005, pml=8, steps=360, n_listeners=8):
    """A sphere and a spinning ellipsoid passing each other (opposite
    velocities of 12 m/s, 4 cells apart across the path), each with its own
    impulse, plus a Gaussian pressure pulse between them; listeners outside the
    box the two paths sweep."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml)
    dt = s["time_step"]
    s["stop_time"] = s["start_time"] + steps * dt
    T = steps * dt
    v = np.array([12.0, 0.0, 0.0])
    off = np.array([0.0, 0.0, 0.014 + 2 * h])
    sph = dict(type=FDTD_OBJ_SPHERE, center=tuple(-0.5 * v * T - off), semi_axes=(0.012, 0.012, 0.012),
               rotation=np.eye(3), velocity=tuple(v))
    ell = dict(type=FDTD_OBJ_ELLIPSOID, center=tuple(0.5 * v * T + off), semi_axes=(0.016, 0.009, 0.012),
               rotation=random_rotation(rng), velocity=tuple(-v), angular_velocity=(0.0, 250.0, 120.0))
    srcs = [dict(object=0, amplitude=1e3, width=1e-4, onset=2e-4), dict(object=1, amplitude=-8e2, width=1e-4, onset=5e-4)]
    ps = gaussian_pressure_source((0.0, 0.01, 0.0), 1e-4, offset=4e-4)
    half = np.array([0.5 * 12.0 * T + 0.016, 0.016, 0.014 + 2 * h + 0.016]) + 4.5 * h
    pts = fibonacci_shell(np.zeros(3), float(np.max(half)) + 1.5 * h, 16 * n_listeners)
    keep = [p for p in pts if np.any(np.abs(p) > half)]
    listeners = np.array(keep[::max(1, len(keep) // n_listeners)][:n_listeners])
    return SceneBuffers(s, objects=[sph, ell], impulse_sources=srcs, pressure_sources=[ps], listeners=listeners), steps


def tone_burst(fs=44100, freq=2000.0, cycles=4, amplitude=0.8):
    """A Hann-windowed sine burst (the samples of a speaker file)."""
    n = int(round(cycles * fs / freq))
    t = np.arange(n) / fs
    return amplitude * np.sin(2 * np.pi * freq * t) * np.sin(np.pi * np.arange(n) / n) ** 2


def speaker_scene(seed=15, n=48, h=0.005, pml=8, steps=300, n_listeners=8, with_sphere=False):
    """A speaker pressure source (SpeakerPressureSource, static position,
    a 2 kHz tone burst at 44.1 kHz starting at 1e-4 s, width 8 mm) at a seeded
    interior point, seeded listeners; optionally a rigid sphere next to it."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml)
    lo, hi = interior_bounds(s, 6)
    pos = rng.uniform(lo, hi)
    spk = dict(position=tuple(pos), width_space=0.008, start_time=1e-4, sample_rate=44100.0, samples=tone_burst())
    objs, imps = [], []
    if with_sphere:
        c = pos + np.array([0.035, 0.0, 0.0])
        objs = [dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(0.015,) * 3, rotation=np.eye(3))]
    listeners = _random_listeners(rng, s, objs, n_listeners)
    return SceneBuffers(s, objects=objs, impulse_sources=imps, listeners=listeners, speaker_sources=[spk]), steps


def speaker_tie_scene(seed=17, n=40, pml=8, steps=400, n_listeners=8):
    """A speaker on a cell centre whose sample index lands on (or within an ulp
    of) an integer at every step: h = 2^-8 m on an even grid, so cell centres
    are exact in binary and the speaker's distance to its own cell is zero;
    the sample period is dt / 2 and the samples alternate in sign.  Which
    sample a step takes then depends on the last bits of the step's time, so
    the series matches the reference's only when the time accumulates as its
    clock does (one addition of dt per step)."""
    rng = np.random.default_rng(seed)
    h = 2.0 ** -8
    s = base_settings((n, n, n), h, pml)
    c = n // 2
    pos = tuple(float(s["min_bound"][d] + (0.5 + c) * h) for d in range(3))
    m = 2 * steps + 16
    samples = np.array([(-1.0) ** i * math.sin(math.pi * (i + 0.5) / m) for i in range(m)])
    spk = dict(position=pos, width_space=1.5 * h, start_time=0.0, sample_rate=2.0 / s["time_step"], samples=samples)
    listeners = _random_listeners(rng, s, [], n_listeners)
    return SceneBuffers(s, objects=[], impulse_sources=[], listeners=listeners, speaker_sources=[spk]), steps


def config2(seed=2, n=128, steps=2000, boundary=FDTD_BOUNDARY_RASTERIZE, pml=12, radius=0.05, h=0.005,
            n_listeners=64, shell=0.2):
    """Rigid sphere (radius 0.05 m, seeded centre) with a Gaussian-impulse
    normal acceleration (1e3 m/s^2, 1e-4 s, onset 4e-4 s); listeners on a shell."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml, boundary)
    c = rng.uniform(-0.05, 0.05, size=3) * (n * h / 0.64)
    obj = dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(radius, radius, radius), rotation=np.eye(3))
    src = dict(object=0, amplitude=1e3, width=1e-4, onset=4e-4)
    listeners = fibonacci_shell(c, shell, n_listeners)
    return SceneBuffers(s, objects=[obj], impulse_sources=[src], listeners=listeners), steps


def config3(seed=3, n=256, steps=4000, boundary=FDTD_BOUNDARY_RASTERIZE, h=0.0025, pml=16, n_obj=5,
            n_listeners=128, axis_lo=0.01, axis_hi=0.04):
    """Five rigid ellipsoids (semi-axes U(0.01, 0.04) m, random orientation,
    non-overlapping) each with its own impulse at a seeded onset in [0, 2 ms]."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml, boundary)
    objs = _place_ellipsoids(rng, s, n_obj, axis_lo, axis_hi)
    srcs = [dict(object=i, amplitude=1e3, width=1e-4, onset=float(rng.uniform(0, 2e-3))) for i in range(n_obj)]
    listeners = _random_listeners(rng, s, objs, n_listeners)
    return SceneBuffers(s, objects=objs, impulse_sources=srcs, listeners=listeners), steps


def _small_object_scene(rng, div, h, pml, n_obj, n_listeners, boundary=FDTD_BOUNDARY_RASTERIZE,
                        sphere_or_ellipsoid=True):
    s = base_settings(div, h, pml, boundary)
    lo, hi = interior_bounds(s, 2)
    half = float(np.min(hi - lo)) / 2.0
    amax = min(0.04, half / 4.0)
    amin = min(0.01, 0.5 * amax)
    objs = []
    if sphere_or_ellipsoid and n_obj == 1 and rng.random() < 0.5:
        r = float(rng.uniform(amin, amax))
        c = rng.uniform(lo + r, hi - r)
        objs.append(dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(r, r, r), rotation=np.eye(3)))
    else:
        objs = _place_ellipsoids(rng, s, n_obj, amin, amax)
    srcs = [dict(object=i, amplitude=1e3, width=1e-4, onset=float(rng.uniform(0, 1e-3))) for i in range(n_obj)]
    listeners = _random_listeners(rng, s, objs, n_listeners)
    return SceneBuffers(s, objects=objs, impulse_sources=srcs, listeners=listeners)


def config4(seed=4, n_scenes=64, edge_lo=32, edge_hi=128, steps_lo=1000, steps_hi=3000):
    """Batch: independent scenes with per-axis edges U{edge_lo..edge_hi},
    h = 5 mm, 8-cell PML, one sphere or ellipsoid with an impulse, 16 listeners."""
    rng = np.random.default_rng(seed)
    scenes, steps = [], []
    for _ in range(n_scenes):
        div = tuple(int(x) for x in rng.integers(edge_lo, edge_hi + 1, size=3))
        scenes.append(_small_object_scene(rng, div, 0.005, 8, 1, 16))
        steps.append(int(rng.integers(steps_lo, steps_hi + 1)))
    return scenes, steps


def config5_stream(seed=5, n_requests=64, rate=200.0, edge_lo=48, edge_hi=192, dur_lo=0.01, dur_hi=0.05):
    """Serving: Poisson arrivals (rate req/s) of impact-sound requests with
    1-3 ellipsoids, 32 listeners, simulated duration U(dur_lo, dur_hi) s."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    h = 0.005
    dt = cfl_dt(h)
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        div = tuple(int(x) for x in rng.integers(edge_lo, edge_hi + 1, size=3))
        n_obj = int(rng.integers(1, 4))
        sc = _small_object_scene(rng, div, h, 8, n_obj, 32, sphere_or_ellipsoid=False)
        dur = float(rng.uniform(dur_lo, dur_hi))
        out.append((t, sc, int(math.ceil(dur / dt))))
    return out


def floor_scene(seed=16, n=48, h=0.005, pml=8, direction=1, sign=1, layers=6, frac=0.0, obj="sphere",
                boundary=FDTD_BOUNDARY_RASTERIZE, steps=300, n_listeners=10, source=True):
    """A plane constraint (XML <constraint direction sign height>) `layers`
    cells inside the interior from the PML on the side its sign makes solid,
    `frac` of a cell past a cell face, with a rigid object resting two cells
    above it (a sphere or an ellipsoid with a Gaussian impulse), an optional
    Gaussian pressure pulse (spatial width 2 h) and seeded listeners on the
    fluid side."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml, boundary)
    mb = s["min_bound"][direction]
    if sign > 0:
        height = mb + (pml + layers + frac) * h
    else:
        height = mb + (n - pml - layers - frac) * h
    cons = [dict(direction=direction, sign=sign, height=height)]
    up = np.zeros(3)
    up[direction] = float(sign)
    objs, imps = [], []
    r = 0.012
    c = rng.uniform(-0.2, 0.2, size=3) * h * 4
    c[direction] = height + sign * (r + 2 * h)
    if obj == "sphere":
        objs.append(dict(type=FDTD_OBJ_SPHERE, center=tuple(c), semi_axes=(r, r, r), rotation=np.eye(3)))
    elif obj == "ellipsoid":
        objs.append(dict(type=FDTD_OBJ_ELLIPSOID, center=tuple(c), semi_axes=(r, 0.7 * r, 0.85 * r),
                         rotation=random_rotation(rng)))
    if objs:
        imps.append(dict(object=0, amplitude=1e3, width=1e-4, onset=3e-4))
    ps = []
    if source:
        pos = c + up * 0.03 + np.array([0.011, -0.007, 0.004])
        src = gaussian_pressure_source(pos, 1e-4, offset=4e-4)
        src["width_space"] = 2 * h
        src["normalize_constant"] = 1.0 / (math.sqrt(2 * math.pi) * 2 * h) ** 3
        ps.append(src)
    lo, hi = interior_bounds(s, 2)
    if sign > 0:
        lo[direction] = height + 3 * h
    else:
        hi[direction] = height - 3 * h
    pts = []
    while len(pts) < n_listeners:
        p = rng.uniform(lo, hi)
        if _object_clear(p, objs, h, 3):
            pts.append(p)
    sb = SceneBuffers(s, objects=objs, impulse_sources=imps, pressure_sources=ps, listeners=np.array(pts),
                      constraints=cons)
    return sb, steps


def uniform_solid_inv_inertia(mass, semi_axes):
    """Inverse inertia tensor (object frame) of a uniform solid ellipsoid:
    I = m diag(b^2 + c^2, a^2 + c^2, a^2 + b^2) / 5 (a sphere: 2/5 m r^2)."""
    a, b, c = semi_axes
    return np.diag([5.0 / (mass * (b * b + c * c)), 5.0 / (mass * (a * a + c * c)), 5.0 / (mass * (a * a + b * b))])


def vibration_scene(seed=17, n=48, h=0.005, pml=8, steps=300, n_obj=2, n_listeners=10, harmonic=True,
                    acc_noise=True, gaussian=False, chirp=False, boundary=FDTD_BOUNDARY_RASTERIZE):
    """Rotated ellipsoids whose surfaces carry harmonic pulsations (seeded
    frequency 1-4 kHz, phase, decay and time range) and acceleration noise
    (three impacts each at seeded off-centre points, overlapping supports of
    0.1-0.4 ms), optionally a Gaussian impulse too."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml, boundary)
    objs = _place_ellipsoids(rng, s, n_obj, 0.012, 0.03)
    dt = s["time_step"]
    T = steps * dt
    isrc, ans = [], []
    for i, o in enumerate(objs):
        if harmonic:
            om = 2 * math.pi * rng.uniform(1e3, 4e3)
            t0 = float(rng.uniform(0, 0.2)) * T
            isrc.append(dict(object=i, kind=FDTD_VIB_HARMONIC, omega=om, phase=float(rng.uniform(0, 2 * math.pi)),
                             amplitude=float(rng.uniform(0.5, 2.0)), decay_rate=float(rng.choice([-1.0, 1.5 / T])),
                             decay_center=0.3 * T, t0=t0, t1=t0 + 0.6 * T))
        if gaussian:
            isrc.append(dict(object=i, amplitude=1e3, width=1e-4, onset=float(rng.uniform(0, 0.4)) * T))
        if chirp:
            c0 = float(rng.uniform(0.05, 0.2)) * T
            isrc.append(dict(object=i, kind=FDTD_VIB_CHIRP, omega=2 * math.pi * rng.uniform(1e3, 2e3),
                             omega1=2 * math.pi * rng.uniform(2.5e3, 4e3), t0=c0, t1=c0 + 0.5 * T))
        if acc_noise:
            m = float(rng.uniform(0.05, 0.5))
            imps = []
            for q in range(3):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                imps.append(dict(timestamp=float(rng.uniform(0.05, 0.5)) * T, support=float(rng.uniform(1e-4, 4e-4)),
                                 impulse=tuple(rng.normal(size=3) * 2e-3), position=tuple(np.asarray(o["semi_axes"]) * d)))
            ans.append(dict(object=i, mass=m, center_of_mass=tuple(rng.normal(size=3) * 1e-3),
                            inv_inertia=uniform_solid_inv_inertia(m, o["semi_axes"]), impacts=imps))
    listeners = _random_listeners(rng, s, objs, n_listeners)
    return SceneBuffers(s, objects=objs, impulse_sources=isrc, listeners=listeners, acc_noise_sources=ans), steps


def quat_mul(a, b):
    """Hamilton product of quaternions (w, x, y, z)."""
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by + ay * bw + az * bx - ax * bz, aw * bz + az * bw + ax * by - ay * bx])


def quat_from_matrix(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    R = np.asarray(R, dtype=np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = 2.0 * math.sqrt(1.0 + tr)
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * math.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
        q = [0.0] * 4
        q[0] = (R[k, j] - R[j, k]) / s
        q[1 + i] = 0.25 * s
        q[1 + j] = (R[j, i] + R[i, j]) / s
        q[1 + k] = (R[k, i] + R[i, k]) / s
    q = np.array(q)
    return q / np.linalg.norm(q)


# a kinematics file's rotation is Blender's z-up one: the reader turns it by
# -90 degrees about x (q <- q * corr); a pose q is written as q * corr^-1
BLENDER_CORR = np.array([math.cos(-math.pi / 4), math.sin(-math.pi / 4), 0.0, 0.0])


def file_frame(center, R):
    """One kinematics-file row (tx ty tz qw qx qy qz) that the reader turns
    into the pose (center, R)."""
    inv = BLENDER_CORR * np.array([1.0, -1.0, -1.0, -1.0])
    q = quat_mul(quat_from_matrix(R), inv)
    return [*center, *q]


def rodrigues(axis, angle):
    k = np.asarray(axis, dtype=np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def as_keyframes(sb, frame_rate):
    """The same scene with every moving object's constant motion written as
    kinematics-file frames (frame f at time f / frame_rate, to one frame past
    stop_time) and its own velocities zeroed.  Linear blending of a constant
    velocity and slerp of a constant-rate turn reproduce the motion."""
    s = dict(sb.settings_dict)
    step = 1.0 / frame_rate
    nf = int(math.ceil(s["stop_time"] / step)) + 2
    objs, kins = [], []
    for i, o in enumerate(sb.objects_list):
        o = dict(o)
        v = np.asarray(o.get("velocity", (0.0, 0.0, 0.0)), dtype=np.float64)
        w = np.asarray(o.get("angular_velocity", (0.0, 0.0, 0.0)), dtype=np.float64)
        if np.any(v != 0) or np.any(w != 0):
            R0 = np.asarray(o.get("rotation", np.eye(3)), dtype=np.float64).reshape(3, 3)
            frames, prev = [], None
            for f in range(nf):
                t = f * step
                c = np.asarray(o["center"]) + v * t
                R = (rodrigues(w, np.linalg.norm(w) * t) if np.any(w != 0) else np.eye(3)) @ R0
                row = np.array(file_frame(c, R))
                if prev is not None and np.dot(row[3:], prev[3:]) < 0:
                    row[3:] = -row[3:]          # the same turn, on the side nearer the last frame
                frames.append(row)
                prev = row
            kins.append(dict(object=i, step_size=step, frames=np.array(frames)))
            o["velocity"] = (0.0, 0.0, 0.0)
            o["angular_velocity"] = (0.0, 0.0, 0.0)
        objs.append(o)
    return SceneBuffers(s, objects=objs, impulse_sources=sb.impulse_list, pressure_sources=sb.pressure_list,
                        listeners=sb.listeners_np, kinematics=kins)


def keyframed_object(seed=22, n=56, h=0.005, pml=8, steps=360, frame_rate=4000.0, n_listeners=8):
    """An ellipsoid on a curved, accelerating path with a varying turn, given
    as kinematics-file frames (a swing: centre on an arc, rotation about a
    wobbling axis), with a Gaussian impulse; listeners outside the swept box."""
    rng = np.random.default_rng(seed)
    s = base_settings((n, n, n), h, pml)
    dt = s["time_step"]
    T = steps * dt
    s["stop_time"] = s["start_time"] + T
    axes = (0.02, 0.011, 0.015)
    R0 = random_rotation(rng)
    step = 1.0 / frame_rate
    nf = int(math.ceil(T / step)) + 2
    rad, om = 0.02, 2.5 / T
    frames, prev = [], None
    for f in range(nf):
        t = f * step
        c = np.array([rad * (math.cos(om * t) - 0.1), 0.4 * rad * math.sin(2 * om * t), 0.2 * rad * t / T])
        ax = np.array([math.sin(3 * om * t), 1.0, 0.3 * math.cos(om * t)])
        R = rodrigues(ax, 1.2 * om * t + 0.8 * (om * t) ** 2) @ R0
        row = np.array(file_frame(c, R))
        if prev is not None and np.dot(row[3:], prev[3:]) < 0:
            row[3:] = -row[3:]
        frames.append(row)
        prev = row
    obj = dict(type=FDTD_OBJ_ELLIPSOID, center=(0.0, 0.0, 0.0), semi_axes=axes, rotation=np.eye(3))
    src = dict(object=0, amplitude=1e3, width=1e-4, onset=3e-4)
    fr = np.array(frames)
    lo, hi = fr[:, :3].min(0), fr[:, :3].max(0)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo) + max(axes) + 4.5 * h
    pts = fibonacci_shell(mid, float(np.max(half)) + 1.5 * h, 16 * n_listeners)
    keep = [p for p in pts if np.any(np.abs(p - mid) > half)]
    listeners = np.array(keep[::max(1, len(keep) // n_listeners)][:n_listeners])
    return SceneBuffers(s, objects=[obj], impulse_sources=[src], listeners=listeners,
                        kinematics=[dict(object=0, step_size=step, frames=fr)]), steps


def with_settings(sb, **changes):
    """The same scene with some settings changed (e.g. sampling=FDTD_SAMPLE_SEVEN)."""
    s = dict(sb.settings_dict)
    s.update(changes)
    return SceneBuffers(s, objects=sb.objects_list, impulse_sources=sb.impulse_list, pressure_sources=sb.pressure_list,
                        listeners=sb.listeners_np, speaker_sources=sb.speaker_list, constraints=sb.constraint_list,
                        acc_noise_sources=sb.acc_noise_list, kinematics=sb.kinematics_list)
