"""Pins of the CPU checker's participating-media code.

The checker restates the reference's media code: the majorant-grid DDA
(MajorantIterator, src/render/media.h:72-98, behind sampleRay :41-106),
delta tracking with medium emission and scattering (sampleT_maj,
src/render/wavefront/wavefront.h:34-78; sampleMediumInteraction,
medium.cpp:13-103), the phase-function continuation with next-event
estimation (sampleMediumScattering, medium.cpp:105-153) and ratio tracking
of shadow rays (traceTransmittance, wavefront.h:80-139).  Each test below
holds one of these against something the mathematics fixes:

* DDA walk: the majorant segments equal a brute-force walk over all 64^3
  majorant cells (per-cell slab intervals of the ray, sorted), with each
  segment's majorant the cell's.
* Ratio tracking: the mean of the shadow-ray estimator equals exp(-tau),
  tau the optical depth - exp(-sigma d) for homogeneous and constant-grid
  media, the numerically integrated trilinear density for a heterogeneous
  grid - also per channel for a chromatic medium (hero channel averaged),
  through a null-material surface, and deep enough for its Russian roulette.
* Delta tracking + emission + scattering: inside an unbounded homogeneous
  medium that emits Le and scatters with albedo a, a camera ray carries
  Le (1 - a^D) at max depth D, per channel, whatever sigma_t.
* Next-event estimation in a heterogeneous scattering medium: images with
  NEE on and off have the same expectation at a max depth deep enough for
  the truncated tail to vanish (the MIS weights of light samples and of
  phase/BSDF-sampled light hits sum to one only if the ratio bookkeeping of
  delta and ratio tracking is right).

Each pin is also run against the checker built with a planted mistake
(oracle/Makefile "plant", ORACLE_PLANT in oracle/core.h) and must fail
there: a dropped null-collision ratio, a wrong DDA axis, a wrong roulette
compensation, sigma_s for sigma_a in the emission, a dropped phase pdf in
the light sample's MIS ratio, and sigma_n for sigma_maj in the light-path
ratio of a null collision.

Statistical tolerances are 5 standard errors.
"""
import numpy as np
import pytest

FAR_QUAD = {"type": "quad", "corners": [[90, 90, 90], [91, 90, 90], [91, 91, 90], [90, 91, 90]], "material": 0}
BLACK = {"bsdf": "diffuse", "diffuse": [0, 0, 0]}


def _scene(krr, media, models=(FAR_QUAD,), materials=(BLACK,), res=(8, 8), camera=None):
    cfg = {"media": list(media), "model": list(models), "materials": list(materials), "resolution": list(res),
           "camera": camera or {"eye": [0.5, 0.5, 3], "target": [0.5, 0.5, 0], "medium": 0}}
    return krr.Scene(cfg)


# ---------------------------------------------------------------------------
# Majorant DDA against a brute-force walk
# ---------------------------------------------------------------------------

def _bruteforce_walk(med, o, d, tmax, majorant):
    """Every majorant cell the ray o + t d, t in [0, tmax], passes through,
    as (t0, t1, cell) sorted by t0: the slab intervals of all 64^3 cells in
    the grid's [0,1]^3 space (the world box scaled per axis)."""
    lo = np.array(med.bounds_min[:3], np.float64)
    hi = np.array(med.bounds_max[:3], np.float64)
    oo = (np.asarray(o, np.float64) - lo) / (hi - lo)
    dd = np.asarray(d, np.float64) / (hi - lo)
    edges = np.arange(65, dtype=np.float64) / 64.0
    t_lo, t_hi = [], []
    for a in range(3):
        if dd[a] != 0.0:
            ta, tb = (edges[:-1] - oo[a]) / dd[a], (edges[1:] - oo[a]) / dd[a]
            t_lo.append(np.minimum(ta, tb))
            t_hi.append(np.maximum(ta, tb))
        else:  # parallel to this axis: the cells whose slab holds the origin
            inside = (edges[:-1] <= oo[a]) & (oo[a] < edges[1:])
            t_lo.append(np.where(inside, -np.inf, np.inf))
            t_hi.append(np.where(inside, np.inf, -np.inf))
    t0 = np.maximum(np.maximum(t_lo[0][:, None, None], t_lo[1][None, :, None]), t_lo[2][None, None, :])
    t1 = np.minimum(np.minimum(t_hi[0][:, None, None], t_hi[1][None, :, None]), t_hi[2][None, None, :])
    t0 = np.maximum(t0, 0.0)
    t1 = np.minimum(t1, tmax)
    ix, iy, iz = np.nonzero(t1 > t0)
    a0, a1 = t0[ix, iy, iz], t1[ix, iy, iz]
    order = np.argsort(a0)
    cells = np.stack([ix, iy, iz], 1)[order]
    sig = np.array([med.sigma_t[c] for c in range(3)]) * med.scale
    maj = majorant.reshape(64, 64, 64)[cells[:, 2], cells[:, 1], cells[:, 0]]  # x fastest
    return np.stack([a0[order], a1[order]], 1), maj[:, None] * sig[None, :]


def _dda_rays(rng, med, n):
    lo = np.array(med.bounds_min[:3]); hi = np.array(med.bounds_max[:3])
    rays = []
    for i in range(n):
        if i % 3 == 0:  # from inside the box
            o = lo + rng.uniform(0.05, 0.95, 3) * (hi - lo)
        else:  # from outside, aimed at a point inside
            o = lo + rng.uniform(-0.6, 1.6, 3) * (hi - lo)
        target = lo + rng.uniform(0.1, 0.9, 3) * (hi - lo)
        d = (target - o) * rng.uniform(0.3, 3.0)
        if i % 5 == 1:
            d[rng.integers(3)] = 0.0  # parallel to one axis
        if not np.any(d):
            d[0] = 1.0
        tmax = np.inf if i % 4 else rng.uniform(0.2, 1.5)  # some rays end inside the box
        rays.append((o.astype(np.float32), d.astype(np.float32), float(np.float32(tmax))))  # the walk takes a binary32 tmax
    return rays


def pin_dda(oracle, krr, lib=None, precision=64, n_rays=40):
    """Largest disagreement between the checker's majorant walk and the
    brute-force walk over random rays, as (ok, detail)."""
    rng = np.random.default_rng(5)
    dens = rng.uniform(0.0, 1.0, 10 * 7 * 12).round(4).tolist()
    s = _scene(krr, [{"type": "dense_grid", "res": [10, 7, 12], "density": dens, "sigma_t": [1.5, 2.0, 0.5],
                      "scale": 1.25, "bounds": [[-1.0, 0.0, 0.2], [2.0, 0.5, 3.0]]}])
    med = s.desc.media[0]
    majorant = np.zeros(64 ** 3, np.float32)
    assert oracle.load().oracle_majorant_grid(med, oracle._ptr(majorant)) == 0
    tol = 1e-9 if precision == 64 else 2e-5
    for k, (o, d, tmax) in enumerate(_dda_rays(rng, med, n_rays)):
        seg = oracle.majorant_segments(med, o, d, tmax, precision=precision, lib=lib)
        bf_t, bf_sig = _bruteforce_walk(med, o.astype(np.float64), d.astype(np.float64), tmax, majorant)
        # zero-length segments (a walk through an exact cell corner) carry no medium
        keep = seg[:, 1] - seg[:, 0] > tol * 8
        seg = seg[keep]
        keep_bf = bf_t[:, 1] - bf_t[:, 0] > tol * 8
        bf_t, bf_sig = bf_t[keep_bf], bf_sig[keep_bf]
        if len(seg) != len(bf_t):
            return False, f"ray {k}: {len(seg)} segments, brute force {len(bf_t)}"
        if len(seg) == 0:
            continue
        scale = max(1.0, float(np.abs(bf_t[np.isfinite(bf_t)]).max()))
        dt = np.abs(seg[:, :2] - bf_t).max() / scale
        ds = np.abs(seg[:, 2:] - bf_sig).max()
        if dt > tol or ds > 1e-6 * max(1.0, bf_sig.max()):
            return False, f"ray {k}: interval error {dt:.3e}, majorant error {ds:.3e}"
    return True, "ok"


@pytest.mark.parametrize("precision", [64, 32])
def test_majorant_walk_matches_bruteforce(oracle, krr, precision):
    ok, detail = pin_dda(oracle, krr, precision=precision)
    assert ok, detail


def test_majorant_walk_homogeneous_is_one_segment(oracle, krr):
    """A homogeneous medium has one segment [0, tmax] with sigma_t
    (media.h:121-126: the iterator of a HomogeneousMedium)."""
    s = _scene(krr, [{"type": "homogeneous", "sigma_t": [0.5, 1.0, 2.0], "albedo": [0.5] * 3}])
    seg = oracle.majorant_segments(s.desc.media[0], (0, 0, 0), (0, 0, -1), 3.5)
    assert seg.shape == (1, 5)
    assert np.allclose(seg[0], [0, 3.5, 0.5, 1.0, 2.0])


# ---------------------------------------------------------------------------
# Ratio tracking against the optical depth
# ---------------------------------------------------------------------------

def _trilinear(dens, res, lo, hi, p):
    """Textbook trilinear interpolation of voxel values sitting at integer
    index coordinates, zero outside the grid (index = (p - lo)/(hi - lo) * res)."""
    g = dens.reshape(res[2], res[1], res[0])  # z, y, x
    idx = (p - lo) / (hi - lo) * np.asarray(res, np.float64)
    i0 = np.floor(idx).astype(np.int64)
    f = idx - i0
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                x, y, z = i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz
                ok = (x >= 0) & (y >= 0) & (z >= 0) & (x < res[0]) & (y < res[1]) & (z < res[2])
                v = np.zeros(len(p))
                v[ok] = g[z[ok], y[ok], x[ok]]
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out += w * v
    return out


def _optical_depth(dens, res, lo, hi, sigma, o, d, tmax, n=400001):
    """Simpson integral of sigma * density along o + t d, t in [0, tmax],
    over the part of the ray inside the grid's bounds (the medium exists
    only there: sampleRay clips the ray to the bounds, media.h:41-106, while
    the trilinear field itself reaches half a voxel past the low faces).  The
    field is piecewise smooth; n points make the error far below the
    statistical tolerance."""
    o, d = np.asarray(o, np.float64), np.asarray(d, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ta, tb = (lo - o) / d, (hi - o) / d
    t0 = max(0.0, float(np.nanmax(np.minimum(ta, tb))))
    t1 = min(tmax, float(np.nanmin(np.maximum(ta, tb))))
    if t1 <= t0:
        return 0.0
    t = np.linspace(t0, t1, n)
    p = o[None, :] + t[:, None] * d[None, :]
    f = _trilinear(np.asarray(dens, np.float64), res, lo, hi, p)
    h = t[1] - t[0]
    integral = h / 3 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    return sigma * integral * np.linalg.norm(d)


def _transmittance_cases(krr):
    """(name, scene, o, d, tmax, expected per-channel transmittance, n)."""
    cases = []
    # homogeneous, gray, unnormalised direction (the walk runs in units of |d|)
    s = _scene(krr, [{"type": "homogeneous", "sigma_t": [0.8] * 3, "albedo": [0.5] * 3}])
    cases.append(("homogeneous", s, (0, 0, 0), (0, 0, -2.0), 1.0, np.exp(-0.8 * 2.0) * np.ones(3), 200000))
    # homogeneous, chromatic: each channel exp(-sigma_c d), hero channel uniform
    sig = np.array([0.3, 0.9, 2.0])
    s = _scene(krr, [{"type": "homogeneous", "sigma_t": sig.tolist(), "albedo": [0.5] * 3}])
    cases.append(("chromatic", s, (0, 0, 0), (1.5, 0, 0), 1.0, np.exp(-sig * 1.5), 200000))
    # constant-density grid: exp(-sigma rho len) inside index range [0, res-1]
    s = _scene(krr, [{"type": "dense_grid", "res": [8, 8, 8], "density": [0.7] * 512, "sigma_t": [2.0] * 3}])
    o, e = np.array([0.05, 0.1, 0.2]), np.array([0.8, 0.75, 0.6])
    cases.append(("constant_grid", s, tuple(o), tuple(e - o), 1.0,
                  np.exp(-2.0 * 0.7 * np.linalg.norm(e - o)) * np.ones(3), 200000))
    # heterogeneous grid, entering from outside and leaving again
    rng = np.random.default_rng(3)
    dens = rng.uniform(0.0, 1.0, 6 * 5 * 7).round(3)
    lo, hi = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
    for name, sigma, n in (("heterogeneous", 3.0, 200000), ("heterogeneous_deep", 12.0, 400000)):
        s = _scene(krr, [{"type": "dense_grid", "res": [6, 5, 7], "density": dens.tolist(), "sigma_t": [sigma] * 3,
                          "bounds": [lo, hi]}])
        o, d = (-0.2, 0.3, 0.4), (1.5, 0.3, 0.15)
        tau = _optical_depth(dens, (6, 5, 7), lo, hi, sigma, o, d, 1.0)
        cases.append((name, s, o, d, 1.0, np.exp(-tau) * np.ones(3), n))
    # a null-material surface across the ray: the loop steps through it
    s = _scene(krr, [{"type": "homogeneous", "sigma_t": [0.6] * 3, "albedo": [0.5] * 3}],
               models=[FAR_QUAD, {"type": "quad", "corners": [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]],
                                  "material": -1}])
    cases.append(("null_surface", s, (0, 0, 0), (0, 0, -2.0), 1.0, np.exp(-0.6 * 2.0) * np.ones(3), 200000))
    return cases


def pin_ratio_tracking(oracle, krr, lib=None, only=None, precision=64):
    """Worst z-score of the ratio-tracking mean against exp(-tau) over the
    cases (hero channel averaged), as (ok, detail)."""
    worst = (0.0, "")
    for name, s, o, d, tmax, expected, n in _transmittance_cases(krr):
        if only and name not in only:
            continue
        means, ses = [], []
        for ch in range(3):
            m, se = oracle.transmittance(s.desc, o, d, tmax, 0, ch, n // 3, seed=17 + ch, precision=precision, lib=lib)
            means.append(m)
            ses.append(se)
        mean = np.mean(means, 0)
        se = np.sqrt(np.sum(np.square(ses), 0)) / 3
        z = np.abs(mean - expected) / (se + 1e-12)
        print(f"ratio tracking {name}: mean {mean} expected {expected} z {z}", flush=True)
        if z.max() > worst[0]:
            worst = (float(z.max()), name)
    return worst[0] <= 5.0, f"worst z {worst[0]:.2f} ({worst[1]})"


@pytest.mark.parametrize("precision", [64, 32])
def test_ratio_tracking_matches_optical_depth(oracle, krr, precision):
    ok, detail = pin_ratio_tracking(oracle, krr, precision=precision)
    assert ok, detail


def test_ratio_tracking_roulette_runs(oracle, krr):
    """The deep case drives the estimate below 0.05, where the roulette of
    wavefront.h:127-131 runs: some runs end at 0, the survivors carry 4x."""
    case = [c for c in _transmittance_cases(krr) if c[0] == "heterogeneous_deep"][0]
    _, s, o, d, tmax, expected, _ = case
    assert expected[0] < 0.05
    m, se = oracle.transmittance(s.desc, o, d, tmax, 0, 0, 100000, seed=3)
    assert m[0] > 0 and abs(m[0] - expected[0]) <= 5 * se[0]


# ---------------------------------------------------------------------------
# Delta tracking: unbounded emissive, scattering homogeneous medium
# ---------------------------------------------------------------------------

EMISSIVE_MEDIUM = {"type": "homogeneous", "sigma_t": [1.0, 1.5, 2.5], "albedo": [0.5, 0.7, 0.9], "Le": [1.0, 2.0, 0.5],
                   "g": 0.3}


def pin_emissive_medium(oracle, krr, lib=None, depths=(1, 3, 6), precision=64, spp=64):
    """Image mean per channel against Le (1 - a^D), as (ok, detail)."""
    s = _scene(krr, [EMISSIVE_MEDIUM], res=(32, 32))
    le = np.array(EMISSIVE_MEDIUM["Le"])
    a = np.array(EMISSIVE_MEDIUM["albedo"])
    worst = (0.0, 0)
    for D in depths:
        img, st = oracle.render(s.desc, s.camera(0), 32, 32, 1, spp=spp, max_depth=D, precision=precision, lib=lib)
        v = img[..., :3].reshape(-1, 3).astype(np.float64)
        mean, se = v.mean(0), v.std(0) / np.sqrt(len(v))
        expected = le * (1 - a ** D)
        z = np.abs(mean - expected) / (se + 1e-12)
        print(f"emissive medium D={D}: mean {mean} expected {expected} z {z}", flush=True)
        if z.max() > worst[0]:
            worst = (float(z.max()), D)
    return worst[0] <= 5.0, f"worst z {worst[0]:.2f} at D={worst[1]}"


@pytest.mark.parametrize("precision", [64, 32])
def test_emissive_scattering_medium_closed_form(oracle, krr, precision):
    ok, detail = pin_emissive_medium(oracle, krr, precision=precision)
    assert ok, detail


# ---------------------------------------------------------------------------
# NEE in a heterogeneous scattering medium: on and off agree
# ---------------------------------------------------------------------------

def _nee_scene(krr):
    rng = np.random.default_rng(9)
    dens = rng.uniform(0.3, 1.0, 5 * 4 * 5).round(3).tolist()
    return krr.Scene({
        "materials": [{"bsdf": "diffuse", "diffuse": [0.6, 0.6, 0.6]}],
        "media": [{"type": "dense_grid", "res": [5, 4, 5], "density": dens, "sigma_t": [2.5] * 3,
                   "albedo": [0.85, 0.85, 0.85], "g": 0.4, "bounds": [[-1.5, -0.1, -1.5], [1.5, 1.7, 1.5]]}],
        "model": [{"type": "quad", "corners": [[-1.5, 0, -1.5], [-1.5, 0, 1.5], [1.5, 0, 1.5], [1.5, 0, -1.5]],
                   "material": 0},
                  {"type": "quad", "corners": [[-0.5, 1.2, -0.5], [0.5, 1.2, -0.5], [0.5, 1.2, 0.5], [-0.5, 1.2, 0.5]],
                   "material": 0, "emission": [4.0, 4.0, 4.0]}],
        "resolution": [24, 24],
        "camera": {"eye": [0.0, 0.7, 1.4], "target": [0.0, 0.4, 0.0], "medium": 0}})


def pin_medium_nee(oracle, krr, lib=None, spp=4096, precision=64):
    """NEE-on minus NEE-off image, both from the same seeds: the mean of the
    per-pixel differences over the image and over 4 x 4 blocks, in standard
    errors of that mean, as (ok, detail).

    max_depth is 60, where the truncated tail is negligible: at a finite max
    depth the two differ by design, because a medium interaction at the last
    depth queues no emitter hit (medium.cpp:74-75 returns at w.depth ==
    maxDepth before the hitLightRayQueue push) while next-event estimation
    from the depth before adds that same light - at max depth 5 the NEE
    image of this scene is 4 % brighter."""
    s = _nee_scene(krr)
    imgs = []
    for nee in (1, 0):
        img, _ = oracle.render(s.desc, s.camera(0), 24, 24, 1, spp=spp, max_depth=60, nee=nee, precision=precision,
                               lib=lib)
        imgs.append(img[..., 0].astype(np.float64))  # gray medium and materials: one channel
    d = imgs[0] - imgs[1]
    whole = d.mean() / (d.std() / np.sqrt(d.size))
    blocks = d.reshape(4, 6, 4, 6).transpose(0, 2, 1, 3).reshape(16, 36)
    zb = np.abs(blocks.mean(1) / (blocks.std(1) / 6.0 + 1e-12))
    print(f"medium NEE on/off: image means {imgs[0].mean():.5f} / {imgs[1].mean():.5f}, paired z {whole:.2f}, "
          f"block z max {zb.max():.2f}", flush=True)
    return abs(whole) <= 5.0 and zb.max() <= 5.5, f"paired z {whole:.2f}, block z max {zb.max():.2f}"


def test_medium_nee_on_off_agree(oracle, krr):
    ok, detail = pin_medium_nee(oracle, krr)
    assert ok, detail


# ---------------------------------------------------------------------------
# Each pin fails on the mistake it is meant to catch
# ---------------------------------------------------------------------------

PLANTS = [
    (1, "ratio tracking drops the null-collision ratio", lambda o, k, lib: pin_ratio_tracking(o, k, lib, only={"homogeneous", "heterogeneous"})),
    (2, "DDA steps the wrong axis's delta", lambda o, k, lib: pin_dda(o, k, lib, n_rays=12)),
    (3, "ratio-tracking roulette divides by 0.75", lambda o, k, lib: pin_ratio_tracking(o, k, lib, only={"heterogeneous_deep"})),
    (4, "medium emission weighted by sigma_s", lambda o, k, lib: pin_emissive_medium(o, k, lib, depths=(3,), spp=16)),
    (5, "light sample's MIS ratio drops the phase pdf", lambda o, k, lib: pin_medium_nee(o, k, lib)),
    (6, "null collision scales the light-path ratio by sigma_n", lambda o, k, lib: pin_medium_nee(o, k, lib)),
]


@pytest.mark.parametrize("plant,what,pin", PLANTS, ids=[str(p[0]) for p in PLANTS])
def test_pin_fails_on_planted_mistake(oracle, krr, plant, what, pin):
    lib = oracle.load_planted(plant)
    ok, detail = pin(oracle, krr, lib)
    print(f"plant {plant} ({what}): {detail}", flush=True)
    assert not ok, f"the pin did not catch: {what} ({detail})"


# ---------------------------------------------------------------------------
# Medium interfaces on the camera path
# ---------------------------------------------------------------------------

def test_absorber_behind_medium_interface(oracle, krr):
    """A purely absorbing medium inside a null-material box (a medium
    interface: "bound", krrscene.cpp:90-111), the camera outside it, an
    emitter behind it: a pixel sees Le exp(-sigma_t l), l the length of its
    camera ray inside the box.  The interface routing (closest hit on a
    null-material triangle -> next ray in the medium on the far side,
    device.cu:52-65, medium.cpp:84-90) is what puts the medium on the ray."""
    sigma, le, w, h = 0.9, 2.0, 32, 32
    s = krr.Scene({"materials": [{"bsdf": "diffuse", "diffuse": [0.0, 0.0, 0.0]}],
                   "media": [{"type": "homogeneous", "sigma_t": [sigma] * 3, "albedo": [0.0] * 3,
                              "bound": [[-1, -1, -1], [1, 1, 1]]}],
                   "model": [{"type": "quad", "corners": [[-6, -6, -3], [6, -6, -3], [6, 6, -3], [-6, 6, -3]],
                              "material": 0, "emission": [le] * 3}],
                   "resolution": [w, h], "camera": {"eye": [0.3, 0.2, 3.5], "target": [0, 0, 0]}})
    cam = s.camera(0)
    img, _ = oracle.render(s.desc, cam, w, h, 1, spp=128, max_depth=4, precision=64)
    expected = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            o, d = oracle.camera_ray(cam, x, y, w, h, (0.5, 0.5))
            o, d = o.astype(np.float64), d.astype(np.float64)
            with np.errstate(divide="ignore"):
                ta, tb = (-1 - o) / d, (1 - o) / d
            t0, t1 = max(np.minimum(ta, tb).max(), 0.0), np.maximum(ta, tb).min()
            inside = max(t1 - t0, 0.0) * np.linalg.norm(d)
            hits = (-3.0 - o[2]) / d[2] > 0
            expected[h - 1 - y, x] = le * np.exp(-sigma * inside) if hits else 0.0
    v = img[..., 0].astype(np.float64)
    # per pixel a mean of 128 Bernoulli(le) samples around the pixel's footprint
    var = (expected * (le - expected)).sum() / (128 * v.size ** 2)
    print(f"interface absorber: mean {v.mean():.5f} expected {expected.mean():.5f} (5 se {5 * np.sqrt(var):.2e})")
    assert abs(v.mean() - expected.mean()) <= 5 * np.sqrt(var) + 3e-3 * expected.mean()
    # rays that miss the box see the emitter unattenuated
    clear = expected == le
    assert clear.any() and np.all(img[..., 0][clear[::1]] > 0)
