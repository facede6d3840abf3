"""Per-zone agreement checks shared by the GPU parity tests.

A device array a agrees with the oracle's b when, on every zone and variable v,

    |a - b| <= rtol |b| + atol_v,   atol_v = floor * max|b_v|,

with rtol = 1e-12 (the specification's relative tolerance, now per zone) and
floor = 1e-14: an absolute floor of a few dozen ulps of the variable's largest
magnitude, there only for components that pass through zero (velocities,
fields, fluxes, conserved momenta), where no relative bound is meaningful.
Variables listed in `positive` (the density and internal energy of a
primitive array) get no floor at all, so atmosphere zones 1e-7 below the
torus's peak are held to 1e-12 of their own value.
"""
import numpy as np

RTOL = 1e-12
FLOOR = 1e-14
PRIM_POSITIVE = (0, 1)


def zone_ratio(a, b, rtol=RTOL, floor=FLOOR, positive=()):
    """max over zones and variables of |a - b| / (rtol |b| + atol_v); the
    arrays agree when this is at most 1.  The last axis is the variable."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, np.shape(a)[-1])
    b = np.asarray(b, dtype=np.float64).reshape(-1, np.shape(b)[-1])
    worst = 0.0
    for v in range(a.shape[1]):
        av, bv = a[:, v], b[:, v]
        if not (np.all(np.isfinite(av)) and np.all(np.isfinite(bv))):
            if not np.array_equal(np.isfinite(av), np.isfinite(bv)) or not np.array_equal(av[~np.isfinite(av)],
                                                                                         bv[~np.isfinite(bv)]):
                return np.inf
            ok = np.isfinite(bv)
            av, bv = av[ok], bv[ok]
        if av.size == 0:
            continue
        atol = 0.0 if v in positive else floor * np.abs(bv).max()
        den = rtol * np.abs(bv) + atol
        diff = np.abs(av - bv)
        same = diff == 0
        if np.any(~same & (den == 0)):
            return np.inf
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(same, 0.0, diff / np.where(den > 0, den, 1.0))
        worst = max(worst, float(r.max()))
    return worst


def zone_rel(a, b, positive=()):
    """the largest per-zone relative difference |a - b| / |b| (with the floor
    of zone_ratio for the signed variables), for reporting"""
    return zone_ratio(a, b, rtol=1.0, floor=FLOOR, positive=positive)


def assert_zones_close(a, b, what="", rtol=RTOL, floor=FLOOR, positive=()):
    r = zone_ratio(a, b, rtol=rtol, floor=floor, positive=positive)
    print(what, "per-zone ratio", r, "(<= 1 passes)")
    assert r <= 1.0, (what, r)
    return r
