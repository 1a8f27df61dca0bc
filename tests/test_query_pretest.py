"""Pins the float pre-test of the FP64 query decisions (tri_clear_miss32 in mrg_traverse.h) from both
sides, on the CPU: a float32 restatement of the pre-test (numpy float32 arithmetic, one rounding per
operation as the device's -ffp-contract=off build does; the two FMAs of the difference of products
through long double) is run on rays aimed at, just inside and just outside the edges and corners of
triangles of many sizes, positions and origin distances, and compared with the FP64 oracle's
decision (RayQuery::intersect, intersection.cpp:108-152).

* sound: the pre-test never rejects a triangle that the FP64 test hits (a plausible slip in the
  bound, a dropped factor or the A and M terms swapped, rejects hits here);
* tight: every ray that passes outside an edge by more than 2^-12 of its origin distance is rejected
  (the earlier 2^-17 M_j M_k bound fails this for small triangles seen from far away).
"""
import numpy as np

from minirender_amd.scene import concat, material

F = np.float32
LD = np.longdouble


def _fma(a, b, c):
    # a * b is exact in long double (48-bit product); one rounding to float after the sum
    return (a.astype(LD) * b.astype(LD) + c.astype(LD)).astype(F)


def _dop(a, b, c, d):
    cd = c * d
    corr = _fma(-c, d, cd)
    return _fma(a, b, -cd) + corr


def _axes(d):
    ad = np.abs(d)
    kz = np.where(ad[:, 0] > ad[:, 1], 0, 1)
    kz = np.where(ad[:, 2] > ad[np.arange(len(d)), kz], 2, kz)
    kx = np.where(kz == 2, 0, kz + 1)
    ky = np.where(kx == 2, 0, kx + 1)
    neg = d[np.arange(len(d)), kz] < 0
    kx, ky = np.where(neg, ky, kx), np.where(neg, kx, ky)
    return kx, ky, kz


def clear_miss32(tri, o, d):
    """The pre-test for one triangle (3x3 float32) and float32 rays o, d ([n][3]): True where it
    proves an FP64 miss."""
    n = len(o)
    idx = np.arange(n)
    kx, ky, kz = _axes(d)
    sx = d[idx, kx] / d[idx, kz]
    sy = d[idx, ky] / d[idx, kz]
    smax = np.maximum(np.abs(sx), np.abs(sy))
    p, m = [], []
    for v in tri:
        t = v[None, :] - o
        tx, ty, tz = t[idx, kx], t[idx, ky], t[idx, kz]
        p.append((tx - sx * tz, ty - sy * tz))
        m.append(np.maximum(np.abs(tx), np.abs(ty)) + smax * np.abs(tz))
    mlo = np.minimum(np.minimum(m[0], m[1]), m[2])
    mhi = np.maximum(np.maximum(m[0], m[1]), m[2])
    ok = (mlo >= F(2.0 ** -40)) & (mhi <= F(2.0 ** 40))
    a = [np.abs(px) + np.abs(py) for px, py in p]
    s = [F(2.0 ** -20) * mi for mi in m]

    def bound(e, j, k):
        return F(0.5) * (s[j] * a[k] + s[k] * a[j]) + (F(2.0 ** -21) * np.abs(e) + s[j] * s[k])

    e = [_dop(p[1][0], p[2][1], p[1][1], p[2][0]), _dop(p[2][0], p[0][1], p[2][1], p[0][0]),
         _dop(p[0][0], p[1][1], p[0][1], p[1][0])]
    eb = [bound(e[0], 1, 2), bound(e[1], 2, 0), bound(e[2], 0, 1)]
    neg = (e[0] < -eb[0]) | (e[1] < -eb[1]) | (e[2] < -eb[2])
    pos = (e[0] > eb[0]) | (e[1] > eb[1]) | (e[2] > eb[2])
    return ok & neg & pos


def _cases(rng, n):
    """One triangle and n rays: aim points on its edges (90%) or corners, pushed in or out of the
    triangle by a log-uniform fraction of the origin distance, seen from 1 to 10^4 triangle sizes."""
    size = 10 ** rng.uniform(-3, 1)
    centre = rng.normal(size=3) * size * 10 ** rng.uniform(-1, 4)
    while True:
        corners = rng.normal(size=(3, 3))
        if np.linalg.norm(np.cross(corners[1] - corners[0], corners[2] - corners[0])) > 0.3:
            break
    tri = (centre + size * corners).astype(F)
    v = tri.astype(np.float64)
    nrm = np.cross(v[1] - v[0], v[2] - v[0])
    nrm /= np.linalg.norm(nrm)
    i = rng.integers(0, 3, n)
    j = (i + 1) % 3
    k = (i + 2) % 3
    s = np.where(rng.random(n) < 0.9, rng.uniform(0.02, 0.98, n), 0.0)
    aim = v[i] + s[:, None] * (v[j] - v[i])
    # in-plane direction away from the opposite corner (outward across edge i-j, or out of corner i)
    edge = v[j] - v[i]
    out = np.cross(edge, nrm)
    out *= np.sign(np.einsum("ij,ij->i", out, v[i] - v[k]))[:, None]
    corner = s == 0.0
    out[corner] = (v[i] - (v[j] + v[k]) / 2)[corner]
    out /= np.linalg.norm(out, axis=1, keepdims=True)
    dist = size * 10 ** rng.uniform(0, 4, n)
    push = rng.choice([-1.0, 1.0], n) * 10 ** rng.uniform(-12, -1, n)
    target = aim + out * (push * dist)[:, None]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.where((d @ nrm)[:, None] * np.sign(d @ nrm)[:, None] < 0.2, d - (d @ nrm)[:, None] * nrm + 0.5 * nrm, d)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = target - d * dist[:, None]
    return tri, o.astype(F), d.astype(F), push, dist


def test_pretest_sound_and_tight(oracle):
    rng = np.random.default_rng(11)
    rejected_hits = 0
    loose = 0
    tight_cases = 0
    for _ in range(48):
        tri, o, d, push, dist = _cases(rng, 1024)
        miss32 = clear_miss32(tri, o, d)
        scene = concat([(tri[None], np.tile(np.array([0, 0, 1], F), (1, 3, 1)))], [material(double_sided=True)])
        rays = np.concatenate([o, d], 1).astype(np.float64)
        hit64 = oracle.OracleScene(scene).intersect(rays, 0.0, np.inf, precision=64)["prim"] == 0
        rejected_hits += int((miss32 & hit64).sum())
        # outside by more than 2^-12 of the distance from the origin, beyond the float rounding of the
        # origin itself: the pre-test has to reject it
        reach = dist + np.abs(o.astype(np.float64)).max(1)
        far_out = (push * dist > 2.0 ** -12 * reach) & ~hit64
        tight_cases += int(far_out.sum())
        loose += int((far_out & ~miss32).sum())
    assert rejected_hits == 0
    assert tight_cases > 3000
    assert loose == 0, f"{loose} of {tight_cases} clear misses not rejected"
