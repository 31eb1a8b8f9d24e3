"""CPU restatement of ffdas's `interpolate` (structured-grid to scattered-point interpolation).

TEST INFRASTRUCTURE ONLY: imported by tests/, smoke() and bench.py's CPU leg as the checker;
the library never calls it.

What it restates (reference files cited per function):
- the plan: the 6-tetrahedra split of every grid cell (src/interpolation/spatial_hash_kernels.cuh:147-203),
  the bounding box and its 128-bin quantisation with round-to-nearest and saturation at 0
  (spatial_hash.cu:31-60, spatial_hash_kernels.cuh:96-145), each tetrahedron registered in
  every bin of its quantised bounding box (spatial_hash_kernels.cuh:206-264);
- point location: a query outside the 128^3 bin range gets no tetrahedron; otherwise the
  candidates of its bin are tested in ascending tetrahedron id with the FP32 barycentric test
  (spatial_hash_kernels.cuh:62-90, 267-320; the reference sorts each bin's list stably by id
  and returns the first that passes, i.e. the lowest passing id); nearest mode takes the
  vertex of largest barycentric weight with the reference's tie order (:51-59);
- evaluation: fill for unlocated points, else the vertex value (nearest) or
  w0 v0 + w1 v1 + w2 v2 + w3 v3 (linear; FP32 products, FP64 for double data)
  (interpolation_kernels.cuh:56-116), batched over a leading frame axis.

Pinning (tests/test_interpolation.py): the reference's own tests
(bindings/python/tests/test_interpolation.py: exact at grid points, the midpoint, fill outside,
batch, output shape); linear mode reproduces affine functions exactly up to rounding (a closed
form independent of which tetrahedron is chosen); nearest mode on a regular grid returns the
nearest grid vertex away from ties; and the binned search equals a brute-force search over all
tetrahedra on small curvilinear grids.
"""
import numpy as np

NBINS = 128

# corners of a cell as (dx, dy, dz) bit codes: bit 0 -> +x, bit 1 -> +y, bit 2 -> +z
# (the reference's v000, v100, v010, v110, v001, v101, v011, v111)
TETS = np.array([[0, 1, 2, 4], [1, 3, 2, 7], [1, 2, 4, 7], [2, 4, 6, 7], [1, 4, 5, 7],
                 [4, 6, 5, 7]], dtype=np.int64)


def tet_vertices(tid, nx, ny):
    """Vertex ids (..., 4) of tetrahedron ids tid (cell * 6 + t)."""
    tid = np.asarray(tid, dtype=np.int64)
    cell, t = tid // 6, tid % 6
    cx = cell % (nx - 1)
    cy = (cell // (nx - 1)) % (ny - 1)
    cz = cell // ((nx - 1) * (ny - 1))
    base = cx + cy * nx + cz * ny * nx
    codes = TETS[t]
    return (base[..., None] + (codes & 1) + ((codes >> 1) & 1) * nx
            + ((codes >> 2) & 1) * ny * nx)


def _quantize(p, mn, scale):
    """round-to-nearest of (p - min) * scale in FP32, negatives saturated to 0, as int."""
    q = (p.astype(np.float32) - mn) * scale
    q = np.where(np.isnan(q), np.float32(np.inf), q)
    return np.rint(np.maximum(q, np.float32(0)))


class Plan:
    """The point-location structure of a (nz, ny, nx, 3) float32 grid."""

    def __init__(self, gridpos):
        g = np.ascontiguousarray(gridpos, dtype=np.float32)
        self.nz, self.ny, self.nx = g.shape[:3]
        self.v = g.reshape(-1, 3)
        self.mn = self.v.min(axis=0)
        mx = self.v.max(axis=0)
        self.scale = np.float32(NBINS - 1) / (mx - self.mn)
        ncell = max(self.nz - 1, 0) * max(self.ny - 1, 0) * max(self.nx - 1, 0)
        self.ntet = 6 * ncell
        if self.ntet:
            qv = _quantize(self.v, self.mn, self.scale).astype(np.int64)
            vid = tet_vertices(np.arange(self.ntet), self.nx, self.ny)
            qt = qv[vid]                                     # (ntet, 4, 3)
            self.lo = np.clip(qt.min(axis=1), 0, NBINS - 1)
            self.hi = np.clip(qt.max(axis=1), 0, NBINS - 1)
            self.vid = vid

    def bin_of(self, p):
        """Quantised bin (3,) of a point, or None outside the bin range."""
        q = _quantize(np.asarray(p, np.float32)[None], self.mn, self.scale)[0]
        if np.any(q > NBINS - 1):
            return None
        return q.astype(np.int64)

    def candidates(self, b):
        if not self.ntet:
            return np.zeros(0, np.int64)
        inside = np.all((self.lo <= b) & (b <= self.hi), axis=1)
        return np.nonzero(inside)[0]

    def barycentric(self, p, tids):
        """FP32 barycentric coordinates (n, 4) and pass flags (n,) of p in tetrahedra tids,
        in the reference's operation order."""
        f = np.float32
        vv = self.v[self.vid[tids]]                          # (n, 4, 3)
        v0, v1, v2, v3 = vv[:, 0], vv[:, 1], vv[:, 2], vv[:, 3]
        p = np.asarray(p, f)
        v10, v20, v30, vp = v1 - v0, v2 - v0, v3 - v0, p - v0

        def cross(a, b):
            return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                             a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                             a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)

        def dot(a, b):
            return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]

        with np.errstate(divide="ignore", invalid="ignore"):
            denom = dot(v10, cross(v20, v30))
            by = dot(vp, cross(v20, v30)) / denom
            bz = dot(v10, cross(vp, v30)) / denom
            bw = dot(v10, cross(v20, vp)) / denom
            bx = f(1) - by - bz - bw
            ok = (np.abs(denom) >= f(1e-12)) & (bx >= 0) & (by >= 0) & (bz >= 0) & (bw >= 0)
        return np.stack([bx, by, bz, bw], axis=1), ok

    def locate(self, p, brute=False):
        """(tetrahedron id or -1, vertex ids (4,), barycentric (4,)) of one point: the lowest
        id that passes among the candidates of its bin (brute: among all tetrahedra)."""
        if brute:
            if self.bin_of(p) is None or not self.ntet:
                return -1, None, None
            cand = np.arange(self.ntet)
        else:
            b = self.bin_of(p)
            if b is None:
                return -1, None, None
            cand = self.candidates(b)
        if cand.size == 0:
            return -1, None, None
        bary, ok = self.barycentric(p, cand)
        hit = np.nonzero(ok)[0]
        if hit.size == 0:
            return -1, None, None
        i = hit[0]
        return int(cand[i]), self.vid[cand[i]], bary[i]


def nearest_of(vids, bary):
    """The reference's float4_max_by_key: strict > within (x, y) and (z, w), then between."""
    ixy = vids[0] if bary[0] > bary[1] else vids[1]
    mxy = bary[0] if ixy == vids[0] else bary[1]
    izw = vids[2] if bary[2] > bary[3] else vids[3]
    mzw = bary[2] if izw == vids[2] else bary[3]
    return int(ixy if mxy > mzw else izw)


def lookup(plan, querypos, mode):
    """Per-query lookup arrays: nearest -> idx (n,), linear -> (idx (n, 4), w (n, 4))."""
    q = np.asarray(querypos, np.float32).reshape(-1, 3)
    n = q.shape[0]
    if mode == "nearest":
        idx = np.full(n, -1, np.int64)
        for i in range(n):
            t, vids, bary = plan.locate(q[i])
            if t >= 0:
                idx[i] = nearest_of(vids, bary)
        return idx
    idx = np.full((n, 4), -1, np.int64)
    w = np.zeros((n, 4), np.float32)
    for i in range(n):
        t, vids, bary = plan.locate(q[i])
        if t >= 0:
            idx[i], w[i] = vids, bary
    return idx, w


def interpolate(gridpos, x, querypos, mode="linear", fill=0.0):
    """The reference's Interpolator(gridpos, mode)(x, querypos, fill=fill), on the CPU."""
    plan = Plan(gridpos)
    x = np.asarray(x)
    batched = x.ndim == 4
    xs = x.reshape((x.shape[0] if batched else 1), -1)
    qshape = np.asarray(querypos).shape[:-1]
    out = np.empty((xs.shape[0], int(np.prod(qshape, dtype=np.int64))), dtype=x.dtype)
    if mode == "nearest":
        idx = lookup(plan, querypos, mode)
        for b in range(xs.shape[0]):
            out[b] = np.where(idx >= 0, xs[b][np.maximum(idx, 0)], fill)
    else:
        idx, w = lookup(plan, querypos, mode)
        valid = idx[:, 0] >= 0
        wi = w.astype(np.float64) if x.real.dtype == np.float64 else w
        for b in range(xs.shape[0]):
            v = xs[b][np.maximum(idx, 0)]                     # (n, 4)
            acc = wi[:, 0] * v[:, 0] + wi[:, 1] * v[:, 1] + wi[:, 2] * v[:, 2] + wi[:, 3] * v[:, 3]
            out[b] = np.where(valid, acc.astype(x.dtype), fill)
    shape = ((xs.shape[0],) if batched else ()) + tuple(qshape)
    return out.reshape(shape)
