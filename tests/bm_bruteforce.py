"""Pure-Python restatement of the reference block matcher for tiny volumes.

Used only to pin the C++ oracle. Written from the reference's semantics
(itkBlockMatcher.txx:46-101 metric, :631-848 scan, :850-944 filter) as
directly as possible, with explicit loops in ITK raster order.
"""
import numpy as np


def ncc(base, bmask, search, smask, b, s, bw):
    """NCC^2 with means over all voxels and masked sums (double precision)."""
    n = 0
    mb = 0.0
    ms = 0.0
    for z in range(bw):
        for y in range(bw):
            for x in range(bw):
                mb += float(base[b[2] + z, b[1] + y, b[0] + x])
                ms += float(search[s[2] + z, s[1] + y, s[0] + x])
                n += 1
    mb /= n
    ms /= n
    p = q = r = 0.0
    for z in range(bw):
        for y in range(bw):
            for x in range(bw):
                on = True
                if bmask is not None and bmask[b[2] + z, b[1] + y, b[0] + x] == 0:
                    on = False
                if smask is not None and smask[s[2] + z, s[1] + y, s[0] + x] == 0:
                    on = False
                if not on:
                    continue
                u = float(base[b[2] + z, b[1] + y, b[0] + x]) - mb
                v = float(search[s[2] + z, s[1] + y, s[0] + x]) - ms
                p += u * v
                q += u * u
                r += v * v
    if q <= 0.0 or r <= 0.0:
        return -2.0
    return p * p / q / r


def candidates(p, L, bw, bg, nw):
    """Candidate starts on one axis, following itkBlockMatcher.txx:744-756 and
    the sparse iterator's end rule."""
    start = p - (nw // 2) * bg
    end = start + nw * bg
    while start < 0:
        start += bg
    while end > L - bw + 1:
        end -= bg
    size = end - start
    count = 1 if size <= 0 else (size - 1) // bg + 1
    return [start + k * bg for k in range(count)]


def get_offsets(base, bmask, search, smask, bw, bg, nw, ng):
    """Per-block (block_idx, match_idx, best, second) in raster order."""
    nz, ny, nx = base.shape
    dims = (nx, ny, nz)
    nb = [(d - bw) // ng + 1 for d in dims]
    st = [(d - (ng * (n - 1) + bw)) // 2 for d, n in zip(dims, nb)]
    out = []
    for kz in range(nb[2]):
        for ky in range(nb[1]):
            for kx in range(nb[0]):
                p = (st[0] + kx * ng, st[1] + ky * ng, st[2] + kz * ng)
                cx = candidates(p[0], nx, bw, bg, nw)
                cy = candidates(p[1], ny, bw, bg, nw)
                cz = candidates(p[2], nz, bw, bg, nw)
                best = np.float32(-3.0)
                second = np.float32(-3.0)
                bidx = (cx[0], cy[0], cz[0])
                for z in cz:
                    for y in cy:
                        for x in cx:
                            v = ncc(base, bmask, search, smask, p, (x, y, z), bw)
                            if float(best) < v:
                                second = best
                                best = np.float32(v)
                                bidx = (x, y, z)
                            elif float(second) < v:
                                second = np.float32(v)
                out.append((p, bidx, best, second))
    return out
