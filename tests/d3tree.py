"""An independent, pure-Python statement of d3-force-3d's Barnes-Hut many-body force, written
to check the C oracle (TEST INFRASTRUCTURE; small inputs only).

It follows the published algorithms of the three tree modules that src/manyBody.js:1-3 imports
(d3-binarytree@1.0.1, d3-quadtree@3.0.1, d3-octree@1.0.1, pinned in the reference's
yarn.lock:310-328), which the reference does not vendor:

  * addAll: the extent of all points, cover(min) then cover(max); cover() starts from an
    integer cube [floor(v), floor(v) + 1) and doubles it towards the point;
  * add: descend by `v >= midpoint` per axis (child index bit a for axis a), then split a
    leaf holding another position until the two part; a point at exactly the position of
    a leaf joins it as the head of its `next` chain;
  * visitAfter: post-order over the cells; visit: pre-order, children 0 .. 2^d - 1;

and the per-cell arithmetic of src/manyBody.js:42-115 (accumulate with |value| weights and
sqrt(4 / numChildren) per level, apply with the theta test, distanceMin/Max and the leaf's
direct sum).  The coincident-point jiggle (`random`) is not modelled: a displacement of
exactly 0 on an axis is returned as None, so callers compare only other nodes.

Here the tree is kept as nested lists, recursion instead of an explicit stack: the shape
and order of the arithmetic are d3's, the code is not.
"""
import math


class Cell:
    __slots__ = ("kids", "data", "x", "value")

    def __init__(self, dims):
        self.kids = [None] * (1 << dims)
        self.data = None     # a leaf: list of node indices at one position (d3's chain)
        self.x = None        # centre of charge, per axis
        self.value = 0.0


def _build(pos, dims):
    n = len(pos)
    lo = [min(p[a] for p in pos) for a in range(dims)]
    hi = [max(p[a] for p in pos) for a in range(dims)]
    x0 = [math.floor(v) for v in lo]
    size = 1.0
    while any(not (hi[a] < x0[a] + size) for a in range(dims)):
        size *= 2.0
    root = None
    for i in range(n):
        p = pos[i]
        if root is None:
            root = Cell(dims)
            root.data = [i]
            continue
        parent, slot, node = None, None, root
        c0, w = list(x0), size
        while node.data is None:
            w /= 2.0
            k = 0
            for a in range(dims):
                if p[a] >= c0[a] + w:
                    c0[a] += w
                    k |= 1 << a
            parent, slot = node, k
            node = node.kids[k]
            if node is None:
                leaf = Cell(dims)
                leaf.data = [i]
                parent.kids[k] = leaf
                break
        else:
            q = pos[node.data[0]]
            if all(p[a] == q[a] for a in range(dims)):
                node.data.insert(0, i)  # d3: the new point heads the chain
                continue
            old = node
            leaf = Cell(dims)
            leaf.data = [i]
            while True:
                split = Cell(dims)
                if parent is None:
                    root = split
                else:
                    parent.kids[slot] = split
                w /= 2.0
                ki = kj = 0
                for a in range(dims):
                    m = c0[a] + w
                    if p[a] >= m:
                        ki |= 1 << a
                    if q[a] >= m:
                        kj |= 1 << a
                for a in range(dims):
                    if p[a] >= c0[a] + w:
                        c0[a] += w
                if ki != kj:
                    split.kids[ki] = leaf
                    split.kids[kj] = old
                    break
                parent, slot = split, ki
    return root, x0, size


def _accumulate(cell, pos, strengths, dims):
    if cell.data is not None:
        cell.x = list(pos[cell.data[0]][:dims])
        cell.value = sum(strengths[i] for i in cell.data)
        return
    s = wt = 0.0
    g = [0.0] * dims
    for k in cell.kids:
        if k is None:
            continue
        _accumulate(k, pos, strengths, dims)
        c = abs(k.value)
        if c:
            s += k.value
            wt += c
            for a in range(dims):
                g[a] += c * k.x[a]
    cell.value = s * math.sqrt(4.0 / len(cell.kids))
    cell.x = [v / wt if wt else float("nan") for v in g]


def manybody(pos, strengths, dims, alpha=1.0, theta=0.9, dmin=1.0, dmax=math.inf, nodes=None):
    """Velocity increments [n][3] of d3's forceManyBody for the points `pos` ([n][3], only
    the first `dims` axes used); entries of nodes that meet a zero displacement are None."""
    n = len(pos)
    pos = [tuple(float(v) for v in p) for p in pos]
    root, x0, size = _build(pos, dims)
    _accumulate(root, pos, strengths, dims)
    theta2, dmin2, dmax2 = theta * theta, dmin * dmin, dmax * dmax
    out = []
    for i in (range(n) if nodes is None else nodes):
        me = pos[i]
        v = [0.0, 0.0, 0.0]
        ok = [True]

        def apply(cell, width):
            if not cell.value:
                return
            d = [cell.x[a] - me[a] for a in range(dims)]
            l = sum(t * t for t in d)
            if width * width / theta2 < l:
                if l < dmax2:
                    if any(t == 0 for t in d):
                        ok[0] = False
                        return
                    if l < dmin2:
                        l = math.sqrt(dmin2 * l)
                    for a in range(dims):
                        v[a] += d[a] * cell.value * alpha / l
                return
            if cell.data is None:
                for k in cell.kids:
                    if k is not None:
                        apply(k, width / 2.0)
                return
            if l >= dmax2:
                return
            if cell.data != [i] and any(t == 0 for t in d):
                ok[0] = False
                return
            if cell.data != [i] and l < dmin2:
                l = math.sqrt(dmin2 * l)
            for j in cell.data:
                if j != i:
                    w = strengths[j] * alpha / l
                    for a in range(dims):
                        v[a] += d[a] * w

        apply(root, size)
        out.append(v if ok[0] else None)
    return out


def cells(pos, dims):
    """Every cell of d3's tree over `pos`, pre-order: (depth, sorted tuple of its node indices)."""
    pos = [tuple(float(v) for v in p) for p in pos]
    root, _, _ = _build(pos, dims)
    found = []

    def collect(cell, depth):
        if cell.data is not None:
            pts = sorted(cell.data)
            found.append((depth, tuple(pts)))
            return pts
        mark = len(found)
        found.append(None)
        pts = []
        for k in cell.kids:
            if k is not None:
                pts += collect(k, depth + 1)
        pts.sort()
        found[mark] = (depth, tuple(pts))
        return pts

    collect(root, 0)
    return found
