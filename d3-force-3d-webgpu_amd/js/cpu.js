// The forces on the CPU: d3-force-3d's own CPU physics in FP64, applied to the node objects
// directly from a force's gpu() descriptor.  Every force of this package runs through here
// when it is called as d3 calls a force, force(alpha) -- which is how forceSimulationGPU
// runs its first cpuTicks ticks and keeps simulating when no GPU can take the forces (the
// reference's CPU path, src/gpu/simulation.js:21-26, 72-108):
//
//   manyBody  Barnes-Hut over d3's tree (src/manyBody.js:42-115): cells valued with
//             sqrt(4 / 2^dims) per level and |value| weights, theta test, distanceMin/Max
//   link      links in order, each seeing the velocities the earlier ones changed
//             (src/link.js:34-54)
//   collide   each node against the tree of predicted positions x + vx, pairs taken once
//             by the higher index, iterations in turn (src/collide.js:47-103)
//   center, radial, x/y/z   src/center.js:16-26, src/radial.js:16-28, src/x.js:11-16
//
// The tree is d3-binarytree / -quadtree / -octree's (their addAll, cover, add, visit,
// visitAfter: an integer cube doubled to cover the points, midpoint splits until every
// leaf holds one position, coincident points chained) kept in typed arrays; cells are
// visited in d3's order, so sums come out in d3's order, and the jiggle draws on the
// simulation's random source where d3 does.

function jiggle(random) {
  return (random() - 0.5) * 1e-6;
}

// A d3 spatial tree over 1, 2 or 3 coordinates.
class SpatialTree {
  constructor(dims) {
    this.dims = dims;
    this.fan = 1 << dims;
    this.cap = 0;
    this.cells = 0;
    this.grow(64);
    this.next = new Int32Array(0);
  }

  grow(cap) {
    const copy = (Type, old, size) => {
      const a = new Type(size);
      if (old) a.set(old.subarray(0, Math.min(old.length, size)));
      return a;
    };
    this.kids = copy(Int32Array, this.kids, cap * this.fan);
    this.kids.fill(-1, this.cap * this.fan);
    this.head = copy(Int32Array, this.head, cap);  // a leaf's first point; -1: inner cell
    this.value = copy(Float64Array, this.value, cap);
    this.radius = copy(Float64Array, this.radius, cap);
    this.cx = copy(Float64Array, this.cx, cap);
    this.cy = copy(Float64Array, this.cy, cap);
    this.cz = copy(Float64Array, this.cz, cap);
    this.cap = cap;
  }

  cell(head) {
    if (this.cells === this.cap) this.grow(2 * this.cap);
    const c = this.cells++;
    this.head[c] = head;
    this.kids.fill(-1, c * this.fan, (c + 1) * this.fan);
    return c;
  }

  // addAll over n points with coordinates px/py/pz (FP64).
  build(px, py, pz, n) {
    const dims = this.dims, P = [px, py, pz];
    this.cells = 0;
    if (this.next.length < n) this.next = new Int32Array(n);
    this.next.fill(-1, 0, n);
    const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity];
    for (let i = 0; i < n; ++i) {
      for (let a = 0; a < dims; ++a) {
        const v = P[a][i];
        if (v < lo[a]) lo[a] = v;
        if (v > hi[a]) hi[a] = v;
      }
    }
    this.origin = [0, 0, 0];
    let size = 1;
    for (let a = 0; a < dims; ++a) this.origin[a] = Math.floor(lo[a]);
    for (;;) {
      let covered = true;
      for (let a = 0; a < dims; ++a) if (!(hi[a] < this.origin[a] + size)) covered = false;
      if (covered) break;
      size *= 2;
    }
    this.size = size;
    this.depth = 0;  // the deepest cell's depth (a traversal stack needs (depth + 1) (fan - 1) + 1)
    this.root = n ? this.cell(0) : -1;
    for (let i = 1; i < n; ++i) this.add(P, i);
  }

  // Entries a depth-first traversal's stack can hold at most.
  stackBound() {
    return (this.depth + 2) * this.fan;
  }

  slot(P, i, corner, half) {
    let k = 0;
    for (let a = 0; a < this.dims; ++a) {
      if (P[a][i] >= corner[a] + half) corner[a] += half, k |= 1 << a;
    }
    return k;
  }

  add(P, i) {
    const dims = this.dims, fan = this.fan, corner = this.origin.slice();
    let half = this.size, parent = -1, at = -1, c = this.root, level = 0;
    while (this.head[c] < 0) {
      half /= 2;
      ++level;
      const k = this.slot(P, i, corner, half);
      parent = c, at = parent * fan + k;
      c = this.kids[at];
      if (c < 0) {
        const leaf = this.cell(i);  // (cell() may grow the arrays: take it before writing)
        this.kids[at] = leaf;
        if (level > this.depth) this.depth = level;
        return;
      }
    }
    const j = this.head[c];
    let same = true;
    for (let a = 0; a < dims; ++a) if (P[a][i] !== P[a][j]) same = false;
    if (same) {  // d3 chains the newcomer in front
      this.next[i] = j;
      this.head[c] = i;
      return;
    }
    for (;;) {  // split until i and j part
      const s = this.cell(-1);
      if (parent < 0) this.root = s;
      else this.kids[at] = s;
      half /= 2;
      ++level;
      let kj = 0;
      for (let a = 0; a < dims; ++a) if (P[a][j] >= corner[a] + half) kj |= 1 << a;
      const ki = this.slot(P, i, corner, half);
      if (ki !== kj) {
        const leaf = this.cell(i);
        this.kids[s * fan + kj] = c;
        this.kids[s * fan + ki] = leaf;
        if (level > this.depth) this.depth = level;
        return;
      }
      parent = s, at = s * fan + ki;
    }
  }

  // Every cell after its children (d3's visitAfter order within each cell's children).
  postOrder(fn) {
    if (this.root < 0) return;
    const fan = this.fan, order = [], stack = [this.root];
    while (stack.length) {
      const c = stack.pop();
      order.push(c);
      if (this.head[c] < 0) for (let k = 0; k < fan; ++k) if (this.kids[c * fan + k] >= 0) stack.push(this.kids[c * fan + k]);
    }
    for (let q = order.length - 1; q >= 0; --q) fn(order[q]);
  }
}

// ---- forces -----------------------------------------------------------------------------

function manyBody(eng, d, alpha) {
  const {nodes, nDim: dims, random} = eng, n = nodes.length, s = d.strengths;
  const theta2 = d.theta2 == null ? 0.81 : d.theta2, dmin2 = d.distanceMin2 == null ? 1 : d.distanceMin2,
        dmax2 = d.distanceMax2 == null ? Infinity : d.distanceMax2;
  const t = eng.treeFor(dims), px = eng.px, py = eng.py, pz = eng.pz;
  for (let i = 0; i < n; ++i) px[i] = nodes[i].x, py[i] = dims > 1 ? nodes[i].y : 0, pz[i] = dims > 2 ? nodes[i].z : 0;
  t.build(px, py, pz, n);
  const fan = t.fan, damp = Math.sqrt(4 / fan), C = [t.cx, t.cy, t.cz], P = [px, py, pz];
  t.postOrder((c) => {
    if (t.head[c] >= 0) {
      let v = 0;
      for (let j = t.head[c]; j >= 0; j = t.next[j]) v += s[nodes[j].index];
      const h = t.head[c];
      for (let a = 0; a < dims; ++a) C[a][c] = P[a][h];
      t.value[c] = v;
      return;
    }
    let v = 0, w = 0, g0 = 0, g1 = 0, g2 = 0;
    for (let k = 0; k < fan; ++k) {
      const q = t.kids[c * fan + k];
      if (q < 0) continue;
      const cw = Math.abs(t.value[q]);
      if (!cw) continue;
      v += t.value[q], w += cw;
      g0 += cw * t.cx[q], g1 += cw * t.cy[q], g2 += cw * t.cz[q];
    }
    t.value[c] = v * damp;
    t.cx[c] = g0 / w, t.cy[c] = g1 / w, t.cz[c] = g2 / w;
  });
  // pre-order over the cells with typed stacks; each term is added to the node's velocity
  // in turn (a local copy of it, written back after the walk: the same roundings as d3's
  // node.vx += term)
  const head = t.head, kids = t.kids, next = t.next, val = t.value, cx = t.cx, cy = t.cy, cz = t.cz;
  const stack = eng.stack(t.stackBound()), widths = eng.widths(t.stackBound());
  for (let i = 0; i < n; ++i) {
    const node = nodes[i], xi = node.x, yi = dims > 1 ? node.y : 0, zi = dims > 2 ? node.z : 0;
    let vx = node.vx, vy = dims > 1 ? node.vy : 0, vz = dims > 2 ? node.vz : 0, sp = 0;
    stack[0] = t.root, widths[0] = t.size, sp = 1;
    while (sp > 0) {
      --sp;
      const c = stack[sp], width = widths[sp], value = val[c];
      if (!value) continue;
      let x = cx[c] - xi, y = dims > 1 ? cy[c] - yi : 0, z = dims > 2 ? cz[c] - zi : 0;
      let l = x * x + y * y + z * z;
      if (width * width / theta2 < l) {
        if (l < dmax2) {
          if (x === 0) x = jiggle(random), l += x * x;
          if (dims > 1 && y === 0) y = jiggle(random), l += y * y;
          if (dims > 2 && z === 0) z = jiggle(random), l += z * z;
          if (l < dmin2) l = Math.sqrt(dmin2 * l);
          vx += x * value * alpha / l;
          if (dims > 1) vy += y * value * alpha / l;
          if (dims > 2) vz += z * value * alpha / l;
        }
        continue;
      }
      const h = head[c];
      if (h < 0) {
        const half = width / 2;
        for (let k = c * fan + fan - 1; k >= c * fan; --k) {
          const q = kids[k];
          if (q >= 0) stack[sp] = q, widths[sp] = half, ++sp;
        }
        continue;
      }
      if (l >= dmax2) continue;
      if (h !== i || next[h] >= 0) {
        if (x === 0) x = jiggle(random), l += x * x;
        if (dims > 1 && y === 0) y = jiggle(random), l += y * y;
        if (dims > 2 && z === 0) z = jiggle(random), l += z * z;
        if (l < dmin2) l = Math.sqrt(dmin2 * l);
      }
      for (let j = h; j >= 0; j = next[j]) {
        if (j === i) continue;
        const w = s[nodes[j].index] * alpha / l;
        vx += x * w;
        if (dims > 1) vy += y * w;
        if (dims > 2) vz += z * w;
      }
    }
    node.vx = vx;
    if (dims > 1) node.vy = vy;
    if (dims > 2) node.vz = vz;
  }
}

function link(eng, d, alpha) {
  const {nDim: dims, random} = eng, links = d.links, iters = d.iterations == null ? 1 : d.iterations;
  for (let k = 0; k < iters; ++k) {
    for (let i = 0, m = links.length; i < m; ++i) {
      const src = links[i].source, dst = links[i].target;
      let x = dst.x + dst.vx - src.x - src.vx || jiggle(random);
      let y = dims > 1 ? dst.y + dst.vy - src.y - src.vy || jiggle(random) : 0;
      let z = dims > 2 ? dst.z + dst.vz - src.z - src.vz || jiggle(random) : 0;
      let l = Math.sqrt(x * x + y * y + z * z);
      l = (l - d.distances[i]) / l * alpha * d.strengths[i];
      x *= l, y *= l, z *= l;
      const b = d.bias[i];
      dst.vx -= x * b;
      src.vx += x * (1 - b);
      if (dims > 1) dst.vy -= y * b, src.vy += y * (1 - b);
      if (dims > 2) dst.vz -= z * b, src.vz += z * (1 - b);
    }
  }
}

function collide(eng, d) {
  const {nodes, nDim: dims, random} = eng, n = nodes.length, radii = d.radii;
  const strength = d.strength == null ? 1 : d.strength, iters = d.iterations == null ? 1 : d.iterations;
  const t = eng.treeFor(dims), px = eng.px, py = eng.py, pz = eng.pz, fan = t.fan;
  for (let k = 0; k < iters; ++k) {
    for (let i = 0; i < n; ++i) {
      const q = nodes[i];
      px[i] = q.x + q.vx, py[i] = dims > 1 ? q.y + q.vy : 0, pz[i] = dims > 2 ? q.z + q.vz : 0;
    }
    t.build(px, py, pz, n);
    t.postOrder((c) => {
      if (t.head[c] >= 0) {
        t.radius[c] = radii[nodes[t.head[c]].index];
        return;
      }
      let r = 0;
      for (let q = 0; q < fan; ++q) {
        const kid = t.kids[c * fan + q];
        if (kid >= 0 && t.radius[kid] > r) r = t.radius[kid];
      }
      t.radius[c] = r;
    });
    const kidsNow = t.kids, headNow = t.head, radNow = t.radius;
    const depth = t.stackBound(), stack = eng.stack(depth), widths = eng.widths(depth), corners = eng.corners(3 * depth);
    for (let i = 0; i < n; ++i) {
      const node = nodes[i], ri = radii[node.index], ri2 = ri * ri;
      const xi = node.x + node.vx, yi = dims > 1 ? node.y + node.vy : 0, zi = dims > 2 ? node.z + node.vz : 0;
      let sp = 1;
      stack[0] = t.root, widths[0] = t.size;
      corners[0] = t.origin[0], corners[1] = t.origin[1], corners[2] = t.origin[2];
      while (sp > 0) {
        --sp;
        const c = stack[sp], w = widths[sp], x0 = corners[3 * sp], y0 = corners[3 * sp + 1], z0 = corners[3 * sp + 2];
        let rj = radNow[c], r = ri + rj;
        const h = headNow[c];
        if (h >= 0) {
          const data = nodes[h];
          if (data.index > node.index) {
            let x = xi - data.x - data.vx, y = dims > 1 ? yi - data.y - data.vy : 0, z = dims > 2 ? zi - data.z - data.vz : 0;
            let l = x * x + y * y + z * z;
            if (l < r * r) {
              if (x === 0) x = jiggle(random), l += x * x;
              if (dims > 1 && y === 0) y = jiggle(random), l += y * y;
              if (dims > 2 && z === 0) z = jiggle(random), l += z * z;
              l = Math.sqrt(l);
              l = (r - l) / l * strength;
              rj *= rj;
              let share = rj / (ri2 + rj);
              x *= l, y *= l, z *= l;
              node.vx += x * share;
              if (dims > 1) node.vy += y * share;
              if (dims > 2) node.vz += z * share;
              share = 1 - share;
              data.vx -= x * share;
              if (dims > 1) data.vy -= y * share;
              if (dims > 2) data.vz -= z * share;
            }
          }
          continue;
        }
        if (x0 > xi + r || x0 + w < xi - r) continue;
        if (dims > 1 && (y0 > yi + r || y0 + w < yi - r)) continue;
        if (dims > 2 && (z0 > zi + r || z0 + w < zi - r)) continue;
        const half = w / 2;
        for (let q = fan - 1; q >= 0; --q) {
          const kid = kidsNow[c * fan + q];
          if (kid < 0) continue;
          stack[sp] = kid, widths[sp] = half;
          corners[3 * sp] = q & 1 ? x0 + half : x0;
          corners[3 * sp + 1] = q & 2 ? y0 + half : y0;
          corners[3 * sp + 2] = q & 4 ? z0 + half : z0;
          ++sp;
        }
      }
    }
  }
}

function center(eng, d) {
  const {nodes, nDim: dims} = eng, n = nodes.length, strength = d.strength == null ? 1 : d.strength;
  let sx = 0, sy = 0, sz = 0;
  for (let i = 0; i < n; ++i) {
    const q = nodes[i];
    sx += q.x || 0, sy += q.y || 0, sz += q.z || 0;
  }
  sx = (sx / n - (d.x || 0)) * strength, sy = (sy / n - (d.y || 0)) * strength, sz = (sz / n - (d.z || 0)) * strength;
  for (let i = 0; i < n; ++i) {
    const q = nodes[i];
    if (sx) q.x -= sx;
    if (sy && dims > 1) q.y -= sy;
    if (sz && dims > 2) q.z -= sz;
  }
}

function radial(eng, d, alpha) {
  const {nodes, nDim: dims} = eng, cx = d.x || 0, cy = d.y || 0, cz = d.z || 0;
  for (let i = 0, n = nodes.length; i < n; ++i) {
    const q = nodes[i];
    // every axis enters the distance, unused ones as (0 - centre) || 1e-6, as in d3
    const dx = q.x - cx || 1e-6, dy = (q.y || 0) - cy || 1e-6, dz = (q.z || 0) - cz || 1e-6;
    const r = Math.sqrt(dx * dx + dy * dy + dz * dz), k = (d.radiuses[i] - r) * d.strengths[i] * alpha / r;
    q.vx += dx * k;
    if (dims > 1) q.vy += dy * k;
    if (dims > 2) q.vz += dz * k;
  }
}

function position(eng, d, alpha) {
  const {nodes} = eng, p = "xyz"[d.axis], v = "v" + p;
  for (let i = 0, n = nodes.length; i < n; ++i) {
    const q = nodes[i];
    q[v] += (d.targets[i] - q[p]) * d.strengths[i] * alpha;
  }
}

const APPLY = {manyBody: manyBody, link: link, collide: collide, center: center, radial: radial, position: position};

// Scratch shared by every force's CPU application (JavaScript runs one at a time): a tree
// per dimensionality and the coordinates the trees are built over.
const scratch = {
  trees: [null, null, null, null],
  px: new Float64Array(0), py: new Float64Array(0), pz: new Float64Array(0),
  treeFor(dims) {
    return this.trees[dims] || (this.trees[dims] = new SpatialTree(dims));
  },
  // traversal stacks, grown to `size` entries when needed
  cellStack: new Int32Array(0), widthStack: new Float64Array(0), cornerStack: new Float64Array(0),
  stack(size) {
    if (this.cellStack.length < size) this.cellStack = new Int32Array(size);
    return this.cellStack;
  },
  widths(size) {
    if (this.widthStack.length < size) this.widthStack = new Float64Array(size);
    return this.widthStack;
  },
  corners(size) {
    if (this.cornerStack.length < size) this.cornerStack = new Float64Array(size);
    return this.cornerStack;
  }
};

// Applies the force `desc` (a gpu() descriptor) to `nodes` once at `alpha`, with d3's CPU
// semantics; `random` feeds the jiggle.
export function applyForce(desc, nodes, nDim, random, alpha) {
  const n = nodes.length;
  if (!n) return;
  if (scratch.px.length < n) {
    scratch.px = new Float64Array(n);
    scratch.py = new Float64Array(n);
    scratch.pz = new Float64Array(n);
  }
  const ctx = {nodes: nodes, nDim: nDim, random: random || Math.random, px: scratch.px, py: scratch.py, pz: scratch.pz,
               treeFor: (dims) => scratch.treeFor(dims), stack: (k) => scratch.stack(k),
               widths: (k) => scratch.widths(k), corners: (k) => scratch.corners(k)};
  APPLY[desc.type](ctx, desc, alpha);
}

// The integration of one tick: v *= decay, x += v, or x = fx, v = 0 on pinned axes
// (src/simulation.js:61-73).
export function integrate(nodes, nDim, decay) {
  for (let i = 0, n = nodes.length; i < n; ++i) {
    const node = nodes[i];
    if (node.fx == null) node.x += node.vx *= decay;
    else node.x = node.fx, node.vx = 0;
    if (nDim > 1) {
      if (node.fy == null) node.y += node.vy *= decay;
      else node.y = node.fy, node.vy = 0;
    }
    if (nDim > 2) {
      if (node.fz == null) node.z += node.vz *= decay;
      else node.z = node.fz, node.vz = 0;
    }
  }
}
