// The JS side of the engine: mirrors a node array in the HIP backend through the N-API
// addon, with the reference engine's behaviour (src/gpu/engine.js):
//
//   * the GPU holds the truth for positions and velocities while the simulation runs;
//   * nodes are written back asynchronously (readback), one readback at a time;
//   * edits made on the JS side in the meantime -- x/y/z moved, fx/fy/fz pinned -- are found
//     by comparing against what was last exchanged and uploaded before the next step
//     (syncFromJS, engine.js:171-194); a node the user has touched is never overwritten by
//     a readback that was already in flight (land, engine.js:198-216);
//   * force descriptors are converted once per version: arrays cross to the backend only
//     when a force's `version` changes (engine.js:357, passes/*.update).
import {native} from "./native.js";

var TYPES = {center: 1, collide: 2, link: 3, manyBody: 4, position: 5, radial: 6};

function num(v) {
  return (v = +v) === v ? v : 0; // undefined / NaN -> 0
}

function float64(values, n) {
  var out = new Float64Array(n);
  if (values) for (var i = 0; i < n; ++i) out[i] = +values[i];
  return out;
}

export function supported(desc) {
  return !!(desc && TYPES[desc.type]);
}

export default function createEngine(device) {
  var addon = native(),
      handle = addon.create(device || 0),
      nodes = [], n = 0, nDim = 2, alive = true, seed = 0,
      enqueued = 0, landed = 0, epoch = 0, stale = false, inflight = null,
      shadow = new Float64Array(0), pinned = new Uint8Array(0), pinnedAt = new Float64Array(0),
      staging = null, forceIds = new Map(), nextId = 1, known = new Map();

  function idOf(force) {
    var id = forceIds.get(force);
    if (!id) forceIds.set(force, id = nextId++);
    return id;
  }

  // --- JS -> GPU ---------------------------------------------------------------------

  function uploadAll() {
    var pos = new Float64Array(3 * n), vel = new Float64Array(3 * n),
        fix = new Float64Array(3 * n), mask = new Uint8Array(n), i, j, node, m;
    for (i = 0; i < n; ++i) {
      node = nodes[i], j = 3 * i;
      shadow[j] = pos[j] = num(node.x);
      shadow[j + 1] = pos[j + 1] = nDim > 1 ? num(node.y) : 0;
      shadow[j + 2] = pos[j + 2] = nDim > 2 ? num(node.z) : 0;
      vel[j] = num(node.vx);
      if (nDim > 1) vel[j + 1] = num(node.vy);
      if (nDim > 2) vel[j + 2] = num(node.vz);
      m = 0;
      if (node.fx != null) m |= 1, pinnedAt[j] = fix[j] = +node.fx;
      if (nDim > 1 && node.fy != null) m |= 2, pinnedAt[j + 1] = fix[j + 1] = +node.fy;
      if (nDim > 2 && node.fz != null) m |= 4, pinnedAt[j + 2] = fix[j + 2] = +node.fz;
      pinned[i] = mask[i] = m;
    }
    addon.setNodes(handle, n, nDim, pos, vel, fix, mask);
    known = new Map(); // fresh node buffers: every force's arrays go up again
    ++epoch;
    stale = false;
    landed = enqueued;
  }

  function uploadPositions(a, b) {
    var out = new Float64Array(3 * (b - a)), i, j, k, node;
    for (i = a; i < b; ++i) {
      node = nodes[i], j = 3 * i, k = 3 * (i - a);
      shadow[j] = out[k] = num(node.x);
      shadow[j + 1] = out[k + 1] = nDim > 1 ? num(node.y) : 0;
      shadow[j + 2] = out[k + 2] = nDim > 2 ? num(node.z) : 0;
    }
    addon.updatePositions(handle, a, out);
  }

  function uploadPinned(a, b) {
    var fix = new Float64Array(3 * (b - a)), mask = new Uint8Array(b - a), i, j, k, node, m;
    for (i = a; i < b; ++i) {
      node = nodes[i], j = 3 * i, k = 3 * (i - a), m = 0;
      if (node.fx != null) m |= 1, pinnedAt[j] = fix[k] = +node.fx;
      if (nDim > 1 && node.fy != null) m |= 2, pinnedAt[j + 1] = fix[k + 1] = +node.fy;
      if (nDim > 2 && node.fz != null) m |= 4, pinnedAt[j + 2] = fix[k + 2] = +node.fz;
      pinned[i] = mask[i - a] = m;
    }
    addon.updatePins(handle, a, fix, mask);
  }

  function syncFromJS() {
    var posRun = -1, pinRun = -1, i, j, node, x, y, z, m, posDirty, pinDirty;
    for (i = 0; i <= n; ++i) {
      posDirty = pinDirty = false;
      if (i < n) {
        node = nodes[i], j = 3 * i;
        x = num(node.x), y = nDim > 1 ? num(node.y) : 0, z = nDim > 2 ? num(node.z) : 0;
        posDirty = x !== shadow[j] || y !== shadow[j + 1] || z !== shadow[j + 2];
        m = (node.fx != null ? 1 : 0) | (nDim > 1 && node.fy != null ? 2 : 0) | (nDim > 2 && node.fz != null ? 4 : 0);
        pinDirty = m !== pinned[i]
            || (m & 1) !== 0 && +node.fx !== pinnedAt[j]
            || (m & 2) !== 0 && +node.fy !== pinnedAt[j + 1]
            || (m & 4) !== 0 && +node.fz !== pinnedAt[j + 2];
      }
      if (posDirty) { if (posRun < 0) posRun = i; }
      else if (posRun >= 0) uploadPositions(posRun, i), posRun = -1;
      if (pinDirty) { if (pinRun < 0) pinRun = i; }
      else if (pinRun >= 0) uploadPinned(pinRun, i), pinRun = -1;
    }
  }

  // Descriptors in the backend's form; arrays only for a force whose version is new.
  function describe(forces) {
    return forces.map(function(force) {
      var d = force.gpu(), id = idOf(force), fresh = known.get(id) !== d.version, out = {
        type: TYPES[d.type], id: id, version: d.version || 0, iterations: d.iterations == null ? 1 : d.iterations,
        x: num(d.x), y: num(d.y), z: num(d.z), strength: d.strength == null ? 1 : +d.strength,
        theta2: d.theta2 == null ? 0.81 : +d.theta2, distanceMin2: d.distanceMin2 == null ? 1 : +d.distanceMin2,
        distanceMax2: d.distanceMax2 == null ? Infinity : +d.distanceMax2, axis: d.axis | 0
      };
      known.set(id, d.version);
      if (!fresh) return out;
      if (d.type === "manyBody") out.strengths = float64(d.strengths, n);
      else if (d.type === "collide") out.radii = float64(d.radii, n);
      else if (d.type === "radial") out.radii = float64(d.radiuses, n), out.strengths = float64(d.strengths, n);
      else if (d.type === "position") out.targets = float64(d.targets, n), out.strengths = float64(d.strengths, n);
      else if (d.type === "link") {
        var m = d.links.length, links = new Uint32Array(2 * m);
        for (var k = 0; k < m; ++k) links[2 * k] = d.links[k].source.index, links[2 * k + 1] = d.links[k].target.index;
        out.links = links;
        out.strengths = float64(d.strengths, m);
        out.distances = float64(d.distances, m);
        out.bias = float64(d.bias, m);
      }
      return out;
    });
  }

  // --- GPU -> JS ---------------------------------------------------------------------

  function land(pos, vel) {
    var i, j, node, x, y, z;
    for (i = 0; i < n; ++i) {
      node = nodes[i], j = 3 * i;
      x = num(node.x), y = nDim > 1 ? num(node.y) : 0, z = nDim > 2 ? num(node.z) : 0;
      if (x === shadow[j] && y === shadow[j + 1] && z === shadow[j + 2]) {
        node.x = shadow[j] = pos[j];
        if (nDim > 1) node.y = shadow[j + 1] = pos[j + 1];
        if (nDim > 2) node.z = shadow[j + 2] = pos[j + 2];
      }
      node.vx = vel[j];
      if (nDim > 1) node.vy = vel[j + 1];
      if (nDim > 2) node.vz = vel[j + 2];
    }
  }

  function readback() {
    if (inflight) return inflight;
    if (!alive || !n || landed === enqueued) return Promise.resolve();
    var target = enqueued, era = epoch;
    if (!staging || staging.n !== n) staging = {n: n, pos: new Float64Array(3 * n), vel: new Float64Array(3 * n)};
    var s = staging;
    return inflight = addon.readbackAsync(handle, s.pos, s.vel).then(function(ok) {
      inflight = null;
      if (ok && era === epoch && alive) land(s.pos, s.vel), landed = target;
    }, function(error) {
      inflight = null;
      if (era === epoch && alive) throw error;
    });
  }

  var engine = {
    get n() { return n; },
    get nDim() { return nDim; },

    setNodes: function(_nodes, _nDim) {
      nodes = _nodes, n = nodes.length, nDim = _nDim;
      enqueued = landed = 0;
      shadow = new Float64Array(3 * n);
      pinned = new Uint8Array(n);
      pinnedAt = new Float64Array(3 * n);
      uploadAll();
      return engine;
    },

    // Build the pass list from the forces (in application order). False when a force has
    // no GPU description (the caller then reports it).
    setForces: function(forces) {
      var ok = forces.every(function(f) { return typeof f.gpu === "function" && supported(f.gpu()); });
      if (!ok) return false;
      this.forces = forces;
      addon.setForces(handle, describe(forces));
      return true;
    },

    fits: function(count) {
      return addon.fits(handle, count, 3);
    },

    setSeed: function(random) {
      seed = Math.floor(random() * 4294967296) >>> 0;
      addon.setSeed(handle, seed);
    },

    markStale: function() {
      stale = true;
    },

    // Submit alphas.length steps; decay is the velocity multiplier (1 - velocityDecay).
    enqueue: function(alphas, decay) {
      if (!alive || !n) return;
      if (stale) uploadAll();
      else syncFromJS();
      if (this.forces) addon.setForces(handle, describe(this.forces));
      addon.enqueue(handle, Float64Array.from(alphas), decay);
      enqueued += alphas.length;
    },

    busy: function() {
      return alive && addon.busy(handle);
    },

    readback: readback,

    idle: function() {
      return alive ? addon.idleAsync(handle).then(function() {}) : Promise.resolve();
    },

    sync: function() {
      function next() {
        return readback().then(function() {
          if (alive && n && landed !== enqueued) return next();
        });
      }
      return next();
    },

    buffers: function() {
      return alive && n ? addon.buffers(handle) : null;
    },

    stats: function() {
      return addon.stats(handle);
    },

    destroy: function() {
      if (!alive) return;
      alive = false;
      addon.destroy(handle);
      nodes = [], shadow = pinnedAt = new Float64Array(0), pinned = new Uint8Array(0), staging = null;
    }
  };
  return engine;
}
