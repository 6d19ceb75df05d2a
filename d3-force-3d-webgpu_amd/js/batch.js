// Many layouts on one GPU, without a simulation object each (no counterpart in the
// reference, which runs one simulation per engine):
//
//   createLayoutBatch(options)   jobs resident together, ticked by one launch sequence
//   createLayoutServer(options)  a request loop: layout() resolves when the job's alpha
//                                falls below alphaMin; jobs that arrive while others run
//                                join them as memory frees up (continuous batching)
//
// A job is (nodes, nDim, forces, params): the nodes as forceSimulation takes them (missing
// positions get d3's phyllotaxis start), d3 force objects with gpu() (initialized here with
// the job's own random source), and d3's schedule (alpha, alphaMin, alphaDecay,
// alphaTarget, velocityDecay).  Forces run in the order link, manyBody, collide, center,
// radial, x, y, z; each job advances exactly as it would alone.
import {native} from "./native.js";
import lcg from "./lcg.js";
import {initializeNodes} from "./nodes.js";

var TYPES = {center: 1, collide: 2, link: 3, manyBody: 4, position: 5, radial: 6};
var ORDER = {link: 0, manyBody: 1, collide: 2, center: 3, radial: 4, position: 5};

function num(v) {
  return (v = +v) === v ? v : 0;
}

function float64(values, n) {
  var out = new Float64Array(n);
  if (values) for (var i = 0; i < n; ++i) out[i] = +values[i];
  return out;
}

function descriptor(force, n) {
  var d = force.gpu();
  if (!d || !TYPES[d.type]) throw new Error("forceBatch: a force without a supported gpu() descriptor");
  var out = {
    type: TYPES[d.type], id: 0, version: d.version || 0, iterations: d.iterations == null ? 1 : d.iterations,
    x: num(d.x), y: num(d.y), z: num(d.z), strength: d.strength == null ? 1 : +d.strength,
    theta2: d.theta2 == null ? 0.81 : +d.theta2, distanceMin2: d.distanceMin2 == null ? 1 : +d.distanceMin2,
    distanceMax2: d.distanceMax2 == null ? Infinity : +d.distanceMax2, axis: d.axis | 0
  };
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
}

// The job object the addon takes, from d3-style inputs.
export function prepareJob(nodes, nDim, forces, params) {
  params = params || {};
  nDim = Math.min(3, Math.max(1, Math.round(nDim || 2)));
  var random = params.randomSource || lcg(), n = nodes.length;
  initializeNodes(nodes, nDim);
  var list = Array.isArray(forces) ? forces : Object.keys(forces || {}).map(function(k) { return forces[k]; });
  list.forEach(function(f) { if (f.initialize) f.initialize(nodes, random, nDim); });
  var descs = list.map(function(f) { return descriptor(f, n); });
  var rank = descs.map(function(d) { var t = Object.keys(TYPES).find(function(k) { return TYPES[k] === d.type; });
    return ORDER[t] + (t === "position" ? d.axis : 0); });
  for (var k = 1; k < rank.length; ++k) {
    if (rank[k] <= rank[k - 1]) throw new Error("forceBatch: forces must come in the order link, manyBody, collide, center, radial, x, y, z");
  }
  var pos = new Float64Array(3 * n), vel = new Float64Array(3 * n), fix = null, fixMask = null;
  for (var i = 0; i < n; ++i) {
    var node = nodes[i];
    pos[3 * i] = +node.x, vel[3 * i] = +node.vx || 0;
    if (nDim > 1) pos[3 * i + 1] = +node.y, vel[3 * i + 1] = +node.vy || 0;
    if (nDim > 2) pos[3 * i + 2] = +node.z, vel[3 * i + 2] = +node.vz || 0;
    var m = (node.fx != null ? 1 : 0) | (nDim > 1 && node.fy != null ? 2 : 0) | (nDim > 2 && node.fz != null ? 4 : 0);
    if (m) {
      if (!fix) fix = new Float64Array(3 * n), fixMask = new Uint8Array(n);
      fixMask[i] = m;
      fix[3 * i] = +node.fx || 0, fix[3 * i + 1] = +node.fy || 0, fix[3 * i + 2] = +node.fz || 0;
    }
  }
  var alphaMin = params.alphaMin == null ? 0.001 : +params.alphaMin;
  return {
    n: n, nDim: nDim, pos: pos, vel: vel, fix: fix, fixMask: fixMask, forces: descs,
    alpha: params.alpha == null ? 1 : +params.alpha, alphaMin: alphaMin,
    alphaDecay: params.alphaDecay == null ? 1 - Math.pow(alphaMin, 1 / 300) : +params.alphaDecay,
    alphaTarget: params.alphaTarget == null ? 0 : +params.alphaTarget,
    velocityDecay: params.velocityDecay == null ? 0.4 : +params.velocityDecay,
    seed: Math.floor(random() * 4294967296) >>> 0, maxTicks: params.maxTicks | 0
  };
}

// Writes a job's final state into its nodes.
function land(nodes, nDim, pos, vel) {
  for (var i = 0, n = nodes.length; i < n; ++i) {
    var node = nodes[i];
    node.x = pos[3 * i], node.vx = vel[3 * i];
    if (nDim > 1) node.y = pos[3 * i + 1], node.vy = vel[3 * i + 1];
    if (nDim > 2) node.z = pos[3 * i + 2], node.vz = vel[3 * i + 2];
  }
}

export function createLayoutBatch(options) {
  options = options || {};
  var addon = native(), handle = addon.batchCreate(options.device || 0, options.budget || 0), jobs = new Map();
  return {
    // The job's id, or null when the HBM budget has no room for it now.
    add: function(nodes, nDim, forces, params) {
      var job = prepareJob(nodes, nDim, forces, params), id = addon.batchSubmit(handle, job);
      if (id !== null) jobs.set(id, {nodes: nodes, nDim: job.nDim});
      return id;
    },
    step: function(ticks) { addon.batchStep(handle, ticks == null ? 1 : ticks); return this; },
    autotune: function(cachePath) { return addon.batchAutotune(handle, cachePath || null); },
    jobs: function() { return addon.batchJobs(handle); },
    done: function() { return addon.batchJobs(handle).every(function(j) { return j.done; }); },
    // Writes the job's current state into its nodes; release frees it as well.
    land: function(id, release) {
      var j = jobs.get(id), n = j.nodes.length, pos = new Float64Array(3 * n), vel = new Float64Array(3 * n);
      addon.batchRead(handle, id, pos, vel);
      land(j.nodes, j.nDim, pos, vel);
      if (release) addon.batchRelease(handle, id), jobs.delete(id);
      return j.nodes;
    },
    destroy: function() { addon.batchDestroy(handle); jobs.clear(); }
  };
}

export function createLayoutServer(options) {
  options = options || {};
  var addon = native(), handle = addon.serveCreate(options.device || 0, {
    budget: options.budget || 0, ticksPerRound: options.ticksPerRound || 0
  });
  return {
    // Resolves with {nodes, ticks, submitted, admitted, finished} once the job is done (its
    // nodes hold the final positions and velocities; times in seconds on the server's clock).
    layout: function(nodes, nDim, forces, params) {
      var job = prepareJob(nodes, nDim, forces, params), n = nodes.length;
      var id = addon.serveSubmit(handle, job), pos = new Float64Array(3 * n), vel = new Float64Array(3 * n);
      var timeout = params && params.timeout || 600;
      return addon.serveWaitAsync(handle, id, timeout, pos, vel).then(function(info) {
        land(nodes, job.nDim, pos, vel);
        info.nodes = nodes;
        return info;
      });
    },
    stats: function() { return addon.serveStats(handle); },
    destroy: function() { addon.serveDestroy(handle); }
  };
}
