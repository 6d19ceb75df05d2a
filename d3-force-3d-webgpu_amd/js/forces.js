// The d3-force-3d forces, with the same constructors and accessors.  Each one computes its
// per-node (or per-link) values on initialize(), like the originals, and describes itself
// with gpu() -> {type, version, ...} exactly as the reference descriptors do
// (src/manyBody.js:142-144, src/link.js:127-129, src/collide.js:134-136,
// src/center.js:49-51, src/radial.js:69-71, src/x.js:43-45); `version` changes whenever
// those values are recomputed, which tells the GPU engine to upload them again.  Called as
// d3 calls a force, force(alpha), each runs d3's CPU physics from that same description
// (cpu.js), so they work in any d3 force simulation as well.
import constant from "./constant.js";
import {applyForce} from "./cpu.js";

// d3-force-3d's initialize(nodes, ...args): the random source is the function among the
// arguments and the dimensionality the 1, 2 or 3 (src/manyBody.js:117-121).
function contextOf(args) {
  return {
    random: args.find(function(arg) { return typeof arg === "function"; }) || Math.random,
    nDim: args.find(function(arg) { return [1, 2, 3].includes(arg); }) || 2
  };
}

// A force function that applies its own description (force.gpu()) to the nodes, random
// source and dimensionality its initialize() received (kept aside, not on the function).
const contexts = new WeakMap();

function cpuForce() {
  function force(alpha) {
    var ctx = contexts.get(force);
    if (ctx) applyForce(force.gpu(), ctx.nodes, ctx.nDim, ctx.random, alpha);
  }
  return force;
}

function attach(force, nodes, args) {
  var c = contextOf(args);
  contexts.set(force, {nodes: nodes || [], random: c.random, nDim: c.nDim});
}

function findNode(nodeById, nodeId) {
  var node = nodeById.get(nodeId);
  if (!node) throw new Error("node not found: " + nodeId);
  return node;
}

export function forceManyBody() {
  var nodes, strength = constant(-30), strengths, distanceMin2 = 1, distanceMax2 = Infinity,
      theta2 = 0.81, version = 0, force = cpuForce();

  function initialize() {
    if (!nodes) return;
    var i, n = nodes.length, node;
    strengths = new Array(n);
    for (i = 0; i < n; ++i) node = nodes[i], strengths[node.index] = +strength(node, i, nodes);
    ++version;
  }

  force.initialize = function(_nodes, ...args) { nodes = _nodes; attach(force, _nodes, args); initialize(); };
  force.strength = function(_) {
    return arguments.length ? (strength = typeof _ === "function" ? _ : constant(+_), initialize(), force) : strength;
  };
  force.distanceMin = function(_) {
    return arguments.length ? (distanceMin2 = _ * _, force) : Math.sqrt(distanceMin2);
  };
  force.distanceMax = function(_) {
    return arguments.length ? (distanceMax2 = _ * _, force) : Math.sqrt(distanceMax2);
  };
  force.theta = function(_) {
    return arguments.length ? (theta2 = _ * _, force) : Math.sqrt(theta2);
  };
  force.gpu = function() {
    return {type: "manyBody", version: version, strengths: strengths, distanceMin2: distanceMin2,
            distanceMax2: distanceMax2, theta2: theta2};
  };
  return force;
}

export function forceLink(links) {
  var id = function(d) { return d.index; }, strength = defaultStrength, strengths,
      distance = constant(30), distances, nodes, count, bias, iterations = 1, version = 0,
      force = cpuForce();

  if (links == null) links = [];

  function defaultStrength(link) {
    return 1 / Math.min(count[link.source.index], count[link.target.index]);
  }

  function initialize() {
    if (!nodes) return;
    var i, n = nodes.length, m = links.length, link,
        nodeById = new Map(nodes.map(function(d, i) { return [id(d, i, nodes), d]; }));
    for (i = 0, count = new Array(n); i < m; ++i) {
      link = links[i], link.index = i;
      if (typeof link.source !== "object") link.source = findNode(nodeById, link.source);
      if (typeof link.target !== "object") link.target = findNode(nodeById, link.target);
      count[link.source.index] = (count[link.source.index] || 0) + 1;
      count[link.target.index] = (count[link.target.index] || 0) + 1;
    }
    for (i = 0, bias = new Array(m); i < m; ++i) {
      link = links[i], bias[i] = count[link.source.index] / (count[link.source.index] + count[link.target.index]);
    }
    strengths = new Array(m), initializeStrength();
    distances = new Array(m), initializeDistance();
  }

  function initializeStrength() {
    if (!nodes) return;
    for (var i = 0, n = links.length; i < n; ++i) strengths[i] = +strength(links[i], i, links);
    ++version;
  }

  function initializeDistance() {
    if (!nodes) return;
    for (var i = 0, n = links.length; i < n; ++i) distances[i] = +distance(links[i], i, links);
    ++version;
  }

  force.initialize = function(_nodes, ...args) { nodes = _nodes; attach(force, _nodes, args); initialize(); };
  force.links = function(_) { return arguments.length ? (links = _, initialize(), force) : links; };
  force.id = function(_) { return arguments.length ? (id = _, force) : id; };
  force.iterations = function(_) { return arguments.length ? (iterations = +_, force) : iterations; };
  force.strength = function(_) {
    return arguments.length ? (strength = typeof _ === "function" ? _ : constant(+_), initializeStrength(), force) : strength;
  };
  force.distance = function(_) {
    return arguments.length ? (distance = typeof _ === "function" ? _ : constant(+_), initializeDistance(), force) : distance;
  };
  force.gpu = function() {
    return {type: "link", version: version, links: links, iterations: iterations, strengths: strengths,
            distances: distances, bias: bias};
  };
  return force;
}

export function forceCollide(radius) {
  var nodes, radii, strength = 1, iterations = 1, version = 0, force = cpuForce();
  if (typeof radius !== "function") radius = constant(radius == null ? 1 : +radius);

  function initialize() {
    if (!nodes) return;
    var i, n = nodes.length, node;
    radii = new Array(n);
    for (i = 0; i < n; ++i) node = nodes[i], radii[node.index] = +radius(node, i, nodes);
    ++version;
  }

  force.initialize = function(_nodes, ...args) { nodes = _nodes; attach(force, _nodes, args); initialize(); };
  force.iterations = function(_) { return arguments.length ? (iterations = +_, force) : iterations; };
  force.strength = function(_) { return arguments.length ? (strength = +_, force) : strength; };
  force.radius = function(_) {
    return arguments.length ? (radius = typeof _ === "function" ? _ : constant(+_), initialize(), force) : radius;
  };
  force.gpu = function() {
    return {type: "collide", version: version, radii: radii, strength: strength, iterations: iterations};
  };
  return force;
}

export function forceCenter(x, y, z) {
  var strength = 1, force = cpuForce();
  if (x == null) x = 0;
  if (y == null) y = 0;
  if (z == null) z = 0;
  force.initialize = function(_nodes, ...args) { attach(force, _nodes, args); };
  force.x = function(_) { return arguments.length ? (x = +_, force) : x; };
  force.y = function(_) { return arguments.length ? (y = +_, force) : y; };
  force.z = function(_) { return arguments.length ? (z = +_, force) : z; };
  force.strength = function(_) { return arguments.length ? (strength = +_, force) : strength; };
  force.gpu = function() {
    return {type: "center", version: 0, x: x, y: y, z: z, strength: strength};
  };
  return force;
}

export function forceRadial(radius, x, y, z) {
  var nodes, strength = constant(0.1), strengths, radiuses, version = 0, force = cpuForce();
  if (typeof radius !== "function") radius = constant(+radius);
  if (x == null) x = 0;
  if (y == null) y = 0;
  if (z == null) z = 0;

  function initialize() {
    if (!nodes) return;
    var i, n = nodes.length;
    strengths = new Array(n);
    radiuses = new Array(n);
    for (i = 0; i < n; ++i) {
      radiuses[i] = +radius(nodes[i], i, nodes);
      strengths[i] = isNaN(radiuses[i]) ? 0 : +strength(nodes[i], i, nodes);
    }
    ++version;
  }

  force.initialize = function(_nodes, ...args) { nodes = _nodes; attach(force, _nodes, args); initialize(); };
  force.strength = function(_) {
    return arguments.length ? (strength = typeof _ === "function" ? _ : constant(+_), initialize(), force) : strength;
  };
  force.radius = function(_) {
    return arguments.length ? (radius = typeof _ === "function" ? _ : constant(+_), initialize(), force) : radius;
  };
  force.x = function(_) { return arguments.length ? (x = +_, force) : x; };
  force.y = function(_) { return arguments.length ? (y = +_, force) : y; };
  force.z = function(_) { return arguments.length ? (z = +_, force) : z; };
  force.gpu = function() {
    return {type: "radial", version: version, radiuses: radiuses, strengths: strengths, x: x, y: y, z: z};
  };
  return force;
}

function positionForce(axis, name) {
  return function(target) {
    var strength = constant(0.1), nodes, strengths, targets, version = 0, force = cpuForce();
    if (typeof target !== "function") target = constant(target == null ? 0 : +target);

    function initialize() {
      if (!nodes) return;
      var i, n = nodes.length;
      strengths = new Array(n);
      targets = new Array(n);
      for (i = 0; i < n; ++i) {
        strengths[i] = isNaN(targets[i] = +target(nodes[i], i, nodes)) ? 0 : +strength(nodes[i], i, nodes);
      }
      ++version;
    }

    force.initialize = function(_nodes, ...args) { nodes = _nodes; attach(force, _nodes, args); initialize(); };
    force.strength = function(_) {
      return arguments.length ? (strength = typeof _ === "function" ? _ : constant(+_), initialize(), force) : strength;
    };
    force["xyz"[axis]] = function(_) {
      return arguments.length ? (target = typeof _ === "function" ? _ : constant(+_), initialize(), force) : target;
    };
    force.gpu = function() {
      return {type: "position", axis: axis, version: version, targets: targets, strengths: strengths};
    };
    return force;
  };
}

export var forceX = positionForce(0, "forceX");
export var forceY = positionForce(1, "forceY");
export var forceZ = positionForce(2, "forceZ");
