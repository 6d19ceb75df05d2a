// forceSimulationGPU: d3-force-3d's forceSimulation API with the physics in the HIP
// backend on an AMD GPU (src/gpu/simulation.js of the reference is the behavioural spec).
//
//   * Same methods as forceSimulation (tick, restart, stop, numDimensions, nodes, alpha,
//     alphaMin, alphaDecay, alphaTarget, velocityDecay, randomSource, force, find, on) plus
//     tickAsync, sync, isGPUEnabled, gpuReady, gpuBuffers, destroy.
//   * tick() queues steps on the GPU and returns; the nodes follow one readback later.
//     A tick() per frame is skipped while the GPU is 3 submissions behind; further calls
//     in the same synchronous run are a batch and always run (simulation.js:220-236).
//   * `await tickAsync(k)` runs k ticks and resolves once the nodes reflect them.
//   * options: {device: GPU index, readback: false to leave the nodes alone until sync()}.
//
// Unlike the reference there is no CPU physics to fall back to: without a usable GPU,
// gpuReady() resolves false and stepping throws.  A force without gpu() is an error.
import dispatch from "./dispatch.js";
import timer from "./timer.js";
import lcg from "./lcg.js";
import createEngine from "./engine.js";
import {deviceCount} from "./native.js";
import {initializeNodes as initializeNodesIn} from "./nodes.js";

var MAX_DIMENSIONS = 3;

export function isBackendAvailable() {
  try {
    return deviceCount() > 0;
  } catch (error) {
    return false;
  }
}

export default function(nodes, numDimensions, options) {
  numDimensions = numDimensions || 2;
  options = options || {};
  var readback = options.readback !== false;

  var nDim = Math.min(MAX_DIMENSIONS, Math.max(1, Math.round(numDimensions))),
      simulation,
      alpha = 1,
      alphaMin = 0.001,
      alphaDecay = 1 - Math.pow(alphaMin, 1 / 300),
      alphaTarget = 0,
      velocityDecay = 0.6,
      forces = new Map(),
      stepper = timer(step),
      event = dispatch("tick", "end"),
      random = lcg();

  var engine = null, ready, usable = false, running = true, frame = false, flushing = null,
      batch = false, seeded = false, destroyed = false, failure = null;

  if (nodes == null) nodes = [];

  function fail(error) {
    failure = error;
    if (engine) engine.destroy();
    engine = null, usable = false, frame = false;
    if (!destroyed) console.warn("forceSimulationGPU: the GPU backend failed.", error);
  }

  function requireGPU() {
    if (usable) return;
    throw failure || new Error("forceSimulationGPU: no usable GPU (the HIP backend could not start)");
  }

  // (Re)build the pass list from the forces; a force without gpu() is an error.
  function refresh() {
    if (!engine) return;
    var list = Array.from(forces.values()), names = Array.from(forces.keys());
    for (var i = 0; i < list.length; ++i) {
      if (typeof list[i].gpu !== "function") {
        throw new Error("Force \"" + names[i] + "\" cannot run on the GPU (it has no force.gpu())");
      }
    }
    if (!engine.setForces(list)) throw new Error("forceSimulationGPU: unsupported force type");
  }

  function advance(iterations) {
    if (!seeded) seeded = true, engine.setSeed(random);
    var alphas = new Array(iterations);
    for (var k = 0; k < iterations; ++k) {
      alpha += (alphaTarget - alpha) * alphaDecay;
      alphas[k] = alpha;
    }
    engine.enqueue(alphas, velocityDecay);
  }

  function step() {
    if (!usable) return;
    if (frame) return;
    frame = true;
    advance(1);
    (readback ? engine.sync() : engine.idle()).then(function() {
      frame = false;
      if (!running) return;
      event.call("tick", simulation);
      if (alpha < alphaMin) {
        running = false;
        stepper.stop();
        event.call("end", simulation);
      }
    }, fail);
  }

  function flush() {
    if (flushing) return;
    flushing = engine.sync().then(function() { flushing = null; }, function(error) { flushing = null, fail(error); });
  }

  function tick(iterations) {
    if (iterations === undefined) iterations = 1;
    requireGPU();
    var frameCall = !batch;
    if (frameCall) batch = true, Promise.resolve().then(function() { batch = false; });
    if (frameCall && iterations === 1 && engine.busy()) return simulation;
    advance(iterations);
    if (readback) flush();
    return simulation;
  }

  function initializeNodes() {
    initializeNodesIn(nodes, nDim);
  }

  function initializeForce(force) {
    if (force.initialize) force.initialize(nodes, random, nDim);
    return force;
  }

  function reset() {
    if (!engine) return;
    if (!engine.fits(nodes.length)) {
      fail(new Error(nodes.length + " nodes exceed this GPU's memory"));
      return;
    }
    engine.setNodes(nodes, nDim);
    refresh();
  }

  function init() {
    return Promise.resolve().then(function() {
      if (destroyed) return false;
      if (deviceCount() < 1) throw new Error("forceSimulationGPU: no GPU visible to the HIP backend");
      engine = createEngine(options.device || 0);
      usable = true;
      reset();
      return usable;
    }).catch(function(error) {
      fail(error);
      return false;
    });
  }

  initializeNodes();
  ready = init();

  return simulation = {
    tick: tick,

    restart: function() {
      return running = true, stepper.restart(step), simulation;
    },

    stop: function() {
      return running = false, stepper.stop(), simulation;
    },

    numDimensions: function(_) {
      return arguments.length
          ? (nDim = Math.min(MAX_DIMENSIONS, Math.max(1, Math.round(_))), forces.forEach(initializeForce), reset(), simulation)
          : nDim;
    },

    nodes: function(_) {
      return arguments.length ? (nodes = _, initializeNodes(), forces.forEach(initializeForce), reset(), simulation) : nodes;
    },

    alpha: function(_) {
      return arguments.length ? (alpha = +_, simulation) : alpha;
    },

    alphaMin: function(_) {
      return arguments.length ? (alphaMin = +_, simulation) : alphaMin;
    },

    alphaDecay: function(_) {
      return arguments.length ? (alphaDecay = +_, simulation) : +alphaDecay;
    },

    alphaTarget: function(_) {
      return arguments.length ? (alphaTarget = +_, simulation) : alphaTarget;
    },

    velocityDecay: function(_) {
      return arguments.length ? (velocityDecay = 1 - _, simulation) : 1 - velocityDecay;
    },

    randomSource: function(_) {
      return arguments.length ? (random = _, forces.forEach(initializeForce), engine && (seeded = true, engine.setSeed(random)), simulation) : random;
    },

    force: function(name, _) {
      if (arguments.length < 2) return forces.get(name);
      if (_ == null) forces.delete(name);
      else forces.set(name, initializeForce(_));
      refresh();
      return simulation;
    },

    find: function() {
      var args = Array.prototype.slice.call(arguments);
      var x = args.shift() || 0,
          y = (nDim > 1 ? args.shift() : null) || 0,
          z = (nDim > 2 ? args.shift() : null) || 0,
          radius = args.shift() || Infinity;
      var i = 0, n = nodes.length, dx, dy, dz, d2, node, closest;
      radius *= radius;
      for (i = 0; i < n; ++i) {
        node = nodes[i];
        dx = x - node.x;
        dy = y - (node.y || 0);
        dz = z - (node.z || 0);
        d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < radius) closest = node, radius = d2;
      }
      return closest;
    },

    on: function(name, _) {
      return arguments.length > 1 ? (event.on(name, _), simulation) : event.on(name);
    },

    // --- additions to the forceSimulation API (src/gpu/simulation.js:384-427) ---

    tickAsync: function(iterations) {
      if (iterations === undefined) iterations = 1;
      return ready.then(function() {
        requireGPU();
        advance(iterations);
        return engine.sync().then(function() { return simulation; });
      });
    },

    sync: function() {
      return ready.then(function() {
        return usable && engine.sync();
      }).then(function() { return simulation; });
    },

    isGPUEnabled: function() {
      return usable;
    },

    gpuReady: function() {
      return ready.then(function() { return usable; });
    },

    // {positions, velocities, count, stride}: device pointers (BigInt) of float4 per node,
    // xyz used, kept current after every step.  null until the GPU is ready.
    gpuBuffers: function() {
      return engine && usable ? engine.buffers() : null;
    },

    destroy: function() {
      destroyed = true, running = false, usable = false;
      stepper.stop();
      if (engine) engine.destroy(), engine = null;
      return simulation;
    }
  };
}
