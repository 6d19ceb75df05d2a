// forceSimulationGPU: d3-force-3d's forceSimulation with the physics on an AMD GPU, through
// the HIP backend (the reference's src/gpu/simulation.js is the behavioural specification).
//
// The physics is in one of two places, `where`:
//   "cpu"  the forces run as d3 runs them, force(alpha) on the node objects (cpu.js);
//   "gpu"  the GPU engine (engine.js) holds positions and velocities and the nodes follow
//          one readback behind; tick() queues steps and returns.
// A simulation starts on the CPU while the backend comes up (gpuReady() settles once it
// has), spends its first cpuTicks timer ticks there (default: 120 up to 1,000 nodes when it
// reads back, else none -- the tick at which the GPU takes over shapes the layout, so the
// count is fixed rather than "until the device is ready"), then moves to the GPU if every
// force has a GPU description, the nodes fit and the device works.  It goes back to the
// CPU after one last readback when a force without one is added, and for good when the GPU
// fails; both with a warning, or, with options.fallback === false, a failure is thrown by
// the next tick() / tickAsync() instead.
//
// Options: device (GPU index), readback (false: leave the nodes alone until sync() or
// tickAsync(); gpuBuffers() has the state), cpuTicks, fallback.
import dispatch from "./dispatch.js";
import timer from "./timer.js";
import lcg from "./lcg.js";
import GpuEngine, {describable} from "./engine.js";
import {integrate} from "./cpu.js";
import {deviceCount} from "./native.js";
import {initializeNodes as placeUnplaced} from "./nodes.js";

const MAX_DIMENSIONS = 3;
const WARM_TICKS = 120, WARM_NODES = 1000;

export function isBackendAvailable() {
  try {
    return deviceCount() > 0;
  } catch (error) {
    return false;
  }
}

// Test hook: {available() -> boolean, create(device) -> engine} replaces how simulations made
// after the call find and make their GPU engine (a fake one lets the hand-overs and
// fallbacks run where there is no GPU); null restores the HIP backend.
let backend = null;
export function useBackendForTests(b) {
  backend = b || null;
}

function dimsOf(value) {
  return Math.min(MAX_DIMENSIONS, Math.max(1, Math.round(value)));
}

export default function forceSimulationGPU(nodes, numDimensions, options) {
  options = options || {};
  const readback = options.readback !== false, fallback = options.fallback !== false;
  const forces = new Map(), events = dispatch("tick", "end");
  let nDim = dimsOf(numDimensions || 2);
  let alpha = 1, alphaMin = 0.001, alphaDecay = 1 - Math.pow(alphaMin, 1 / 300), alphaTarget = 0,
      velocityDecay = 0.6, random = lcg();
  let simulation = null;

  // where the physics is, and the GPU side's state
  let where = "cpu", gpu = null, starting = true, oversized = false, destroyed = false;
  let failure = null;      // why the GPU is gone (thrown when fallback is off)
  let leaving = null;      // the last readback before the physics moves to the CPU
  let frameOut = false;    // a timer frame's GPU step + readback is under way
  let following = null;    // the readback loop tick() starts
  let inRun = false;       // tick() already ran in this synchronous run of the caller
  let warm = 0;            // timer ticks still to run on the CPU
  let seeded = false;      // the device's jiggle seed has been drawn from `random`
  let noticed = null;      // the force last reported as CPU-only
  let running = true;

  if (nodes == null) nodes = [];

  function warn(message, error) {
    if (!destroyed) console.warn("forceSimulationGPU: " + message, error === undefined ? "" : error);
  }

  // ---- moving the physics ---------------------------------------------------------------

  // The GPU is gone for good: continue on the CPU (the nodes hold the last readback).
  function gpuLost(error) {
    if (gpu) gpu.destroy();
    gpu = null, where = "cpu", frameOut = false, failure = error || failure;
    if (fallback) warn("the GPU backend failed; continuing on the CPU.", error);
  }

  // To the CPU while the GPU stays up: first bring the nodes up to date.
  function leaveGpu() {
    if (where !== "gpu") return;
    where = "cpu";
    leaving = gpu.sync().then(() => { leaving = null; }, (error) => { leaving = null; gpuLost(error); });
  }

  // To the GPU: it takes the nodes as the CPU left them.
  function enterGpu() {
    if (where === "gpu") return;
    where = "gpu";
    gpu.markStale();
  }

  // The pass list from the current forces, or back to the CPU for a force without gpu().
  function configure() {
    if (!gpu || oversized) return;
    const list = Array.from(forces.values());
    const cpuOnly = Array.from(forces.keys()).find((name) => !describable(forces.get(name)));
    if (cpuOnly === undefined) {
      noticed = null;
      gpu.setForces(list);
      enterGpu();
      return;
    }
    if (noticed !== cpuOnly) {
      noticed = cpuOnly;
      warn("force \"" + cpuOnly + "\" has no GPU description (force.gpu()): simulating on the CPU.");
    }
    leaveGpu();
  }

  // The node set on the device, if it fits there.
  function placeNodes() {
    if (!gpu) return;
    oversized = !gpu.fits(nodes.length);
    if (oversized) {
      gpu.setNodes([], nDim); // frees the device buffers
      where = "cpu";
      warn(nodes.length + " nodes exceed this GPU's memory: simulating on the CPU.");
      return;
    }
    gpu.setNodes(nodes, nDim);
    configure();
  }

  function start() {
    const source = backend;
    return Promise.resolve().then(() => {
      if (destroyed) return false;
      if (!(source ? source.available() : isBackendAvailable())) {
        failure = new Error("no GPU visible to the HIP backend");
        if (fallback) warn("no GPU visible to the HIP backend; simulating on the CPU.");
        return false;
      }
      gpu = source ? source.create(options.device || 0) : new GpuEngine(options.device || 0);
      placeNodes();
      return where === "gpu";
    }).catch((error) => {
      gpuLost(error);
      return false;
    }).then((onGpu) => {
      starting = false;
      return onGpu;
    });
  }

  // ---- stepping -------------------------------------------------------------------------

  function nextAlpha() {
    return alpha += (alphaTarget - alpha) * alphaDecay;
  }

  // d3's tick, on the CPU.
  function stepCpu(iterations, warming) {
    if (leaving) return simulation;
    if (!fallback && !warming && !starting && failure) throw failure;
    for (let k = 0; k < iterations; ++k) {
      const a = nextAlpha();
      forces.forEach((force) => force(a));
      integrate(nodes, nDim, velocityDecay);
    }
    return simulation;
  }

  // Queue `iterations` steps on the GPU; if the GPU refuses them, they run on the CPU.
  function stepGpu(iterations) {
    // the seed is drawn at the first GPU step: CPU ticks may draw from `random` before, and
    // how many they are must not depend on when the device came up
    if (!seeded) seeded = true, gpu.setSeed(random);
    const before = alpha, alphas = new Array(iterations);
    for (let k = 0; k < iterations; ++k) alphas[k] = nextAlpha();
    try {
      gpu.enqueue(alphas, velocityDecay);
    } catch (error) {
      alpha = before;
      gpuLost(error);
      stepCpu(iterations, false);
    }
  }

  function afterTick() {
    events.call("tick", simulation);
    if (alpha < alphaMin) {
      running = false;
      stepper.stop();
      events.call("end", simulation);
    }
  }

  // One timer frame.
  function frame() {
    if (leaving) return;
    if (warm > 0) {
      --warm;
      if (gpu) gpu.markStale(); // the GPU takes over from wherever the CPU leaves the nodes
      stepCpu(1, true);
      afterTick();
      return;
    }
    if (starting) return; // the backend is still coming up
    if (where === "gpu") {
      if (frameOut) return;
      frameOut = true;
      stepGpu(1);
      if (where !== "gpu") {  // the GPU refused the step: it ran on the CPU
        frameOut = false;
        afterTick();
        return;
      }
      (readback ? gpu.sync() : gpu.idle()).then(() => {
        frameOut = false;
        if (running) afterTick();
      }, gpuLost);
      return;
    }
    stepCpu(1, false);
    afterTick();
  }

  const stepper = timer(frame);

  // Keep the nodes following the GPU after tick(): one readback loop at a time.
  function follow() {
    if (following) return;
    following = gpu.sync().then(() => { following = null; }, (error) => { following = null; gpuLost(error); });
  }

  function tick(iterations) {
    if (iterations === undefined) iterations = 1;
    warm = 0;
    if (where !== "gpu") return stepCpu(iterations, false);
    // one tick() per frame must not pile up work when the GPU falls behind, so such a call
    // is skipped while it is busy; more calls in the same synchronous run (a warm-up loop)
    // are a batch and always run
    const first = !inRun;
    if (first) inRun = true, Promise.resolve().then(() => { inRun = false; });
    if (first && iterations === 1 && gpu.busy()) return simulation;
    stepGpu(iterations);
    if (readback && where === "gpu") follow();
    return simulation;
  }

  function initializeForce(force) {
    if (force.initialize) force.initialize(nodes, random, nDim);
    return force;
  }

  // the timer's first ticks on a node set run on the CPU
  function armWarmup() {
    if (options.cpuTicks != null) warm = Math.max(0, Math.floor(options.cpuTicks)) || 0;
    else warm = readback && nodes.length <= WARM_NODES ? WARM_TICKS : 0;
  }

  placeUnplaced(nodes, nDim);
  armWarmup();
  const ready = start();
  const settled = () => ready.then(() => leaving);

  return simulation = {
    tick: tick,

    restart() {
      running = true;
      stepper.restart(frame);
      return simulation;
    },

    stop() {
      running = false;
      stepper.stop();
      return simulation;
    },

    numDimensions(_) {
      if (!arguments.length) return nDim;
      nDim = dimsOf(_);
      forces.forEach(initializeForce);
      placeNodes();
      return simulation;
    },

    nodes(_) {
      if (!arguments.length) return nodes;
      nodes = _;
      placeUnplaced(nodes, nDim);
      armWarmup();
      forces.forEach(initializeForce);
      placeNodes();
      return simulation;
    },

    alpha(_) {
      return arguments.length ? (alpha = +_, simulation) : alpha;
    },

    alphaMin(_) {
      return arguments.length ? (alphaMin = +_, simulation) : alphaMin;
    },

    alphaDecay(_) {
      return arguments.length ? (alphaDecay = +_, simulation) : +alphaDecay;
    },

    alphaTarget(_) {
      return arguments.length ? (alphaTarget = +_, simulation) : alphaTarget;
    },

    velocityDecay(_) {
      return arguments.length ? (velocityDecay = 1 - _, simulation) : 1 - velocityDecay;
    },

    randomSource(_) {
      if (!arguments.length) return random;
      random = _;
      forces.forEach(initializeForce);
      if (gpu) seeded = true, gpu.setSeed(random);
      return simulation;
    },

    force(name, _) {
      if (arguments.length < 2) return forces.get(name);
      if (_ == null) forces.delete(name);
      else forces.set(name, initializeForce(_));
      configure();
      return simulation;
    },

    // the node closest to (x, y, z) within `radius` (d3's find, over 1..3 dimensions)
    find(...args) {
      const x = args.shift() || 0, y = (nDim > 1 ? args.shift() : null) || 0,
            z = (nDim > 2 ? args.shift() : null) || 0;
      let best2 = Math.pow(args.shift() || Infinity, 2), closest;
      for (const node of nodes) {
        const dx = x - node.x, dy = y - (node.y || 0), dz = z - (node.z || 0), d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best2) closest = node, best2 = d2;
      }
      return closest;
    },

    on(name, _) {
      return arguments.length > 1 ? (events.on(name, _), simulation) : events.on(name);
    },

    // ---- beyond forceSimulation (src/gpu/simulation.js:384-427) ----

    // Promise for the simulation once `iterations` ticks have run and the nodes show them.
    tickAsync(iterations) {
      if (iterations === undefined) iterations = 1;
      warm = 0;
      return settled().then(() => {
        if (where !== "gpu") return stepCpu(iterations, false);
        stepGpu(iterations);
        if (where !== "gpu") return simulation;  // the GPU refused them: they ran on the CPU
        return gpu.sync().then(() => simulation, (error) => {
          gpuLost(error);
          if (!fallback) throw error;
          return simulation;
        });
      });
    },

    // Promise for the simulation once the nodes show everything the GPU has computed.
    sync() {
      return settled().then(() => where === "gpu" && gpu.sync()).then(() => simulation, (error) => {
        gpuLost(error);
        if (!fallback) throw error;
        return simulation;
      });
    },

    isGPUEnabled() {
      return where === "gpu";
    },

    // Promise for whether the physics runs on the GPU once the backend is up.
    gpuReady() {
      return ready.then(() => where === "gpu");
    },

    // {positions, velocities, count, stride, epoch}: device pointers (BigInt) to a float4
    // per node (xyz used), current after every step; null while the physics is on the CPU.
    gpuBuffers() {
      return gpu && where === "gpu" ? gpu.buffers() : null;
    },

    destroy() {
      destroyed = true, running = false, where = "cpu";
      stepper.stop();
      if (gpu) gpu.destroy(), gpu = null;
      return simulation;
    }
  };
}
