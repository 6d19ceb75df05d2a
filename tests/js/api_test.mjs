// The JavaScript host without a GPU: exports, keys, initial positions, descriptors, the CPU
// path against the reference's own stored results (test/collide-test.js, center-test.js,
// x-test.js, find-test.js, simulation-test.js), and the hand-overs and fallbacks of
// test/gpu-simulation-test.js driven by a fake GPU engine.  Every simulation here is made
// with no GPU visible (useBackendForTests), whatever the machine has.
import assert from "assert";
import * as d3 from "../../d3-force-3d-webgpu_amd/js/index.js";
import {useBackendForTests} from "../../d3-force-3d-webgpu_amd/js/simulation.js";
import {integrate} from "../../d3-force-3d-webgpu_amd/js/cpu.js";
import {it, run} from "./harness.mjs";

const SIM_KEYS = ["alpha", "alphaDecay", "alphaMin", "alphaTarget", "find", "force", "nodes", "numDimensions", "on",
                  "randomSource", "restart", "stop", "tick", "velocityDecay"];
const GPU_KEYS = ["destroy", "gpuBuffers", "gpuReady", "isGPUEnabled", "sync", "tickAsync"];
const NO_GPU = {available: () => false, create: () => { throw new Error("no GPU"); }};

// The warnings `fn` causes (but, unless `all`, those of the CPU-only reference simulations,
// which come when they find no GPU).
async function withWarnings(fn, all) {
  const warn = console.warn, warnings = [];
  console.warn = function() {
    const w = Array.prototype.join.call(arguments, " ");
    if (all || !/no GPU visible/.test(w)) warnings.push(w);
  };
  try { await fn(warnings); } finally { console.warn = warn; }
  return warnings;
}

function cpuSimulation(nodes, dims, options) {
  useBackendForTests(NO_GPU);
  const warn = console.warn;
  console.warn = function() {};
  try {
    return d3.forceSimulationGPU(nodes, dims, options).stop();
  } finally {
    console.warn = warn;
    useBackendForTests(null);
  }
}

// test/asserts.js: index exact, positions and velocities within 1e-6
function assertNode(actual, expected) {
  assert.strictEqual(actual.index, expected.index);
  for (const k of ["x", "y", "vx", "vy", "fx", "fy"]) {
    if (expected[k] === undefined) continue;
    assert(Math.abs(actual[k] - expected[k]) < 1e-6, k + ": " + actual[k] + " vs " + expected[k]);
  }
}

function makeNodes(n, dims, seed) {
  let s = seed || 12345;
  const rnd = () => (s = (s * 1664525 + 1013904223) % 4294967296) / 4294967296, nodes = [];
  for (let i = 0; i < n; ++i) {
    const node = {x: (rnd() - 0.5) * 200};
    if (dims > 1) node.y = (rnd() - 0.5) * 200;
    if (dims > 2) node.z = (rnd() - 0.5) * 200;
    nodes.push(node);
  }
  return nodes;
}

const state = (nodes) => nodes.map((n) => [n.x, n.y, n.z, n.vx, n.vy, n.vz]);

// A GPU engine that runs the forces on the nodes at once (so a run through it must equal the
// CPU path bit for bit), with positions and velocities rounded to f32 when `f32` is set (so
// its ticks can be told apart), and failures and a capacity on demand.
function fakeBackend(opts) {
  opts = opts || {};
  const made = [];
  return {
    made: made,
    available: () => true,
    create() {
      const e = {
        nodes: [], nDim: 2, forces: null, enqueues: 0, syncs: 0, stale: 0, alive: true,
        fits: (n) => opts.capacity == null || n <= opts.capacity,
        setNodes(nodes, nDim) { e.nodes = nodes, e.nDim = nDim; return e; },
        markStale() { e.stale++; },
        setSeed() {},
        setForces(forces) { e.forces = forces; return true; },
        enqueue(alphas, decay) {
          if (++e.enqueues === opts.failEnqueueAt) throw new Error("injected enqueue failure");
          for (const a of alphas) {
            e.forces.forEach((f) => f(a));
            integrate(e.nodes, e.nDim, decay);
            if (opts.f32) for (const n of e.nodes) for (const k of ["x", "y", "z", "vx", "vy", "vz"]) if (k in n) n[k] = Math.fround(n[k]);
          }
        },
        busy: () => false,
        sync() {
          return ++e.syncs === opts.failSyncAt ? Promise.reject(new Error("injected wait failure")) : Promise.resolve();
        },
        idle: () => Promise.resolve(),
        buffers: () => ({count: e.nodes.length, stride: 16}),
        destroy() { e.alive = false; }
      };
      made.push(e);
      return e;
    }
  };
}

function withBackend(b, make) {
  useBackendForTests(b);
  try { return make(); } finally { useBackendForTests(null); }
}

it("exports the forces and the GPU simulation", function() {
  for (const name of ["forceCenter", "forceCollide", "forceLink", "forceManyBody", "forceRadial", "forceSimulationGPU",
                      "forceX", "forceY", "forceZ", "isWebGPUAvailable", "checkWebGPUSupport"]) {
    assert.strictEqual(typeof d3[name], "function", name);
  }
});

it("has the forceSimulation API plus the GPU additions", function() {
  const sim = cpuSimulation();
  assert.deepStrictEqual(Object.keys(sim).sort(), SIM_KEYS.concat(GPU_KEYS).sort());
  sim.destroy();
});

it("initializes nodes with indices and phyllotaxis positions (simulation-test.js)", function() {
  const sim = cpuSimulation(), a = {}, b = {}, c = {};
  sim.nodes([a, b, c]);
  assertNode(a, {index: 0, x: 7.0710678118654755, y: 0, vy: 0, vx: 0});
  assertNode(b, {index: 1, x: -9.03088751750192, y: 8.27303273571596, vy: 0, vx: 0});
  assertNode(c, {index: 2, x: 1.3823220809823638, y: -15.750847141167634, vy: 0, vx: 0});
  sim.destroy();
});

it("without a GPU it is the CPU simulation, with one warning", async function() {
  let sim;
  const warnings = await withWarnings(async function() {
    useBackendForTests(NO_GPU);
    try { sim = d3.forceSimulationGPU(makeNodes(20, 2), 2).stop().force("charge", d3.forceManyBody()); }
    finally { useBackendForTests(null); }
    assert.strictEqual(await sim.gpuReady(), false);
    await sim.tickAsync(3);
  }, true);
  assert.strictEqual(sim.isGPUEnabled(), false);
  assert.strictEqual(sim.gpuBuffers(), null);
  assert.strictEqual(warnings.length, 1, warnings.join("; "));
  assert(/no GPU/.test(warnings[0]), warnings[0]);
  assert(sim.nodes().every((n) => n.vx !== 0 && isFinite(n.x)));
  sim.destroy();
});

it("with fallback: false, stepping without a GPU throws", async function() {
  const sim = cpuSimulation([{}, {}], 3, {fallback: false});
  assert.strictEqual(await sim.gpuReady(), false);
  assert.throws(() => sim.tick(), /GPU/);
  await assert.rejects(() => sim.tickAsync(1), /GPU/);
  sim.destroy();
});

it("collides nodes on the CPU (collide-test.js:5-19)", function() {
  const collide = d3.forceCollide(1), f = cpuSimulation().force("collide", collide), a = {}, b = {}, c = {};
  f.nodes([a, b, c]);
  f.tick(10);
  assertNode(a, {index: 0, x: 7.0710678118654755, y: 0, vy: 0, vx: 0});
  assertNode(b, {index: 1, x: -9.03088751750192, y: 8.27303273571596, vy: 0, vx: 0});
  assertNode(c, {index: 2, x: 1.3823220809823638, y: -15.750847141167634, vy: 0, vx: 0});
  collide.radius(100);
  f.tick(10);
  assertNode(a, {index: 0, x: 174.08616723117228, y: 66.51743051995625, vy: 0.26976816231064354, vx: 0.677346615710878});
  assertNode(b, {index: 1, x: -139.73606544743998, y: 95.69860503079263, vy: 0.3545632444404687, vx: -0.5300880593105067});
  assertNode(c, {index: 2, x: -34.9275994083864, y: -169.69384995620052, vy: -0.6243314067511122, vx: -0.1472585564003713});
});

it("collide respects fixed positions (collide-test.js:22-32)", function() {
  const collide = d3.forceCollide(1), f = cpuSimulation().force("collide", collide);
  const a = {fx: 0, fy: 0}, b = {}, c = {};
  f.nodes([a, b, c]);
  f.tick(10);
  assertNode(a, {fx: 0, fy: 0, index: 0, x: 0, y: 0, vy: 0, vx: 0});
  collide.radius(100);
  f.tick(10);
  assertNode(a, {fx: 0, fy: 0, index: 0, x: 0, y: 0, vy: 0, vx: 0});
});

it("collide jiggles equal positions, reproducibly (collide-test.js:34-50)", function() {
  const f = cpuSimulation().force("collide", d3.forceCollide(1)), a = {x: 0, y: 0}, b = {x: 0, y: 0};
  f.nodes([a, b]);
  f.tick();
  assert(a.x !== b.x && a.y !== b.y);
  assert.strictEqual(a.vx, -b.vx);
  assert.strictEqual(a.vy, -b.vy);
  const nodes = Array.from({length: 10}, () => ({x: 0, y: 0}));
  cpuSimulation().nodes(nodes).force("collide", d3.forceCollide()).tick(50);
  assertNode(nodes[0], {x: -5.371433857229194, y: -2.6644608278592576, index: 0, vy: 0, vx: 0});
});

it("center repositions nodes, fixed ones stay (center-test.js)", function() {
  const f = cpuSimulation().force("center", d3.forceCenter(0, 0));
  const a = {x: 100, y: 0, z: 0}, b = {x: 200, y: 0, z: 0}, c = {x: 300, y: 0, z: 0};
  f.nodes([a, b, c]);
  f.tick();
  assertNode(a, {index: 0, x: -100, y: 0, vy: 0, vx: 0});
  assertNode(b, {index: 1, x: 0, y: 0, vy: 0, vx: 0});
  assertNode(c, {index: 2, x: 100, y: 0, vy: 0, vx: 0});
  const g = cpuSimulation().force("center", d3.forceCenter()), p = {fx: 0, fy: 0, fz: 0}, q = {}, r = {};
  g.nodes([p, q, r]);
  g.tick();
  assertNode(p, {fx: 0, fy: 0, index: 0, x: 0, y: 0, vy: 0, vx: 0});
});

it("forceX / forceY centre nodes, with accessors (x-test.js)", function() {
  for (const [axis, make] of [["x", d3.forceX], ["y", d3.forceY]]) {
    const other = axis === "x" ? "y" : "x", v = "v" + axis;
    const f = cpuSimulation().force(axis, make(200));
    const a = {[axis]: 100, [other]: 0}, b = {[axis]: 200, [other]: 0}, c = {[axis]: 300, [other]: 0};
    f.nodes([a, b, c]);
    f.tick(30);
    assert(a[axis] > 190 && a[v] > 0);
    assert(b[axis] === 200 && b[v] === 0);
    assert(c[axis] < 210 && c[v] < 0);
    const g = cpuSimulation().force(axis, make().strength(0.1)[axis]((d) => d.target));
    const p = {[axis]: 100, [other]: 0, target: 300}, q = {[axis]: 200, [other]: 0, target: 200}, r = {[axis]: 300, [other]: 0, target: 100};
    g.nodes([p, q, r]);
    g.tick(30);
    assert(p[axis] > 290 && p[v] > 0);
    assert(q[axis] === 200 && q[v] === 0);
    assert(r[axis] < 110 && r[v] < 0);
  }
  const h = cpuSimulation().force("x", d3.forceX(200)), a = {fx: 0, fy: 0}, b = {}, c = {};
  h.nodes([a, b, c]);
  h.tick();
  assertNode(a, {fx: 0, fy: 0, index: 0, x: 0, y: 0, vy: 0, vx: 0});
});

it("find() finds the closest node, within a radius (find-test.js)", function() {
  const f = cpuSimulation(), a = {x: 5, y: 0}, b = {x: 10, y: 16}, c = {x: -10, y: -4};
  f.nodes([a, b, c]);
  assert.strictEqual(f.find(0, 0), a);
  assert.strictEqual(f.find(0, 0, 1), undefined);
  assert.strictEqual(f.find(0, 20), b);
});

it("calls forces as d3 does: force(alpha) runs its CPU physics", function() {
  const nodes = makeNodes(30, 3);
  nodes.forEach((n, i) => { n.index = i, n.vx = n.vy = n.vz = 0; });
  const charge = d3.forceManyBody();
  charge.initialize(nodes, Math.random, 3);
  charge(0.5);
  assert(nodes.every((n) => n.vx !== 0 && n.vz !== 0), "the force moved every node's velocity");
});

it("describes every force with gpu() like the reference descriptors", function() {
  const nodes = [{index: 0}, {index: 1}, {index: 2}];
  const mb = d3.forceManyBody().strength(-10).theta(0.5).distanceMin(2).distanceMax(100);
  mb.initialize(nodes);
  assert.deepStrictEqual(mb.gpu(), {type: "manyBody", version: 1, strengths: [-10, -10, -10], distanceMin2: 4,
                                    distanceMax2: 10000, theta2: 0.25});
  const ln = d3.forceLink([{source: 0, target: 1}, {source: 1, target: 2}]);
  ln.initialize(nodes);
  const g = ln.gpu();
  assert.strictEqual(g.type, "link");
  assert.deepStrictEqual(g.bias, [1 / 3, 2 / 3]);  // count[source] / (count[source] + count[target])
  assert.deepStrictEqual(g.strengths, [1, 1]);      // 1 / min(count)
  assert.deepStrictEqual(g.distances, [30, 30]);
  const v0 = g.version;
  ln.distance(50);
  assert(ln.gpu().version > v0, "changing a parameter bumps the version");
  const r = d3.forceRadial(10, 1, 2, 3).strength(0.3);
  r.initialize(nodes);
  assert.deepStrictEqual(r.gpu(), {type: "radial", version: 1, radiuses: [10, 10, 10], strengths: [0.3, 0.3, 0.3],
                                   x: 1, y: 2, z: 3});
  assert.deepStrictEqual(d3.forceCenter(1, 2, 3).strength(0.5).gpu(), {type: "center", version: 0, x: 1, y: 2, z: 3,
                                                                          strength: 0.5});
  const fz = d3.forceZ(7);
  fz.initialize(nodes);
  assert.strictEqual(fz.gpu().type, "position");
  assert.strictEqual(fz.gpu().axis, 2);
});

// ---- hand-overs and fallbacks (test/gpu-simulation-test.js:147-346, 384-406) ----------

function gravity() {
  let nodes;
  function force(alpha) { for (const n of nodes) n.vy -= 0.5 * alpha; }
  force.initialize = function(_) { nodes = _; };
  return force;
}

function chargedLinks(sim) {
  const n = sim.nodes().length, links = Array.from({length: n - 1}, (_, i) => ({source: i + 1, target: i >> 1}));
  return sim.force("link", d3.forceLink(links)).force("charge", d3.forceManyBody());
}

it("a force without gpu() moves the physics to the CPU and back, losing nothing", async function() {
  const b = fakeBackend(), nodes = makeNodes(40, 3), cpuNodes = makeNodes(40, 3);
  const warnings = await withWarnings(async function() {
    const sim = chargedLinks(withBackend(b, () => d3.forceSimulationGPU(nodes, 3).stop()));
    const cpu = chargedLinks(cpuSimulation(cpuNodes, 3));
    assert.strictEqual(await sim.gpuReady(), true);
    await sim.tickAsync(3), cpu.tick(3);
    sim.force("gravity", gravity()), cpu.force("gravity", gravity());
    assert.strictEqual(sim.isGPUEnabled(), false);
    await sim.tickAsync(3), cpu.tick(3);
    assert(nodes.every((n) => isFinite(n.vy)));
    sim.force("gravity", null), cpu.force("gravity", null);
    assert.strictEqual(sim.isGPUEnabled(), true);
    await sim.tickAsync(3), cpu.tick(3);
    sim.destroy();
  });
  assert.strictEqual(warnings.length, 1);
  assert(/gravity/.test(warnings[0]), warnings[0]);
  assert.deepStrictEqual(state(nodes), state(cpuNodes));
  assert(b.made[0].stale >= 2, "each arrival on the GPU uploads the nodes the CPU moved");
});

it("a failure while stepping continues on the CPU from where the GPU was", async function() {
  for (const opts of [{failEnqueueAt: 3}, {failSyncAt: 2}]) {
    const b = fakeBackend(opts), nodes = makeNodes(30, 2), cpuNodes = makeNodes(30, 2);
    const warnings = await withWarnings(async function() {
      const sim = chargedLinks(withBackend(b, () => d3.forceSimulationGPU(nodes, 2).stop()));
      const cpu = chargedLinks(cpuSimulation(cpuNodes, 2));
      for (let k = 0; k < 4; ++k) await sim.tickAsync(2), cpu.tick(2);
      assert.strictEqual(sim.isGPUEnabled(), false);
      assert.strictEqual(b.made[0].alive, false, "the failed engine is released");
      sim.destroy();
    });
    assert(warnings.some((w) => /failed/.test(w)), warnings.join("; "));
    assert.deepStrictEqual(state(nodes), state(cpuNodes), JSON.stringify(opts));
  }
});

it("with fallback: false, a GPU failure is thrown", async function() {
  const b = fakeBackend({failEnqueueAt: 2}), nodes = makeNodes(10, 2);
  const sim = chargedLinks(withBackend(b, () => d3.forceSimulationGPU(nodes, 2, {fallback: false}).stop()));
  await sim.tickAsync(1);
  await assert.rejects(() => sim.tickAsync(1), /injected/);
  assert.throws(() => sim.tick(), /injected/);
  sim.destroy();
});

it("nodes beyond the device's capacity run on the CPU, with a warning", async function() {
  const b = fakeBackend({capacity: 10}), nodes = makeNodes(30, 2);
  let sim;
  const warnings = await withWarnings(async function() {
    sim = withBackend(b, () => d3.forceSimulationGPU(nodes, 2).stop().force("charge", d3.forceManyBody()));
    assert.strictEqual(await sim.gpuReady(), false);
    sim.force("center", d3.forceCenter());  // must not bring the GPU back
    await sim.tickAsync(2);
  });
  assert.strictEqual(sim.isGPUEnabled(), false);
  assert(/exceed/.test(warnings.join()), warnings.join());
  assert(nodes.every((n) => isFinite(n.x) && n.vx !== 0));
  sim.destroy();
});

function record(sim, nodes, count) {
  const states = [];
  return new Promise(function(resolve) {
    sim.on("tick", function() {
      states.push(state(nodes));
      if (states.length === count) sim.on("tick", null).stop(), resolve(states);
    });
  });
}

const isF32 = (s) => s.every((d) => d.every((v) => v === undefined || Math.fround(v) === v));

it("the timer's first cpuTicks ticks are the CPU's, then the GPU's", async function() {
  const b = fakeBackend({f32: true}), nodes = makeNodes(60, 3), cpuNodes = makeNodes(60, 3);
  const sim = chargedLinks(withBackend(b, () => d3.forceSimulationGPU(nodes, 3, {cpuTicks: 10})));
  const states = await record(sim, nodes, 11);
  chargedLinks(cpuSimulation(cpuNodes, 3)).tick(10);
  assert.deepStrictEqual(states[9], state(cpuNodes), "ticks 1 to 10 are the CPU simulation's, to the last bit");
  assert(!isF32(states[9]));
  assert(isF32(states[10]), "tick 11 ran on the GPU");
  sim.destroy();
});

it("cpuTicks: 120 by default up to 1,000 nodes, none above, none once tick() is called", async function() {
  const small = fakeBackend({f32: true}), big = fakeBackend({f32: true}), a = makeNodes(30, 3), c = makeNodes(1001, 3);
  const simA = withBackend(small, () => d3.forceSimulationGPU(a, 3).force("charge", d3.forceManyBody()));
  assert(!isF32((await record(simA, a, 2)).slice(-1)), "default warm-up ticks on the CPU");
  simA.tick();
  assert((await record(simA.restart(), a, 3)).every(isF32), "after tick(), the timer's ticks run on the GPU");
  const simC = withBackend(big, () => d3.forceSimulationGPU(c, 3).force("charge", d3.forceManyBody()));
  assert(isF32((await record(simC, c, 1))[0]), "above 1,000 nodes the first tick is the GPU's");
  simA.destroy(), simC.destroy();
});

run();
