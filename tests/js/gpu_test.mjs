// The JavaScript host on a GPU: lifecycle, async stepping, edits, pins, events, buffers
// and determinism (test/gpu-simulation-test.js is the model).
import assert from "assert";
import * as d3 from "../../d3-force-3d-webgpu_amd/js/index.js";
import {native} from "../../d3-force-3d-webgpu_amd/js/native.js";
import {it, run} from "./harness.mjs";

function makeNodes(n, dims, size, seed) {
  var s = seed || 12345, nodes = [];
  function rnd() { return Math.fround((s = (s * 1664525 + 1013904223) % 4294967296) / 4294967296); }
  for (var i = 0; i < n; ++i) {
    var node = {x: Math.fround((rnd() - 0.5) * (size || 200))};
    if (dims > 1) node.y = Math.fround((rnd() - 0.5) * (size || 200));
    if (dims > 2) node.z = Math.fround((rnd() - 0.5) * (size || 200));
    nodes.push(node);
  }
  return nodes;
}

function graph(n) {
  var s = 42, links = [];
  function rnd() { return (s = (s * 1664525 + 1013904223) % 4294967296) / 4294967296; }
  for (var i = 1; i < n; ++i) links.push({source: i, target: Math.floor(rnd() * i)});
  return links;
}

it("runs on the GPU when a device is available", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(10, 3), 3).stop();
  assert.strictEqual(await sim.gpuReady(), true);
  assert.strictEqual(sim.isGPUEnabled(), true);
  sim.destroy();
});

it("advances alpha exactly like forceSimulation", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(20, 2), 2).stop().alphaTarget(0.1).alphaDecay(0.05);
  sim.force("charge", d3.forceManyBody());
  await sim.tickAsync(17);
  let alpha = 1;
  for (let k = 0; k < 17; ++k) alpha += (0.1 - alpha) * 0.05;
  assert.strictEqual(sim.alpha(), alpha);
  sim.destroy();
});

it("tickAsync() resolves with the simulation once the nodes are up to date", async function() {
  const nodes = makeNodes(50, 3), before = nodes.map(n => n.x);
  const sim = d3.forceSimulationGPU(nodes, 3).stop().force("charge", d3.forceManyBody());
  const r = await sim.tickAsync(3);
  assert.strictEqual(r, sim);
  assert(nodes.some((n, i) => n.x !== before[i]), "nodes moved");
  nodes.forEach(n => assert(isFinite(n.x + n.y + n.z + n.vx + n.vy + n.vz)));
  sim.destroy();
});

it("tick() queues steps; sync() lands them in the nodes", async function() {
  const nodes = makeNodes(40, 2), before = nodes.map(n => n.y);
  const sim = d3.forceSimulationGPU(nodes, 2).stop().force("charge", d3.forceManyBody());
  await sim.gpuReady();
  for (let k = 0; k < 5; ++k) sim.tick();
  await sim.sync();
  assert(nodes.some((n, i) => n.y !== before[i]));
  sim.destroy();
});

it("honours fixed positions (fx, fy, fz)", async function() {
  const nodes = makeNodes(120, 3);
  nodes[0].fx = 5, nodes[0].fy = -6, nodes[0].fz = 7;
  nodes[1].fx = 100;
  const sim = d3.forceSimulationGPU(nodes, 3).stop().force("charge", d3.forceManyBody().theta(0)).force("center", d3.forceCenter());
  await sim.tickAsync(5);
  assert.deepStrictEqual([nodes[0].x, nodes[0].y, nodes[0].z, nodes[0].vx], [5, -6, 7, 0]);
  assert.strictEqual(nodes[1].x, 100);
  sim.destroy();
});

it("picks up pins and moves made between steps (dragging)", async function() {
  const nodes = makeNodes(60, 2);
  const sim = d3.forceSimulationGPU(nodes, 2).stop().force("charge", d3.forceManyBody());
  await sim.tickAsync(2);
  nodes[3].fx = 42, nodes[3].fy = -17;
  nodes[4].x = 500;
  await sim.tickAsync(1);
  assert.strictEqual(nodes[3].x, 42);
  assert.strictEqual(nodes[3].y, -17);
  assert(Math.abs(nodes[4].x - 500) < 50, "the moved node continues from where it was put: " + nodes[4].x);
  delete nodes[3].fx, delete nodes[3].fy;
  await sim.tickAsync(3);
  assert.notStrictEqual(nodes[3].x, 42, "unpinned nodes move again");
  sim.destroy();
});

it("readback: false leaves the nodes alone until sync()", async function() {
  const nodes = makeNodes(30, 3), before = nodes.map(n => n.x);
  const sim = d3.forceSimulationGPU(nodes, 3, {readback: false}).stop().force("charge", d3.forceManyBody());
  await sim.gpuReady();
  sim.tick(4);
  await new Promise(r => setTimeout(r, 50));
  assert.deepStrictEqual(nodes.map(n => n.x), before);
  await sim.sync();
  assert(nodes.some((n, i) => n.x !== before[i]));
  sim.destroy();
});

it("gpuBuffers() points at the device-resident state, bit-equal to the nodes", async function() {
  const nodes = makeNodes(33, 3);
  const sim = d3.forceSimulationGPU(nodes, 3).stop().force("charge", d3.forceManyBody());
  assert.strictEqual(sim.gpuBuffers(), null, "null until the GPU is ready");
  await sim.gpuReady();
  await sim.tickAsync(2);
  const b = sim.gpuBuffers();
  assert.strictEqual(b.count, 33);
  assert.strictEqual(b.stride, 16);
  assert.strictEqual(typeof b.positions, "bigint");
  // read the float4 per node through the pointer (test/gpu-simulation-test.js:421-443)
  const addon = native(), h = addon.create(0);
  try {
    const p = addon.readDevice(h, b.positions, 4 * 33);
    nodes.forEach(function(n, i) {
      assert.strictEqual(p[4 * i], n.x);
      assert.strictEqual(p[4 * i + 1], n.y);
      assert.strictEqual(p[4 * i + 2], n.z);
    });
    const v = addon.readDevice(h, b.velocities, 4 * 33);
    nodes.forEach(function(n, i) { assert.strictEqual(v[4 * i], n.vx); });
  } finally {
    addon.destroy(h);
  }
  sim.destroy();
});

it("is deterministic for given inputs", async function() {
  async function layout() {
    const nodes = Array.from({length: 300}, () => ({}));
    const sim = d3.forceSimulationGPU(nodes, 3).stop()
        .force("link", d3.forceLink(graph(300))).force("charge", d3.forceManyBody()).force("center", d3.forceCenter());
    await sim.tickAsync(60);
    sim.destroy();
    return nodes.map(n => [n.x, n.y, n.z]);
  }
  assert.deepStrictEqual(await layout(), await layout());
});

it("lays out a graph: links settle near their distance", async function() {
  const nodes = Array.from({length: 500}, () => ({})), links = graph(500);
  const sim = d3.forceSimulationGPU(nodes, 3).stop()
      .force("link", d3.forceLink(links).distance(30)).force("charge", d3.forceManyBody()).force("center", d3.forceCenter());
  await sim.tickAsync(300);
  let mean = 0;
  for (const l of links) mean += Math.hypot(l.source.x - l.target.x, l.source.y - l.target.y, l.source.z - l.target.z);
  mean /= links.length;
  assert(mean > 15 && mean < 120, "mean link length " + mean);
  sim.destroy();
});

it("fires tick and end events from the timer", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(20, 2), 2).alphaDecay(0.3).force("charge", d3.forceManyBody());
  let ticks = 0;
  await new Promise(function(resolve, reject) {
    sim.on("tick", function() { ticks++; });
    sim.on("end", resolve);
    setTimeout(() => reject(new Error("no end event")), 20000);
  });
  assert(ticks > 5, "ticks: " + ticks);
  sim.destroy();
});

it("takes forces added, changed and removed while it runs", async function() {
  const nodes = makeNodes(80, 3);
  const sim = d3.forceSimulationGPU(nodes, 3).stop().force("charge", d3.forceManyBody());
  await sim.tickAsync(2);
  sim.force("x", d3.forceX(300).strength(0.5));
  await sim.tickAsync(30);
  const meanX = nodes.reduce((s, n) => s + n.x, 0) / nodes.length;
  assert(meanX > 100, "forceX pulled the nodes: " + meanX);
  sim.force("x", null);
  sim.force("charge").strength(-50);
  await sim.tickAsync(2);
  sim.destroy();
});

async function withWarnings(fn) {
  const warn = console.warn, warnings = [];
  console.warn = function() { warnings.push(Array.prototype.join.call(arguments, " ")); };
  try { await fn(); } finally { console.warn = warn; }
  return warnings;
}

function gravity() {
  let nodes;
  function force(alpha) { for (const n of nodes) n.vy -= 0.5 * alpha; }
  force.initialize = function(_) { nodes = _; };
  return force;
}

it("a force without gpu() moves the physics to the CPU (with a warning) and back", async function() {
  const nodes = makeNodes(40, 3);
  let sim;
  const warnings = await withWarnings(async function() {
    sim = d3.forceSimulationGPU(nodes, 3).stop().force("charge", d3.forceManyBody());
    await sim.tickAsync(3);
    assert.strictEqual(sim.isGPUEnabled(), true);
    sim.force("gravity", gravity());
    assert.strictEqual(sim.isGPUEnabled(), false);
    const vy = nodes.map((n) => n.vy);
    await sim.tickAsync(3);
    assert(nodes.some((n, i) => n.vy !== vy[i]), "the CPU ran the forces, the custom one too");
    sim.force("gravity", null);
    assert.strictEqual(sim.isGPUEnabled(), true);
    await sim.tickAsync(3);
  });
  assert.strictEqual(warnings.length, 1);
  assert(/gravity/.test(warnings[0]), warnings[0]);
  assert(nodes.every((n) => isFinite(n.x + n.y + n.z)));
  assert(isF32(state(nodes)), "the last ticks ran on the GPU");
  sim.destroy();
});

const state = (nodes) => nodes.map((n) => [n.x, n.y, n.z, n.vx, n.vy, n.vz]);
const isF32 = (s) => s.every((d) => d.every((v) => v === undefined || Math.fround(v) === v));

function record(sim, nodes, count) {
  const states = [];
  return new Promise(function(resolve) {
    sim.on("tick", function() {
      states.push(state(nodes));
      if (states.length === count) sim.on("tick", null).stop(), resolve(states);
    });
  });
}

it("the timer's first cpuTicks ticks are the CPU simulation's, then the GPU's", async function() {
  const links = () => Array.from({length: 59}, (_, i) => ({source: i + 1, target: i >> 1}));
  const nodes = makeNodes(60, 3), cpuNodes = makeNodes(60, 3);
  const sim = d3.forceSimulationGPU(nodes, 3, {cpuTicks: 10}).force("link", d3.forceLink(links())).force("charge", d3.forceManyBody());
  const states = await record(sim, nodes, 11);
  const cpu = d3.forceSimulationGPU(cpuNodes, 3, {cpuTicks: 0}).stop().force("link", d3.forceLink(links())).force("charge", d3.forceManyBody());
  // the reference run on the CPU: the same forces called as d3 calls them
  let alpha = 1;
  const decay = 1 - Math.pow(0.001, 1 / 300), fs = [cpu.force("link"), cpu.force("charge")];
  for (let k = 0; k < 10; ++k) {
    alpha += (0 - alpha) * decay;
    fs.forEach((f) => f(alpha));
    for (const n of cpuNodes) n.x += n.vx *= 0.6, n.y += n.vy *= 0.6, n.z += n.vz *= 0.6;
  }
  assert.deepStrictEqual(states[9], state(cpuNodes), "ticks 1 to 10 ran on the CPU, to the last bit");
  assert(!isF32(states[9]));
  assert(isF32(states[10]), "tick 11 ran on the GPU");
  assert.strictEqual(sim.isGPUEnabled(), true);
  sim.destroy(), cpu.destroy();
});

it("a GPU failure while stepping continues on the CPU, with a warning", async function() {
  const nodes = makeNodes(50, 2);
  let sim;
  process.env.D3FGPU_FAIL = "enqueue:3";  // the engine's third enqueue reports a failure
  const warnings = await withWarnings(async function() {
    sim = d3.forceSimulationGPU(nodes, 2).stop().force("charge", d3.forceManyBody());
    assert.strictEqual(await sim.gpuReady(), true);
    delete process.env.D3FGPU_FAIL;
    await sim.tickAsync(1);
    await sim.tickAsync(1);
    await sim.tickAsync(2);  // refused: runs on the CPU
    assert.strictEqual(sim.isGPUEnabled(), false);
    await sim.tickAsync(2);
  });
  assert(warnings.some((w) => /failed/.test(w) && /injected/.test(w)), warnings.join("; "));
  assert(nodes.every((n) => isFinite(n.x + n.y) && !isF32([[n.x]])), "the last ticks ran on the CPU");
  sim.destroy();
});

it("nodes beyond the device's capacity run on the CPU, with a warning", async function() {
  const nodes = makeNodes(30, 2);
  let sim;
  process.env.D3FGPU_CAPACITY = "10";
  const warnings = await withWarnings(async function() {
    sim = d3.forceSimulationGPU(nodes, 2).stop().force("charge", d3.forceManyBody());
    assert.strictEqual(await sim.gpuReady(), false);
    delete process.env.D3FGPU_CAPACITY;
    sim.force("center", d3.forceCenter());  // must not bring the GPU back
    await sim.tickAsync(2);
  });
  assert.strictEqual(sim.isGPUEnabled(), false);
  assert(/exceed/.test(warnings.join()), warnings.join());
  assert(nodes.every((n) => isFinite(n.x) && n.vx !== 0));
  sim.destroy();
});

it("lays out graphs through a server and a batch, the same either way", async function() {
  function job(n, seed) {
    const nodes = [];
    for (let i = 0; i < n; ++i) nodes.push({});
    const links = graph(n).map((l) => ({source: l.source, target: (l.target + seed) % l.source}));
    return [nodes, [d3.forceLink(links), d3.forceManyBody(), d3.forceCollide(3), d3.forceCenter()]];
  }
  const server = d3.createLayoutServer({ticksPerRound: 7});
  const served = await Promise.all([0, 1, 2].map((k) => server.layout(...job(300 + 200 * k, k + 1).slice(0, 1), 3,
                                                                      job(300 + 200 * k, k + 1)[1])));
  const stats = server.stats();
  assert.strictEqual(stats.finished, 3);
  assert.ok(stats.p99 >= stats.p50 && stats.p50 > 0);
  server.destroy();
  for (let k = 0; k < 3; ++k) {
    const batch = d3.createLayoutBatch();
    const [nodes, forces] = job(300 + 200 * k, k + 1);
    const id = batch.add(nodes, 3, forces);
    while (!batch.done()) batch.step(25);
    batch.land(id, true);
    batch.destroy();
    assert.ok(served[k].ticks >= 300);
    for (let i = 0; i < nodes.length; ++i) {
      assert.strictEqual(served[k].nodes[i].x, nodes[i].x);
      assert.strictEqual(served[k].nodes[i].z, nodes[i].z);
    }
  }
});

it("destroy() releases the GPU for good", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(5, 2), 2).stop().force("charge", d3.forceManyBody());
  await sim.tickAsync(1);
  assert.strictEqual(sim.isGPUEnabled(), true);
  sim.destroy();
  assert.strictEqual(sim.isGPUEnabled(), false);
  assert.strictEqual(sim.gpuBuffers(), null);
});

run();
