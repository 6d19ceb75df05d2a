// The JavaScript host on a GPU: lifecycle, async stepping, edits, pins, events, buffers
// and determinism (test/gpu-simulation-test.js is the model).
import assert from "assert";
import * as d3 from "../../d3-force-3d-webgpu_amd/js/index.js";
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

it("gpuBuffers() describes the device-resident state", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(33, 3), 3).stop().force("charge", d3.forceManyBody());
  assert.strictEqual(sim.gpuBuffers(), null, "null until the GPU is ready");
  await sim.gpuReady();
  const b = sim.gpuBuffers();
  assert.strictEqual(b.count, 33);
  assert.strictEqual(b.stride, 16);
  assert.strictEqual(typeof b.positions, "bigint");
  assert(b.positions !== 0n && b.velocities !== 0n);
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

it("rejects a force it cannot run", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(5, 2), 2).stop();
  await sim.gpuReady();
  assert.throws(() => sim.force("custom", function() {}), /cannot run on the GPU/);
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

it("destroy() stops it for good", async function() {
  const sim = d3.forceSimulationGPU(makeNodes(5, 2), 2).stop().force("charge", d3.forceManyBody());
  await sim.tickAsync(1);
  sim.destroy();
  assert.strictEqual(sim.isGPUEnabled(), false);
  assert.throws(() => sim.tick());
});

run();
