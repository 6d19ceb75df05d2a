// Times forceSimulationGPU on the HIP backend through the Node host the way the reference's
// benchmark page times its GPU simulation (examples/benchmark.html:119-132): link
// (distance 30) + manyBody + center on the page's graph, three warm-up ticks, then
// await tickAsync(k) for k = 10, 20, 40, ... until one call takes over half the budget
// (1.5 s) or k reaches 640; ms per tick is that call's wall time over k.  A call's time
// holds its ticks, the readback of the nodes and their update in JavaScript.
//
//   node gpu-benchmark.mjs <n> [<n> ...]
//
// Prints one JSON line per size: n, links, ms_per_tick, ticks, node.
import {performance} from "perf_hooks";
import * as d3 from "./index.js";

// The page's graph (examples/benchmark.html:85-93): a random tree (node i > 0 joins a
// uniformly drawn earlier node) plus n / 2 random pairs, from one LCG seeded with 42.
function pageGraph(n) {
  let s = 42;
  const rnd = () => (s = (s * 1664525 + 1013904223) % 4294967296) / 4294967296;
  const nodes = Array.from({length: n}, () => ({})), links = [];
  for (let i = 1; i < n; ++i) links.push({source: i, target: Math.floor(rnd() * i)});
  for (let i = 0; i < n / 2; ++i) links.push({source: Math.floor(rnd() * n), target: Math.floor(rnd() * n)});
  return {nodes, links};
}

async function timeGPU(n, budget) {
  const graph = pageGraph(n);
  const sim = d3.forceSimulationGPU(graph.nodes, 3).stop();
  sim.force("link", d3.forceLink(graph.links).distance(30));
  sim.force("charge", d3.forceManyBody());
  sim.force("center", d3.forceCenter());
  if (!(await sim.gpuReady())) throw new Error("the simulation fell back to the CPU");
  await sim.tickAsync(3);  // warm-up
  for (let ticks = 10; ; ticks *= 2) {
    const t0 = performance.now();
    await sim.tickAsync(ticks);
    const dt = performance.now() - t0;
    if (dt > budget / 2 || ticks >= 640) {
      sim.destroy();
      return {n, links: graph.links.length, ms_per_tick: dt / ticks, ticks, node: process.version};
    }
  }
}

(async () => {
  for (const arg of process.argv.slice(2)) console.log(JSON.stringify(await timeGPU(Number(arg), 1500)));
})().catch((e) => {
  console.error(e && e.stack ? e.stack : String(e));
  process.exit(1);
});
