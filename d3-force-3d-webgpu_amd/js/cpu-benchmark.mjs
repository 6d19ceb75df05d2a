// Times forceSimulationGPU's CPU path -- d3-force-3d's physics in FP64 on one thread, what
// the simulation runs when no GPU can take it -- the way the reference's benchmark page
// times d3-force-3d's tick (examples/benchmark.html:105-117): one untimed tick(1), then
// timed tick(1) calls, at least `min` of them, until `budget` seconds have gone.
//
//   node cpu-benchmark.mjs <dir>
//
// <dir> holds the workload: case.json {n, dims, forces: [names in order], min, budget,
// ticks (a fixed count instead of the budget)}, links.u32 (source, target pairs),
// radii.f64 (collide radii), pins.f64 (index, fx, fy, fz per pinned node); positions
// start at d3's phyllotaxis.  Prints one JSON line: the mean ms per tick, the ticks
// timed, and the host (CPU model, os.cpus().length, Node version).
import fs from "fs";
import os from "os";
import path from "path";
import {performance} from "perf_hooks";
import * as d3 from "./index.js";
import {useBackendForTests} from "./simulation.js";

const dir = process.argv[2];
const c = JSON.parse(fs.readFileSync(path.join(dir, "case.json"), "utf8"));
const read = (name, Type) => {
  const file = path.join(dir, name);
  if (!fs.existsSync(file)) return null;
  const b = fs.readFileSync(file);
  return new Type(b.buffer, b.byteOffset, b.byteLength / Type.BYTES_PER_ELEMENT);
};
const ends = read("links.u32", Uint32Array), radii = read("radii.f64", Float64Array), pins = read("pins.f64", Float64Array);

const nodes = Array.from({length: c.n}, () => ({}));
if (pins) {
  for (let k = 0; k < pins.length; k += 4) {
    const node = nodes[pins[k]];
    node.fx = pins[k + 1], node.fy = pins[k + 2], node.fz = pins[k + 3];
  }
}
console.warn = () => {};
useBackendForTests({available: () => false, create: () => null});  // the CPU path, whatever the host has
const sim = d3.forceSimulationGPU(nodes, c.dims).stop();
useBackendForTests(null);
for (const name of c.forces) {
  if (name === "link") {
    const links = [];
    for (let k = 0; k < ends.length; k += 2) links.push({source: ends[k], target: ends[k + 1]});
    sim.force("link", d3.forceLink(links));
  } else if (name === "manyBody") {
    sim.force("charge", d3.forceManyBody());
  } else if (name === "center") {
    sim.force("center", d3.forceCenter());
  } else if (name === "collide") {
    sim.force("collide", d3.forceCollide((d) => radii[d.index]).iterations(c.collideIterations || 1));
  } else if (name === "radial") {
    sim.force("radial", d3.forceRadial(c.radialRadius).strength(c.radialStrength));
  }
}

sim.tick(1);  // untimed, as the reference's page does
let ticks = 0;
const t0 = performance.now();
if (c.ticks) {
  for (; ticks < c.ticks; ++ticks) sim.tick(1);
} else {
  while (ticks < (c.min || 3) || performance.now() - t0 < 1e3 * (c.budget || 1.5)) sim.tick(1), ++ticks;
}
const ms = (performance.now() - t0) / ticks;
const cpus = os.cpus();
console.log(JSON.stringify({
  ms_per_tick: ms, ticks_timed: ticks, finite: nodes.every((d) => isFinite(d.x + d.y + (d.z || 0))),
  cpu_model: cpus.length ? cpus[0].model : "unknown", os_cpus: cpus.length, node: process.version
}));
