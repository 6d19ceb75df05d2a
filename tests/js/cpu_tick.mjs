// Runs forceSimulationGPU on its CPU path (no GPU visible) for the case in argv[2] (JSON:
// n, dims, ticks, positions, pins, links) with link + manyBody + collide + center + radial
// + x, and writes the final positions and velocities to argv[3] (JSON).  tests/test_js_host.py
// compares them with the C oracle's d3 tick.
import fs from "fs";
import * as d3 from "../../d3-force-3d-webgpu_amd/js/index.js";
import {useBackendForTests} from "../../d3-force-3d-webgpu_amd/js/simulation.js";

const c = JSON.parse(fs.readFileSync(process.argv[2], "utf8"));
const nodes = c.positions.map((p) => {
  const node = {x: p[0]};
  if (c.dims > 1) node.y = p[1];
  if (c.dims > 2) node.z = p[2];
  return node;
});
for (const [i, fx, fy, fz] of c.pins) {
  nodes[i].fx = fx;
  if (c.dims > 1) nodes[i].fy = fy;
  if (c.dims > 2) nodes[i].fz = fz;
}
console.warn = () => {};
useBackendForTests({available: () => false, create: () => null});
const sim = d3.forceSimulationGPU(nodes, c.dims).stop();
useBackendForTests(null);
sim.force("link", d3.forceLink(c.links.map(([s, t]) => ({source: s, target: t}))))
   .force("charge", d3.forceManyBody())
   .force("collide", d3.forceCollide((d, i) => 1 + (i % 5)))
   .force("center", d3.forceCenter())
   .force("radial", d3.forceRadial(80).strength(0.05))
   .force("x", d3.forceX(10).strength(0.02));
sim.tick(c.ticks);
fs.writeFileSync(process.argv[3], JSON.stringify({
  pos: nodes.map((n) => [n.x, n.y || 0, n.z || 0]),
  vel: nodes.map((n) => [n.vx, n.vy || 0, n.vz || 0]),
  alpha: sim.alpha()
}));
