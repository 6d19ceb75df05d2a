// The GPU engine of forceSimulationGPU: one HIP context (through the N-API addon) holding
// the simulation's nodes, with the behaviour of the reference engine (src/gpu/engine.js):
//
//   * while the simulation runs, the device holds the truth for positions and velocities;
//   * the JS nodes follow by readbacks, one in flight at a time; a readback that an upload
//     overtook is dropped (the C layer's epoch, fg_readback_fetch);
//   * x/y/z or fx/fy/fz set on the JS side between submissions are sent up before the next
//     one, and a readback never overwrites a node moved in the meantime.
//
// How: the nodes are packed into one Float64Array (addon.cc: x, y, z, vx, vy, vz, fx, fy,
// fz per node) and the addon does the rest in native code -- it keeps the shadow of what
// the device was told, uploads the runs of nodes that differ from it (push) and merges
// readbacks into the packed state (land) -- so the JS side only packs and unpacks.
// Force arrays cross to the device only when a force's `version` changes.
import {native} from "./native.js";

const STATE = 9;
const TYPES = {center: 1, collide: 2, link: 3, manyBody: 4, position: 5, radial: 6};

function finite(v) {
  v = +v;
  return v === v ? v : 0; // undefined and NaN count as 0, as the reference's floats() does
}

function column(values, count) {
  const out = new Float64Array(count);
  if (values) for (let i = 0; i < count; ++i) out[i] = +values[i];
  return out;
}

export function describable(force) {
  if (typeof force.gpu !== "function") return false;
  const d = force.gpu();
  return !!(d && TYPES[d.type]);
}

// The packed state of `nodes` (positions, velocities, pins) into `state`.
function pack(nodes, nDim, state) {
  for (let i = 0, n = nodes.length; i < n; ++i) {
    const node = nodes[i], o = STATE * i;
    state[o] = finite(node.x);
    state[o + 1] = nDim > 1 ? finite(node.y) : 0;
    state[o + 2] = nDim > 2 ? finite(node.z) : 0;
    state[o + 3] = finite(node.vx);
    state[o + 4] = nDim > 1 ? finite(node.vy) : 0;
    state[o + 5] = nDim > 2 ? finite(node.vz) : 0;
    state[o + 6] = node.fx == null ? NaN : +node.fx;
    state[o + 7] = nDim > 1 && node.fy != null ? +node.fy : NaN;
    state[o + 8] = nDim > 2 && node.fz != null ? +node.fz : NaN;
  }
}

// Positions and velocities from the packed state back into `nodes`.
function unpack(nodes, nDim, state) {
  for (let i = 0, n = nodes.length; i < n; ++i) {
    const node = nodes[i], o = STATE * i;
    node.x = state[o];
    node.vx = state[o + 3];
    if (nDim > 1) node.y = state[o + 1], node.vy = state[o + 4];
    if (nDim > 2) node.z = state[o + 2], node.vz = state[o + 5];
  }
}

export default class GpuEngine {
  constructor(device) {
    this.addon = native();
    this.handle = this.addon.create(device || 0);
    this.nodes = [];
    this.nDim = 2;
    this.state = new Float64Array(0);
    this.incoming = null;        // {pos, vel} buffers the readbacks land in
    this.forces = null;
    this.ids = new WeakMap();    // force object -> id for the backend's pass cache
    this.nextId = 1;
    this.sent = new Map();       // id -> version whose arrays the device has
    this.submitted = 0;          // steps queued since the nodes were set
    this.current = 0;            // steps the JS nodes reflect
    this.inflight = null;
    this.reload = false;         // the JS side moved every node: upload all before the next step
    this.alive = true;
  }

  get count() {
    return this.nodes.length;
  }

  fits(count) {
    return this.addon.fits(this.handle, count, 3);
  }

  setNodes(nodes, nDim) {
    this.nodes = nodes;
    this.nDim = nDim;
    this.state = new Float64Array(STATE * nodes.length);
    this.incoming = null;
    this.upload();
    return this;
  }

  upload() {
    pack(this.nodes, this.nDim, this.state);
    this.addon.load(this.handle, this.state, this.nDim);
    this.sent.clear(); // fresh node buffers: every force's arrays go up again
    this.reload = false;
    this.current = this.submitted;
  }

  // The next submission uploads every node first (the CPU moved them).
  markStale() {
    this.reload = true;
  }

  setSeed(random) {
    this.addon.setSeed(this.handle, Math.floor(random() * 4294967296) >>> 0);
  }

  // The pass list, in application order; false when a force has no GPU description.
  setForces(forces) {
    if (!forces.every(describable)) return false;
    this.forces = forces;
    this.addon.setForces(this.handle, this.describe(forces));
    return true;
  }

  idOf(force) {
    let id = this.ids.get(force);
    if (!id) this.ids.set(force, id = this.nextId++);
    return id;
  }

  describe(forces) {
    const n = this.nodes.length;
    return forces.map((force) => {
      const d = force.gpu(), id = this.idOf(force), version = d.version || 0;
      const out = {
        type: TYPES[d.type], id: id, version: version, iterations: d.iterations == null ? 1 : d.iterations,
        x: finite(d.x), y: finite(d.y), z: finite(d.z), strength: d.strength == null ? 1 : +d.strength,
        theta2: d.theta2 == null ? 0.81 : +d.theta2, distanceMin2: d.distanceMin2 == null ? 1 : +d.distanceMin2,
        distanceMax2: d.distanceMax2 == null ? Infinity : +d.distanceMax2, axis: d.axis | 0
      };
      if (this.sent.get(id) === version) return out; // the device has this version's arrays
      this.sent.set(id, version);
      switch (d.type) {
        case "manyBody": out.strengths = column(d.strengths, n); break;
        case "collide": out.radii = column(d.radii, n); break;
        case "radial": out.radii = column(d.radiuses, n), out.strengths = column(d.strengths, n); break;
        case "position": out.targets = column(d.targets, n), out.strengths = column(d.strengths, n); break;
        case "link": {
          const m = d.links.length, ends = new Uint32Array(2 * m);
          d.links.forEach((link, k) => { ends[2 * k] = link.source.index; ends[2 * k + 1] = link.target.index; });
          out.links = ends;
          out.strengths = column(d.strengths, m);
          out.distances = column(d.distances, m);
          out.bias = column(d.bias, m);
          break;
        }
      }
      return out;
    });
  }

  // Queue alphas.length steps; decay multiplies the velocities (1 - velocityDecay).
  enqueue(alphas, decay) {
    if (!this.alive || !this.nodes.length) return;
    if (this.reload) {
      this.upload();
    } else {
      pack(this.nodes, this.nDim, this.state);
      this.addon.push(this.handle, this.state);
    }
    if (this.forces) this.addon.setForces(this.handle, this.describe(this.forces));
    this.addon.enqueue(this.handle, Float64Array.from(alphas), decay);
    this.submitted += alphas.length;
  }

  busy() {
    return this.alive && this.addon.busy(this.handle);
  }

  // One readback: resolves once it has landed in the nodes (or was dropped as stale).
  readback() {
    if (this.inflight) return this.inflight;
    if (!this.alive || !this.nodes.length || this.current === this.submitted) return Promise.resolve();
    const n = this.nodes.length, target = this.submitted;
    if (!this.incoming || this.incoming.pos.length !== 3 * n) {
      this.incoming = {pos: new Float64Array(3 * n), vel: new Float64Array(3 * n)};
    }
    const buf = this.incoming, nodes = this.nodes;
    this.inflight = this.addon.readbackAsync(this.handle, buf.pos, buf.vel).then((fresh) => {
      this.inflight = null;
      if (!fresh || !this.alive || nodes !== this.nodes) return;
      pack(nodes, this.nDim, this.state);
      this.addon.land(this.handle, this.state, buf.pos, buf.vel);
      unpack(nodes, this.nDim, this.state);
      this.current = target;
    }, (error) => {
      this.inflight = null;
      if (this.alive && nodes === this.nodes) throw error;
    });
    return this.inflight;
  }

  // Resolves once the nodes reflect every queued step.
  sync() {
    const more = () => this.readback().then(() => {
      if (this.alive && this.nodes.length && this.current !== this.submitted) return more();
    });
    return more();
  }

  // Resolves once the device has run every queued step (the nodes are left alone).
  idle() {
    return this.alive ? this.addon.idleAsync(this.handle).then(() => undefined) : Promise.resolve();
  }

  buffers() {
    return this.alive && this.nodes.length ? this.addon.buffers(this.handle) : null;
  }

  stats() {
    return this.addon.stats(this.handle);
  }

  // A copy of `floats` floats of device memory at `pointer` (a gpuBuffers() pointer).
  readDevice(pointer, floats) {
    return this.addon.readDevice(this.handle, pointer, floats);
  }

  destroy() {
    if (!this.alive) return;
    this.alive = false;
    this.addon.destroy(this.handle);
    this.nodes = [];
    this.state = new Float64Array(0);
    this.incoming = null;
  }
}
