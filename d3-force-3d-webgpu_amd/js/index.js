// d3-force-3d's force API with the simulation running on an AMD GPU (HIP backend).
// The exports of the reference (src/index.js:1-13) that belong to the GPU path.
export {forceCenter, forceCollide, forceLink, forceManyBody, forceRadial, forceX, forceY, forceZ} from "./forces.js";
export {default as forceSimulationGPU, isBackendAvailable} from "./simulation.js";
export {isBackendAvailable as isWebGPUAvailable} from "./simulation.js";
// Beyond the reference: many layouts per GPU (a batch, and a serving loop).
export {createLayoutBatch, createLayoutServer} from "./batch.js";

// Support probe with the shape of the reference's checkWebGPUSupport (src/gpu/device.js).
import {isBackendAvailable} from "./simulation.js";
export function checkWebGPUSupport() {
  var ok = isBackendAvailable();
  return Promise.resolve({supported: ok, reason: ok ? null : "no GPU visible to the HIP backend"});
}
