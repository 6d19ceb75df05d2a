// Loads the N-API addon (lib/d3fgpu.node, built by `make addon`) that wraps the HIP
// library.  There is no fallback: a missing build is an error.
import {createRequire} from "module";

const require = createRequire(import.meta.url);
let addon = null, loadError = null;
try {
  addon = require("../lib/d3fgpu.node");
} catch (error) {
  loadError = error;
}

export function native() {
  if (!addon) throw new Error("d3fgpu: the native backend is not built (lib/d3fgpu.node): " + (loadError && loadError.message));
  return addon;
}

export function deviceCount() {
  return native().deviceCount();
}
