// The simulation's default random source: a linear congruential generator with the
// constants and seed of d3-force-3d (src/lcg.js:1-9): s = (1664525 s + 1013904223) mod 2^32.
const a = 1664525;
const c = 1013904223;
const m = 4294967296;

export default function() {
  let s = 1;
  return () => (s = (a * s + c) % m) / m;
}
