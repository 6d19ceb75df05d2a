// A minimal test runner for the Node host tests (mocha is not available here): it()
// registers async tests, run() executes them in order and exits non-zero on a failure.
const tests = [];
export function it(name, fn) { tests.push({name, fn}); }
export async function run() {
  let failed = 0;
  for (const t of tests) {
    try {
      await t.fn();
      console.log("ok   " + t.name);
    } catch (error) {
      failed++;
      console.log("FAIL " + t.name + "\n     " + (error && error.stack || error));
    }
  }
  console.log(`${tests.length - failed} passed, ${failed} failed`);
  process.exit(failed ? 1 : 0);
}
