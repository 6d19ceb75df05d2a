// The simulation's stepper: calls `callback` about every frame (16 ms) until stopped,
// like d3-timer's timer() does with requestAnimationFrame in a browser.
const FRAME = 16;

export default function timer(callback) {
  let handle = null, fn = callback;

  function loop() {
    handle = setTimeout(loop, FRAME);
    fn();
  }

  const t = {
    restart: function(cb) {
      if (cb) fn = cb;
      if (handle) clearTimeout(handle);
      handle = setTimeout(loop, 0);
        return t;
    },
    stop: function() {
      if (handle) clearTimeout(handle);
      handle = null;
      return t;
    }
  };
  return t.restart();
}
