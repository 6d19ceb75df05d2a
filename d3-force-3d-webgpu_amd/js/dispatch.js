// A small event dispatcher with the call shape of d3-dispatch, for the simulation's
// "tick" and "end" events: on("tick.name", fn), on("tick") to read, call(type, that).
export default function dispatch(...types) {
  const listeners = {};
  types.forEach(function(t) { listeners[t] = new Map(); });

  function parse(typename) {
    let name = "", i = typename.indexOf(".");
    if (i >= 0) name = typename.slice(i + 1), typename = typename.slice(0, i);
    if (typename && !listeners[typename]) throw new Error("unknown type: " + typename);
    return {type: typename, name: name};
  }

  return {
    on: function(typename, callback) {
      const t = parse(typename);
      if (arguments.length < 2) return t.type ? listeners[t.type].get(t.name) : undefined;
      (t.type ? [t.type] : types).forEach(function(type) {
        if (callback == null) listeners[type].delete(t.name);
        else if (t.type) listeners[type].set(t.name, callback);
      });
      return this;
    },
    call: function(type, that, ...args) {
      listeners[type].forEach(function(fn) { fn.apply(that, args); });
    }
  };
}
