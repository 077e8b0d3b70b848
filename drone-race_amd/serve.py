"""ZMQ endpoint of the engine: the drop-in for the reference's C++ executable (src/main.cpp).

    python drone-race_amd/serve.py [tcp://*:5555] [--max-steps 500]

Binds a ZMQ PAIR socket (main.cpp:77-83), serves one simulation session (runEngine,
main.cpp:9-44) to the front-end (python-src/main.py + protocol.py's ZMQRecv) and exits.
Needs pyzmq; without it this entry point stops with an error.
"""
import argparse
import importlib.util
import os
import sys
import time


def _package():
    name = "drone_race_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("addr", nargs="?", default="tcp://*:5555")
    ap.add_argument("--max-steps", type=int, default=500)
    args = ap.parse_args(argv)
    try:
        import zmq
    except ImportError as exc:
        raise SystemExit(f"serve.py needs pyzmq for the ZMQ socket ({exc}); drone_race_amd.wire.serve works with any socket object")
    pkg = _package()
    ctx = zmq.Context()
    sock = ctx.socket(zmq.PAIR)
    sock.setsockopt(zmq.LINGER, 0)
    print(f"Binding ZMQ to addr {args.addr}...", flush=True)
    sock.bind(args.addr)
    t0 = time.perf_counter()
    event = pkg.wire.serve(sock, max_steps=args.max_steps)
    print(f"Simulation took {1e3 * (time.perf_counter() - t0):.0f} ms; event {event}", flush=True)
    sock.close()
    ctx.term()


if __name__ == "__main__":
    main()
