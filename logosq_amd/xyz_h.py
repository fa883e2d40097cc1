"""XYZ Heisenberg evolution harness: the dense backend of examples/xyz_h.rs.

Reads the same environment variables (XYZ_QUBITS, XYZ_STEPS, XYZ_DT, XYZ_JX,
XYZ_JY, XYZ_JZ, XYZ_FIELD, XYZ_BACKEND, XYZ_OUTPUT_FILE; xyz_h.rs:22-62,
146-149), starts from State::one_state, evolves with
xyz_heisenberg::create_circuit and writes the same JSON record
(xyz_h.rs:107-143): initial and final calculate_energy_efficient energies,
their difference, the execution time and the operation count.  Only the
dense backend exists here (XYZ_BACKEND=mps is refused).

    python -m logosq_amd.xyz_h
"""
import json
import os
import sys
import time


def _env(name, cast, default):
    v = os.environ.get(name)
    if v is None:
        return default
    try:
        return cast(v)
    except ValueError:
        return default


def run(num_qubits=8, time_steps=10, dt=0.1, jx=1.0, jy=1.0, jz=1.0, external_field=0.0, dtype=None):
    """One dense run; returns the record xyz_h.rs writes."""
    from . import algorithms, states
    from ._lib import call
    params = algorithms.HeisenbergParameters(jx, jy, jz, external_field, time_steps, dt)
    circuit = algorithms.xyz_heisenberg_circuit(num_qubits, params)
    st = states.State.one_state(num_qubits, states.C128 if dtype is None else dtype)
    e0 = algorithms.xyz_energy_efficient(st, params)
    call("lq_synchronize")
    t0 = time.perf_counter()
    circuit.execute(st)
    call("lq_synchronize")
    ms = (time.perf_counter() - t0) * 1e3
    e1 = algorithms.xyz_energy_efficient(st, params)
    return {"framework": "logosq_amd (HIP, MI355X)", "backend": "dense", "qubits": num_qubits,
            "time_steps": time_steps, "dt": dt, "jx": jx, "jy": jy, "jz": jz, "external_field": external_field,
            "initial_energy": e0, "final_energy": e1, "energy_change": e1 - e0, "runtime_ms": round(ms, 2),
            "num_operations": circuit.num_operations()}


def main():
    backend = os.environ.get("XYZ_BACKEND", "dense").lower()
    if backend in ("mps", "tensor", "mps_backend"):
        print("XYZ_BACKEND=%s: only the dense state-vector backend is provided" % backend, file=sys.stderr)
        return 2
    rec = run(num_qubits=_env("XYZ_QUBITS", int, 8), time_steps=_env("XYZ_STEPS", int, 10),
              dt=_env("XYZ_DT", float, 0.1), jx=_env("XYZ_JX", float, 1.0), jy=_env("XYZ_JY", float, 1.0),
              jz=_env("XYZ_JZ", float, 1.0), external_field=_env("XYZ_FIELD", float, 0.0))
    out = os.environ.get("XYZ_OUTPUT_FILE", "logosq_xyz_heisenberg.json")
    with open(out, "w") as f:
        json.dump(rec, f, indent=2)
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
