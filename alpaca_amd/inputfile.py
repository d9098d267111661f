"""Run ALPACA XML input files on the GPU path (single-phase, uniform level).

Reads the subset of the reference's input format that the single-phase
uniform path uses (input_output/input_reader/*: ``domain/nodeSize``,
``domain/nodeRatio``, ``domain/boundaryConditions/material``,
``domain/initialConditions/material1``, ``materials`` with one stiffened gas,
``timeControl``), places the initial prime states at the cell centres the way
PrimeStateInitializer::GetInitialPrimeStates does
(prime_state_initializer.cpp:59-100: ``origin + (i + 0.5) * cell_size`` with
the block origin summed over the node id's bits as in
DomainCoordinatesOfId, id_information.h:118-142), converts them with the
stiffened-gas energy and advances the case to the end time with the HIP path.

What the input can ask for and this path does not do is refused with an
error rather than ignored: more than one material, another equation of
state, gravity or viscous/heat-conduction properties, side conditions other
than ZeroGradient/Symmetry/Periodic or different conditions on the two sides
of a direction, non-unit dimensionalization references, a start time other
than zero.  Multiresolution is replaced by a uniform level: ``level`` L gives
``nodeRatio * 2^L`` blocks per direction with cell size
``nodeSize / 16 / 2^L`` (CellSizeOfLevel, id_information.h:73-77).

    python -m alpaca_amd.inputfile SodX.xml --dim 2 --flux hllc --out sod.npz
    python -m alpaca_amd.inputfile SodX.xml --out sod.h5   # reference HDF5 + XDMF layout
    python -m alpaca_amd.inputfile SodX.xml --restore SodX/restart/restart_0.100000.h5

As the reference does, a run writes into a folder named after the input file
(``--output-folder`` overrides it): the standard output the input's
``<output><standardOutput>`` asks for in ``domain/`` (data_<t>.h5 + .xdmf in
the reference's layout and data_time_series.xdmf), and the restart snapshots
its ``<restart>`` section asks for (restore mode and file, snapshot stamps /
wall-clock interval / intervals to keep) in ``restart/`` (alpaca_amd.restart).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import xml.etree.ElementTree as ET

import numpy as np

from . import cases, gpu
from .expression import Program

IC = 16
_BC_KINDS = {"zerogradient": gpu.BC_ZERO_GRADIENT, "symmetry": gpu.BC_SYMMETRY,
             "periodic": gpu.BC_PERIODIC}
_SIDES = (("west", "east"), ("south", "north"), ("bottom", "top"))
PRIME_NAMES = ("density", "velocityX", "velocityY", "velocityZ", "pressure")


class InputError(ValueError):
    pass


@dataclasses.dataclass
class InputCase:
    path: str
    dim: int
    node_size: float
    node_ratio: list
    bc: list
    gamma: float
    background_pressure: float
    ic_text: str
    end_time: float
    cfl: float
    maximum_level: int
    restore_mode: str = "off"          # restart/restore/mode: off, soft or forced
    restore_file: str = ""
    snapshot_stamps: list = dataclasses.field(default_factory=list)
    snapshot_interval: int = 0         # wall seconds, 0 = no interval snapshots
    snapshots_to_keep: int = 0
    time_naming_factor: float = 1.0
    output_times: list = dataclasses.field(default_factory=list)  # standard output


def _text(root, path, default=None):
    el = root.find(path)
    if el is None or el.text is None:
        if default is not None:
            return default
        raise InputError(f"missing <{path}>")
    return el.text.strip()


def _float(root, path, default=None):
    return float(_text(root, path, None if default is None else str(default)))


def read_input(path: str, dim: int | None = None) -> InputCase:
    """Parse an input file.  dim is the compile-time dimension of the
    reference build the file is meant for; when None it is 3 if the node
    ratio has a z entry above one, else 2."""
    root = ET.parse(path).getroot()
    ratio_el = root.find("domain/nodeRatio")
    if ratio_el is None:
        raise InputError("missing <domain/nodeRatio>")
    ratio = [int(_text(root, f"domain/nodeRatio/{a}", "1")) for a in "xyz"]
    if dim is None:
        dim = 3 if ratio[2] > 1 else 2
    if dim not in (2, 3):
        raise InputError("only 2D and 3D builds are supported")
    if dim == 2:
        ratio[2] = 1
    nmat = int(_text(root, "materials/numberOfMaterials", "1"))
    if nmat != 1:
        raise InputError(f"{nmat} materials: only single-phase inputs run on this path")
    eos_type = _text(root, "materials/material1/equationOfState/type")
    if eos_type != "StiffenedGas":
        raise InputError(f"equation of state {eos_type}: only StiffenedGas is implemented")
    gamma = _float(root, "materials/material1/equationOfState/gamma")
    bpress = _float(root, "materials/material1/equationOfState/backgroundPressure", 0.0)
    for prop in ("shearViscosity", "bulkViscosity", "thermalConductivity"):
        v = root.find(f"materials/material1/properties/{prop}")
        if v is not None and v.text and float(v.text) != 0.0:
            raise InputError(f"{prop} = {v.text.strip()}: viscous and heat terms are off on "
                             "this path (Euler)")
    grav = root.find("sourceTerms/gravity")
    if grav is not None:
        for c in grav:
            if c.text and float(c.text) != 0.0:
                raise InputError("gravity source term is off on this path")
    bc = []
    for d in range(dim):
        kinds = []
        for side in _SIDES[d]:
            name = _text(root, f"domain/boundaryConditions/material/{side}").lower()
            if name not in _BC_KINDS:
                raise InputError(f"{side} boundary {name}: only ZeroGradient, Symmetry and "
                                 "Periodic are implemented")
            kinds.append(_BC_KINDS[name])
        if kinds[0] != kinds[1]:
            raise InputError(f"direction {'xyz'[d]}: both sides must carry the same condition")
        bc.append(kinds[0])
    while len(bc) < 3:
        bc.append(gpu.BC_ZERO_GRADIENT)
    dimz = root.find("dimensionalization")
    if dimz is not None:
        for c in dimz:
            if c.text and float(c.text) != 1.0:
                raise InputError("dimensionalization references other than 1 are not "
                                 "supported")
    if _float(root, "timeControl/startTime", 0.0) != 0.0:
        raise InputError("start time must be 0")
    end_time = _float(root, "timeControl/endTime")
    restore_mode = _text(root, "restart/restore/mode", "Off").replace(" ", "").lower()
    if restore_mode not in ("off", "soft", "forced"):
        raise InputError(f"restore mode {restore_mode} not known")
    restore_file = _text(root, "restart/restore/fileName", "") if restore_mode != "off" else ""
    snap_type = _text(root, "restart/snapshots/type", "Off").replace(" ", "").lower()
    if snap_type not in ("off", "interval", "stamps", "intervalstamps"):
        raise InputError(f"snapshot type {snap_type} not known")
    stamps = []
    if snap_type in ("stamps", "intervalstamps"):
        # ComputeSnapshotTimes (instantiation_input_output_manager.cpp:168-230):
        # <ts1>, <ts2>, ... inside [start, end], then the end time itself
        k = 1
        while root.find(f"restart/snapshots/stamps/ts{k}") is not None:
            stamps.append(_float(root, f"restart/snapshots/stamps/ts{k}"))
            k += 1
        stamps = [t for t in stamps if 0.0 <= t <= end_time] + [end_time]
        if stamps != sorted(stamps):
            raise InputError("snapshot timestamps must be ascending")
    interval = keep = 0
    if snap_type in ("interval", "intervalstamps"):
        interval = int(_text(root, "restart/snapshots/interval"))
        keep = int(_text(root, "restart/snapshots/intervalsToKeep"))
    out_type = _text(root, "output/standardOutput/type", "Off").replace(" ", "").lower()
    if out_type not in ("off", "interval", "stamps", "intervalstamps"):
        raise InputError(f"standard output type {out_type} not known")
    out_stamps = []
    k = 1
    while root.find(f"output/standardOutput/stamps/ts{k}") is not None:
        out_stamps.append(_float(root, f"output/standardOutput/stamps/ts{k}"))
        k += 1
    from . import h5output

    out_times = h5output.output_times(
        out_type, _float(root, "output/standardOutput/interval", 0.0) if out_type in
        ("interval", "intervalstamps") else 0.0, out_stamps, 0.0, end_time)
    return InputCase(
        path=path, dim=dim, node_size=_float(root, "domain/nodeSize"), node_ratio=ratio, bc=bc,
        gamma=gamma, background_pressure=bpress,
        ic_text=_text(root, "domain/initialConditions/material1"),
        end_time=end_time, cfl=_float(root, "timeControl/CFLNumber"),
        maximum_level=int(_text(root, "multiResolution/maximumLevel", "0")),
        restore_mode=restore_mode, restore_file=restore_file, snapshot_stamps=stamps,
        snapshot_interval=interval, snapshots_to_keep=keep,
        time_naming_factor=_float(root, "output/timeNamingFactor", 1.0), output_times=out_times)


def block_origin(index: int, block_size: float) -> float:
    """Origin of the block with integer index along one axis: the sum of
    bit_k * block_size * 2^k accumulated from the lowest bit up, the order
    DomainCoordinatesOfId (id_information.h:118-142) adds them in."""
    x = 0.0
    size = block_size
    while index:
        x += (index & 1) * size
        index >>= 1
        size *= 2.0
    return x


def cell_centres(case: InputCase, level: int = 0):
    """Per-axis cell-centre coordinates of the uniform level (1D arrays)."""
    block_size = case.node_size / float(1 << level)
    cell_size = case.node_size / float(IC) / float(1 << level)
    axes = []
    for d in range(3):
        nb = case.node_ratio[d] << level if d < case.dim else 1
        if d >= case.dim:
            axes.append(np.zeros(1))
            continue
        c = np.empty(nb * IC)
        for b in range(nb):
            o = block_origin(b, block_size)
            c[b * IC:(b + 1) * IC] = o + (np.arange(IC, dtype=np.float64) + 0.5) * cell_size
        axes.append(c)
    return axes, cell_size


def initial_state(case: InputCase, level: int = 0):
    """Conservatives [NE][NX][NY][NZ] of the initial condition."""
    (xs, ys, zs), _ = cell_centres(case, level)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    prog = Program(case.ic_text)
    env = prog.run({"x": X, "y": Y, "z": Z}, list(PRIME_NAMES))
    rho = env["density"]
    vel = (env["velocityX"], env["velocityY"], env["velocityZ"])
    p = env["pressure"]
    cons = cases.primes_to_conservatives(case.dim, case.gamma, case.background_pressure, rho,
                                         vel[:case.dim], p)
    if case.dim == 2:
        cons = cons[:, :, :, :1]
    return np.ascontiguousarray(cons)


def case_desc(case: InputCase, level: int = 0, flux: int = gpu.FLUX_HLLC,
              precision: int = gpu.FP64):
    nb = [r << level if d < case.dim else 1 for d, r in enumerate(case.node_ratio)]
    _, cell_size = cell_centres(case, level)
    return gpu.make_desc(case.dim, nb[:case.dim], case.bc[:case.dim], flux=flux,
                         precision=precision, gamma=case.gamma, b=case.background_pressure,
                         cell_size=cell_size)


def restore_path(case: InputCase, restore: str | None = None):
    """The restart file to start from, or None (RestoreSimulationFromSnapshot,
    input_output_manager.cpp:477-520): an explicit path must exist; from the
    input file, mode Soft starts fresh when the file is missing and Forced
    fails."""
    import os

    if restore:
        if not os.path.exists(restore):
            raise InputError(f"restart file {restore} does not exist")
        return restore
    if case.restore_mode == "off":
        return None
    if os.path.exists(case.restore_file):
        return case.restore_file
    if case.restore_mode == "forced":
        raise InputError("Force restore not possible: restart file "
                         f"{case.restore_file} does not exist")
    return None


def run(case: InputCase, level: int = 0, flux: int = gpu.FLUX_HLLC, precision: int = gpu.FP64,
        max_steps: int = 10 ** 9, restart_folder: str | None = None, restore: str | None = None,
        output_folder: str | None = None):
    """Advance the case to its end time on the GPU; returns a dict with the
    final conservatives, the primes, the step count, every dt, the time and
    the restart files written.

    Like ModularAlgorithmAssembler::Initialization and ComputeLoop
    (modular_algorithm_assembler.cpp:97-240): the state comes from a restart
    file when restore_path() names one (its time becomes the start time),
    else from the initial condition; with restart_folder set, snapshots are
    written there after initialisation and after each macro step as the
    input's <restart><snapshots> asks (restart.SnapshotSchedule), forced
    when the time step turns unhealthy; with output_folder, the standard
    output the input's <output><standardOutput> asks is written there
    (h5output.OutputSeries: density, velocity and pressure in the reference's
    HDF5 + XDMF layout, forced at the start and when the step turns
    unhealthy, plus the data_time_series.xdmf file)."""
    from . import restart as rs

    desc = case_desc(case, level, flux, precision)
    path = restore_path(case, restore)
    if (path or restart_folder) and level != 0:
        raise InputError("restart files hold level 0 only")
    time0 = 0.0
    if path:
        time0, fdim, fnb, blocks = rs.read(path, case.dim)
        want = [case.node_ratio[d] for d in range(case.dim)]
        if list(fnb) != want:
            raise InputError(f"restart file has {fnb} blocks, the input {want}")
        st0 = rs.domain_from_blocks(blocks, case.dim, fnb)
    else:
        st0 = initial_state(case, level)
    gc = gpu.GpuCase(desc, st0, cfl=case.cfl, t_end=case.end_time, time0=time0)
    outputs = None
    if output_folder is not None:
        from . import h5output

        outputs = h5output.OutputSeries(output_folder, case.output_times, case.time_naming_factor,
                                        restored_time=time0 if path else None)
        outputs.begin()
        _, cell_size = cell_centres(case, level)
        nb_out = [desc.nb[d] if d < case.dim else 1 for d in range(3)]

        def write_output(t, force=False):
            def one(p, series):
                pr = primes(case, gc.download())
                h5output.write(p, t, case.dim, nb_out, cell_size, pr["density"], pr["velocity"],
                               pr["pressure"], series=series)
            outputs.maybe_write(t, one, force)

        write_output(time0, force=True)
    schedule = None
    if restart_folder is not None:
        schedule = rs.SnapshotSchedule(restart_folder, case.snapshot_stamps,
                                       case.snapshot_interval, case.snapshots_to_keep,
                                       case.time_naming_factor,
                                       restored_time=time0 if path else None)

        def write(t, force=False):
            schedule.maybe_write(t, lambda f: rs.write(
                f, t, case.dim, desc.nb[:case.dim], gc.download_blocks(), case.gamma,
                case.background_pressure), force)

        write(time0)
    hooks = []
    if schedule is not None:   # a restart snapshot comes before the output
        hooks.append(lambda rec: write(rec.time, bool(rec.flags & 2)))
    if outputs is not None:
        hooks.append(lambda rec: write_output(rec.time, bool(rec.flags & 2)))

    def on_step(rec):
        for h in hooks:
            h(rec)

    steps, dts = gc.run(max_steps, on_step=on_step if hooks else None)
    if outputs is not None:
        outputs.end()
    rec = gc.record()
    cons = gc.download()
    return {"conservatives": cons, "steps": steps, "dts": np.array(dts), "time": rec.time,
            "start_time": time0, "flags": rec.flags,
            "restart_files": list(schedule.written) if schedule else [],
            "output_files": list(outputs.written) if outputs else [], **primes(case, cons)}


def primes(case: InputCase, cons):
    """Density, velocity and pressure of a conservative state (the
    ConservativesToPrimeStates arithmetic: u = m * (1/rho), stiffened-gas
    pressure)."""
    rho, vel, p = cases.conservatives_to_primes(case.dim, case.gamma, case.background_pressure,
                                                cons)
    return {"density": rho, "velocity": np.stack(vel), "pressure": p}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("input")
    ap.add_argument("--dim", type=int, default=None)
    ap.add_argument("--level", type=int, default=0, help="uniform level (multiresolution off)")
    ap.add_argument("--flux", choices=["hllc", "roe"], default="hllc")
    ap.add_argument("--precision", choices=["fp64", "mixed", "fp64-fma"], default="fp64")
    ap.add_argument("--max-steps", type=int, default=10 ** 9)
    ap.add_argument("--out", default=None,
                    help="write the final state: .npz (all arrays) or .h5 (the reference's "
                         "HDF5 + XDMF output layout)")
    ap.add_argument("--output-folder", default=None,
                    help="folder for the standard output (domain/) and restart snapshots "
                         "(restart/) the input file asks for; default: the input file's name "
                         "without extension, in the working directory, as the reference does")
    ap.add_argument("--restore", default=None,
                    help="start from this restart file (overrides <restart><restore>)")
    a = ap.parse_args(argv)
    case = read_input(a.input, a.dim)
    prec = {"fp64": gpu.FP64, "mixed": gpu.MIXED, "fp64-fma": gpu.FP64_FMA}[a.precision]
    import os

    base = a.output_folder or os.path.splitext(os.path.basename(a.input))[0]
    from . import h5output

    try:
        h5output._lib()
        have_hdf5 = True
    except h5output.Hdf5Error:
        have_hdf5 = False
    res = run(case, a.level, gpu.FLUX_HLLC if a.flux == "hllc" else gpu.FLUX_ROE, prec,
              a.max_steps,
              restart_folder=os.path.join(base, "restart") if a.level == 0 and have_hdf5 else None,
              restore=a.restore,
              output_folder=os.path.join(base, "domain") if have_hdf5 else None)
    if a.out and a.out.endswith(".h5"):
        from . import h5output

        nb = [r << a.level for r in case.node_ratio]
        _, cell_size = cell_centres(case, a.level)
        h5output.write(a.out[:-3], res["time"], case.dim, nb, cell_size, res["density"],
                       res["velocity"], res["pressure"])
    elif a.out:
        np.savez(a.out, **{k: v for k, v in res.items()})
    print(json.dumps({"input": a.input, "dim": case.dim, "level": a.level,
                      "blocks": [r << a.level for r in case.node_ratio[:case.dim]],
                      "steps": res["steps"], "start_time": res["start_time"],
                      "time": res["time"], "flags": res["flags"],
                      "restart_files": res["restart_files"],
                      "output_files": len(res["output_files"])}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
