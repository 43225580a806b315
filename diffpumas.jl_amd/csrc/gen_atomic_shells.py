"""Writes dp_atomic_shells.hpp: the per-element atomic shell occupancies and
binding energies (eV) of Geant4 10.7 G4AtomicShells, in the layout the
reference uses (src/AtomicShellData.jl with the per-Z start offsets of
src/Materials.jl:347-357), for the electronic stopping power's Fano density
effect.  Run once in a checkout that has the reference next to it:

    python gen_atomic_shells.py /root/reference/src > dp_atomic_shells.hpp
"""
import os
import re
import sys


def _array(text, name):
    m = re.search(name + r"\s*=\s*\w+\[(.*?)\]", text, re.S)
    return [v for v in re.split(r"[\s,]+", m.group(1).strip()) if v]


def main(src):
    data = open(os.path.join(src, "AtomicShellData.jl")).read()
    mats = open(os.path.join(src, "Materials.jl")).read()
    index = [int(v) for v in _array(mats, "_ATOMIC_SHELL_INDEX")]
    occ = [int(v) for v in _array(data, "_ATOMIC_SHELL_OCCUPANCY")]
    energy = _array(data, "_ATOMIC_SHELL_ENERGY")
    assert len(occ) == len(energy) == index[-1], (len(occ), len(energy), index[-1])
    out = ["// Generated by gen_atomic_shells.py: Geant4 10.7 G4AtomicShells data,",
           "// shells of element Z at [SHELL_INDEX[Z-1], SHELL_INDEX[Z]).  Do not edit.",
           "#pragma once", "#include <stdint.h>", "", "namespace dpshell {",
           f"constexpr int SHELL_N_Z = {len(index) - 1};",
           f"constexpr uint16_t SHELL_INDEX[{len(index)}] = {{"]
    for i in range(0, len(index), 12):
        out.append("    " + ", ".join(str(v) for v in index[i:i + 12]) + ",")
    out.append("};")
    out.append(f"constexpr uint8_t SHELL_OCCUPANCY[{len(occ)}] = {{")
    for i in range(0, len(occ), 18):
        out.append("    " + ", ".join(str(v) for v in occ[i:i + 18]) + ",")
    out.append("};")
    out.append(f"constexpr float SHELL_ENERGY_EV[{len(energy)}] = {{")
    for i in range(0, len(energy), 8):
        out.append("    " + ", ".join(v + ("f" if "." in v else ".0f") for v in energy[i:i + 8]) + ",")
    out.append("};")
    out.append("}  // namespace dpshell")
    print("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
