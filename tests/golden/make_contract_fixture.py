"""Copy the reference's committed Mobile V2 evidence scalars into a fixture.

Reads the reference's own output data files (JSON written by its contract
tools) and keeps only the numbers the parity tests compare against.  Run in
the development container where the reference is mounted:

    python tests/golden/make_contract_fixture.py /root/reference
"""

from __future__ import annotations

import json
import sys
from pathlib import Path


def main(root: str) -> None:
    ref = Path(root)
    contracts = json.loads((ref / "outputs/mobile_v2_reference/mobile_v2_cpu_contracts.json").read_text())
    gate = json.loads((ref / "outputs/mobile_v2_production/mobile_v2_production_gate_ab.json").read_text())
    cases = {}
    for name, case in contracts["cases"].items():
        if "energy" not in case:
            continue
        keep = {k: case[k] for k in ("substeps", "maximum_CFL", "mass_residual_m3",
                                      "momentum_accounting_residual_m4_s", "minimum_h_m")}
        keep["energy"] = case["energy"]
        if "low_to_forward_face_delta_q_m2_s" in case:
            keep["low_to_forward_face_delta_q_m2_s"] = case["low_to_forward_face_delta_q_m2_s"]
        cases[name] = keep
    fixture = {
        "source": "outputs/mobile_v2_reference/mobile_v2_cpu_contracts.json; "
                  "outputs/mobile_v2_production/mobile_v2_production_gate_ab.json",
        "contract_cases": cases,
        "entrainment_energy_j_per_density": contracts["cases"]["ENTRAINMENT"]["energy_transfer_j_per_density"],
        "deposition_energy_j_per_density": contracts["cases"]["DEPOSITION"]["energy_transfer_j_per_density"],
        "gate_a": {k: gate["gate_a"][k] for k in (
            "entrained_volume_m3", "deposited_volume_m3", "deposition_removed_momentum_kg_m_s",
            "entrainment_H_free_max_error_m", "deposition_H_free_max_error_m", "ledger_error_m3")},
    }
    out = Path(__file__).resolve().parent / "mobile_v2_cpu_contracts.json"
    out.write_text(json.dumps(fixture, indent=1) + "\n")
    print(out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
