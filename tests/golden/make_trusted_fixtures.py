"""Extract the reference's trusted 30-digit expression values into JSON fixtures.

The reference stores, next to each equation module, a ``tests/<name>.py`` file
holding ``trusted_dict = {"key": mpf("..."), ...}``: values of each
expression with every free symbol set from ``md5(symbol name)``
(``nrpy/validate_expressions/validate_expressions.py:368-400``).  This script
reads those files as text (regular expressions only; nothing is imported or
executed) and writes ``tests/golden/trusted/<name>.json`` as
``{"key": "<decimal string>"}``; complex entries store ``[real, imag]``.

Run in a container that has the reference checked out at /root/reference:

    python tests/golden/make_trusted_fixtures.py
"""

from __future__ import annotations

import json
import os
import re

REF = "/root/reference/nrpy"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trusted")

FILES = {
    # full BSSN RHS with KO, SinhSpherical, rfm precompute
    "rhs_eval_RbarSymbolic": "infrastructures/BHaH/general_relativity/tests/"
    "rhs_eval_OnePlusLog_GammaDriving2ndOrder_Covariant_SinhSpherical_RbargfsTrue_T4munuFalse_ImprovementsFalse.py",
    "rhs_eval_RbarExpanded": "infrastructures/BHaH/general_relativity/tests/"
    "rhs_eval_OnePlusLog_GammaDriving2ndOrder_Covariant_SinhSpherical_RbargfsFalse_T4munuFalse_ImprovementsFalse.py",
    "BSSN_quantities_SinhSpherical_rfm_precompute": "equations/general_relativity/tests/"
    "BSSN_quantities_SinhSpherical_rfm_precompute.py",
    "BSSN_constraints_SinhSpherical_rfm_precompute": "equations/general_relativity/tests/"
    "BSSN_constraints_SinhSpherical_rfm_precompute.py",
    "BSSN_gauge_RHSs_SinhSpherical_rfm_precompute": "equations/general_relativity/tests/"
    "BSSN_gauge_RHSs_OnePlusLog_GammaDriving2ndOrder_Covariant_SinhSpherical_rfm_precompute.py",
    "InitialData_Cartesian_BrillLindquist": "equations/general_relativity/tests/"
    "InitialData_Cartesian_BrillLindquist.py",
    "WaveEquation_RHSs": "equations/wave_equation/tests/WaveEquation_RHSs_WaveEquation.py",
}

_MPF = re.compile(r'"([^"]+)":\s*mpf\("([^"]+)"\)')
_MPC = re.compile(r'"([^"]+)":\s*mpc\(\s*real="([^"]+)",\s*imag="([^"]+)"\s*\)', re.S)


def parse(text: str) -> dict:
    """Return {key: value-string or [re, im]} from a trusted_dict file's text."""
    out: dict = {}
    for k, v in _MPF.findall(text):
        out[k] = v
    for k, re_, im in _MPC.findall(text):
        out[k] = [re_, im]
    return out


def main() -> None:
    os.makedirs(OUT, exist_ok=True)
    for name, rel in FILES.items():
        path = os.path.join(REF, rel)
        if not os.path.exists(path):
            print(f"skip (absent): {rel}")
            continue
        with open(path, encoding="utf-8") as f:
            data = parse(f.read())
        with open(os.path.join(OUT, name + ".json"), "w", encoding="utf-8") as f:
            json.dump({"source": "nrpy/" + rel, "values": data}, f, indent=1, sort_keys=True)
        print(f"{name}: {len(data)} values")


if __name__ == "__main__":
    main()
