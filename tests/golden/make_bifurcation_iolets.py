"""Writes bifurcation_lores_iolets.json from the case's input.xml."""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from hemepure_amd.config import load_case  # noqa: E402


def main(xml_path):
    c = load_case(xml_path)

    def d(x):
        return dict(position=list(x.position), normal=list(x.normal), density_mean=x.density_mean,
                    density_amp=x.density_amp, period=x.period, phase=x.phase)

    out = {"source": "cases/bifurcation/bifurcation_lores/input.xml (iolets converted with UnitConverter)",
           "tau": c.tau, "rho0": c.rho0, "inlets": [d(x) for x in c.inlets], "outlets": [d(x) for x in c.outlets]}
    Path(__file__).with_name("bifurcation_lores_iolets.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1])
