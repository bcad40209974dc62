"""Reads a HemeLB case (input.xml + .gmy) into the arguments of hlb_create.

Covers the parts of SimConfig (src/configuration/SimConfig.cc) that feed the
collide-and-stream path: the time step and voxel size (relaxation time,
LbmParameters.h:28-37), the geometry file, the initial pressure, and cosine
pressure / parabolic velocity iolets with their position and normal.
Pressures convert with UnitConverter (src/util/UnitConverter.cc:15-45):
lattice pressure = Cs2 + dp[mmHg] * 133.3223874 / (1000 dx^2 / dt^2), and the
iolet densities are pressure / Cs2 (InOutLetCosine.h:64-80).  The XML is read
with xml.etree (data only).
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

from . import COSINE, PARABOLIC, Geometry, iolet_desc, params

CS2 = 1.0 / 3.0
MMHG_TO_PASCAL = 133.3223874         # src/constants.h:24
BLOOD_DENSITY = 1000.0               # src/constants.h:25
BLOOD_VISCOSITY = 0.004              # src/constants.h:27
REFERENCE_PRESSURE_MMHG = 0.0        # src/constants.h:23


def _value(el, name, units=None):
    child = el.find(name)
    if child is None:
        raise ValueError(f"<{el.tag}> has no <{name}>")
    if units is not None and child.get("units") not in (units, None):
        raise ValueError(f"<{name}> in {child.get('units')}, expected {units}")
    return child.get("value")


def _vector(text: str):
    nums = re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", text)
    if len(nums) != 3:
        raise ValueError(f"not a 3-vector: {text}")
    return tuple(float(x) for x in nums)


@dataclass
class UnitConverter:
    """UnitConverter (src/util/UnitConverter.cc:15-45)."""
    time_step: float
    voxel_size: float

    @property
    def lattice_pressure(self) -> float:
        mass = BLOOD_DENSITY * self.voxel_size ** 3
        return mass / (self.voxel_size * self.time_step ** 2)

    def pressure_difference(self, mmhg: float) -> float:
        return mmhg * MMHG_TO_PASCAL / self.lattice_pressure

    def pressure(self, mmhg: float) -> float:
        return CS2 + self.pressure_difference(mmhg - REFERENCE_PRESSURE_MMHG)

    def time(self, seconds: float) -> float:
        return seconds / self.time_step

    def speed(self, m_per_s: float) -> float:
        return m_per_s / (self.voxel_size / self.time_step)


@dataclass
class CaseConfig:
    geometry_path: Path
    steps: int
    tau: float
    rho0: float
    inlets: list = field(default_factory=list)
    outlets: list = field(default_factory=list)
    units: UnitConverter | None = None

    def geometry(self) -> Geometry:
        return Geometry.read_gmy(self.geometry_path, str(self.geometry_path).endswith("+"))

    def params(self, **kw):
        return params(tau=self.tau, **kw)


def _iolet(el, units: UnitConverter):
    cond = el.find("condition")
    kind, sub = cond.get("type"), cond.get("subtype")
    position = _vector(_value(el, "position", "lattice"))
    normal = _vector(_value(el, "normal", "dimensionless"))
    if kind == "pressure" and sub == "cosine":
        # SimConfig.cc:725-754: amplitude is a difference, mean absolute.
        amp = units.pressure_difference(float(_value(cond, "amplitude", "mmHg"))) / CS2
        mean = units.pressure(float(_value(cond, "mean", "mmHg"))) / CS2
        phase = float(_value(cond, "phase", "rad"))
        period = units.time(float(_value(cond, "period", "s")))
        return iolet_desc(COSINE, position, normal, density_mean=mean, density_amp=amp, phase=phase,
                          period=period)
    if kind == "velocity" and sub == "parabolic":
        radius = float(_value(cond, "radius", "lattice"))
        vmax = units.speed(float(_value(cond, "maximum", "m/s")))
        return iolet_desc(PARABOLIC, position, normal, radius=radius, max_speed=vmax)
    raise ValueError(f"iolet condition {kind}/{sub} is not supported by this reader")


def load_case(xml_path) -> CaseConfig:
    """The collide-and-stream inputs of a HemeLB input.xml."""
    xml_path = Path(xml_path)
    root = ET.parse(xml_path).getroot()
    sim = root.find("simulation")
    dt = float(_value(sim, "step_length", "s"))
    dx = float(_value(sim, "voxel_size", "m"))
    steps = int(float(_value(sim, "steps", "lattice")))
    units = UnitConverter(dt, dx)
    tau = 0.5 + (dt * BLOOD_VISCOSITY / BLOOD_DENSITY) / (CS2 * dx * dx)
    gmy = root.find("geometry").find("datafile").get("path")
    rho0 = 1.0
    ic = root.find("initialconditions")
    if ic is not None and ic.find("pressure") is not None:
        uni = ic.find("pressure").find("uniform")
        if uni is not None:
            rho0 = units.pressure(float(uni.get("value"))) / CS2
    inlets = [_iolet(e, units) for e in root.find("inlets").findall("inlet")]
    outlets = [_iolet(e, units) for e in root.find("outlets").findall("outlet")]
    return CaseConfig(xml_path.parent / gmy, steps, tau, rho0, inlets, outlets, units)
