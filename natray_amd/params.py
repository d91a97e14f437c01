"""The reference's scene-construction API: SimParams, Shot, BulletType, presets, set_kv.

Mirrors src/params.rs:1-269 and src/presets.rs:1-116 field for field (same names, defaults and
error behaviour), so a caller that builds a `SimParams` for the reference builds the same one
here. LookParams only carries the renderer's settings through `set_kv`.
"""
from __future__ import annotations

import copy
import dataclasses as dc
from enum import Enum


class Nose(Enum):
    ROUND = "Round"
    SPITZER = "Spitzer"
    FLAT = "Flat"
    SPHERE = "Sphere"


class Metal(Enum):
    LEAD = "Lead"
    COPPER = "Copper"
    STEEL = "Steel"


@dc.dataclass(frozen=True)
class BulletType:
    """src/params.rs:19-37"""

    name: str
    caliber_mm: float
    mass_g: float
    speed: float
    twist_in: float
    nose: Nose
    metal: Metal

    def spin_rpm(self) -> float:
        if self.twist_in <= 0.0:
            return 0.0
        return self.speed / (self.twist_in * 0.0254) * 60.0


# src/params.rs:39-47
LR22 = BulletType(".22 LR, 40 gr lead round nose", 5.7, 2.59, 330.0, 16.0, Nose.ROUND, Metal.LEAD)
LUGER9 = BulletType("9 mm Luger, 115 gr FMJ", 9.0, 7.45, 360.0, 10.0, Nose.ROUND, Metal.COPPER)
ACP45 = BulletType(".45 ACP, 230 gr FMJ", 11.5, 14.9, 255.0, 16.0, Nose.ROUND, Metal.COPPER)
M855 = BulletType("5.56 NATO, 62 gr M855", 5.7, 4.0, 940.0, 7.0, Nose.SPITZER, Metal.COPPER)
WIN308 = BulletType(".308 Win, 150 gr spitzer", 7.8, 9.72, 860.0, 10.0, Nose.SPITZER, Metal.COPPER)
WADCUTTER = BulletType(".177 air rifle, 8.4 gr wadcutter", 4.5, 0.54, 240.0, 16.0, Nose.FLAT, Metal.LEAD)
BB = BulletType(".177 steel BB", 4.4, 0.34, 150.0, 0.0, Nose.SPHERE, Metal.STEEL)
BULLETS = [LR22, LUGER9, ACP45, M855, WIN308, WADCUTTER, BB]

MAX_SHOTS = 4  # src/params.rs:49


@dc.dataclass
class Shot:
    """src/params.rs:51-110"""

    bullet: BulletType = LR22
    speed: float = LR22.speed
    caliber_mm: float = LR22.caliber_mm
    mass_g: float = LR22.mass_g
    spin_rpm: float = LR22.spin_rpm()
    range_m: float = 10.0
    muzzle_yaw_deg: float = 0.5
    azimuth_deg: float = 0.0
    elevation_deg: float = 0.0
    impact_offset_cm: float = 0.0
    impact_side_cm: float = 0.0
    delay_ms: float = 0.0

    @classmethod
    def with_bullet(cls, b: BulletType, **kw) -> "Shot":
        s = cls()
        s.set_bullet(b)
        for k, v in kw.items():
            setattr(s, k, v)
        return s

    def set_bullet(self, b: BulletType) -> None:
        self.bullet = b
        self.caliber_mm = b.caliber_mm
        self.mass_g = b.mass_g
        self.speed = b.speed
        self.spin_rpm = b.spin_rpm()


@dc.dataclass
class SimParams:
    """src/params.rs:112-160"""

    shots: list = dc.field(default_factory=lambda: [Shot()])
    balloon_diameter_cm: float = 16.0
    balloon_height: float = 0.55
    rubber_modulus_mpa: float = 0.42
    rubber_rest_diameter_cm: float = 6.0
    rubber_break_stretch: float = 7.0
    rubber_thickness_mm: float = 0.25
    gravity: float = 9.81
    spray_amount: float = 1.0
    mist_amount: float = 1.0
    resolution: int = 64

    def add_shot(self) -> None:
        """A copy of the last shot, a quarter turn further round (src/params.rs:134-140)."""
        s = copy.copy(self.shots[-1]) if self.shots else Shot()
        # Rust's % keeps the dividend's sign, like math.fmod.
        import math

        s.azimuth_deg = math.fmod(s.azimuth_deg + 270.0, 360.0) - 180.0
        self.shots.append(s)


@dc.dataclass
class LookParams:
    """src/params.rs:162-201 (renderer settings; carried only so set_kv behaves the same)."""

    balloon_color: tuple = (0.85, 0.06, 0.05)
    sun_intensity: float = 1.0
    sky_intensity: float = 1.0
    env_rotation_deg: float = 0.0
    exposure_ev: float = 0.0
    dof: bool = True
    f_stop: float = 4.0
    focal_mm: float = 50.0
    bloom: float = 0.04
    caustics: bool = True
    mist: bool = True
    render_scale: float = 1.0
    max_bounces: int = 8
    shutter_us: float = 20.0


class ParamError(ValueError):
    pass


def _parse(v: str, ty):
    try:
        if ty is bool:
            if v not in ("true", "false"):
                raise ValueError
            return v == "true"
        if ty is int:
            n = int(v)
            if n < 0:
                raise ValueError
            return n
        return ty(v)
    except ValueError:
        raise ParamError(f"bad value '{v}'") from None


_SHOT_KEYS = {
    "bullet_speed": "speed",
    "caliber": "caliber_mm",
    "mass": "mass_g",
    "spin": "spin_rpm",
    "range": "range_m",
    "yaw": "muzzle_yaw_deg",
    "azimuth": "azimuth_deg",
    "elevation": "elevation_deg",
    "impact_offset": "impact_offset_cm",
    "impact_side": "impact_side_cm",
    "delay": "delay_ms",
}
_SIM_KEYS = {
    "diameter": ("balloon_diameter_cm", float),
    "height": ("balloon_height", float),
    "modulus": ("rubber_modulus_mpa", float),
    "rest_diameter": ("rubber_rest_diameter_cm", float),
    "break_stretch": ("rubber_break_stretch", float),
    "thickness": ("rubber_thickness_mm", float),
    "gravity": ("gravity", float),
    "spray": ("spray_amount", float),
    "mist": ("mist_amount", float),
    "resolution": ("resolution", int),
}
_LOOK_KEYS = {
    "sun": ("sun_intensity", float),
    "sky": ("sky_intensity", float),
    "env_rotation": ("env_rotation_deg", float),
    "exposure": ("exposure_ev", float),
    "dof": ("dof", bool),
    "fstop": ("f_stop", float),
    "focal": ("focal_mm", float),
    "bloom": ("bloom", float),
    "caustics": ("caustics", bool),
    "mist_render": ("mist", bool),
    "shutter": ("shutter_us", float),
    "max_bounces": ("max_bounces", int),
}


def set_kv(sim: SimParams, look: LookParams, key: str, v: str) -> None:
    """Applies a `key=value` override (src/params.rs:207-269). Bullet keys go to the first
    shot, or to shot N with a `shotN.` prefix; `shots=N` sets how many there are."""
    idx, name = 0, key
    if key.startswith("shot") and "." in key[4:]:
        n, name = key[4:].split(".", 1)
        k = _parse(n, int)
        if k == 0:
            raise ParamError("shots count from 1")
        idx = k - 1
    if idx >= MAX_SHOTS:
        raise ParamError(f"at most {MAX_SHOTS} shots")
    count = _parse(v, int) if name == "shots" else max(len(sim.shots), idx + 1)
    if not 1 <= count <= MAX_SHOTS:
        raise ParamError(f"shots 1..{MAX_SHOTS}")
    del sim.shots[count:]
    while len(sim.shots) < count:
        sim.add_shot()
    s = sim.shots[idx]
    if name == "shots":
        return
    if name == "bullet":
        k = _parse(v, int)
        if k >= len(BULLETS):
            raise ParamError(f"bullet 0..{len(BULLETS) - 1}")
        s.set_bullet(BULLETS[k])
    elif name in _SHOT_KEYS:
        setattr(s, _SHOT_KEYS[name], _parse(v, float))
    elif name in _SIM_KEYS:
        attr, ty = _SIM_KEYS[name]
        setattr(sim, attr, _parse(v, ty))
    elif name in _LOOK_KEYS:
        attr, ty = _LOOK_KEYS[name]
        setattr(look, attr, _parse(v, ty))
    elif name == "color":
        c = [_parse(x, float) for x in v.split(",")]
        if len(c) != 3:
            raise ParamError("color needs r,g,b")
        look.balloon_color = tuple(c)
    else:
        raise ParamError(f"unknown parameter '{key}'")


# ---- presets (src/presets.rs:5-116) --------------------------------------------------------


@dc.dataclass
class View:
    yaw: float
    pitch: float
    dist: float


@dc.dataclass
class Preset:
    name: str
    about: str
    sim: SimParams
    look: LookParams
    view: View


def all_presets() -> list[Preset]:
    """src/presets.rs:20-108"""
    L = LookParams
    return [
        Preset("Backlit", "The default shot with the sun behind the balloon, so the spray and mist light up.",
               SimParams(), L(env_rotation_deg=95.0, bloom=0.08, exposure_ev=0.3), View(-35.0, 6.0, 0.9)),
        Preset("Tumbling rifle round",
               "A long rifle bullet turns sideways in a big, thick-walled balloon, so the exit blows out far wider than the entry.",
               SimParams(shots=[Shot.with_bullet(M855, speed=500.0)], balloon_diameter_cm=22.0,
                         rubber_rest_diameter_cm=8.0, rubber_thickness_mm=0.5),
               L(balloon_color=(0.05, 0.25, 0.9)), View(-20.0, 5.0, 2.0)),
        Preset("Crossfire", "Two bullets at right angles, at the same moment, one just above the other.",
               SimParams(shots=[Shot(impact_offset_cm=1.5), Shot(azimuth_deg=90.0, impact_offset_cm=-1.5)]),
               L(balloon_color=(0.95, 0.45, 0.05)), View(-45.0, 12.0, 1.3)),
        Preset("Four sides", "Four bullets half a millisecond apart, each from a quarter turn further round.",
               SimParams(shots=[Shot(azimuth_deg=a, delay_ms=t) for a, t in
                                [(0.0, 0.0), (90.0, 0.5), (-180.0, 1.0), (-90.0, 1.5)]]),
               L(balloon_color=(0.1, 0.7, 0.2)), View(-35.0, 20.0, 1.2)),
        Preset("Moon gravity", "One sixth of Earth's gravity: the water hangs in the air and comes apart slowly.",
               SimParams(gravity=1.62, balloon_diameter_cm=20.0), L(balloon_color=(0.95, 0.8, 0.05)),
               View(-35.0, 4.0, 1.1)),
        Preset("Weightless", "No gravity. The balloon is a sphere and the water flies apart without falling.",
               SimParams(gravity=0.0), L(balloon_color=(0.55, 0.1, 0.8)), View(-35.0, 8.0, 1.0)),
        Preset("Slow pellet",
               "A slow airgun pellet into a small balloon filled far past its size. The hole tears all the way round and the water keeps the balloon's shape for a moment.",
               SimParams(shots=[Shot.with_bullet(WADCUTTER, speed=120.0)], balloon_diameter_cm=12.0,
                         rubber_rest_diameter_cm=3.2),
               L(balloon_color=(0.9, 0.1, 0.45)), View(-35.0, 8.0, 0.8)),
        Preset("Bounce", "A steel BB at 40 m/s against a thick, stiff balloon. The wall gives a millimetre and throws it back.",
               SimParams(shots=[Shot.with_bullet(BB, speed=40.0)], rubber_thickness_mm=1.5, rubber_modulus_mpa=4.0),
               L(balloon_color=(0.05, 0.55, 0.55)), View(-150.0, 6.0, 0.5)),
    ]


def named(name: str) -> Preset:
    """src/presets.rs:110-116"""
    presets = all_presets()
    for p in presets:
        if p.name.lower() == name.lower():
            return p
    names = ", ".join(p.name for p in presets)
    raise ParamError(f"no preset '{name}', try one of: {names}")
