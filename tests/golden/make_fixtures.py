"""Copy the reference's golden-image fixtures into tests/golden/.

The reference's golden-image regression (crates/ps-app/tests/golden.rs:
27-48) renders nine scene files at fixed times, exposures and camera pitches
and compares each frame with a PNG held in tests/golden/. This script copies
those inputs and expected outputs as data, for the scenes whose frames are
made of the ground, atmosphere and cloud passes and the post-process chain
only: the hail frame (the precipitation subsystem's particles, enabled for
that scene alone, golden.rs:96-102) is left out. The thunderstorm frame is
kept: besides this path it holds one thin lightning bolt drawn by the
lightning subsystem, a few hundred pixels. Run it in a tree where the reference is readable:

    python tests/golden/make_fixtures.py /root/reference
"""
import shutil
import sys
from pathlib import Path

# (scene, UTC time, EV100, camera pitch in degrees): golden.rs:36-48
SCENES = [
    ("clear_summer_noon", "2026-06-21T11:00:00Z", 14.0, 5.0),
    ("broken_cumulus_afternoon", "2026-05-10T14:30:00Z", 14.0, 5.0),
    ("overcast_drizzle", "2026-04-12T10:00:00Z", 14.0, 45.0),
    ("high_cirrus_sunset", "2026-09-22T17:30:00Z", 11.0, 5.0),
    ("winter_overcast_snow", "2026-01-08T12:00:00Z", 14.0, 60.0),
    ("twilight_civil", "2026-12-21T08:05:00Z", 8.0, 5.0),
    ("mountain_wave_clouds", "2026-03-15T13:00:00Z", 14.0, 5.0),
    ("thunderstorm", "2026-08-16T16:00:00Z", 14.0, 5.0),
]
COVERAGE_GRIDS = ["mountain_wave_lozenges_128x128.bin"]


def main(ref: Path) -> None:
    out = Path(__file__).resolve().parent
    (out / "scenes").mkdir(exist_ok=True)
    (out / "images").mkdir(exist_ok=True)
    for name, *_ in SCENES:
        shutil.copyfile(ref / "tests" / "scenes" / f"{name}.toml", out / "scenes" / f"{name}.toml")
        shutil.copyfile(ref / "tests" / "golden" / f"{name}.png", out / "images" / f"{name}.png")
    for g in COVERAGE_GRIDS:
        shutil.copyfile(ref / "tests" / "scenes" / "presets" / g, out / "scenes" / g)
    print(f"copied {len(SCENES)} scenes into {out}")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
