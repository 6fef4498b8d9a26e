"""Writes hosek_rgb_dataset.json: the RGB coefficient and radiance datasets of
the Hosek-Wilkie sky model (version 1.4a), the numbers lrt_sky_cook
interpolates.

The datasets are third-party data published by Lukas Hosek and Alexander
Wilkie under the 3-clause BSD license; the JSON carries that license's notice
and attribution (field "license").  This script reads them from the
reference tree's copy of the model (third_party/hosek/ArHosekSkyModelData_RGB.h,
arrays datasetRGB1..3 and datasetRGBRad1..3) and stores each number as the
double the C compiler makes of its decimal literal (Python's float() and
repr() round-trip it exactly).

  python ray-tracing_amd/data/make_hosek_dataset.py [/root/reference]

Layout (as the model indexes it): per channel, coefficients
[albedo 0..1][turbidity 1..10][control point 0..5][parameter 0..8] (1080
doubles) and radiances [albedo][turbidity][control point] (120 doubles).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
NUM = re.compile(r"[-+]?\d+\.\d+[eE][-+]?\d+")


def _array(text: str, name: str) -> list[float]:
    m = re.search(r"double\s+" + re.escape(name) + r"\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise ValueError(f"{name} not found")
    body = "\n".join(line.split("//")[0] for line in m.group(1).splitlines())
    return [float(x) for x in NUM.findall(body)]


def _license(text: str) -> str:
    m = re.search(r"/\*\s*(This source is published under the following 3-clause BSD license\..*?)\*/", text, re.S)
    if not m:
        raise ValueError("license block not found")
    return m.group(1).strip()


def main(root: str = "/root/reference") -> None:
    src = Path(root) / "third_party" / "hosek" / "ArHosekSkyModelData_RGB.h"
    text = src.read_text(encoding="latin-1")
    coeff = [_array(text, f"datasetRGB{c}") for c in (1, 2, 3)]
    rad = [_array(text, f"datasetRGBRad{c}") for c in (1, 2, 3)]
    assert all(len(a) == 2 * 10 * 6 * 9 for a in coeff), [len(a) for a in coeff]
    assert all(len(a) == 2 * 10 * 6 for a in rad), [len(a) for a in rad]
    doc = {
        "model": "Hosek-Wilkie sky model, RGB datasets, version 1.4a",
        "license": _license(text),
        "layout": {"coefficients": "[channel][albedo 0..1][turbidity 1..10][control point 0..5][parameter 0..8]",
                   "radiances": "[channel][albedo 0..1][turbidity 1..10][control point 0..5]"},
        "coefficients": coeff,
        "radiances": rad,
    }
    out = HERE / "hosek_rgb_dataset.json"
    out.write_text(json.dumps(doc, indent=0))
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main(*sys.argv[1:])
