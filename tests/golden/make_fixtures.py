"""Writes the golden fixtures of tests/test_golden.py and
tests/test_gpu_golden.py from the CPU oracle (oracle/lem_oracle.c).

Each small case stores its inputs (DEM, rain, parameters) and the oracle's
state after a few timesteps; the 512^2 cases store SHA-256 digests of the
generated input and of the outputs (the arrays themselves would be several
MB).  The oracle is pinned by the independent checks in tests/test_oracle*.py;
these fixtures freeze what it computed then, so a later change to the oracle
or to the HIP path that moves any bit is caught.

    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from parallem_amd import synth  # noqa: E402

FIELDS_F = ["dem", "soilT", "stone", "fines", "fa", "depo", "soilM", "TotB"]
FIELDS_I = ["fd", "SFD", "mask"]
SCALARS = ["FA_max", "totE", "totI", "totG", "totD", "totwC", "totwP", "totbio", "sinkcount",
           "plateau_rounds1", "plateau_rounds2", "fa_uncomputed", "sed_uncomputed", "totwater"]


def valley():
    """A valley whose lowest cell (the outlet) sits mid-row, so that
    fd[outlet] = WEST closes a 2-cycle with its west neighbour (D10)."""
    H, W = 40, 41
    z = np.full((H, W), -9999.0)
    r, c = np.mgrid[0:H, 0:W]
    inner = (r > 0) & (r < H - 1) & (c > 0) & (c < W - 1)
    z[inner] = (10.0 * np.abs(c - 20) + 3.0 * r + 0.01 * ((r * 7 + c * 3) % 5))[inner]
    return z, np.full(z.shape, 1000.0)


def cases():
    d = synth.fbm_512(side=48, seed_offset=101)
    f = synth.fbm_512(side=40, seed_offset=102)
    z, rain = valley()
    return [
        ("fbm48_sfd", d.z, d.rain, 0, "nonneg", 10, 3),
        ("fbm48_mfd", d.z, d.rain, 1, "nonneg", 10, 3),
        ("valley_outlet_cycle", z, rain, 0, "nonneg", 10, 2),
        ("fbm40_formula_nan", f.z, f.rain, 0, "formula", 10, 2),
        ("fbm48_runin0", d.z, d.rain, 0, "nonneg", 0, 3),
    ]


def run_oracle(z, rain, flowtype, mode, runiniter, steps):
    H, W = z.shape
    orc = O.Oracle(z, rain, O.default_params(W, H, flowtype, runiniter=runiniter), soilT_mode=mode)
    for k in range(steps):
        orc.step(-1000 + k)
    return orc


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


BIG = [("fbm512_sfd", 0), ("fbm512_mfd", 1)]
BIG_STEPS = 3


def main():
    manifest = {"small": {}, "big": {}}
    for name, z, rain, ft, mode, runin, steps in cases():
        orc = run_oracle(z, rain, ft, mode, runin, steps)
        out = {f"in_{k}": v for k, v in (("dem", z), ("rain", rain))}
        out.update({f: orc[f] for f in FIELDS_F + FIELDS_I})
        sc = orc.scalars()
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        manifest["small"][name] = {"flowtype": ft, "soilT_mode": mode, "runiniter": runin, "steps": steps,
                                   "outlet": int(orc.s.outlet), "scalars": {k: sc[k] for k in SCALARS}}
    for name, ft in BIG:
        d = synth.fbm_512()
        orc = run_oracle(d.z, d.rain, ft, "nonneg", 10, BIG_STEPS)
        manifest["big"][name] = {"flowtype": ft, "steps": BIG_STEPS, "input_dem": digest(d.z),
                                 "input_rain": digest(d.rain), "outlet": int(orc.s.outlet),
                                 "outputs": {f: digest(orc[f]) for f in FIELDS_F + FIELDS_I}}
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
