"""Writes tests/golden/head_on_sh_coefficients.json: the seeded real
spherical-harmonic coefficients (2 <= l <= 8, one set per droplet) of the
head-on collision workload, as lbm_amd.configs.sh_coefficients draws them.

    python tests/golden/make_head_on_coefficients.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lbm_amd import configs  # noqa: E402


def main():
    coeffs = configs.sh_coefficients(2)
    rec = dict(seed=configs.SH_SEED, lmin=configs.SH_LMIN, lmax=configs.SH_LMAX,
               generator="numpy.random.default_rng(seed).standard_normal, one draw per droplet",
               modes=[list(lm) for lm in configs.sh_modes()],
               coefficients=[[float.hex(float(x)) for x in a] for a in coeffs])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "head_on_sh_coefficients.json")
    with open(out, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
