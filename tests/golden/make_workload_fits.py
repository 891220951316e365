"""Writes the wall materials of the seeded ADE workloads as .fit files (the reference's
section grammar: # R_inf / # K L / # A_k zeta_k / # B_l C_l alpha_l beta_l), one per wall
slot, into tests/golden/workload_fits/<workload>/wall<i>.fit. Slots without a material (the
cPML face, R = 0) get no file. Run from the repository root:

    python tests/golden/make_workload_fits.py
"""
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fases_amd import workloads as W          # noqa: E402
from fases_amd.solver import write_fit        # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent / "workload_fits"
CONFIGS = {"shoebox_256x192x128_seed1001": 1, "shoebox_512x384x256_seed1002": 2, "shoebox_192x160x128_seed1004": 4}


def materials_of(n):
    return W.config(n).materials


def main():
    for name, n in CONFIGS.items():
        m = materials_of(n)
        d = OUT / name
        d.mkdir(parents=True, exist_ok=True)
        for i in range(m.n):
            if m.Rinf[i] == 0.0 and not m.A[i].any() and not m.B[i].any():
                continue
            write_fit(str(d / ("wall%d.fit" % i)), m.Rinf[i], m.A[i], m.zeta[i], m.B[i], m.C[i], m.alpha[i],
                      m.beta[i])
        print("wrote", d)


if __name__ == "__main__":
    main()
