"""Write the reference-held data fixtures used by the CPU tests.

Reads (as text, nothing executed) data files that ship in the reference tree
and writes small CSV subsets next to this script:

  sod_exact_t0p1.csv       <- Exec/RegTests/EB-C7/exact_soln
      exact Sod solution at t = 0.1 (x, pressure, density, Temp, u),
      interface at x = 0, left (rho, p) = (1, 1), right (0.125, 0.1)
  pmf_lidryer_rows.csv     <- Exec/RegTests/PMF/LiDryer_H2_p1_phi0_4000tu0300.dat
      premixed H2/air flame (LiDryer), p = 1 atm: x, T, u, rho and the nine
      species columns (mole fractions)
  tgreact_hp_equilibrium.csv <- Exec/RegTests/TGReact/profiles.dat
      unburnt H2/O2/N2 mass fractions at 300 K, 1 atm, and the HP
      equilibrium (T_adiabatic, nine mass fractions) computed with the
      LiDryer mechanism

Run from the repository root:  python tests/golden/make_reference_fixtures.py
"""
import csv
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def sod():
    src = os.path.join(REF, "Exec/RegTests/EB-C7/exact_soln")
    rows = []
    with open(src) as f:
        hdr = f.readline().split()
        for line in f:
            if line.strip():
                rows.append(line.split())
    with open(os.path.join(HERE, "sod_exact_t0p1.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(hdr)
        for r in rows[::2]:
            w.writerow(r)


def pmf():
    src = os.path.join(REF, "Exec/RegTests/PMF/LiDryer_H2_p1_phi0_4000tu0300.dat")
    with open(src) as f:
        lines = f.read().splitlines()
    names = [s.strip('"') for s in lines[0].split("=")[1].split()]
    data = [l.split() for l in lines[2:] if l.strip()]
    pick = list(range(0, len(data), 40)) + [len(data) - 1]
    with open(os.path.join(HERE, "pmf_lidryer_rows.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(names)
        for i in sorted(set(pick)):
            w.writerow(data[i])


def tgreact():
    src = os.path.join(REF, "Exec/RegTests/TGReact/profiles.dat")
    with open(src) as f:
        rows = list(csv.reader(f))
    with open(os.path.join(HERE, "tgreact_hp_equilibrium.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(rows[0])
        for r in rows[1::8]:
            w.writerow(r)


if __name__ == "__main__":
    sod()
    pmf()
    tgreact()
