"""Extract equilibrium-initialisation vectors from LUMA's own test-suite outputs.

Source: LUMA/cases/testsuite/base/case00N/io_lite.Lev0.Reg0.Rnk0.0.dat (the t = 0 dumps
written after f = feq(rho, u), GridObj_init_grids.cpp:315-333).  Each row there holds
RHO, UX, UY and the nine D2Q9 populations F0..F8 in LUMA's c_opt order, printed with 8
decimals (header at line 3 of each file).  This script keeps the distinct
(rho, ux, uy, F0..F8) tuples and writes them to t0_feq_d2q9.json as data.  It reads the
reference tree only when run by hand; tests read the committed JSON.
"""
import json
from pathlib import Path

BASE = Path("/root/reference/LUMA/cases/testsuite/base")
OUT = Path(__file__).resolve().parent / "t0_feq_d2q9.json"
# cases 001 and 009 restart from a previous case's restart file, so their t = 0 state is
# not an equilibrium initialisation (LUMA/cases/testsuite/scripts/runTestSuite.sh:108-114)
RESTARTS = {"case001", "case009"}


def main():
    rows = {}
    for case in sorted(BASE.glob("case00*")):
        if case.name in RESTARTS:
            continue
        p = case / "io_lite.Lev0.Reg0.Rnk0.0.dat"
        if not p.exists():
            continue
        with open(p) as fh:
            lines = fh.read().splitlines()
        hdr = lines[2].split()
        ci = {name: i for i, name in enumerate(hdr)}
        for ln in lines[3:]:
            t = ln.split()
            if len(t) < len(hdr) - 1:
                continue
            key = tuple(t[ci[c]] for c in ["RHO", "UX", "UY"] + [f"F{v}" for v in range(9)])
            rows.setdefault(key, case.name)
    vecs = [{"case": c, "rho": float(k[0]), "ux": float(k[1]), "uy": float(k[2]),
             "f": [float(x) for x in k[3:]]} for k, c in rows.items()]
    OUT.write_text(json.dumps({"source": "LUMA/cases/testsuite/base/case00N/io_lite.Lev0.Reg0.Rnk0.0.dat",
                               "decimals": 8, "vectors": vecs}, indent=1))
    print(f"{len(vecs)} vectors -> {OUT}")


if __name__ == "__main__":
    main()
