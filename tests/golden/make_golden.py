"""Generate the golden sweep fixtures: tests/golden/sweeps.npz and
tests/golden/sweeps.json.

The reference ships no golden vectors (its tests only compare the brick
loops against its own array loop at 1e-6 on unseeded random inputs), so the
fixtures come from the oracle, cross-checked before they are written:

  * every case, both evaluation orders, against a numpy.roll brute force of
    the same periodic stencil on the dense field (a textbook periodic
    stencil, independent of bricks and adjacency): relative error <= 1e-13
    after 1 sweep and <= 1e-12 after 8;
  * when oracle/_ref is built (this container), the reference-order outputs
    of the 7-point star and the reference-form cond bit for bit against the
    reference's own accessor loops (stencils/3axis.cpp:51-66, :123-140)
    compiled from its headers.

Cases: grids 16^3 (2x2x2 bricks) and 24x32x40 (3x4x5 bricks, non-cubic),
periodic, one ghost-brick layer, input U[-1,1] from the seeded fill (seed
20240); stencils 7pt, 13pt, 25pt, 125pt, cond (reference form) and cond
(threshold form), coefficients from Stencil.make(name, seed=77) stored in
the fixture; outputs after 1 and 8 ping-pong sweeps (periodic refresh before
each) in the library's FMA order ("fma": what every GPU schedule reproduces
bit for bit) and in the reference's textual order ("ref").

sweeps.json holds, per (case, grid, sweeps, order), the SHA-256 of the dense
interior output (float64, little-endian, z-y-x order), its L2 norm and max
|x|; sweeps.npz holds the coefficients, theta, the 16^3 input and the 16^3
one-sweep FMA outputs as dense arrays.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from bricklib_amd import stencils as S  # noqa: E402

SEED_IN = 20240
SEED_COEFF = 77
GRIDS = {"16": (2, 2, 2), "24x32x40": (3, 4, 5)}
CASES = {  # name -> (stencil name, cond_mode)
    "7pt": ("7pt", 1), "13pt": ("13pt", 1), "25pt": ("25pt", 1), "125pt": ("125pt", 1),
    "cond_ref": ("cond", 0), "cond_theta": ("cond", 1)}
SWEEPS = (1, 8)
ORDERS = {"fma": O.MODE_FMA, "ref": O.MODE_REF}


def digest(dense: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(dense, "<f8").tobytes()).hexdigest()


def dense_ref(field, kind, coeff, theta, cond_mode):
    """numpy.roll brute force on a periodic (z, y, x) field."""
    relu = kind == O.COND7
    acc = np.zeros_like(field)
    for c, (di, dj, dk) in zip(coeff, S.offsets(kind)):
        x = np.roll(field, shift=(-dk, -dj, -di), axis=(0, 1, 2))
        if relu:
            x = np.where(x < 0, 0.0, x)
        acc += c * x
    if relu:
        sel = acc > 0 if cond_mode == 0 else field > theta
        acc = np.where(sel, acc, -acc)
    return acc


def run_case(nb, kind, coeff, theta, cond_mode, nsweeps, mode):
    g = O.Grid(nb)
    a, b = g.storage(), g.storage()
    g.fill_uniform(a, SEED_IN)
    res = g.sweeps(kind, coeff, a, b, nsweeps, mode=mode, theta=theta, cond_mode=cond_mode)
    return g, g.to_dense(a if nsweeps == 0 else res)


def ref_check(nb, kind, coeff, theta, cond_mode, nsweeps, dense):
    """Bit-for-bit against the reference's accessor loops (7pt, cond reference form)."""
    if not O.ref_available() or not (kind == O.STAR7 or (kind == O.COND7 and cond_mode == 0)):
        return None
    g = O.Grid(nb)
    a, b = g.storage(), g.storage()
    g.fill_uniform(a, SEED_IN)
    src, dst = a, b
    fn = O.ref().ref_sweep_7pt if kind == O.STAR7 else O.ref().ref_sweep_cond
    for _ in range(nsweeps):
        g.refresh(src)
        fn(O.P(g.adj), g.nbricks, O.P(g.interior), g.interior.size, O.P(src), C.c_size_t(512),
           O.P(dst), C.c_size_t(512), O.P(np.ascontiguousarray(coeff, np.float64)))
        src, dst = dst, src
    return bool(np.array_equal(g.to_dense(src), dense))


def main():
    meta = {"seed_in": SEED_IN, "seed_coeff": SEED_COEFF, "grids": {k: list(v) for k, v in
                                                                     GRIDS.items()},
            "cases": {}, "ref_checked": []}
    arrays = {}
    for cname, (sname, cond_mode) in CASES.items():
        st = S.Stencil.make(sname, seed=SEED_COEFF, cond_mode=cond_mode)
        coeff = np.ascontiguousarray(st.coeff, np.float64)
        arrays[f"coeff_{cname}"] = coeff
        arrays[f"theta_{cname}"] = np.array([st.theta])
        meta["cases"][cname] = {"stencil": sname, "kind": int(st.kind), "cond_mode": cond_mode,
                                "theta": float(st.theta), "out": {}}
        for gname, nb in GRIDS.items():
            _, inp = run_case(nb, st.kind, coeff, st.theta, cond_mode, 0, O.MODE_FMA)
            if gname == "16":
                arrays["in_16"] = inp
            for ns in SWEEPS:
                truth = inp
                for _ in range(ns):
                    truth = dense_ref(truth, st.kind, coeff, st.theta, cond_mode)
                for oname, mode in ORDERS.items():
                    _, out = run_case(nb, st.kind, coeff, st.theta, cond_mode, ns, mode)
                    rel = np.max(np.abs(out - truth)) / max(np.max(np.abs(truth)), 1e-300)
                    bound = 1e-13 if ns == 1 else 1e-12
                    if not rel <= bound:
                        raise SystemExit(f"{cname} {gname} {ns} {oname}: brute force rel {rel}")
                    if oname == "ref":
                        ok = ref_check(nb, st.kind, coeff, st.theta, cond_mode, ns, out)
                        if ok is not None:
                            if not ok:
                                raise SystemExit(f"{cname} {gname} {ns}: differs from _ref")
                            meta["ref_checked"].append(f"{cname}/{gname}/{ns}")
                    key = f"{gname}/{ns}/{oname}"
                    meta["cases"][cname]["out"][key] = {
                        "sha256": digest(out), "l2": float(np.sqrt(np.sum(out * out))),
                        "maxabs": float(np.max(np.abs(out))), "brute_force_rel": float(rel)}
                    if gname == "16" and ns == 1 and oname == "fma":
                        arrays[f"out_{cname}_16_1_fma"] = out
    np.savez_compressed(os.path.join(HERE, "sweeps.npz"), **arrays)
    with open(os.path.join(HERE, "sweeps.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"wrote {len(arrays)} arrays, {sum(len(c['out']) for c in meta['cases'].values())} "
          f"digests; checked against _ref: {len(meta['ref_checked'])}")


if __name__ == "__main__":
    main()
