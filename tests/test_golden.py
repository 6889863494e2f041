"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py
from the oracle after a brute-force and reference cross-check): the oracle
still reproduces every digest, and the fixture's coefficients and input are
what the library's generators produce."""
import importlib.util
import json
import os

import numpy as np
import pytest

import oracle as O
from bricklib_amd import stencils as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gen():
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(HERE, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


G = _gen()
META = json.load(open(os.path.join(HERE, "sweeps.json")))
ARR = np.load(os.path.join(HERE, "sweeps.npz"))


def test_fixture_inputs_match_generators():
    for cname, (sname, cond_mode) in G.CASES.items():
        st = S.Stencil.make(sname, seed=G.SEED_COEFF, cond_mode=cond_mode)
        np.testing.assert_array_equal(ARR[f"coeff_{cname}"], st.coeff)
        assert ARR[f"theta_{cname}"][0] == st.theta
    _, inp = G.run_case((2, 2, 2), O.STAR7, ARR["coeff_7pt"], 0.0, 1, 0, O.MODE_FMA)
    np.testing.assert_array_equal(inp, ARR["in_16"])


@pytest.mark.parametrize("cname", list(G.CASES))
def test_oracle_reproduces_digests(cname):
    case = META["cases"][cname]
    coeff, theta = ARR[f"coeff_{cname}"], float(ARR[f"theta_{cname}"][0])
    for key, want in case["out"].items():
        gname, ns, oname = key.split("/")
        _, out = G.run_case(tuple(META["grids"][gname]), case["kind"], coeff, theta,
                            case["cond_mode"], int(ns), G.ORDERS[oname])
        assert G.digest(out) == want["sha256"], key
        if gname == "16" and ns == "1" and oname == "fma":
            np.testing.assert_array_equal(out, ARR[f"out_{cname}_16_1_fma"])


def test_reference_cross_check_recorded():
    # the 7-point and reference-form cond cases were checked against the
    # reference's own loops when the fixture was made
    assert len(META["ref_checked"]) == 8
    for c in META["cases"].values():
        for v in c["out"].values():
            assert v["brute_force_rel"] <= 1e-12
