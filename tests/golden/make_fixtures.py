"""Writes the small golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/make_fixtures.py

Fixtures (data only):
  fft_vectors.npz   numpy.fft forward/inverse transforms of seeded complex
                    vectors, including lengths with prime factors above 61
                    (the oracle's Bluestein path) and the padded lengths of the
                    reference runner tests (6, 12, 16).
  reference_kats.json
                    known-answer values copied from the reference's own tests:
                    crates/fullmag-fdm-demag/src/newell.rs:676-907 and
                    crates/fullmag-engine/src/lib.rs:3001-3028, and the
                    explicit 3x3x1 masked plan of
                    crates/fullmag-runner/src/native_fdm.rs:859-900.
  xorshift_states.json
                    xorshift64 (<<13, >>7, <<17) states for seed 42, computed
                    here with Python's arbitrary-precision integers masked to
                    64 bits, and the unit vectors the planner generator
                    (crates/fullmag-plan/src/lib.rs:1927-1957) makes from them.
"""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK64 = (1 << 64) - 1


def fft_fixture():
    rng = np.random.default_rng(20260501)
    out = {}
    for n in [1, 2, 3, 5, 6, 7, 12, 16, 24, 60, 97, 101, 128, 194, 256, 384, 1000]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x_{n}"] = x
        out[f"fwd_{n}"] = np.fft.fft(x)
        out[f"inv_{n}"] = np.fft.ifft(x) * n
    x3 = rng.standard_normal((4, 6, 10)) + 1j * rng.standard_normal((4, 6, 10))
    out["x3"] = x3
    out["fwd3"] = np.fft.fftn(x3)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def xorshift_fixture():
    state = 42
    states = []
    for _ in range(12):
        state ^= (state << 13) & MASK64
        state ^= state >> 7
        state ^= (state << 17) & MASK64
        states.append(state)
    vecs = []
    for i in range(4):
        c = [float(s) / float(MASK64) * 2.0 - 1.0 for s in states[3 * i:3 * i + 3]]
        nrm = float(np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]))
        vecs.append([c[0] / nrm, c[1] / nrm, c[2] / nrm])
    with open(os.path.join(HERE, "xorshift_states.json"), "w") as f:
        json.dump({"seed": 42, "states": [str(s) for s in states], "unit_vectors": vecs}, f, indent=1)


def kat_fixture():
    kats = {
        "newell": {
            "source": "crates/fullmag-fdm-demag/src/newell.rs:676-907",
            "nxy_10x10x1_unit": {"grid": [10, 10, 1], "cell": [1.0, 1.0, 1.0],
                                  "points": [[1, 1, 0, -4.55648226e-02, 1e-5],
                                             [2, 2, 0, -5.28968456e-03, 1e-6],
                                             [5, 5, 0, -3.37642932e-04, 5e-8]]},
            "trace_tol": 1e-10,
            "cubic_nxx": [1.0 / 3.0, 1e-6],
            "asymptotic_8_8_rel": 1e-3,
        },
        "exchange": {
            "source": "crates/fullmag-engine/src/lib.rs:3001-3028",
            "stencil_center": {"m": [[1, 0, 0], [0, 1, 0], [1, 0, 0]], "expect": [2.0, -2.0, 0.0]},
            "masked": {"mask": [1, 1, 0], "m": [[1, 0, 0], [0, 1, 0], [0.7, 0.3, 0.0]],
                       "expect1": [1.0, -1.0, 0.0], "expect2": [0.0, 0.0, 0.0]},
        },
        "runner_masked_plan": {
            "source": "crates/fullmag-runner/src/native_fdm.rs:859-900",
            "grid": [3, 3, 1], "cell": [5e-9, 5e-9, 10e-9],
            "active_mask": [1, 1, 1, 1, 0, 1, 1, 1, 0],
            "m0": [[1.0, 0.0, 0.0],
                   [0.9950041652780258, 0.09983341664682815, 0.0],
                   [0.9800665778412416, 0.19866933079506122, 0.0],
                   [0.9992009587217894, 0.0, 0.03996803834887158],
                   [0.9937606691655043, 0.09970865087213879, 0.04972948160146045],
                   [0.9778332467629838, 0.19771314245924698, 0.06988589031642899],
                   [0.9968017063026194, -0.039904089712529575, 0.06972124896577284],
                   [0.9892364775387807, 0.05946310942269411, 0.1338082836649087],
                   [0.9711213242426827, 0.15730105252897553, 0.17902957342582418]],
            "ms": 800e3, "a": 13e-12, "alpha": 0.1, "gamma": 2.211e5,
            "dt": 2.5e-13, "h_ext": [1.5e3, -2.0e3, 7.5e2],
            "tolerances": {"m_rel": 5e-6, "m_abs": 1e-8, "field_rel": 5e-5, "field_abs": 1e-2,
                           "energy_rel": 5e-6, "single_vs_double_m": 1e-5,
                           "single_vs_double_e_rel": 1e-4},
        },
        "thin_film_plan": {
            "source": "crates/fullmag-runner/src/native_fdm.rs:902-947,1422-1492",
            "grid": [8, 6, 1], "h_demag_rel": 5e-4, "h_demag_abs": 0.1,
        },
        "tier_a": {
            "source": "native/backends/fdm/tests/tier_a_compare.cu:138-283",
            "grid": [8, 8, 4], "cell": [2e-9, 2e-9, 2e-9], "ms": 8e5, "a": 13e-12, "alpha": 0.5,
            "gamma": 2.211e5, "dt": 1e-14, "steps": 100, "m_tol": 1e-12, "e_rel_tol": 1e-10,
        },
        "tier_b": {
            "source": "native/backends/fdm/tests/tier_b_compare.cu:57-127",
            "m_tol": 1e-5, "e_rel_tol": 1e-4,
        },
    }
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)


if __name__ == "__main__":
    fft_fixture()
    xorshift_fixture()
    kat_fixture()
    print("fixtures written to", HERE)
